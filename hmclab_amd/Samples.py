"""Sample files with the reference's NPY layout (hmclab/Samples.py:16-335).

A ``.npy`` samples file holds an (N, d+1) float64 array on disk (row k =
sample k, last column = misfit) and is read back transposed to (d+1, N)
(Samples.py:161,325).  Attributes live in a pickle side-file
``<filename>.pkl`` (Samples.py:165-174).  The writer appends rows and
rewrites a fixed-size header in place, so the file is a valid ``.npy``
after every append.  HDF5 (``.h5``) output needs h5py, which this image
does not provide; asking for it raises a clear error.
"""
from __future__ import annotations

import os
import pickle
import struct

import numpy as np

_HEADER_BYTES = 256   # magic + length + padded dict; room for any shape


def _header(rows: int, cols: int) -> bytes:
    d = "{'descr': '<f8', 'fortran_order': False, 'shape': (%d, %d), }" % (rows, cols)
    body_len = _HEADER_BYTES - 10
    d = d.ljust(body_len - 1) + "\n"
    return b"\x93NUMPY" + bytes([1, 0]) + struct.pack("<H", body_len) + d.encode("latin1")


class _AppendableNPY:
    """Row-appendable float64 .npy file with a fixed-size version-1.0 header."""

    def __init__(self, filename: str, cols: int):
        self.filename = filename
        self.cols = cols
        self.rows = 0
        self.fp = open(filename, "wb")
        self.fp.write(_header(0, cols))

    def append(self, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype="<f8")
        assert rows.ndim == 2 and rows.shape[1] == self.cols
        self.fp.seek(0, os.SEEK_END)
        rows.tofile(self.fp)
        self.rows += rows.shape[0]
        self.fp.seek(0)
        self.fp.write(_header(self.rows, self.cols))
        self.fp.flush()

    def close(self):
        if self.fp is not None:
            self.fp.close()
            self.fp = None


def _check_extension(filename: str) -> tuple[str, str]:
    stem, ext = os.path.splitext(filename)
    if ext in ("", ".h5"):
        try:
            import h5py  # noqa: F401
        except ModuleNotFoundError:
            raise ModuleNotFoundError(
                "HDF5 samples files need h5py, which is not installed; use a .npy filename"
            ) from None
        raise NotImplementedError("HDF5 samples files are not written by this engine; use .npy")
    if ext != ".npy":
        raise AttributeError(f"Unkown extension `{ext}` for samples file.")
    return stem, ext


class SamplesWriter:
    """Writer side of ``Samples(filename, mode="w", overwrite=...)``."""

    def __init__(self, filename: str, dimensions: int, overwrite: bool = False):
        _check_extension(filename)
        directory = os.path.dirname(filename)
        if directory != "" and not os.path.isdir(directory):
            raise NotADirectoryError(
                f"Trying to write a samples file to a non-existent directory `{directory}`."
            )
        if not overwrite and (os.path.isfile(filename) or os.path.isfile(f"{filename}.pkl")):
            shown = filename
            if os.path.isfile(f"{filename}.pkl"):
                shown += f"` or attributes file `{filename}.pkl"
            raise FileExistsError(f"Trying to write samples to an already existing file `{shown}`.")
        self.filename = filename
        self.attributes = {}
        self._arr = _AppendableNPY(filename, dimensions + 1)
        self.write_attribute("write_index", 0)
        self.write_attribute("last_written_sample", -1)

    def write_attribute(self, name, value):
        self.attributes[name] = value
        with open(f"{self.filename}.pkl", "wb") as f:
            pickle.dump(self.attributes, f)

    def append(self, columns: np.ndarray):
        """Append samples given as columns (d+1, k), the reference's layout."""
        columns = np.asarray(columns, dtype=np.float64)
        self._arr.append(columns.T)
        k = columns.shape[1]
        self.attributes["write_index"] = self.attributes["write_index"] + k
        self.attributes["last_written_sample"] = self.attributes["last_written_sample"] + k
        with open(f"{self.filename}.pkl", "wb") as f:
            pickle.dump(self.attributes, f)

    def close(self):
        self._arr.close()
        with open(f"{self.filename}.pkl", "wb") as f:
            pickle.dump(self.attributes, f)


class Samples:
    """Reader of a samples file (Samples.py:16-278, NPY branch)."""

    def __init__(self, filename: str, burn_in: int = None):
        _check_extension(filename)
        if not os.path.isfile(filename):
            raise FileNotFoundError(f"Trying to read samples file `{filename}` which does not exist.")
        self.filename = filename
        self._numpy_array = np.load(filename, mmap_mode="r").T
        with open(f"{filename}.pkl", "rb") as f:
            self._numpy_attributes = pickle.load(f)
        self.burn_in = 0 if burn_in is None else burn_in
        self.last_sample = self.read_attribute("write_index")
        if self.last_sample <= self.burn_in:
            raise ValueError(
                f"The burn-in phase is longer than the chain itself. "
                f"Total samples before burn in: {self.last_sample}"
            )

    def read_attribute(self, name):
        return self._numpy_attributes[name]

    @property
    def numpy(self):
        return self._numpy_array[:, self.burn_in:]

    @property
    def samples(self):
        return self.numpy[:-1, :]

    @property
    def misfits(self):
        return self.numpy[-1, :]

    def __getitem__(self, key):
        return self.numpy[key]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        self._numpy_array = None

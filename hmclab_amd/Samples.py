"""Sample files with the reference's NPY layout (hmclab/Samples.py:16-335).

A ``.npy`` samples file holds an (N, d+1) float64 array on disk (row k =
sample k, last column = misfit) and is read back transposed to (d+1, N)
(Samples.py:161,325).  Attributes live in a pickle side-file
``<filename>.pkl`` (Samples.py:165-174).  The writer appends rows and
rewrites the header in place, so the file is a valid ``.npy`` after every
append.  The header is NPY version 2.0 with 64 spare bytes for the shape to
grow, the layout the reference's appender reopens (it accepts version 2.0
only, with exactly that header length, hmclab/Helpers/AppendNPY.py:65-85),
so a reference process can keep appending to a file written here.  HDF5
(``.h5``) output needs h5py, which this image does not provide; asking for it
raises a clear error.
"""
from __future__ import annotations

import io
import os
import pickle
import struct

import numpy as np

SPARE_HEADER_BYTES = 64


def npy_header(rows: int, cols: int, spare: bool = True) -> bytes:
    """NPY 2.0 header of a C-order float64 (rows, cols) array; with `spare`,
    64 more bytes of blank padding (counted in the header length) so the
    shape can grow without moving the data."""
    buf = io.BytesIO()
    np.lib.format.write_array_header_2_0(buf, {"descr": "<f8", "fortran_order": False, "shape": (rows, cols)})
    h = bytearray(buf.getvalue())
    if spare:
        h[-1:] = b" " * SPARE_HEADER_BYTES + b"\n"      # the closing newline moves to the end
        struct.pack_into("<I", h, 8, len(h) - 12)
    return bytes(h)


class _AppendableNPY:
    """Row-appendable float64 .npy file whose header length stays fixed."""

    def __init__(self, filename: str, cols: int):
        self.filename = filename
        self.cols = cols
        self.rows = 0
        self.fp = open(filename, "wb")
        self.header = npy_header(0, cols)
        self.fp.write(self.header)

    def _header_for(self, rows: int) -> bytes:
        h = npy_header(rows, self.cols)
        if len(h) != len(self.header):
            # the shape grew into the spare bytes
            h = npy_header(rows, self.cols, spare=False)
        if len(h) != len(self.header):
            raise OSError(f"samples file {self.filename}: no header room left for shape ({rows}, {self.cols})")
        return h

    def append(self, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype="<f8")
        assert rows.ndim == 2 and rows.shape[1] == self.cols
        self.fp.seek(0, os.SEEK_END)
        rows.tofile(self.fp)
        self.rows += rows.shape[0]
        self.fp.seek(0)
        self.fp.write(self._header_for(self.rows))
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

"""CPU tests of the product's host-side logic: tomography builder and ESS
against the oracle, bf16 rounding against torch, the samples file format,
and the reference-mirroring argument checks (which fail before any device
work)."""
import os
import pickle

import numpy as np
import pytest
import scipy.sparse as sp

from hmclab_amd import Distributions as D
from hmclab_amd import MassMatrices as MM
from hmclab_amd import Samplers, diagnostics, tomography
from hmclab_amd.engine import bf16_round, bf16_to_f32
from hmclab_amd.Samples import Samples, SamplesWriter
from oracle import ess as E
from oracle import tomography as T


def test_ray_matrix_matches_walk_oracle():
    rng = np.random.default_rng(3)
    nx, nz, h = 16, 16, 10.0
    S, R = tomography.random_rays(rng, 60, nx, nz, h, h)
    G = tomography.ray_matrix(S, R, nx, nz, h, h)
    Go = T.straight_ray_matrix(S, R, nx, nz, h, h)
    assert np.allclose(G.toarray(), Go.toarray(), atol=1e-9)
    assert np.allclose(np.asarray(G.sum(axis=1)).ravel(), np.hypot(*(R - S).T), rtol=1e-12)


def test_synthetic_tomography_shape():
    prob = tomography.synthetic_problem(seed=1, n_rays=200)
    G = prob["G"]
    assert G.shape == (200, 4096)
    nnz_row = np.diff(G.indptr)
    assert nnz_row.min() > 0 and 60 <= nnz_row.mean() <= 130


def test_ess_vectorised_equals_oracle():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(5, 120, 7)).cumsum(axis=1) * 0.1 + rng.normal(size=(5, 120, 7))
    x[:, 30:40, 2] = x[:, 30:31, 2]          # repeated values (rejections) -> ties
    got = diagnostics.ess_bulk(x)
    ref = np.array([E.ess_bulk(x[:, :, k]) for k in range(7)])
    assert np.allclose(got, ref, rtol=1e-10)


def test_ess_torch_path_equals_numpy():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4, 80, 5)).cumsum(axis=1)
    x[:, 10:20, 1] = x[:, 10:11, 1]
    assert np.allclose(diagnostics.ess_bulk(torch.from_numpy(x)), diagnostics.ess_bulk(x), rtol=1e-9)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32) * 1e-2,
                        np.array([0.0, -0.0, 1.0, 65504.0, 1e-40, np.inf, -np.inf], dtype=np.float32)])
    ours = bf16_to_f32(bf16_round(x))
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)


def test_samples_file_roundtrip(tmp_path):
    fn = str(tmp_path / "s.npy")
    w = SamplesWriter(fn, 3)
    w.write_attribute("proposals", 4)
    cols = np.arange(16, dtype=float).reshape(4, 4)
    w.append(cols[:, :1])
    w.append(cols[:, 1:4])
    w.close()
    arr = np.load(fn)
    assert arr.shape == (4, 4) and np.array_equal(arr.T, cols)
    with open(fn + ".pkl", "rb") as f:
        attrs = pickle.load(f)
    assert attrs["write_index"] == 4 and attrs["last_written_sample"] == 3
    with Samples(fn, burn_in=1) as s:
        assert s.numpy.shape == (4, 3)
        assert np.array_equal(s.misfits, cols[-1, 1:])
    with pytest.raises(FileExistsError):
        SamplesWriter(fn, 3)
    SamplesWriter(fn, 3, overwrite=True).close()
    with pytest.raises(ModuleNotFoundError, match="h5py"):
        SamplesWriter(str(tmp_path / "x.h5"), 3)


def test_samples_file_header_is_npy_2_0_with_spare_room(tmp_path):
    """The samples file is NPY 2.0 and its header is as long as a fresh 2.0
    header of the current shape plus 64 spare bytes, the two conditions under
    which the reference's appender reopens a file (AppendNPY.py:65-85); when
    the shape outgrows the spare bytes the header keeps its length."""
    from hmclab_amd.Samples import _AppendableNPY, npy_header

    fn = str(tmp_path / "v2.npy")
    w = SamplesWriter(fn, 4)
    cols = np.arange(15.0).reshape(5, 3)
    w.append(cols)
    w.append(cols + 100)
    w.close()
    with open(fn, "rb") as f:
        assert np.lib.format.read_magic(f) == (2, 0)
        shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
        assert (shape, fortran, dtype) == ((6, 5), False, np.dtype("<f8"))
        assert f.tell() == len(npy_header(6, 5))
        assert f.tell() % 64 == 0
    arr = np.load(fn)
    assert np.array_equal(arr.T, np.hstack([cols, cols + 100]))
    # growth: find a row count whose shape no longer fits the spare bytes
    a = _AppendableNPY(str(tmp_path / "g.npy"), 5)
    n0 = len(a.header)
    grew = next(r for r in (10 ** k for k in range(1, 200)) if len(npy_header(r, 5)) != n0)
    assert len(a._header_for(grew)) == n0
    a.close()


def test_distribution_api_validation():
    n = D.Normal(np.zeros((3, 1)), 1.0)
    with pytest.raises(ValueError, match="Bounds vectors are incompatible."):
        n.update_bounds(np.ones((3, 1)), -np.ones((3, 1)))
    with pytest.raises(ValueError):
        D.LinearMatrix(np.ones((4, 3)), np.ones(4), 1.0)          # d must be a column
    with pytest.raises(ValueError):
        D.LinearMatrix(np.ones((5, 3)), np.ones((4, 1)), 1.0)     # shape mismatch
    lm = D.LinearMatrix(sp.eye(3, format="csr"), np.ones((3, 1)), np.ones((3, 1)))
    assert lm.dimensions == 3 and not lm.dense
    br = D.BayesRule([n, D.Uniform(-np.ones(3), np.ones(3)), lm])
    assert np.all(br.lower_bounds == -1) and np.all(br.upper_bounds == 1)
    x, p = np.full((3, 1), 1.5), np.ones((3, 1))
    br.corrector(x, p)
    assert np.allclose(x, 0.5) and np.allclose(p, -1.0)


def test_mass_matrix_host_methods():
    m = MM.Diagonal(np.array([1.0, 4.0]))
    p = np.array([[2.0], [2.0]])
    assert m.kinetic_energy(p) == 0.5 * (4 + 1)
    with pytest.raises(ValueError):
        m.kinetic_energy(np.ones((3, 1)))
    assert np.allclose(m._device_diagonal(), [1.0, 4.0])
    assert MM.Unit(2)._device_diagonal() is None


@pytest.mark.parametrize("kw,msg", [
    (dict(proposals=10, online_thinning=3), "multiple of the online thinning"),
    (dict(proposals=0), "larger than zero"),
    (dict(stepsize=-1.0), "Stepsize should be a float larger than zero."),
    (dict(amount_of_steps=0), "larger than zero"),
    (dict(autotuning=True, learning_rate=0.4), "learning rate"),
    (dict(mass_matrix=MM.Unit(434)), "mass matrix"),
])
def test_sampler_argument_checks(kw, msg):
    post = D.Normal(np.zeros((3, 1)), 1.0)
    with pytest.raises(AssertionError, match=msg):
        Samplers.HMC(seed=1).sample(None, post, **kw)


def test_sampler_unknown_integrator():
    with pytest.raises(ValueError, match="Unknown integrator"):
        Samplers.HMC().sample(None, D.Normal(np.zeros((2, 1)), 1.0), integrator="rk4")


def test_ess_rejects_too_short_chains():
    import pytest

    from hmclab_amd import diagnostics

    with pytest.raises(ValueError, match="at least 4 draws"):
        diagnostics.ess_bulk(np.zeros((3, 3, 2)))


def test_chunk_writer_keeps_order_and_reports_errors():
    import pytest

    from hmclab_amd.Samplers import _ChunkWriter

    seen = []
    w = _ChunkWriter(lambda X, chi: seen.append((X.copy(), chi.copy())))
    for c in range(5):
        w.put((np.full((3, 2, 4), c, dtype=float), np.full((3, 2), c, dtype=float)))
    w.close()
    assert [int(x[0, 0]) for x, _ in seen] == [0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4]

    def boom(X, chi):
        raise OSError("disk full")

    w = _ChunkWriter(boom)
    w.put((np.zeros((1, 1, 1)), np.zeros((1, 1))))
    with pytest.raises(OSError, match="disk full"):
        w.close()

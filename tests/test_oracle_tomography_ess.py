"""Pin the straight-ray oracle (closed forms, brute-force sampling, symmetry)
and the ESS oracle (iid and AR(1) closed forms)."""
import math

import numpy as np
import pytest

from oracle import ess as E
from oracle import tomography as T


def test_axis_aligned_rays():
    nx = nz = 8
    L = T.ray_cells_walk(0.0, 15.0, 80.0, 15.0, nx, nz, 10.0, 10.0)   # row iz=1
    assert sorted(L) == [8 + i for i in range(8)]
    assert np.allclose(list(L.values()), 10.0)
    V = T.ray_cells_walk(35.0, 80.0, 35.0, 0.0, nx, nz, 10.0, 10.0)   # column ix=3, downward
    assert sorted(V) == [iz * nx + 3 for iz in range(8)]
    assert np.allclose(list(V.values()), 10.0)


def test_diagonal_ray():
    nx = nz = 6
    D = T.ray_cells_walk(0.0, 0.0, 60.0, 60.0, nx, nz, 10.0, 10.0)
    assert sorted(D) == [i * nx + i for i in range(nx)]
    assert np.allclose(list(D.values()), 10.0 * math.sqrt(2.0))


def test_random_rays_row_sums_brute_force_symmetry():
    rng = np.random.default_rng(0)
    nx, nz, h = 12, 10, 10.0
    for _ in range(20):
        p = rng.uniform([0, 0], [nx * h, nz * h])
        q = rng.uniform([0, 0], [nx * h, nz * h])
        w = T.ray_cells_walk(*p, *q, nx, nz, h, h)
        assert np.isclose(sum(w.values()), np.hypot(*(q - p)), rtol=1e-12)
        bf = T.brute_force_lengths(*p, *q, nx, nz, h, h)
        dense = np.zeros(nx * nz)
        for j, v in w.items():
            dense[j] = v
        assert np.allclose(dense, bf, atol=2e-2)
        back = T.ray_cells_walk(*q, *p, nx, nz, h, h)
        assert set(back) == set(w) and all(np.isclose(back[j], w[j]) for j in w)


def test_ess_iid_and_ar1():
    rng = np.random.default_rng(1)
    M, N = 4, 2000
    iid = rng.normal(size=(M, N))
    assert abs(E.ess_bulk(iid) / (M * N) - 1.0) < 0.12
    rho = 0.6
    x = np.zeros((M, N))
    e = rng.normal(size=(M, N))
    x[:, 0] = e[:, 0] / math.sqrt(1 - rho * rho)
    for t in range(1, N):
        x[:, t] = rho * x[:, t - 1] + e[:, t]
    expect = M * N * (1 - rho) / (1 + rho)
    assert abs(E.ess_bulk(x) / expect - 1.0) < 0.2
    # a stuck chain (not mixing) has tiny ESS
    stuck = np.concatenate([np.zeros((1, N)), np.ones((1, N))], axis=0) + 1e-3 * rng.normal(size=(2, N))
    assert E.ess_bulk(stuck) < 50

"""Straight-ray travel-time tomography: CSR ray-length matrix and the
synthetic 64x64 problem.

The reference names straight-ray tomography as supported physics
(README.md:34) but ships no builder; its likelihood is the sparse
``LinearMatrix`` backend (LinearMatrix.py:309-426) with a ray-length matrix
G, travel times t = G s for slowness s.  This builder computes, per ray, the
parameters where the segment crosses vertical and horizontal grid lines,
sorts them, and assigns each sub-segment (by its midpoint) to a cell with
length = (t_{k+1} - t_k) * |ray|.  Cell ``j = iz * nx + ix`` covers
``[ix*dx, (ix+1)*dx) x [iz*dz, (iz+1)*dz)``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def ray_matrix(sources: np.ndarray, receivers: np.ndarray, nx: int, nz: int, dx: float,
               dz: float) -> sp.csr_matrix:
    """CSR (n_rays, nx*nz) of per-cell path lengths of straight rays."""
    S = np.asarray(sources, dtype=np.float64).reshape(-1, 2)
    R = np.asarray(receivers, dtype=np.float64).reshape(-1, 2)
    n = S.shape[0]
    gx = np.arange(nx + 1) * dx
    gz = np.arange(nz + 1) * dz
    rows, cols, vals = [], [], []
    for i in range(n):
        x0, z0 = S[i]
        x1, z1 = R[i]
        ddx, ddz = x1 - x0, z1 - z0
        L = np.hypot(ddx, ddz)
        if L == 0.0:
            continue
        ts = [np.array([0.0, 1.0])]
        if ddx != 0.0:
            ts.append((gx - x0) / ddx)
        if ddz != 0.0:
            ts.append((gz - z0) / ddz)
        t = np.concatenate(ts)
        t = np.unique(np.clip(t, 0.0, 1.0))
        seg = np.diff(t)
        keep = seg > 0.0
        tm = 0.5 * (t[:-1] + t[1:])[keep]
        seg = seg[keep]
        px, pz = x0 + tm * ddx, z0 + tm * ddz
        inside = (px >= 0) & (px < nx * dx) & (pz >= 0) & (pz < nz * dz)
        ix = np.floor(px[inside] / dx).astype(np.int64)
        iz = np.floor(pz[inside] / dz).astype(np.int64)
        j = iz * nx + ix
        lengths = seg[inside] * L
        # merge repeated cells (sub-segments split at a grid corner)
        uj, inv = np.unique(j, return_inverse=True)
        lv = np.bincount(inv, weights=lengths)
        rows.append(np.full(uj.size, i, dtype=np.int64))
        cols.append(uj)
        vals.append(lv)
    if rows:
        r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    else:
        r = c = np.zeros(0, dtype=np.int64)
        v = np.zeros(0)
    G = sp.csr_matrix((v, (r, c)), shape=(n, nx * nz))
    G.sort_indices()
    return G


def boundary_points(rng: np.random.Generator, n: int, nx: int, nz: int, dx: float, dz: float,
                    sides: np.ndarray | None = None):
    """Uniform points on the grid boundary; side 0..3 = bottom, right, top, left."""
    W, H = nx * dx, nz * dz
    if sides is None:
        sides = rng.integers(0, 4, size=n)
    u = rng.uniform(0.0, 1.0, size=n)
    x = np.where(sides == 0, u * W, np.where(sides == 1, W, np.where(sides == 2, u * W, 0.0)))
    z = np.where(sides == 0, 0.0, np.where(sides == 1, u * H, np.where(sides == 2, H, u * H)))
    return np.stack([x, z], axis=1), sides


def random_rays(rng: np.random.Generator, n: int, nx: int, nz: int, dx: float, dz: float):
    """Source/receiver pairs on different sides of the boundary."""
    s_side = rng.integers(0, 4, size=n)
    r_side = (s_side + rng.integers(1, 4, size=n)) % 4
    S, _ = boundary_points(rng, n, nx, nz, dx, dz, s_side)
    R, _ = boundary_points(rng, n, nx, nz, dx, dz, r_side)
    return S, R


def gaussian_smooth(field: np.ndarray, sigma_cells: float) -> np.ndarray:
    """Separable Gaussian filter (reflect boundaries) of a 2-D field."""
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(field, sigma=sigma_cells, mode="reflect")


def synthetic_problem(seed: int = 0, nx: int = 64, nz: int = 64, cell: float = 10.0,
                      n_rays: int = 2000, background: float = 1.0 / 3000.0,
                      perturbation_std: float = 5e-6, smoothing_cells: float = 2.0,
                      noise_std: float = 1e-4, prior_std: float = 1e-5):
    """The batched straight-ray tomography configuration: 64 x 64 cells of
    10 m, background slowness 1/3000 s/m plus Gaussian-smoothed
    N(0, (5e-6)^2) perturbations, 2000 straight rays between boundary
    points on different sides, data noise N(0, (1e-4 s)^2).  The smoothing
    width (2 cells) and the prior standard deviation (1e-5 s/m) are not
    fixed by the configuration; these values are this module's choice.
    Returns dict(G, d, s_true, prior_mean, prior_std, noise_std)."""
    rng = np.random.default_rng(seed)
    pert = gaussian_smooth(rng.normal(0.0, perturbation_std, size=(nz, nx)), smoothing_cells)
    s_true = (background + pert).reshape(-1)
    S, R = random_rays(rng, n_rays, nx, nz, cell, cell)
    G = ray_matrix(S, R, nx, nz, cell, cell)
    d = G @ s_true + rng.normal(0.0, noise_std, size=n_rays)
    return dict(G=G, d=d.reshape(-1, 1), s_true=s_true, prior_mean=background,
                prior_std=prior_std, noise_std=noise_std, sources=S, receivers=R)

"""Straight-ray tomography matrix, restated by an incremental cell walk.

Test infrastructure only.  The reference lists straight-ray tomography as
physics (README.md:34) but ships no builder; its likelihood is
``_LinearMatrix_sparse_forward_simple_covariance`` (LinearMatrix.py:309-426)
with a CSR ray-length matrix.  This oracle builds that matrix with an
Amanatides-Woo style traversal (step from cell to cell, choosing the nearer
of the next vertical or horizontal grid line) in plain Python loops.  The
product builder (``hmclab_amd.tomography``) uses a different algorithm
(sorted grid-line crossing parameters, vectorised), so agreement between the
two is a real check; closed forms (axis-aligned and diagonal rays, row sums
equal to ray length) pin both.

Cell ``j = iz * nx + ix`` covers ``[ix*dx, (ix+1)*dx) x [iz*dz, (iz+1)*dz)``.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp


def ray_cells_walk(x0, z0, x1, z1, nx, nz, dx, dz):
    """Return {cell: length} for the segment (x0,z0)->(x1,z1) by cell walking."""
    out = {}
    L = math.hypot(x1 - x0, z1 - z0)
    if L == 0.0:
        return out
    ux, uz = (x1 - x0) / L, (z1 - z0) / L
    xmax, zmax = nx * dx, nz * dz
    # clip the segment to the grid box (slab method)
    t0, t1 = 0.0, L
    for p0, u, hi in ((x0, ux, xmax), (z0, uz, zmax)):
        if u == 0.0:
            if p0 < 0.0 or p0 > hi:
                return out
        else:
            ta, tb = (0.0 - p0) / u, (hi - p0) / u
            if ta > tb:
                ta, tb = tb, ta
            t0, t1 = max(t0, ta), min(t1, tb)
    if t1 <= t0:
        return out
    t = t0
    # starting cell from a point nudged along the ray
    eps = 1e-9 * (t1 - t0)
    px, pz = x0 + ux * (t + eps), z0 + uz * (t + eps)
    ix = min(max(int(math.floor(px / dx)), 0), nx - 1)
    iz = min(max(int(math.floor(pz / dz)), 0), nz - 1)
    sx = 1 if ux > 0 else (-1 if ux < 0 else 0)
    sz = 1 if uz > 0 else (-1 if uz < 0 else 0)
    while t < t1:
        # parameter of the next vertical / horizontal grid line
        if sx > 0:
            tx = ((ix + 1) * dx - x0) / ux
        elif sx < 0:
            tx = (ix * dx - x0) / ux
        else:
            tx = math.inf
        if sz > 0:
            tz = ((iz + 1) * dz - z0) / uz
        elif sz < 0:
            tz = (iz * dz - z0) / uz
        else:
            tz = math.inf
        tn = min(tx, tz, t1)
        if tn > t:
            j = iz * nx + ix
            out[j] = out.get(j, 0.0) + (tn - t)
        t = tn
        if t >= t1:
            break
        if tx <= tz:
            ix += sx
        if tz <= tx:
            iz += sz
        if not (0 <= ix < nx and 0 <= iz < nz):
            break
    return out


def straight_ray_matrix(sources, receivers, nx, nz, dx, dz):
    """CSR (n_rays, nx*nz) of per-cell path lengths, built ray by ray."""
    rows, cols, vals = [], [], []
    for i, ((xs, zs), (xr, zr)) in enumerate(zip(sources, receivers)):
        cells = ray_cells_walk(xs, zs, xr, zr, nx, nz, dx, dz)
        for j in sorted(cells):
            rows.append(i)
            cols.append(j)
            vals.append(cells[j])
    return sp.csr_matrix((np.array(vals), (np.array(rows), np.array(cols))),
                         shape=(len(sources), nx * nz))


def brute_force_lengths(x0, z0, x1, z1, nx, nz, dx, dz, samples=200000):
    """Midpoint-rule estimate of per-cell lengths (fine sampling of the ray)."""
    L = math.hypot(x1 - x0, z1 - z0)
    t = (np.arange(samples) + 0.5) / samples
    px, pz = x0 + t * (x1 - x0), z0 + t * (z1 - z0)
    inside = (px >= 0) & (px < nx * dx) & (pz >= 0) & (pz < nz * dz)
    ix = np.floor(px[inside] / dx).astype(int)
    iz = np.floor(pz[inside] / dz).astype(int)
    j = iz * nx + ix
    lengths = np.bincount(j, minlength=nx * nz) * (L / samples)
    return lengths

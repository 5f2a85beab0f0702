"""Sets of independent dense linear problems sampled together on one GPU.

hmclab samples one target per ``HMC.sample`` call and runs several chains as
separate processes (``ParallelSampleSMP``, hmclab/Samplers.py:1757-1977).
:class:`ProblemSet` holds P problems of uneven size — each a
``BayesRule([Normal|Laplace prior, LinearMatrix(G float32, d, variance)])``
with optional bounds (reflection corrector, base.py:239-374) and Unit or
Diagonal mass (MassMatrices.py:82-238) — and advances 16 chains per problem
in shared launches: the leapfrog, 3-stage or 4-stage trajectory
(Samplers.py:1524-1726) in float32 on MFMA, the endpoint misfit and
Metropolis test (Samplers.py:1471-1492) in FP64.  Each problem has its own
number of steps; a whole proposal round is one replay of a captured
hipGraph.

States, momenta and gradients use a flat host layout: problem after problem,
16 rows of ``d_p`` values.  :meth:`split` and :meth:`join` convert between
that layout and a list of ``(16, d_p)`` arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, iptr, lptr

CHAINS = 16


class ProblemSet:
    def __init__(self, dimensions, data):
        dims = np.ascontiguousarray(np.asarray(dimensions, dtype=np.int64))
        data = np.ascontiguousarray(np.asarray(data, dtype=np.int64))
        if dims.ndim != 1 or dims.shape != data.shape or dims.size == 0:
            raise ValueError("dimensions and data must be equal-length 1-D sequences")
        self.P = int(dims.size)
        self.d = dims.copy()
        self.n = data.copy()
        self.offsets = np.concatenate([[0], np.cumsum(CHAINS * self.d)])
        self.total = int(self.offsets[-1])
        h = C.c_void_p()
        call("hmc_set_create", self.P, lptr(dims), lptr(data), C.byref(h))
        self._h = h

    @property
    def chains(self) -> int:
        return self.P * CHAINS

    # ---- problem data ----------------------------------------------------------
    def set_problem(self, i: int, G, d_obs, variance, prior: str = "normal", means=None, scale=1.0,
                    mass=None, lower=None, upper=None):
        """Problem i: likelihood N(G m, variance) with G (n, d) float32, and a
        Normal (``scale`` = standard deviation) or Laplace (``scale`` =
        dispersion) prior with the given means (default 0).  ``mass``: the
        diagonal of a Diagonal mass matrix (default Unit); ``lower`` /
        ``upper``: bounds of the target (default none)."""
        G = np.ascontiguousarray(np.asarray(G, dtype=np.float32))
        n, d = int(self.n[i]), int(self.d[i])
        if G.shape != (n, d):
            raise ValueError(f"G of problem {i} must have shape {(n, d)}")
        dd = np.ascontiguousarray(np.asarray(d_obs, dtype=np.float64).reshape(n))
        var = np.asarray(variance, dtype=np.float64)
        scalar = var.size == 1
        var = np.ascontiguousarray(var.reshape(1 if scalar else n))
        call("hmc_set_problem_dense", self._h, int(i), G.ctypes.data_as(C.POINTER(C.c_float)), dptr(dd),
             dptr(var), int(scalar))
        a = np.zeros(d) if means is None else np.asarray(means, dtype=np.float64).reshape(d)
        s = np.broadcast_to(np.asarray(scale, dtype=np.float64), (d,))
        if prior == "normal":
            kind, b = _lib.PRIOR_NORMAL, 1.0 / (s * s)
        elif prior == "laplace":
            kind, b = _lib.PRIOR_LAPLACE, 1.0 / s
        else:
            raise ValueError("prior must be 'normal' or 'laplace'")
        call("hmc_set_problem_prior", self._h, int(i), kind, dptr(np.ascontiguousarray(a)),
             dptr(np.ascontiguousarray(b, dtype=np.float64)))
        if mass is not None:
            m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64).reshape(-1), (d,)))
            call("hmc_set_problem_mass", self._h, int(i), _lib.MASS_DIAGONAL, dptr(m))
        if lower is not None or upper is not None:
            lo = None if lower is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(lower, dtype=np.float64).reshape(-1), (d,)))
            hi = None if upper is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(upper, dtype=np.float64).reshape(-1), (d,)))
            call("hmc_set_problem_bounds", self._h, int(i), dptr(lo), dptr(hi))

    def set_option(self, name: str, value: float):
        call("hmc_set_set_option", self._h, name.encode(), float(value))

    def get_option(self, name: str) -> float:
        v = C.c_double(0.0)
        call("hmc_set_get_option", self._h, name.encode(), C.byref(v))
        return v.value

    def tile_config(self):
        r, w = C.c_int(0), C.c_int(0)
        call("hmc_set_tile_config", self._h, C.byref(r), C.byref(w))
        return dict(resid_rows=r.value, grad_waves=w.value, resid_lds=int(self.get_option("resid_lds")),
                    premult=int(self.get_option("premult")), premult_rows=int(self.get_option("premult_rows")),
                    lds_chunk=int(self.get_option("lds_chunk")))

    # ---- layout helpers ----------------------------------------------------------
    def split(self, flat):
        flat = np.asarray(flat)
        return [flat[self.offsets[p]:self.offsets[p + 1]].reshape(CHAINS, int(self.d[p]))
                for p in range(self.P)]

    def join(self, parts):
        return np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in parts]))

    def _flat(self, X):
        X = np.asarray(X, dtype=np.float64)
        if isinstance(X, list) or X.ndim != 1:
            X = self.join(X)
        if X.size != self.total:
            raise ValueError(f"expected {self.total} values in the flat layout")
        return np.ascontiguousarray(X)

    # ---- state, evaluation -----------------------------------------------------------
    def set_state(self, X):
        call("hmc_set_set_state", self._h, dptr(self._flat(X)))

    def get_state(self):
        X = np.empty(self.total)
        chi = np.empty(self.chains)
        call("hmc_set_get_state", self._h, dptr(X), dptr(chi))
        return X, chi

    def misfit(self, X):
        out = np.empty(self.chains)
        call("hmc_set_misfit", self._h, dptr(self._flat(X)), dptr(out))
        return out

    def gradient(self, X):
        out = np.empty(self.total)
        call("hmc_set_gradient", self._h, dptr(self._flat(X)), dptr(out))
        return out

    def _L(self, L):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.int32), (self.P,)))

    def set_integrator(self, integrator: str, problems=None):
        """"lf", "3s" or "4s" (Samplers.py:1524-1726), for every problem or for
        the given problem indices."""
        if integrator not in _lib.INTEGRATORS:
            raise ValueError("integrator must be 'lf', '3s' or '4s'")
        if problems is None:
            self.set_option("integrator", _lib.INTEGRATORS[integrator])
        else:
            for p in np.atleast_1d(problems):
                call("hmc_set_problem_integrator", self._h, int(p), _lib.INTEGRATORS[integrator])

    def propose(self, P0, eps, u, L, integrator=None):
        if integrator is not None:
            self.set_integrator(integrator)
        eps = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(self.chains))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(self.chains))
        acc = np.empty(self.chains, dtype=np.int32)
        h0, h1, X1 = np.empty(self.chains), np.empty(self.chains), np.empty(self.total)
        call("hmc_set_propose", self._h, dptr(self._flat(P0)), dptr(eps), dptr(u), iptr(self._L(L)),
             iptr(acc), dptr(h0), dptr(h1), dptr(X1))
        return dict(accepted=acc.astype(bool), H0=h0, H1=h1, X1=X1)

    def run(self, proposals: int, L, stepsize, randomize=True, seed=0, first_proposal=0, graph=True,
            integrator=None):
        if integrator is not None:
            self.set_integrator(integrator)
        step = np.ascontiguousarray(np.broadcast_to(np.asarray(stepsize, dtype=np.float64), (self.chains,)))
        acc = np.zeros(self.chains, dtype=np.int64)
        call("hmc_set_run", self._h, int(proposals), iptr(self._L(L)), dptr(step), int(bool(randomize)),
             C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), int(first_proposal), int(bool(graph)), lptr(acc))
        return acc

    def set_trace(self, capacity: int):
        call("hmc_set_set_trace", self._h, int(capacity))

    def trace_count(self) -> int:
        n = C.c_int64(0)
        call("hmc_set_trace_count", self._h, C.byref(n))
        return n.value

    def read_trace(self, first: int, count: int):
        X = np.empty((count, self.total))
        chi = np.empty((count, self.chains))
        call("hmc_set_read_trace", self._h, int(first), int(count), dptr(X), dptr(chi))
        return X, chi

    def enable_timing(self, on: bool = True):
        call("hmc_set_enable_timing", self._h, int(bool(on)))

    def timing(self, kind: int):
        ms, n = C.c_double(0.0), C.c_int64(0)
        call("hmc_set_timing", self._h, int(kind), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hmc_set_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_problems(seed: int = 0, problems: int = 256, d_min: int = 50, d_max: int = 2000,
                    data_ratio: float = 1.5, noise: float = 0.05):
    """The uneven-batch workload: d_p log-uniform in [d_min, d_max], n_p =
    round(1.5 d_p), G float32 N(0, 1/d_p), m_true ~ N(0, I), noise sigma 0.05;
    problems in the first half get a Normal prior (std 1), the rest a Laplace
    prior (dispersion 0.5); leapfrog steps per problem uniform in {5..30}."""
    rng = np.random.default_rng(seed)
    d = np.round(np.exp(rng.uniform(np.log(d_min), np.log(d_max), size=problems))).astype(np.int64)
    n = np.round(data_ratio * d).astype(np.int64)
    L = rng.integers(5, 31, size=problems).astype(np.int32)
    specs = []
    for p in range(problems):
        G = (rng.standard_normal(size=(n[p], d[p]), dtype=np.float32) / np.float32(np.sqrt(d[p]))).astype(np.float32)
        m = rng.standard_normal(size=d[p])
        dd = G.astype(np.float64) @ m + rng.normal(0.0, noise, size=n[p])
        prior, scale = ("normal", 1.0) if p < problems // 2 else ("laplace", 0.5)
        specs.append(dict(G=G, d_obs=dd, variance=noise * noise, prior=prior, scale=scale))
    return d, n, L, specs


def curvature_and_start(spec, iters: int = 30, seed: int = 0):
    """Largest eigenvalue of the likelihood Hessian G^T G / var plus the prior's
    (power iteration), and the ridge estimate (G^T G / var + I / s^2)^-1 G^T d / var
    with s the prior scale (a Gaussian stand-in for a Laplace prior) — a start
    near the posterior bulk, where leapfrog energy errors stay O(1)."""
    G = spec["G"].astype(np.float64)
    var = float(spec["variance"])
    s2 = float(spec["scale"]) ** 2 * (2.0 if spec["prior"] == "laplace" else 1.0)
    v = np.random.default_rng(seed).normal(size=G.shape[1])
    lam = 0.0
    for _ in range(iters):
        w = G.T @ (G @ v) / var
        lam = float(np.linalg.norm(w) / np.linalg.norm(v))
        v = w / np.linalg.norm(w)
    H = G.T @ G / var
    H[np.diag_indices_from(H)] += 1.0 / s2
    x = np.linalg.solve(H, G.T @ spec["d_obs"] / var)
    return lam + 1.0 / s2, x

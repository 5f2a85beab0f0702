"""Thin object wrappers over the C-ABI handles (problem = one target,
chains = a batch of Markov chains resident on the GPU)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import call, dptr, iptr, lptr, vptr


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (round to nearest even); return the bits."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16).astype(np.uint16)
    quiet = np.where((u & np.uint64(0xFFFF)) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(nan, ((u >> np.uint64(16)) | quiet).astype(np.uint16), r)
    return r.reshape(f.shape)


def bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32)


def integrator_schedule(integrator: str, L: int):
    """(first drift, stages (S, 3) of kick, drift, drift2) of an engine integrator,
    in units of the local step (host only)."""
    first = C.c_double(0.0)
    n = C.c_int64(0)
    call("hmc_integrator_schedule", _lib.INTEGRATORS[integrator], int(L), C.byref(first), None, 0, C.byref(n))
    st = np.empty((n.value, 3))
    call("hmc_integrator_schedule", _lib.INTEGRATORS[integrator], int(L), C.byref(first), dptr(st), n.value,
         C.byref(n))
    return first.value, st


def philox4x32_10(counter, key):
    """The device RNG's Philox4x32-10 block, evaluated by the library on the host."""
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in counter])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    call("hmc_philox4x32_10", c, k, o)
    return [int(v) for v in o]


class Problem:
    """One BayesRule target on the device (likelihood, priors, bounds, mass)."""

    def __init__(self, dimensions: int, precision: str = "f64", engine: str = "auto"):
        self.dimensions = int(dimensions)
        self.precision = precision
        h = C.c_void_p()
        call("hmc_problem_create", self.dimensions, _lib.PRECISIONS[precision], C.byref(h))
        self._h = h
        self.n_data = 0
        if engine != "auto":
            call("hmc_problem_set_engine", self._h, _lib.ENGINES[engine])

    def set_option(self, name: str, value: float):
        """Engine tuning knob (see hmc_problem_set_option)."""
        call("hmc_problem_set_option", self._h, name.encode(), float(value))

    # -- likelihood ------------------------------------------------------------
    def set_dense(self, G, d_obs, variance, g_dtype: str | None = None):
        G = np.asarray(G)
        n = G.shape[0]
        if g_dtype == "bf16":
            buf = np.ascontiguousarray(bf16_round(G))
            dt = _lib.BF16
        elif G.dtype == np.float32:
            buf = np.ascontiguousarray(G)
            dt = _lib.F32
        elif G.dtype == np.uint16 and g_dtype == "bf16-bits":
            buf = np.ascontiguousarray(G)
            dt = _lib.BF16
        else:
            buf = np.ascontiguousarray(G, dtype=np.float64)
            dt = _lib.F64
        d_obs = np.ascontiguousarray(np.asarray(d_obs, dtype=np.float64).reshape(-1))
        var, scalar = _variance(variance, n)
        call("hmc_problem_set_dense", self._h, vptr(buf), dt, n, dptr(d_obs), dptr(var), scalar)
        self.n_data = n

    def set_csr(self, G, d_obs, variance):
        G = sp.csr_matrix(G)
        G.sort_indices()
        n = G.shape[0]
        indptr = np.ascontiguousarray(G.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(G.indices, dtype=np.int32)
        if G.dtype == np.float32:
            vals, dt = np.ascontiguousarray(G.data, dtype=np.float32), _lib.F32
        else:
            vals, dt = np.ascontiguousarray(G.data, dtype=np.float64), _lib.F64
        d_obs = np.ascontiguousarray(np.asarray(d_obs, dtype=np.float64).reshape(-1))
        var, scalar = _variance(variance, n)
        call("hmc_problem_set_csr", self._h, iptr(indptr), iptr(indices), vptr(vals), dt, n,
             int(G.nnz), dptr(d_obs), dptr(var), scalar)
        self.n_data = n

    def add_prior(self, kind: int, a, b, constant: float = 0.0):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64).reshape(-1), (self.dimensions,)))
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(b, dtype=np.float64).reshape(-1), (self.dimensions,)))
        call("hmc_problem_add_prior", self._h, kind, dptr(a), dptr(b), float(constant))

    def set_bounds(self, lower, upper):
        lo = None if lower is None else np.ascontiguousarray(np.asarray(lower, dtype=np.float64).reshape(-1))
        hi = None if upper is None else np.ascontiguousarray(np.asarray(upper, dtype=np.float64).reshape(-1))
        call("hmc_problem_set_bounds", self._h, dptr(lo), dptr(hi))

    def set_mass(self, diagonal=None):
        if diagonal is None:
            call("hmc_problem_set_mass", self._h, _lib.MASS_UNIT, None)
        else:
            dg = np.ascontiguousarray(np.asarray(diagonal, dtype=np.float64).reshape(-1))
            call("hmc_problem_set_mass", self._h, _lib.MASS_DIAGONAL, dptr(dg))

    # -- evaluation ------------------------------------------------------------
    def misfit_batch(self, M_rows: np.ndarray) -> np.ndarray:
        M = _lib.as_f64_rows(M_rows, -1, self.dimensions)
        out = np.empty(M.shape[0])
        call("hmc_misfit_batch", self._h, dptr(M), M.shape[0], dptr(out))
        return out

    def gradient_batch(self, M_rows: np.ndarray) -> np.ndarray:
        M = _lib.as_f64_rows(M_rows, -1, self.dimensions)
        out = np.empty_like(M)
        call("hmc_gradient_batch", self._h, dptr(M), M.shape[0], dptr(out))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hmc_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _variance(variance, n):
    if np.ndim(variance) == 0:
        return np.array([float(variance)]), 1
    v = np.ascontiguousarray(np.asarray(variance, dtype=np.float64).reshape(-1))
    if v.size != n:
        raise ValueError("data variance must be a scalar or have one value per datum")
    return v, 0


class Chains:
    """B Markov chains of one problem, state resident on the device."""

    def __init__(self, problem: Problem, chains: int):
        self.problem = problem
        self.B = int(chains)
        self.d = problem.dimensions
        h = C.c_void_p()
        call("hmc_chains_create", problem._h, self.B, C.byref(h))
        self._h = h

    def set_state(self, X_rows):
        X = _lib.as_f64_rows(X_rows, self.B, self.d)
        call("hmc_chains_set_state", self._h, dptr(X))

    def get_state(self):
        X = np.empty((self.B, self.d))
        chi = np.empty(self.B)
        call("hmc_chains_get_state", self._h, dptr(X), dptr(chi))
        return X, chi

    def set_anchor(self, x_ref=None):
        """Anchor of bf16 trajectories (see hmc_chains_set_anchor): x_ref (d,)
        or None for the mean of the chains' current positions."""
        x = None if x_ref is None else np.ascontiguousarray(np.asarray(x_ref, dtype=np.float64).reshape(self.d))
        call("hmc_chains_set_anchor", self._h, dptr(x))

    def trajectory_gradient(self, X_rows):
        """Gradient field the trajectories integrate, at B models (B, d)."""
        X = _lib.as_f64_rows(X_rows, self.B, self.d)
        out = np.empty_like(X)
        call("hmc_chains_trajectory_gradient", self._h, dptr(X), dptr(out))
        return out

    def propose(self, P0_rows, eps, u, amount_of_steps: int, integrator: str = "lf",
                trajectory: bool = False):
        P0 = _lib.as_f64_rows(P0_rows, self.B, self.d)
        eps = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(self.B))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(self.B))
        acc = np.empty(self.B, dtype=np.int32)
        h0 = np.empty(self.B)
        h1 = np.empty(self.B)
        X1 = np.empty((self.B, self.d)) if trajectory else None
        P1 = np.empty((self.B, self.d)) if trajectory else None
        call("hmc_chains_propose", self._h, dptr(P0), dptr(eps), dptr(u), int(amount_of_steps),
             _lib.INTEGRATORS[integrator], iptr(acc), dptr(h0), dptr(h1), dptr(X1), dptr(P1))
        out = dict(accepted=acc.astype(bool), H0=h0, H1=h1)
        if trajectory:
            out.update(X1=X1, P1=P1)
        return out

    def run(self, proposals: int, amount_of_steps: int, stepsize, integrator="lf", randomize=True,
            autotune=False, target=0.65, learning_rate=0.75, seed=0, first_proposal=0, thinning=1):
        """Run proposals on device; stepsize (B,) is updated in place when autotuning."""
        step = np.ascontiguousarray(np.broadcast_to(np.asarray(stepsize, dtype=np.float64), (self.B,))).copy()
        accepted = np.zeros(self.B, dtype=np.int64)
        call("hmc_chains_run", self._h, int(proposals), int(amount_of_steps), _lib.INTEGRATORS[integrator],
             dptr(step), int(bool(randomize)), int(bool(autotune)), float(target), float(learning_rate),
             C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), int(first_proposal), int(thinning), lptr(accepted))
        return step, accepted

    def set_tuning_history(self, capacity: int):
        """Record exp(H0 - H1) and the pre-update step of every autotuned
        device proposal (up to `capacity` proposals); 0 frees."""
        call("hmc_chains_set_tuning_history", self._h, int(capacity))

    def tuning_history(self):
        """(rates, steps), each (proposals recorded, B)."""
        n = C.c_int64(0)
        call("hmc_chains_tuning_history_count", self._h, C.byref(n))
        rates = np.empty((n.value, self.B))
        steps = np.empty((n.value, self.B))
        call("hmc_chains_read_tuning_history", self._h, 0, n.value, dptr(rates), dptr(steps))
        return rates, steps

    def set_trace(self, capacity: int):
        call("hmc_chains_set_trace", self._h, int(capacity), self.d)

    def trace_count(self) -> int:
        n = C.c_int64(0)
        call("hmc_chains_trace_count", self._h, C.byref(n))
        return n.value

    def read_trace(self, first: int, count: int):
        X = np.empty((count, self.B, self.d))
        chi = np.empty((count, self.B))
        call("hmc_chains_read_trace", self._h, int(first), int(count), dptr(X), dptr(chi))
        return X, chi

    def reset_trace(self):
        call("hmc_chains_reset_trace", self._h)

    def trace_device(self):
        xp, cp, dt = C.c_void_p(), _lib._dp(), C.c_int(0)
        call("hmc_chains_trace_device", self._h, C.byref(xp), C.byref(cp), C.byref(dt))
        return xp.value, C.cast(cp, C.c_void_p).value, dt.value

    def enable_timing(self, on: bool = True, every: int = 1):
        """Kernel timing with HIP events; every > 1 records only the launches
        of proposals whose index is a multiple of it (hmc_chains_run)."""
        call("hmc_chains_enable_timing", self._h, int(every) if on else 0)

    def timing(self, kind: int):
        ms, n = C.c_double(0.0), C.c_int64(0)
        call("hmc_chains_timing", self._h, int(kind), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def sync(self):
        call("hmc_chains_sync", self._h)

    def set_temperatures(self, temperatures=None):
        """Chain c samples exp(-chi / T_c); None resets every chain to T = 1."""
        if temperatures is None:
            call("hmc_chains_set_temperatures", self._h, None)
            return
        T = np.ascontiguousarray(np.broadcast_to(np.asarray(temperatures, dtype=np.float64), (self.B,)))
        call("hmc_chains_set_temperatures", self._h, dptr(T))

    def exchange(self, pairs, u, proposal: int = 0):
        """One replica-exchange event: pairs (k, 2) with the second chain deciding, u (k,)."""
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(pairs.shape[0]))
        swapped = np.zeros(pairs.shape[0], dtype=np.int32)
        call("hmc_chains_exchange", self._h, iptr(pairs), pairs.shape[0], dptr(u), int(proposal), iptr(swapped))
        return swapped.astype(bool)

    def set_exchange(self, interval=None, schedule=None):
        """Schedule (events, 2k) applied by run() after proposals t % interval == 0."""
        if schedule is None:
            call("hmc_chains_set_exchange", self._h, 0, None, 0, 0)
            return
        sched = np.ascontiguousarray(np.asarray(schedule, dtype=np.int32))
        if sched.ndim != 2 or sched.shape[1] % 2:
            raise ValueError("schedule must have shape (events, 2 * pairs)")
        call("hmc_chains_set_exchange", self._h, int(interval), iptr(sched), sched.shape[0], sched.shape[1] // 2)

    def exchange_counts(self):
        n = np.zeros(self.B, dtype=np.int64)
        call("hmc_chains_exchange_counts", self._h, lptr(n))
        return n

    def set_surrogate(self, surrogate=None):
        """Drive the trajectory with a learned surrogate (None: exact gradient)."""
        call("hmc_chains_set_surrogate", self._h, None if surrogate is None else surrogate._h)
        self._surrogate = surrogate

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hmc_chains_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

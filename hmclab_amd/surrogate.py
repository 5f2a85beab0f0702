"""Learned surrogate potential for the HMC trajectory (HIP kernels in
``csrc/hmc_surrogate.hip``, C-ABI in ``include/hmc_engine.h``).

The network f(x) = w3 . silu(W2 silu(W1 xn + b1) + b2) + b3 with
xn = (x - x_mean) / x_scale approximates the *likelihood* misfit,
chi_lik(x) ~ f_mean + f_scale f(x).  Attached to a set of chains, its input
gradient replaces the likelihood gradient inside the leapfrog, while the
priors stay exact and every Metropolis test uses the exact FP64 Hamiltonian;
leapfrog on any smooth potential is volume preserving and reversible, so the
chain still targets the exact posterior and only the acceptance rate depends
on the surrogate's quality.  hmclab has no counterpart; the trajectory it
replaces is ``HMC._propagate_leapfrog`` (hmclab/Samplers.py:1524-1584).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr

_fp = C.POINTER(C.c_float)


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


class Surrogate:
    """MLP d -> hidden -> hidden -> 1 with SiLU, trained on the GPU."""

    def __init__(self, dimensions: int, hidden: int = 512, seed: int = 0, mode: str = "bf16"):
        self.d = int(dimensions)
        self.hidden = int(hidden)
        h = C.c_void_p()
        call("hmc_surrogate_create", self.d, self.hidden, C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF),
             C.byref(h))
        self._h = h
        self.set_mode(mode)
        self.f_mean, self.f_scale = 0.0, 1.0
        self.x_mean = np.zeros(self.d)
        self.x_scale = np.ones(self.d)

    # ---- configuration --------------------------------------------------------
    def set_mode(self, mode: str):
        """"bf16": bf16 MFMA GEMMs with FP32 accumulation; "f32": FP32 GEMMs."""
        if mode not in ("bf16", "f32"):
            raise ValueError("mode must be 'bf16' or 'f32'")
        call("hmc_surrogate_set_mode", self._h, 0 if mode == "bf16" else 1)
        self.mode = mode

    def set_normalization(self, x_mean, x_scale, f_mean: float, f_scale: float):
        self.x_mean = np.ascontiguousarray(np.asarray(x_mean, dtype=np.float64).reshape(self.d))
        self.x_scale = np.ascontiguousarray(np.asarray(x_scale, dtype=np.float64).reshape(self.d))
        self.f_mean, self.f_scale = float(f_mean), float(f_scale)
        call("hmc_surrogate_set_normalization", self._h, dptr(self.x_mean), dptr(self.x_scale),
             self.f_mean, self.f_scale)

    def fit_normalization(self, X, chi):
        """Per-coordinate mean/std of the training positions; misfit mean/std."""
        X = np.asarray(X, dtype=np.float64)
        chi = np.asarray(chi, dtype=np.float64)
        xs = X.std(axis=0)
        xs = np.where(xs > 0, xs, 1.0)
        fs = float(chi.std()) if chi.std() > 0 else 1.0
        self.set_normalization(X.mean(axis=0), xs, float(chi.mean()), fs)

    # ---- parameters -------------------------------------------------------------
    @property
    def param_count(self) -> int:
        n = C.c_int64(0)
        call("hmc_surrogate_param_count", self._h, C.byref(n))
        return n.value

    def get_params(self) -> dict:
        flat = np.empty(self.param_count, dtype=np.float32)
        call("hmc_surrogate_get_params", self._h, _fptr(flat))
        return self.unpack(flat)

    def set_params(self, P: dict):
        call("hmc_surrogate_set_params", self._h, _fptr(self.pack(P)))

    def unpack(self, flat: np.ndarray) -> dict:
        d, H = self.d, self.hidden
        sizes = [("W1", (H, d)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("w3", (H,)), ("b3", ())]
        out, k = {}, 0
        for name, shape in sizes:
            n = int(np.prod(shape)) if shape else 1
            out[name] = flat[k:k + n].reshape(shape) if shape else np.float32(flat[k])
            k += n
        return out

    def pack(self, P: dict) -> np.ndarray:
        parts = [np.asarray(P[k], dtype=np.float32).ravel() for k in ("W1", "b1", "W2", "b2", "w3", "b3")]
        return np.ascontiguousarray(np.concatenate(parts))

    # ---- evaluation and training ------------------------------------------------
    def evaluate(self, X_rows):
        """Surrogate misfit (B,) and gradient (B, d) at raw positions."""
        X = np.ascontiguousarray(np.asarray(X_rows, dtype=np.float64).reshape(-1, self.d))
        f = np.empty(X.shape[0])
        g = np.empty_like(X)
        call("hmc_surrogate_eval", self._h, dptr(X), X.shape[0], dptr(f), dptr(g))
        return f, g

    def loss_and_grads(self, Xn, y, T, alpha=1.0, beta=1.0):
        """Batch loss and parameter gradients in normalised units (B % 64 == 0)."""
        Xn = np.ascontiguousarray(Xn, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        loss = C.c_double(0.0)
        flat = np.empty(self.param_count, dtype=np.float32)
        call("hmc_surrogate_loss_grad", self._h, dptr(Xn), dptr(y), dptr(T), Xn.shape[0], float(alpha),
             float(beta), C.byref(loss), _fptr(flat))
        return loss.value, self.unpack(flat)

    def train(self, X, chi, grad, epochs: int = 20, batch: int = 256, learning_rate: float = 1e-3,
              alpha: float = 1.0, beta: float = 1.0, seed: int = 0, normalize: bool = True):
        """Adam on (position, likelihood misfit, likelihood gradient) pairs; returns
        the mean loss of every epoch."""
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, self.d)
        chi = np.ascontiguousarray(chi, dtype=np.float64).reshape(-1)
        grad = np.ascontiguousarray(grad, dtype=np.float64).reshape(-1, self.d)
        if normalize:
            self.fit_normalization(X, chi)
        hist = np.empty(int(epochs))
        call("hmc_surrogate_train", self._h, dptr(X), dptr(chi), dptr(grad), X.shape[0], int(batch),
             int(epochs), float(learning_rate), float(alpha), float(beta),
             C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), dptr(hist))
        return hist

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hmc_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def likelihood_of(distribution):
    """The likelihood term of a target: the part a surrogate models (priors and
    bounds stay exact in the trajectory)."""
    from .Distributions import _Spec

    spec = _Spec()
    distribution._contribute(spec)
    if spec.likelihood is None:
        raise ValueError("the target has no likelihood term for a surrogate to model")
    return spec.likelihood


def training_pairs(distribution, X_rows, precision: str = "f64", batch: int = 4096):
    """Exact likelihood misfits (N,) and gradients (N, d) at positions X (N, d),
    evaluated on the GPU in batches — the (m, misfit, gradient) training pairs."""
    from .Distributions import build_problem

    X = np.ascontiguousarray(np.asarray(X_rows, dtype=np.float64))
    prob = build_problem(likelihood_of(distribution), precision=precision, engine="gemm")
    try:
        chi = np.empty(X.shape[0])
        grad = np.empty_like(X)
        for a in range(0, X.shape[0], batch):
            b = min(a + batch, X.shape[0])
            chi[a:b] = prob.misfit_batch(X[a:b])
            grad[a:b] = prob.gradient_batch(X[a:b])
    finally:
        prob.close()
    return chi, grad

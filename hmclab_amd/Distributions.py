"""Target distributions with the reference's interface, evaluated on the GPU.

Mirrors ``hmclab.Distributions`` for the classes on the HMC hot path:
``_AbstractDistribution`` (hmclab/Distributions/base.py:21-374), ``Normal``
(base.py:427-643), ``Laplace`` (base.py:646-744), ``Uniform`` (base.py:747-800),
``AdditiveDistribution``/``BayesRule`` (base.py:1005-1205) and ``LinearMatrix``
(hmclab/Distributions/LinearMatrix.py:15-135).  Constructors, attribute names,
bounds handling and error messages follow the reference; ``misfit`` and
``gradient`` run on the GPU through libhmc.so, and ``misfit_batch`` /
``gradient_batch`` evaluate many models (columns) in one call.

Storage type of ``LinearMatrix``: by default G, d and an array covariance
are stored in float32, as the reference's backends do (their ``dtype``
defaults to ``numpy.single`` and the wrapper never forwards one,
LinearMatrix.py:30-34,148-155,318-329), so ``LinearMatrix(G, d, var)``
defines the same target as in the reference.  Differences by design:
* ``dtype=numpy.float64`` keeps float64 storage (the reference has no way to
  ask for it: its wrapper reinterprets the array's memory instead,
  LinearMatrix.py:30-31); ``dtype="bfloat16"`` defines the target by
  bf16-rounded G with float64 data.
* The dense residual gradient uses the stored G; the reference's uses the
  caller's uncast G (LinearMatrix.py:206-208), about 1e-7 relative apart in
  float32 storage.  The premultiplied b = G^T C^-1 d is built from the
  stored G as well (the reference mixes the caller's G into it, :176), so
  both forms define the same target.
* ``premultiplication`` selects the algebra of a dense likelihood on the
  device as in LinearMatrix.py:163-207: True evaluates the Metropolis misfit
  and the trajectory gradients through H = G^T C^-1 G and b = G^T C^-1 d,
  False in the residual form 0.5*||(G m - d)/sigma||^2 with gradient
  G^T((G m - d)/sigma^2), None (default) premultiplies when n > d, the
  reference's rule.  Host-side ``misfit()``/``gradient()`` use the residual
  form.
* A full data covariance (n x n) is handled by pre-whitening with its
  Cholesky factor (G' = L^-1 G, d' = L^-1 d, unit variance), which gives the
  same misfit and gradient as LinearMatrix.py:226-305 and 444-481.
* A full-covariance ``Normal`` prior (base.py:531,553-574) becomes extra rows
  of the dense likelihood: with the precision P = R R^T (R lower Cholesky),
  0.5 (m - mu)^T P (m - mu) = 0.5 ||R^T m - R^T mu||^2, rows R^T, data R^T mu,
  unit variance.  A sparse likelihood cannot take these dense rows
  (NotImplementedError).
"""
from __future__ import annotations

from typing import List, Union

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

from . import _lib
from .engine import Problem


class _AbstractDistribution:
    """Interface of a target (base.py:21-374): ``dimensions``, ``misfit``,
    ``gradient``, bounds, ``corrector``."""

    name: str = None
    dimensions: int = None
    lower_bounds = None
    upper_bounds = None
    normalized: bool = False
    normalization_constant: float = 0.0

    # ---- GPU evaluation ----------------------------------------------------
    def _gpu(self, precision: str = "f64") -> Problem:
        cache = self.__dict__.setdefault("_gpu_cache", {})
        key = (precision, _bounds_key(self))
        if key not in cache:
            cache.clear()
            cache[key] = build_problem(self, precision=precision)
        return cache[key]

    def _invalidate(self):
        self.__dict__.pop("_gpu_cache", None)

    # Tempering: the distribution is exp(-chi(m) / temperature), as the
    # reference's tempered targets divide misfit and gradient by their
    # temperature (e.g. StandardNormal1D, base.py:392-414).  1 = untempered.
    temperature: float = 1.0

    def misfit(self, coordinates: np.ndarray) -> float:
        """chi(m) / temperature for a (dimensions, 1) column vector, on the GPU."""
        m = np.asarray(coordinates, dtype=np.float64)
        return float(self._gpu().misfit_batch(m.reshape(1, -1))[0]) / self.temperature

    def gradient(self, coordinates: np.ndarray) -> np.ndarray:
        """d chi / dm / temperature as a (dimensions, 1) column vector, on the GPU."""
        m = np.asarray(coordinates, dtype=np.float64)
        return self._gpu().gradient_batch(m.reshape(1, -1)).reshape(-1, 1) / self.temperature

    def misfit_batch(self, M: np.ndarray, precision: str = "f64") -> np.ndarray:
        """Misfits of the columns of M (dimensions, B) -> (B,)."""
        M = np.asarray(M, dtype=np.float64).reshape(self.dimensions, -1)
        return self._gpu(precision).misfit_batch(M.T) / self.temperature

    def gradient_batch(self, M: np.ndarray, precision: str = "f64") -> np.ndarray:
        """Gradients of the columns of M (dimensions, B) -> (dimensions, B)."""
        M = np.asarray(M, dtype=np.float64).reshape(self.dimensions, -1)
        return self._gpu(precision).gradient_batch(M.T).T / self.temperature

    # ---- bounds (host bookkeeping, same rules and messages as base.py) ------
    def corrector(self, coordinates: np.ndarray, momentum: np.ndarray):
        """Reflect at the bounds in place (base.py:239-270)."""
        if self.lower_bounds is not None:
            low = coordinates < self.lower_bounds
            coordinates[low] += 2 * (self.lower_bounds[low] - coordinates[low])
            momentum[low] *= -1.0
        if self.upper_bounds is not None:
            high = coordinates > self.upper_bounds
            coordinates[high] += 2 * (self.upper_bounds[high] - coordinates[high])
            momentum[high] *= -1.0

    def update_bounds(self, lower=None, upper=None):
        """base.py:272-359."""
        old = (self.lower_bounds, self.upper_bounds)
        if type(upper) == list:
            upper = np.array(upper)[:, None]
        if type(lower) == list:
            lower = np.array(lower)[:, None]
        self.upper_bounds, self.lower_bounds = upper, lower
        if lower is not None and type(lower) is not np.ndarray:
            self.lower_bounds, self.upper_bounds = old
            raise ValueError("Lower bounds object not understood.")
        if upper is not None and type(upper) is not np.ndarray:
            self.lower_bounds, self.upper_bounds = old
            raise ValueError("Upper bounds object not understood.")
        if (self.lower_bounds is not None and self.lower_bounds.shape != (self.dimensions, 1)) or (
            self.upper_bounds is not None and self.upper_bounds.shape != (self.dimensions, 1)
        ):
            self.lower_bounds, self.upper_bounds = old
            raise ValueError("Bounds vectors are of incorrect size.")
        if (
            self.lower_bounds is not None
            and self.upper_bounds is not None
            and np.any(self.upper_bounds <= self.lower_bounds)
        ):
            self.lower_bounds, self.upper_bounds = old
            raise ValueError("Bounds vectors are incompatible.")
        self._invalidate()

    def misfit_bounds(self, coordinates: np.ndarray) -> float:
        """base.py:361-374."""
        if (self.lower_bounds is not None and np.any(coordinates < self.lower_bounds)) or (
            self.upper_bounds is not None and np.any(coordinates > self.upper_bounds)
        ):
            return np.inf
        return 0.0

    def normalize(self):
        raise AttributeError("This distribution is not normalizable.")

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        raise NotImplementedError()

    # ---- contribution to a device problem (overridden) ----------------------
    def _contribute(self, spec: "_Spec"):
        raise TypeError(
            f"{type(self).__name__} has no GPU implementation; the engine supports "
            "Normal, Laplace, Uniform, LinearMatrix and BayesRule of those."
        )


def _bounds_key(dist):
    lo, hi = dist.lower_bounds, dist.upper_bounds
    return (None if lo is None else lo.tobytes(), None if hi is None else hi.tobytes())


class Normal(_AbstractDistribution):
    """Gaussian prior (base.py:427-643); diagonal or scalar covariance on the GPU."""

    def __init__(self, means, covariance, inverse_covariance=None, lower_bounds=None,
                 upper_bounds=None):
        self.name = "Gaussian (normal) distribution"
        if type(means) == list:
            means = np.array(means)[:, None]
        if type(covariance) == list:
            covariance = np.array(covariance)[:, None]
        if type(means) == float or type(means) == int:
            self.dimensions = 1
        else:
            self.dimensions = means.size
        self.diagonal = False
        self.normalization_constant = 0.0
        self.generate_ready = False
        if type(means) == float or type(means) == int:
            means = np.ones((self.dimensions, 1)) * means
        else:
            means = np.array(means, dtype=np.float64).reshape(self.dimensions, 1)
        self.means = means
        if type(covariance) in (float, int, np.float64, np.float32):
            covariance = np.float64(covariance)
            self.diagonal = True
        elif covariance.shape == (means.size, means.size):
            self.diagonal = False
        else:
            self.diagonal = True
            covariance = np.array(covariance, dtype=np.float64).reshape(self.dimensions, 1)
        self.covariance = covariance
        if inverse_covariance is not None:
            self.inverse_covariance = inverse_covariance
        elif self.diagonal:
            self.inverse_covariance = 1.0 / self.covariance
        else:
            self.inverse_covariance = np.linalg.inv(self.covariance)
        if self.dimensions == 1:
            self.diagonal = True
        self.update_bounds(lower_bounds, upper_bounds)

    def normalize(self):
        """base.py:576-594."""
        if np.ndim(self.covariance) == 0:
            det = self.covariance ** self.dimensions
        elif self.covariance.shape == (self.means.size, self.means.size):
            det = np.linalg.det(self.covariance)
        elif self.covariance.shape == (self.means.size, 1):
            det = np.prod(self.covariance)
        else:
            raise ValueError("Covariance matrix shape not understood.")
        self.normalization_constant = 0.5 * (np.log(np.abs(det)) + self.dimensions * np.log(2 * np.pi))
        self._invalidate()

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        if self.diagonal:
            return rng.normal(size=(self.dimensions, repeat)) * self.covariance ** 0.5 + self.means
        return np.linalg.cholesky(self.covariance) @ rng.normal(size=(self.dimensions, repeat)) + self.means

    def _contribute(self, spec):
        if not self.diagonal:
            inv = np.asarray(self.inverse_covariance, dtype=np.float64)
            if inv.ndim == 2 and inv.shape == (self.dimensions, self.dimensions) and np.count_nonzero(
                inv - np.diag(np.diag(inv))
            ) == 0:
                spec.priors.append((_lib.PRIOR_NORMAL, self.means, np.diag(inv), self.normalization_constant))
                return
            # 0.5 (m - mu)^T P (m - mu) with P = inverse_covariance = R R^T
            # (base.py:553-574): whitened rows R^T, data R^T mu
            R = np.linalg.cholesky(0.5 * (inv + inv.T))
            spec.rows.append((np.ascontiguousarray(R.T), R.T @ self.means[:, 0]))
            if self.normalization_constant != 0.0:
                # the constant rides on a zero-weight diagonal term
                spec.priors.append((_lib.PRIOR_NORMAL, self.means, np.zeros(self.dimensions),
                                    self.normalization_constant))
            return
        inv = np.broadcast_to(np.asarray(self.inverse_covariance, dtype=np.float64).reshape(-1), (self.dimensions,))
        spec.priors.append((_lib.PRIOR_NORMAL, self.means, inv, self.normalization_constant))

    @staticmethod
    def create_default(dimensions: int, diagonal=True) -> "Normal":
        means = np.random.rand(dimensions, 1)
        sd = np.random.rand(dimensions, 1) + 1
        return Normal(means, sd * sd)


class Laplace(_AbstractDistribution):
    """Uncorrelated Laplace prior (base.py:646-744)."""

    def __init__(self, means, dispersions, lower_bounds=None, upper_bounds=None):
        if type(means) == list:
            means = np.array(means)[:, None]
        if type(dispersions) == list:
            dispersions = np.array(dispersions)[:, None]
        self.dimensions = means.size
        self.means = np.array(means, dtype=np.float64).reshape(self.dimensions, 1)
        dispersions = np.array(dispersions, dtype=np.float64)
        dispersions.shape = (self.dimensions, 1)   # base.py:674 (ndarray required)
        self.dispersions = dispersions
        self.inverse_dispersions = 1.0 / dispersions
        self.normalization_constant = 0.0
        self.update_bounds(lower_bounds, upper_bounds)

    def normalize(self):
        """base.py:712-725."""
        self.normalization_constant = np.log(
            1.0 / (2.0 * (np.prod(self.dispersions) ** (1.0 / self.dimensions)))
        )
        self._invalidate()

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        return rng.laplace(loc=self.means, scale=self.dispersions, size=(self.dimensions, repeat))

    def _contribute(self, spec):
        spec.priors.append((_lib.PRIOR_LAPLACE, self.means, self.inverse_dispersions, self.normalization_constant))

    @staticmethod
    def create_default(dimensions: int) -> "Laplace":
        return Laplace(np.random.rand(dimensions, 1), 10 ** np.random.rand(dimensions, 1))


class Uniform(_AbstractDistribution):
    """Bounded flat prior (base.py:747-800)."""

    def __init__(self, lower_bounds, upper_bounds):
        self.name = "uniform distribution"
        lower_bounds = np.asarray(lower_bounds)
        lower_bounds = np.resize(lower_bounds, (lower_bounds.size, 1))
        upper_bounds = np.asarray(upper_bounds)
        upper_bounds = np.resize(upper_bounds, (upper_bounds.size, 1))
        self.dimensions = lower_bounds.size
        self.update_bounds(lower_bounds, upper_bounds)

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        return rng.uniform(self.lower_bounds, self.upper_bounds, (self.dimensions, repeat))

    def _contribute(self, spec):
        pass   # bounds only; collapsed into the target's bounds

    @staticmethod
    def create_default(dimensions: int) -> "Uniform":
        return Uniform(np.random.rand(dimensions, 1) * 5 - 10, np.random.rand(dimensions, 1) * 5 + 10)


class AdditiveDistribution(_AbstractDistribution):
    """Sum of component misfits with collapsed bounds (base.py:1005-1198)."""

    def __init__(self, list_of_distributions: List[_AbstractDistribution], lower_bounds=None,
                 upper_bounds=None):
        self.name = "additive distribution"
        self.dimensions = list_of_distributions[0].dimensions
        self.separate_distributions = list_of_distributions
        for dist in self.separate_distributions:
            assert dist.dimensions == self.dimensions
        self.lower_bounds = lower_bounds
        self.upper_bounds = upper_bounds
        self.collapse_bounds()

    def collapse_bounds(self):
        """base.py:1061-1101: maximum of lower bounds, minimum of upper bounds."""
        for dist in self.separate_distributions:
            assert dist.dimensions == self.dimensions
            if dist.lower_bounds is not None:
                assert dist.lower_bounds.shape == (self.dimensions, 1)
                self.lower_bounds = (
                    dist.lower_bounds if self.lower_bounds is None
                    else np.maximum(self.lower_bounds, dist.lower_bounds)
                )
            if dist.upper_bounds is not None:
                assert dist.upper_bounds.shape == (self.dimensions, 1)
                self.upper_bounds = (
                    dist.upper_bounds if self.upper_bounds is None
                    else np.minimum(self.upper_bounds, dist.upper_bounds)
                )
        self._invalidate()

    def add_distribution(self, distribution: _AbstractDistribution):
        assert distribution.dimensions == self.dimensions
        self.separate_distributions.append(distribution)
        self.collapse_bounds()

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        raise NotImplementedError(
            "Generating samples from this distribution is not implemented or supported."
        )

    def _contribute(self, spec):
        for dist in self.separate_distributions:
            dist._contribute(spec)


class BayesRule(AdditiveDistribution):
    """Unnormalised Bayes' rule (base.py:1202-1205)."""


class LinearMatrix(_AbstractDistribution):
    """Gaussian likelihood of d = G m (LinearMatrix.py:15-135).

    G: dense ndarray (n, d) or scipy sparse matrix; d: (n, 1) column;
    data_covariance: Python float, (n, 1) variances, or (n, n) covariance.
    """

    def __init__(self, G, d, data_covariance, dtype=None, premultiplication=None, use_mkl=False,
                 **kwargs):
        if kwargs:
            raise TypeError(f"Unidentified argument(s): {kwargs}")
        self.premultiplication = premultiplication   # device algebra of a dense G (None: n > d)
        self.use_mkl = use_mkl
        self.g_dtype = None
        if dtype is None:
            # the reference's backends store G, d and an array covariance in
            # float32 (dtype=numpy.single, LinearMatrix.py:148-155,318-329)
            dtype = np.float32
        if str(dtype) in ("bfloat16", "bf16"):
            # the target is defined by bf16-rounded G; data and covariance stay float64
            self.g_dtype = "bf16"
            G = G.astype(np.float32)
            dtype = None
        if dtype is not None:
            # LinearMatrix.py:30-34,152-155: G, d and a vector/matrix covariance
            # take the storage type
            G = G.astype(dtype)
            d = np.asarray(d).astype(dtype)
            if type(data_covariance) not in (float, np.float32, np.float64):
                data_covariance = data_covariance.astype(dtype)
        self.dimensions = G.shape[1]
        if not (type(d) is np.ndarray and d.shape == (d.size, 1)):
            raise ValueError(
                "Didn't understand the data vector object. Should be a NumPy column vector "
                f"(ndarray: [datapoints, 1]). {type(d)}, {d.shape}"
            )
        if type(G) is np.ndarray and G.shape == (d.size, self.dimensions):
            self.dense = True
        elif sp.issparse(G) and G.shape == (d.size, self.dimensions):
            self.dense = False
        else:
            raise ValueError(
                "Didn't understand the forward model matrix object. Should either be a NumPy "
                "matrix (ndarray: [n_data, n_params]) or a SciPy sparse matrix."
            )
        if type(data_covariance) in (float, np.float32, np.float64) or (
            type(data_covariance) == np.ndarray and data_covariance.shape == (d.size, 1)
        ):
            self.covariance_simple = True
        elif (type(data_covariance) is np.ndarray or sp.issparse(data_covariance)) and data_covariance.shape == (
            d.size,
            d.size,
        ):
            self.covariance_simple = False
        else:
            raise ValueError(
                "Didn't understand the data covariance object. Should either be a float, "
                "NumPy column vector (ndarray: [n_data, 1]) or NumPy square matrix "
                "(ndarray: [n_data, n_data])."
            )
        self.G = G
        self.d = d
        self.data_covariance = data_covariance

    def forward(self, coordinates: np.ndarray) -> np.ndarray:
        return self.G @ coordinates

    def generate(self, repeat=1, rng=np.random.default_rng()) -> np.ndarray:
        raise NotImplementedError()

    def _whitened(self):
        """Forward model, data and variance actually sent to the device."""
        if self.covariance_simple:
            var = self.data_covariance
            if isinstance(var, np.ndarray):
                var = var.astype(np.float64).reshape(-1)
            return self.G, self.d, var
        C = self.data_covariance.toarray() if sp.issparse(self.data_covariance) else np.asarray(self.data_covariance)
        L = np.linalg.cholesky(C.astype(np.float64))
        Gd = self.G.toarray() if sp.issparse(self.G) else np.asarray(self.G, dtype=np.float64)
        Gw = sla.solve_triangular(L, Gd.astype(np.float64), lower=True)
        dw = sla.solve_triangular(L, np.asarray(self.d, dtype=np.float64), lower=True)
        return Gw, dw, 1.0

    def _contribute(self, spec):
        if spec.likelihood is not None:
            raise NotImplementedError("the GPU engine supports one LinearMatrix likelihood per target")
        spec.likelihood = self

    @staticmethod
    def create_default(dimensions: int, dtype=np.dtype("float64")) -> "LinearMatrix":
        G = np.eye(dimensions, dtype=dtype)
        d = np.ones(dimensions)[:, None]
        return LinearMatrix(G, d, 1.0, dtype=dtype)


class _Spec:
    def __init__(self):
        self.priors = []
        self.likelihood = None
        self.rows = []        # whitened dense rows of full-covariance priors: (R^T, R^T mu)


def _stacked(lik, rows):
    """One dense whitened operator from the likelihood and the prior rows:
    G = [G_w; R_k^T], d = [d_w; R_k^T mu_k], variances [var_w; 1]."""
    d = rows[0][0].shape[1]
    Gs, ds, vs = [], [], []
    if lik is not None:
        if sp.issparse(lik.G):
            raise NotImplementedError(
                "a full-covariance Normal prior with a sparse LinearMatrix is not supported by the GPU engine")
        Gw, dw, varw = lik._whitened()
        n = Gw.shape[0]
        Gs.append(np.asarray(Gw, dtype=np.float64))
        ds.append(np.asarray(dw, dtype=np.float64).reshape(n, 1))
        vs.append(np.broadcast_to(np.asarray(varw, dtype=np.float64).reshape(-1), (n,)))
    for Rt, b in rows:
        Gs.append(Rt)
        ds.append(b.reshape(d, 1))
        vs.append(np.ones(d))
    return np.vstack(Gs), np.vstack(ds), np.concatenate(vs)


def build_problem(dist: _AbstractDistribution, precision: str = "f64", engine: str = "auto",
                  mass_diagonal=None, options: dict | None = None) -> Problem:
    """Compile a target into a device problem."""
    spec = _Spec()
    dist._contribute(spec)
    prob = Problem(dist.dimensions, precision=precision, engine=engine)
    lik = spec.likelihood
    if lik is not None and lik.premultiplication is not None and not sp.issparse(lik.G):
        # LinearMatrix.py:163-167: an explicit choice of the algebra
        pm = int(bool(lik.premultiplication))
        prob.set_option("endpoint_premult", pm)
        prob.set_option("premult_trajectory", pm)
    for k, v in (options or {}).items():
        prob.set_option(k, v)
    if spec.rows:
        # full-covariance priors: extra whitened rows of one dense operator,
        # float64 (bf16- or float32-exact likelihood rows stay exact)
        G, d, var = _stacked(lik, spec.rows)
        prob.set_dense(G, d, var)
    elif lik is not None:
        G, d, var = lik._whitened()
        if sp.issparse(G):
            prob.set_csr(G, d, var)
        else:
            prob.set_dense(G, d, var, g_dtype=lik.g_dtype)
    for kind, a, b, const in spec.priors:
        prob.add_prior(kind, a, b, const)
    if dist.lower_bounds is not None or dist.upper_bounds is not None:
        prob.set_bounds(dist.lower_bounds, dist.upper_bounds)
    if mass_diagonal is not None:
        prob.set_mass(mass_diagonal)
    return prob

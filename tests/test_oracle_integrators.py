"""Pin the oracle's integrators, Metropolis rule and autotuning.

Leapfrog: on a separable quadratic potential each mode evolves by the
position-Verlet step matrix, whose L-th power has the Chebyshev closed form
[[cos L t, (e/m)(1 - e^2 w/4) sin L t / sin t], [-e lam sin L t / sin t, cos L t]]
with cos t = 1 - e^2 w / 2, w = lam/m.  Multi-stage schemes: time
reversibility, consistency with the exact oscillator (second order), unit
Jacobian.  Metropolis: strict '>' and inf/NaN cases.  Autotune: the
Robbins-Monro sum in closed form.
"""
import numpy as np
import pytest

from oracle import hmc_oracle as O


class _Quad:
    """U(x) = 0.5 * sum lam x^2 - b x, no bounds."""

    def __init__(self, lam, b):
        self.lam = lam.reshape(-1, 1)
        self.b = b.reshape(-1, 1)
        self.dimensions = lam.size

    def gradient(self, x):
        return self.lam * x - self.b

    def misfit(self, x):
        return float(0.5 * np.sum(self.lam * x * x) - np.sum(self.b * x))

    def corrector(self, x, p):
        pass


def _verlet_power(lam, m, eps, L):
    w = lam / m
    c = 1.0 - eps * eps * w / 2.0
    t = np.arccos(c)
    sLt, s1 = np.sin(L * t), np.sin(t)
    return np.cos(L * t), (eps / m) * (1 - eps * eps * w / 4) * sLt / s1, -eps * lam * sLt / s1


@pytest.mark.parametrize("L", [1, 2, 7, 25])
@pytest.mark.parametrize("diag_mass", [False, True])
def test_leapfrog_matches_chebyshev_closed_form(L, diag_mass):
    rng = np.random.default_rng(10)
    d = 6
    lam = rng.uniform(0.5, 4.0, size=d)
    b = rng.normal(size=d)
    mvec = rng.uniform(0.5, 2.0, size=d) if diag_mass else np.ones(d)
    mass = O.Diagonal(mvec) if diag_mass else O.Unit(d)
    eps = 0.17
    x0 = rng.normal(size=(d, 1))
    p0 = rng.normal(size=(d, 1))
    x1, p1 = O.leapfrog(_Quad(lam, b), mass, x0, p0, eps, L)
    xs = b / lam                                      # minimum of U
    a11, a12, a21 = _verlet_power(lam, mvec, eps, L)
    y = x0[:, 0] - xs
    ex = xs + a11 * y + a12 * p0[:, 0]
    ep = a21 * y + a11 * p0[:, 0]
    assert np.allclose(x1[:, 0], ex, rtol=1e-11, atol=1e-12)
    assert np.allclose(p1[:, 0], ep, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("kind", ["lf", "3s", "4s"])
def test_reversibility(kind):
    rng = np.random.default_rng(11)
    d = 5
    q = _Quad(rng.uniform(0.5, 3.0, size=d), rng.normal(size=d))
    m = O.Unit(d)
    x0, p0 = rng.normal(size=(d, 1)), rng.normal(size=(d, 1))
    x1, p1 = O.integrate(q, m, x0, p0, 0.2, 9, kind)
    x2, p2 = O.integrate(q, m, x1, -p1, 0.2, 9, kind)
    assert np.allclose(x2, x0, atol=1e-12) and np.allclose(-p2, p0, atol=1e-12)


@pytest.mark.parametrize("kind", ["lf", "3s", "4s"])
def test_second_order_consistency(kind):
    # harmonic oscillator x'' = -x: exact solution at T = 1
    q = _Quad(np.array([1.0]), np.array([0.0]))
    m = O.Unit(1)
    x0, p0 = np.array([[1.0]]), np.array([[0.3]])
    T = 1.0
    ex = np.cos(T) * 1.0 + np.sin(T) * 0.3
    errs = []
    for L in (20, 40, 80):
        x1, _ = O.integrate(q, m, x0, p0, T / L, L, kind)
        errs.append(abs(x1[0, 0] - ex))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5


@pytest.mark.parametrize("kind", ["lf", "3s", "4s"])
def test_unit_jacobian_nonlinear(kind):
    class _Anh:
        dimensions = 2

        def gradient(self, x):
            return np.array([[x[0, 0] ** 3 + 0.5 * x[1, 0]], [x[1, 0] + 0.5 * x[0, 0]]])

        def corrector(self, x, p):
            pass

    m = O.Unit(2)
    z0 = np.array([0.4, -0.3, 0.2, 0.7])

    def flow(z):
        x1, p1 = O.integrate(_Anh(), m, z[:2].reshape(2, 1), z[2:].reshape(2, 1), 0.1, 5, kind)
        return np.concatenate([x1[:, 0], p1[:, 0]])

    h = 1e-6
    J = np.stack([(flow(z0 + h * e) - flow(z0 - h * e)) / (2 * h) for e in np.eye(4)], axis=1)
    assert abs(np.linalg.det(J) - 1.0) < 1e-6


def test_metropolis_rule():
    assert O.metropolis(1.0, 1.0 - np.log(0.5), 0.5)[1] is False     # a == u rejects (strict)
    assert O.metropolis(1.0, 1.0 - np.log(0.5), 0.4999)[1] is True
    assert O.metropolis(1.0, np.inf, 0.0)[1] is False                # exp(-inf)=0, not > 0
    assert O.metropolis(np.inf, 1.0, 0.99)[1] is True                # overflow accepts
    assert O.metropolis(np.nan, 1.0, 0.0)[1] is False                # NaN rejects


def test_autotune_closed_form():
    eps0, a, kappa, delta = 0.3, 0.9, 0.75, 0.65
    e = eps0
    for t in range(50):
        e = O.autotune_step(e, t, a, delta, kappa)
    expect = eps0 - sum((k + 1) ** (-kappa) * (delta - min(a, 1)) for k in range(50))
    assert np.isclose(e, expect, rtol=1e-14)
    # NaN acceptance counts as 0, the step is floored at the minimal step
    assert O.autotune_step(0.1, 0, np.nan, 0.65, 0.75) == 1e-18
    assert np.isclose(O.autotune_step(0.1, 3, np.inf, 0.65, 0.75), 0.1 + 4 ** -0.75 * 0.35)

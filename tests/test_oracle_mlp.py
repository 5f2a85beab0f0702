"""Pins the surrogate-MLP oracle (oracle/mlp_oracle.py) by what the
mathematics fixes: central finite differences of f in the input, of the
loss in every weight (which exercises the hand-written double
backpropagation), of silu' and silu''; Adam's first step is -lr*sign(g)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _params(rng, D, H, scale=0.5):
    return dict(W1=rng.normal(size=(H, D)) * scale, b1=rng.normal(size=H) * 0.1,
                W2=rng.normal(size=(H, H)) * scale / np.sqrt(H), b2=rng.normal(size=H) * 0.1,
                w3=rng.normal(size=H) * scale, b3=np.float64(0.3))


def test_silu_derivatives_match_finite_differences():
    z = np.linspace(-6, 6, 101)
    h = 1e-5
    assert np.allclose(M.dsilu(z), (M.silu(z + h) - M.silu(z - h)) / (2 * h), atol=1e-9)
    assert np.allclose(M.d2silu(z), (M.dsilu(z + h) - M.dsilu(z - h)) / (2 * h), atol=1e-9)


def test_input_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    D, H, B = 7, 16, 5
    P = _params(rng, D, H)
    X = rng.normal(size=(B, D))
    _, G, _ = M.forward(P, X)
    h = 1e-6
    for j in range(D):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, j] += h
        Xm[:, j] -= h
        fd = (M.forward(P, Xp)[0] - M.forward(P, Xm)[0]) / (2 * h)
        assert np.allclose(G[:, j], fd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.0, 1.0), (0.7, 1.3)])
def test_loss_gradients_match_finite_differences(alpha, beta):
    rng = np.random.default_rng(1)
    D, H, B = 6, 8, 4
    P = _params(rng, D, H)
    X = rng.normal(size=(B, D))
    y = rng.normal(size=B)
    T = rng.normal(size=(B, D))
    _, grads = M.loss_and_grads(P, X, y, T, alpha, beta)
    h = 1e-6
    for k, v in P.items():
        flat = np.atleast_1d(np.asarray(v, dtype=np.float64)).ravel()
        for i in range(flat.size):
            Pp = {kk: np.array(vv, dtype=np.float64, copy=True) for kk, vv in P.items()}
            Pm = {kk: np.array(vv, dtype=np.float64, copy=True) for kk, vv in P.items()}
            Pp[k].reshape(-1)[i] += h
            Pm[k].reshape(-1)[i] -= h
            fd = (M.loss_and_grads(Pp, X, y, T, alpha, beta)[0] - M.loss_and_grads(Pm, X, y, T, alpha, beta)[0]) / (2 * h)
            an = np.atleast_1d(grads[k]).ravel()[i]
            assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (k, i, an, fd)


def test_dropped_second_derivative_is_caught():
    """A plausible slip (dropping the silu'' term of the double backprop)
    fails the finite-difference check above."""
    rng = np.random.default_rng(2)
    D, H, B = 5, 6, 3
    P = _params(rng, D, H)
    X, y, T = rng.normal(size=(B, D)), rng.normal(size=B), rng.normal(size=(B, D))
    saved = M.d2silu
    try:
        M.d2silu = lambda z: np.zeros_like(z)
        _, bad = M.loss_and_grads(P, X, y, T, 0.0, 1.0)
    finally:
        M.d2silu = saved
    _, good = M.loss_and_grads(P, X, y, T, 0.0, 1.0)
    assert np.max(np.abs(bad["W1"] - good["W1"])) > 1e-3


def test_adam_first_step_is_signed_learning_rate():
    rng = np.random.default_rng(3)
    P = dict(a=rng.normal(size=10))
    g = dict(a=rng.normal(size=10))
    m, v = dict(a=np.zeros(10)), dict(a=np.zeros(10))
    P0 = P["a"].copy()
    M.adam_step(P, g, m, v, 1, lr=0.01)
    assert np.allclose(P["a"], P0 - 0.01 * np.sign(g["a"]), atol=1e-9)

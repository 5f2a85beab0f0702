"""Pin the oracle's target terms (priors, BayesRule, LinearMatrix, mass
matrices) by closed forms, an independent library (scipy.stats), the
residual/premultiplied identity and directional finite differences
(the reference's own check, tests/test_distributions.py:172-215)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.stats as st

from oracle import hmc_oracle as O


def _fd_rel(dist, x, delta):
    g = dist.gradient(x)
    m1 = dist.misfit(x)
    m2 = dist.misfit(x + delta * x)
    return (m2 - m1 - (g.T @ x).item() * delta) / (m2 - m1)


def test_normal_diag_closed_form_and_scipy():
    rng = np.random.default_rng(1)
    d = 7
    mu = rng.normal(size=(d, 1))
    var = rng.uniform(0.5, 2.0, size=(d, 1))
    x = rng.normal(size=(d, 1))
    n = O.Normal(mu, var)
    assert np.isclose(n.misfit(x), 0.5 * np.sum((x - mu) ** 2 / var), rtol=1e-14)
    assert np.allclose(n.gradient(x), (x - mu) / var, rtol=1e-14)
    n.normalize()
    ref = -st.multivariate_normal(mu[:, 0], np.diag(var[:, 0])).logpdf(x[:, 0])
    assert np.isclose(n.misfit(x), ref, rtol=1e-12)


def test_normal_scalar_and_full_covariance():
    rng = np.random.default_rng(2)
    d = 5
    mu = rng.normal(size=(d, 1))
    x = rng.normal(size=(d, 1))
    n = O.Normal(mu, 2.5)
    assert np.isclose(n.misfit(x), 0.5 * np.sum((x - mu) ** 2) / 2.5)
    A = rng.normal(size=(d, d))
    C = A @ A.T + d * np.eye(d)
    nf = O.Normal(mu, C)
    nf.normalize()
    ref = -st.multivariate_normal(mu[:, 0], C).logpdf(x[:, 0])
    assert np.isclose(nf.misfit(x), ref, rtol=1e-12)
    assert np.allclose(nf.gradient(x), np.linalg.solve(C, x - mu), rtol=1e-10)


def test_laplace_closed_form_and_sign_at_mean():
    rng = np.random.default_rng(3)
    d = 6
    mu = rng.normal(size=(d, 1))
    b = rng.uniform(0.2, 3.0, size=(d, 1))
    x = rng.normal(size=(d, 1))
    lp = O.Laplace(mu, b)
    # scipy's Laplace density gives the same data term up to the constant
    ref = -np.sum(st.laplace(mu[:, 0], b[:, 0]).logpdf(x[:, 0])) - np.sum(np.log(2 * b))
    assert np.isclose(lp.misfit(x), ref, rtol=1e-12)
    g = lp.gradient(mu.copy())
    assert np.all(g == 0.0)                     # sign(0) = 0 (base.py:709)
    assert np.allclose(lp.gradient(x), np.sign(x - mu) / b)


@pytest.mark.parametrize("delta", [1e-6, -1e-6, 1e-3])
def test_fd_gradients(delta):
    rng = np.random.default_rng(4)
    d, n = 8, 12
    G = rng.normal(size=(n, d))
    dd = rng.normal(size=(n, 1))
    x = rng.normal(size=(d, 1)) + 2.0
    dists = [
        O.Normal(rng.normal(size=(d, 1)), rng.uniform(0.5, 2, size=(d, 1))),
        O.Laplace(rng.normal(size=(d, 1)) - 5.0, rng.uniform(0.5, 2, size=(d, 1))),
        O.LinearMatrixSimple(G, dd, 0.3, dtype=np.float64, premultiplication=False),
        O.LinearMatrixSimple(sp.csr_matrix(G), dd, rng.uniform(0.5, 2, size=(n, 1)),
                             dtype=np.float64, premultiplication=False),
    ]
    dists.append(O.BayesRule(list(dists)))
    for dist in dists:
        assert abs(_fd_rel(dist, x, delta)) < 1e-3


def test_bounds_validation_messages_and_reflection():
    u = O.Uniform(-np.ones(3), np.ones(3))
    with pytest.raises(ValueError, match="Bounds vectors are incompatible."):
        u.update_bounds(np.ones((3, 1)), -np.ones((3, 1)))
    with pytest.raises(ValueError, match="Bounds vectors are of incorrect size."):
        u.update_bounds(np.ones((4, 1)), None)
    x = np.array([[-1.3], [0.2], [1.1]])
    p = np.array([[1.0], [2.0], [3.0]])
    assert u.misfit(x) == np.inf
    u.corrector(x, p)
    assert np.allclose(x, [[-0.7], [0.2], [0.9]])
    assert np.allclose(p, [[-1.0], [2.0], [-3.0]])
    assert u.misfit(x) == 0.0
    assert np.all(u.gradient(x) == 0.0)


def test_bayes_collapse_bounds_max_min():
    a = O.Uniform([-1.0, -2.0], [3.0, 1.0])
    b = O.Uniform([-3.0, 0.0], [2.0, 5.0])
    br = O.BayesRule([O.Normal(np.zeros(2), 1.0), a, b])
    assert np.allclose(br.lower_bounds[:, 0], [-1.0, 0.0])
    assert np.allclose(br.upper_bounds[:, 0], [2.0, 1.0])
    assert br.misfit(np.array([[0.5], [-0.5]])) == np.inf


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("scalar", [False, True])
def test_linear_residual_equals_premultiplied(sparse, scalar):
    rng = np.random.default_rng(5)
    n, d = 40, 15
    G = rng.normal(size=(n, d))
    if sparse:
        G[rng.uniform(size=G.shape) < 0.7] = 0.0
        G = sp.csr_matrix(G)
    dd = rng.normal(size=(n, 1))
    var = 0.7 if scalar else rng.uniform(0.5, 2.0, size=(n, 1))
    r = O.LinearMatrixSimple(G, dd, var, dtype=np.float64, premultiplication=False)
    pm = O.LinearMatrixSimple(G, dd, var, dtype=np.float64, premultiplication=True)
    for _ in range(3):
        x = rng.normal(size=(d, 1))
        assert np.isclose(r.misfit(x), pm.misfit(x), rtol=1e-9)
        assert np.allclose(r.gradient(x), np.asarray(pm.gradient(x)), rtol=1e-9, atol=1e-12)


def test_linear_identity_closed_form_and_float32_storage():
    rng = np.random.default_rng(6)
    d = 9
    dd = rng.normal(size=(d, 1))
    x = rng.normal(size=(d, 1))
    lin = O.LinearMatrixSimple(np.eye(d), dd, 2.0, dtype=np.float64, premultiplication=False)
    assert np.isclose(lin.misfit(x), 0.25 * np.sum((x - dd) ** 2))
    assert np.allclose(lin.gradient(x), (x - dd) / 2.0)
    # float32 storage: arithmetic is float64 on float32-rounded G and d
    G = rng.normal(size=(20, d))
    d2 = rng.normal(size=(20, 1))
    l32 = O.LinearMatrixSimple(G, d2, 0.5, dtype=np.float32, premultiplication=False)
    G32 = G.astype(np.float32).astype(np.float64)
    d32 = d2.astype(np.float32).astype(np.float64)
    assert np.isclose(l32.misfit(x), 0.5 * np.sum((G32 @ x - d32) ** 2) / 0.5, rtol=1e-12)


def test_float32_storage_gradient_keeps_the_callers_dense_g():
    """Dense residual form: the misfit uses the float32 G, the gradient the
    caller's G (LinearMatrix.py:148,206-208); sparse: both use the cast G
    (:323,359)."""
    rng = np.random.default_rng(16)
    G = rng.normal(size=(30, 12))
    dd = rng.normal(size=(30, 1))
    x = rng.normal(size=(12, 1))
    l32 = O.LinearMatrixSimple(G, dd, 0.5, dtype=np.float32, premultiplication=False)
    G32 = G.astype(np.float32).astype(np.float64)
    r = G32 @ x - dd.astype(np.float32).astype(np.float64)
    assert np.allclose(l32.gradient(x), G.T @ r / 0.5, rtol=1e-12)
    s32 = O.LinearMatrixSimple(sp.csr_matrix(G), dd, 0.5, dtype=np.float32, premultiplication=False)
    assert np.allclose(s32.gradient(x), G32.T @ r / 0.5, rtol=1e-6)
    assert not np.allclose(s32.gradient(x), G.T @ r / 0.5, rtol=1e-12, atol=0)


def _full_cov_case(seed=17, n=25, d=8):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, d))
    dd = rng.normal(size=(n, 1))
    A = rng.normal(size=(n, n)) / np.sqrt(n)
    C = A @ A.T + 0.5 * np.eye(n)
    return rng, G, dd, C


@pytest.mark.parametrize("premult", [False, True])
def test_full_covariance_misfit_is_the_gaussian_log_density(premult):
    """chi = -log N(d; G m, C) - 0.5 log|2 pi C| (scipy.stats), the residual and
    premultiplied forms agree, and the gradient matches central differences."""
    from scipy.stats import multivariate_normal

    rng, G, dd, C = _full_cov_case()
    lm = O.LinearMatrixDenseCovariance(G, dd, C, dtype=np.float64, premultiplication=premult)
    const = 0.5 * np.linalg.slogdet(2 * np.pi * C)[1]
    for _ in range(3):
        x = rng.normal(size=(G.shape[1], 1))
        want = -multivariate_normal(mean=(G @ x)[:, 0], cov=C).logpdf(dd[:, 0]) - const
        assert np.isclose(lm.misfit(x), want, rtol=1e-10)
        h = 1e-6
        fd = np.array([(lm.misfit(x + h * e[:, None]) - lm.misfit(x - h * e[:, None])) / (2 * h)
                       for e in np.eye(G.shape[1])])
        assert np.allclose(lm.gradient(x)[:, 0], fd, rtol=1e-6, atol=1e-6)


def test_full_covariance_forms_sparse_whitened_and_diagonal_agree():
    """The sparse-covariance backend, the pre-whitened simple form
    (L^-1 G, L^-1 d, unit variance with C = L L^T) and, for a diagonal C, the
    vector-variance backend all give the dense-covariance misfit and gradient."""
    import scipy.linalg as sla

    rng, G, dd, C = _full_cov_case(18)
    dense = O.LinearMatrixDenseCovariance(G, dd, C, dtype=np.float64, premultiplication=False)
    sparse = O.LinearMatrixSparseCovariance(sp.csr_matrix(G), dd, sp.csr_matrix(C), dtype=np.float64)
    L = np.linalg.cholesky(C)
    white = O.LinearMatrixSimple(sla.solve_triangular(L, G, lower=True), sla.solve_triangular(L, dd, lower=True),
                                 1.0, dtype=np.float64, premultiplication=False)
    v = rng.uniform(0.5, 2.0, size=(G.shape[0], 1))
    diag = O.LinearMatrixDenseCovariance(G, dd, np.diag(v[:, 0]), dtype=np.float64, premultiplication=False)
    simple = O.LinearMatrixSimple(G, dd, v, dtype=np.float64, premultiplication=False)
    for _ in range(3):
        x = rng.normal(size=(G.shape[1], 1))
        for other in (sparse, white):
            assert np.isclose(other.misfit(x), dense.misfit(x), rtol=1e-10)
            assert np.allclose(other.gradient(x), dense.gradient(x), rtol=1e-9, atol=1e-12)
        assert np.isclose(diag.misfit(x), simple.misfit(x), rtol=1e-12)
        assert np.allclose(diag.gradient(x), simple.gradient(x), rtol=1e-12)


def test_mass_matrices_closed_form_and_shape_errors():
    rng = np.random.default_rng(7)
    d = 6
    p = rng.normal(size=(d, 1))
    diag = rng.uniform(0.5, 3.0, size=(d, 1))
    u = O.Unit(d, rng)
    m = O.Diagonal(diag, rng)
    assert np.isclose(u.kinetic_energy(p), 0.5 * np.sum(p ** 2))
    assert np.isclose(m.kinetic_energy(p), 0.5 * np.sum(p ** 2 / diag))
    assert np.allclose(m.kinetic_energy_gradient(p), p / diag)
    with pytest.raises(ValueError):
        u.kinetic_energy(np.ones((d + 1, 1)))
    with pytest.raises(ValueError):
        m.kinetic_energy_gradient(np.ones((d + 1, 1)))
    # momentum draw p ~ N(0, M): sample variance matches the diagonal
    m.rng = np.random.default_rng(8)
    P = np.hstack([m.generate_momentum() for _ in range(20000)])
    assert np.allclose(P.var(axis=1), diag[:, 0], rtol=0.05)

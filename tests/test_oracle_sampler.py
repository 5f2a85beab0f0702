"""Pin the oracle's seeded proposal loop: the draw order read from the
reference source (momentum normal, then step jitter, then accept uniform,
one Generator per chain), thinning arithmetic (Samplers.py:675-678) and
convergence to the analytic linear-Gaussian posterior (notebook formula)."""
import numpy as np
import pytest

from oracle import hmc_oracle as O


def _c1_like(seed=0, n=30, d=6):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, d))
    m_true = rng.normal(size=(d, 1))
    dd = G @ m_true + rng.normal(0, 0.1, size=(n, 1))
    prior = O.Normal(np.zeros((d, 1)), 1.0)
    lik = O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)
    return G, dd, O.BayesRule([prior, lik])


@pytest.mark.parametrize("randomize", [True, False])
def test_draw_order(randomize):
    _, _, post = _c1_like()
    d = post.dimensions
    out, rec, _ = O.sample_chain(post, O.Unit(d), seed=42, stepsize=0.01, amount_of_steps=3,
                                 proposals=20, randomize_stepsize=randomize, record=True)
    rng = np.random.default_rng(42)
    for r in rec:
        p0 = rng.normal(size=(d, 1))
        assert np.array_equal(p0, r["p0"])
        if randomize:
            assert r["local_eps"] == rng.uniform(0.5, 1.5) * 0.01
        else:
            assert r["local_eps"] == 0.01
        assert r["u"] == rng.uniform(0, 1)


def test_teacher_forced_proposal_equals_chain_step():
    _, _, post = _c1_like()
    d = post.dimensions
    out, rec, _ = O.sample_chain(post, O.Unit(d), seed=3, stepsize=0.01, amount_of_steps=5,
                                 proposals=10, record=True)
    x = np.zeros((d, 1))
    chi = post.misfit(x)
    for k, r in enumerate(rec):
        res = O.propose(post, O.Unit(d), x, chi, r["p0"], r["local_eps"], 5, r["u"])
        assert res["accepted"] == r["accepted"]
        if res["accepted"]:
            x, chi = res["x1"], res["chi1"]
        assert np.array_equal(out[:-1, k:k + 1], x)
        assert out[-1, k] == chi


@pytest.mark.parametrize("proposals,thin", [(10, 1), (12, 3), (100, 10)])
def test_thinning_count(proposals, thin):
    _, _, post = _c1_like()
    out = O.sample_chain(post, O.Unit(post.dimensions), seed=1, stepsize=0.01,
                         amount_of_steps=2, proposals=proposals, online_thinning=thin)
    assert out.shape == (post.dimensions + 1, (proposals - 1) // thin + 1)


def test_posterior_moments_match_analytic():
    G, dd, post = _c1_like(seed=5)
    d = post.dimensions
    mu, Sigma = O.analytic_linear_gaussian(G, dd, 0.01, np.zeros(d), 1.0)
    out = O.sample_chain(post, O.Unit(d), seed=9, stepsize=0.02, amount_of_steps=10,
                         proposals=3000, initial_model=mu)
    S = out[:-1, 300:]
    sd = np.sqrt(np.diag(Sigma))
    # Monte-Carlo tolerance: several standard errors with a conservative ESS of N/5
    se = sd / np.sqrt(S.shape[1] / 5.0)
    assert np.all(np.abs(S.mean(axis=1) - mu) < 5 * se)
    assert np.allclose(S.std(axis=1), sd, rtol=0.15)

"""Tempered chains and replica exchange on the GPU against the oracle
(oracle.Tempered, oracle.exchange_event, oracle.sample_parallel).

Tolerances as in test_gpu_parity: FP64 1e-10 relative to the vector's
largest magnitude; exchange decisions identical."""
import numpy as np
import pytest

from hmclab_amd import Distributions as D
from hmclab_amd import MassMatrices as MM
from hmclab_amd.Distributions import build_problem
from hmclab_amd.engine import Chains
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _case(seed, n=120, d=40):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    mu, b = rng.normal(size=(d, 1)) * 0.1, rng.uniform(0.3, 1.0, size=(d, 1))
    post = D.BayesRule([D.Laplace(mu.copy(), b.copy()), D.LinearMatrix(G, dd, 0.01, dtype=np.float64)])
    orc = O.BayesRule([O.Laplace(mu, b), O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    return post, orc, rng


@pytest.mark.parametrize("engine", ["gemm", "fused"])
def test_tempered_teacher_forced_proposal(engine):
    post, orc, rng = _case(1)
    d, B = post.dimensions, 6
    T = np.array([1.0, 2.0, 4.0, 1.5, 8.0, 3.0])
    ch = Chains(build_problem(post, precision="f64", engine=engine), B)
    ch.set_temperatures(T)
    ch.set_state(rng.normal(size=(B, d)) * 0.3)
    X0, chi0 = ch.get_state()
    P0 = rng.normal(size=(B, d))
    e = 0.01 * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    res = ch.propose(P0, e, u, 10, "lf", trajectory=True)
    unit = O.Unit(d, np.random.default_rng(0))
    for c in range(B):
        t = O.Tempered(orc, T[c])
        ref = O.propose(t, unit, X0[c].reshape(d, 1), None, P0[c].reshape(d, 1), e[c], 10, u[c])
        assert _rel(chi0[c], ref["chi0"]) < 1e-10                      # tempered misfit
        assert _rel(res["X1"][c], ref["x1"][:, 0]) < 1e-10
        assert _rel(res["H0"][c], ref["h0"]) < 1e-10
        assert _rel(res["H1"][c], ref["h1"]) < 1e-10
        if abs(np.log(u[c]) - (ref["h0"] - ref["h1"])) > 1e-6:
            assert bool(res["accepted"][c]) == ref["accepted"]


def test_exchange_event_matches_oracle():
    post, orc, rng = _case(2)
    d, B = post.dimensions, 8
    T = np.array([1.0, 1.3, 1.7, 2.2, 2.9, 3.8, 5.0, 6.5])
    ch = Chains(build_problem(post, precision="f64", engine="gemm"), B)
    ch.set_temperatures(T)
    ch.set_state(rng.normal(size=(B, d)) * 0.3)
    X0, chi0 = ch.get_state()
    row = np.random.default_rng(5).choice(B, B, replace=False)
    u = np.random.default_rng(6).uniform(size=B // 2)
    sw = ch.exchange(row.reshape(-1, 2), u)
    X1, chi1 = ch.get_state()
    dists = [O.Tempered(orc, t) for t in T]
    xs = [X0[c].reshape(d, 1).copy() for c in range(B)]
    chis = [dists[c].misfit(xs[c]) for c in range(B)]
    ref = O.exchange_event(dists, xs, chis, row, u)
    assert list(sw) == ref
    for c in range(B):
        assert _rel(X1[c], xs[c][:, 0]) < 1e-15
        assert _rel(chi1[c], chis[c]) < 1e-10
    assert ch.exchange_counts().sum() == 2 * sum(ref)


def test_parallel_sample_smp_matches_oracle():
    from hmclab_amd import Samplers

    post, orc, rng = _case(3, n=80, d=12)
    d, N = post.dimensions, 4
    T = [1.0, 2.0, 4.0, 8.0]
    posteriors = []
    for t in T:
        p = D.BayesRule(list(post.separate_distributions))
        p.temperature = t
        posteriors.append(p)
    samplers = [Samplers.HMC(seed=100 + c) for c in range(N)]
    m0 = [rng.normal(size=(d, 1)) * 0.2 for _ in range(N)]
    par = Samplers.ParallelSampleSMP(seed=7)
    par.sample(samplers, [None] * N, posteriors, overwrite_existing_files=True, proposals=30, exchange=True,
               exchange_interval=2, initial_model=m0,
               kwargs={"stepsize": 0.01, "amount_of_steps": 5, "disable_progressbar": True})
    got = par.sampler.results                                  # (N, d+1, 30)
    sched = O.exchange_schedule(np.random.default_rng(7), N, 30, 2)
    assert np.array_equal(sched, par.exchange_schedule)
    rngs = [np.random.default_rng(100 + c) for c in range(N)]
    ref, dec = O.sample_parallel([O.Tempered(orc, t) for t in T], [O.Unit(d) for _ in range(N)], rngs, sched, 2,
                                 stepsize=0.01, amount_of_steps=5, initial_models=m0, proposals=30)
    assert any(any(x) for x in dec)                            # some exchanges happened
    for c in range(N):
        assert _rel(got[c], ref[c]) < 1e-9


def test_device_parallel_tempering_targets_tempered_posteriors():
    rng = np.random.default_rng(4)
    n, d = 60, 10
    G = rng.normal(size=(n, d))
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.5, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.25, dtype=np.float64)])
    mu, Sigma = O.analytic_linear_gaussian(G, dd, 0.25, np.zeros(d), 1.0)
    levels = np.array([1.0, 2.0, 4.0, 8.0])
    per = 64
    B = per * len(levels)
    T = np.repeat(levels, per)
    ch = Chains(build_problem(post, precision="f64", engine="gemm"), B)
    ch.set_temperatures(T)
    ch.set_state(np.repeat(mu.reshape(1, d), B, axis=0))
    sched_rng = np.random.default_rng(9)
    # exchanges between neighbouring temperature levels
    def row():
        perm = []
        for k in range(per):
            a = int(sched_rng.integers(0, len(levels) - 1))
            perm += [a * per + k, (a + 1) * per + k]
        return np.array(perm, dtype=np.int32)
    schedule = np.stack([row() for _ in range(300)])
    ch.set_exchange(1, schedule)
    lam = np.linalg.eigvalsh(np.linalg.inv(Sigma)).max()
    step = 0.6 / np.sqrt(lam) * np.sqrt(T)
    ch.run(50, 10, step, seed=3)
    ch.set_trace(250)
    ch.run(250, 10, step, seed=3, first_proposal=50)
    assert ch.exchange_counts().sum() > 0
    X, chi = ch.read_trace(0, 250)
    sd = np.sqrt(np.diag(Sigma))
    for i, t in enumerate(levels):
        s = X[:, i * per:(i + 1) * per, :].reshape(-1, d)
        assert np.sqrt(np.mean(((s.mean(axis=0) - mu.ravel()) / (sd * np.sqrt(t))) ** 2)) < 0.15, t
        assert np.all(np.abs(s.std(axis=0) / (sd * np.sqrt(t)) - 1.0) < 0.15), t

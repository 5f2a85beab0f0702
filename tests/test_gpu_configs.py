"""The measured configurations at their stated sizes, through the C-ABI.

* straight-ray tomography (64 x 64 cells of 10 m, 2000 rays from the ray
  builder, CSR G, Normal prior, Diagonal mass = diagonal of the posterior
  precision), 1024 chains resident in float32 and in float64: teacher-forced
  20-step proposals of 8 chains against the oracle (positions 1e-4 float32,
  1e-10 float64; Hamiltonians the FP64 Hamiltonians of the returned end
  points, 1e-10), and a device-RNG run from exact posterior draws whose
  pooled moments match the analytic FP64 posterior within Monte-Carlo error;
* the uneven batch (256 problems, d log-uniform in [50, 2000], n = 1.5 d,
  16 chains each, L in [5, 30]): teacher-forced proposals of the smallest,
  the median and the largest problem against the oracle (1e-4) in the
  residual form and both premultiplied layouts, finite end points for all
  256, and a device-RNG run in which every problem accepts;
* the dense 20000 x 8192 operator with 512 chains: Hamiltonians of one
  25-step bf16 proposal are the FP64 Hamiltonians (torch float64 products of
  the same bf16-exact G) of its start and end points (1e-9);
* the learned-surrogate configuration (same operator, Laplace prior of
  dispersion 0.3, MLP 8192 -> 512 -> 512 -> 1): one surrogate-driven 25-step
  proposal of 4 chains against the oracle leapfrog driven by the oracle MLP
  (1e-4, float32 surrogate arithmetic), Hamiltonians against torch FP64
  (1e-9), and the bf16-GEMM surrogate's displacement within 3e-2 of it.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from hmclab_amd import Distributions as D
from hmclab_amd import tomography
from hmclab_amd.Distributions import build_problem
from hmclab_amd.engine import Chains
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------------- #
# straight-ray tomography
# --------------------------------------------------------------------------- #
@pytest.fixture(scope="module")
def tomo():
    pd = tomography.synthetic_problem(seed=0)
    G, dd = pd["G"], pd["d"]
    n, d = G.shape
    sd2, sm2, m0 = pd["noise_std"] ** 2, pd["prior_std"] ** 2, pd["prior_mean"]
    A = (G.T @ G).toarray() / sd2
    A[np.diag_indices(d)] += 1.0 / sm2
    rhs = G.T @ dd[:, 0] / sd2 + m0 / sm2
    Lc = np.linalg.cholesky(A)
    mu = sla.cho_solve((Lc, True), rhs)
    Linv = sla.solve_triangular(Lc, np.eye(d), lower=True)
    sd = np.sqrt(np.sum(Linv * Linv, axis=0))
    mass = np.diag(A).copy()
    Ms = 1.0 / np.sqrt(mass)
    v = np.random.default_rng(0).normal(size=d)
    lam = 0.0
    for _ in range(50):
        w = Ms * (A @ (Ms * v))
        lam = float(np.linalg.norm(w) / np.linalg.norm(v))
        v = w / np.linalg.norm(w)
    z = np.random.default_rng(1).normal(size=(d, 1024))
    X0 = (mu[:, None] + sla.solve_triangular(Lc.T, z, lower=False)).T.copy()
    del A, Linv
    return dict(G=G, d=dd, sd2=sd2, sm2=sm2, m0=m0, mu=mu, sd=sd, mass=mass, lam=lam, X0=X0)


def _tomo_targets(t, precision):
    dtype = np.float32 if precision == "f32" else np.float64
    n, d = t["G"].shape
    Gp = t["G"].astype(dtype)
    dd = t["d"].astype(dtype) if dtype == np.float32 else t["d"]
    post = D.BayesRule([D.Normal(np.full((d, 1), t["m0"]), t["sm2"]),
                        D.LinearMatrix(Gp, dd, t["sd2"], dtype=dtype)])
    orc = O.BayesRule([O.Normal(np.full((d, 1), t["m0"]), t["sm2"]),
                       O.LinearMatrixSimple(Gp, dd, t["sd2"], dtype=dtype, premultiplication=False)])
    return post, orc


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_tomography_1024_chains_teacher_forced(tomo, precision):
    t = tomo
    n, d = t["G"].shape
    B, L = 1024, 20
    post, orc = _tomo_targets(t, precision)
    prob = build_problem(post, precision=precision, mass_diagonal=t["mass"])
    ch = Chains(prob, B)
    ch.set_state(t["X0"])
    X0, chi0 = ch.get_state()
    rng = np.random.default_rng(3)
    P0 = np.sqrt(t["mass"])[None, :] * rng.normal(size=(B, d))
    if precision == "f32":
        P0 = P0.astype(np.float32).astype(np.float64)
    eps = 0.5 / np.sqrt(t["lam"]) * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    res = ch.propose(P0, eps, u, L, "lf", trajectory=True)
    assert np.all(np.isfinite(res["X1"])) and np.all(np.isfinite(res["H1"]))
    mass = O.Diagonal(t["mass"])
    tol = 1e-4 if precision == "f32" else 1e-10
    near = 0
    for c in (0, 1, 2, 255, 511, 512, 777, 1023):
        x0 = X0[c].reshape(d, 1)
        assert _rel(chi0[c], orc.misfit(x0)) < (1e-6 if precision == "f32" else 1e-10), c
        ref = O.propose(orc, mass, x0, None, P0[c].reshape(d, 1), eps[c], L, u[c])
        assert _rel(res["X1"][c], ref["x1"][:, 0]) < tol, c
        assert _rel(res["P1"][c], ref["p1"][:, 0]) < tol, c
        # the end-point Hamiltonian is the FP64 Hamiltonian of the returned state
        x1, p1 = res["X1"][c].reshape(d, 1), res["P1"][c].reshape(d, 1)
        h1 = orc.misfit(x1) + mass.kinetic_energy(p1)
        assert abs(res["H1"][c] - h1) <= (1e-6 if precision == "f32" else 1e-10) * abs(h1), c
        if abs(np.log(u[c]) - (ref["h0"] - ref["h1"])) < 1e-3:
            near += 1
        else:
            assert bool(res["accepted"][c]) == ref["accepted"], c
    assert near <= 2
    ch.close()
    prob.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_tomography_1024_chains_posterior_moments(tomo, precision):
    """100 device-RNG proposals of 1024 chains started at exact posterior
    draws: pooled mean within 0.03 posterior sd (RMS over the 4096 cells) and
    pooled std within 3 % of the analytic FP64 posterior (Monte-Carlo error
    with about 5e4 effective draws: ~0.005 sd and ~0.3 %)."""
    t = tomo
    n, d = t["G"].shape
    B, L, K = 1024, 20, 100
    post, _ = _tomo_targets(t, precision)
    prob = build_problem(post, precision=precision, mass_diagonal=t["mass"])
    ch = Chains(prob, B)
    ch.set_state(t["X0"])
    ch.set_trace(K)
    _, acc = ch.run(K, L, np.full(B, 0.5 / np.sqrt(t["lam"])), seed=11)
    assert acc.mean() / K > 0.5
    s1 = np.zeros(d)
    s2 = np.zeros(d)
    for k in range(K):
        X, chi = ch.read_trace(k, 1)
        assert np.all(np.isfinite(chi))
        s1 += X[0].sum(axis=0)
        s2 += (X[0] ** 2).sum(axis=0)
    mean = s1 / (K * B)
    std = np.sqrt(s2 / (K * B) - mean ** 2)
    assert np.sqrt(np.mean(((mean - t["mu"]) / t["sd"]) ** 2)) < 0.03
    assert np.sqrt(np.mean((std / t["sd"] - 1.0) ** 2)) < 0.03
    ch.close()
    prob.close()


# --------------------------------------------------------------------------- #
# uneven batch
# --------------------------------------------------------------------------- #
@pytest.fixture(scope="module")
def uneven():
    from hmclab_amd.problem_set import ProblemSet, uneven_problems

    d, n, L, specs = uneven_problems(seed=0, problems=256)
    ps = ProblemSet(d, n)
    for p, spec in enumerate(specs):
        ps.set_problem(p, spec["G"], spec["d_obs"], spec["variance"], prior=spec["prior"], scale=spec["scale"])
    ps.specs = specs
    return ps, d, n, L


def _uneven_oracle(spec):
    d = spec["G"].shape[1]
    prior = (O.Normal(np.zeros((d, 1)), spec["scale"] ** 2) if spec["prior"] == "normal"
             else O.Laplace(np.zeros((d, 1)), np.full((d, 1), spec["scale"])))
    lik = O.LinearMatrixSimple(spec["G"].astype(np.float64), spec["d_obs"].reshape(-1, 1), spec["variance"],
                               dtype=np.float64, premultiplication=False)
    return O.BayesRule([prior, lik])


@pytest.mark.parametrize("form", ["residual", "premult_strips", "premult_rows"])
def test_uneven_256_problems_teacher_forced(uneven, form):
    from hmclab_amd.problem_set import CHAINS

    ps, d, n, L = uneven
    ps.set_option("premult", int(form != "residual"))
    ps.set_option("premult_rows", int(form == "premult_rows"))
    rng = np.random.default_rng(4)
    ps.set_state(ps.join([rng.normal(size=(CHAINS, int(dp))) * 0.2 for dp in d]))
    X0, chi0 = ps.get_state()
    P0 = ps.join([rng.normal(size=(CHAINS, int(dp))).astype(np.float32).astype(np.float64) for dp in d])
    eps = 0.01 * rng.uniform(0.5, 1.5, size=ps.chains)
    u = rng.uniform(size=ps.chains)
    res = ps.propose(P0, eps, u, L)
    assert np.all(np.isfinite(res["X1"])) and np.all(np.isfinite(res["H1"]))
    order = np.argsort(d, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    assert d[picks[-1]] > 1900
    xs, pz, x1 = ps.split(X0), ps.split(P0), ps.split(res["X1"])
    near = 0
    for p in picks:
        o = _uneven_oracle(ps.specs[p])
        dp = int(d[p])
        unit = O.Unit(dp, np.random.default_rng(0))
        for c in (0, 7, 15):
            gc = p * CHAINS + c
            ref = O.propose(o, unit, xs[p][c].reshape(dp, 1), None, pz[p][c].reshape(dp, 1), eps[gc], int(L[p]),
                            u[gc])
            assert _rel(chi0[gc], ref["chi0"]) < 1e-10, (p, c)
            assert _rel(res["H0"][gc], ref["h0"]) < 1e-10, (p, c)
            assert _rel(x1[p][c], ref["x1"][:, 0]) < 1e-4, (p, c, dp)
            assert _rel(res["H1"][gc], ref["h1"]) < 1e-4, (p, c)
            if abs(np.log(u[gc]) - (ref["h0"] - ref["h1"])) < 1e-3:
                near += 1
            else:
                assert bool(res["accepted"][gc]) == ref["accepted"], (p, c)
    assert near <= 2


def test_uneven_256_problems_device_run_accepts_everywhere(uneven):
    """Ridge starts, per-problem steps 0.5 / sqrt(lambda_max), 20 graph-replayed
    rounds: every problem accepts (more than 20 % of its 16 x 20 proposals)
    and every state stays finite."""
    from hmclab_amd.problem_set import CHAINS, curvature_and_start

    ps, d, n, L = uneven
    ps.set_option("premult", -1)
    starts, steps = [], []
    for spec in ps.specs:
        lam, x = curvature_and_start(spec)
        starts.append(np.repeat(x[None, :], CHAINS, axis=0))
        steps.append(np.full(CHAINS, 0.5 / np.sqrt(lam)))
    ps.set_state(ps.join(starts))
    K = 20
    acc = ps.run(K, L, np.concatenate(steps), seed=5, graph=True)
    rate = acc.reshape(-1, CHAINS).sum(axis=1) / (CHAINS * K)
    X, chi = ps.get_state()
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(chi))
    assert rate.min() > 0.2, (int(np.argmin(rate)), float(rate.min()))


# --------------------------------------------------------------------------- #
# dense 20000 x 8192: 512 chains, and the learned-surrogate configuration
# --------------------------------------------------------------------------- #
def _full():
    # the module pytest imported (tests/ is on sys.path), so the cached problem is shared
    import test_gpu_full_size as F

    return F.full_problem_cache()


def _fp64_hamiltonian_terms(torch, Gt, dd, X, prior, sigma2):
    Xt = torch.from_numpy(X).to("cuda")
    r = Gt @ Xt.T - torch.from_numpy(dd).to("cuda")
    chi = 0.5 * (r * r).sum(dim=0) / sigma2
    if prior == "normal":
        chi = chi + 0.5 * (Xt * Xt).sum(dim=1)
    else:
        chi = chi + Xt.abs().sum(dim=1) / 0.3
    return chi.cpu().numpy()


def test_dense_512_chains_hamiltonians():
    from test_gpu_full_size import N_PARAMS, SIGMA

    torch, Gt, dd, post, X4 = _full()
    B = 512
    rng = np.random.default_rng(9)
    X = (X4[0][None, :] + 0.01 * rng.standard_normal(size=(B, N_PARAMS))).astype(np.float32).astype(np.float64)
    prob = build_problem(post, precision="bf16", engine="gemm", options={"gemm_config": 3, "bf16_split_grad": 1})
    ch = Chains(prob, B)
    ch.set_state(X)
    P0 = rng.standard_normal(size=(B, N_PARAMS)).astype(np.float32).astype(np.float64)
    u = rng.uniform(size=B)
    res = ch.propose(P0, np.full(B, 1e-3), u, 25, "lf", trajectory=True)
    h0 = _fp64_hamiltonian_terms(torch, Gt, dd, X, "normal", SIGMA ** 2) + 0.5 * np.sum(P0 ** 2, axis=1)
    h1 = (_fp64_hamiltonian_terms(torch, Gt, dd, res["X1"], "normal", SIGMA ** 2)
          + 0.5 * np.sum(res["P1"] ** 2, axis=1))
    assert _rel(res["H0"], h0) < 1e-9
    assert _rel(res["H1"], h1) < 1e-9
    with np.errstate(over="ignore"):
        a = np.exp(res["H0"] - res["H1"])
    ok = np.abs(np.log(np.maximum(u, 1e-300)) - (res["H0"] - res["H1"])) > 1e-9
    assert np.array_equal(res["accepted"][ok], (a > u)[ok])
    ch.close()
    prob.close()


class _SurrogateTarget:
    """Oracle trajectory target: exact Laplace prior + the oracle MLP's
    likelihood gradient on normalised inputs."""

    def __init__(self, prior, P, xm, xs, fm, fs):
        self.prior, self.P, self.xm, self.xs, self.fm, self.fs = prior, P, xm, xs, fm, fs

    def gradient(self, x):
        from oracle import mlp_oracle as M

        _, g, _ = M.forward(self.P, ((x[:, 0] - self.xm) / self.xs)[None, :])
        return self.prior.gradient(x) + (self.fs * g[0] / self.xs)[:, None]

    def corrector(self, x, p):
        pass


def test_surrogate_configuration_full_size():
    from test_gpu_full_size import N_PARAMS, SIGMA

    from hmclab_amd.surrogate import Surrogate

    torch, Gt, dd, post_n, X4 = _full()
    d, H, B, L = N_PARAMS, 512, 4, 25
    lik = post_n.separate_distributions[1]
    post = D.BayesRule([D.Laplace(np.zeros((d, 1)), np.full((d, 1), 0.3)), lik])
    prior_o = O.Laplace(np.zeros((d, 1)), np.full((d, 1), 0.3))
    rng = np.random.default_rng(10)
    P = dict(W1=rng.normal(size=(H, d)) / np.sqrt(d), b1=rng.normal(size=H) * 0.1,
             W2=rng.normal(size=(H, H)) / np.sqrt(H), b2=rng.normal(size=H) * 0.1,
             w3=rng.normal(size=H) / np.sqrt(H), b3=np.float64(0.2))
    P = {k: np.asarray(v, dtype=np.float32).astype(np.float64) for k, v in P.items()}
    # normalisation: inputs (x - x_ref) / 0.01, likelihood gradients of about 100
    # per coordinate (the trajectory moves each coordinate by a few 1e-2)
    xm, xs, fm, fs = X4[0].copy(), np.full(d, 0.01), 1.0e4, 100.0
    X = X4.copy()
    P0 = rng.standard_normal(size=(B, d)).astype(np.float32).astype(np.float64)
    eps = np.full(B, 1e-3)
    u = rng.uniform(size=B)
    ends = {}
    for mode in ("f32", "bf16"):
        prob = build_problem(post, precision="bf16", engine="gemm")
        ch = Chains(prob, B)
        s = Surrogate(d, H, seed=1, mode=mode)
        s.set_params(P)
        s.set_normalization(xm, xs, fm, fs)
        ch.set_surrogate(s)
        ch.set_state(X)
        res = ch.propose(P0, eps, u, L, "lf", trajectory=True)
        h0 = _fp64_hamiltonian_terms(torch, Gt, dd, X, "laplace", SIGMA ** 2) + 0.5 * np.sum(P0 ** 2, axis=1)
        h1 = (_fp64_hamiltonian_terms(torch, Gt, dd, res["X1"], "laplace", SIGMA ** 2)
              + 0.5 * np.sum(res["P1"] ** 2, axis=1))
        assert _rel(res["H0"], h0) < 1e-9, mode
        assert _rel(res["H1"], h1) < 1e-9, mode
        ends[mode] = res["X1"]
        ch.close()
        s.close()
        prob.close()
    target = _SurrogateTarget(prior_o, P, xm, xs, fm, fs)
    unit = O.Unit(d, np.random.default_rng(0))
    for c in range(B):
        x1, _ = O.leapfrog(target, unit, X[c].reshape(d, 1), P0[c].reshape(d, 1), eps[c], L)
        assert _rel(ends["f32"][c], x1[:, 0]) < 1e-4, c
        # bf16 GEMMs: the displacement from the start within 3e-2
        assert _rel(ends["bf16"][c] - X[c], x1[:, 0] - X[c]) < 3e-2, c

"""GPU parity of the problem-set engine (many uneven dense problems, 16
chains each) against the oracle, problem by problem, through the C-ABI.

Tolerances, relative to the oracle's float64 values and scaled by each
array's largest magnitude: misfits 1e-10 (FP64 endpoint on the float32
state and float32 G, which the oracle holds exactly); gradients, trajectory
end points and end energies 1e-4 (float32 trajectory arithmetic).  Accept
decisions must agree except for near-ties |log u - (H0 - H1)| < 1e-3."""
import numpy as np
import pytest

from hmclab_amd.problem_set import CHAINS, ProblemSet, uneven_problems
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _build(seed, dims, ratios=None, noise=0.05):
    rng = np.random.default_rng(seed)
    dims = np.asarray(dims)
    ratios = np.full(len(dims), 1.5) if ratios is None else np.asarray(ratios)
    n = np.maximum(1, np.round(ratios * dims)).astype(int)
    ps = ProblemSet(dims, n)
    orc, specs = [], []
    for p, (d, nn) in enumerate(zip(dims, n)):
        G = (rng.standard_normal((nn, d)) / np.sqrt(d)).astype(np.float32)
        dd = G.astype(np.float64) @ rng.normal(size=d) + rng.normal(0, noise, size=nn)
        mu = rng.normal(size=d) * 0.1
        if p % 2 == 0:
            ps.set_problem(p, G, dd, noise ** 2, prior="normal", means=mu, scale=1.0)
            prior = O.Normal(mu.reshape(d, 1), 1.0)
            specs.append(dict(G=G, d_obs=dd, variance=noise ** 2, prior="normal", scale=1.0))
        else:
            ps.set_problem(p, G, dd, noise ** 2, prior="laplace", means=mu, scale=0.5)
            prior = O.Laplace(mu.reshape(d, 1), np.full((d, 1), 0.5))
            specs.append(dict(G=G, d_obs=dd, variance=noise ** 2, prior="laplace", scale=0.5))
        lik = O.LinearMatrixSimple(G.astype(np.float64), dd.reshape(nn, 1), noise ** 2, dtype=np.float64,
                                   premultiplication=False)
        orc.append(O.BayesRule([prior, lik]))
    ps.specs = specs
    return ps, orc, rng


def _start_and_steps(ps):
    """Ridge start (16 copies per problem) and per-chain steps 0.5/sqrt(lam_max)."""
    from hmclab_amd.problem_set import curvature_and_start

    starts, steps = [], []
    for spec in ps.specs:
        lam, x = curvature_and_start(spec)
        starts.append(np.repeat(x[None, :], CHAINS, axis=0))
        steps.append(np.full(CHAINS, 0.5 / np.sqrt(lam)))
    return ps.join(starts), np.concatenate(steps)


DIMS = [3, 17, 50, 130, 257, 64]
RATIOS = [1.5, 1.5, 0.4, 1.5, 1.5, 3.0]


def test_misfit_and_gradient_match_oracle():
    ps, orc, rng = _build(1, DIMS, RATIOS)
    X = ps.join([rng.normal(size=(CHAINS, d)) * 0.3 for d in DIMS])
    Xf = X.astype(np.float32).astype(np.float64)
    chi = ps.misfit(X)
    g = ps.split(ps.gradient(X))
    xs = ps.split(Xf)
    for p, o in enumerate(orc):
        for c in range(CHAINS):
            x = xs[p][c].reshape(-1, 1)
            assert _rel(chi[p * CHAINS + c], o.misfit(x)) < 1e-10, (p, c)
            assert _rel(g[p][c], o.gradient(x)[:, 0]) < 1e-4, (p, c)


@pytest.mark.parametrize("form", ["residual", "premult_strips", "premult_rows"])
@pytest.mark.parametrize("Ls", [[1, 1, 1, 1, 1, 1], [3, 12, 1, 7, 5, 9]])
def test_teacher_forced_proposal_matches_oracle(Ls, form):
    """Residual-form and premultiplied trajectories (H = G^T C^-1 G built in
    FP64, stored fp32, read as column strips or as rows; the oracle keeps the
    residual form) against the oracle; odd and even L exercise both state
    buffers of the premultiplied form."""
    ps, orc, rng = _build(2, DIMS, RATIOS)
    premult = int(form != "residual")
    ps.set_option("premult", premult)
    ps.set_option("premult_rows", int(form == "premult_rows"))
    ps.set_state(ps.join([rng.normal(size=(CHAINS, d)) * 0.2 for d in DIMS]))
    X0, chi0 = ps.get_state()
    # momenta live in float32 on the device: give the oracle the same values
    P0 = ps.join([rng.normal(size=(CHAINS, d)).astype(np.float32).astype(np.float64) for d in DIMS])
    eps = 0.01 * rng.uniform(0.5, 1.5, size=ps.chains)
    u = rng.uniform(size=ps.chains)
    res = ps.propose(P0, eps, u, Ls)
    assert ps.get_option("premult") == premult
    xs, pz, x1 = ps.split(X0), ps.split(P0), ps.split(res["X1"])
    near = 0
    for p, o in enumerate(orc):
        d = DIMS[p]
        unit = O.Unit(d, np.random.default_rng(0))
        for c in range(CHAINS):
            gc = p * CHAINS + c
            ref = O.propose(o, unit, xs[p][c].reshape(d, 1), None, pz[p][c].reshape(d, 1), eps[gc], Ls[p], u[gc])
            assert _rel(chi0[gc], ref["chi0"]) < 1e-10
            assert _rel(res["H0"][gc], ref["h0"]) < 1e-10
            assert _rel(x1[p][c], ref["x1"][:, 0]) < 1e-4, (p, c)
            assert _rel(res["H1"][gc], ref["h1"]) < 1e-4, (p, c)
            if abs(np.log(u[gc]) - (ref["h0"] - ref["h1"])) < 1e-3:
                near += 1
            else:
                assert bool(res["accepted"][gc]) == ref["accepted"], (p, c)
    assert near <= 4


def _build_ext(seed, dims, ratios, mass_problems, bound_problems, noise=0.05, bound=0.3):
    """Problems as in _build, with a Diagonal mass (diagonal uniform in
    [0.5, 2]) on mass_problems and bounds [-bound, bound] on bound_problems;
    the oracle holds the same mass matrices and bounds (BayesRule's collapsed
    bounds, whose corrector reflects)."""
    rng = np.random.default_rng(seed)
    n = np.maximum(1, np.round(np.asarray(ratios) * np.asarray(dims))).astype(int)
    ps = ProblemSet(dims, n)
    orc, masses = [], []
    for p, (d, nn) in enumerate(zip(dims, n)):
        G = (rng.standard_normal((nn, d)) / np.sqrt(d)).astype(np.float32)
        dd = G.astype(np.float64) @ rng.normal(size=d) * 0.1 + rng.normal(0, noise, size=nn)
        mu = rng.normal(size=d) * 0.05
        diag = rng.uniform(0.5, 2.0, size=d) if p in mass_problems else None
        lo, hi = (np.full(d, -bound), np.full(d, bound)) if p in bound_problems else (None, None)
        kind, scale = ("normal", 1.0) if p % 2 == 0 else ("laplace", 0.5)
        ps.set_problem(p, G, dd, noise ** 2, prior=kind, means=mu, scale=scale, mass=diag, lower=lo, upper=hi)
        prior = (O.Normal(mu.reshape(d, 1), 1.0) if kind == "normal"
                 else O.Laplace(mu.reshape(d, 1), np.full((d, 1), 0.5)))
        lik = O.LinearMatrixSimple(G.astype(np.float64), dd.reshape(nn, 1), noise ** 2, dtype=np.float64,
                                   premultiplication=False)
        orc.append(O.BayesRule([prior, lik], lower_bounds=None if lo is None else lo.reshape(d, 1),
                               upper_bounds=None if hi is None else hi.reshape(d, 1)))
        masses.append(O.Diagonal(diag) if diag is not None else O.Unit(d))
    return ps, orc, masses, rng


@pytest.mark.parametrize("form", ["residual", "premult_rows"])
@pytest.mark.parametrize("integ", ["lf", "3s", "4s", "mixed"])
def test_integrators_mass_and_bounds_match_oracle(integ, form):
    """The 3-stage and 4-stage schedules (Samplers.py:1586-1726), Diagonal mass
    (MassMatrices.py:159-238: p = sqrt(M) z, K = p^T M^-1 p / 2, drift M^-1 p)
    on problems 1, 3, 4 and bounds [-0.3, 0.3] with the reflection corrector
    (base.py:239-270) on problems 1, 2, 5 against the oracle's proposal with
    the same float32 momenta, local steps and uniforms.  Starts within 0.27 of
    0 and trajectories of length ~0.05 make some coordinates reflect."""
    ps, orc, masses, rng = _build_ext(7, DIMS, RATIOS, mass_problems={1, 3, 4}, bound_problems={1, 2, 5})
    ps.set_option("premult", int(form != "residual"))
    ps.set_option("premult_rows", 1)
    Ls = [3, 7, 1, 5, 2, 4]
    ps.set_state(ps.join([np.clip(rng.normal(size=(CHAINS, d)) * 0.15, -0.27, 0.27) for d in DIMS]))
    X0, chi0 = ps.get_state()
    P0 = ps.join([(rng.normal(size=(CHAINS, d)) * (np.sqrt(m.diagonal[:, 0]) if hasattr(m, "diagonal") else 1.0))
                  .astype(np.float32).astype(np.float64) for d, m in zip(DIMS, masses)])
    eps = 0.01 * rng.uniform(0.5, 1.5, size=ps.chains)
    u = rng.uniform(size=ps.chains)
    if integ == "mixed":        # one integrator per problem (each problem is its own HMC.sample)
        per = ["lf", "3s", "4s", "4s", "lf", "3s"]
        for p, k in enumerate(per):
            ps.set_integrator(k, problems=[p])
        res = ps.propose(P0, eps, u, Ls)
        assert ps.get_option("integrator") == -1
    else:
        per = [integ] * len(DIMS)
        res = ps.propose(P0, eps, u, Ls, integrator=integ)
        assert ps.get_option("integrator") == {"lf": 0, "3s": 1, "4s": 2}[integ]
    xs, pz, x1 = ps.split(X0), ps.split(P0), ps.split(res["X1"])
    near, reflected = 0, [0]

    def counting(corrector):
        def corr(x, p):
            before = x.copy()
            corrector(x, p)
            reflected[0] += int(np.sum(before != x))
        return corr

    for p, o in enumerate(orc):
        d = DIMS[p]
        o.corrector = counting(o.corrector)
        for c in range(CHAINS):
            gc = p * CHAINS + c
            ref = O.propose(o, masses[p], xs[p][c].reshape(d, 1), None, pz[p][c].reshape(d, 1), eps[gc], Ls[p],
                            u[gc], integrator=per[p])
            assert _rel(chi0[gc], ref["chi0"]) < 1e-10, (p, c)
            assert _rel(res["H0"][gc], ref["h0"]) < 1e-10, (p, c)
            assert _rel(x1[p][c], ref["x1"][:, 0]) < 1e-4, (p, c)
            assert _rel(res["H1"][gc], ref["h1"]) < 1e-4, (p, c)
            if abs(np.log(u[gc]) - (ref["h0"] - ref["h1"])) < 1e-3:
                near += 1
            else:
                assert bool(res["accepted"][gc]) == ref["accepted"], (p, c)
    assert near <= 4
    assert reflected[0] > 0          # the bounds were hit


def test_bounds_give_infinite_misfit_and_gradient_outside():
    """Outside its bounds a chain's misfit is inf (base.py:361-374) and so is
    every component of its gradient (BayesRule adds the scalar bounds misfit,
    base.py:1045-1054); chains inside keep the oracle's values."""
    ps, orc, masses, rng = _build_ext(9, [40, 33], [1.5, 1.5], mass_problems=set(), bound_problems={0})
    X = [np.clip(rng.normal(size=(CHAINS, d)) * 0.1, -0.25, 0.25) for d in (40, 33)]
    X[0][3, 7] = 0.31        # chain 3 of problem 0 outside
    X[0][5, 0] = -0.4        # chain 5 too
    flat = ps.join(X)
    chi = ps.misfit(flat)
    g = ps.split(ps.gradient(flat))
    for p in range(2):
        for c in range(CHAINS):
            x = X[p][c].astype(np.float32).astype(np.float64).reshape(-1, 1)
            if p == 0 and c in (3, 5):
                assert np.isinf(chi[c]) and np.all(np.isinf(g[0][c]))
                assert np.isinf(orc[0].misfit(x))
            else:
                assert _rel(chi[p * CHAINS + c], orc[p].misfit(x)) < 1e-10
                assert _rel(g[p][c], orc[p].gradient(x)[:, 0]) < 1e-4
    with pytest.raises(ValueError, match="incompatible"):
        ps2 = ProblemSet([4], [6])
        ps2.set_problem(0, np.ones((6, 4), np.float32), np.zeros(6), 1.0, lower=np.ones(4), upper=np.zeros(4))


@pytest.mark.parametrize("integ,mass", [("lf", False), ("4s", True)])
def test_device_rng_multistage_diagonal_mass_posterior(integ, mass):
    """Device-RNG chains of a Normal-prior problem with the 4-stage integrator
    and a Diagonal mass (momenta sqrt(M) z from the Philox draws) keep the
    analytic posterior (6400 correlated draws: mean error and sd ratio within
    0.15 of the posterior sd)."""
    dims = [20, 12]
    rng = np.random.default_rng(13)
    n = [30, 18]
    ps = ProblemSet(dims, n)
    Gs, ds = [], []
    for p, (d, nn) in enumerate(zip(dims, n)):
        G = (rng.standard_normal((nn, d)) / np.sqrt(d)).astype(np.float32)
        dd = G.astype(np.float64) @ rng.normal(size=d) + rng.normal(0, 0.3, size=nn)
        H = G.astype(np.float64).T @ G.astype(np.float64) / 0.09 + np.eye(d)
        diag = np.diag(H).copy() if mass else None
        ps.set_problem(p, G, dd, 0.09, prior="normal", scale=1.0, mass=diag)
        Gs.append(G.astype(np.float64))
        ds.append(dd)
    G, dd = Gs[0], ds[0]
    H = G.T @ G / 0.09 + np.eye(dims[0])
    Sigma = np.linalg.inv(H)
    mu = Sigma @ (G.T @ dd / 0.09)
    ps.set_state(ps.join([np.zeros((CHAINS, d)) for d in dims]))
    # the diagonal mass rescales the curvature: steps from the preconditioned spectrum
    Mi = 1.0 / np.diag(H) if mass else np.ones(dims[0])
    lam = np.linalg.eigvalsh(np.sqrt(Mi)[:, None] * H * np.sqrt(Mi)[None, :]).max()
    eps = (1.5 if integ == "4s" else 0.5) / np.sqrt(lam)
    step = np.full(ps.chains, eps)
    ps.run(50, [10, 10], step, seed=3, integrator=integ)
    ps.set_trace(400)
    acc = ps.run(400, [10, 10], step, seed=3, first_proposal=50, integrator=integ)
    assert acc[:CHAINS].mean() / 400 > 0.5
    X, _ = ps.read_trace(0, 400)
    flat = np.stack([ps.split(x)[0] for x in X]).reshape(-1, dims[0])
    sd = np.sqrt(np.diag(Sigma))
    err = np.abs(flat.mean(axis=0) - mu) / sd
    assert np.sqrt(np.mean(err ** 2)) < 0.15
    assert np.all(np.abs(flat.std(axis=0) / sd - 1) < 0.15)


@pytest.mark.parametrize("premult", [0, 1])
def test_staggered_stage_windows_are_bitwise_identical(premult):
    """Problems' stage windows staggered over a round's launches (to balance
    the launches' bytes) against every trajectory starting at launch 0: the
    same draws bit for bit, with mixed integrators and trajectory lengths."""
    ps, orc, rng = _build(12, DIMS, RATIOS)
    ps.set_option("premult", premult)
    ps.set_integrator("4s", problems=[2])
    ps.set_integrator("3s", problems=[4])
    X0 = ps.join([rng.normal(size=(CHAINS, d)) * 0.2 for d in DIMS])
    Ls = [3, 12, 1, 7, 5, 9]
    step = np.full(ps.chains, 0.01)
    out = {}
    for stagger in (0, 1):
        ps.set_option("stagger", stagger)
        assert ps.get_option("stagger") == stagger
        ps.set_state(X0)
        ps.set_trace(3)
        acc = ps.run(3, Ls, step, seed=5)
        out[stagger] = ps.read_trace(0, 3) + (acc,)
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("premult", [0, 1])
def test_graph_replays_match_direct_launches_and_tile_shapes(premult):
    """Graph replays against direct launches (bitwise), then other tile
    shapes: rows per tile and waves per strip, and for the premultiplied form
    rows of H against strips of H (the first two use rows of H)."""
    dims = [40, 200, 75, 310]
    out = []
    for graph, rows, waves, hrow in [(True, 64, 4, 1), (False, 64, 4, 1), (True, 128, 8, 0), (True, 256, 4, 1)]:
        ps, _, rng = _build(3, dims)
        ps.set_option("premult", premult)
        ps.set_option("premult_rows", hrow)
        ps.set_option("resid_rows", rows)
        ps.set_option("grad_waves", waves)
        X0, step = _start_and_steps(ps)
        ps.set_state(X0)
        acc = ps.run(5, [6, 11, 3, 8], step, seed=9, graph=graph)
        X1, _ = ps.get_state()                   # eager work between the replays
        acc2 = ps.run(5, [6, 11, 3, 8], step, seed=9, first_proposal=5, graph=graph)
        X2, chi2 = ps.get_state()
        out.append((acc, acc2, X1, X2, chi2, ps.tile_config()))
        ps.close()
    ref = out[0]
    assert ref[0].sum() > 0.3 * 5 * len(ref[0]) and ref[1].sum() > 0.3 * 5 * len(ref[1])
    # graph replays and direct launches of the same shapes are bitwise identical
    assert np.array_equal(out[1][2], ref[2]) and np.array_equal(out[1][3], ref[3])
    assert np.array_equal(out[1][0], ref[0]) and np.array_equal(out[1][1], ref[1])
    # other tile shapes change only the summation order of the gradient
    for o in out[2:]:
        assert _rel(o[3], ref[3]) < 1e-3
        assert np.abs(o[1] - ref[1]).sum() <= 4


@pytest.mark.parametrize("rows", [64, 128, 256])
def test_lds_staged_residual_is_bitwise_identical(rows):
    """Residual tiles staged through LDS (64- or 128-column chunks; 256-row
    tiles use 64) run the same MFMA sequence per accumulator as the
    register-fed tiles: states, misfits and decisions agree bit for bit
    (d = 75 and 310 end in partial chunks, n = 465 in a partial tile)."""
    dims = [40, 200, 75, 310]
    out = []
    for lds, chunk in ((0, 64), (1, 64), (1, 128)):
        ps, _, _ = _build(3, dims)
        ps.set_option("resid_rows", rows)
        ps.set_option("grad_waves", 4)
        ps.set_option("resid_lds", lds)
        ps.set_option("lds_chunk", chunk)
        ps.set_option("premult", 0)
        X0, step = _start_and_steps(ps)
        ps.set_state(X0)
        acc = ps.run(4, [6, 11, 3, 8], step, seed=4)
        X1, chi1 = ps.get_state()
        out.append((acc, X1, chi1))
        ps.close()
    for o in out[1:]:
        assert np.array_equal(out[0][0], o[0])
        assert np.array_equal(out[0][1], o[1]) and np.array_equal(out[0][2], o[2])


@pytest.mark.parametrize("rows", [64, 128, 256])
def test_premultiplied_row_tiles_chunk_widths_are_bitwise_identical(rows):
    """Rows of H through LDS in 64- or 128-column chunks: same MFMA sequence,
    same bits (d = 75 and 310 end in partial chunks and tiles)."""
    dims = [40, 200, 75, 310]
    out = []
    for chunk in (64, 128):
        ps, _, _ = _build(3, dims)
        ps.set_option("premult", 1)
        ps.set_option("premult_rows", 1)
        ps.set_option("resid_rows", rows)
        ps.set_option("lds_chunk", chunk)
        X0, step = _start_and_steps(ps)
        ps.set_state(X0)
        acc = ps.run(4, [6, 11, 3, 8], step, seed=4)
        X1, chi1 = ps.get_state()
        out.append((acc, X1, chi1))
        ps.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("premult", [0, 1])
def test_device_rng_posterior_matches_analytic_for_normal_problems(premult):
    dims = [20, 48]
    ps, orc, rng = _build(4, dims, noise=0.3)
    ps.set_option("premult", premult)
    # start at the analytic posterior means of the Normal-prior problem 0; problem 1 (Laplace) near 0
    ps.set_state(ps.join([np.zeros((CHAINS, d)) for d in dims]))
    o = orc[0]
    lik = o.separate_distributions[1]
    G = lik.G.astype(np.float64)
    prior_mu = o.separate_distributions[0].means[:, 0]
    H = G.T @ G / 0.3 ** 2 + np.eye(dims[0])
    Sigma = np.linalg.inv(H)
    mu = Sigma @ (G.T @ lik.d[:, 0] / 0.3 ** 2 + prior_mu)
    eps = 0.5 / np.sqrt(np.linalg.eigvalsh(H).max())
    ps.set_trace(400)
    step = np.full(ps.chains, eps)
    ps.run(50, [10, 10], step, seed=11)
    ps.set_trace(400)
    acc = ps.run(400, [10, 10], step, seed=11, first_proposal=50)
    assert acc[:CHAINS].mean() / 400 > 0.5
    X, _ = ps.read_trace(0, 400)
    s0 = np.stack([ps.split(x)[0] for x in X])          # (400, 16, d0)
    flat = s0.reshape(-1, dims[0])
    sd = np.sqrt(np.diag(Sigma))
    err = np.abs(flat.mean(axis=0) - mu) / sd
    assert np.sqrt(np.mean(err ** 2)) < 0.15            # 6400 correlated draws
    assert np.all(np.abs(flat.std(axis=0) / sd - 1) < 0.15)


def test_uneven_workload_builder_and_autotune():
    d, n, L, specs = uneven_problems(seed=0, problems=12, d_max=300)
    assert np.all((d >= 50) & (d <= 300)) and np.all(n == np.round(1.5 * d))
    assert np.all((L >= 5) & (L <= 30))
    from hmclab_amd.problem_set import curvature_and_start

    ps = ProblemSet(d, n)
    starts, steps = [], []
    for p, s in enumerate(specs):
        ps.set_problem(p, s["G"], s["d_obs"], s["variance"], prior=s["prior"], scale=s["scale"])
        lam, x = curvature_and_start(s)
        starts.append(np.repeat(x[None, :], CHAINS, axis=0))
        steps.append(np.full(CHAINS, 0.5 / np.sqrt(lam)))
    ps.set_state(ps.join(starts))
    acc = ps.run(6, L, np.concatenate(steps), seed=1)
    cfg = ps.tile_config()
    assert cfg["resid_rows"] in (64, 128, 256) and cfg["grad_waves"] in (4, 8)
    assert cfg["premult"] == 1                     # n = 1.5 d: one pass over H reads fewer bytes
    rate = acc.reshape(12, CHAINS).sum(axis=1) / (6 * CHAINS)
    assert np.all(rate > 0.3), rate

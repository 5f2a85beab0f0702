"""GPU parity against the oracle, through the C-ABI.

Tolerances (relative to the oracle's float64 values, per element, scaled by
the vector's largest magnitude): float64 1e-10, float32 1e-4; bfloat16
trajectories (G given in bf16, hi/lo split operands, FP32 accumulation):
gradients 2e-4, misfits (always FP64) 1e-10.  Accept/reject decisions must be
identical except for near-ties |log u - (H0 - H1)| < 1e-6 (reported, not
failed).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from hmclab_amd import Distributions as D
from hmclab_amd import MassMatrices as MM
from hmclab_amd.Distributions import build_problem
from hmclab_amd.engine import Chains, bf16_round, bf16_to_f32
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-10, "f32": 1e-4}


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _dense_case(seed, n, d, prior="normal", bounds=False, sparse=False, density=0.1, g_dtype=np.float64):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    if sparse:
        G[rng.uniform(size=G.shape) > density] = 0.0
    G = G.astype(g_dtype)
    m_true = rng.normal(size=(d, 1))
    dd = G.astype(np.float64) @ m_true + rng.normal(0, 0.1, size=(n, 1))
    var = rng.uniform(0.005, 0.02, size=(n, 1))
    Gp = sp.csr_matrix(G) if sparse else G
    comps_gpu, comps_or = [], []
    if prior in ("normal", "both"):
        mu, pv = rng.normal(size=(d, 1)) * 0.1, rng.uniform(0.5, 2.0, size=(d, 1))
        comps_gpu.append(D.Normal(mu.copy(), pv.copy()))
        comps_or.append(O.Normal(mu, pv))
    if prior in ("laplace", "both"):
        mu, b = rng.normal(size=(d, 1)) * 0.1, rng.uniform(0.3, 1.0, size=(d, 1))
        comps_gpu.append(D.Laplace(mu.copy(), b.copy()))
        comps_or.append(O.Laplace(mu, b))
    if bounds:
        lo, hi = -3.0 * np.ones(d), 3.0 * np.ones(d)
        comps_gpu.append(D.Uniform(lo, hi))
        comps_or.append(O.Uniform(lo, hi))
    comps_gpu.append(D.LinearMatrix(Gp, dd, var, dtype=g_dtype))
    comps_or.append(O.LinearMatrixSimple(Gp, dd, var, dtype=g_dtype, premultiplication=False))
    return D.BayesRule(comps_gpu), O.BayesRule(comps_or), rng


# --------------------------------------------------------------------------- #
@pytest.mark.parametrize("engine", ["gemm", "fused"])
@pytest.mark.parametrize("n,d", [(100, 50), (300, 200), (130, 257)])
def test_dense_misfit_gradient_f64(engine, n, d):
    post, orc, rng = _dense_case(1, n, d, prior="both", bounds=True)
    prob = build_problem(post, precision="f64", engine=engine)
    X = 0.5 * rng.normal(size=(5, d))  # inside the +-3 bounds
    X[4, 0] = 3.5                      # out of bounds -> inf misfit, inf gradient
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(4):
        x = X[b].reshape(d, 1)
        assert _rel(chi[b], orc.misfit(x)) < TOL["f64"]
        assert _rel(g[b], orc.gradient(x)[:, 0]) < TOL["f64"]
    assert np.isinf(chi[4]) and np.all(np.isinf(g[4]))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_sparse_misfit_gradient(precision):
    dtype = np.float64 if precision == "f64" else np.float32
    post, orc, rng = _dense_case(2, 400, 300, prior="normal", sparse=True, g_dtype=dtype)
    prob = build_problem(post, precision=precision)
    # positions are stored in the trajectory precision; give both sides the same x
    X = rng.normal(size=(7, 300)).astype(dtype).astype(np.float64)
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(7):
        x = X[b].reshape(-1, 1)
        # the endpoint misfit is FP64; in float32 storage the reference also rounds
        # sigma = sqrt(var) to float32 (LinearMatrix.py:326-329), a 1e-7 effect
        assert _rel(chi[b], orc.misfit(x)) < (1e-10 if precision == "f64" else 1e-6)
        assert _rel(g[b], orc.gradient(x)[:, 0]) < TOL[precision]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_sparse_many_parameter_slices(precision):
    """d = 5000: 79 parameter slices, so each wave of the fused sparse kernel
    owns up to 5 (float32, 16 waves: the 8-slice build) or 10 (float64, 8
    waves: the 16-slice build); misfit, gradient and one 4-step proposal with
    host momenta against the oracle."""
    dtype = np.float64 if precision == "f64" else np.float32
    d = 5000
    post, orc, rng = _dense_case(12, 300, d, prior="normal", sparse=True, density=0.02, g_dtype=dtype)
    prob = build_problem(post, precision=precision)
    X = (rng.normal(size=(3, d)) * 0.3).astype(dtype).astype(np.float64)
    chi, g = prob.misfit_batch(X), prob.gradient_batch(X)
    for b in range(3):
        x = X[b].reshape(-1, 1)
        assert _rel(chi[b], orc.misfit(x)) < (1e-10 if precision == "f64" else 1e-6)
        assert _rel(g[b], orc.gradient(x)[:, 0]) < TOL[precision]
    ch = Chains(prob, 3)
    ch.set_state(X)
    p0 = rng.normal(size=(3, d)).astype(dtype).astype(np.float64)
    eps, u = 0.002, 0.5
    res = ch.propose(p0, np.full(3, eps), np.full(3, u), 4, trajectory=True)
    for b in range(3):
        ref = O.propose(orc, O.Unit(d), X[b].reshape(-1, 1), None, p0[b].reshape(-1, 1), eps, 4, u)
        assert _rel(res["X1"][b], ref["x1"][:, 0]) < TOL[precision]
        assert abs(res["H0"][b] - ref["h0"]) <= (1e-10 if precision == "f64" else 1e-6) * abs(ref["h0"])


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_sparse_degenerate_rows_and_residues(precision):
    """Edge cases of the fused sparse layout: empty data rows, parameters no
    ray touches, and rows whose columns all share one residue modulo 16 (the
    entry ordering then cannot avoid repeated LDS bank groups and falls back
    to its conflict path); misfit, gradient and one proposal against the
    oracle."""
    dtype = np.float64 if precision == "f64" else np.float32
    rng = np.random.default_rng(21)
    n, d = 200, 300
    G = np.zeros((n, d))
    for i in range(n):
        if i % 7 == 3:
            continue                                   # empty data row
        if i % 3 == 0:
            cols = rng.choice(np.arange(0, 288, 16), size=12, replace=False)   # one residue mod 16
        else:
            cols = rng.choice(d - 40, size=rng.integers(1, 60), replace=False)  # parameters >= 260 untouched
        G[i, cols] = rng.normal(size=cols.size)
    G = G.astype(dtype)
    dd = G.astype(np.float64) @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    Gs = sp.csr_matrix(G)
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(Gs, dd, 0.01, dtype=dtype)])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(Gs, dd, 0.01, dtype=dtype, premultiplication=False)])
    prob = build_problem(post, precision=precision)
    X = (rng.normal(size=(5, d)) * 0.5).astype(dtype).astype(np.float64)
    chi, g = prob.misfit_batch(X), prob.gradient_batch(X)
    for b in range(5):
        x = X[b].reshape(-1, 1)
        assert _rel(chi[b], orc.misfit(x)) < (1e-10 if precision == "f64" else 1e-6)
        assert _rel(g[b], orc.gradient(x)[:, 0]) < TOL[precision]
    ch = Chains(prob, 5)
    ch.set_state(X)
    p0 = rng.normal(size=(5, d)).astype(dtype).astype(np.float64)
    res = ch.propose(p0, np.full(5, 0.003), np.full(5, 0.5), 5, trajectory=True)
    for b in range(5):
        ref = O.propose(orc, O.Unit(d), X[b].reshape(-1, 1), None, p0[b].reshape(-1, 1), 0.003, 5, 0.5)
        assert _rel(res["X1"][b], ref["x1"][:, 0]) < TOL[precision]


def test_sparse_model_changes_after_launches():
    """The fused sparse engine keeps prior, mass and bounds arrays in its own
    slot order, gathered again after any change: a Laplace prior, bounds and
    a Diagonal mass added after earlier evaluations and proposals must show up
    in the next launch (closed-form prior terms, bounds as infinite misfit,
    and the oracle's Diagonal-mass proposal)."""
    from hmclab_amd import _lib

    post, orc, rng = _dense_case(11, 300, 200, prior="normal", sparse=True)
    d = 200
    prob = build_problem(post, precision="f64")
    X = rng.normal(size=(4, d)) * 0.5
    chi0, g0 = prob.misfit_batch(X), prob.gradient_batch(X)
    ch = Chains(prob, 4)
    ch.set_state(X)
    p0 = rng.normal(size=(4, d))
    ch.propose(p0, np.full(4, 0.01), np.full(4, 0.5), 3)
    # a second prior term: Laplace with dispersions s (device b = 1/s), no constant
    mu2, s2 = rng.normal(size=d) * 0.1, rng.uniform(0.3, 1.0, size=d)
    prob.add_prior(_lib.PRIOR_LAPLACE, mu2, 1.0 / s2, 0.0)
    chi1, g1 = prob.misfit_batch(X), prob.gradient_batch(X)
    for b in range(4):
        assert _rel(chi1[b] - chi0[b], np.sum(np.abs(X[b] - mu2) / s2)) < 1e-10
        assert _rel(g1[b] - g0[b], np.sign(X[b] - mu2) / s2) < 1e-10
    # bounds: unchanged inside, infinite outside
    lo, hi = -3.0 * np.ones(d), 3.0 * np.ones(d)
    prob.set_bounds(lo, hi)
    assert _rel(prob.misfit_batch(X), chi1) < 1e-14
    Xo = X.copy()
    Xo[2, 17] = 3.5
    assert np.isinf(prob.misfit_batch(Xo)[2]) and np.isfinite(prob.misfit_batch(Xo)[1])
    # a Diagonal mass after the chains have proposed: the next proposal uses it
    m = rng.uniform(0.5, 2.0, size=d)
    prob.set_mass(m)
    ch.set_state(X)
    eps, u = 0.004, 0.5
    res = ch.propose(p0, np.full(4, eps), np.full(4, u), 6, trajectory=True)
    orc_full = O.BayesRule([orc.separate_distributions[0], O.Laplace(mu2.reshape(-1, 1), s2.reshape(-1, 1)),
                            O.Uniform(lo, hi), orc.separate_distributions[-1]])
    for b in range(4):
        ref = O.propose(orc_full, O.Diagonal(m), X[b].reshape(-1, 1), None, p0[b].reshape(-1, 1), eps, 6, u)
        assert _rel(res["X1"][b], ref["x1"][:, 0]) < 1e-10
        assert abs(res["H0"][b] - ref["h0"]) <= 1e-10 * abs(ref["h0"])
        assert abs(res["H1"][b] - ref["h1"]) <= 1e-10 * abs(ref["h1"])


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("n,d,B", [(520, 384, 3), (300, 700, 130)])
def test_bf16_gemm_gradient_and_fp64_misfit(cfg, n, d, B):
    rng = np.random.default_rng(3)
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = rng.normal(size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype="bfloat16")])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    prob = build_problem(post, precision="bf16", options={"gemm_config": cfg})
    X = rng.normal(size=(B, d)).astype(np.float32).astype(np.float64)   # float32 state
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(B):
        x = X[b].reshape(d, 1)
        assert _rel(chi[b], orc.misfit(x)) < 1e-10
        assert _rel(g[b], orc.gradient(x)[:, 0]) < 2e-4


@pytest.mark.parametrize("incremental,split_grad,premult,integ", [(0, 2, 0, "lf"), (1, 2, 0, "lf"), (1, 1, 0, "lf"),
                                                                (0, 2, 1, "lf"), (0, 2, 1, "4s")])
def test_bf16_trajectory_close_to_fp64(incremental, split_grad, premult, integ):
    """bf16 trajectories (hi/lo planes of x in every residual GEMM; with
    incremental_residual the anchored residual r = (G x_ref - d) + G_hi
    bf16(x - x_ref); premultiplied: g = (H x_ref - b) + H_hi bf16(x - x_ref))
    stay within 2e-3 of the FP64 oracle trajectory (relative to the largest
    coordinate) over 10 leapfrog steps; the Hamiltonians are exact FP64 of
    those end points."""
    rng = np.random.default_rng(11)
    n, d, B, L = 520, 384, 8, 10
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype="bfloat16")])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    prob = build_problem(post, precision="bf16",
                         options={"incremental_residual": incremental, "bf16_split_grad": split_grad,
                                  "premult_trajectory": premult})
    ch = Chains(prob, B)
    ch.set_state((rng.normal(size=(B, d)) * 0.3).astype(np.float32).astype(np.float64))
    X0, _ = ch.get_state()
    P0 = rng.normal(size=(B, d)).astype(np.float32).astype(np.float64)
    e = 0.002 * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    res = ch.propose(P0, e, u, L, integ, trajectory=True)
    unit = O.Unit(d, np.random.default_rng(0))
    for c in range(B):
        ref = O.propose(orc, unit, X0[c].reshape(d, 1), None, P0[c].reshape(d, 1), e[c], L, u[c], integ)
        assert _rel(res["X1"][c], ref["x1"][:, 0]) < 2e-3, c
        assert _rel(res["H0"][c], ref["h0"]) < 1e-10
        chi1 = orc.misfit(res["X1"][c].reshape(d, 1))
        k1 = 0.5 * float(np.sum(res["P1"][c] ** 2))
        assert abs(res["H1"][c] - (chi1 + k1)) <= 1e-6 * abs(chi1 + k1), c


def _bf16(a):
    return bf16_to_f32(bf16_round(np.asarray(a, dtype=np.float32))).astype(np.float64)


def _anchored_case(seed=12, n=520, d=384):
    rng = np.random.default_rng(seed)
    G = _bf16(rng.normal(size=(n, d)) / np.sqrt(d))
    m_true = rng.normal(size=(d, 1))
    dd = G @ m_true + rng.normal(0, 0.1, size=(n, 1))
    var = 0.01
    pv = 1.0
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), pv), D.LinearMatrix(G, dd, var, dtype="bfloat16")])
    H = G.T @ G / var
    b = (G.T @ dd / var)[:, 0]
    mu = np.linalg.solve(H + np.eye(d) / pv, b)
    return rng, G, dd, var, pv, post, H, b, mu


@pytest.mark.parametrize("premult", [1, 0])
def test_bf16_anchored_gradient_field_matches_its_definition(premult):
    """The gradient field of anchored bf16 trajectories, evaluated through the
    C-ABI at arbitrary positions, equals its definition computed here in
    float64 from the same bf16 roundings:
      premultiplied  g = f32(H x_ref - b) + bf16(H) bf16(x - x_ref) + prior'(x)
      residual form  r = f32(G x_ref - d) + G bf16(x - x_ref),
                     g = G^T (hi + lo planes of r / var) + prior'(x)
    to float32 accumulation (1e-5 of the largest entry).  At x = x_ref the
    offset plane is zero and the field is the FP64 gradient rounded once to
    float32 (1e-6); away from it the field stays within 1e-2 of the FP64
    gradient (bf16 rounding of H, or of G and r / var, and of the offset: each
    2^-9 relative; a float64 emulation of this case gives 2-6e-3).  The anchor
    is set from the host."""
    rng, G, dd, var, pv, post, H, b, mu = _anchored_case()
    n, d = G.shape
    B = 6
    prob = build_problem(post, precision="bf16", engine="gemm",
                         options={"premult_trajectory": premult, "incremental_residual": 1, "bf16_split_grad": 2})
    ch = Chains(prob, B)
    x_ref = (mu + 0.01 * rng.normal(size=d)).astype(np.float32).astype(np.float64)
    ch.set_state(np.tile(x_ref, (B, 1)))
    ch.set_anchor(x_ref)
    X = (x_ref[None, :] + 0.05 * rng.normal(size=(B, d))).astype(np.float32).astype(np.float64)
    X[0] = x_ref
    g = ch.trajectory_gradient(X)
    g64 = X @ H - b[None, :] + X / pv
    for c in range(B):
        off = _bf16(X[c] - x_ref)
        prior = X[c] / pv
        if premult:
            yref = (H @ x_ref - b).astype(np.float32).astype(np.float64)
            want = yref + _bf16(H) @ off + prior
        else:
            rref = (G @ x_ref - dd[:, 0]).astype(np.float32).astype(np.float64)
            w = ((rref + G @ off) / var).astype(np.float32)
            hi = _bf16(w)
            lo = _bf16(w.astype(np.float64) - hi)
            want = G.T @ (hi + lo) + prior
        assert _rel(g[c], want) < 1e-5, c
        # at x_ref: one float32 rounding of the FP64 gradient (premultiplied), or
        # the hi/lo planes of r / var (16 bits, 2^-17 relative; residual form)
        assert _rel(g[c], g64[c]) < ((1e-6 if premult else 2e-5) if c == 0 else 1e-2), c


@pytest.mark.parametrize("premult,integ", [(1, "lf"), (0, "lf"), (1, "4s")])
def test_bf16_anchored_trajectory_is_reversible(premult, integ):
    """A bf16 trajectory whose stage gradient depends on the position alone is
    reversible: integrating (x1, -p1) with the same step returns to x0.  A
    chain returns to float32 rounding of the state (below 1e-6 of the largest
    coordinate after 25 steps) unless float32 rounding on the way back moves a
    coordinate of x - x_ref across a bf16 rounding boundary; the field then
    differs by one bf16 ulp of that offset and the chain returns to within
    5e-4 (measured: at most 1.3e-4).  A float64 emulation of this case with a
    gradient carried along the path instead (y += H_hi bf16(dx) from the
    trajectory's start) misses by 2.5-3.7e-3 on every chain."""
    rng, G, dd, var, pv, post, H, b, mu = _anchored_case(seed=13)
    n, d = G.shape
    B, L = 8, 25
    lam = np.linalg.eigvalsh(H + np.eye(d) / pv).max()
    prob = build_problem(post, precision="bf16", engine="gemm",
                         options={"premult_trajectory": premult, "incremental_residual": 1, "bf16_split_grad": 1})
    ch = Chains(prob, B)
    X0 = (mu[None, :] + 0.05 * rng.normal(size=(B, d))).astype(np.float32).astype(np.float64)
    ch.set_state(X0)
    ch.set_anchor()
    P0 = rng.normal(size=(B, d)).astype(np.float32).astype(np.float64)
    e = np.full(B, 0.5 / np.sqrt(lam))
    u = np.zeros(B)          # accept every finite proposal
    fwd = ch.propose(P0, e, u, L, integ, trajectory=True)
    assert fwd["accepted"].all()
    X1, P1 = fwd["X1"], fwd["P1"]
    assert _rel(X1, X0) > 1e-2                 # the trajectory moved
    ch.set_state(X1)
    back = ch.propose(-P1, e, u, L, integ, trajectory=True)
    err = np.array([_rel(back["X1"][c], X0[c]) for c in range(B)])
    assert err.min() < 1e-6, err
    assert err.max() < 5e-4, err


@pytest.mark.parametrize("precision,cfg,B,pm,order", [
    ("bf16", 1, 256, 0, 0), ("bf16", 3, 256, 0, 0), ("bf16", 3, 200, 0, 0), ("f64", 1, 256, 0, 0),
    ("bf16", 3, 256, 1, 0), ("f64", 1, 200, 1, 0), ("bf16", 4, 256, 1, 0), ("bf16", 4, 200, 0, 0),
    ("bf16", 4, 256, 1, 1), ("bf16", 3, 200, 0, 1)])
def test_split_tail_matches_unsplit(precision, cfg, B, pm, order):
    """Planning for 16 compute units makes both GEMMs of a 2500 x 2200 problem
    with 256 chains leave a partial last round (residual 40 tiles of 128 rows
    or 20 of 256, gradient 36 or 18; 200 chains leave a partly empty chain
    tile), so the tail tiles run as float32
    k-range partials summed by the split epilogue.  The trajectory equals the
    unsplit one up to float32 summation order (1e-5 of the largest
    coordinate after 5 leapfrog steps), its end-point Hamiltonians are FP64
    of those end points, and gradients stay within the bf16 tolerance of the
    oracle.  The FP64 endpoint GEMM (128-row tiles: 80 of them) splits its
    tail the same way, and so do both GEMMs of FP64 trajectories (to 1e-12).
    pm = 1: premultiplied trajectories, whose stage GEMMs (M = d = 2200)
    split their tails the same way.  order = 1 deals the tail's k-ranges one
    per XCD (when the split count divides 8) instead of by tile; cfg = 4 is
    the 256x128 tile of 8 waves of 128x32."""
    rng = np.random.default_rng(21)
    n, d, L = 2500, 2200, 5
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype="bfloat16")])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    X0 = (rng.normal(size=(B, d)) * 0.3).astype(np.float32).astype(np.float64)
    P0 = rng.normal(size=(B, d)).astype(np.float32).astype(np.float64)
    e = 0.002 * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    res, grad = {}, {}
    for name, opts in (("split", {"gemm_split_cus": 16}), ("plain", {"gemm_split_tail": 0})):
        base = {"gemm_config": cfg, "incremental_residual": 1, "tile_order": order} if precision == "bf16" else {}
        base["premult_trajectory"] = pm
        prob = build_problem(post, precision=precision, engine="gemm", options={**base, **opts})
        ch = Chains(prob, B)
        ch.set_state(X0)
        res[name] = ch.propose(P0, e, u, L, "lf", trajectory=True)
        grad[name] = prob.gradient_batch(X0[:4])
    # bf16: float32 summation order, and where it moves a coordinate of
    # x - x_ref across a bf16 rounding boundary, one bf16 ulp of that offset
    same = 1e-4 if precision == "bf16" else 1e-12
    assert _rel(res["split"]["X1"], res["plain"]["X1"]) < same
    assert _rel(res["split"]["P1"], res["plain"]["P1"]) < same
    assert _rel(res["split"]["H0"], res["plain"]["H0"]) < 1e-12
    for c in range(0, B, 37):
        chi1 = orc.misfit(res["split"]["X1"][c].reshape(d, 1))
        k1 = 0.5 * float(np.sum(res["split"]["P1"][c] ** 2))
        assert abs(res["split"]["H1"][c] - (chi1 + k1)) <= (1e-6 if precision == "bf16" else 1e-10) * abs(chi1 + k1), c
    for b in range(4):
        ref = orc.gradient(X0[b].reshape(d, 1))[:, 0]
        assert _rel(grad["split"][b], ref) < (2e-4 if precision == "bf16" else 1e-10)
        assert _rel(grad["split"][b], grad["plain"][b]) < same


@pytest.mark.parametrize("B,order", [(256, 0), (200, 0), (256, 1)])
def test_split_fixup_equals_split_epilogue(B, order):
    """Premultiplied bf16 stages whose tiles are all cut into two k-ranges
    (256x128 tiles of d = 2200 planned for 36 compute units: 18 tiles x 2):
    summing the two partials inside the GEMM (the second workgroup of a tile
    adds the first one's published partial) gives the same bits as the split
    epilogue kernel, whichever workgroup arrives first; repeated proposals
    reuse the reset tile counters.  200 chains leave a partly empty chain
    tile; order 1 deals the two k-ranges to different XCDs."""
    rng = np.random.default_rng(23)
    n, d, L = 2500, 2200, 5
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype="bfloat16")])
    X0 = (rng.normal(size=(B, d)) * 0.3).astype(np.float32).astype(np.float64)
    res = {}
    for fix in (0, 1):
        prob = build_problem(post, precision="bf16", engine="gemm",
                             options={"gemm_config": 3, "premult_trajectory": 1, "gemm_split_cus": 36,
                                      "tile_order": order, "split_fixup": fix})
        ch = Chains(prob, B)
        ch.set_state(X0)
        ch.set_anchor()
        r = np.random.default_rng(5)
        outs = []
        for k in range(3):
            P0 = r.normal(size=(B, d)).astype(np.float32).astype(np.float64)
            e = 0.002 * r.uniform(0.5, 1.5, size=B)
            u = r.uniform(size=B)
            outs.append(ch.propose(P0, e, u, L, "lf", trajectory=True))
        res[fix] = outs
        ch.close()
        prob.close()
    for a, b in zip(res[0], res[1]):
        for key in ("X1", "P1", "H0", "H1"):
            assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("precision,g_dtype,d,sym", [("bf16", "bfloat16", 300, 1), ("f64", None, 300, 1),
                                                     ("f64", None, 300, 0), ("f64", None, 100, 1),
                                                     ("f64", None, 520, 1), ("bf16", "bfloat16", 640, 1)])
def test_premultiplied_endpoint_matches_residual_form(precision, g_dtype, d, sym):
    """With n > d the GEMM engine evaluates the FP64 endpoint misfit as
    0.5 (x^T H x - 2 b^T x + d^T S^-1 d), H = G^T S^-1 G built once in FP64
    (the reference's premultiplied algebra, LinearMatrix.py:164-191), over the
    full H (sym 0) or its upper block triangle (sym 1, the default; d = 100,
    300, 520, 640 give 1, 3, 5 and 5 blocks of 128 rows with partial last
    blocks).  It equals the residual form and the oracle to 1e-11 relative, for
    misfits and for the H0/H1 of a proposal; vector variances enter H and b."""
    rng = np.random.default_rng(31)
    n, B = 700, 6
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    var = rng.uniform(0.005, 0.02, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0),
                        D.LinearMatrix(G, dd, var, dtype=g_dtype if g_dtype is not None else np.float64)])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, var, dtype=np.float64, premultiplication=False)])
    X = (rng.normal(size=(B, d)) * 0.5).astype(np.float32).astype(np.float64)
    P0 = rng.normal(size=(B, d)).astype(np.float32).astype(np.float64)
    e = 0.002 * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    chi, res = {}, {}
    # trajectories in the residual form in both runs: only the endpoint differs
    for name, opt in (("premult", {"endpoint_symmetric": sym, "premult_trajectory": 0}),
                      ("residual", {"endpoint_premult": 0})):
        prob = build_problem(post, precision=precision, engine="gemm", options=opt)
        chi[name] = prob.misfit_batch(X)
        ch = Chains(prob, B)
        ch.set_state(X)
        res[name] = ch.propose(P0, e, u, 4, "lf", trajectory=True)
    for b in range(B):
        assert _rel(chi["premult"][b], orc.misfit(X[b].reshape(d, 1))) < 1e-11
    assert _rel(chi["premult"], chi["residual"]) < 1e-11
    for k in ("H0", "H1"):
        assert _rel(res["premult"][k], res["residual"][k]) < 1e-11, k
    assert np.array_equal(res["premult"]["X1"], res["residual"]["X1"])


def test_bf16_single_plane_gradient_gemm_tolerance():
    """bf16_split_grad = 1 rounds the scaled residual to one bf16 plane in the
    gradient GEMM: gradients within 1e-2 of the oracle (stated tolerance of
    that mode), misfits still exact FP64."""
    rng = np.random.default_rng(4)
    n, d = 520, 384
    G = bf16_to_f32(bf16_round(rng.normal(size=(n, d)) / np.sqrt(d))).astype(np.float64)
    dd = rng.normal(size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype="bfloat16")])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    prob = build_problem(post, precision="bf16", options={"bf16_split_grad": 1})
    X = rng.normal(size=(4, d)).astype(np.float32).astype(np.float64)
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(4):
        x = X[b].reshape(d, 1)
        assert _rel(chi[b], orc.misfit(x)) < 1e-10
        assert _rel(g[b], orc.gradient(x)[:, 0]) < 1e-2


# --------------------------------------------------------------------------- #
def _teacher(post, orc, mass_g, mass_o, rng, B, L, integ, precision, engine, eps=0.01, x_scale=0.3,
             options=None):
    d = post.dimensions
    prob = build_problem(post, precision=precision, engine=engine,
                         mass_diagonal=mass_g._device_diagonal(), options=options)
    ch = Chains(prob, B)
    X0 = rng.normal(size=(B, d)) * x_scale
    ch.set_state(X0)
    X0s, chi0 = ch.get_state()
    P0 = np.stack([mass_o.generate_momentum()[:, 0] for _ in range(B)])
    e = eps * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    res = ch.propose(P0, e, u, L, integ, trajectory=True)
    tol = TOL[precision]
    near = 0
    for b in range(B):
        x0 = X0s[b].reshape(d, 1)
        ref = O.propose(orc, mass_o, x0, None, P0[b].reshape(d, 1), e[b], L, u[b], integ)
        assert _rel(chi0[b], ref["chi0"]) < max(tol, 1e-10)
        assert _rel(res["X1"][b], ref["x1"][:, 0]) < tol
        assert _rel(res["P1"][b], ref["p1"][:, 0]) < tol
        assert _rel(res["H1"][b], ref["h1"]) < tol
        assert _rel(res["H0"][b], ref["h0"]) < tol
        margin = abs(np.log(u[b]) - (ref["h0"] - ref["h1"]))
        if margin < 1e-6:
            near += 1
        else:
            assert bool(res["accepted"][b]) == ref["accepted"]
    return near


@pytest.mark.parametrize("engine", ["gemm", "fused"])
@pytest.mark.parametrize("integ", ["lf", "3s", "4s"])
@pytest.mark.parametrize("L", [1, 10])
def test_teacher_forced_proposal_f64(engine, integ, L):
    post, orc, rng = _dense_case(4, 150, 70, prior="both", bounds=True)
    d = post.dimensions
    _teacher(post, orc, MM.Unit(d), O.Unit(d, np.random.default_rng(5)), rng, 6, L, integ, "f64", engine)


@pytest.mark.parametrize("integ", ["lf", "3s", "4s"])
@pytest.mark.parametrize("L", [1, 10])
def test_teacher_forced_premultiplied_trajectory_f64(integ, L):
    """FP64 trajectories through H = G^T C^-1 G (one GEMM per stage, g = H x - b,
    LinearMatrix.py:203-207) against the oracle's residual-form trajectory
    (n = 150 > d = 70, both priors, bounds), FP64 tolerance."""
    post, orc, rng = _dense_case(4, 150, 70, prior="both", bounds=True)
    d = post.dimensions
    _teacher(post, orc, MM.Unit(d), O.Unit(d, np.random.default_rng(5)), rng, 6, L, integ, "f64", "gemm",
             options={"premult_trajectory": 1})


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("premult", [0, 1])
@pytest.mark.parametrize("integ", ["lf", "4s"])
def test_teacher_forced_fused_dense_forms(precision, premult, integ):
    """Fused dense engine (n = 150 > d = 70, both priors, bounds): trajectories
    in the residual form (premult_trajectory 0) and through H = G^T C^-1 G
    (the default for n > d: one pass over H per stage, the gradient of every
    owned parameter before any update) against the oracle's residual form."""
    post, orc, rng = _dense_case(4, 150, 70, prior="both", bounds=True)
    d = post.dimensions
    _teacher(post, orc, MM.Unit(d), O.Unit(d, np.random.default_rng(5)), rng, 6, 10, integ, precision, "fused",
             options={"premult_trajectory": premult})


@pytest.mark.parametrize("engine", ["gemm", "fused"])
def test_teacher_forced_diagonal_mass_f64(engine):
    post, orc, rng = _dense_case(6, 120, 64, prior="normal")
    d = post.dimensions
    diag = rng.uniform(0.5, 3.0, size=(d, 1))
    _teacher(post, orc, MM.Diagonal(diag), O.Diagonal(diag, np.random.default_rng(7)), rng, 4, 12,
             "lf", "f64", engine)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_teacher_forced_sparse(precision):
    dtype = np.float64 if precision == "f64" else np.float32
    post, orc, rng = _dense_case(8, 300, 200, prior="normal", sparse=True, g_dtype=dtype)
    d = post.dimensions
    diag = rng.uniform(0.5, 2.0, size=(d, 1))
    _teacher(post, orc, MM.Diagonal(diag), O.Diagonal(diag, np.random.default_rng(9)), rng, 9, 8,
             "lf", precision, "auto", eps=0.02)


@pytest.mark.parametrize("precision,tiles", [("f32", (1, 2, 4)), ("f64", (1, 2))])
def test_fused_tile_chains_do_not_change_results(precision, tiles):
    """Each chain's sums run in the same order whatever the number of chains
    a fused-engine workgroup carries, so 1, 2 or 4 chains per workgroup give
    bitwise identical proposals (sparse operator, host randoms)."""
    dtype = np.float64 if precision == "f64" else np.float32
    post, _, rng = _dense_case(8, 300, 200, prior="normal", sparse=True, g_dtype=dtype)
    d = post.dimensions
    B = 9
    X0 = rng.normal(size=(B, d)) * 0.3
    P0 = rng.normal(size=(B, d))
    e = 0.02 * rng.uniform(0.5, 1.5, size=B)
    u = rng.uniform(size=B)
    out = []
    for ct in tiles:
        prob = build_problem(post, precision=precision, engine="fused", options={"fused_tile_chains": ct})
        ch = Chains(prob, B)
        ch.set_state(X0)
        out.append(ch.propose(P0, e, u, 8, "lf", trajectory=True))
    for r in out[1:]:
        for k in ("X1", "P1", "H0", "H1"):
            assert np.array_equal(r[k], out[0][k]), k


def test_bounds_reflection_in_trajectory():
    # tight bounds force reflections inside the fused kick/drift epilogue
    rng = np.random.default_rng(10)
    d = 20
    lo, hi = -0.2 * np.ones(d), 0.25 * np.ones(d)
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.Uniform(lo, hi)])
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0), O.Uniform(lo, hi)])
    _teacher(post, orc, MM.Unit(d), O.Unit(d, np.random.default_rng(11)), rng, 5, 15, "lf", "f64",
             "auto", eps=0.05, x_scale=0.05)


# --------------------------------------------------------------------------- #
def _c1(seed=0):
    """The dense 1-chain configuration: G 100x50 N(0,1), noise 0.1, prior N(0,1)."""
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(100, 50))
    m_true = rng.normal(size=(50, 1))
    dd = G @ m_true + rng.normal(0, 0.1, size=(100, 1))
    post = D.BayesRule([D.Normal(np.zeros((50, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype=np.float64)])
    orc = O.BayesRule([O.Normal(np.zeros((50, 1)), 1.0),
                       O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
    return G, dd, post, orc


@pytest.mark.parametrize("engine", ["gemm", "fused"])
@pytest.mark.parametrize("stepsize", [0.004, 0.05])
def test_seeded_chain_matches_oracle(engine, stepsize):
    from hmclab_amd import Samplers

    G, dd, post, orc = _c1()
    mu, _ = O.analytic_linear_gaussian(G, dd, 0.01, np.zeros(50), 1.0)
    sam = Samplers.HMC(seed=7)
    sam.sample(None, post, stepsize=stepsize, amount_of_steps=10, proposals=60, engine=engine,
               initial_model=mu, disable_progressbar=True)
    got = sam.results[0]
    ref, rec, accepted = O.sample_chain(orc, O.Unit(50), seed=7, stepsize=stepsize, amount_of_steps=10,
                                        proposals=60, initial_model=mu, record=True)
    assert got.shape == ref.shape
    assert _rel(got[:-1], ref[:-1]) < 1e-10           # positions
    assert _rel(got[-1], ref[-1]) < 1e-10             # misfits
    assert sam.accepted_proposals == accepted
    if stepsize < 0.01:
        assert accepted > 20                          # the chain moves (stable step)
    else:
        assert accepted == 0                          # dt=0.05 exceeds the leapfrog stability limit


@pytest.mark.parametrize("engine", ["gemm", "fused"])
def test_device_rng_posterior_matches_analytic(engine):
    rng = np.random.default_rng(12)
    n, d, B = 40, 6, 256
    G = rng.normal(size=(n, d))
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.3, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.09, dtype=np.float64)])
    mu, Sigma = O.analytic_linear_gaussian(G, dd, 0.09, np.zeros(d), 1.0)
    prob = build_problem(post, precision="f64", engine=engine)
    ch = Chains(prob, B)
    ch.set_state(np.broadcast_to(mu.reshape(1, d), (B, d)))
    lam = np.linalg.eigvalsh(np.linalg.inv(Sigma)).max()
    eps = 1.0 / np.sqrt(lam)
    ch.run(50, 8, eps, seed=3)                      # burn-in
    ch.set_trace(200)
    step, acc = ch.run(200, 8, eps, seed=3, first_proposal=50)
    X, chi = ch.read_trace(0, 200)
    S = X.reshape(-1, d)
    assert acc.mean() / 200 > 0.5
    sd = np.sqrt(np.diag(Sigma))
    se = sd / np.sqrt(S.shape[0] / 10.0)
    assert np.all(np.abs(S.mean(axis=0) - mu) < 5 * se)
    C = np.cov(S.T)
    assert np.allclose(np.sqrt(np.diag(C)), sd, rtol=0.05)
    corr = C / np.outer(np.sqrt(np.diag(C)), np.sqrt(np.diag(C)))
    ref_corr = Sigma / np.outer(sd, sd)
    assert np.max(np.abs(corr - ref_corr)) < 0.05


def test_device_rng_is_reproducible_and_engine_independent():
    G, dd, post, orc = _c1(1)
    outs = []
    G, dd, _, _ = _c1(1)
    mu, _ = O.analytic_linear_gaussian(G, dd, 0.01, np.zeros(50), 1.0)
    accs = []
    for engine in ("gemm", "fused", "fused"):
        prob = build_problem(post, precision="f64", engine=engine)
        ch = Chains(prob, 4)
        ch.set_state(np.broadcast_to(mu.reshape(1, 50), (4, 50)))
        ch.set_trace(30)
        _, acc = ch.run(30, 10, 0.004, seed=99)
        outs.append(ch.read_trace(0, 30)[0])
        accs.append(acc)
    assert np.array_equal(outs[1], outs[2])            # bitwise reproducible
    assert np.array_equal(accs[0], accs[1]) and accs[0].sum() > 20
    assert _rel(outs[0], outs[1]) < 1e-9               # same draws in both engines


def test_set_state_rejects_infinite_misfit():
    d = 4
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.Uniform(-np.ones(d), np.ones(d))])
    ch = Chains(build_problem(post), 2)
    with pytest.raises(ValueError, match="inf"):
        ch.set_state(np.full((2, d), 2.0))


def test_premultiplication_keyword_selects_the_algebra():
    """LinearMatrix(premultiplication=True/False) selects the device algebra
    as the reference's keyword does (LinearMatrix.py:163-167): True on an
    n < d problem (where the default keeps the residual form) gives the same
    misfits and proposal as the explicit premultiplied options, False on an
    n > d problem the same as the explicit residual options; both match the
    oracle."""
    rng = np.random.default_rng(41)
    B = 4
    for n, d, flag in ((60, 90, True), (150, 70, False)):
        G = rng.normal(size=(n, d)) / np.sqrt(d)
        dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
        orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                           O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False)])
        X = rng.normal(size=(B, d)) * 0.3
        P0 = rng.normal(size=(B, d))
        e = np.full(B, 0.003)
        u = rng.uniform(size=B)
        out = []
        for kw, opts in ((dict(premultiplication=flag), None),
                         ({}, {"endpoint_premult": int(flag), "premult_trajectory": int(flag)})):
            post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype=np.float64, **kw)])
            prob = build_problem(post, precision="f64", engine="gemm", options=opts)
            chi = prob.misfit_batch(X)
            ch = Chains(prob, B)
            ch.set_state(X)
            res = ch.propose(P0, e, u, 5, "lf", trajectory=True)
            out.append((chi, res["X1"], res["H1"]))
            ch.close()
            prob.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        for b in range(B):
            assert _rel(out[0][0][b], orc.misfit(X[b].reshape(d, 1))) < 1e-11


# --------------------------------------------------------------------------- #
def test_reference_default_float32_storage():
    """LinearMatrix(G, d, var) without dtype stores G and d in float32, as the
    reference's backends do (LinearMatrix.py:148-155,318-329): misfits equal
    the oracle with float32 storage to 1e-10 (FP64 arithmetic on the same
    float32 values), gradients to 1e-6 (the reference's dense gradient uses the
    caller's uncast G, :206-208), and differ from float64 storage by more than
    1e-9; a dense and a sparse G, scalar and vector variance."""
    rng = np.random.default_rng(41)
    n, d = 90, 40
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    X = rng.normal(size=(5, d))
    for Gp, var in ((G, 0.01), (sp.csr_matrix(np.where(np.abs(G) > 0.1, G, 0.0)), rng.uniform(0.005, 0.02, (n, 1)))):
        post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(Gp, dd, var)])
        orc32 = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                             O.LinearMatrixSimple(Gp, dd, var, dtype=np.float32, premultiplication=False)])
        orc64 = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                             O.LinearMatrixSimple(Gp, dd, var, dtype=np.float64, premultiplication=False)])
        prob = build_problem(post, precision="f64")
        chi = prob.misfit_batch(X)
        g = prob.gradient_batch(X)
        for b in range(5):
            x = X[b].reshape(d, 1)
            # vector variances: the reference takes sigma = sqrt(float32 var) in float32
            tol = 1e-10 if np.ndim(var) == 0 else 1e-6
            assert _rel(chi[b], orc32.misfit(x)) < tol
            assert _rel(g[b], orc32.gradient(x)[:, 0]) < 1e-6
            assert abs(chi[b] - orc64.misfit(x)) > 1e-9 * abs(chi[b])
        prob.close()


def _full_cov_case(seed, n=120, d=40, sparse=False):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    if sparse:
        G = sp.csr_matrix(np.where(np.abs(G) > 0.12, G, 0.0))
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    A = rng.normal(size=(n, n)) / np.sqrt(n)
    C = 0.01 * (A @ A.T + 0.5 * np.eye(n))
    return rng, G, dd, C


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_full_data_covariance_matches_oracle(kind):
    """A full data covariance, pre-whitened on the host (G' = L^-1 G,
    d' = L^-1 d), against the oracle's dense-covariance backend
    (LinearMatrix.py:226-305) or sparse-covariance backend (:444-481) in
    FP64: misfits and gradients 1e-10, a teacher-forced 10-step proposal 1e-10."""
    rng, G, dd, C = _full_cov_case(43, sparse=(kind == "sparse"))
    d = G.shape[1]
    Cp = sp.csr_matrix(C) if kind == "sparse" else C
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, Cp, dtype=np.float64)])
    lik_o = (O.LinearMatrixSparseCovariance(G, dd, Cp, dtype=np.float64) if kind == "sparse"
             else O.LinearMatrixDenseCovariance(G, dd, C, dtype=np.float64, premultiplication=False))
    orc = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0), lik_o])
    prob = build_problem(post, precision="f64")
    X = 0.5 * rng.normal(size=(4, d))
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(4):
        x = X[b].reshape(d, 1)
        assert _rel(chi[b], orc.misfit(x)) < 1e-10
        assert _rel(g[b], orc.gradient(x)[:, 0]) < 1e-10
    ch = Chains(prob, 4)
    ch.set_state(X)
    P0 = rng.normal(size=(4, d))
    e = 0.002 * rng.uniform(0.5, 1.5, size=4)
    u = rng.uniform(size=4)
    res = ch.propose(P0, e, u, 10, "lf", trajectory=True)
    unit = O.Unit(d, np.random.default_rng(0))
    for c in range(4):
        ref = O.propose(orc, unit, X[c].reshape(d, 1), None, P0[c].reshape(d, 1), e[c], 10, u[c])
        assert _rel(res["X1"][c], ref["x1"][:, 0]) < 1e-10, c
        assert _rel(res["H1"][c], ref["h1"]) < 1e-10, c
        assert bool(res["accepted"][c]) == ref["accepted"], c


@pytest.mark.parametrize("with_likelihood", [True, False])
def test_full_covariance_normal_prior_matches_oracle(with_likelihood):
    """A full-covariance Normal prior (base.py:531,553-574), normalised, alone or
    with a dense likelihood, against the oracle: misfits and gradients 1e-10,
    a teacher-forced 10-step proposal 1e-10."""
    rng = np.random.default_rng(44)
    n, d = 70, 30
    A = rng.normal(size=(d, d)) / np.sqrt(d)
    Cm = A @ A.T + 0.3 * np.eye(d)
    mu = rng.normal(size=(d, 1)) * 0.2
    prior = D.Normal(mu.copy(), Cm.copy())
    prior.normalize()
    prior_o = O.Normal(mu, Cm)
    prior_o.normalize()
    comps, comps_o = [prior], [prior_o]
    if with_likelihood:
        G = rng.normal(size=(n, d)) / np.sqrt(d)
        dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
        comps.append(D.LinearMatrix(G, dd, 0.01, dtype=np.float64))
        comps_o.append(O.LinearMatrixSimple(G, dd, 0.01, dtype=np.float64, premultiplication=False))
    post, orc = D.BayesRule(comps), O.BayesRule(comps_o)
    prob = build_problem(post, precision="f64")
    X = mu[:, 0][None, :] + 0.3 * rng.normal(size=(4, d))
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    for b in range(4):
        x = X[b].reshape(d, 1)
        assert _rel(chi[b], orc.misfit(x)) < 1e-10
        assert _rel(g[b], orc.gradient(x)[:, 0]) < 1e-10
    ch = Chains(prob, 4)
    ch.set_state(X)
    P0 = rng.normal(size=(4, d))
    e = 0.01 * rng.uniform(0.5, 1.5, size=4)
    u = rng.uniform(size=4)
    res = ch.propose(P0, e, u, 10, "lf", trajectory=True)
    unit = O.Unit(d, np.random.default_rng(0))
    for c in range(4):
        ref = O.propose(orc, unit, X[c].reshape(d, 1), None, P0[c].reshape(d, 1), e[c], 10, u[c])
        assert _rel(res["X1"][c], ref["x1"][:, 0]) < 1e-10, c
        assert _rel(res["H1"][c], ref["h1"]) < 1e-10, c

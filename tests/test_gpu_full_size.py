"""The dense configuration at its full size, through the C-ABI.

G is 20000 x 8192, generated N(0, 1/8192) and rounded to bf16 (the target's
definition), noise sigma 0.01, Normal prior std 1.  The NumPy oracle is too
slow at this size, so the references are torch float64 products on the GPU
of the same bf16-exact G:

* misfit (FP64 endpoint): residual form 1e-10 relative; premultiplied form
  (the default since n > d) 1e-9 relative: 0.5 (x^T H x - 2 b^T x + d^T S^-1 d)
  cancels from about 2e8 down to chi ~ 1e4, so its FP64 rounding is ~1e-6
  absolute, far below what moves a Metropolis decision;
* gradient of bf16 trajectories (hi/lo split operands, FP32 accumulation):
  2e-3 relative to the largest component (the bench's trajectory gradient
  error is about 2e-3 with one-plane operands);
* one 25-step proposal: the Hamiltonians are the FP64 Hamiltonians of the
  start and returned end points (1e-9 relative, premultiplied endpoint).
"""
import numpy as np
import pytest

from hmclab_amd import Distributions as D
from hmclab_amd.Distributions import build_problem
from hmclab_amd.engine import Chains, bf16_round, bf16_to_f32

pytestmark = pytest.mark.gpu

N_DATA, N_PARAMS, SIGMA = 20000, 8192, 0.01


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


_CACHE = {}


def full_problem_cache():
    """The full-size target, built once per test session."""
    if "p" not in _CACHE:
        _CACHE["p"] = _make_full_problem()
    return _CACHE["p"]


@pytest.fixture(scope="module")
def full_problem():
    return full_problem_cache()


def _make_full_problem():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    G = rng.standard_normal(size=(N_DATA, N_PARAMS), dtype=np.float32) * np.float32(1.0 / np.sqrt(N_PARAMS))
    Gf = bf16_to_f32(bf16_round(G))
    del G
    m_true = rng.standard_normal(size=(N_PARAMS, 1))
    Gt = torch.from_numpy(Gf).to("cuda", torch.float64)
    dd = (Gt @ torch.from_numpy(m_true).to("cuda")).cpu().numpy() + rng.normal(0.0, SIGMA, size=(N_DATA, 1))
    post = D.BayesRule([D.Normal(np.zeros((N_PARAMS, 1)), 1.0),
                        D.LinearMatrix(Gf, dd, SIGMA ** 2, dtype="bfloat16")])
    # chains near the posterior: the ridge estimate plus noise
    X = (m_true.T + 0.01 * rng.standard_normal(size=(4, N_PARAMS))).astype(np.float32).astype(np.float64)
    return torch, Gt, dd, post, X


def _reference(torch, Gt, dd, X):
    Xt = torch.from_numpy(X).to("cuda")
    r = Gt @ Xt.T - torch.from_numpy(dd).to("cuda")                 # (n, B)
    chi = 0.5 * (r * r).sum(dim=0) / SIGMA ** 2 + 0.5 * (Xt * Xt).sum(dim=1)
    grad = (Gt.T @ r).T / SIGMA ** 2 + Xt
    return chi.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("premult", [1, 0])
def test_full_size_misfit_and_gradient(full_problem, premult):
    torch, Gt, dd, post, X = full_problem
    prob = build_problem(post, precision="bf16", engine="gemm", options={"endpoint_premult": premult})
    chi = prob.misfit_batch(X)
    g = prob.gradient_batch(X)
    chi_ref, g_ref = _reference(torch, Gt, dd, X)
    assert _rel(chi, chi_ref) < (1e-9 if premult else 1e-10)
    for b in range(X.shape[0]):
        assert _rel(g[b], g_ref[b]) < 2e-3, b
    prob.close()


def test_full_size_proposal_hamiltonian(full_problem):
    torch, Gt, dd, post, X = full_problem
    prob = build_problem(post, precision="bf16", engine="gemm",
                         options={"gemm_config": 3, "incremental_residual": 1, "bf16_split_grad": 1})
    B = X.shape[0]
    ch = Chains(prob, B)
    ch.set_state(X)
    rng = np.random.default_rng(6)
    P0 = rng.standard_normal(size=(B, N_PARAMS)).astype(np.float32).astype(np.float64)
    res = ch.propose(P0, np.full(B, 1e-3), rng.uniform(size=B), 25, "lf", trajectory=True)
    chi1, _ = _reference(torch, Gt, dd, res["X1"])
    k1 = 0.5 * np.sum(res["P1"] ** 2, axis=1)
    assert _rel(res["H1"], chi1 + k1) < 1e-9
    chi0, _ = _reference(torch, Gt, dd, X)
    k0 = 0.5 * np.sum(P0 ** 2, axis=1)
    assert _rel(res["H0"], chi0 + k0) < 1e-9
    ch.close()
    prob.close()


def test_full_size_premultiplied_trajectory():
    """bf16 trajectories through H (y = H x - b from hi/lo planes of H and x,
    then y += H_hi dx per stage) against residual-form bf16 trajectories from
    the same start and momenta: end points within 5e-3 of the largest
    coordinate (both are bf16 approximations of one FP64 trajectory), and
    the Hamiltonians are the FP64 Hamiltonians of the returned end points."""
    torch, Gt, dd, post, X = full_problem_cache()
    B = X.shape[0]
    rng = np.random.default_rng(7)
    P0 = rng.standard_normal(size=(B, N_PARAMS)).astype(np.float32).astype(np.float64)
    ends = []
    for premult in (0, 1):
        prob = build_problem(post, precision="bf16", engine="gemm",
                             options={"gemm_config": 3, "incremental_residual": 1, "bf16_split_grad": 1,
                                      "premult_trajectory": premult})
        ch = Chains(prob, B)
        ch.set_state(X)
        res = ch.propose(P0, np.full(B, 1e-3), rng.uniform(size=B), 25, "lf", trajectory=True)
        chi1, _ = _reference(torch, Gt, dd, res["X1"])
        k1 = 0.5 * np.sum(res["P1"] ** 2, axis=1)
        assert _rel(res["H1"], chi1 + k1) < 1e-9
        ends.append(res["X1"])
        ch.close()
        prob.close()
    assert _rel(ends[1], ends[0]) < 5e-3

"""hmc_misfit_batch_noalloc (device-resident models and outputs, scratch in a
static device array) against hmc_misfit_batch through the C-ABI: two calls
in a row, the second with fewer models, for a dense target on the GEMM engine
(premultiplied FP64 endpoint, n > d), a dense bf16 target in the residual
form, and a sparse target on the fused engine.  FP64 misfits agree to 1e-12
relative (the no-allocation call runs the endpoint GEMM without the split
tail, so only the summation order differs)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from hmclab_amd import Distributions as D
from hmclab_amd import _lib
from hmclab_amd.Distributions import build_problem

pytestmark = pytest.mark.gpu


def _dev_ptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def _noalloc(torch, prob, M):
    Md = torch.from_numpy(np.ascontiguousarray(M)).to("cuda")
    out = torch.full((M.shape[0],), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.call("hmc_misfit_batch_noalloc", prob._h, _dev_ptr(Md), M.shape[0], _dev_ptr(out))
    return out.cpu().numpy()


def _targets():
    rng = np.random.default_rng(5)
    out = []
    n, d = 2500, 2200
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    dd = G @ rng.normal(size=(d, 1)) + rng.normal(0, 0.1, size=(n, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01)])
    out.append(("dense_f64_premult", build_problem(post, precision="f64", engine="gemm"), d, 0.3))
    n2, d2 = 600, 900
    G2 = rng.normal(size=(n2, d2)) / np.sqrt(d2)
    dd2 = G2 @ rng.normal(size=(d2, 1))
    post2 = D.BayesRule([D.Laplace(np.zeros((d2, 1)), np.full((d2, 1), 0.5)),
                         D.LinearMatrix(G2, dd2, 0.05, dtype="bfloat16")])
    out.append(("dense_bf16_residual", build_problem(post2, precision="bf16", engine="gemm"), d2, 0.5))
    Gs = sp.random(400, 300, density=0.05, random_state=3, format="csr")
    ds = Gs @ rng.normal(size=(300, 1))
    post3 = D.BayesRule([D.Normal(np.zeros((300, 1)), 2.0), D.LinearMatrix(Gs, ds, 0.02)])
    out.append(("sparse_f64_fused", build_problem(post3, precision="f64"), 300, 1.0))
    return out


def test_noalloc_misfit_matches_allocating_entry_point():
    import torch

    rng = np.random.default_rng(6)
    for name, prob, d, scale in _targets():
        print(f"[no-alloc] {name}", flush=True)
        for B in (300, 70):          # two calls in a row, the second smaller
            M = scale * rng.normal(size=(B, d))
            if prob.precision == "bf16":
                M = M.astype(np.float32).astype(np.float64)   # float32 state
            ref = prob.misfit_batch(M)
            got = _noalloc(torch, prob, M)
            assert np.all(np.isfinite(got)), name
            assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12, (name, B)
        prob.close()


def test_noalloc_rejects_oversized_batches_and_unbuilt_operator():
    import torch

    rng = np.random.default_rng(7)
    n, d = 3000, 2200
    G = rng.normal(size=(n, d)) / np.sqrt(d)
    dd = G @ rng.normal(size=(d, 1))
    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0), D.LinearMatrix(G, dd, 0.01)])
    prob = build_problem(post, precision="f64", engine="gemm")
    M = 0.1 * rng.normal(size=(8, d))
    # H (n > d: premultiplied endpoint) is not built before the first misfit
    with pytest.raises(AssertionError, match="premultiplied operator"):
        _noalloc(torch, prob, M)
    ref = prob.misfit_batch(M)
    assert np.max(np.abs(_noalloc(torch, prob, M) - ref) / np.abs(ref)) < 1e-12
    # 4096 rows of 2240 doubles need 73 MB of scratch: rejected before any work
    big = torch.zeros((4096, d), dtype=torch.float64, device="cuda")
    out = torch.zeros(4096, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(ValueError, match="32 MiB"):
        _lib.call("hmc_misfit_batch_noalloc", prob._h, _dev_ptr(big), 4096, _dev_ptr(out))
    prob.close()

"""Benchmark: min-ESS/sec of batched HMC on one MI355X.

Default workload (the dense configuration the metric is quoted on):
G 20000 x 8192 generated FP64 N(0, 1/8192) and rounded to bf16, m_true ~
N(0, I), noise sigma 0.01, Normal prior std 1, Unit mass, 512 chains,
25 leapfrog steps; trajectory GEMMs on bf16 MFMA (hi/lo split operands, FP32
accumulation), endpoint Hamiltonian and Metropolis in FP64.  A "step" is
one proposal for all 512 chains.  Chains start from exact posterior draws;
W warm-up proposals run untimed, K proposals are timed; min-ESS is the
rank-normalised split bulk ESS minimised over all 8192 parameters, pooled
over chains.  Posterior mean/std errors are against the analytic FP64
posterior of the bf16-rounded G.

The roofline line is measured live with HIP events on the engine's stream
around every trajectory GEMM launch (algorithmic FLOP per launch =
2 n d B).  The CPU baseline times the oracle (the reference's op sequence:
L gradients + 2 misfits per proposal, residual form, float64, one core) on
a bounded sample and converts with the GPU's measured ESS per proposal.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"bf16": 2.5e15, "f64": 78.6e12, "f32": 157.3e12, "hbm": 8.0e12}
METRIC = "min-ESS/sec per GPU: surrogate+batched vs plain HMC at matched posterior error"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_dense_problem(seed, n, d, sigma):
    rng = np.random.default_rng(seed)
    from hmclab_amd.engine import bf16_round, bf16_to_f32

    G = rng.standard_normal(size=(n, d), dtype=np.float32) * np.float32(1.0 / np.sqrt(d))
    Gb = bf16_round(G)                    # bits, the target's definition
    Gf = bf16_to_f32(Gb)                  # exact values as float32
    m_true = rng.standard_normal(size=(d, 1))
    dd = Gf.astype(np.float64) @ m_true + rng.normal(0.0, sigma, size=(n, 1))
    return Gb, Gf, dd


def analytic_posterior(torch, Gf, dd, sigma2, prior_var=1.0):
    """Sigma = (G^T G / s2 + I / pv)^-1, mu = Sigma G^T d / s2 (FP64 on the GPU)."""
    G = torch.from_numpy(Gf).to("cuda", torch.float64)
    H = G.T @ G / sigma2
    H.diagonal().add_(1.0 / prior_var)
    rhs = G.T @ torch.from_numpy(dd[:, 0]).to("cuda", torch.float64) / sigma2
    L = torch.linalg.cholesky(H)
    mu = torch.cholesky_solve(rhs[:, None], L)[:, 0]
    Linv = torch.linalg.solve_triangular(L, torch.eye(H.shape[0], device="cuda", dtype=torch.float64), upper=False)
    var = (Linv * Linv).sum(dim=0)        # diag(H^-1) = column norms of L^-1
    lam_max = torch.linalg.eigvalsh(H)[-1].item()
    del G, H
    return mu, var, L, lam_max


def posterior_draws(torch, mu, L, B, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    z = torch.randn(mu.shape[0], B, device="cuda", dtype=torch.float64, generator=g)
    x = torch.linalg.solve_triangular(L.T, z, upper=True)    # x ~ N(0, H^-1)
    return (mu[:, None] + x).T.contiguous().cpu().numpy()


def cpu_baseline(Gf, dd, sigma2, L, eps, proposals, seed):
    """Reference-faithful single-chain proposals with the oracle on one core."""
    try:
        from threadpoolctl import threadpool_limits
    except Exception:   # pragma: no cover
        threadpool_limits = None
    from oracle import hmc_oracle as O

    d = Gf.shape[1]
    G64 = Gf.astype(np.float64)
    post = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                        O.LinearMatrixSimple(G64, dd, sigma2, dtype=np.float64, premultiplication=False)])
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    if ctx:
        ctx.__enter__()
    try:
        t0 = time.perf_counter()
        O.sample_chain(post, O.Unit(d), seed=seed, stepsize=eps, amount_of_steps=L, proposals=proposals)
        dt = time.perf_counter() - t0
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
    return proposals / dt, dt


def trace_to_torch(torch, ch, K):
    """Device trace (K, B, d) -> torch float32 tensor (B, K, d) on the GPU."""
    out = torch.empty((ch.B, K, ch.d), dtype=torch.float32, device="cuda")
    step = max(1, min(K, (1 << 27) // (ch.B * ch.d)))
    for k0 in range(0, K, step):
        k1 = min(K, k0 + step)
        X, _ = ch.read_trace(k0, k1 - k0)
        out[:, k0:k1, :] = torch.from_numpy(X.astype(np.float32)).to("cuda").permute(1, 0, 2)
    return out


def run_dense(args):
    import torch

    from hmclab_amd import Distributions as D
    from hmclab_amd import diagnostics
    from hmclab_amd.Distributions import build_problem
    from hmclab_amd.engine import Chains

    n, d, B, L, sigma = args.n, args.d, args.chains, args.steps_per_traj, 0.01
    t_setup = time.perf_counter()
    Gb, Gf, dd = make_dense_problem(args.seed, n, d, sigma)
    log(f"problem {n}x{d} generated in {time.perf_counter() - t_setup:.1f}s")
    mu, var, Lc, lam_max = analytic_posterior(torch, Gf, dd, sigma * sigma)
    eps = args.eps_scale / np.sqrt(lam_max)
    X0 = posterior_draws(torch, mu, Lc, B, args.seed + 1)
    mu_np, sd_np = mu.cpu().numpy(), np.sqrt(var.cpu().numpy())
    del Lc
    torch.cuda.empty_cache()
    log(f"analytic posterior: lam_max={lam_max:.4g} eps={eps:.4g} ({time.perf_counter() - t_setup:.1f}s)")

    post = D.BayesRule([D.Normal(np.zeros((d, 1)), 1.0),
                        D.LinearMatrix(Gf.astype(np.float64), dd, sigma * sigma, dtype="bfloat16")])
    prob = build_problem(post, precision=args.precision, engine="gemm")
    ch = Chains(prob, B)
    ch.set_state(X0)
    log(f"device problem ready ({time.perf_counter() - t_setup:.1f}s)")

    step = np.full(B, eps)
    if args.warmup:
        ch.run(args.warmup, L, step, seed=args.seed, first_proposal=0)
    ch.set_trace(args.steps)
    ch.enable_timing(True)
    ch.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, acc = ch.run(args.steps, L, step, seed=args.seed, first_proposal=args.warmup)
    ch.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gemm_ms, gemm_n = ch.timing(0)
    end_ms, end_n = ch.timing(1)
    ew_ms, ew_n = ch.timing(2)
    log(f"timed {args.steps} proposals in {elapsed:.3f}s; gemm {gemm_ms:.1f} ms/{gemm_n}, "
        f"endpoint {end_ms:.1f} ms/{end_n}, elementwise {ew_ms:.1f} ms/{ew_n}")

    draws = trace_to_torch(torch, ch, args.steps)
    ess = diagnostics.ess_bulk(draws)
    min_ess = float(np.min(ess))
    pooled = draws.to(torch.float64).reshape(-1, d)
    mean = pooled.mean(dim=0).cpu().numpy()
    std = pooled.std(dim=0).cpu().numpy()
    del draws, pooled
    err_mean = float(np.sqrt(np.mean(((mean - mu_np) / sd_np) ** 2)))
    err_std = float(np.sqrt(np.mean((std / sd_np - 1.0) ** 2)))
    acc_rate = float(acc.mean() / args.steps)
    value = min_ess / elapsed

    # roofline of the dominant kernel (each trajectory GEMM launch = 2 n d B flops)
    flop_launch = 2.0 * n * d * B
    avg_s = gemm_ms / 1e3 / max(gemm_n, 1)
    achieved = flop_launch / avg_s
    peak = PEAK["bf16"] if args.precision == "bf16" else PEAK[args.precision]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_dense_gemm.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "ESS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": f"dense linear-Gaussian {n}x{d}, G stored bf16, {B} chains, L={L}, "
                        "plain batched HMC (no surrogate), bf16 trajectory + fp64 endpoint",
            "chains": B, "n_data": n, "n_params": d, "leapfrog_steps": L, "stepsize": eps,
            "min_ess": min_ess, "acceptance_rate": acc_rate,
            "posterior_rms_mean_err_over_sd": err_mean, "posterior_rms_std_rel_err": err_std,
            "proposals_per_s": args.steps / elapsed,
        },
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": "k_gemm_bf16 (trajectory GEMMs)", "avg_launch_us": avg_s * 1e6},
    }

    if args.cpu_baseline:
        ess_per_prop_chain = min_ess / (args.steps * B)
        props_s, dt = cpu_baseline(Gf, dd, sigma * sigma, L, eps, args.cpu_proposals, args.seed)
        result["cpu_baseline"] = {
            "value": ess_per_prop_chain * props_s, "unit": "ESS/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_proposals} proposals of one chain (oracle, float64 G, residual form, "
                      f"L={L}) in {dt:.1f}s = {props_s:.4g} proposals/s, times the GPU's measured "
                      f"min-ESS per proposal per chain {ess_per_prop_chain:.4g}",
        }
    ch.close()
    prob.close()
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--d", type=int, default=8192)
    ap.add_argument("--chains", type=int, default=512)
    ap.add_argument("--steps-per-traj", type=int, default=25)
    ap.add_argument("--eps-scale", type=float, default=1.0)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "f64"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-proposals", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); chains shard across GPUs "
              "without communication, so the one-GPU figure is the per-GPU rate.", file=sys.stderr)
        sys.exit(2)
    res = run_dense(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: min-ESS/sec of batched HMC on one MI355X.

Default workload (the dense configuration the metric is quoted on):
G 20000 x 8192 generated FP64 N(0, 1/8192) and rounded to bf16, m_true ~
N(0, I), noise sigma 0.01, Normal prior std 1, Unit mass, 512 chains,
25 leapfrog steps.  Two legs run on the same target and the same exact
posterior draws as starting states:

* plain batched FP64 GPU HMC (the comparator): trajectory GEMMs on FP64 MFMA
  with the measured leg's trajectory form, and (default) also the
  residual-form FP64 comparator;
* surrogate+batched (the reported value): trajectory GEMMs on bf16 MFMA
  (FP32 accumulation), endpoint Hamiltonian and Metropolis in FP64; an MLP
  surrogate of the likelihood (8192 -> 512 -> 512 -> 1, SiLU) is trained on
  the GPU on (m, misfit, gradient) pairs from warm-up proposals.  Whether it
  drives the timed trajectories is decided during warm-up, before timing
  starts: it is attached only if its held-out gradient error is below
  --sur-max-grad-err and a warm-up probe of --sur-probe proposals keeps
  acceptance above the fallback threshold; attached, it still falls back to
  the exact gradient when a chunk's acceptance drops below the threshold.
  The line reports the fraction of timed proposals the surrogate drove.

Both legs default to premultiplied trajectories, the reference's algebra for
n > d (LinearMatrix.py:163-177): H = G^T C^-1 G and b = G^T C^-1 d are built
once in FP64; the FP64 leg computes g = H x - b per stage, the bf16 leg the
anchored field g = (H x_ref - b) + bf16(H) bf16(x - x_ref) with x_ref the
mean of the starting draws (a function of the position alone, so the
trajectories stay reversible).  --premult-trajectory 0 runs both legs in the
residual form (g = G^T (G x - d) / var).

A "step" is one proposal for all chains.  W warm-up proposals run untimed,
K proposals are timed.  value = (min-ESS per proposal) x (timed proposals per
second): min-ESS is the rank-normalised split bulk ESS minimised over all
8192 parameters, pooled over chains, estimated from E = max(K, --ess-draws)
draws per chain (the K timed draws continued untimed by the same chains; the
split-chain estimator overstates ESS on short chains: 1.32x at 20 draws,
1.05x at 200, against 2000-draw runs of both legs), and min-ESS of the K
timed draws alone is reported beside it.  Posterior mean/std errors of both
legs are against the analytic FP64 posterior of the bf16-rounded G
("matched posterior error").

The roofline line is measured live with HIP events on the engine's stream
around every trajectory stage GEMM launch and its split-tail epilogue
(algorithmic FLOP per stage = 2 d^2 B premultiplied, 2 n d B per GEMM in the
residual form).  The CPU baseline times the oracle (the reference's op sequence:
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
    lam_max = _power_lam_max(torch, H)
    del G, H
    return mu, var, L, lam_max


def _power_lam_max(torch, H, iters=60):
    """Largest eigenvalue of a symmetric positive definite H by power iteration
    (only the initial step size depends on it; the warm-up tunes the step)."""
    g = torch.Generator(device=H.device).manual_seed(0)
    v = torch.randn(H.shape[0], device=H.device, dtype=H.dtype, generator=g)
    lam = 0.0
    for _ in range(iters):
        w = H @ v
        lam = float(torch.dot(v, w) / torch.dot(v, v))
        v = w / torch.linalg.vector_norm(w)
    return lam


def posterior_draws(torch, mu, L, B, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    z = torch.randn(mu.shape[0], B, device="cuda", dtype=torch.float64, generator=g)
    x = torch.linalg.solve_triangular(L.T, z, upper=True)    # x ~ N(0, H^-1)
    return (mu[:, None] + x).T.contiguous().cpu().numpy()


def cpu_baseline_child(args):
    """Runs in a fresh process with BLAS pinned to one thread: regenerate the
    problem, then time single-chain oracle proposals (the reference's op
    sequence: L gradients + 2 misfits, residual form, float64 G)."""
    from oracle import hmc_oracle as O

    _, Gf, dd = make_dense_problem(args.seed, args.n, args.d, 0.01)
    d = args.d
    G64 = Gf.astype(np.float64)
    del Gf
    post = O.BayesRule([O.Normal(np.zeros((d, 1)), 1.0),
                        O.LinearMatrixSimple(G64, dd, 1e-4, dtype=np.float64, premultiplication=False)])
    t0 = time.perf_counter()
    O.sample_chain(post, O.Unit(d), seed=args.seed, stepsize=args.cpu_eps, amount_of_steps=args.steps_per_traj,
                   proposals=args.cpu_proposals)
    dt = time.perf_counter() - t0
    print(json.dumps({"proposals_per_s": args.cpu_proposals / dt, "seconds": dt}), flush=True)


def cpu_baseline(args, eps):
    """Runs `cpu_procs` single-chain oracle processes at once (one core each:
    BLAS pinned to one thread, no GPU), as independent CPU chains would run
    one process per chain; returns the per-process proposals/s and seconds."""
    import subprocess

    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "BLIS_NUM_THREADS"):
        env[k] = "1"
    env["HIP_VISIBLE_DEVICES"] = ""
    procs = []
    for i in range(max(1, args.cpu_procs)):
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child", "--seed", str(args.seed),
               "--n", str(args.n), "--d", str(args.d), "--steps-per-traj", str(args.steps_per_traj),
               "--cpu-proposals", str(args.cpu_proposals), "--cpu-eps", repr(float(eps))]
        procs.append(subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    rates, secs = [], []
    for pr in procs:
        out, err = pr.communicate(timeout=900)
        if pr.returncode != 0:
            raise RuntimeError(f"cpu baseline failed: {err[-2000:]}")
        res = json.loads(out.strip().splitlines()[-1])
        rates.append(res["proposals_per_s"])
        secs.append(res["seconds"])
    return rates, secs


def _cpu_model():
    """Host CPU model name and logical CPU count (from /proc/cpuinfo)."""
    try:
        names = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs visible)" if names else None
    except OSError:
        return None


def trace_to_torch(torch, ch, K):
    """Device trace (K, B, d) -> torch float32 tensor (B, K, d) on the GPU."""
    out = torch.empty((ch.B, K, ch.d), dtype=torch.float32, device="cuda")
    step = max(1, min(K, (1 << 27) // (ch.B * ch.d)))
    for k0 in range(0, K, step):
        k1 = min(K, k0 + step)
        X, _ = ch.read_trace(k0, k1 - k0)
        out[:, k0:k1, :] = torch.from_numpy(X.astype(np.float32)).to("cuda").permute(1, 0, 2)
    return out


def _posterior_errors(torch, ch, K, mu_np, sd_np, d, K_timed=None):
    """min-ESS and the RMS errors of the pooled mean (in units of the reference
    sd) and std (relative) against the reference moments (mu_np, sd_np; None:
    no reference), from the K recorded draws per chain.  Also returns the
    pooled mean and std, and the min-ESS of the first K_timed draws alone."""
    from hmclab_amd import diagnostics

    draws = trace_to_torch(torch, ch, K)
    ess = diagnostics.ess_bulk(draws) if K >= 4 else np.full(d, np.nan)   # short profiling runs
    if K_timed is not None and K_timed < K:
        ess_timed = diagnostics.ess_bulk(draws[:, :K_timed].contiguous()) if K_timed >= 4 else np.full(d, np.nan)
    else:
        ess_timed = ess
    pooled = draws.to(torch.float64).reshape(-1, d)
    mean = pooled.mean(dim=0).cpu().numpy()
    std = pooled.std(dim=0).cpu().numpy()
    del draws, pooled
    err_mean = err_std = None
    if mu_np is not None:
        err_mean = float(np.sqrt(np.mean(((mean - mu_np) / sd_np) ** 2)))
        err_std = float(np.sqrt(np.mean((std / sd_np - 1.0) ** 2)))
    return float(np.min(ess)), err_mean, err_std, mean, std, float(np.min(ess_timed))


def _kernel_times(ch):
    return {k: ch.timing(k) for k in range(8)}


def _trajectory_gradient_error(post, ch, X, precision):
    """Relative error (max-norm per row; worst row and RMS over rows) of the
    gradient field the chains' trajectories integrate (for bf16 the anchored
    field, hmc_chains_trajectory_gradient) against the FP64 gradient of the
    same target, at positions X (B, d)."""
    from hmclab_amd.Distributions import build_problem

    if precision == "f64":
        return 0.0, 0.0
    ref = build_problem(post, precision="f64", engine="gemm", options={"premult_trajectory": 0})
    try:
        g64 = ref.gradient_batch(X)
    finally:
        ref.close()
    g = ch.trajectory_gradient(X)
    e = np.max(np.abs(g - g64), axis=1) / np.max(np.abs(g64), axis=1)
    return float(np.max(e)), float(np.sqrt(np.mean(e ** 2)))


def _plain_fp64_leg(args, torch, post, X0, lam_max, mu_np, sd_np, premult):
    """Comparator: plain batched FP64 GPU HMC (exact gradients on FP64 MFMA),
    trajectories premultiplied (g = H x - b, the reference's form for n > d)
    or in the residual form (g = G^T (G x - d) / var)."""
    from hmclab_amd.Distributions import build_problem
    from hmclab_amd.engine import Chains

    B, L, d = args.chains, args.steps_per_traj, args.d
    prob = build_problem(post, precision="f64", engine="gemm", options={"premult_trajectory": int(premult)})
    ch = Chains(prob, B)
    ch.set_state(X0)
    sub = argparse.Namespace(**vars(args))
    sub.warmup = min(args.warmup, 5)
    eps, done = _tune(ch, B, L, args.eps_scale / np.sqrt(lam_max), sub, args.seed + 7, "fp64 ")
    K = args.plain_steps
    E = max(K, args.ess_draws)
    elapsed, acc = _timed(ch, B, L, eps, K, args.seed + 7, done, torch, trace=E, every=args.timing_every)
    kt = _kernel_times(ch)
    if E > K:   # untimed continuation of the same chains for the ESS estimate
        ch.run(E - K, L, np.full(B, eps), seed=args.seed + 7, first_proposal=done + K)
    min_ess, em, es, mean, std, min_ess_timed = _posterior_errors(torch, ch, E, mu_np, sd_np, d, K)
    ch.close()
    prob.close()
    value = min_ess / E * K / elapsed
    log(f"plain fp64 ({'premultiplied' if premult else 'residual form'}): {K} proposals in "
        f"{elapsed:.2f}s, acceptance {acc.mean() / K:.3f}, min ESS {min_ess:.0f} over {E} draws "
        f"-> {value:.1f} ESS/s")
    return dict(trajectory_form="premultiplied" if premult else "residual",
                value=value, min_ess=min_ess, ess_draws=E, min_ess_timed_draws=min_ess_timed,
                seconds=elapsed, proposals=args.plain_steps,
                acceptance_rate=float(acc.mean() / args.plain_steps), stepsize=eps,
                posterior_rms_mean_err_over_sd=em, posterior_rms_std_rel_err=es,
                ms_per_proposal=elapsed / args.plain_steps * 1e3, _mean=mean, _std=std,
                stage_fp64_avg_us=(kt[4][0] / max(kt[4][1], 1) * 1e3 if premult
                                   else (kt[0][0] + kt[3][0]) / max(kt[0][1] + kt[3][1], 1) * 1e3),
                endpoint_fp64_avg_us=kt[1][0] / max(kt[1][1], 1) * 1e3)


def _train_surrogate(args, torch, ch, post, B, L, eps, done):
    """Warm-up plain HMC samples -> exact (m, misfit, gradient) pairs -> Adam."""
    from hmclab_amd.surrogate import Surrogate, training_pairs

    d = args.d
    t0 = time.perf_counter()
    K = args.sur_warmup
    ch.set_trace(K)
    ch.run(K, L, np.full(B, eps), seed=args.seed + 3, first_proposal=done)
    X, _ = ch.read_trace(0, K)
    ch.set_trace(0)            # the probe and the timed phase set their own traces
    X = X.reshape(K * B, d)
    chi, grad = training_pairs(post, X)
    t_pairs = time.perf_counter() - t0
    sur = Surrogate(d, args.sur_hidden, seed=args.seed, mode="bf16")
    t1 = time.perf_counter()
    hist = sur.train(X, chi, grad, epochs=args.sur_epochs, batch=256, learning_rate=args.sur_lr)
    t_train = time.perf_counter() - t1
    # held-out check at the chains' current states
    Xc, _ = ch.get_state()
    chi_c, g_c = training_pairs(post, Xc)
    f_s, g_s = sur.evaluate(Xc)
    gc = g_c - g_c.mean(axis=0)
    log(f"surrogate held-out set: {Xc.shape[0]} chain states, position spread {float(np.std(Xc, axis=0).mean()):.4g} "
        f"(training {float(np.std(X, axis=0).mean()):.4g}), misfit sd {float(np.std(chi_c)):.4g} "
        f"(training {float(np.std(chi)):.4g}), gradient spread {float(np.linalg.norm(gc)):.4g}")
    grad_err = float(np.linalg.norm(g_s - g_c) / np.linalg.norm(gc))
    misfit_err = float(np.sqrt(np.mean((f_s - chi_c) ** 2)) / max(np.std(chi_c), 1e-300))
    del X, chi, grad
    log(f"surrogate: {K * B} pairs in {t_pairs:.1f}s, {args.sur_epochs} epochs in {t_train:.1f}s, "
        f"loss {hist[0]:.4g} -> {hist[-1]:.4g}; held-out gradient error {grad_err:.3f} (relative to "
        f"the gradient's spread), misfit error {misfit_err:.3f} sd")
    return sur, dict(pairs=K * B, warmup_and_pairs_s=t_pairs, train_s=t_train, epochs=args.sur_epochs,
                     hidden=args.sur_hidden, loss_first=float(hist[0]), loss_last=float(hist[-1]),
                     heldout_grad_rel_err=grad_err, heldout_misfit_err_over_sd=misfit_err,
                     proposals_done=done + K)


def run_dense(args):
    import torch

    from hmclab_amd import Distributions as D
    from hmclab_amd.Distributions import build_problem
    from hmclab_amd.engine import Chains

    n, d, B, L, sigma = args.n, args.d, args.chains, args.steps_per_traj, 0.01
    t_setup = time.perf_counter()
    Gb, Gf, dd = make_dense_problem(args.seed, n, d, sigma)
    log(f"problem {n}x{d} generated in {time.perf_counter() - t_setup:.1f}s")
    mu, var, Lc, lam_max = analytic_posterior(torch, Gf, dd, sigma * sigma)
    X0 = posterior_draws(torch, mu, Lc, B, args.seed + 1)
    mu_np, sd_np = mu.cpu().numpy(), np.sqrt(var.cpu().numpy())
    del Lc
    torch.cuda.empty_cache()
    log(f"analytic posterior: lam_max={lam_max:.4g} ({time.perf_counter() - t_setup:.1f}s)")

    if args.prior == "laplace":
        # Laplace prior (dispersion 0.3): no closed-form posterior; both legs are
        # compared with each other (the plain FP64 leg is the reference run) and
        # start from the Gaussian-prior posterior, which the warm-up moves off
        prior = D.Laplace(np.zeros((d, 1)), np.full((d, 1), 0.3))
        mu_np = sd_np = None
    else:
        prior = D.Normal(np.zeros((d, 1)), 1.0)
    post = D.BayesRule([prior, D.LinearMatrix(Gf.astype(np.float64), dd, sigma * sigma, dtype="bfloat16")])
    # comparator with the measured leg's trajectory form; with the
    # premultiplied form also the residual-form FP64 comparator (--plain-residual)
    plain = (_plain_fp64_leg(args, torch, post, X0, lam_max, mu_np, sd_np, args.premult_trajectory)
             if args.plain_steps > 0 else None)
    plain_res = (_plain_fp64_leg(args, torch, post, X0, lam_max, mu_np, sd_np, 0)
                 if args.plain_steps > 0 and args.premult_trajectory and args.plain_residual else None)
    if plain is not None and mu_np is None:
        mu_np, sd_np = plain["_mean"], plain["_std"]

    opts = {}
    if args.precision == "bf16":
        opts = {"bf16_split": args.bf16_split, "gemm_config": args.gemm_config}
        if args.bf16_split_grad:
            opts["bf16_split_grad"] = args.bf16_split_grad
        if args.gemm_config_grad is not None:
            opts["gemm_config_grad"] = args.gemm_config_grad
        if args.incremental_residual:
            opts["incremental_residual"] = args.incremental_residual
        if args.tile_order:
            opts["tile_order"] = args.tile_order
        if not args.split_tail:
            opts["gemm_split_tail"] = 0
    opts["premult_trajectory"] = args.premult_trajectory
    prob = build_problem(post, precision=args.precision, engine="gemm", options=opts)
    ch = Chains(prob, B)
    ch.set_state(X0)
    # bf16 trajectories: the anchor of the position-only gradient field is the
    # mean of the starting draws, fixed for the whole run (warm-up included)
    ch.set_anchor()
    log(f"device problem ready ({time.perf_counter() - t_setup:.1f}s)")
    eps, done = _tune(ch, B, L, args.eps_scale / np.sqrt(lam_max), args, args.seed)

    # surrogate: trained on warm-up pairs, then attached for the timed phase only
    # if its held-out gradient error and a warm-up probe pass (decided here,
    # before timing, so the timed chain is one fixed Markov kernel)
    sur, sur_info = None, None
    if args.surrogate:
        sur, sur_info = _train_surrogate(args, torch, ch, post, B, L, eps, done)
        done = sur_info["proposals_done"]
        decision = "rejected: held-out gradient error %.3f > %.3g" % (sur_info["heldout_grad_rel_err"],
                                                                      args.sur_max_grad_err)
        if sur_info["heldout_grad_rel_err"] <= args.sur_max_grad_err:
            ch.set_surrogate(sur)
            _, a = ch.run(args.sur_probe, L, np.full(B, eps), seed=args.seed + 5, first_proposal=done)
            done += args.sur_probe
            rate = float(a.sum()) / (args.sur_probe * B)
            sur_info["probe_acceptance"] = rate
            if rate < args.sur_fallback:
                ch.set_surrogate(None)
                decision = f"rejected: warm-up probe acceptance {rate:.3f} < {args.sur_fallback}"
            else:
                decision = f"attached: warm-up probe acceptance {rate:.3f}"
        sur_info["decision"] = decision
        log(f"surrogate {decision}")
    attached = sur is not None and sur_info["decision"].startswith("attached")

    # timed sampling phase: chunks of proposals; an attached surrogate is
    # detached when a chunk's acceptance falls below the fallback threshold
    E = max(args.steps, args.ess_draws)
    ch.set_trace(E)
    ch.enable_timing(True, every=args.timing_every)
    ch.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc = np.zeros(B, dtype=np.int64)
    fallback_at, sur_acc, sur_props, k = None, [], 0, 0
    while k < args.steps:
        m = min(args.chunk, args.steps - k)
        driving = attached and fallback_at is None
        _, a = ch.run(m, L, np.full(B, eps), seed=args.seed, first_proposal=done + k)
        acc += a
        k += m
        if driving:
            sur_props += m
            rate = float(a.sum()) / (m * B)
            sur_acc.append(rate)
            if rate < args.sur_fallback:
                ch.set_surrogate(None)
                fallback_at = k
    ch.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kt = _kernel_times(ch)
    ch.enable_timing(False)
    if E > args.steps:   # untimed continuation of the same chains for the ESS estimate
        ch.run(E - args.steps, L, np.full(B, eps), seed=args.seed, first_proposal=done + args.steps)
    if fallback_at is not None:
        log(f"surrogate fallback after {fallback_at} timed proposals (chunk acceptance {sur_acc[-1]:.3f} < "
            f"{args.sur_fallback})")
    names = {0: "residual gemm", 1: "endpoint", 2: "elementwise", 3: "gradient gemm", 4: "premultiplied stage gemm",
             5: "surrogate gemms", 6: "surrogate kick/drift", 7: "premultiplied stage split-tail epilogue"}
    log(f"timed {args.steps} proposals in {elapsed:.3f}s; " +
        ", ".join(f"{names[k_]} {kt[k_][0]:.1f} ms/{kt[k_][1]}" for k_ in names if kt[k_][1]))
    if args.premult_trajectory:
        slot, gemm_key = 4, "premult_gemm"
        gemm_name = ("k_gemm_bf16<256,128,64,64,1,3,EpiLeap<float>> alone (premultiplied stage GEMM: "
                     "g = y_ref + bf16(H) bf16(x - x_ref) with the kick/drift epilogue; its split-tail epilogue "
                     "kernel is timed apart, in 'stage')")
    elif kt[3][0] >= kt[0][0]:
        slot, gemm_key = 3, "gradient"
        gemm_name = ("k_gemm_bf16<...,EpiLeap> + k_split_epilogue (gradient GEMM + kick/drift; "
                     "the timed region holds both kernels)")
    else:
        slot, gemm_key = 0, "residual"
        gemm_name = ("k_gemm_bf16<...,EpiResidAnchor> + k_split_epilogue (residual GEMM; the timed region "
                     "holds both kernels)")
    gemm_ms, gemm_n = kt[slot]
    min_ess, err_mean, err_std, _, _, min_ess_timed = _posterior_errors(torch, ch, E, mu_np, sd_np, d, args.steps)
    acc_rate = float(acc.mean() / args.steps)
    value = min_ess / E * args.steps / elapsed
    # the gradient field the timed trajectories integrated, at the chains' final positions
    Xf, _ = ch.get_state()
    grad_err, grad_err_rms = (_trajectory_gradient_error(post, ch, Xf, args.precision)
                              if not (attached and fallback_at is None) else (None, None))

    # algorithmic FLOP of one stage GEMM (K = n residual form, K = d premultiplied)
    flop_launch = 2.0 * d * d * B if args.premult_trajectory else 2.0 * n * d * B
    avg_s = gemm_ms / 1e3 / max(gemm_n, 1)
    achieved = flop_launch / avg_s if gemm_n else 0.0
    peak = PEAK["bf16"] if args.precision == "bf16" else PEAK[args.precision]
    # HBM bytes per launch of the dominant kernel from the committed counter
    # profile (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel); only for
    # the configuration and GEMM options it was measured with
    traffic = stage_traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_dense_gemm.json")
    if os.path.exists(tfile) and args.precision == "bf16" and n == 20000 and d == 8192 and B == 512:
        try:
            tj = json.load(open(tfile))
            ran = {"gemm_config": args.gemm_config, "gemm_config_grad": args.gemm_config_grad,
                   "bf16_split": args.bf16_split, "bf16_split_grad": args.bf16_split_grad,
                   "incremental_residual": args.incremental_residual, "tile_order": args.tile_order,
                   "split_tail": args.split_tail, "prior": args.prior,
                   "premult_trajectory": args.premult_trajectory, "gradient_field": "anchored"}
            for prof in tj.get("profiles", []):
                if all(prof.get("options", {}).get(k_, object()) == v for k_, v in ran.items()):
                    traffic = prof["bytes_per_launch"].get(gemm_key)
                    stage_traffic = prof["bytes_per_launch"].get("premult_stage")
        except Exception:
            traffic = stage_traffic = None
    # the whole premultiplied stage: GEMM kernel + split-tail epilogue kernel
    stage = None
    if args.premult_trajectory and args.precision == "bf16" and gemm_n:
        st_s = (kt[4][0] + kt[7][0]) / 1e3 / gemm_n
        stage = {"kernels": "k_gemm_bf16<EpiLeap<float>> + k_split_epilogue<EpiLeap<float>>",
                 "avg_us": st_s * 1e6, "epilogue_launches": kt[7][1],
                 "achieved": flop_launch / st_s / 1e12, "frac": flop_launch / st_s / peak,
                 "traffic": stage_traffic,
                 "algorithmic_bytes": 2.0 * d * d + 2 * 2.0 * d * B + 4 * 4.0 * d * B}

    cfg = {
        "prior": "Laplace(0, 0.3); errors are against the plain fp64 leg" if args.prior == "laplace"
                 else "Normal(0, 1); errors are against the analytic fp64 posterior",
        "workload": f"dense linear {n}x{d} ({args.prior} prior), G stored bf16, {B} chains, L={L}: surrogate+batched "
                    f"({args.precision} trajectory GEMMs, fp64 endpoint Metropolis, MLP surrogate decided in "
                    f"warm-up) vs plain batched fp64 GPU HMC; both legs "
                    + ("premultiplied (H = G^T C^-1 G, the reference's form for n > d)" if args.premult_trajectory
                       else "in the residual form"),
        "chains": B, "n_data": n, "n_params": d, "leapfrog_steps": L, "stepsize": eps,
        "min_ess": min_ess, "ess_draws_per_chain": E, "min_ess_per_proposal": min_ess / E,
        "min_ess_timed_draws": min_ess_timed,
        "value_from_timed_draws_only": min_ess_timed / elapsed, "acceptance_rate": acc_rate,
        "posterior_rms_mean_err_over_sd": err_mean, "posterior_rms_std_rel_err": err_std,
        "proposals_per_s": args.steps / elapsed,
        "surrogate_fraction_of_timed_proposals": sur_props / args.steps,
        "trajectory_gradient_rel_err_vs_fp64": grad_err,
        "trajectory_gradient_rel_err_vs_fp64_rms": grad_err_rms,
        "trajectory_gradient_field": ("anchored: y_ref + bf16(H) bf16(x - x_ref), x_ref = mean of the starting draws"
                                      if args.premult_trajectory and args.precision == "bf16" else
                                      "exact" if args.precision == "f64" else
                                      "residual form (" + ("anchored" if args.incremental_residual else "hi/lo x planes")
                                      + ")"),
        "gemm": {"config": args.gemm_config, "config_grad": args.gemm_config_grad,
                 "bf16_split": args.bf16_split, "bf16_split_grad": args.bf16_split_grad,
                 "incremental_residual": args.incremental_residual, "tile_order": args.tile_order,
                 "split_tail": args.split_tail, "premult_trajectory": args.premult_trajectory},
        "kernel_avg_us": {names[k_]: kt[k_][0] / kt[k_][1] * 1e3 for k_ in names if kt[k_][1]},
    }
    if sur_info is not None:
        sur_info.update(timed_chunk_acceptance=sur_acc, timed_fallback_after_proposals=fallback_at,
                        fallback_threshold=args.sur_fallback, max_heldout_grad_err=args.sur_max_grad_err)
        if sur_props:
            sur_info["value_including_training_s"] = min_ess / E * args.steps / (
                elapsed + sur_info["train_s"] + sur_info["warmup_and_pairs_s"])
        cfg["surrogate"] = sur_info
    if plain is not None:
        plain.pop("_mean", None)
        plain.pop("_std", None)
        cfg["plain_fp64"] = plain
        cfg["speedup_vs_plain_fp64"] = value / plain["value"]
        cfg["ms_per_proposal_ratio_vs_plain_fp64"] = plain["ms_per_proposal"] / (elapsed / args.steps * 1e3)
    if plain_res is not None:
        plain_res.pop("_mean", None)
        plain_res.pop("_std", None)
        cfg["plain_fp64_residual_form"] = plain_res
        cfg["speedup_vs_plain_fp64_residual_form"] = value / plain_res["value"]
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
        "config": cfg,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": gemm_name, "avg_launch_us": avg_s * 1e6, "launches": gemm_n,
                     "algorithmic_flop_per_launch": flop_launch,
                     "event_sample": f"HIP events around the launches of every {args.timing_every}th timed "
                                     f"proposal ({gemm_n} stage launches); an event pair costs the stream a few us",
                     "endpoint_fp64_avg_us": kt[1][0] / max(kt[1][1], 1) * 1e3, "stage": stage},
    }

    if args.cpu_baseline:
        ess_per_prop_chain = min_ess / (E * B)
        rates, secs = cpu_baseline(args, eps)
        agg = float(np.sum(rates))
        result["cpu_baseline"] = {
            "value": ess_per_prop_chain * agg, "unit": "ESS/s", "cores": len(rates), "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"{len(rates)} single-chain oracle processes at once (one core each, BLAS pinned to 1 "
                      f"thread, float64 G, residual form, L={L}), {args.cpu_proposals} proposals each in "
                      f"{min(secs):.1f}-{max(secs):.1f}s: {agg:.4g} proposals/s in total "
                      f"({np.mean(rates):.4g} per process), times the GPU's measured min-ESS per proposal "
                      f"per chain {ess_per_prop_chain:.4g}",
        }
    ch.close()
    prob.close()
    if sur is not None:
        sur.close()
    return result


def _tune(ch, B, L, eps, args, seed, log_prefix=""):
    """Warm-up: multiplicative step tuning towards acceptance 0.65, then W
    untimed proposals.  Returns (eps, proposals done)."""
    done = 0
    for _ in range(args.tune_rounds):
        _, acc_t = ch.run(args.tune_props, L, np.full(B, eps), seed=seed, first_proposal=done)
        done += args.tune_props
        rate = acc_t.mean() / args.tune_props
        eps *= float(np.exp(1.5 * (rate - 0.65)))
        log(f"{log_prefix}tune: acceptance {rate:.3f} -> eps {eps:.4g}")
    if args.warmup:
        ch.run(args.warmup, L, np.full(B, eps), seed=seed, first_proposal=done)
    return eps, done + args.warmup


def _timed(ch, B, L, eps, steps, seed, first, torch=None, trace=None, every=1):
    ch.set_trace(steps if trace is None else trace)
    ch.enable_timing(True, every=every)
    ch.sync()
    if torch is not None:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, acc = ch.run(steps, L, np.full(B, eps), seed=seed, first_proposal=first)
    ch.sync()
    if torch is not None:
        torch.cuda.synchronize()
    return time.perf_counter() - t0, acc


def run_tomography(args):
    """Batched straight-ray tomography (64x64 cells, 2000 rays, CSR G) on the
    fused engine: one kernel per proposal per tile of chains."""
    import torch

    from hmclab_amd import Distributions as D
    from hmclab_amd import MassMatrices as MM
    from hmclab_amd import diagnostics, tomography
    from hmclab_amd.Distributions import build_problem
    from hmclab_amd.engine import Chains

    B, L = args.chains, args.steps_per_traj
    prob_d = tomography.synthetic_problem(seed=args.seed)
    G, dd = prob_d["G"], prob_d["d"]
    n, d = G.shape
    nnz = G.nnz
    sd2, sm2 = prob_d["noise_std"] ** 2, prob_d["prior_std"] ** 2
    Gt = torch.from_numpy(G.toarray()).to("cuda", torch.float64)
    H = Gt.T @ Gt / sd2
    H.diagonal().add_(1.0 / sm2)
    rhs = Gt.T @ torch.from_numpy(dd[:, 0]).to("cuda", torch.float64) / sd2 + prob_d["prior_mean"] / sm2
    Lc = torch.linalg.cholesky(H)
    mu = torch.cholesky_solve(rhs[:, None], Lc)[:, 0]
    Linv = torch.linalg.solve_triangular(Lc, torch.eye(d, device="cuda", dtype=torch.float64), upper=False)
    sd = torch.sqrt((Linv * Linv).sum(dim=0)).cpu().numpy()
    mass = H.diagonal().cpu().numpy().copy()          # Diagonal mass = diag of the posterior precision
    Ms = torch.from_numpy(1.0 / np.sqrt(mass)).to("cuda")
    lam_max = _power_lam_max(torch, Ms[:, None] * H * Ms[None, :])
    X0 = posterior_draws(torch, mu, Lc, B, args.seed + 1)
    mu_np = mu.cpu().numpy()
    del Gt, H, Linv, Lc
    torch.cuda.empty_cache()
    dtype = np.float32 if args.precision == "f32" else np.float64
    post = D.BayesRule([D.Normal(np.full((d, 1), prob_d["prior_mean"]), sm2),
                        D.LinearMatrix(G.astype(dtype), dd.astype(dtype) if dtype == np.float32 else dd,
                                       sd2, dtype=dtype)])
    prob = build_problem(post, precision=args.precision, mass_diagonal=mass,
                         options={"fused_tile_chains": args.fused_tile_chains} if args.fused_tile_chains else None)
    ch = Chains(prob, B)
    ch.set_state(X0)
    eps, done = _tune(ch, B, L, args.eps_scale / np.sqrt(lam_max), args, args.seed)
    elapsed, acc = _timed(ch, B, L, eps, args.steps, args.seed, done, torch)
    k_ms, k_n = ch.timing(0)
    draws = trace_to_torch(torch, ch, args.steps)
    ess = diagnostics.ess_bulk(draws) if args.steps >= 4 else np.full(d, np.nan)
    pooled = draws.to(torch.float64).reshape(-1, d)
    mean, std = pooled.mean(dim=0).cpu().numpy(), pooled.std(dim=0).cpu().numpy()
    del draws, pooled
    sx = 4 if args.precision == "f32" else 8
    sv = sx
    step_bytes = nnz * (sv + 4) + (n + 1) * 4 + 4 * d * B * sx
    launch_bytes = L * step_bytes + (nnz * (sv + 4) + d * B * sx) + d * B * sx
    avg_s = k_ms / 1e3 / max(k_n, 1)
    achieved = launch_bytes / avg_s
    # HBM bytes per launch from the committed counter profile (FETCH_SIZE +
    # WRITE_SIZE, calibrated on a copy kernel), for the configuration it was
    # measured with; the matrix streams from L2, so the profile's L2-to-L1
    # read bytes are reported beside it
    traffic = l2_bytes = None
    tfile = os.path.join(ROOT, "profiles", "traffic_tomography.json")
    if os.path.exists(tfile):
        try:
            for prof in json.load(open(tfile)).get("profiles", []):
                o = prof.get("options", {})
                if (o.get("precision") == args.precision and o.get("chains") == B and o.get("steps_per_traj") == L
                        and args.fused_tile_chains in o.get("fused_tile_chains", [])):
                    traffic = prof["bytes_per_launch"]["fused_propose"]
                    l2_bytes = prof.get("l2_to_l1_read_bytes_per_launch")
        except Exception:
            traffic = l2_bytes = None
    result = {
        "metric": METRIC, "value": float(np.min(ess)) / elapsed, "unit": "ESS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": done, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"straight-ray tomography 64x64 cells, {n} rays (nnz {nnz}), {B} chains, "
                               f"L={L}, Diagonal mass, fused sparse engine",
                   "min_ess": float(np.min(ess)), "acceptance_rate": float(acc.mean() / args.steps),
                   "stepsize": eps, "proposals_per_s": args.steps / elapsed,
                   "posterior_rms_mean_err_over_sd": float(np.sqrt(np.mean(((mean - mu_np) / sd) ** 2))),
                   "posterior_rms_std_rel_err": float(np.sqrt(np.mean((std / sd - 1.0) ** 2)))},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK["hbm"], "traffic": traffic,
                     "kernel": "k_fused_propose<sparse> (whole proposal)", "avg_launch_us": avg_s * 1e6,
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "l2_to_l1_read_bytes_per_launch": l2_bytes},
    }
    ch.close()
    prob.close()
    return result


def _jump_pilot(ps, args, L, d, step, done, K):
    """Per problem, the smallest over coordinates of the mean squared jump
    between consecutive draws over the draws' variance (1 - lag-1
    autocorrelation, times 2), from K proposals of 16 chains, and the time
    per proposal round."""
    import torch

    from hmclab_amd.problem_set import CHAINS

    ps.set_trace(K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ps.run(K, L, step, seed=args.seed + 17, first_proposal=done)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / K * 1e3
    X, _ = ps.read_trace(0, K)
    ps.set_trace(0)
    J = np.zeros(ps.P)
    for p in range(ps.P):
        x = X[:, ps.offsets[p]:ps.offsets[p + 1]].reshape(K, CHAINS, int(d[p]))
        jump = np.mean((x[1:] - x[:-1]) ** 2, axis=(0, 1))
        var = np.var(x.reshape(-1, int(d[p])), axis=0) + 1e-300
        J[p] = float(np.min(jump / var))
    return J, ms, done + K


def _adapt_uneven_integrators(ps, args, L, d, step, done, tune):
    """Warm-up adaptation of each problem's integrator (each problem is its own
    HMC.sample, so each may use its own, Samplers.py:1728-1738): the problems
    whose leapfrog draws move least (pilot jump statistic) switch to the
    4-stage integrator, whose step at the same acceptance is about twice as
    long, so their fixed L steps cover a longer trajectory.  Fractions 0..0.4
    of the problems are tried, each after re-tuning the steps, and the one
    with the largest predicted min over problems of ESS per second (ESS per
    draw ~ J / (4 - J) for lag-1 jump J, over the pilot's time per round) is
    kept.  Everything here is warm-up, before the timed phase."""
    from hmclab_amd.problem_set import CHAINS

    J0, ms0, done = _jump_pilot(ps, args, L, d, step, done, args.adapt_pilot)
    order = np.argsort(J0)
    lf_step = step.copy()
    best = None
    scores = {}
    for frac in (0.0, 0.1, 0.2, 0.3, 0.4):
        k = int(round(frac * ps.P))
        sel = order[:k]
        ps.set_integrator("lf")
        if k:
            ps.set_integrator("4s", problems=sel)
        mask = np.zeros(ps.P, bool)
        mask[sel] = True
        st = lf_step * np.repeat(np.where(mask, 2.0, 1.0), CHAINS)
        if k:
            st, done, _ = tune(st, args.adapt_tune_rounds, done)
            J, ms, done = _jump_pilot(ps, args, L, d, st, done, args.adapt_pilot)
        else:
            J, ms = J0, ms0
        score = float(np.min(J / (4.0 - np.minimum(J, 3.99)))) / (ms / 1e3)
        scores[f"{frac:.1f}"] = round(score, 2)
        if best is None or score > best[0]:
            best = (score, k, sel.copy(), st.copy())
    _, k, sel, st = best
    ps.set_integrator("lf")
    if k:
        ps.set_integrator("4s", problems=sel)
    return st, done, dict(problems_4s=int(k), scores=scores, rule="largest predicted min-ESS/s over fractions "
                          "0..0.4 of the least-moving problems on the 4-stage integrator",
                          L_of_4s_problems=sorted(int(x) for x in L[sel]), problems=sorted(int(x) for x in sel))


def run_uneven(args):
    """256 independent dense problems of uneven size (d log-uniform in
    [50, 2000], n = 1.5 d, G float32), 16 chains each, leapfrog steps per
    problem uniform in {5..30}; one hipGraph replay per proposal round."""
    import torch

    from hmclab_amd import diagnostics
    from hmclab_amd.problem_set import CHAINS, ProblemSet, curvature_and_start, uneven_problems

    t0 = time.perf_counter()
    d, n, L, specs = uneven_problems(seed=args.seed, problems=args.problems)
    ps = ProblemSet(d, n)
    starts, steps = [], []
    for p, sp in enumerate(specs):
        ps.set_problem(p, sp["G"], sp["d_obs"], sp["variance"], prior=sp["prior"], scale=sp["scale"])
        lam, x = curvature_and_start(sp)
        starts.append(np.repeat(x[None, :], CHAINS, axis=0))
        steps.append(np.full(CHAINS, args.eps_scale / np.sqrt(lam)))
    ps.set_state(ps.join(starts))
    step = np.concatenate(steps)
    log(f"{args.problems} problems (sum n*d = {int(np.sum(n * d)):,}) ready in {time.perf_counter() - t0:.1f}s")
    # tune each problem's step towards acceptance 0.65 (multiplicative, per problem)
    done = 0

    def tune(step, rounds, done):
        for r in range(rounds):
            acc = ps.run(args.tune_props, L, step, seed=args.seed, first_proposal=done)
            done += args.tune_props
            rate = acc.reshape(-1, CHAINS).sum(axis=1) / (CHAINS * args.tune_props)
            gain = 1.5 / (1.0 + r / 5.0)           # shrinking steps average out the acceptance noise
            step = step * np.repeat(np.exp(gain * (rate - 0.65)), CHAINS)
        return step, done, rate

    step, done, rate = tune(step, args.tune_rounds, done)
    log(f"tuned: acceptance per problem {rate.min():.2f}..{rate.max():.2f}; tiles {ps.tile_config()}")
    ps.run(args.warmup, L, step, seed=args.seed, first_proposal=done)
    done += args.warmup
    adapt = None
    if args.adapt_integrators:
        step, done, adapt = _adapt_uneven_integrators(ps, args, L, d, step, done, tune)
        log(f"adapted: {adapt['problems_4s']} problems on the 4-stage integrator (scores {adapt['scores']})")
    # kernel times from one directly launched round with events
    ps.enable_timing(True)
    ps.run(2, L, step, seed=args.seed, first_proposal=done, graph=False)
    kt = {k: ps.timing(k) for k in range(4)}
    ps.enable_timing(False)
    done += 2
    ps.set_trace(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    acc = ps.run(args.steps, L, step, seed=args.seed, first_proposal=done, graph=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    X, _ = ps.read_trace(0, args.steps)
    ess_min = []
    for p in range(args.problems):
        a, b = ps.offsets[p], ps.offsets[p + 1]
        draws = X[:, a:b].reshape(args.steps, CHAINS, int(d[p])).transpose(1, 0, 2)
        ess_min.append(float(np.min(diagnostics.ess_bulk(draws))) if args.steps >= 4 else float("nan"))
    ess_min = np.array(ess_min)
    rates = acc.reshape(-1, CHAINS).sum(axis=1) / (CHAINS * args.steps)
    # gradient stages of a round per problem: L, or 4L on the 4-stage integrator
    S = L.astype(np.int64).copy()
    if adapt:
        S[adapt["problems"]] *= 4
    r_ms, r_n = kt[0]
    g_ms, g_n = kt[1]
    e_ms, e_n = kt[2]
    c_ms, c_n = kt[3]
    step_ms = (r_ms + g_ms) / 2.0                      # resid + grad launches of one round
    tiles = ps.tile_config()
    if tiles["premult"]:
        # one pass over H (d x d fp32) per step; x and p read and written
        traj_bytes = float(np.sum(S * (4.0 * d * d + 256.0 * d)))
        actual = traj_bytes / (step_ms / 1e3)
        kernel = ((f"k_set_resid_lds<{tiles['resid_rows'] // 64},1,{tiles['lds_chunk']}> (rows of H)"
                   if tiles["premult_rows"] else f"k_set_grad<{tiles['grad_waves']},2> (column strips of H)")
                  + ": g = H x - b, kick/drift; all trajectory launches of one round")
    else:
        # the algorithm needs one pass of G per step; the residual and gradient launches make two
        traj_bytes = float(np.sum(S * (4.0 * n * d + 256.0 * d)))
        actual = float(np.sum(S * (8.0 * n * d + 256.0 * d + 2 * 64.0 * n))) / (step_ms / 1e3)
        kernel = "k_set_resid + k_set_grad (all trajectory launches of one round)"
    achieved = traj_bytes / (step_ms / 1e3)
    result = {
        "metric": METRIC, "value": float(ess_min.min()) / elapsed, "unit": "ESS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": done, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.problems} uneven dense problems (d in [50,2000] log-uniform, n=1.5d, "
                               f"G fp32), {CHAINS} chains each, L in [5,30] (leapfrog; the 4-stage integrator "
                               f"for the least-moving problems when the warm-up adaptation picks it); "
                               f"fp32 MFMA trajectories "
                               f"({'premultiplied H = G^T C^-1 G' if tiles['premult'] else 'residual form'}), "
                               f"fp64 residual-form endpoint; one hipGraph replay per proposal round",
                   "problems": args.problems, "chains": args.problems * CHAINS,
                   "sum_nd": int(np.sum(n * d)), "value_is": "min over problems of min-ESS, per second",
                   "median_problem_min_ess_per_s": float(np.median(ess_min)) / elapsed,
                   "integrator_adaptation": adapt,
                   "worst_problems": [{"problem": int(q), "d": int(d[q]), "n": int(n[q]), "L": int(L[q]),
                                       "prior": specs[q]["prior"], "acceptance": float(rates[q]),
                                       "min_ess_per_s": float(ess_min[q]) / elapsed}
                                      for q in np.argsort(ess_min)[:6]],
                   "acceptance_min": float(rates.min()), "acceptance_median": float(np.median(rates)),
                   "rounds_per_s": args.steps / elapsed, "tile_config": tiles,
                   "kernel_ms_per_round": {"resid": r_ms / 2, "grad": g_ms / 2, "endpoint": e_ms / 2,
                                           "per_chain": c_ms / 2}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                     "frac": achieved / PEAK["hbm"], "traffic": None,
                     "kernel": kernel,
                     "algorithmic_bytes_per_round": traj_bytes, "actual_stream_GBps": actual / 1e9,
                     "avg_launch_us": {"resid": r_ms / max(r_n, 1) * 1e3, "grad": g_ms / max(g_n, 1) * 1e3}},
    }
    ps.close()
    return result


def run_copy(args):
    """Calibration for HBM byte counters: a plain device copy of a known size
    (torch's copy kernel, 1 GiB read + 1 GiB written per launch)."""
    import torch

    nbytes = 1 << 30
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"metric": "device copy bandwidth (counter calibration)", "value": 2 * nbytes / dt / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 2, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "copy 1 GiB -> 1 GiB", "bytes_read_per_launch": nbytes,
                                            "bytes_written_per_launch": nbytes}}


def run_small(args):
    """The 1-chain dense case (G 100x50): latency per proposal on the fused
    engine at the stated step 0.05 and ESS/s at a tuned step."""
    import torch

    from hmclab_amd import Distributions as D
    from hmclab_amd import diagnostics
    from hmclab_amd.Distributions import build_problem
    from hmclab_amd.engine import Chains
    from oracle import hmc_oracle as O

    rng = np.random.default_rng(args.seed)
    G = rng.normal(size=(100, 50))
    dd = G @ rng.normal(size=(50, 1)) + rng.normal(0, 0.1, size=(100, 1))
    post = D.BayesRule([D.Normal(np.zeros((50, 1)), 1.0), D.LinearMatrix(G, dd, 0.01, dtype=np.float64)])
    mu, Sigma = O.analytic_linear_gaussian(G, dd, 0.01, np.zeros(50), 1.0)
    prob = build_problem(post, precision="f64")
    ch = Chains(prob, 1)
    ch.set_state(mu.reshape(1, 50))
    L = 10
    # stated step 0.05: latency only (it exceeds the stability limit 2/sqrt(lam_max));
    # one untimed proposal first builds what the first call builds (H when premultiplied)
    ch.run(1, L, np.full(1, 0.05), seed=args.seed + 1)
    ch.sync()
    ch.enable_timing(True)
    t0 = time.perf_counter()
    ch.run(args.steps, L, np.full(1, 0.05), seed=args.seed)
    ch.sync()
    lat_stated = (time.perf_counter() - t0) / args.steps
    eps0 = args.eps_scale / np.sqrt(np.linalg.eigvalsh(np.linalg.inv(Sigma)).max())
    eps, done = _tune(ch, 1, L, eps0, args, args.seed)
    elapsed, acc = _timed(ch, 1, L, eps, args.steps, args.seed, done + args.steps, torch)
    k_ms, k_n = ch.timing(0)
    X, _ = ch.read_trace(0, args.steps)
    ess = diagnostics.ess_bulk(X.transpose(1, 0, 2))
    result = {
        "metric": METRIC, "value": float(np.min(ess)) / elapsed, "unit": "ESS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": done, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "dense linear-Gaussian 100x50, 1 chain, L=10, fused engine",
                   "us_per_proposal_tuned": elapsed / args.steps * 1e6,
                   "us_per_proposal_dt0.05": lat_stated * 1e6, "stepsize": eps,
                   "acceptance_rate": float(acc.mean() / args.steps), "min_ess": float(np.min(ess))},
        "roofline": {"bound": "hbm", "achieved": None, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                     "frac": None, "traffic": None, "kernel": "k_fused_propose<dense> (latency-bound)",
                     "avg_launch_us": k_ms / max(k_n, 1) * 1e3},
    }
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="dense", choices=["dense", "tomography", "small", "uneven", "copy"])
    ap.add_argument("--problems", type=int, default=256)
    ap.add_argument("--timing-every", type=int, default=5,
                    help="kernel timing events on every k-th timed proposal (dense workload)")
    ap.add_argument("--adapt-integrators", type=int, default=1, choices=[0, 1],
                    help="uneven batch: warm-up choice of the 4-stage integrator for the least-moving problems")
    ap.add_argument("--adapt-pilot", type=int, default=30, help="proposals per adaptation pilot")
    ap.add_argument("--adapt-tune-rounds", type=int, default=10)
    ap.add_argument("--prior", default="normal", choices=["normal", "laplace"],
                    help="dense workload prior: Normal(0,1) (analytic posterior) or Laplace(0,0.3)")
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=20000)
    ap.add_argument("--d", type=int, default=8192)
    ap.add_argument("--chains", type=int, default=None,
                    help="default: 512 (dense), 1024 (tomography)")
    ap.add_argument("--steps-per-traj", type=int, default=None,
                    help="default: 25 (dense), 20 (tomography)")
    ap.add_argument("--eps-scale", type=float, default=0.3)
    ap.add_argument("--tune-rounds", type=int, default=None, help="default: 6 (20 for the uneven batch)")
    ap.add_argument("--tune-props", type=int, default=None, help="default: 3 (5 for the uneven batch)")
    ap.add_argument("--cpu-baseline-child", action="store_true")
    ap.add_argument("--cpu-eps", type=float, default=1e-3)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "f64", "f32"])
    ap.add_argument("--bf16-split", type=int, default=2, choices=[1, 2])
    ap.add_argument("--bf16-split-grad", type=int, default=1, choices=[1, 2],
                    help="bf16 planes of the scaled residual in the gradient GEMM (1: one plane, "
                         "measured to leave acceptance unchanged; 2: hi/lo)")
    ap.add_argument("--gemm-config", type=int, default=3, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--fused-tile-chains", type=int, default=0, choices=[0, 1, 2, 4],
                    help="tomography: largest chains per fused-engine workgroup (0: as LDS allows)")
    ap.add_argument("--split-tail", type=int, default=1, choices=[0, 1],
                    help="bf16 GEMMs: split the last partial round of tiles along K")
    ap.add_argument("--tile-order", type=int, default=0, choices=[0, 1])
    ap.add_argument("--plain-residual", type=int, default=0, choices=[0, 1],
                    help="with premultiplied trajectories, also time the residual-form FP64 comparator")
    ap.add_argument("--ess-draws", type=int, default=200,
                    help="draws per chain for the ESS estimate: the timed proposals continued untimed")
    ap.add_argument("--premult-trajectory", type=int, default=1, choices=[0, 1],
                    help="trajectories of both dense legs through H = G^T C^-1 G (the reference's "
                         "premultiplication, its default when n > d); 0: residual form")
    ap.add_argument("--incremental-residual", type=int, default=1, choices=[0, 1],
                    help="bf16: advance a float32 residual by G dx after the first stage")
    ap.add_argument("--gemm-config-grad", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="tile configuration of the gradient GEMM (default: --gemm-config)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-proposals", type=int, default=6,
                    help="oracle proposals per CPU-baseline process (about 1.4 s each on the dense workload)")
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1),
                    help="concurrent single-core CPU-baseline processes (default: 16, or fewer cores)")
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed proposals of the plain fp64 comparator (default: --steps, so both "
                         "legs' ESS estimates see chains of equal length; 0 skips it)")
    ap.add_argument("--chunk", type=int, default=10, help="proposals per device run call")
    ap.add_argument("--no-surrogate", dest="surrogate", action="store_false")
    ap.add_argument("--sur-warmup", type=int, default=100, help="warm-up proposals giving training pairs")
    ap.add_argument("--sur-hidden", type=int, default=512)
    ap.add_argument("--sur-epochs", type=int, default=8)
    ap.add_argument("--sur-lr", type=float, default=1e-3)
    ap.add_argument("--sur-fallback", type=float, default=0.5)
    ap.add_argument("--sur-probe", type=int, default=4,
                    help="warm-up proposals with the surrogate attached that decide whether it drives the "
                         "timed phase")
    ap.add_argument("--sur-max-grad-err", type=float, default=0.5,
                    help="largest held-out gradient error (relative to the gradient's spread) at which the "
                         "surrogate is tried")
    args = ap.parse_args()
    if args.plain_steps is None:
        args.plain_steps = args.steps
    tomo = args.workload == "tomography"
    uneven = args.workload == "uneven"
    if args.tune_rounds is None:
        args.tune_rounds = 20 if uneven else 6
    if args.tune_props is None:
        args.tune_props = 5 if uneven else 3
    if args.chains is None:
        args.chains = 1024 if tomo else 512
    if args.steps_per_traj is None:
        args.steps_per_traj = 20 if tomo else 25
    if args.cpu_baseline_child:
        cpu_baseline_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); chains shard across GPUs "
              "without communication, so the one-GPU figure is the per-GPU rate.", file=sys.stderr)
        sys.exit(2)
    if tomo:
        if args.precision == "bf16":
            args.precision = "f32"
        res = run_tomography(args)
    elif args.workload == "small":
        res = run_small(args)
    elif args.workload == "uneven":
        res = run_uneven(args)
    elif args.workload == "copy":
        res = run_copy(args)
    else:
        res = run_dense(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

/*
 * hmc_engine.h — C-ABI of the hmclab_amd HIP engine (libhmc.so, gfx950).
 *
 * The reference (larsgeb/hmclab) is pure Python; its only native call site
 * is a ctypes binding of a CSR SpMV (hmclab/Helpers/InterfaceMKL.py:32-35,
 * 119-121, 187-189) that takes raw pointers, 0-based int32 CSR arrays and
 * caller-allocated outputs.  This header follows that precedent: plain
 * pointers and sizes, host arrays in float64 unless stated, int status codes,
 * no exceptions and no aborts across the boundary.
 *
 * Each entry point names the reference interface it replaces.  Arrays of
 * chains are row-major (B, d): chain c's coordinates are contiguous, i.e.
 * the reference's (d, 1) column vectors laid side by side as rows.
 */
#ifndef HMC_ENGINE_H
#define HMC_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HMC_OK 0
#define HMC_ERR_ARG 1       /* invalid argument (Python: ValueError)          */
#define HMC_ERR_STATE 2     /* call out of order (Python: AssertionError)     */
#define HMC_ERR_HIP 3       /* HIP runtime failure (Python: RuntimeError)     */
#define HMC_ERR_NOMEM 4     /* allocation failure (Python: MemoryError)       */
#define HMC_ERR_UNSUPPORTED 5 /* combination not built (NotImplementedError)  */
#define HMC_ERR_INJECTED 6  /* fault injected by HMC_FAULT_INJECT (tests)     */

/* element types */
#define HMC_F64 0
#define HMC_F32 1
#define HMC_BF16 2

/* prior term kinds */
#define HMC_PRIOR_NORMAL 1  /* a = means, b = inverse variances    */
#define HMC_PRIOR_LAPLACE 2 /* a = means, b = inverse dispersions  */

/* mass matrix kinds */
#define HMC_MASS_UNIT 0
#define HMC_MASS_DIAGONAL 1

/* integrators (Samplers.py:1728-1737) */
#define HMC_INT_LEAPFROG 0
#define HMC_INT_3STAGE 1
#define HMC_INT_4STAGE 2

/* engines */
#define HMC_ENGINE_AUTO 0
#define HMC_ENGINE_GEMM 1   /* MFMA GEMM per gradient over the whole batch     */
#define HMC_ENGINE_FUSED 2  /* one kernel per proposal per chain tile          */

typedef struct hmc_problem hmc_problem;
typedef struct hmc_chains hmc_chains;

/* Library version (major*10000 + minor*100 + patch). */
int hmc_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* hmc_last_error(void);

/* Number of visible HIP devices (0 when no GPU is present). */
int hmc_device_count(int* count);

/* ---------------------------------------------------------------------------
 * Problem = one BayesRule target (hmclab/Distributions/base.py:1005-1205):
 * an optional linear-Gaussian likelihood plus element-wise priors, collapsed
 * bounds and the mass matrix.  Host inputs are copied to the device at set
 * time, as the reference copies/casts at construction (LinearMatrix.py:152-159).
 * ------------------------------------------------------------------------- */

/* Create an empty target of `dimensions` parameters.  `precision` is the
 * arithmetic type of trajectory gradients (HMC_F64, HMC_F32 or HMC_BF16);
 * the Metropolis Hamiltonian is always evaluated in float64.
 * Replaces: BayesRule.__init__ (base.py:1012-1034). */
int hmc_problem_create(int64_t dimensions, int precision, hmc_problem** out);

/* Dense forward model G (n_data x dimensions, row-major, element type
 * g_dtype), observed data d (n_data) and data variance (n_data values, or one
 * value when variance_is_scalar).  The target is defined by G as given in
 * g_dtype; trajectory arithmetic rounds it to `precision`.
 * Replaces: _LinearMatrix_dense_forward_simple_covariance.__init__
 * (LinearMatrix.py:139-182), residual form. */
int hmc_problem_set_dense(hmc_problem* p, const void* G, int g_dtype, int64_t n_data,
                          const double* d_obs, const double* data_variance,
                          int variance_is_scalar);

/* Sparse CSR forward model (0-based int32 indptr[n_data+1], indices[nnz],
 * values[nnz] of type dtype).  Replaces:
 * _LinearMatrix_sparse_forward_simple_covariance.__init__ (LinearMatrix.py:309-359)
 * and the CSR SpMV binding sparse_gemv (InterfaceMKL.py:38-64). */
int hmc_problem_set_csr(hmc_problem* p, const int32_t* indptr, const int32_t* indices,
                        const void* values, int dtype, int64_t n_data, int64_t nnz,
                        const double* d_obs, const double* data_variance,
                        int variance_is_scalar);

/* Add one element-wise prior term (kind HMC_PRIOR_*), per-parameter arrays
 * a, b of length `dimensions`; `constant` is its normalization constant.
 * Replaces: Normal.misfit/gradient (base.py:539-574, diagonal case) and
 * Laplace.misfit/gradient (base.py:689-710). */
int hmc_problem_add_prior(hmc_problem* p, int kind, const double* a, const double* b,
                          double constant);

/* Collapsed bounds (either may be NULL).  Replaces: update_bounds /
 * collapse_bounds / corrector / misfit_bounds (base.py:239-374, 1061-1142). */
int hmc_problem_set_bounds(hmc_problem* p, const double* lower, const double* upper);

/* Mass matrix (HMC_MASS_UNIT with NULL, or HMC_MASS_DIAGONAL with the
 * diagonal).  Replaces: Unit / Diagonal (MassMatrices.py:82-238). */
int hmc_problem_set_mass(hmc_problem* p, int kind, const double* diagonal);

/* Choose the trajectory engine (HMC_ENGINE_*). */
int hmc_problem_set_engine(hmc_problem* p, int engine);

/* Engine tuning knobs (no counterpart in the reference):
 *   "bf16_split"  1 or 2: bf16 planes of the FP32 operand of both trajectory
 *                 GEMMs (2 = hi/lo split, ~16-bit operand precision; default 2)
 *   "bf16_split_grad" 1 or 2: the same for the gradient GEMM only (its operand
 *                 is the scaled residual)
 *   "incremental_residual" 0 or 1: bf16 residual-form trajectories take
 *                 r = (G x_ref - d) + G_hi bf16(x - x_ref) with a float32
 *                 residual at the chain set's anchor x_ref (one bf16 plane of
 *                 x - x_ref per stage instead of hi/lo planes of x: half the
 *                 MFMA work of the residual GEMM; default 0)
 *   "gemm_config" 0-4: bf16 GEMM tile of both trajectory GEMMs
 *                 (0 = 128x128, 4 waves of 64x64; 1 = 128x128, 8 waves of
 *                 64x32; 2 = 256x128, 4 waves of 128x64; 3 = 256x128, 8 waves
 *                 of 64x64; 4 = 256x128, 8 waves of 128x32; 3 and 4 for
 *                 one-plane operands, config 1 for two; default 1)
 *   "gemm_config_grad" 0-4: the same for the gradient GEMM only
 *   "split_fixup" 0 or 1: tiles of the premultiplied bf16 stage cut into two
 *                 k-ranges sum their float32 partials inside the GEMM kernel
 *                 (the second workgroup of a tile adds the first one's partial
 *                 and applies the kick/drift epilogue; same results as the
 *                 split epilogue kernel, one launch per stage; default 0: at
 *                 d = 8192, 512 chains it measured 145 us per stage against
 *                 105 us for GEMM + split epilogue kernel)
 *   "gemm_split_tail" 0 or 1: when the tiles of a bf16 GEMM leave a last
 *                 partial round on the compute units, split those tiles along
 *                 K into float32 partials summed in fixed order (default 1)
 *   "endpoint_premult" 0 or 1: the dense GEMM engine's FP64 endpoint misfit
 *                 through H = G^T S^-1 G and b = G^T S^-1 d, built once in FP64
 *                 (chi = 0.5 (x^T H x - 2 b^T x + d^T S^-1 d), a GEMM with K = d
 *                 instead of n); default: when n > d, the reference's
 *                 premultiplication rule (LinearMatrix.py:164-167), and H
 *                 takes at most 16 GiB
 *   "endpoint_symmetric" 0 or 1: the premultiplied endpoint over the upper
 *                 block triangle of the symmetric H (x_a^T H_aa x_a + 2 x_a^T
 *                 H_ac x_c, c > a: half the FP64 work; default 1) or over all
 *                 of H (0)
 *   "premult_trajectory" -1, 0 or 1: trajectory gradients of dense targets
 *                 through H as well (LinearMatrix.py:203-207, g = H x - b):
 *                 the GEMM engine runs one GEMM with K = d per stage, the
 *                 fused engine one pass over H per stage (d <= 1024).  FP64: g from
 *                 H directly; bf16: g = (H x_ref - b) + H_hi bf16(x - x_ref)
 *                 with the anchor's FP64 product kept in float32 and one bf16
 *                 plane of H and of x - x_ref.  1 needs the
 *                 premultiplied endpoint (chains creation fails otherwise);
 *                 -1 (default) follows the endpoint (the reference's n > d
 *                 rule); 0 keeps the residual form
 *   "fused_tile_chains" 0, 1, 2 or 4: chains one workgroup of the fused
 *                 engine carries (0 = automatic: float32 sparse operators
 *                 with at least 4 chains per compute unit 4, else 2 if the
 *                 LDS tile fits, else 4 (float32) or 1; a value that does
 *                 not fit the LDS tile or exceeds the chains is an error at
 *                 the first call)
 *   "gemm_split_cus" n: compute units the split-tail plan assumes (0 = the
 *                 device's count, the default; a smaller n plans for a share
 *                 of the device)
 *   "tile_order"  0 or 1: bf16 GEMM tiles dealt to the 8 XCDs in m-major
 *                 (tiles sharing G rows share an L2; default) or n-major
 *                 (tiles sharing chains share an L2) ranges; for tiles split
 *                 along K into 2, 4 or 8 k-ranges, 1 gives each XCD a single
 *                 k-range (its L2 holds that k-range of the chain operand)
 * Must be set before hmc_chains_create. */
int hmc_problem_set_option(hmc_problem* p, const char* name, double value);

int hmc_problem_destroy(hmc_problem* p);

/* Batched misfit chi(m) for B models M (B x dimensions) -> out[B].
 * Replaces: BayesRule.misfit (base.py:1036-1043). */
int hmc_misfit_batch(hmc_problem* p, const double* M, int64_t B, double* out);

/* hmc_misfit_batch for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  M_dev (B x dimensions,
 * float64, row-major) and out_dev (B) are DEVICE pointers owned by the
 * caller; every other buffer lives in a static 32 MiB device array per GPU,
 * so the call makes no hipMalloc/hipFree.  Fails with HMC_ERR_ARG when the
 * batch needs more than 32 MiB (the models in the padded state layout plus
 * the endpoint partial sums: use fewer models per call), and with
 * HMC_ERR_STATE when the problem's premultiplied operator H (dense, n > d)
 * has not been built yet by an earlier misfit evaluation.  Runs on the
 * problem's stream and returns after it finished.  Calls that overlap on one
 * GPU share the array and race. */
int hmc_misfit_batch_noalloc(hmc_problem* p, const double* M_dev, int64_t B, double* out_dev);

/* Batched gradient for B models -> G_out (B x dimensions).
 * Replaces: BayesRule.gradient (base.py:1045-1054). */
int hmc_gradient_batch(hmc_problem* p, const double* M, int64_t B, double* out);

/* ---------------------------------------------------------------------------
 * Chains = B independent Markov chains on one problem, state resident on the
 * device.  Replaces the per-chain HMC sampler state of Samplers.HMC
 * (Samplers.py:1105-1178) and ParallelSampleSMP's one-process-per-chain
 * layout (Samplers.py:1816-1977) by columns of one batch.
 * ------------------------------------------------------------------------- */
int hmc_chains_create(hmc_problem* p, int64_t B, hmc_chains** out);
int hmc_chains_destroy(hmc_chains* c);

/* Set positions (B x dimensions) and evaluate their misfits.  Fails with
 * HMC_ERR_ARG when a misfit is NaN or inf (Samplers.py:419-426). */
int hmc_chains_set_state(hmc_chains* c, const double* X);

/* Copy positions (B x dimensions, may be NULL) and misfits (B, may be NULL). */
int hmc_chains_get_state(hmc_chains* c, double* X, double* chi);

/* One proposal for every chain with host-supplied randoms (parity mode):
 * momenta P0 (B x dimensions, already scaled by sqrt(M)), local step sizes
 * eps[B] (stepsize * uniform(0.5,1.5) as the reference draws them), accept
 * uniforms u[B].  Outputs (any may be NULL): accepted[B], H0[B], H1[B] and
 * the trajectory end point X1/P1 (B x dimensions).
 * Replaces: HMC._propose + _evaluate_acceptance (Samplers.py:1463-1492) with
 * the integrator of Samplers.py:1524-1726. */
int hmc_chains_propose(hmc_chains* c, const double* P0, const double* eps, const double* u,
                       int amount_of_steps, int integrator, int32_t* accepted, double* H0,
                       double* H1, double* X1, double* P1);

/* Anchor of the bf16 trajectories (premultiplied, or residual form with
 * "incremental_residual"): the stage gradient is
 *   g(x) = g(x_ref) + A_hi bf16(x - x_ref) + prior'(x)
 * (A = H, or G and G^T through the residual) with g(x_ref) computed in FP64,
 * a function of the position alone, so trajectories stay reversible and
 * volume preserving.  x_ref (dimensions values) or NULL: the mean of the
 * chains' current positions.  The anchor stays until it is set again; a
 * chain set without one takes the mean at its first trajectory.  Set it
 * before a sampling phase, not between its proposals (a kernel that changes
 * with the chain's own state is not a Markov kernel of the target).  A no-op
 * for other precisions and engines.  No counterpart in the reference (its
 * trajectories are FP64). */
int hmc_chains_set_anchor(hmc_chains* c, const double* x_ref);

/* The gradient field the trajectories of this chain set integrate, at B
 * models X (B x dimensions) -> out (B x dimensions): the anchored bf16 field,
 * the premultiplied FP64 field H x - b + prior', or the residual form.  For
 * measuring the trajectory's gradient error against hmc_gradient_batch of
 * an FP64 problem.  Replaces: the gradient calls inside
 * _propagate_leapfrog (Samplers.py:1549-1569). */
int hmc_chains_trajectory_gradient(hmc_chains* c, const double* X, double* out);

/* Run `proposals` proposals entirely on the device with the Philox RNG
 * (seed, counter offset `first_proposal`).  Per-chain step sizes stepsize[B]
 * (host, in/out when autotuning).  randomize: multiply by uniform(0.5,1.5).
 * autotune != 0 applies the Robbins-Monro update of Samplers.py:1494-1522
 * with (target, learning_rate).  Every `thinning`-th proposal (index i with
 * i % thinning == 0) is recorded into the device trace (see below).
 * accepted_total[B] (may be NULL) receives accept counts of this call.
 * Replaces: _AbstractSampler._sample_loop (Samplers.py:529-704). */
int hmc_chains_run(hmc_chains* c, int64_t proposals, int amount_of_steps, int integrator,
                   double* stepsize, int randomize, int autotune, double target,
                   double learning_rate, uint64_t seed, int64_t first_proposal,
                   int64_t thinning, int64_t* accepted_total);

/* Autotuning history of hmc_chains_run calls with autotune != 0: for every
 * proposal (up to `capacity`, then HMC_ERR_STATE) and chain, the acceptance
 * value exp(H0 - H1) (may be inf or NaN) and the step size before its update.
 * Setting a capacity (0 frees) discards the recorded history.  read copies
 * proposals [first, first+count) as (count x B) arrays; either may be NULL.
 * Replaces: the acceptance_rates / stepsizes arrays of HMC.autotune
 * (Samplers.py:1494-1500) written by _close_sampler_specific (:1419-1426). */
int hmc_chains_set_tuning_history(hmc_chains* c, int64_t capacity);
int hmc_chains_tuning_history_count(hmc_chains* c, int64_t* count);
int hmc_chains_read_tuning_history(hmc_chains* c, int64_t first, int64_t count, double* rates, double* steps);

/* Stage coefficients of an integrator (host only, no device work), in units of
 * the local step: the first drift, then per stage (kick, drift, drift2) as
 * the engines apply them; L steps give L (leapfrog), 3L (3-stage) or 4L
 * (4-stage) stages; the closing drift of a multi-stage step and the opening
 * drift of the next share a stage (drift, drift2).  `count` receives the
 * number of stages; `stages` (3 * capacity values) may be NULL.
 * Replaces: the coefficient tables and stage order of _propagate_leapfrog,
 * _propagate_3_stage_simplified and _propagate_4_stage_simplified
 * (Samplers.py:1524-1726). */
int hmc_integrator_schedule(int integrator, int L, double* first_drift, double* stages, int64_t capacity,
                            int64_t* count);

/* Philox4x32-10 block of the device RNG (host evaluation of the same code):
 * out = philox(counter[4], key[2]).  No counterpart in the reference, whose
 * draws come from numpy's default_rng; for known-answer tests. */
int hmc_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out);

/* Device trace of recorded samples: capacity in slots; each slot holds
 * B x dimensions coordinates (stored in the trajectory precision) and B
 * misfits.  Setting a capacity discards recorded samples. */
int hmc_chains_set_trace(hmc_chains* c, int64_t capacity_slots, int64_t param_stride);
/* Number of recorded slots. */
int hmc_chains_trace_count(hmc_chains* c, int64_t* slots);
/* Copy slots [first, first+count) to host as float64: X (count x B x dims)
 * and chi (count x B); either may be NULL.  */
int hmc_chains_read_trace(hmc_chains* c, int64_t first, int64_t count, double* X, double* chi);
/* Device pointer and element type of the trace (for device-side analysis). */
int hmc_chains_trace_device(hmc_chains* c, void** x_ptr, double** chi_ptr, int* dtype);
int hmc_chains_reset_trace(hmc_chains* c);

/* Accumulated kernel times since the last reset, from HIP events recorded on
 * the engine's stream around each launch (in kinds 0-3 and 5 a launch and
 * its split-tail epilogue kernel count as one).  kind: 0 = residual GEMM
 * (fused engine: the whole-proposal kernel), 1 = endpoint misfit,
 * 2 = element-wise / Metropolis, 3 = gradient GEMM with the kick/drift
 * epilogue (residual form), 4 = premultiplied stage GEMM kernel with the
 * kick/drift epilogue (bf16: the GEMM kernel alone; FP64: the whole stage),
 * 5 = surrogate forward and input-gradient GEMMs, 6 = surrogate kick/drift,
 * 7 = the bf16 premultiplied stage's split-tail epilogue kernel.
 * total_ms and launches may be NULL.  Timing is off until
 * hmc_chains_enable_timing(c, k): k = 1 records every launch, k > 1 only the
 * launches of proposals whose index is a multiple of k in hmc_chains_run
 * (each event pair costs the stream a few microseconds), 0 turns it off. */
int hmc_chains_enable_timing(hmc_chains* c, int on);
int hmc_chains_timing(hmc_chains* c, int kind, double* total_ms, int64_t* launches);

/* Synchronise the engine's stream. */
int hmc_chains_sync(hmc_chains* c);

/* ---- tempering and replica exchange ----------------------------------------
 * Replaces ParallelSampleSMP with exchange (Samplers.py:1816-1977, exchange
 * step :589-669) for chains whose posteriors differ by a temperature T_c
 * (misfit / T_c, as the reference's tempered distributions do, e.g.
 * base.py:392-414): chain c samples exp(-chi(m) / T_c).  NULL resets all
 * chains to T = 1.  Misfits returned by hmc_chains_get_state and recorded in
 * the trace are the tempered ones, chi / T_c. */
int hmc_chains_set_temperatures(hmc_chains* c, const double* temperatures);
/* One exchange event with host uniforms: pairs[2k], pairs[2k+1] (the second
 * decides, as the reference's odd schedule position does), u[k] its uniform;
 * the pair swaps positions and misfits when
 * exp(chi_b/T_b - chi_a/T_b + chi_a/T_a - chi_b/T_a) > u[k].  swapped[k] out (may be NULL). */
int hmc_chains_exchange(hmc_chains* c, const int32_t* pairs, int64_t npairs, const double* u,
                        int64_t proposal, int32_t* swapped);
/* Exchange schedule for hmc_chains_run: after proposal t with t % interval
 * == 0 and t / interval < events, row t / interval of schedule (events x
 * 2*npairs chain indices) is applied with device uniforms.  NULL clears it. */
int hmc_chains_set_exchange(hmc_chains* c, int64_t interval, const int32_t* schedule, int64_t events,
                            int64_t npairs);
/* Accepted exchanges per chain since creation. */
int hmc_chains_exchange_counts(hmc_chains* c, int64_t* counts);

/* ---- learned surrogate potential ------------------------------------------
 * No counterpart in the reference: hmclab's HMC always calls the target's
 * gradient (Samplers.py:1524-1584, `self.target.gradient(position)` inside
 * _propagate_leapfrog).  A surrogate is an MLP
 *   f(x) = w3 . silu(W2 silu(W1 xn + b1) + b2) + b3,  xn = (x - x_mean)/x_scale
 * of the likelihood misfit, chi_lik(x) ~ f_mean + f_scale f(x).  Attached to
 * chains of a dense problem (GEMM engine), its input gradient replaces the
 * likelihood gradient inside the trajectory; priors stay exact and the
 * Metropolis test keeps the exact FP64 Hamiltonian.  Parameters are float32,
 * packed as W1 (hidden x d, row-major) | b1 | W2 (hidden x hidden) | b2 | w3 | b3. */
typedef struct hmc_surrogate hmc_surrogate;

/* Glorot-uniform weights from `seed`; hidden must be a multiple of 64. */
int hmc_surrogate_create(int64_t dimensions, int64_t hidden, uint64_t seed, hmc_surrogate** out);
int hmc_surrogate_destroy(hmc_surrogate* s);
/* 0: bf16 MFMA GEMMs with FP32 accumulation (default); 1: FP32 GEMMs. */
int hmc_surrogate_set_mode(hmc_surrogate* s, int mode);
int hmc_surrogate_set_normalization(hmc_surrogate* s, const double* x_mean, const double* x_scale,
                                    double f_mean, double f_scale);
int hmc_surrogate_param_count(hmc_surrogate* s, int64_t* n);
int hmc_surrogate_get_params(hmc_surrogate* s, float* params);
/* Replaces the parameters and resets the Adam moments. */
int hmc_surrogate_set_params(hmc_surrogate* s, const float* params);
/* Misfit approximation f_mean + f_scale f(x) [B] and its gradient [B x d]
 * for rows X [B x d] (either output may be NULL). */
int hmc_surrogate_eval(hmc_surrogate* s, const double* X, int64_t B, double* f, double* g);
/* Mean training loss over a batch of NORMALISED rows Xn [B x d] with targets
 * y [B] (of f) and t [B x d] (of df/dxn), and its packed parameter gradient
 * (param_count floats; NULL to skip); B a multiple of 64. */
int hmc_surrogate_loss_grad(hmc_surrogate* s, const double* Xn, const double* y, const double* t,
                            int64_t B, double alpha, double beta, double* loss, float* grad);
/* Adam training on N raw pairs (X [N x d], chi [N], grad [N x d]) of the
 * likelihood misfit: `epochs` passes over shuffled batches of `batch` rows
 * (a multiple of 64, <= N); loss_history[epochs] receives each epoch's mean
 * loss (may be NULL).  Normalisation must be set first. */
int hmc_surrogate_train(hmc_surrogate* s, const double* X, const double* chi, const double* grad,
                        int64_t N, int64_t batch, int64_t epochs, double lr, double alpha,
                        double beta, uint64_t seed, double* loss_history);
/* Attach (or with NULL detach) a surrogate to chains of a dense problem. */
int hmc_chains_set_surrogate(hmc_chains* c, hmc_surrogate* s);

/* ---- sets of independent problems -----------------------------------------
 * Replaces running one HMC.sample per problem and per chain in separate
 * processes (ParallelSampleSMP, Samplers.py:1757-1977, around the
 * integrators :1524-1726 and HMC._evaluate_acceptance :1471-1492): P dense
 * linear problems of uneven size (G float32, n_p x d_p), each with one Normal
 * or Laplace prior (base.py:451-574, 653-710), Unit or Diagonal mass
 * (MassMatrices.py:82-238), optional bounds with the reflection corrector
 * (base.py:239-374), the leapfrog, 3-stage or 4-stage integrator with
 * per-problem step counts, and 16 chains per problem, advance together.
 * Trajectories run in float32 on MFMA; the endpoint misfit and Metropolis
 * test are FP64.
 * Host layout of states, momenta and gradients: problem after problem, 16
 * rows (chains) of d_p doubles each.  Chain index gc = p * 16 + c. */
typedef struct hmc_set hmc_set;

int hmc_set_create(int64_t problems, const int64_t* dimensions, const int64_t* data, hmc_set** out);
int hmc_set_destroy(hmc_set* s);
/* G row-major n x d float32; variance scalar (scalar_variance=1) or [n]. */
int hmc_set_problem_dense(hmc_set* s, int64_t problem, const float* G, const double* d_obs,
                          const double* variance, int scalar_variance);
/* HMC_PRIOR_NORMAL (a = means, b = inverse variances) or HMC_PRIOR_LAPLACE
 * (a = means, b = inverse dispersions), [d] each. */
int hmc_set_problem_prior(hmc_set* s, int64_t problem, int kind, const double* a, const double* b);
/* Diagonal mass (MassMatrices.py:159-238): kind HMC_MASS_DIAGONAL with
 * diagonal[d] > 0, or HMC_MASS_UNIT (diagonal ignored; the default). */
int hmc_set_problem_mass(hmc_set* s, int64_t problem, int kind, const double* diagonal);
/* Integrator of one problem (HMC_INT_*; Samplers.py:1524-1726; default
 * leapfrog); the option "integrator" sets every problem's. */
int hmc_set_problem_integrator(hmc_set* s, int64_t problem, int integrator);
/* Collapsed bounds of the problem's target (base.py:1061-1101): lower and
 * upper [d] or NULL for no bound on that side; lower < upper elementwise
 * ("Bounds vectors are incompatible.", base.py:359).  Trajectories reflect at
 * the bounds (base.py:257-270); the misfit is inf outside them (:361-374). */
int hmc_set_problem_bounds(hmc_set* s, int64_t problem, const double* lower, const double* upper);
/* "integrator" HMC_INT_LEAPFROG | HMC_INT_3STAGE | HMC_INT_4STAGE (all
 * problems; Samplers.py:1524-1726; default leapfrog),
 * "stagger" 1|0 (default 1: each problem's stage window starts at the launch of
 * the round that best balances the launches' bytes; 0: every trajectory starts
 * at the first launch; same results),
 * "autotune" 1|0 (time the tile shapes on this set at first run; default 1),
 * "resid_rows" 64|128|256, "grad_waves" 4|8, "resid_lds" 0|1 (residual tiles
 * read into registers or staged through LDS; same results) (fix a shape, no
 * autotune); "premult" -1|0|1: trajectories from H = G^T C^-1 G and
 * b = G^T C^-1 d (built once in FP64, stored fp32) instead of the residual
 * form, as LinearMatrix.py:163-177/203-207 does when premultiplication is on;
 * -1 (default) chooses it when one pass over every H reads fewer bytes than
 * the two passes over every G.  The FP64 accept test keeps the residual form;
 * "premult_rows" 1|0: the premultiplied step streams rows of H (H symmetric)
 * or 64-column strips of it; "lds_chunk" 64|128: columns per LDS chunk of the
 * staged row tiles (128 for 64- and 128-row tiles; same results) (fix them,
 * no autotune). */
int hmc_set_set_option(hmc_set* s, const char* name, double value);
/* Resolved "resid_rows", "grad_waves", "resid_lds", "premult_rows", "lds_chunk", "autotune", "stagger", "integrator" (-1 when
 * the problems' integrators differ), "premult"
 * (the last is decided at the first run or proposal).  No reference
 * counterpart (engine configuration). */
int hmc_set_get_option(hmc_set* s, const char* name, double* value);
int hmc_set_tile_config(hmc_set* s, int* resid_rows, int* grad_waves);
int hmc_set_set_state(hmc_set* s, const double* X);
int hmc_set_get_state(hmc_set* s, double* X, double* chi);
/* misfit [P*16] / gradient (host layout) at X (host layout, rounded to float32) */
int hmc_set_misfit(hmc_set* s, const double* X, double* out);
int hmc_set_gradient(hmc_set* s, const double* X, double* out);
/* One proposal with host momenta P0, local steps eps [P*16], accept uniforms
 * u [P*16] and step counts L [P]; outputs may be NULL. */
int hmc_set_propose(hmc_set* s, const double* P0, const double* eps, const double* u, const int32_t* L,
                    int32_t* accepted, double* H0, double* H1, double* X1);
/* Device-RNG proposals; each one a replay of a captured hipGraph when
 * use_graph != 0.  stepsize [P*16], L [P]; accepted [P*16] counts accepts. */
int hmc_set_run(hmc_set* s, int64_t proposals, const int32_t* L, const double* stepsize, int randomize,
                uint64_t seed, int64_t first_proposal, int use_graph, int64_t* accepted);
int hmc_set_set_trace(hmc_set* s, int64_t capacity);
int hmc_set_trace_count(hmc_set* s, int64_t* n);
int hmc_set_read_trace(hmc_set* s, int64_t first, int64_t count, double* X, double* chi);
int hmc_set_enable_timing(hmc_set* s, int on);
/* kind 0 residual, 1 gradient + kick/drift, 2 endpoint misfit, 3 per-chain kernels */
int hmc_set_timing(hmc_set* s, int kind, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* HMC_ENGINE_H */

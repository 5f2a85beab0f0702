// Host-side objects behind the C-ABI handles and the engine entry points.
#pragma once

#include <new>
#include <string>
#include <vector>

#include "hmc_common.hpp"

namespace hmc {

// Sliced ELL with 64-row slices (one wavefront per slice), rows sorted by
// descending length inside the whole matrix so each slice is nearly uniform.
// Element k of slice s, lane r lives at base[s] + k*64 + r (coalesced).
struct Sell {
  int64_t rows = 0, cols = 0, nslices = 0, total = 0;
  int32_t* base = nullptr;   // [nslices] element offset of the slice
  int32_t* width = nullptr;  // [nslices] entries per lane in the slice
  int32_t* perm = nullptr;   // [nslices*64] original row of (slice, lane), -1 = padding
  int32_t* inv = nullptr;    // [rows] position (slice * 64 + lane) of each original row
  uint16_t* col = nullptr;   // [total] column indices (the fused engine's LDS tile keeps them below 65536)
  void* val = nullptr;       // [total] trajectory precision
  void* val_tgt = nullptr;   // [total] target precision (aliases val when equal)
  int max_width = 0;
};

struct Problem {
  int64_t d = 0;            // parameters
  int64_t n = 0;            // data (0: no likelihood)
  int precision = HMC_F64;  // trajectory arithmetic
  int state_dtype = HMC_F64;// storage of positions/momenta (F64 or F32)
  int op = 0;               // 0 none, 1 dense, 2 csr
  int engine = HMC_ENGINE_AUTO;
  int bf16_split = 2;       // bf16 planes of the positions in the residual GEMM (1 or 2)
  int bf16_split_grad = 2;  // bf16 planes of the residuals in the gradient GEMM (1 or 2)
  // bf16 residual-form trajectories: r = r_ref + G_hi bf16(x - x_ref) with the
  // anchor's residual r_ref kept in float32 (one bf16 plane of x - x_ref per
  // stage) instead of hi/lo planes of x
  int resid_update = 0;
  int tile_order = 0;       // bf16 GEMM tiles per XCD: 0 m-major, 1 n-major
  int gemm_cfg = 1;         // bf16 GEMM tile configuration of the residual GEMM
  int gemm_cfg_grad = 1;    // ... of the gradient GEMM
  int split_tail = 1;       // bf16 GEMMs: split the last partial round of tiles along K
  int split_fixup = 0;      // premultiplied bf16 stage: two k-ranges summed inside the GEMM
  int n_cu = 0;             // compute units of the device (split-tail planning)
  int fused_ct = 0;         // fused engine: largest chains per workgroup (0: as LDS allows)
  // FP64 endpoint misfit through H = G^T S^-1 G (dense GEMM engine): -1 when
  // n > d (the reference's premultiplication rule), 0 never, 1 always
  int premult = -1;
  double* H = nullptr;      // [d_pad][ldh] FP64, built at the first endpoint
  double* hb = nullptr;     // [d_pad] G^T S^-1 d
  int64_t ldh = 0;
  // trajectories through H as well (dense GEMM engine; needs the
  // premultiplied endpoint): -1 with the endpoint, 0 residual form, 1 premultiplied
  int premult_traj = -1;
  // premultiplied endpoint over the upper block triangle of H (half the work)
  int premult_sym = 1;
  uint16_t* Hh = nullptr;   // bf16 plane of H [d_pad][ldh] (bf16 trajectories)
  float* Hf = nullptr;      // float32 H (fused float32 trajectories)
  double dtd = 0.0;         // d^T S^-1 d
  int tgt_dtype = HMC_F64;  // element type defining the target's G

  // dense (row-major, zero padded): G [n_pad][ldg], Gt [d_pad][ldgt]
  int64_t n_pad = 0, d_pad = 0, ldg = 0, ldgt = 0;
  void* G = nullptr;
  void* Gt = nullptr;
  void* G_tgt = nullptr;    // target-precision copy of G (aliases G when equal)
  void* Gt_tgt = nullptr;   // target-precision copy of Gt (aliases Gt when equal)

  // sparse
  Sell sg, sgt;

  double* dobs = nullptr;   // [n]
  double* var = nullptr;    // [n] data variance

  // fused sparse engine: data and model arrays in SELL position order (data
  // row of forward-slice position p = sg.perm[p], parameter of backward-slice
  // position q = sgt.perm[q]); padding positions hold neutral values
  double* pos_dobs = nullptr;   // [sg.nslices * 64], padding 0
  double* pos_var = nullptr;    // [sg.nslices * 64], padding 1
  double* pos_model = nullptr;  // [kPosModelArrays][sgt.nslices * 64]
  int model_version = 0;        // bumped by the prior, bounds and mass setters
  int pos_model_version = -1;   // model_version pos_model was gathered at

  // element-wise part
  int n_priors = 0;
  int prior_kind[kMaxPriors] = {0, 0, 0, 0};
  double* prior_a[kMaxPriors] = {nullptr, nullptr, nullptr, nullptr};
  double* prior_b[kMaxPriors] = {nullptr, nullptr, nullptr, nullptr};
  double misfit_const = 0.0;
  double* lo = nullptr;
  double* hi = nullptr;
  double* minv = nullptr;
  double* msqrt = nullptr;
  int mass_kind = HMC_MASS_UNIT;

  hipStream_t stream = nullptr;
  std::vector<void*> owned;   // device allocations freed at destroy

  DevModel model() const {
    DevModel m{};
    m.n_priors = n_priors;
    for (int k = 0; k < n_priors; ++k) m.pr[k] = DevPrior{prior_kind[k], prior_a[k], prior_b[k]};
    m.lo = lo;
    m.hi = hi;
    m.minv = minv;
    m.msqrt = msqrt;
    return m;
  }
  int resolved_engine() const;
};

struct Chains {
  Problem* p = nullptr;
  int64_t B = 0, B_pad = 0;
  int64_t ldx = 0;          // row stride of position/momentum arrays
  int64_t ldr = 0;          // row stride of residual arrays
  void* x0 = nullptr;       // current positions [B_pad][ldx] (state dtype)
  double* chi0 = nullptr;   // [B_pad]
  void* xt = nullptr;       // trajectory positions
  void* pt = nullptr;       // trajectory momenta
  double* p_host = nullptr; // host-supplied momenta staging [B_pad][ldx] f64
  uint16_t* xh = nullptr;   // bf16 hi/lo planes of xt (GEMM engine, BF16)
  uint16_t* xl = nullptr;
  uint16_t* rh = nullptr;   // bf16 hi/lo planes of the residual
  uint16_t* rl = nullptr;
  void* r = nullptr;        // residual [B_pad][ldr] (state dtype, F64/F32 GEMM)
  // Anchored bf16 trajectories (premultiplied, or residual form with
  // resid_update): the stage gradient is a function of the position alone,
  //   premultiplied  g = (H x_ref - b) + H_hi bf16(x - x_ref) + prior'(x)
  //   residual       r = (G x_ref - d) + G_hi bf16(x - x_ref),  g = G^T (r / var) + prior'(x)
  // with one anchor x_ref per chain set, fixed until it is set again, so a
  // trajectory is reversible and volume preserving (a deterministic gradient
  // field).  xa = x_ref [ldx]; yref = H x_ref - b [ldx] or G x_ref - d [ldr],
  // float32 values of FP64 products.
  float* xa = nullptr;
  float* yref = nullptr;
  bool anchor_ready = false;
  // premultiplied bf16: the planes of x - x_ref alternate between two buffers
  // (a stage's GEMM reads one while its epilogue writes the other); FP64 reads
  // a copy xs of the positions while the epilogue updates xt
  uint16_t* dxp[2] = {nullptr, nullptr};
  int dx_cur = 0;
  void* xs = nullptr;
  double* partial = nullptr;// endpoint partial sums [mblocks][B_pad]
  float* split_ws = nullptr;// bf16 GEMM split-tail partial planes
  size_t split_ws_elems = 0;
  int* fix = nullptr;       // two-k-range fix-up: [0] error flag, [1 + tile] counters (kFixTiles)
  bool fix_used = false;
  double* split_ws64 = nullptr;  // FP64 GEMM split-tail partial planes
  size_t split_ws64_elems = 0;
  int64_t n_mblocks = 0;
  double* k0 = nullptr;     // kinetic energy at the start of the trajectory
  double* eps = nullptr;    // local step per chain
  double* u = nullptr;      // accept uniform per chain
  double* h0 = nullptr;
  double* h1 = nullptr;
  int32_t* acc = nullptr;   // accept flag per chain
  double* step = nullptr;   // device per-chain step size (run mode)
  int64_t* acc_count = nullptr;
  int32_t* oob = nullptr;   // per-chain out-of-bounds flag (gradient_batch)

  // autotuning history of device runs (Samplers.py:1419-1426,1494-1522): per
  // proposal and chain the acceptance value exp(H0 - H1) and the step before
  // the update, [hist_cap][B]
  double* hist_rate = nullptr;
  double* hist_step = nullptr;
  int64_t hist_cap = 0, hist_n = 0;

  // trace
  void* tr_x = nullptr;
  double* tr_chi = nullptr;
  int64_t tr_cap = 0, tr_n = 0;

  // timing
  bool timing = false;
  int timing_every = 1;     // events only on proposals with index % timing_every == 0
  bool timing_skip = false; // the current proposal records no events
  // kinds: 0 residual GEMM (fused engine: the whole-proposal kernel),
  // 1 endpoint misfit, 2 element-wise / Metropolis, 3 gradient GEMM with the
  // kick/drift epilogue (residual form), 4 premultiplied stage GEMM with its
  // kick/drift epilogue, 5 surrogate forward + input-gradient GEMMs,
  // 6 surrogate kick/drift
  double t_ms[kTimingKinds] = {};
  int64_t t_n[kTimingKinds] = {};
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  size_t ev_next = 0;

  // learned surrogate driving the trajectory gradient (GEMM engine), or null
  void* surrogate = nullptr;

  // tempering: chain c samples exp(-beta[c] chi); null = all beta 1
  double* beta = nullptr;
  // replica exchange schedule for device runs: event e (after proposal e*ex_interval)
  // pairs ex_pairs[e][2k], ex_pairs[e][2k+1] (the second decides)
  int32_t* ex_pairs = nullptr;
  int64_t ex_events = 0, ex_npairs = 0, ex_interval = 0;
  int64_t* ex_count = nullptr;  // accepted exchanges per chain
  int32_t* ex_tmp = nullptr;    // staging of one host-supplied event: pairs, flags
  double* ex_u = nullptr;       // staging of its uniforms

  std::vector<void*> owned;
};

// error plumbing of the C-ABI: every entry point runs its body through
// guarded(), which turns engine errors into status codes and the thread's
// last-error message.
void set_error(const std::string& msg);
void maybe_inject(const char* name);
template <class F>
int guarded(const char* name, F&& f) {
  try {
    maybe_inject(name);
    f();
    return HMC_OK;
  } catch (const Error& e) {
    set_error(std::string(name) + ": " + e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(std::string(name) + ": host allocation failed");
    return HMC_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(std::string(name) + ": " + e.what());
    return HMC_ERR_STATE;
  } catch (...) {
    set_error(std::string(name) + ": unknown failure");
    return HMC_ERR_STATE;
  }
}

// allocation helpers (zero-filled; the fill has completed on return)
void* dev_alloc(std::vector<void*>& owned, size_t bytes);

// Set-up transfers and fills on the legacy null stream.  A device fill or a
// pageable host-to-device copy there can return before the device has
// finished it, and the engines' non-blocking streams are not ordered after
// null-stream work, so each helper waits for the null stream: the data is in
// place before any later launch on any stream.
void h2d(void* dst, const void* src, size_t bytes);
void d2h(void* dst, const void* src, size_t bytes);
void dzero(void* dst, size_t bytes);
size_t dtype_size(int dt);

// kernel timing brackets (no-ops unless c.timing)
void time_begin(Chains& c, int kind, hipEvent_t& a);
void time_end(Chains& c, int kind, hipEvent_t a);
void time_collect(Chains& c);

// ---- GEMM engine (hmc_gemm.hip) -------------------------------------------
// Row padding of G and G^T: every GEMM configuration's M-tile divides it.
int64_t gemm_row_pad();
size_t gemm_split_workspace(const Problem& p, int64_t B);      // float elements
size_t gemm_split_workspace_f64(const Problem& p, int64_t B);  // double elements
bool endpoint_premult(const Problem& p);      // endpoint misfit uses H (dense GEMM engine)
int64_t endpoint_blocks(const Chains& c);     // rows of the endpoint partial sums
double endpoint_const(const Problem& p);      // constant added to every misfit
// Trajectory gradient + fused kick/drift for every chain:
//   p -= kick*eps*g(xt); xt += drift*eps*Minv p; corrector; [xt += drift2*eps*Minv p; corrector]
void gemm_kick_drift(Chains& c, double kick, double drift, double drift2, bool first);
bool traj_premult(const Problem& p);
void premult_build(Chains& c);            // H, b in FP64 (dense), once
// bf16 trajectories of this problem use an anchor (see Chains::xa)
bool gemm_anchored(const Problem& p);
// set the anchor of a chain set: x_ref from the host (d values) or, when
// null, the mean of the chains' current positions; computes yref in FP64
void gemm_set_anchor(Chains& c, const double* x_ref);
// the gradient a trajectory stage applies at model rows X (state dtype; for
// bf16 the anchored field) into out (device, f64, row stride ldx)
void gemm_traj_gradient(Chains& c, const void* X, double* out);
bool fused_premult(const Problem& p);
// Endpoint partial sums of the data misfit in FP64 into c.partial.
void gemm_endpoint(Chains& c, const void* X, int x_dtype);
// gradient (B x d) of model rows X into out (device, f64, row stride ldx)
void gemm_gradient(Chains& c, const void* X, double* out);

// ---- element-wise kernels (hmc_elem.hip) -----------------------------------
// momentum refresh + first drift (drift0*eps), K0.  host_p: use c.p_host.
void elem_refresh(Chains& c, bool host_p, double drift0, uint64_t seed, int64_t prop,
                  bool device_scalars, int randomize);
// standalone drift + corrector: xt += drift*eps*Minv pt
void elem_drift(Chains& c, double drift);
// endpoint Hamiltonian + Metropolis + state update (+ autotune, trace record)
void elem_metropolis(Chains& c, bool have_partials, int autotune, double target,
                     double learning_rate, int64_t prop_index, int64_t trace_slot,
                     int64_t hist_slot = -1);
// misfit of X (device, state dtype) -> out[B] using c.partial (or none)
void elem_finish_misfit(Chains& c, const void* X, bool have_partials, double* out);
// convert f64 host-layout rows into state dtype rows (+bf16 planes when given)
void elem_load_rows(Chains& c, const double* src, void* dst, uint16_t* hi, uint16_t* lo);
void elem_store_rows(Chains& c, const void* src, int src_dtype, int64_t ld, double* dst);
void elem_oob_flags(Chains& c, const void* X);
// one replica-exchange event over npairs pairs (device array [2*npairs]); u: device
// uniforms per pair, or null for the Philox exchange stream keyed by (seed, decider, prop)
// slot >= 0 rewrites that trace slot's rows of swapped chains (the recorded
// sample of a proposal is its post-exchange state)
void elem_exchange(Chains& c, const int32_t* pairs, int64_t npairs, const double* u, uint64_t seed,
                   int64_t prop, int32_t* swapped, int64_t slot);

// ---- fused engine (hmc_fused.hip) ------------------------------------------
bool fused_supported(const Problem& p);
// Whole proposal (refresh, L steps, endpoint, Metropolis) in one launch.
void fused_propose(Chains& c, bool host_p, int L, int integrator, uint64_t seed, int64_t prop,
                   int randomize, int autotune, double target, double learning_rate,
                   int64_t trace_slot, int64_t hist_slot = -1);
void fused_misfit(Chains& c, const void* X, double* out);
void fused_gradient(Chains& c, const void* X, double* out);

// ---- learned surrogate stage on the bf16 GEMM pipeline (hmc_gemm.hip) -------
// Operands of one surrogate-driven trajectory stage for the chains' rows:
// xh = bf16 normalised positions [rows][Dp]; weight planes w1h [H][Dp],
// w1t [Dp][H], w2h [H][H], w2t [H][H] (row padding to 256 zero filled);
// per-stage buffers a1h, u2h, u1h (bf16 [rows][H]) and s1 (float [rows][H]).
struct SurGemm {
  int64_t D = 0, Dp = 0, H = 0;
  const uint16_t *xh = nullptr, *w1h = nullptr, *w1t = nullptr, *w2h = nullptr, *w2t = nullptr;
  const float *b1 = nullptr, *b2 = nullptr, *w3 = nullptr, *gscale = nullptr;
  const float *xm = nullptr, *xsinv = nullptr;   // input normalisation (x - xm) * xsinv
  uint16_t *a1h = nullptr, *u2h = nullptr, *u1h = nullptr;
  float* s1 = nullptr;
  float* ws = nullptr;        // split-tail partials
  size_t ws_elems = 0;
};
size_t surrogate_split_workspace(int64_t D, int64_t H, int64_t B, int n_cu);
// Z1 = W1 xn + b1, A1 = silu(Z1); U2 = silu'(W2 A1 + b2) w3; V1 = W2^T U2;
// U1 = silu'(Z1) V1; g = gscale (W1^T U1) + prior'(x) -> kick/drift of the
// chains (four GEMMs; the last one's epilogue is the leapfrog stage)
void surrogate_stage_bf16(Chains& c, const SurGemm& g, double kick, double drift, double drift2);

// ---- learned surrogate (hmc_surrogate.hip) ---------------------------------
// Same contract as gemm_kick_drift with the likelihood gradient taken from
// the surrogate network attached to the chains.
void surrogate_kick_drift(Chains& c, double kick, double drift, double drift2, bool first);

// integrator schedules: coefficients of the first drift and of each
// (kick, drift, drift2) stage; returned in units of the local step.
struct Stage {
  double kick, drift, drift2;
};
void integrator_schedule(int integrator, int L, double& first_drift, std::vector<Stage>& stages);

}  // namespace hmc

struct hmc_problem {
  hmc::Problem p;
};
struct hmc_chains {
  hmc::Chains c;
};

// Many independent dense linear problems of uneven size, sampled together.
//
// hmclab samples one target per process (ParallelSampleSMP runs one
// HMC.sample per chain, Samplers.py:1757-1977; the trajectory is
// HMC._propagate_leapfrog, Samplers.py:1524-1584, and the accept test
// HMC._evaluate_acceptance, Samplers.py:1471-1492).  Here P problems x C=16
// chains share every launch:
//
//   refresh      one workgroup per chain: momentum, first half drift, K0
//   resid        tiles (problem, 64*RW rows): w = (G x - d) / var       (fp32 MFMA)
//   grad_leap    tiles (problem, 64 columns): g = G^T w + prior'(x);
//                p -= eps g; x += c eps p                               (fp32 MFMA)
//   endpoint     tiles (problem, 64 rows): sum (G x - d)^2 / var         (fp64 MFMA)
//   metropolis   one workgroup per chain: exact FP64 H0/H1, accept, trace
//
// Step s of the trajectory launches only the tiles of problems with L > s:
// problems are ordered by descending L so that those tiles form a prefix of
// one tile list.  A whole proposal round is captured once as a hipGraph and
// replayed; the RNG counter and the run parameters live in device memory so
// replays need no new arguments.  Two kernel shapes (rows per residual tile,
// waves per gradient tile) are chosen per set by timing every combination
// on the set itself.
//
// Premultiplied trajectories (LinearMatrix.py:163-177, 203-207: the
// reference's default when n > d): H = G^T C^-1 G and b = G^T C^-1 d are
// built once in FP64 (stored fp32), and each step is one pass over H,
//   grad_leap    tiles (problem, 64 columns): g = H x - b + prior'(x); kick/drift
// with no residual launch.  Every workgroup of a problem then reads the whole
// x, so x alternates between two buffers: step s reads buffer (L - s) mod 2
// and writes the other, and the last step always writes xt.  The endpoint
// stays in the residual form above (exact for the fp32 state).
//
// Layouts in HBM (fp32 unless noted), per problem p with d, n:
//   G   [round_up(n,16)][ldg = round_up(d,16)]  row-major, zero padded
//   x0, xt, pt  [C][ldg]   chain-major states (fp32), zero padded
//   w   [C][ldr = round_up(n,16)]   scaled residuals
//   dobs, ivar  [n] fp64;   prior a, b  [d] fp64
//   premultiplied: H [ldg][ldg], xt2 [C][ldg] (fp32);  b [d] fp64
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "hmc_internal.hpp"

namespace hmc {

// k blocks (residual) or row groups (gradient) whose loads are issued together
constexpr int SET_U = 4;

constexpr int SC = 16;   // chains per problem: one MFMA tile width

struct SetDesc {
  int32_t d, n, ldg, ldr;
  int64_t g_off, x_off, w_off, v_off, a_off, h_off;   // h_off: offset in host-layout rows
  int64_t hm_off;                                     // premultiplied G^T C^-1 G [ldg][ldg]
  int32_t prior_kind, L;
  int32_t ptile0, n_ptiles;                           // endpoint tiles of this problem
  int32_t S, c_off;       // gradient stages of a trajectory; first stage's coefficients
  int32_t mass, bounded;  // diagonal mass; bounds (reflection corrector)
  double fd;              // first drift coefficient of the problem's integrator
  int32_t so, pad_;       // launch of the round that runs the trajectory's first stage
};

// Per-set model arrays beside the descriptors: stage coefficients {kick,
// drift, drift2} (the integrator's schedule for each problem's L, as the
// chain engine's), inverse / square-root mass diagonals and bounds (a_off
// layout; present when some problem has them).
struct SetModel {
  const double* coef;
  const double* minv;
  const double* msqrt;
  const double* lo;
  const double* hi;
};

__device__ inline bool set_oob(const SetModel& m, const SetDesc& D, int64_t aj, double x) {
  return D.bounded && (x < m.lo[aj] || x > m.hi[aj]);
}

// dK/dp (MassMatrices.py:133, 218)
__device__ inline float set_kgrad(const SetModel& m, const SetDesc& D, int64_t aj, float p) {
  return D.mass ? (float)m.minv[aj] * p : p;
}

// reflection at the problem's bounds, lower first then upper (base.py:257-270)
__device__ inline void set_correct(const SetModel& m, const SetDesc& D, int64_t aj, float& x, float& p) {
  if (!D.bounded) return;
  const float lo = (float)m.lo[aj];
  if (x < lo) {
    x += 2.f * (lo - x);
    p *= -1.f;
  }
  const float hi = (float)m.hi[aj];
  if (x > hi) {
    x += 2.f * (hi - x);
    p *= -1.f;
  }
}

// Stage `step` of one coordinate: gradient g (likelihood + prior), kick,
// drift(s) with the corrector after each (Samplers.py:1524-1726).
__device__ inline void set_leap(const SetModel& m, const SetDesc& D, int64_t aj, int step, double e, float g,
                                float& xv, float& pv) {
  const double* cf = m.coef + 3 * ((int64_t)D.c_off + step);
  const float kk = (float)(cf[0] * e), dd = (float)(cf[1] * e), dd2 = (float)(cf[2] * e);
  if (set_oob(m, D, aj, (double)xv)) g = INFINITY;
  pv -= kk * g;
  xv += dd * set_kgrad(m, D, aj, pv);
  set_correct(m, D, aj, xv, pv);
  if (cf[2] != 0.0) {
    xv += dd2 * set_kgrad(m, D, aj, pv);
    set_correct(m, D, aj, xv, pv);
  }
}

struct SetRun {
  uint64_t seed;
  int64_t prop;        // proposal index of the next refresh
  int64_t trace_base;  // proposal index recorded in trace slot 0
  int64_t trace_cap;
  int32_t randomize;
  int32_t host_p;      // momenta/eps/u supplied by the host
};

__device__ inline float set_prior_grad(int kind, double a, double b, float x) {
  if (kind == HMC_PRIOR_NORMAL) return -(float)b * ((float)a - x);
  if (kind == HMC_PRIOR_LAPLACE) {
    const float af = (float)a;
    const float s = (x > af) ? 1.f : ((x < af) ? -1.f : 0.f);
    return s * (float)b;
  }
  return 0.f;
}

__device__ inline double set_prior_misfit(int kind, double a, double b, double x) {
  if (kind == HMC_PRIOR_NORMAL) {
    const double r = a - x;
    return 0.5 * (r * (b * r));
  }
  if (kind == HMC_PRIOR_LAPLACE) return fabs(x - a) * b;
  return 0.0;
}

// ---------------------------------------------------------------------------
// residual: w[c][row] = (sum_k G[row][k] x[c][k] - dobs[row]) * ivar[row]
// 4 waves; wave wv owns RW MFMA row tiles of 16 rows.  MFMA 16x16x4 f32:
// lane (i = l%16, q = l/16) feeds A = G[row i][k], B = x[chain i][k] with
// k = k0 + 4q + e in instruction e (one float4 load per operand feeds four
// instructions); the accumulator of lane (q, i) holds rows 4q+v of chain i.
// ---------------------------------------------------------------------------
template <int RW>
__global__ __launch_bounds__(256) void k_set_resid(const SetDesc* __restrict__ desc,
                                                   const int2* __restrict__ tiles,
                                                   const float* __restrict__ G,
                                                   const float* __restrict__ X,
                                                   const double* __restrict__ dobs,
                                                   const double* __restrict__ ivar,
                                                   float* __restrict__ W) {
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = lane & 15, q = lane >> 4;
  const int rbase = t.y * (64 * RW) + wv * (16 * RW);
  const float* xrow = X + D.x_off + (int64_t)i * D.ldg + 4 * q;
  floatx4 acc[RW];
  const float* grow[RW];
  bool act[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    acc[r] = floatx4{0.f, 0.f, 0.f, 0.f};
    act[r] = rbase + 16 * r < D.n;
    grow[r] = G + D.g_off + (int64_t)(rbase + 16 * r + i) * D.ldg + 4 * q;
  }
  // the loads of SET_U consecutive 16-wide k blocks are issued before their
  // MFMAs (same accumulation order)
  int k0 = 0;
  for (; k0 + 16 * SET_U <= D.ldg; k0 += 16 * SET_U) {
    floatx4 bb[SET_U], aa[SET_U][RW];
#pragma unroll
    for (int u = 0; u < SET_U; ++u) {
      bb[u] = *(const floatx4*)(xrow + k0 + 16 * u);
#pragma unroll
      for (int r = 0; r < RW; ++r)
        aa[u][r] = act[r] ? *(const floatx4*)(grow[r] + k0 + 16 * u) : floatx4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < SET_U; ++u)
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        if (!act[r]) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[u][r][e], bb[u][e], acc[r], 0, 0, 0);
      }
  }
  for (; k0 < D.ldg; k0 += 16) {
    const floatx4 b = *(const floatx4*)(xrow + k0);
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      if (!act[r]) continue;
      const floatx4 a = *(const floatx4*)(grow[r] + k0);
      acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], acc[r], 0, 0, 0);
      acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], acc[r], 0, 0, 0);
      acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], acc[r], 0, 0, 0);
      acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], acc[r], 0, 0, 0);
    }
  }
  const double* dp = dobs + D.v_off;
  const double* vp = ivar + D.v_off;
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    if (!act[r]) continue;
    const int row0 = rbase + 16 * r + 4 * q;
    floatx4 o;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = row0 + v;
      o[v] = (row < D.n) ? (acc[r][v] - (float)dp[row]) * (float)vp[row] : 0.f;
    }
    *(floatx4*)(W + D.w_off + (int64_t)i * D.ldr + row0) = o;
  }
}


// ---------------------------------------------------------------------------
// residual, LDS-staged variant: the tile's 64*RW rows of G are read KC (64 or
// 128) columns at a time, KC/4 lanes per row (256 or 512 contiguous bytes per
// row), into LDS through registers (the next chunk's loads are in flight
// while the current chunk's MFMAs run); the MFMA sequence per accumulator is
// the one of k_set_resid (k = 16 s + 4 q + e), so all give identical results.
//
// KICK = 1: the premultiplied step on rows of H instead of G.  H is symmetric,
// so the tile's rows j give (H x)_j = sum_k H[j][k] x[k] for its 64*RW
// parameters, streamed row by row; the epilogue is the kick/drift of
// k_set_grad (g = H x - b + prior'(x), x read from buffer (L - step) mod 2 of
// X / X2 and written to the other).
// ---------------------------------------------------------------------------
template <int RW, int KICK, int KC>
__global__ __launch_bounds__(256) void k_set_resid_lds(const SetDesc* __restrict__ desc,
                                                       const int2* __restrict__ tiles,
                                                       const float* __restrict__ G,
                                                       float* __restrict__ X,
                                                       const double* __restrict__ dobs,
                                                       const double* __restrict__ ivar,
                                                       float* __restrict__ W,
                                                       float* __restrict__ X2, float* __restrict__ P,
                                                       const double* __restrict__ pa,
                                                       const double* __restrict__ pb,
                                                       const double* __restrict__ hb,
                                                       const double* __restrict__ eps, int step, SetModel md) {
  constexpr int ROWS = 64 * RW, LD = KC + 4;
  constexpr int LPR = KC / 4, RPP = 256 / LPR;   // loader lanes per row, rows per pass
  constexpr int GPT = ROWS / RPP;                // float4 loads of G per thread per chunk
  constexpr int XPT = SC / RPP;                  // ... of x (KC = 64: 1, 128: 2)
  static_assert(KC == 64 || KC == 128, "chunk width");
  __shared__ __attribute__((aligned(16))) float Gs[ROWS * LD];
  __shared__ __attribute__((aligned(16))) float Xs[SC * LD];
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i = lane & 15, q = lane >> 4;
  const int r0 = t.y * ROWS;
  const int nr = KICK ? D.d : D.n;               // rows of the operator
  const int nrow = (nr + 15) & ~15;              // rows present in memory (zero padded)
  const float* Gp = G + (KICK ? D.hm_off : D.g_off);
  const int ls = step - D.so;   // the problem's own stage (KICK)
  const bool flip = KICK && ((D.S - ls) & 1);
  const float* Xin = flip ? X2 : X;
  const float* Xp = Xin + D.x_off;
  // loader mapping: thread tid covers column 4 * (tid % LPR) of rows tid / LPR + RPP j
  const int lc = 4 * (tid % LPR), lr = tid / LPR;
  floatx4 gr[GPT], xr[XPT];
  auto gload = [&](int k0) {
    const int col = k0 + lc;
#pragma unroll
    for (int j = 0; j < GPT; ++j) {
      const int row = r0 + lr + RPP * j;
      gr[j] = (row < nrow && col < D.ldg) ? *(const floatx4*)(Gp + (int64_t)row * D.ldg + col)
                                           : floatx4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < XPT; ++j)
      xr[j] = (col < D.ldg) ? *(const floatx4*)(Xp + (int64_t)(lr + RPP * j) * D.ldg + col)
                            : floatx4{0.f, 0.f, 0.f, 0.f};
  };
  auto sstore = [&]() {
#pragma unroll
    for (int j = 0; j < GPT; ++j) *(floatx4*)(&Gs[(lr + RPP * j) * LD + lc]) = gr[j];
#pragma unroll
    for (int j = 0; j < XPT; ++j) *(floatx4*)(&Xs[(lr + RPP * j) * LD + lc]) = xr[j];
  };
  floatx4 acc[RW];
  bool act[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    acc[r] = floatx4{0.f, 0.f, 0.f, 0.f};
    act[r] = r0 + wv * 16 * RW + 16 * r < nr;
  }
  gload(0);
  for (int k0 = 0; k0 < D.ldg; k0 += KC) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + KC < D.ldg) gload(k0 + KC);
    const int kmax = D.ldg - k0 < KC ? D.ldg - k0 : KC;   // a multiple of 16
    for (int s16 = 0; s16 < kmax; s16 += 16) {
      const floatx4 b = *(const floatx4*)(&Xs[i * LD + s16 + 4 * q]);
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        if (!act[r]) continue;
        const floatx4 a = *(const floatx4*)(&Gs[(wv * 16 * RW + 16 * r + i) * LD + s16 + 4 * q]);
        acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], acc[r], 0, 0, 0);
        acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], acc[r], 0, 0, 0);
      }
    }
  }
  if (KICK) {
    // lane (i, q) holds parameters row0 .. row0 + 3 of chain i
    if (ls < 0 || ls >= D.S) return;
    const double ec = eps[(int64_t)t.x * SC + i];
    float* Xout = flip ? X : X2;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      if (!act[r]) continue;
      const int row0 = r0 + wv * 16 * RW + 16 * r + 4 * q;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = row0 + v;
        if (j >= D.d) continue;
        const int64_t o = D.x_off + (int64_t)i * D.ldg + j;
        const float xv0 = Xin[o];
        const float gv = (acc[r][v] - (float)hb[D.a_off + j]) +
                         set_prior_grad(D.prior_kind, pa[D.a_off + j], pb[D.a_off + j], xv0);
        float xv = xv0, pv = P[o];
        set_leap(md, D, D.a_off + j, ls, ec, gv, xv, pv);
        Xout[o] = xv;
        P[o] = pv;
      }
    }
    return;
  }
  const double* dp = dobs + D.v_off;
  const double* vp = ivar + D.v_off;
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    if (!act[r]) continue;
    const int row0 = r0 + wv * 16 * RW + 16 * r + 4 * q;
    floatx4 o;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = row0 + v;
      o[v] = (row < D.n) ? (acc[r][v] - (float)dp[row]) * (float)vp[row] : 0.f;
    }
    *(floatx4*)(W + D.w_off + (int64_t)i * D.ldr + row0) = o;
  }
}

// ---------------------------------------------------------------------------
// gradient + kick/drift on 64 columns of one problem.  NW waves split the row
// groups of 16; lane (i, q) loads G[row][col0 + 4i .. 4i+3] (one float4 per
// row, 256 contiguous bytes per wave) and feeds component t to MFMA tile t,
// so tile t's A row i is column col0 + 4i + t; B = w[chain i][row].  The
// accumulator of lane (q, i) then holds columns col0 + 16q + 4v + t of chain i.
// MODE 0: kick/drift; MODE 1: store the gradient (f64) for gradient queries;
// MODE 2: kick/drift from the premultiplied form: G is H, the rows summed are
// H's ldg rows and the weights are the chain states x (H symmetric, so
// sum_k H[k][j] x[k] = (H x)[j]), read from buffer (L - step) mod 2 of
// X / X2 and written to the other; b is subtracted.
// ---------------------------------------------------------------------------
template <int NW, int MODE>
__global__ __launch_bounds__(NW * 64) void k_set_grad(const SetDesc* __restrict__ desc,
                                                      const int2* __restrict__ tiles,
                                                      const float* __restrict__ G,
                                                      const float* __restrict__ W,
                                                      float* __restrict__ X, float* __restrict__ P,
                                                      const double* __restrict__ pa,
                                                      const double* __restrict__ pb,
                                                      const double* __restrict__ eps, int step,
                                                      double* __restrict__ gout,
                                                      float* __restrict__ X2,
                                                      const double* __restrict__ hb, SetModel md,
                                                      const int32_t* __restrict__ oob) {
  __shared__ float red[NW][SC][64];
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = lane & 15, q = lane >> 4;
  const int col0 = t.y * 64;
  const bool cin = col0 + 4 * i < D.ldg;
  floatx4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int ls = step - D.so;   // the problem's own stage
  const bool flip = MODE == 2 && ((D.S - ls) & 1);
  const float* Xin = flip ? X2 : X;
  float* Xout = MODE == 2 ? (flip ? X : X2) : X;
  const float* wrow = MODE == 2 ? Xin + D.x_off + (int64_t)i * D.ldg + 4 * q
                                : W + D.w_off + (int64_t)i * D.ldr + 4 * q;
  const float* gcol = G + (MODE == 2 ? D.hm_off : D.g_off) + col0 + 4 * i;
  const int ngroups = MODE == 2 ? D.ldg >> 4 : (D.n + 15) >> 4;
  int g = wv;
  // SET_U row groups of this wave at a time: loads first, then the MFMAs
  for (; g + NW * (SET_U - 1) < ngroups; g += NW * SET_U) {
    floatx4 bb[SET_U], aa[SET_U][4];
#pragma unroll
    for (int v = 0; v < SET_U; ++v) {
      const int r0 = (g + NW * v) * 16;
      bb[v] = *(const floatx4*)(wrow + r0);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        aa[v][e] = cin ? *(const floatx4*)(gcol + (int64_t)(r0 + 4 * q + e) * D.ldg) : floatx4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int v = 0; v < SET_U; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[v][e][u], bb[v][e], acc[u], 0, 0, 0);
  }
  for (; g < ngroups; g += NW) {
    const int r0 = g * 16;
    const floatx4 b = *(const floatx4*)(wrow + r0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const floatx4 a = cin ? *(const floatx4*)(gcol + (int64_t)(r0 + 4 * q + e) * D.ldg)
                            : floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u)
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[e], acc[u], 0, 0, 0);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wv][i][16 * q + 4 * v + u] = acc[u][v];
  __syncthreads();
  for (int e = threadIdx.x; e < SC * 64; e += NW * 64) {
    const int c = e >> 6, col = e & 63;
    const int j = col0 + col;
    if (j >= D.d) continue;
    float gs = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) gs += red[w][c][col];
    const int64_t o = D.x_off + (int64_t)c * D.ldg + j;
    const float xv0 = Xin[o];
    if (MODE == 2) gs -= (float)hb[D.a_off + j];
    const float gv = gs + set_prior_grad(D.prior_kind, pa[D.a_off + j], pb[D.a_off + j], xv0);
    if (MODE == 1) {
      // BayesRule gradient: + inf everywhere when the chain is outside the
      // bounds (base.py:1045-1054)
      gout[D.h_off + (int64_t)c * D.d + j] = oob[(int64_t)t.x * SC + c] ? INFINITY : (double)gv;
    } else {
      if (ls < 0 || ls >= D.S) continue;
      float xv = xv0, pv = P[o];
      set_leap(md, D, D.a_off + j, ls, eps[(int64_t)t.x * SC + c], gv, xv, pv);
      Xout[o] = xv;
      P[o] = pv;
    }
  }
}

// ---------------------------------------------------------------------------
// premultiplied operator of one problem, 64 x 64 tile t.y of H (row-major
// tiles of ceil(ldg/64) per side): H[j][k] = sum_r G[r][j] (1/var_r) G[r][k]
// in FP64 (MFMA 16x16x4 f64, rows r in order), stored fp32.  Wave wv covers
// rows j0 + 16 wv .. + 15; accumulator u holds columns k0 + 16 u + i of rows
// j0 + 16 wv + q + 4 v in lane (i, q).  G's zero padding gives zero padding.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_set_build_h(const SetDesc* __restrict__ desc,
                                                     const int2* __restrict__ tiles,
                                                     const float* __restrict__ G,
                                                     const double* __restrict__ ivar,
                                                     float* __restrict__ H) {
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = lane & 15, q = lane >> 4;
  const int nt = (D.ldg + 63) / 64;
  const int j0 = (t.y / nt) * 64 + 16 * wv, k0 = (t.y % nt) * 64;
  const float* Gp = G + D.g_off;
  const double* wp = ivar + D.v_off;
  const bool ja = j0 + i < D.ldg;
  bool kb[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) kb[u] = k0 + 16 * u + i < D.ldg;
  doublex4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = doublex4{0.0, 0.0, 0.0, 0.0};
  for (int r0 = 0; r0 < D.n; r0 += 4) {
    const int r = r0 + q;
    const bool rv = r < D.n;
    const float* grow = Gp + (int64_t)r * D.ldg;
    const double a = (rv && ja) ? (double)grow[j0 + i] : 0.0;
    const double w = rv ? wp[r] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double b = (rv && kb[u]) ? w * (double)grow[k0 + 16 * u + i] : 0.0;
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u], 0, 0, 0);
    }
  }
  float* Hp = H + D.hm_off;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!kb[u]) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = j0 + q + 4 * v;
      if (j < D.ldg) Hp[(int64_t)j * D.ldg + k0 + 16 * u + i] = (float)acc[u][v];
    }
  }
}

// b[j] = sum_r G[r][j] d_r / var_r in FP64 for the 64 columns of a gradient
// tile: four row phases of 64 threads, summed in fixed order.
__global__ __launch_bounds__(256) void k_set_build_b(const SetDesc* __restrict__ desc,
                                                     const int2* __restrict__ tiles,
                                                     const float* __restrict__ G,
                                                     const double* __restrict__ dobs,
                                                     const double* __restrict__ ivar,
                                                     double* __restrict__ hb) {
  __shared__ double part[4][64];
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int col = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int j = t.y * 64 + col;
  double acc = 0.0;
  if (j < D.d)
    for (int r = ph; r < D.n; r += 4)
      acc += (double)G[D.g_off + (int64_t)r * D.ldg + j] * (dobs[D.v_off + r] * ivar[D.v_off + r]);
  part[ph][col] = acc;
  __syncthreads();
  if (ph == 0 && j < D.d)
    hb[D.a_off + j] = ((part[0][col] + part[1][col]) + part[2][col]) + part[3][col];
}

// ---------------------------------------------------------------------------
// endpoint data misfit in FP64: partial[tile][c] = sum over the tile's 64 rows
// of (G x - d)^2 / var.  MFMA 16x16x4 f64 with the same k order as resid; the
// accumulator of lane (q, i) holds rows q + 4v of chain i.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_set_endpoint(const SetDesc* __restrict__ desc,
                                                      const int2* __restrict__ tiles,
                                                      const float* __restrict__ G,
                                                      const float* __restrict__ X,
                                                      const double* __restrict__ dobs,
                                                      const double* __restrict__ ivar,
                                                      double* __restrict__ partial) {
  __shared__ double part[4][SC];
  const int2 t = tiles[blockIdx.x];
  const SetDesc D = desc[t.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = lane & 15, q = lane >> 4;
  const int rbase = t.y * 64 + wv * 16;
  double s = 0.0;
  if (rbase < D.n) {
    const float* grow = G + D.g_off + (int64_t)(rbase + i) * D.ldg + 4 * q;
    const float* xrow = X + D.x_off + (int64_t)i * D.ldg + 4 * q;
    doublex4 acc = {0.0, 0.0, 0.0, 0.0};
    // the loads of SET_U consecutive 16-wide k blocks are issued before their
    // MFMAs (same accumulation order)
    int k0 = 0;
    for (; k0 + 16 * SET_U <= D.ldg; k0 += 16 * SET_U) {
      floatx4 aa[SET_U], bb[SET_U];
#pragma unroll
      for (int u = 0; u < SET_U; ++u) {
        aa[u] = *(const floatx4*)(grow + k0 + 16 * u);
        bb[u] = *(const floatx4*)(xrow + k0 + 16 * u);
      }
#pragma unroll
      for (int u = 0; u < SET_U; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)aa[u][e], (double)bb[u][e], acc, 0, 0, 0);
    }
    for (; k0 < D.ldg; k0 += 16) {
      const floatx4 a = *(const floatx4*)(grow + k0);
      const floatx4 b = *(const floatx4*)(xrow + k0);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a[e], (double)b[e], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rbase + q + 4 * v;
      if (row < D.n) {
        const double r = acc[v] - dobs[D.v_off + row];
        s += r * (r * ivar[D.v_off + row]);
      }
    }
  }
  s += __shfl_xor(s, 16, 64);
  s += __shfl_xor(s, 32, 64);
  if (q == 0) part[wv][i] = s;
  __syncthreads();
  if (threadIdx.x < SC) {
    const double v = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                     part[3][threadIdx.x];
    partial[(int64_t)blockIdx.x * SC + threadIdx.x] = v;
  }
}

// ---------------------------------------------------------------------------
// per-chain kernels (one workgroup of 256 per chain gc = problem * 16 + c)
// ---------------------------------------------------------------------------
constexpr int SW = 256;

__global__ __launch_bounds__(SW) void k_set_refresh(const SetDesc* __restrict__ desc,
                                                    const float* __restrict__ X0, float* __restrict__ XT,
                                                    float* __restrict__ XT2, float* __restrict__ PT,
                                                    const double* __restrict__ p_host,
                                                    double* __restrict__ k0, double* __restrict__ eps,
                                                    double* __restrict__ u,
                                                    const double* __restrict__ stepv,
                                                    const SetRun* __restrict__ run, SetModel md) {
  __shared__ double scratch[SW / 64 + 1];
  const int64_t gc = blockIdx.x;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  const SetRun R = *run;
  double e;
  if (!R.host_p) {
    double ue, ua;
    philox_scalars(R.seed, (uint32_t)gc, (uint64_t)R.prop, ue, ua);
    e = R.randomize ? ue * stepv[gc] : stepv[gc];
    if (threadIdx.x == 0) {
      eps[gc] = e;
      u[gc] = ua;
    }
  } else {
    e = eps[gc];
  }
  const float dd = (float)(D.fd * e);
  double ks = 0.0;
  const int64_t base = D.x_off + (int64_t)c * D.ldg;
  // premultiplied trajectories start in buffer S mod 2 (their last stage writes XT)
  float* xs = (XT2 && (D.S & 1)) ? XT2 : XT;
  const int npairs = (D.d + 1) / 2;
  for (int qp = threadIdx.x; qp < npairs; qp += SW) {
    double z[2] = {0.0, 0.0};
    if (!R.host_p) philox_normal2(R.seed, (uint32_t)gc, (uint64_t)R.prop, (uint32_t)qp, z[0], z[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = 2 * qp + h;
      if (j >= D.d) break;
      const int64_t aj = D.a_off + j;
      float pv = R.host_p ? (float)p_host[D.h_off + (int64_t)c * D.d + j]
                          : (float)(D.mass ? md.msqrt[aj] * z[h] : z[h]);
      ks += D.mass ? (double)pv * (md.minv[aj] * (double)pv) : (double)pv * (double)pv;
      float xv = X0[base + j];
      xv += dd * set_kgrad(md, D, aj, pv);
      set_correct(md, D, aj, xv, pv);
      xs[base + j] = xv;
      PT[base + j] = pv;
    }
  }
  const double s = block_sum<SW>(ks, scratch);
  if (threadIdx.x == 0) k0[gc] = 0.5 * s;
}

// misfit of chain c (prior + likelihood partials; inf outside the bounds,
// base.py:361-374) and, with p, its kinetic energy (MassMatrices.py:113, 198)
__device__ inline double set_chain_misfit(const SetDesc& D, const float* x, const double* partial,
                                          int c, const double* pa, const double* pb, double* scratch,
                                          const float* p, double* kin, const SetModel& md) {
  double pm = 0.0, kk = 0.0, ob = 0.0;
  for (int j = threadIdx.x; j < D.d; j += SW) {
    const int64_t aj = D.a_off + j;
    pm += set_prior_misfit(D.prior_kind, pa[aj], pb[aj], (double)x[j]);
    if (set_oob(md, D, aj, (double)x[j])) ob = 1.0;
    if (p) kk += D.mass ? (double)p[j] * (md.minv[aj] * (double)p[j]) : (double)p[j] * (double)p[j];
  }
  pm = block_sum<SW>(pm, scratch);
  if (p) kk = block_sum<SW>(kk, scratch);
  if (D.bounded) ob = block_sum<SW>(ob, scratch);
  double lik = 0.0;
  for (int k = 0; k < D.n_ptiles; ++k) lik += partial[(int64_t)(D.ptile0 + k) * SC + c];
  if (kin) *kin = 0.5 * kk;
  return ob > 0.0 ? INFINITY : 0.5 * lik + pm;
}

__global__ __launch_bounds__(SW) void k_set_metropolis(
    const SetDesc* __restrict__ desc, float* __restrict__ X0, const float* __restrict__ XT,
    const float* __restrict__ PT, const double* __restrict__ partial, const double* __restrict__ pa,
    const double* __restrict__ pb, double* __restrict__ chi0, const double* __restrict__ k0,
    const double* __restrict__ u, double* __restrict__ h0o, double* __restrict__ h1o,
    int32_t* __restrict__ acc, int64_t* __restrict__ acc_count, const SetRun* __restrict__ run,
    float* __restrict__ tr_x, double* __restrict__ tr_chi, int64_t tr_row, int64_t nchains, SetModel md) {
  __shared__ double scratch[SW / 64 + 1];
  __shared__ int accepted;
  __shared__ double chi_new;
  const int64_t gc = blockIdx.x;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  const int64_t base = D.x_off + (int64_t)c * D.ldg;
  double k1;
  const double chi1 = set_chain_misfit(D, XT + base, partial, c, pa, pb, scratch, PT + base, &k1, md);
  if (threadIdx.x == 0) {
    const double H0 = chi0[gc] + k0[gc];
    const double H1 = chi1 + k1;
    const double a = exp(H0 - H1);
    const int ok = (a > u[gc]) ? 1 : 0;   // strict, Samplers.py:1486
    accepted = ok;
    chi_new = ok ? chi1 : chi0[gc];
    if (h0o) h0o[gc] = H0;
    if (h1o) h1o[gc] = H1;
    if (acc) acc[gc] = ok;
    acc_count[gc] += ok;
    chi0[gc] = chi_new;
  }
  __syncthreads();
  if (accepted)
    for (int j = threadIdx.x; j < D.d; j += SW) X0[base + j] = XT[base + j];
  const SetRun R = *run;
  const int64_t slot = R.prop - R.trace_base;
  if (tr_x && slot >= 0 && slot < R.trace_cap) {
    const float* src = accepted ? XT : X0;
    float* dst = tr_x + slot * tr_row + D.h_off + (int64_t)c * D.d;
    for (int j = threadIdx.x; j < D.d; j += SW) dst[j] = src[base + j];
    if (threadIdx.x == 0) tr_chi[slot * nchains + gc] = chi_new;
  }
}

__global__ void k_set_advance(SetRun* run) {
  if (threadIdx.x == 0 && blockIdx.x == 0) run->prop += 1;
}

__global__ __launch_bounds__(SW) void k_set_finish(const SetDesc* __restrict__ desc,
                                                   const float* __restrict__ X,
                                                   const double* __restrict__ partial,
                                                   const double* __restrict__ pa,
                                                   const double* __restrict__ pb,
                                                   double* __restrict__ out, SetModel md) {
  __shared__ double scratch[SW / 64 + 1];
  const int64_t gc = blockIdx.x;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  const double chi = set_chain_misfit(D, X + D.x_off + (int64_t)c * D.ldg, partial, c, pa, pb, scratch,
                                      nullptr, nullptr, md);
  if (threadIdx.x == 0) out[gc] = chi;
}

// per chain: 1 when its state lies outside its problem's bounds
__global__ __launch_bounds__(SW) void k_set_oob(const SetDesc* __restrict__ desc, const float* __restrict__ X,
                                                int32_t* __restrict__ oob, SetModel md) {
  __shared__ double scratch[SW / 64 + 1];
  const int64_t gc = blockIdx.x;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  double ob = 0.0;
  if (D.bounded)
    for (int j = threadIdx.x; j < D.d; j += SW)
      if (set_oob(md, D, D.a_off + j, (double)X[D.x_off + (int64_t)c * D.ldg + j])) ob = 1.0;
  ob = block_sum<SW>(ob, scratch);
  if (threadIdx.x == 0) oob[gc] = ob > 0.0 ? 1 : 0;
}

// host-layout f64 rows <-> padded device fp32 states
__global__ void k_set_load(const SetDesc* __restrict__ desc, const double* __restrict__ src,
                           float* __restrict__ dst, int64_t total_h) {
  const int64_t gc = blockIdx.y;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < D.d; j += gridDim.x * blockDim.x)
    dst[D.x_off + (int64_t)c * D.ldg + j] = (float)src[D.h_off + (int64_t)c * D.d + j];
}

__global__ void k_set_store(const SetDesc* __restrict__ desc, const float* __restrict__ src,
                            double* __restrict__ dst) {
  const int64_t gc = blockIdx.y;
  const SetDesc D = desc[gc / SC];
  const int c = (int)(gc % SC);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < D.d; j += gridDim.x * blockDim.x)
    dst[D.h_off + (int64_t)c * D.d + j] = (double)src[D.x_off + (int64_t)c * D.ldg + j];
}

// ---------------------------------------------------------------------------
// host object
// ---------------------------------------------------------------------------
struct SetProblemHost {
  int64_t d = 0, n = 0;
  bool has_data = false;
  int prior_kind = 0;
  bool mass = false, bounded = false;
  int integ = HMC_INT_LEAPFROG;
};

struct ProblemSet {
  int64_t P = 0;
  std::vector<SetProblemHost> hp;
  std::vector<SetDesc> hd;         // host copy of descriptors
  SetDesc* desc = nullptr;
  int64_t g_total = 0, x_total = 0, w_total = 0, v_total = 0, a_total = 0, h_total = 0;
  float* G = nullptr;
  double* dobs = nullptr;
  double* ivar = nullptr;
  double* pa = nullptr;
  double* pb = nullptr;
  float *x0 = nullptr, *xt = nullptr, *pt = nullptr, *w = nullptr;
  double *chi0 = nullptr, *k0 = nullptr, *eps = nullptr, *u = nullptr, *h0 = nullptr, *h1 = nullptr,
         *step = nullptr, *p_host = nullptr, *partial = nullptr, *gout = nullptr, *xin = nullptr;
  int32_t* acc = nullptr;
  int64_t* acc_count = nullptr;
  SetRun* run = nullptr;
  // tiles
  int rw = 1, nw = 4;             // resid rows per wave (x16), grad waves per tile
  int rlds = 0;                   // residual tiles staged through LDS (k_set_resid_lds)
  int hrow = 0;                   // premultiplied step: 0 column strips (k_set_grad), 1 rows of H
  int kc = 64;                    // columns per LDS chunk of the staged row tiles (64 or 128)
  // premultiplied trajectories: -1 automatic, 0 off, 1 on; pm = resolved, built
  int premult = -1;
  bool pm = false, pm_built = false;
  int64_t hm_total = 0;
  float *H = nullptr, *xt2 = nullptr;
  double* hb = nullptr;
  bool tuned = false;
  int autotune = 1;
  int2 *rtiles = nullptr, *gtiles = nullptr, *htiles = nullptr, *ptiles = nullptr;
  // per step s: tiles of the problems with L > s at [off[s], off[s] + count[s])
  std::vector<int64_t> roff, rprefix, goff, gprefix, hoff, hprefix;
  int64_t n_ptiles = 0;
  // per problem (in hp) the integrator (HMC_INT_*); per problem the stage
  // count of a trajectory (L, 3L or 4L); Smax the largest
  int Smax = 0;
  std::vector<int32_t> Lcur, Scur, Soff;
  // stage windows of the problems staggered over the round's launches to
  // balance their bytes (option "stagger"); off while the tile shapes are timed
  int stagger = 1;
  bool tuning = false;
  double* coef = nullptr;           // stage coefficients (in tile_mem)
  double *minv = nullptr, *msqrt = nullptr, *lo = nullptr, *hi = nullptr;
  SetModel model() const { return SetModel{coef, minv, msqrt, lo, hi}; }
  bool finalized = false, state_set = false;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // trace
  float* tr_x = nullptr;
  double* tr_chi = nullptr;
  int64_t tr_cap = 0, tr_n = 0;
  // timing: 0 resid, 1 grad, 2 endpoint, 3 per-chain kernels
  bool timing = false;
  double t_ms[4] = {0, 0, 0, 0};
  int64_t t_n[4] = {0, 0, 0, 0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  hipStream_t stream = nullptr;
  std::vector<void*> owned;
  std::vector<void*> tile_mem;

  int64_t chains() const { return P * SC; }
};

static void set_time(ProblemSet& s, int kind, bool begin, hipEvent_t& e0) {
  if (!s.timing) return;
  if (begin) {
    HMC_HIP(hipEventCreate(&e0));
    HMC_HIP(hipEventRecord(e0, s.stream));
  } else {
    hipEvent_t e1;
    HMC_HIP(hipEventCreate(&e1));
    HMC_HIP(hipEventRecord(e1, s.stream));
    s.ev.push_back({kind, {e0, e1}});
  }
}

static void set_collect(ProblemSet& s) {
  if (s.ev.empty()) return;
  HMC_HIP(hipStreamSynchronize(s.stream));
  for (auto& e : s.ev) {
    float ms = 0.f;
    HMC_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    s.t_ms[e.first] += ms;
    s.t_n[e.first] += 1;
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  s.ev.clear();
}

static void free_tiles(ProblemSet& s) {
  for (void* p : s.tile_mem) (void)hipFree(p);
  s.tile_mem.clear();
  s.rtiles = s.gtiles = s.htiles = nullptr;
  s.coef = nullptr;
}

static void drop_graph(ProblemSet& s) {
  if (s.gexec) (void)hipGraphExecDestroy(s.gexec);
  if (s.graph) (void)hipGraphDestroy(s.graph);
  s.gexec = nullptr;
  s.graph = nullptr;
}

// Tile lists for the current L and tile shapes, one per trajectory step: the
// tiles of the problems with L > step, problems by decreasing tile cost (the
// length of the rows a tile streams), so that the longest tiles start first
// and short ones fill the end of the launch.  Step 0 covers every problem.
static void build_tiles(ProblemSet& s) {
  free_tiles(s);
  // each problem's stage schedule for its L (the chain engine's integrator
  // tables, Samplers.py:1524-1726)
  std::vector<double> coef;
  std::vector<Stage> sched;
  s.Scur.assign(s.P, 0);
  s.Smax = 0;
  for (int64_t p = 0; p < s.P; ++p) {
    integrator_schedule(s.hp[p].integ, s.Lcur[p], s.hd[p].fd, sched);
    s.hd[p].c_off = (int32_t)(coef.size() / 3);
    s.hd[p].S = (int32_t)sched.size();
    s.Scur[p] = (int32_t)sched.size();
    s.Smax = std::max(s.Smax, (int)sched.size());
    for (const Stage& g : sched) {
      coef.push_back(g.kick);
      coef.push_back(g.drift);
      coef.push_back(g.drift2);
    }
  }
  s.coef = (double*)dev_alloc(s.tile_mem, std::max<size_t>(coef.size(), 3) * 8);
  h2d(s.coef, coef.data(), coef.size() * 8);
  // stage windows [Soff_p, Soff_p + S_p) within the Smax launches of a round:
  // problems by decreasing trajectory bytes, each where the heaviest launch
  // of its window stays lightest (then the least loaded window, then the
  // earliest); the longest trajectories have no choice
  s.Soff.assign(s.P, 0);
  if (s.stagger && !s.tuning) {
    std::vector<double> w(s.P), load(s.Smax, 0.0);
    for (int64_t p = 0; p < s.P; ++p)
      w[p] = s.pm ? (double)s.hd[p].ldg * s.hd[p].ldg : 2.0 * (double)round_up(s.hp[p].n, 16) * s.hd[p].ldg;
    std::vector<int32_t> ord(s.P);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return w[a] * s.Scur[a] > w[b] * s.Scur[b]; });
    for (int32_t p : ord) {
      const int S = s.Scur[p];
      int best = 0;
      double best_max = 1e300, best_sum = 1e300;
      for (int o = 0; o + S <= s.Smax; ++o) {
        double mx = 0.0, sm = 0.0;
        for (int k = o; k < o + S; ++k) {
          mx = std::max(mx, load[k] + w[p]);
          sm += load[k];
        }
        if (mx < best_max || (mx == best_max && sm < best_sum)) {
          best_max = mx;
          best_sum = sm;
          best = o;
        }
      }
      s.Soff[p] = best;
      for (int k = best; k < best + S; ++k) load[k] += w[p];
    }
  }
  for (int64_t p = 0; p < s.P; ++p) s.hd[p].so = s.Soff[p];
  std::vector<int2> rt, gt, ht;
  const int rows = 64 * s.rw;
  auto build = [&](std::vector<int2>& out, std::vector<int64_t>& off, std::vector<int64_t>& cnt,
                   auto cost, auto ntiles) {
    std::vector<int32_t> ord(s.P);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return cost(a) > cost(b); });
    // list st < Smax: launch st of a round; list Smax: every problem
    off.assign(s.Smax + 1, 0);
    cnt.assign(s.Smax + 1, 0);
    for (int st = 0; st <= s.Smax; ++st) {
      off[st] = (int64_t)out.size();
      for (int32_t p : ord) {
        if (st < s.Smax && (st < s.Soff[p] || st >= s.Soff[p] + s.Scur[p])) continue;
        const int64_t nb = ntiles(p);
        for (int64_t b = 0; b < nb; ++b) out.push_back(make_int2(p, (int)b));
      }
      cnt[st] = (int64_t)out.size() - off[st];
    }
  };
  auto ldg = [&](int32_t p) { return (int64_t)s.hd[p].ldg; };
  // residual tiles stream rows of G (ldg long); gradient strips stream the
  // rows of G (n of them) or of H (ldg); row tiles of H stream rows of H
  build(rt, s.roff, s.rprefix, ldg, [&](int32_t p) { return (s.hp[p].n + rows - 1) / rows; });
  build(gt, s.goff, s.gprefix,
        [&](int32_t p) { return s.pm ? ldg(p) : round_up(s.hp[p].n, 16); },
        [&](int32_t p) { return (s.hp[p].d + 63) / 64; });
  build(ht, s.hoff, s.hprefix, ldg, [&](int32_t p) { return (s.hp[p].d + rows - 1) / rows; });
  s.rtiles = (int2*)dev_alloc(s.tile_mem, std::max<size_t>(rt.size(), 1) * sizeof(int2));
  s.gtiles = (int2*)dev_alloc(s.tile_mem, std::max<size_t>(gt.size(), 1) * sizeof(int2));
  s.htiles = (int2*)dev_alloc(s.tile_mem, std::max<size_t>(ht.size(), 1) * sizeof(int2));
  h2d(s.rtiles, rt.data(), rt.size() * sizeof(int2));
  h2d(s.gtiles, gt.data(), gt.size() * sizeof(int2));
  h2d(s.htiles, ht.data(), ht.size() * sizeof(int2));
  for (int64_t p = 0; p < s.P; ++p) {
    s.hd[p].L = s.Lcur[p];
    s.hd[p].mass = s.hp[p].mass ? 1 : 0;
    s.hd[p].bounded = s.hp[p].bounded ? 1 : 0;
  }
  h2d(s.desc, s.hd.data(), s.P * sizeof(SetDesc));
  drop_graph(s);
}

static void launch_grad(ProblemSet& s, float* X, int list, int step, int mode);

template <int RW, int KICK, int KC>
static void launch_rows_kc(ProblemSet& s, int64_t ntiles, const int2* tiles, const float* Gop, float* X,
                           int step) {
  hipLaunchKernelGGL((k_set_resid_lds<RW, KICK, KC>), dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc,
                     tiles, Gop, X, s.dobs, s.ivar, s.w, KICK ? s.xt2 : nullptr, KICK ? s.pt : nullptr, s.pa,
                     s.pb, KICK ? s.hb : nullptr, s.eps, step, s.model());
}

// chunk width s.kc (64 or 128 columns); 128 only for RW <= 2 (LDS size)
template <int RW, int KICK>
static void launch_rows_lds(ProblemSet& s, int64_t ntiles, const int2* tiles, const float* Gop, float* X,
                            int step) {
  if (s.kc == 128 && RW <= 2)
    launch_rows_kc<RW, KICK, (RW <= 2 ? 128 : 64)>(s, ntiles, tiles, Gop, X, step);
  else
    launch_rows_kc<RW, KICK, 64>(s, ntiles, tiles, Gop, X, step);
}

// tile list `list` (launch of a round, or s.Smax for every problem)
static void launch_resid(ProblemSet& s, int list, float* X) {
  const int64_t ntiles = s.rprefix[list];
  const int2* tiles = s.rtiles + s.roff[list];
  if (ntiles <= 0) return;
  hipEvent_t e;
  set_time(s, 0, true, e);
  if (s.rlds && s.rw == 1)
    launch_rows_lds<1, 0>(s, ntiles, tiles, s.G, X, 0);
  else if (s.rlds && s.rw == 2)
    launch_rows_lds<2, 0>(s, ntiles, tiles, s.G, X, 0);
  else if (s.rlds)
    launch_rows_lds<4, 0>(s, ntiles, tiles, s.G, X, 0);
  else if (s.rw == 1)
    hipLaunchKernelGGL(k_set_resid<1>, dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                       s.G, X, s.dobs, s.ivar, s.w);
  else if (s.rw == 2)
    hipLaunchKernelGGL(k_set_resid<2>, dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                       s.G, X, s.dobs, s.ivar, s.w);
  else
    hipLaunchKernelGGL(k_set_resid<4>, dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                       s.G, X, s.dobs, s.ivar, s.w);
  HMC_HIP(hipGetLastError());
  set_time(s, 0, false, e);
}

// Premultiplied step of the trajectory: column strips of H (k_set_grad mode 2)
// or rows of H (k_set_resid_lds with the kick epilogue), 64 * rw rows a tile.
static void launch_premult_step(ProblemSet& s, int step) {
  if (!s.hrow) {
    launch_grad(s, s.xt, step, step, 2);
    return;
  }
  const int64_t ntiles = s.hprefix[step];
  const int2* tiles = s.htiles + s.hoff[step];
  if (ntiles <= 0) return;
  hipEvent_t e;
  set_time(s, 1, true, e);
  if (s.rw == 1)
    launch_rows_lds<1, 1>(s, ntiles, tiles, s.H, s.xt, step);
  else if (s.rw == 2)
    launch_rows_lds<2, 1>(s, ntiles, tiles, s.H, s.xt, step);
  else
    launch_rows_lds<4, 1>(s, ntiles, tiles, s.H, s.xt, step);
  HMC_HIP(hipGetLastError());
  set_time(s, 1, false, e);
}

static void launch_grad(ProblemSet& s, float* X, int list, int step, int mode) {
  const int64_t ntiles = s.gprefix[list];
  const int2* tiles = s.gtiles + s.goff[list];
  if (ntiles <= 0) return;
  hipEvent_t e;
  set_time(s, 1, true, e);
  if (mode == 0) {
    if (s.nw == 4)
      hipLaunchKernelGGL((k_set_grad<4, 0>), dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                         s.G, s.w, X, s.pt, s.pa, s.pb, s.eps, step, nullptr, nullptr, nullptr, s.model(), nullptr);
    else
      hipLaunchKernelGGL((k_set_grad<8, 0>), dim3((unsigned)ntiles), dim3(512), 0, s.stream, s.desc, tiles,
                         s.G, s.w, X, s.pt, s.pa, s.pb, s.eps, step, nullptr, nullptr, nullptr, s.model(), nullptr);
  } else if (mode == 2) {
    // X is xt; the premultiplied step alternates between xt and xt2
    if (s.nw == 4)
      hipLaunchKernelGGL((k_set_grad<4, 2>), dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                         s.H, nullptr, X, s.pt, s.pa, s.pb, s.eps, step, nullptr, s.xt2, s.hb, s.model(), nullptr);
    else
      hipLaunchKernelGGL((k_set_grad<8, 2>), dim3((unsigned)ntiles), dim3(512), 0, s.stream, s.desc, tiles,
                         s.H, nullptr, X, s.pt, s.pa, s.pb, s.eps, step, nullptr, s.xt2, s.hb, s.model(), nullptr);
  } else {
    hipLaunchKernelGGL((k_set_grad<4, 1>), dim3((unsigned)ntiles), dim3(256), 0, s.stream, s.desc, tiles,
                       s.G, s.w, X, s.pt, s.pa, s.pb, s.eps, step, s.gout, nullptr, nullptr, s.model(), s.acc);
  }
  HMC_HIP(hipGetLastError());
  set_time(s, 1, false, e);
}

static void launch_endpoint(ProblemSet& s, const float* X) {
  hipEvent_t e;
  set_time(s, 2, true, e);
  hipLaunchKernelGGL(k_set_endpoint, dim3((unsigned)s.n_ptiles), dim3(256), 0, s.stream, s.desc, s.ptiles,
                     s.G, X, s.dobs, s.ivar, s.partial);
  HMC_HIP(hipGetLastError());
  set_time(s, 2, false, e);
}

// One proposal round on the stream (captured into a graph by run).
static void enqueue_proposal(ProblemSet& s, bool want_energies, bool trace) {
  const int64_t B = s.chains();
  hipEvent_t e;
  set_time(s, 3, true, e);
  hipLaunchKernelGGL(k_set_refresh, dim3((unsigned)B), dim3(SW), 0, s.stream, s.desc, s.x0, s.xt,
                     s.pm ? s.xt2 : nullptr, s.pt, s.p_host, s.k0, s.eps, s.u, s.step, s.run, s.model());
  HMC_HIP(hipGetLastError());
  set_time(s, 3, false, e);
  for (int st = 0; st < s.Smax; ++st) {
    if (s.pm) {
      launch_premult_step(s, st);
    } else {
      launch_resid(s, st, s.xt);
      launch_grad(s, s.xt, st, st, 0);
    }
  }
  launch_endpoint(s, s.xt);
  set_time(s, 3, true, e);
  hipLaunchKernelGGL(k_set_metropolis, dim3((unsigned)B), dim3(SW), 0, s.stream, s.desc, s.x0, s.xt, s.pt,
                     s.partial, s.pa, s.pb, s.chi0, s.k0, s.u, want_energies ? s.h0 : nullptr,
                     want_energies ? s.h1 : nullptr, s.acc, s.acc_count, s.run, trace ? s.tr_x : nullptr,
                     s.tr_chi, s.h_total, B, s.model());
  hipLaunchKernelGGL(k_set_advance, dim3(1), dim3(64), 0, s.stream, s.run);
  HMC_HIP(hipGetLastError());
  set_time(s, 3, false, e);
}

static void set_misfit_of(ProblemSet& s, const float* X, double* out) {
  launch_endpoint(s, X);
  hipLaunchKernelGGL(k_set_finish, dim3((unsigned)s.chains()), dim3(SW), 0, s.stream, s.desc, X, s.partial,
                     s.pa, s.pb, out, s.model());
  HMC_HIP(hipGetLastError());
}

static void finalize(ProblemSet& s) {
  if (s.finalized) return;
  for (int64_t p = 0; p < s.P; ++p)
    HMC_REQUIRE(s.hp[p].has_data, HMC_ERR_STATE, "problem " + std::to_string(p) + " has no data");
  // endpoint tiles: 64 rows each, problem order
  std::vector<int2> pt;
  for (int64_t p = 0; p < s.P; ++p) {
    s.hd[p].ptile0 = (int32_t)pt.size();
    const int64_t nr = (s.hp[p].n + 63) / 64;
    for (int64_t b = 0; b < nr; ++b) pt.push_back(make_int2((int)p, (int)b));
    s.hd[p].n_ptiles = (int32_t)nr;
    s.hd[p].prior_kind = s.hp[p].prior_kind;
    s.hd[p].mass = s.hp[p].mass ? 1 : 0;
    s.hd[p].bounded = s.hp[p].bounded ? 1 : 0;
  }
  s.n_ptiles = (int64_t)pt.size();
  s.ptiles = (int2*)dev_alloc(s.owned, pt.size() * sizeof(int2));
  h2d(s.ptiles, pt.data(), pt.size() * sizeof(int2));
  s.partial = (double*)dev_alloc(s.owned, (size_t)s.n_ptiles * SC * 8);
  h2d(s.desc, s.hd.data(), s.P * sizeof(SetDesc));
  s.finalized = true;
}

static void set_L(ProblemSet& s, const int32_t* L) {
  bool changed = s.rtiles == nullptr;
  for (int64_t p = 0; p < s.P; ++p) {
    HMC_REQUIRE(L[p] >= 1, HMC_ERR_ARG, "amount_of_steps must be >= 1 for every problem");
    if (s.Lcur[p] != L[p]) changed = true;
  }
  if (changed) {
    std::copy(L, L + s.P, s.Lcur.begin());
    build_tiles(s);
  }
}

// Resolve the trajectory form and build H, b on first use.  Automatic: on when
// one pass over every H (sum d^2) reads fewer bytes than the residual and
// gradient passes over every G (2 sum n d) and H fits in 16 GiB.
static void ensure_premult(ProblemSet& s) {
  bool want = s.premult > 0;
  if (s.premult < 0) {
    double hd = 0.0, gd = 0.0;
    for (int64_t p = 0; p < s.P; ++p) {
      hd += (double)s.hd[p].ldg * s.hd[p].ldg;
      gd += 2.0 * (double)round_up(s.hp[p].n, 16) * s.hd[p].ldg;
    }
    want = hd < gd && hd * 4.0 <= 16.0 * (1ull << 30);
  }
  if (want != s.pm) {
    s.pm = want;
    s.tuned = false;
    build_tiles(s);   // strip costs follow the form; drops the graph
  }
  if (!s.pm || s.pm_built) return;
  s.H = (float*)dev_alloc(s.owned, s.hm_total * 4);
  s.hb = (double*)dev_alloc(s.owned, s.a_total * 8);
  s.xt2 = (float*)dev_alloc(s.owned, s.x_total * 4);   // zero filled, padding included
  std::vector<int2> ht;
  for (int64_t p = 0; p < s.P; ++p) {
    const int64_t nt = (s.hd[p].ldg + 63) / 64;
    for (int64_t k = 0; k < nt * nt; ++k) ht.push_back(make_int2((int)p, (int)k));
  }
  void* tl = nullptr;
  HMC_HIP(hipMalloc(&tl, ht.size() * sizeof(int2)));
  try {
    h2d(tl, ht.data(), ht.size() * sizeof(int2));
    hipLaunchKernelGGL(k_set_build_h, dim3((unsigned)ht.size()), dim3(256), 0, s.stream, s.desc,
                       (const int2*)tl, s.G, s.ivar, s.H);
    HMC_HIP(hipGetLastError());
    // every problem's 64-column tiles (all L >= 1, so the step-0 prefix is the whole list)
    hipLaunchKernelGGL(k_set_build_b, dim3((unsigned)s.gprefix[s.Smax]), dim3(256), 0, s.stream, s.desc,
                       s.gtiles + s.goff[s.Smax],
                       s.G, s.dobs, s.ivar, s.hb);
    HMC_HIP(hipGetLastError());
    HMC_HIP(hipStreamSynchronize(s.stream));
  } catch (...) {
    (void)hipFree(tl);
    throw;
  }
  HMC_HIP(hipFree(tl));
  s.pm_built = true;
}

// Time every (rw, nw) combination on one trajectory step of this set.
static void autotune_tiles(ProblemSet& s) {
  if (s.tuned || !s.autotune) return;
  const bool saved = s.timing;
  s.timing = false;
  s.tuning = true;   // unstaggered windows: launch 0 covers every problem
  std::vector<float> xs;   // scratch states: use xt/pt copies of x0
  HMC_HIP(hipMemcpyAsync(s.xt, s.x0, s.x_total * 4, hipMemcpyDeviceToDevice, s.stream));
  HMC_HIP(hipMemsetAsync(s.pt, 0, s.x_total * 4, s.stream));
  HMC_HIP(hipMemsetAsync(s.eps, 0, s.chains() * 8, s.stream));   // zero step: states unchanged
  hipEvent_t a, b;
  HMC_HIP(hipEventCreate(&a));
  HMC_HIP(hipEventCreate(&b));
  float best = 1e30f;
  int brw = s.rw, bnw = 4, blds = s.rlds, bh = s.hrow, bkc = s.kc;
  // residual form: (rows staged or not, LDS chunk, rows per tile, waves per
  // gradient tile); premultiplied: (strips or rows of H, LDS chunk, rows per
  // tile, waves per strip)
  struct Cand {
    int lds, kc, rw, nw;
  };
  std::vector<Cand> cands;
  for (int lds : {0, 1})
    for (int kc : {64, 128})
      for (int rw : {1, 2, 4})
        for (int nw : {4, 8}) {
          if (kc == 128 && (!lds || rw > 2)) continue;
          if (s.pm && ((lds == 0 && rw != 1) || (lds == 1 && nw != 4))) continue;   // unused knobs
          cands.push_back({lds, kc, rw, nw});
        }
  for (const Cand& cd : cands) {
      const int lds = cd.lds, rw = cd.rw, nw = cd.nw;
      if (s.pm) s.hrow = lds;
      else s.rlds = lds;
      s.kc = cd.kc;
      s.rw = rw;
      s.nw = nw;
      build_tiles(s);
      for (int rep = 0; rep < 2; ++rep) {   // first pass warms caches and code
        HMC_HIP(hipEventRecord(a, s.stream));
        for (int k = 0; k < 3; ++k) {
          if (s.pm) {
            launch_premult_step(s, 0);
          } else {
            launch_resid(s, 0, s.xt);
            launch_grad(s, s.xt, 0, 0, 0);
          }
        }
        HMC_HIP(hipEventRecord(b, s.stream));
      }
      HMC_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      HMC_HIP(hipEventElapsedTime(&ms, a, b));
      if (ms < best) {
        best = ms;
        brw = rw;
        bnw = nw;
        bkc = cd.kc;
        if (s.pm) bh = lds;
        else blds = lds;
      }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  s.rlds = blds;
  s.hrow = bh;
  s.kc = bkc;
  s.rw = brw;
  s.nw = bnw;
  s.tuning = false;
  build_tiles(s);
  s.tuned = true;
  s.timing = saved;
}

}  // namespace hmc

struct hmc_set {
  hmc::ProblemSet s;
};

using namespace hmc;

extern "C" {

int hmc_set_create(int64_t problems, const int64_t* dimensions, const int64_t* data, hmc_set** out) {
  return guarded("hmc_set_create", [&] {
    HMC_REQUIRE(out && dimensions && data && problems > 0, HMC_ERR_ARG, "invalid arguments");
    for (int64_t p = 0; p < problems; ++p)
      HMC_REQUIRE(dimensions[p] > 0 && data[p] > 0 && dimensions[p] < (1 << 24) && data[p] < (1 << 24),
                  HMC_ERR_ARG, "problem sizes must be positive (and below 2^24)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      throw Error{HMC_ERR_HIP, "no HIP device is visible"};
    }
    auto* h = new hmc_set();
    ProblemSet& s = h->s;
    try {
      HMC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
      s.P = problems;
      s.hp.resize(problems);
      s.hd.resize(problems);
      s.Lcur.assign(problems, 0);
      for (int64_t p = 0; p < problems; ++p) {
        SetProblemHost& q = s.hp[p];
        q.d = dimensions[p];
        q.n = data[p];
        SetDesc& D = s.hd[p];
        std::memset(&D, 0, sizeof(D));
        D.d = (int32_t)q.d;
        D.n = (int32_t)q.n;
        D.ldg = (int32_t)round_up(q.d, 16);
        D.ldr = (int32_t)round_up(q.n, 16);
        D.g_off = s.g_total;
        s.g_total += round_up(q.n, 16) * D.ldg;
        D.x_off = s.x_total;
        s.x_total += (int64_t)SC * D.ldg;
        D.w_off = s.w_total;
        s.w_total += (int64_t)SC * D.ldr;
        D.v_off = s.v_total;
        s.v_total += q.n;
        D.a_off = s.a_total;
        s.a_total += q.d;
        D.h_off = s.h_total;
        s.h_total += (int64_t)SC * q.d;
        D.hm_off = s.hm_total;
        s.hm_total += (int64_t)D.ldg * D.ldg;
      }
      const int64_t B = s.chains();
      // G gets 64 floats of slack: the last gradient tile's guarded float4 loads stay inside
      s.G = (float*)dev_alloc(s.owned, (s.g_total + 64) * 4);
      s.dobs = (double*)dev_alloc(s.owned, s.v_total * 8);
      s.ivar = (double*)dev_alloc(s.owned, s.v_total * 8);
      s.pa = (double*)dev_alloc(s.owned, s.a_total * 8);
      s.pb = (double*)dev_alloc(s.owned, s.a_total * 8);
      s.x0 = (float*)dev_alloc(s.owned, s.x_total * 4);
      s.xt = (float*)dev_alloc(s.owned, s.x_total * 4);
      s.pt = (float*)dev_alloc(s.owned, s.x_total * 4);
      s.w = (float*)dev_alloc(s.owned, s.w_total * 4);
      s.chi0 = (double*)dev_alloc(s.owned, B * 8);
      s.k0 = (double*)dev_alloc(s.owned, B * 8);
      s.eps = (double*)dev_alloc(s.owned, B * 8);
      s.u = (double*)dev_alloc(s.owned, B * 8);
      s.h0 = (double*)dev_alloc(s.owned, B * 8);
      s.h1 = (double*)dev_alloc(s.owned, B * 8);
      s.step = (double*)dev_alloc(s.owned, B * 8);
      s.acc = (int32_t*)dev_alloc(s.owned, B * 4);
      s.acc_count = (int64_t*)dev_alloc(s.owned, B * 8);
      s.p_host = (double*)dev_alloc(s.owned, s.h_total * 8);
      s.xin = (double*)dev_alloc(s.owned, s.h_total * 8);
      s.gout = (double*)dev_alloc(s.owned, s.h_total * 8);
      s.run = (SetRun*)dev_alloc(s.owned, sizeof(SetRun));
      s.desc = (SetDesc*)dev_alloc(s.owned, problems * sizeof(SetDesc));
    } catch (...) {
      for (void* p : s.owned) (void)hipFree(p);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      delete h;
      throw;
    }
    *out = h;
  });
}

int hmc_set_destroy(hmc_set* h) {
  return guarded("hmc_set_destroy", [&] {
    if (!h) return;
    ProblemSet& s = h->s;
    (void)hipStreamSynchronize(s.stream);
    drop_graph(s);
    free_tiles(s);
    for (auto& e : s.ev) {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    for (void* p : s.owned) (void)hipFree(p);
    (void)hipStreamDestroy(s.stream);
    delete h;
  });
}

int hmc_set_problem_dense(hmc_set* h, int64_t i, const float* G, const double* d_obs, const double* variance,
                          int scalar_variance) {
  return guarded("hmc_set_problem_dense", [&] {
    HMC_REQUIRE(h && G && d_obs && variance, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(i >= 0 && i < s.P, HMC_ERR_ARG, "problem index out of range");
    HMC_REQUIRE(!s.finalized, HMC_ERR_STATE, "the set is already in use; data cannot change");
    const SetDesc& D = s.hd[i];
    const int64_t n = s.hp[i].n, d = s.hp[i].d;
    std::vector<float> g((size_t)round_up(n, 16) * D.ldg, 0.f);
    for (int64_t r = 0; r < n; ++r) std::memcpy(&g[r * D.ldg], G + r * d, d * 4);
    std::vector<double> iv(n);
    for (int64_t r = 0; r < n; ++r) {
      const double v = scalar_variance ? variance[0] : variance[r];
      HMC_REQUIRE(v > 0.0 && std::isfinite(v), HMC_ERR_ARG, "data variances must be positive and finite");
      iv[r] = 1.0 / v;
    }
    h2d(s.G + D.g_off, g.data(), g.size() * 4);
    h2d(s.dobs + D.v_off, d_obs, n * 8);
    h2d(s.ivar + D.v_off, iv.data(), n * 8);
    s.hp[i].has_data = true;
  });
}

// kind: HMC_PRIOR_NORMAL (a = means, b = inverse variances) or
// HMC_PRIOR_LAPLACE (a = means, b = inverse dispersions); one prior per problem.
int hmc_set_problem_prior(hmc_set* h, int64_t i, int kind, const double* a, const double* b) {
  return guarded("hmc_set_problem_prior", [&] {
    HMC_REQUIRE(h && a && b, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(i >= 0 && i < s.P, HMC_ERR_ARG, "problem index out of range");
    HMC_REQUIRE(kind == HMC_PRIOR_NORMAL || kind == HMC_PRIOR_LAPLACE, HMC_ERR_ARG, "unknown prior kind");
    HMC_REQUIRE(!s.finalized, HMC_ERR_STATE, "the set is already in use; priors cannot change");
    const int64_t d = s.hp[i].d;
    h2d(s.pa + s.hd[i].a_off, a, d * 8);
    h2d(s.pb + s.hd[i].a_off, b, d * 8);
    s.hp[i].prior_kind = kind;
  });
}

// Diagonal mass of problem i (MassMatrices.py:159-238): kind HMC_MASS_DIAGONAL
// with diag[d] > 0, or HMC_MASS_UNIT (diag ignored) to return to Unit mass.
int hmc_set_problem_mass(hmc_set* h, int64_t i, int kind, const double* diag) {
  return guarded("hmc_set_problem_mass", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(i >= 0 && i < s.P, HMC_ERR_ARG, "problem index out of range");
    HMC_REQUIRE(kind == HMC_MASS_UNIT || kind == HMC_MASS_DIAGONAL, HMC_ERR_ARG, "unknown mass kind");
    HMC_REQUIRE(!s.finalized, HMC_ERR_STATE, "the set is already in use; masses cannot change");
    const int64_t d = s.hp[i].d;
    if (kind == HMC_MASS_UNIT) {
      s.hp[i].mass = false;
      return;
    }
    HMC_REQUIRE(diag, HMC_ERR_ARG, "NULL mass diagonal");
    std::vector<double> mi(d), ms(d);
    for (int64_t j = 0; j < d; ++j) {
      HMC_REQUIRE(diag[j] > 0.0 && std::isfinite(diag[j]), HMC_ERR_ARG, "mass diagonal must be positive");
      mi[j] = 1.0 / diag[j];
      ms[j] = std::sqrt(diag[j]);
    }
    if (!s.minv) {
      s.minv = (double*)dev_alloc(s.owned, s.a_total * 8);
      s.msqrt = (double*)dev_alloc(s.owned, s.a_total * 8);
    }
    h2d(s.minv + s.hd[i].a_off, mi.data(), d * 8);
    h2d(s.msqrt + s.hd[i].a_off, ms.data(), d * 8);
    s.hp[i].mass = true;
  });
}

// Integrator of problem i (HMC_INT_*; Samplers.py:1524-1726): each problem is
// its own HMC.sample, so each may have its own; the set option "integrator"
// sets all of them.
int hmc_set_problem_integrator(hmc_set* h, int64_t i, int integrator) {
  return guarded("hmc_set_problem_integrator", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(i >= 0 && i < s.P, HMC_ERR_ARG, "problem index out of range");
    HMC_REQUIRE(integrator == HMC_INT_LEAPFROG || integrator == HMC_INT_3STAGE || integrator == HMC_INT_4STAGE,
                HMC_ERR_ARG, "integrator must be 0 (lf), 1 (3s) or 2 (4s)");
    if (s.hp[i].integ != integrator) {
      s.hp[i].integ = integrator;
      if (s.rtiles) build_tiles(s);
    }
  });
}

// Bounds of problem i (the BayesRule's collapsed bounds, base.py:1061-1101):
// lower / upper [d] or NULL (no bound on that side); lower < upper
// elementwise (base.py:349-359).  Both NULL removes the bounds.
int hmc_set_problem_bounds(hmc_set* h, int64_t i, const double* lower, const double* upper) {
  return guarded("hmc_set_problem_bounds", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(i >= 0 && i < s.P, HMC_ERR_ARG, "problem index out of range");
    HMC_REQUIRE(!s.finalized, HMC_ERR_STATE, "the set is already in use; bounds cannot change");
    const int64_t d = s.hp[i].d;
    if (!lower && !upper) {
      s.hp[i].bounded = false;
      return;
    }
    std::vector<double> lo(d, -INFINITY), hi(d, INFINITY);
    for (int64_t j = 0; j < d; ++j) {
      if (lower) lo[j] = lower[j];
      if (upper) hi[j] = upper[j];
      HMC_REQUIRE(!(std::isnan(lo[j]) || std::isnan(hi[j])) && lo[j] < hi[j], HMC_ERR_ARG,
                  "Bounds vectors are incompatible.");
    }
    if (!s.lo) {
      s.lo = (double*)dev_alloc(s.owned, s.a_total * 8);
      s.hi = (double*)dev_alloc(s.owned, s.a_total * 8);
    }
    h2d(s.lo + s.hd[i].a_off, lo.data(), d * 8);
    h2d(s.hi + s.hd[i].a_off, hi.data(), d * 8);
    s.hp[i].bounded = true;
  });
}

// "integrator" HMC_INT_LEAPFROG | HMC_INT_3STAGE | HMC_INT_4STAGE (every
// problem; a trajectory of L steps makes L, 3L or 4L gradient stages),
// "autotune" (1: time the tile shapes on first use, 0: keep the set ones),
// "resid_rows" (64, 128, 256 rows per residual tile), "grad_waves" (4 or 8),
// "resid_lds" (1: residual tiles staged through LDS, 0: read into registers),
// "premult" (-1 automatic, 0 residual form, 1 premultiplied trajectories),
// "premult_rows" (premultiplied step over 1: rows of H, 0: column strips),
// "lds_chunk" (64 or 128 columns per LDS chunk of the staged row tiles; 128
// applies to 64- and 128-row tiles).
int hmc_set_set_option(hmc_set* h, const char* name, double value) {
  return guarded("hmc_set_set_option", [&] {
    HMC_REQUIRE(h && name, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    const std::string k(name);
    if (k == "integrator") {
      HMC_REQUIRE(value == HMC_INT_LEAPFROG || value == HMC_INT_3STAGE || value == HMC_INT_4STAGE, HMC_ERR_ARG,
                  "integrator must be 0 (lf), 1 (3s) or 2 (4s)");
      for (auto& q : s.hp) q.integ = (int)value;
    } else if (k == "stagger") {
      HMC_REQUIRE(value == 0 || value == 1, HMC_ERR_ARG, "stagger must be 0 or 1");
      s.stagger = (int)value;
    } else if (k == "autotune") {
      s.autotune = value != 0.0;
    } else if (k == "resid_rows") {
      HMC_REQUIRE(value == 64 || value == 128 || value == 256, HMC_ERR_ARG, "resid_rows must be 64, 128 or 256");
      s.rw = (int)value / 64;
      s.autotune = 0;
    } else if (k == "premult") {
      HMC_REQUIRE(value == -1 || value == 0 || value == 1, HMC_ERR_ARG, "premult must be -1, 0 or 1");
      s.premult = (int)value;
    } else if (k == "lds_chunk") {
      HMC_REQUIRE(value == 64 || value == 128, HMC_ERR_ARG, "lds_chunk must be 64 or 128");
      s.kc = (int)value;
      s.autotune = 0;
    } else if (k == "premult_rows") {
      HMC_REQUIRE(value == 0 || value == 1, HMC_ERR_ARG, "premult_rows must be 0 or 1");
      s.hrow = (int)value;
      s.autotune = 0;
    } else if (k == "resid_lds") {
      HMC_REQUIRE(value == 0 || value == 1, HMC_ERR_ARG, "resid_lds must be 0 or 1");
      s.rlds = (int)value;
      s.autotune = 0;
    } else if (k == "grad_waves") {
      HMC_REQUIRE(value == 4 || value == 8, HMC_ERR_ARG, "grad_waves must be 4 or 8");
      s.nw = (int)value;
      s.autotune = 0;
    } else {
      throw Error{HMC_ERR_ARG, "unknown option '" + k + "'"};
    }
    if (s.rtiles) build_tiles(s);
  });
}

// Resolved values: "resid_rows", "grad_waves", "resid_lds", "autotune" and
// "premult" (1 when trajectories use H; decided at the first run or proposal).
int hmc_set_get_option(hmc_set* h, const char* name, double* value) {
  return guarded("hmc_set_get_option", [&] {
    HMC_REQUIRE(h && name && value, HMC_ERR_ARG, "NULL argument");
    const ProblemSet& s = h->s;
    const std::string k(name);
    if (k == "resid_rows") *value = 64.0 * s.rw;
    else if (k == "grad_waves") *value = s.nw;
    else if (k == "resid_lds") *value = s.rlds;
    else if (k == "premult_rows") *value = s.hrow;
    else if (k == "lds_chunk") *value = s.kc;
    else if (k == "autotune") *value = s.autotune;
    else if (k == "stagger") *value = s.stagger;
    else if (k == "premult") *value = s.pm ? 1.0 : 0.0;
    else if (k == "integrator") {
      // the common integrator, or -1 when the problems differ
      *value = s.hp[0].integ;
      for (const auto& q : s.hp)
        if (q.integ != s.hp[0].integ) *value = -1.0;
    }
    else throw Error{HMC_ERR_ARG, "unknown option '" + k + "'"};
  });
}

int hmc_set_tile_config(hmc_set* h, int* resid_rows, int* grad_waves) {
  return guarded("hmc_set_tile_config", [&] {
    HMC_REQUIRE(h && resid_rows && grad_waves, HMC_ERR_ARG, "NULL argument");
    *resid_rows = 64 * h->s.rw;
    *grad_waves = h->s.nw;
  });
}

// X: host layout, problem after problem, 16 rows (chains) of d_p values each.
int hmc_set_set_state(hmc_set* h, const double* X) {
  return guarded("hmc_set_set_state", [&] {
    HMC_REQUIRE(h && X, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    finalize(s);
    HMC_HIP(hipMemcpyAsync(s.xin, X, s.h_total * 8, hipMemcpyHostToDevice, s.stream));
    hipLaunchKernelGGL(k_set_load, dim3(8, (unsigned)s.chains()), dim3(256), 0, s.stream, s.desc, s.xin, s.x0,
                       s.h_total);
    HMC_HIP(hipGetLastError());
    set_misfit_of(s, s.x0, s.chi0);
    std::vector<double> chi(s.chains());
    HMC_HIP(hipMemcpyAsync(chi.data(), s.chi0, chi.size() * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
    for (double c : chi)
      HMC_REQUIRE(std::isfinite(c), HMC_ERR_ARG, "initial model has a non-finite misfit");
    s.state_set = true;
  });
}

int hmc_set_get_state(hmc_set* h, double* X, double* chi) {
  return guarded("hmc_set_get_state", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(s.state_set, HMC_ERR_STATE, "state not set");
    if (X) {
      hipLaunchKernelGGL(k_set_store, dim3(8, (unsigned)s.chains()), dim3(256), 0, s.stream, s.desc, s.x0,
                         s.xin);
      HMC_HIP(hipGetLastError());
      HMC_HIP(hipMemcpyAsync(X, s.xin, s.h_total * 8, hipMemcpyDeviceToHost, s.stream));
    }
    if (chi) HMC_HIP(hipMemcpyAsync(chi, s.chi0, s.chains() * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
  });
}

// misfit (P*16) and gradient (host layout) at positions X (host layout);
// positions are rounded to the fp32 state first.
int hmc_set_misfit(hmc_set* h, const double* X, double* out) {
  return guarded("hmc_set_misfit", [&] {
    HMC_REQUIRE(h && X && out, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    finalize(s);
    HMC_HIP(hipMemcpyAsync(s.xin, X, s.h_total * 8, hipMemcpyHostToDevice, s.stream));
    hipLaunchKernelGGL(k_set_load, dim3(8, (unsigned)s.chains()), dim3(256), 0, s.stream, s.desc, s.xin, s.xt,
                       s.h_total);
    set_misfit_of(s, s.xt, s.h0);
    HMC_HIP(hipMemcpyAsync(out, s.h0, s.chains() * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
  });
}

int hmc_set_gradient(hmc_set* h, const double* X, double* out) {
  return guarded("hmc_set_gradient", [&] {
    HMC_REQUIRE(h && X && out, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    finalize(s);
    if (!s.rtiles) {
      std::vector<int32_t> ones(s.P, 1);
      set_L(s, ones.data());
    }
    HMC_HIP(hipMemcpyAsync(s.xin, X, s.h_total * 8, hipMemcpyHostToDevice, s.stream));
    hipLaunchKernelGGL(k_set_load, dim3(8, (unsigned)s.chains()), dim3(256), 0, s.stream, s.desc, s.xin, s.xt,
                       s.h_total);
    // out-of-bounds flags of the chains (in the accept-flag buffer)
    hipLaunchKernelGGL(k_set_oob, dim3((unsigned)s.chains()), dim3(SW), 0, s.stream, s.desc, s.xt, s.acc,
                       s.model());
    HMC_HIP(hipGetLastError());
    launch_resid(s, s.Smax, s.xt);
    launch_grad(s, s.xt, s.Smax, 0, 1);
    HMC_HIP(hipMemcpyAsync(out, s.gout, s.h_total * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
  });
}

// One proposal with host randoms (parity with the single-chain reference loop):
// P0 host layout, eps/u per chain, L per problem.
int hmc_set_propose(hmc_set* h, const double* P0, const double* eps, const double* u, const int32_t* L,
                    int32_t* accepted, double* H0, double* H1, double* X1) {
  return guarded("hmc_set_propose", [&] {
    HMC_REQUIRE(h && P0 && eps && u && L, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(s.state_set, HMC_ERR_STATE, "state not set");
    set_L(s, L);
    ensure_premult(s);
    const int64_t B = s.chains();
    SetRun R{};
    R.host_p = 1;
    R.trace_cap = 0;
    R.prop = 0;
    HMC_HIP(hipMemcpyAsync(s.run, &R, sizeof(R), hipMemcpyHostToDevice, s.stream));
    HMC_HIP(hipMemcpyAsync(s.p_host, P0, s.h_total * 8, hipMemcpyHostToDevice, s.stream));
    HMC_HIP(hipMemcpyAsync(s.eps, eps, B * 8, hipMemcpyHostToDevice, s.stream));
    HMC_HIP(hipMemcpyAsync(s.u, u, B * 8, hipMemcpyHostToDevice, s.stream));
    enqueue_proposal(s, true, false);
    if (X1) {
      hipLaunchKernelGGL(k_set_store, dim3(8, (unsigned)B), dim3(256), 0, s.stream, s.desc, s.xt, s.xin);
      HMC_HIP(hipGetLastError());
      HMC_HIP(hipMemcpyAsync(X1, s.xin, s.h_total * 8, hipMemcpyDeviceToHost, s.stream));
    }
    if (accepted) HMC_HIP(hipMemcpyAsync(accepted, s.acc, B * 4, hipMemcpyDeviceToHost, s.stream));
    if (H0) HMC_HIP(hipMemcpyAsync(H0, s.h0, B * 8, hipMemcpyDeviceToHost, s.stream));
    if (H1) HMC_HIP(hipMemcpyAsync(H1, s.h1, B * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
    set_collect(s);
  });
}

int hmc_set_set_trace(hmc_set* h, int64_t capacity) {
  return guarded("hmc_set_set_trace", [&] {
    HMC_REQUIRE(h && capacity >= 0, HMC_ERR_ARG, "invalid argument");
    ProblemSet& s = h->s;
    HMC_HIP(hipStreamSynchronize(s.stream));
    if (s.tr_x) {
      auto it = std::find(s.owned.begin(), s.owned.end(), (void*)s.tr_x);
      if (it != s.owned.end()) s.owned.erase(it);
      (void)hipFree(s.tr_x);
      it = std::find(s.owned.begin(), s.owned.end(), (void*)s.tr_chi);
      if (it != s.owned.end()) s.owned.erase(it);
      (void)hipFree(s.tr_chi);
      s.tr_x = nullptr;
      s.tr_chi = nullptr;
    }
    s.tr_cap = capacity;
    s.tr_n = 0;
    if (capacity > 0) {
      s.tr_x = (float*)dev_alloc(s.owned, (size_t)capacity * s.h_total * 4);
      s.tr_chi = (double*)dev_alloc(s.owned, (size_t)capacity * s.chains() * 8);
    }
    drop_graph(s);
  });
}

int hmc_set_trace_count(hmc_set* h, int64_t* n) {
  return guarded("hmc_set_trace_count", [&] {
    HMC_REQUIRE(h && n, HMC_ERR_ARG, "NULL argument");
    *n = h->s.tr_n;
  });
}

// X [count][h_total] (f64, host layout), chi [count][P*16]
int hmc_set_read_trace(hmc_set* h, int64_t first, int64_t count, double* X, double* chi) {
  return guarded("hmc_set_read_trace", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    ProblemSet& s = h->s;
    HMC_REQUIRE(first >= 0 && count >= 0 && first + count <= s.tr_n, HMC_ERR_ARG, "trace range out of bounds");
    HMC_HIP(hipStreamSynchronize(s.stream));
    if (X) {
      std::vector<float> buf((size_t)s.h_total);
      for (int64_t k = 0; k < count; ++k) {
        d2h(buf.data(), s.tr_x + (first + k) * s.h_total, s.h_total * 4);
        for (int64_t j = 0; j < s.h_total; ++j) X[k * s.h_total + j] = (double)buf[j];
      }
    }
    if (chi)
      d2h(chi, s.tr_chi + first * s.chains(), count * s.chains() * 8);
  });
}

// Device-RNG proposals: every proposal is one replay of a captured graph
// (use_graph = 1) or the same launches issued directly (0).  stepsize [P*16]
// per chain; L [P]; accepted [P*16] receives the accepted count of the call.
int hmc_set_run(hmc_set* h, int64_t proposals, const int32_t* L, const double* stepsize, int randomize,
                uint64_t seed, int64_t first_proposal, int use_graph, int64_t* accepted) {
  return guarded("hmc_set_run", [&] {
    HMC_REQUIRE(h && L && stepsize && proposals >= 0, HMC_ERR_ARG, "invalid arguments");
    ProblemSet& s = h->s;
    HMC_REQUIRE(s.state_set, HMC_ERR_STATE, "state not set");
    set_L(s, L);
    ensure_premult(s);
    autotune_tiles(s);
    const int64_t B = s.chains();
    SetRun R{};
    R.seed = seed;
    R.prop = first_proposal;
    R.trace_base = first_proposal;
    R.trace_cap = s.tr_x ? s.tr_cap : 0;
    R.randomize = randomize ? 1 : 0;
    R.host_p = 0;
    HMC_HIP(hipMemcpyAsync(s.run, &R, sizeof(R), hipMemcpyHostToDevice, s.stream));
    HMC_HIP(hipMemcpyAsync(s.step, stepsize, B * 8, hipMemcpyHostToDevice, s.stream));
    HMC_HIP(hipMemsetAsync(s.acc_count, 0, B * 8, s.stream));
    const bool trace = s.tr_x != nullptr;
    if (use_graph && !s.timing) {
      if (!s.gexec) {
        HMC_HIP(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_proposal(s, false, trace);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HMC_HIP(hipStreamEndCapture(s.stream, &s.graph));
        HMC_HIP(hipGraphInstantiate(&s.gexec, s.graph, nullptr, nullptr, 0));
      }
      for (int64_t k = 0; k < proposals; ++k) HMC_HIP(hipGraphLaunch(s.gexec, s.stream));
    } else {
      for (int64_t k = 0; k < proposals; ++k) enqueue_proposal(s, false, trace);
    }
    if (accepted) HMC_HIP(hipMemcpyAsync(accepted, s.acc_count, B * 8, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
    set_collect(s);
    if (trace) s.tr_n = std::min(s.tr_cap, proposals);
  });
}

int hmc_set_enable_timing(hmc_set* h, int on) {
  return guarded("hmc_set_enable_timing", [&] {
    HMC_REQUIRE(h, HMC_ERR_ARG, "NULL argument");
    h->s.timing = on != 0;
    for (int k = 0; k < 4; ++k) {
      h->s.t_ms[k] = 0.0;
      h->s.t_n[k] = 0;
    }
  });
}

// kind 0: residual launches, 1: gradient/kick-drift launches, 2: endpoint, 3: per-chain kernels
int hmc_set_timing(hmc_set* h, int kind, double* total_ms, int64_t* launches) {
  return guarded("hmc_set_timing", [&] {
    HMC_REQUIRE(h && total_ms && launches && kind >= 0 && kind < 4, HMC_ERR_ARG, "invalid argument");
    set_collect(h->s);
    *total_ms = h->s.t_ms[kind];
    *launches = h->s.t_n[kind];
  });
}

}  // extern "C"

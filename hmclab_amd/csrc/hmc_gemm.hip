// MFMA GEMM engine: the trajectory gradient of a dense linear-Gaussian
// likelihood for a whole batch of chains as two NT GEMMs with fused
// epilogues, plus the FP64 endpoint misfit.
//
//   GEMM1  R[c][i] = ((G X^T)[i][c] - d_i) / var_i        (G: n x d, X: B x d)
//   GEMM2  g[c][j] = (Gt R^T)[j][c] + prior'(x) -> kick + drift + corrector
//
// This is the reference's residual form Gt @ ((G @ m - d) / data_variance)
// (hmclab/Distributions/LinearMatrix.py:208) with the leapfrog update of
// Samplers.py:1560-1581 applied in the GEMM2 epilogue, and the misfit
// 0.5*||(G m - d)/sigma||^2 (LinearMatrix.py:193-202) in FP64.
//
// Both operands are K-contiguous ("NT"): A rows are G (or Gt) rows, B rows
// are chain rows of X (or R).  Output element (m, n) is parameter/data index
// m of chain n; MFMA accumulators hold four m of one chain, so
// the epilogue reads/writes 4-element runs of chain-major rows (bf16: 4
// consecutive rows; f64: rows q, q+4, q+8, q+12 of a 16-row tile).
#include "hmc_internal.hpp"

namespace hmc {

// ---------------------------------------------------------------------------
// global loads: 8 contiguous elements
// ---------------------------------------------------------------------------
__device__ inline void ld8(const double* p, double o[8]) {
  const f64x2* q = reinterpret_cast<const f64x2*>(p);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const f64x2 v = q[i];
    o[2 * i] = v[0];
    o[2 * i + 1] = v[1];
  }
}
__device__ inline void ld8(const float* p, double o[8]) {
  const floatx4* q = reinterpret_cast<const floatx4*>(p);
  const floatx4 a = q[0], b = q[1];
  o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3];
  o[4] = b[0]; o[5] = b[1]; o[6] = b[2]; o[7] = b[3];
}
__device__ inline void ld8(const uint16_t* p, double o[8]) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
  uint32_t w[4] = {v[0], v[1], v[2], v[3]};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    o[2 * i] = bf16_bits_to_f32((uint16_t)(w[i] & 0xffffu));
    o[2 * i + 1] = bf16_bits_to_f32((uint16_t)(w[i] >> 16));
  }
}

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs; give each XCD a contiguous range of tiles (m-major) so tiles
// sharing rows of A run on the same L2.
// order 1 deals n-major ranges instead (tiles sharing columns of B, i.e. the
// same chains, run on the same L2).
__device__ inline bool map_tile(int gm, int gn, int& tm, int& tn, int order = 0) {
  const int T = gm * gn;
  const int chunk = (T + 7) / 8;
  const int L = blockIdx.x;
  const int t = (L % 8) * chunk + (L / 8);
  if (t >= T) return false;
  if (order == 0) {
    tm = t / gn;
    tn = t % gn;
  } else {
    tm = t % gm;
    tn = t / gm;
  }
  return true;
}
static inline int grid_for(int gm, int gn) { return ((gm * gn + 7) / 8) * 8; }

// Split tail.  T tiles on n_cu compute units leave a last round of T mod n_cu
// tiles that runs on a fraction of the chip.  The first F tiles (whole rounds,
// F a multiple of 8 and of gn) run over all of K; each of the T - F tail tiles
// is cut into SPL contiguous k-ranges, one workgroup each, that store float32
// partial tiles; k_split_epilogue sums the SPL partials of every tail element
// in fixed order and applies the epilogue.  Workgroup u < F takes full tile
// (u % 8) * (F / 8) + u / 8 (XCD-contiguous, as map_tile); the tail units are
// dealt the same way, the gn chain tiles of one (row block, k-range) adjacent.
struct SplitPlan {
  int F = 0, SPL = 1, grid = 0;
  int64_t TR = 0;   // rows of a partial plane (tail row blocks x BM)
  int64_t mt0 = 0;  // first tail row
  bool split() const { return SPL > 1; }
};

// min_piece: fewest 64-wide k-tiles per piece; max_spl: most pieces per tile
static SplitPlan plan_split(int BM, int BN, int M, int N, int K, int n_cu, bool allow, int BK = 64,
                            int min_piece = 8, int max_spl = 8) {
  SplitPlan s;
  const int gm = (M + BM - 1) / BM, gn = (N + BN - 1) / BN, T = gm * gn;
  const int nk = (K + BK - 1) / BK * BK / 64;
  s.F = T;
  s.grid = grid_for(gm, gn);
  if (!allow || n_cu <= 0) return s;
  int L = 8;
  while (L % gn) L += 8;                       // lcm(8, gn)
  const int F = ((T / n_cu) * n_cu / L) * L;
  const int tail = T - F;
  if (tail <= 0) return s;
  int spl = n_cu / tail;
  spl = spl > max_spl ? max_spl : spl;
  spl = spl > nk / min_piece ? nk / min_piece : spl;
  if (spl < 2) return s;
  s.F = F;
  s.SPL = spl;
  s.mt0 = (int64_t)(F / gn) * BM;
  s.TR = (int64_t)(gm - F / gn) * BM;
  s.grid = F + ((tail * spl + 7) / 8) * 8;
  return s;
}

// order 1 with SPL dividing 8: XCD x takes only k-range x % SPL of the tail
// tiles (8 / SPL XCDs per k-range, each a contiguous range of tiles), so an
// XCD's L2 holds one k-range of the shared B operand instead of all of them.
__device__ inline bool split_job(int gm, int gn, int nk, int F, int SPL, int& tm, int& tn, int& kt0, int& kt1,
                                 int& sp, int order = 0) {
  const int T = gm * gn;
  const int u = blockIdx.x;
  if (u < F) {
    const int t = (u % 8) * (F / 8) + u / 8;
    tm = t / gn;
    tn = t % gn;
    return true;
  }
  const int units = (T - F) * SPL, chunk = (units + 7) / 8, v = u - F;
  if (order == 1 && 8 % SPL == 0) {
    const int x = v % 8, nxs = 8 / SPL, per = (T - F + nxs - 1) / nxs;
    const int t = (x / SPL) * per + v / 8;
    if (v / 8 >= per || t >= T - F) return false;
    tm = F / gn + t / gn;
    tn = t % gn;
    sp = x % SPL;
  } else {
    const int t2 = (v % 8) * chunk + v / 8;
    if (t2 >= units) return false;
    const int g = t2 / gn;
    tn = t2 % gn;
    tm = F / gn + g / SPL;
    sp = g % SPL;
  }
  kt0 = sp * nk / SPL;
  kt1 = (sp + 1) * nk / SPL;
  return true;
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM (v_mfma_f64_16x16x4f64): C = A * Bt^T, block tile BM x BN,
// k-tile 16, wave tile WM x WN.  K is permuted inside a k-tile: lane group
// q at MFMA step s uses k = 4q + s, so each lane reads 4 contiguous doubles.
// Requires A rows up to the tile edge and K padded to 16 (zeros).
// ---------------------------------------------------------------------------
template <int BM, int BN, int WM, int WN, typename TA, typename TB, class Epi>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void k_gemm_f64(
    const TA* __restrict__ A, int64_t lda, const TB* __restrict__ Bt, int64_t ldb, int M, int N,
    int K, int F, int SPL, double* __restrict__ part, int64_t TR, Epi epi) {
  constexpr int BK = 16, LDS = BK + 2;
  constexpr int NT = (BM / WM) * (BN / WN) * 64;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int WPR = BN / WN;  // waves per row of waves
  __shared__ double As[BM * LDS];
  __shared__ double Bs[BN * LDS];

  const int gm = (M + BM - 1) / BM, gn = (N + BN - 1) / BN;
  int tm, tn, kt0 = 0, kt1 = (K + BK - 1) / BK, sp = -1;
  if (F >= gm * gn) {
    if (!map_tile(gm, gn, tm, tn)) return;
  } else if (!split_job(gm, gn, kt1, F, SPL, tm, tn, kt0, kt1, sp)) {
    return;
  }
  const int64_t m_blk = (int64_t)tm * BM, n_blk = (int64_t)tn * BN;
  A += (int64_t)kt0 * BK;
  Bt += (int64_t)kt0 * BK;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w / WPR, wn = w % WPR;
  const int q = lane >> 4, r = lane & 15;

  // global -> LDS mapping: rows x 16 cols in chunks of 8 elements
  constexpr int A_CH = BM * 2 / NT;  // chunks per thread (A)
  constexpr int B_CH = BN * 2 / NT;
  static_assert(A_CH >= 1 && B_CH >= 1, "tile too small for thread count");

  double ra[A_CH][8], rb[B_CH][8];
  auto gload = [&](int kt) {
#pragma unroll
    for (int i = 0; i < A_CH; ++i) {
      const int c = tid + i * NT, row = c >> 1, col = (c & 1) * 8;
      ld8(A + (m_blk + row) * lda + (int64_t)kt * BK + col, ra[i]);
    }
#pragma unroll
    for (int i = 0; i < B_CH; ++i) {
      const int c = tid + i * NT, row = c >> 1, col = (c & 1) * 8;
      ld8(Bt + (n_blk + row) * ldb + (int64_t)kt * BK + col, rb[i]);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < A_CH; ++i) {
      const int c = tid + i * NT, row = c >> 1, col = (c & 1) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) As[row * LDS + col + e] = ra[i][e];
    }
#pragma unroll
    for (int i = 0; i < B_CH; ++i) {
      const int c = tid + i * NT, row = c >> 1, col = (c & 1) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) Bs[row * LDS + col + e] = rb[i][e];
    }
  };

  doublex4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};

  const int nk = kt1 - kt0;
  gload(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (kt + 1 < nk) gload(kt + 1);
    double af[TM][4], bfr[TN][4];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const double* src = &As[(wm * WM + i * 16 + r) * LDS + 4 * q];
#pragma unroll
      for (int s = 0; s < 4; ++s) af[i][s] = src[s];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const double* src = &Bs[(wn * WN + j * 16 + r) * LDS + 4 * q];
#pragma unroll
      for (int s = 0; s < 4; ++s) bfr[j][s] = src[s];
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][s], bfr[j][s], acc[i][j], 0, 0, 0);
  }

  // D layout (measured on gfx950): lane (q, r) holds rows q, q+4, q+8, q+12 of
  // column r of each 16x16 tile
  if (sp >= 0) {
    // partial tile of k-range sp: part[sp][chain][row - first tail row]
    const int64_t mt0 = (int64_t)(F / gn) * BM;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int64_t m0 = m_blk + wm * WM + i * 16 + q;
        const int64_t n = n_blk + wn * WN + j * 16 + r;
        if (n >= N) continue;
        double* dst = part + ((int64_t)sp * N + n) * TR + (m0 - mt0);
#pragma unroll
        for (int v = 0; v < 4; ++v) dst[4 * v] = acc[i][j][v];
      }
    return;
  }
  if constexpr (Epi::kReduce) {
    // per-column sums over the block's rows, deterministic order
    __shared__ double red[(BM / WM)][4][BN];
    double colsum[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      colsum[j] = 0.0;
      const int64_t n = n_blk + wn * WN + j * 16 + r;
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int64_t m0 = m_blk + wm * WM + i * 16 + q;
#pragma unroll
        for (int v = 0; v < 4; ++v) colsum[j] += epi.term(m0 + 4 * v, n, acc[i][j][v]);
      }
      red[wm][q][wn * WN + j * 16 + r] = colsum[j];
    }
    __syncthreads();
    for (int col = tid; col < BN; col += NT) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < BM / WM; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) s += red[a][b][col];
      epi.store_partial(tm, n_blk + col, s);
    }
  } else {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int64_t m0 = m_blk + wm * WM + i * 16 + q;
        const int64_t n = n_blk + wn * WN + j * 16 + r;
        double v[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
        epi.apply(m0, 4, n, v);
      }
  }
}

// Split-tail epilogues of the FP64 GEMM.  Element-wise epilogues: sum the
// SPL partials of 4 consecutive rows of one chain in fixed order, then apply.
template <class Epi>
__global__ __launch_bounds__(256) void k_split_epilogue_f64(const double* __restrict__ part, int SPL, int64_t TR,
                                                            int N, int64_t mt0, Epi epi) {
  const int64_t per_chain = TR / 4, total = (int64_t)N * per_chain;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(i / per_chain);
    const int64_t ml = 4 * (i % per_chain);
    double v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = part[(int64_t)n * TR + ml + e];
    for (int k = 1; k < SPL; ++k)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] += part[((int64_t)k * N + n) * TR + ml + e];
    epi.apply(mt0 + ml, 1, n, v);
  }
}

// Reducing epilogue (endpoint misfit): one wavefront per (128-row block,
// chain) sums the SPL partials of each row, applies term() and reduces the
// block's rows with a fixed butterfly.
template <class Epi>
__global__ __launch_bounds__(256) void k_split_reduce_f64(const double* __restrict__ part, int SPL, int64_t TR,
                                                          int N, int64_t mt0, Epi epi) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int rb = blockIdx.y;
  if (n >= N) return;
  double s = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t row = (int64_t)rb * 128 + lane + 64 * h;
    double v = part[(int64_t)n * TR + row];
    for (int k = 1; k < SPL; ++k) v += part[((int64_t)k * N + n) * TR + row];
    s += epi.term(mt0 + row, n, v);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) epi.store_partial((int)(mt0 / 128) + rb, n, s);
}

// ---------------------------------------------------------------------------
// BF16 MFMA GEMM (v_mfma_f32_32x32x16_bf16, gfx950): C = A * (B0 + B1)^T
// with NSPLIT = 1 or 2 bf16 planes of B (hi/lo split of an FP32 operand),
// FP32 accumulation.  Block tile BM x BN, k-tile 64 (one 128-byte row per
// operand row), wave tile WM x WN.
//
// Operand tiles stream global -> LDS with gfx950's direct-to-LDS loads
// (global_load_lds_dwordx4: each wavefront instruction fills 1 KB = 8 rows)
// through an S-stage ring, so S-1 k-tiles are in flight without register
// staging; each wave waits only for its own loads of the tile it is about to
// use (s_waitcnt vmcnt), then one workgroup barrier per k-tile.  Rows are
// stored with their 16-byte chunks XOR-swizzled by (row & 7) so the
// ds_read_b128 fragment loads of 32 consecutive rows are bank-conflict free.
// K permutation: lane group q (0/1) at MFMA step s uses k = 32q + 8s + e, i.e.
// logical chunk 4q + s of the row.
// ---------------------------------------------------------------------------
#define HMC_LDS __attribute__((address_space(3)))

// s_waitcnt vmcnt(n) leaving expcnt / lgkmcnt unconstrained (gfx9 encoding)
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt out of range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// workgroup barrier without the release/acquire fences of __syncthreads():
// LDS-DMA completion is ordered explicitly with wait_vmcnt.
__device__ inline void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <int BM, int BN, int WM, int WN, int NSPLIT, int S, class Epi>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void k_gemm_bf16(
    const uint16_t* __restrict__ A, int64_t lda, const uint16_t* __restrict__ B0,
    const uint16_t* __restrict__ B1, int64_t ldb, int M, int N, int K, int order, int F, int SPL,
    float* __restrict__ part, int64_t TR, int* __restrict__ fix, Epi epi) {
  constexpr int BK = 64;                 // elements per row per k-tile (128 bytes)
  constexpr int NW = (BM / WM) * (BN / WN);
  constexpr int TM = WM / 32, TN = WN / 32;
  constexpr int WPR = BN / WN;
  constexpr int A_GROUPS = BM / 8, B_GROUPS = NSPLIT * BN / 8;      // 1 KB groups per stage
  constexpr int GROUPS = A_GROUPS + B_GROUPS;
  static_assert(GROUPS % NW == 0, "load groups must divide over waves");
  constexpr int GPW = GROUPS / NW;       // load instructions per wave per k-tile
  constexpr int STAGE = (BM + NSPLIT * BN) * BK;   // elements per stage
  static_assert(GPW * (S - 2) < 64, "vmcnt range");
  // the k ring, reused by the epilogue for the block's FP32 tile C[BN][BM + 4]
  constexpr int CLD = BM + 4;
  constexpr int SMEM = (S * STAGE > BN * CLD * 2) ? S * STAGE : BN * CLD * 2;
  static_assert(SMEM * 2 <= 160 * 1024, "LDS budget");
  __shared__ __attribute__((aligned(1024))) uint16_t smem[SMEM];

  const int gm = (M + BM - 1) / BM, gn = (N + BN - 1) / BN;
  int tm, tn, kt0 = 0, kt1 = (K + BK - 1) / BK, sp = -1;
  if (F >= gm * gn) {
    if (!map_tile(gm, gn, tm, tn, order)) return;
  } else if (!split_job(gm, gn, kt1, F, SPL, tm, tn, kt0, kt1, sp, order)) {
    return;
  }
  const int64_t m_blk = (int64_t)tm * BM, n_blk = (int64_t)tn * BN;
  A += (int64_t)kt0 * BK;
  B0 += (int64_t)kt0 * BK;
  if (NSPLIT == 2) B1 += (int64_t)kt0 * BK;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w / WPR, wn = w % WPR;
  const int q = lane >> 5, r = lane & 31;

  // per-lane part of the load address: row within group, swizzled chunk
  const int lrow = lane >> 3;
  auto issue = [&](int kt, int st) {
    uint16_t* stage = smem + st * STAGE;
#pragma unroll
    for (int i = 0; i < GPW; ++i) {
      const int g = w * GPW + i;
      const uint16_t* src;
      uint16_t* dst;
      if (g < A_GROUPS) {
        const int row = g * 8 + lrow;
        const int c = (lane & 7) ^ (row & 7);
        src = A + (m_blk + row) * lda + (int64_t)kt * BK + c * 8;
        dst = stage + g * 8 * BK;
      } else {
        const int gb = g - A_GROUPS;
        const int sp = gb / (BN / 8), gr = gb % (BN / 8);
        const int row = gr * 8 + lrow;
        const int c = (lane & 7) ^ (row & 7);
        const uint16_t* Bp = (NSPLIT == 2 && sp == 1) ? B1 : B0;
        src = Bp + (n_blk + row) * ldb + (int64_t)kt * BK + c * 8;
        dst = stage + BM * BK + sp * BN * BK + gr * 8 * BK;
      }
      __builtin_amdgcn_global_load_lds((const void*)src, (HMC_LDS void*)dst, 16, 0, 0);
    }
  };

  floatx16 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;

  const int nk = kt1 - kt0;
#pragma unroll
  for (int s0 = 0; s0 < S - 1; ++s0)
    if (s0 < nk) issue(s0, s0);
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's loads of tile kt are done once at most (S-2) later tiles are pending
    if (kt + S - 2 < nk) wait_vmcnt<GPW * (S - 2)>();
    else wait_vmcnt<0>();
    lds_barrier();
    if (kt + S - 1 < nk) issue(kt + S - 1, (kt + S - 1) % S);
    const uint16_t* As = smem + (kt % S) * STAGE;
    const uint16_t* Bs = As + BM * BK;
    // fragments of k-step s+1 are read from LDS while the MFMAs of step s run
    bf16x8 af[2][TM], bfr[2][NSPLIT][TN];
    auto load_frags = [&](int s, int buf) {
      const int c = 4 * q + s;
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int row = wm * WM + i * 32 + r;
        af[buf][i] = *reinterpret_cast<const bf16x8*>(&As[row * BK + ((c ^ (row & 7)) << 3)]);
      }
#pragma unroll
      for (int sp = 0; sp < NSPLIT; ++sp)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int row = wn * WN + j * 32 + r;
          bfr[buf][sp][j] =
              *reinterpret_cast<const bf16x8*>(&Bs[sp * BN * BK + row * BK + ((c ^ (row & 7)) << 3)]);
        }
    };
    load_frags(0, 0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (s < 3) load_frags(s + 1, (s + 1) & 1);
#pragma unroll
      for (int sp = 0; sp < NSPLIT; ++sp)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[s & 1][i], bfr[s & 1][sp][j],
                                                                acc[i][j], 0, 0, 0);
    }
  }

  // Epilogue through LDS.  D layout (32x32): lane (q, r) holds rows 8g + 4q + v
  // (g, v = 0..3) of column r, so direct stores would scatter 16-byte pieces
  // over 32 chains.  The block's tile goes to LDS as C[n][m] (the ring is free
  // once every wave has left the k loop), then each thread applies the
  // epilogue to 4 consecutive rows of one chain: a wavefront covers 256
  // consecutive rows of one chain row in the chain-major outputs.
  float* Cs = reinterpret_cast<float*>(smem);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int ml = wm * WM + i * 32 + 8 * g + 4 * q;
        const int nl = wn * WN + j * 32 + r;
        *reinterpret_cast<floatx4*>(&Cs[nl * CLD + ml]) =
            floatx4{acc[i][j][4 * g], acc[i][j][4 * g + 1], acc[i][j][4 * g + 2], acc[i][j][4 * g + 3]};
      }
  __syncthreads();
  if (sp < 0) {
    for (int idx = tid; idx < BN * (BM / 4); idx += NW * 64) {
      const int nl = idx / (BM / 4), ml = 4 * (idx % (BM / 4));
      const floatx4 c = *reinterpret_cast<const floatx4*>(&Cs[nl * CLD + ml]);
      const float v[4] = {c[0], c[1], c[2], c[3]};
      epi.apply(m_blk + ml, 1, n_blk + nl, v);
    }
  } else {
    // partial tile of k-range sp: part[sp][chain][row - first tail row]
    const int64_t mt0 = (int64_t)(F / gn) * BM;
    auto store_partial = [&]() {
      for (int idx = tid; idx < BN * (BM / 4); idx += NW * 64) {
        const int nl = idx / (BM / 4), ml = 4 * (idx % (BM / 4));
        const int64_t n = n_blk + nl;
        if (n >= N) continue;
        *reinterpret_cast<floatx4*>(&part[((int64_t)sp * N + n) * TR + (m_blk - mt0) + ml]) =
            *reinterpret_cast<const floatx4*>(&Cs[nl * CLD + ml]);
      }
    };
    if (fix == nullptr) {
      store_partial();
      return;
    }
    // Fix-up of a tile cut into two k-ranges (SPL = 2): the two workgroups
    // take roles from the tile's counter (the first gets 0, the second 1 or
    // 2).  The first stores its partial, publishes it at agent scope and adds
    // 1 more, so the counter reaches 3 once both have arrived and the partial
    // is visible, in either order.  The second waits for 3 (the first is
    // already running, so the wait ends; a bounded spin sets fix[0] instead
    // of hanging), resets the counter for the next launch and applies the
    // epilogue to partial 0 + partial 1 — the split epilogue kernel's sum.
    __shared__ int s_role;
    int* cnt = fix + 1 + ((tm - F / gn) * gn + tn);
    if (tid == 0) s_role = __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    // (Fences: release only in the first, one acquire after the wait in the
    // second; the spin itself reads the counter relaxed, so it does not
    // invalidate the caches the other tiles of the XCD are streaming H
    // through.)
    if (s_role == 0) {
      store_partial();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (tid == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    if (tid == 0) {
      int spins = 0;
      while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 3) {
        if (++spins > (1 << 22)) {
          __hip_atomic_fetch_or(fix, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int other = 1 - sp;
    for (int idx = tid; idx < BN * (BM / 4); idx += NW * 64) {
      const int nl = idx / (BM / 4), ml = 4 * (idx % (BM / 4));
      const int64_t n = n_blk + nl;
      if (n >= N) continue;
      const floatx4 own = *reinterpret_cast<const floatx4*>(&Cs[nl * CLD + ml]);
      const floatx4 oth = *reinterpret_cast<const floatx4*>(&part[((int64_t)other * N + n) * TR + (m_blk - mt0) + ml]);
      const floatx4 a = sp == 0 ? own : oth, b = sp == 0 ? oth : own;
      const float v[4] = {a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]};
      epi.apply(m_blk + ml, 1, n, v);
    }
  }
}

// sum of the SPL partial planes of the tail rows (fixed order), then the
// epilogue on 4 consecutive rows of one chain
template <class Epi>
__global__ __launch_bounds__(256) void k_split_epilogue(const float* __restrict__ part, int SPL, int64_t TR, int N,
                                                        int64_t mt0, Epi epi) {
  const int64_t per_chain = TR / 4, total = (int64_t)N * per_chain;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(i / per_chain);
    const int64_t ml = 4 * (i % per_chain);
    floatx4 acc = *reinterpret_cast<const floatx4*>(&part[(int64_t)n * TR + ml]);
    for (int k = 1; k < SPL; ++k) {
      const floatx4 c = *reinterpret_cast<const floatx4*>(&part[((int64_t)k * N + n) * TR + ml]);
      acc[0] += c[0];
      acc[1] += c[1];
      acc[2] += c[2];
      acc[3] += c[3];
    }
    const float v[4] = {acc[0], acc[1], acc[2], acc[3]};
    epi.apply(mt0 + ml, 1, n, v);
  }
}

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
// residual: out[n][m] = (acc - d_m) / var_m   (m < M, n < N)
template <typename T>
struct EpiResid {
  static constexpr bool kReduce = false;
  T* out;           // state dtype [N_pad][ld]
  uint16_t* oh;     // bf16 planes (or nullptr)
  uint16_t* ol;
  int64_t ld;
  const double* dobs;
  const double* var;
  int M, N;
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t m = m0 + (int64_t)e * ms;
      if (m >= M) return;
      const T rr = ((T)v[e] - (T)dobs[m]) / (T)var[m];
      const int64_t o = n * ld + m;
      if (out) out[o] = rr;
      if (oh) {
        const float f = (float)rr;
        const uint16_t h = f32_to_bf16_bits(f);
        oh[o] = h;
        ol[o] = f32_to_bf16_bits(f - bf16_bits_to_f32(h));
      }
    }
  }
};

// anchored residual (bf16 trajectories): r = rref_m + acc with acc = G_hi .
// bf16(x - x_ref); writes the bf16 hi/lo planes of r/var (ol null: the
// gradient GEMM reads the hi plane only)
struct EpiResidAnchor {
  static constexpr bool kReduce = false;
  const float* rref;   // [M] G x_ref - d (float32 of an FP64 product)
  uint16_t* oh;
  uint16_t* ol;
  int64_t ld;
  const double* var;
  int M, N;
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
    if (ms == 1 && m0 + 3 < M) {
      // 4 consecutive rows of one chain: one 8-byte store per bf16 plane
      // (o is a multiple of 4)
      const int64_t o = n * ld + m0;
      const floatx4 r4 = *reinterpret_cast<const floatx4*>(&rref[m0]);
      uint32_t h[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float f = (r4[e] + (float)v[e]) / (float)var[m0 + e];
        h[e] = f32_to_bf16_bits(f);
        l[e] = f32_to_bf16_bits(f - bf16_bits_to_f32((uint16_t)h[e]));
      }
      *reinterpret_cast<uint2*>(&oh[o]) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
      if (ol) *reinterpret_cast<uint2*>(&ol[o]) = uint2{l[0] | (l[1] << 16), l[2] | (l[3] << 16)};
      return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t m = m0 + (int64_t)e * ms;
      if (m >= M) return;
      const int64_t o = n * ld + m;
      const float f = (rref[m] + (float)v[e]) / (float)var[m];
      const uint16_t h = f32_to_bf16_bits(f);
      oh[o] = h;
      if (ol) ol[o] = f32_to_bf16_bits(f - bf16_bits_to_f32(h));
    }
  }
};

// trajectory step: g = acc [+ yref_j] [- bias_j] + prior'(x); p -= kick*eps*g;
// x += drift*eps*kgrad(p); corrector; optional second drift + corrector.
// The bf16 planes written for the next GEMM hold x (hi/lo), or x - xa (one
// plane) for anchored trajectories.
template <typename T>
struct EpiLeap {
  static constexpr bool kReduce = false;
  T* x;
  T* p;
  uint16_t* xh;
  uint16_t* xl;
  int64_t ld;
  const double* eps;
  const double* beta;   // per-chain inverse temperature scaling the kick (null: 1)
  double kick, drift, drift2;
  DevModel md;
  int M, N;
  const float* xa;      // anchor: planes hold x - xa (or null)
  const float* yref;    // premultiplied bf16: g = yref + acc (or null)
  const double* bias;   // premultiplied FP64 trajectories: g = acc - bias[j] (or null)
  const float* gscale;  // surrogate: g = gscale[j] * acc (or null)
  const float* xsinv;   // planes hold (x - xa) * xsinv (surrogate's normalised input) (or null)
  __device__ __forceinline__ void step(int64_t j, T& xv, T& pv, T g, T kk, T dd, T dd2) const {
    g += prior_grad<T>(md, j, xv);
    if (out_of_bounds(md, j, (double)xv)) g = (T)INFINITY;
    pv -= kk * g;
    xv += dd * kgrad<T>(md, j, pv);
    corrector<T>(md, j, xv, pv);
    if (drift2 != 0.0) {
      xv += dd2 * kgrad<T>(md, j, pv);
      corrector<T>(md, j, xv, pv);
    }
  }
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
    const double e = eps[n];
    const T kk = (T)(beta ? kick * e * beta[n] : kick * e), dd = (T)(drift * e), dd2 = (T)(drift2 * e);
    if constexpr (sizeof(T) == 4) {
      if (ms == 1 && m0 + 3 < M) {
        // 4 consecutive parameters of one chain: 16-byte accesses of x and p,
        // one 8-byte store per bf16 plane (o is a multiple of 4)
        const int64_t o = n * ld + m0;
        floatx4 x4 = *reinterpret_cast<const floatx4*>(&x[o]);
        floatx4 p4 = *reinterpret_cast<const floatx4*>(&p[o]);
        floatx4 g4 = {(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
        if (gscale) g4 *= *reinterpret_cast<const floatx4*>(&gscale[m0]);
        if (yref) g4 += *reinterpret_cast<const floatx4*>(&yref[m0]);
        floatx4 a4 = {0.f, 0.f, 0.f, 0.f};
        floatx4 s4 = {1.f, 1.f, 1.f, 1.f};
        if (xa) a4 = *reinterpret_cast<const floatx4*>(&xa[m0]);
        if (xsinv) s4 = *reinterpret_cast<const floatx4*>(&xsinv[m0]);
        uint32_t h[4], l[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          T xv = x4[k], pv = p4[k];
          T g = (T)g4[k];
          if (bias) g -= (T)bias[m0 + k];
          step(m0 + k, xv, pv, g, kk, dd, dd2);
          x4[k] = xv;
          p4[k] = pv;
          const float f = ((float)xv - a4[k]) * s4[k];
          h[k] = f32_to_bf16_bits(f);
          l[k] = f32_to_bf16_bits(f - bf16_bits_to_f32((uint16_t)h[k]));
        }
        *reinterpret_cast<floatx4*>(&x[o]) = x4;
        *reinterpret_cast<floatx4*>(&p[o]) = p4;
        if (xh) {
          *reinterpret_cast<uint2*>(&xh[o]) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
          if (xl) *reinterpret_cast<uint2*>(&xl[o]) = uint2{l[0] | (l[1] << 16), l[2] | (l[3] << 16)};
        }
        return;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t j = m0 + (int64_t)k * ms;
      if (j >= M) return;
      const int64_t o = n * ld + j;
      T xv = x[o], pv = p[o];
      T g = (T)v[k];
      if (gscale) g = (T)((float)v[k] * gscale[j]);
      if (yref) g = (T)((float)v[k] + yref[j]);
      if (bias) g -= (T)bias[j];
      step(j, xv, pv, g, kk, dd, dd2);
      x[o] = xv;
      p[o] = pv;
      if (xh) {
        float f = xa ? (float)xv - xa[j] : (float)xv;
        if (xsinv) f *= xsinv[j];
        const uint16_t h = f32_to_bf16_bits(f);
        xh[o] = h;
        if (xl) xl[o] = f32_to_bf16_bits(f - bf16_bits_to_f32(h));
      }
    }
  }
};

// surrogate hidden layer of chain n, unit m: z = acc + bias_m, s = sigmoid(z);
// mode 0: bf16 plane of silu(z) = z s and silu'(z) = s (1 + z (1 - s)) in
// float; mode 1: bf16 plane of silu'(z) w3_m
struct EpiSurAct {
  static constexpr bool kReduce = false;
  const float* bias;
  const float* w3;
  uint16_t* plane;
  float* sd;
  int64_t ld;
  int mode;
  int M, N;
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t m = m0 + (int64_t)e * ms;
      if (m >= M) return;
      const int64_t o = n * ld + m;
      const float z = (float)v[e] + bias[m];
      const float sg = 1.f / (1.f + __expf(-z));
      const float d1 = sg * (1.f + z * (1.f - sg));
      if (mode == 0) {
        plane[o] = f32_to_bf16_bits(z * sg);
        sd[o] = d1;
      } else {
        plane[o] = f32_to_bf16_bits(d1 * w3[m]);
      }
    }
  }
};

// surrogate backward product: bf16 plane of sd[n][m] * acc
struct EpiSurMul {
  static constexpr bool kReduce = false;
  const float* sd;
  uint16_t* plane;
  int64_t ld;
  int M, N;
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t m = m0 + (int64_t)e * ms;
      if (m >= M) return;
      const int64_t o = n * ld + m;
      plane[o] = f32_to_bf16_bits(sd[o] * (float)v[e]);
    }
  }
};

// gradient output: out[n][m] = acc [+ yref_m] [- bias_m] + prior'(x) (+inf for a
// chain out of bounds), in the arithmetic of EpiLeap's stage gradient
template <typename T>
struct EpiGrad {
  static constexpr bool kReduce = false;
  const T* x;
  double* out;
  int64_t ld;
  const int32_t* oob;
  DevModel md;
  int M, N;
  const float* yref;    // anchored premultiplied bf16 field (or null)
  const double* bias;   // premultiplied FP64 field (or null)
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t j = m0 + (int64_t)k * ms;
      if (j >= M) return;
      const int64_t o = n * ld + j;
      T gv = (T)v[k];
      if (yref) gv = (T)((float)v[k] + yref[j]);
      if (bias) gv -= (T)bias[j];
      double g = (double)(gv + prior_grad<T>(md, j, x[o]));
      if (oob[n]) g += INFINITY;
      out[o] = g;
    }
  }
};

// endpoint misfit partial sums: sum_m (acc - d_m)^2 / var_m over the block rows
struct EpiMisfit {
  static constexpr bool kReduce = true;
  const double* dobs;
  const double* var;
  double* partial;   // [gm][ldp]
  int64_t ldp;
  int M, N;
  __device__ double term(int64_t m, int64_t n, double v) const {
    if (m >= M || n >= N) return 0.0;
    const double rr = v - dobs[m];
    return rr * rr / var[m];
  }
  __device__ void store_partial(int tm, int64_t n, double s) const {
    if (n < N) partial[(int64_t)tm * ldp + n] = s;
  }
};

// plain FP64 store out[n][m] = acc (m < M, n < N)
struct EpiStoreF64 {
  static constexpr bool kReduce = false;
  double* out;
  int64_t ld;
  int M, N;
  template <typename V>
  __device__ __forceinline__ void apply(int64_t m0, int ms, int64_t n, const V v[4]) const {
    if (n >= N) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t m = m0 + (int64_t)e * ms;
      if (m >= M) return;
      out[n * ld + m] = (double)v[e];
    }
  }
};

// premultiplied endpoint: with y = H x (acc), partial sums over the block's
// rows m of x_m (y_m - 2 b_m); chi_data = 0.5 (sum + d^T S^-1 d)
template <typename TX>
struct EpiQuad {
  static constexpr bool kReduce = true;
  const TX* x;       // chain rows [N][ldx]
  int64_t ldx;
  const double* hb;  // G^T S^-1 d
  double* partial;   // [blocks][ldp]
  int64_t ldp;
  int M, N;
  __device__ double term(int64_t m, int64_t n, double v) const {
    if (m >= M || n >= N) return 0.0;
    return (double)x[n * ldx + m] * (v - 2.0 * hb[m]);
  }
  __device__ void store_partial(int tm, int64_t n, double s) const {
    if (n < N) partial[(int64_t)tm * ldp + n] = s;
  }
};

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
// FP64 tiles: 128 x 128, 4 waves of 64 x 64.
constexpr int F64_BM = 128, F64_BN = 128, F64_WM = 64, F64_WN = 64;
// BF16 tiles: 128 x 128 with a 3-stage LDS ring (144 KB); config 0 = 4 waves
// of 64 x 64, config 1 = 8 waves of 64 x 32 (two waves per SIMD).
constexpr int BF_BM = 128, BF_BN = 128, BF_STAGES = 3;

template <typename TA, typename TB, class Epi>
static void launch_f64(const Chains& c, const TA* A, int64_t lda, const TB* Bt, int64_t ldb, int M,
                       int N, int K, const Epi& epi) {
  const Problem& p = *c.p;
  const hipStream_t s = p.stream;
  constexpr int NT = (F64_BM / F64_WM) * (F64_BN / F64_WN) * 64;
  const bool ws = p.split_tail && c.split_ws64;
  SplitPlan pl = plan_split(F64_BM, F64_BN, M, N, K, p.n_cu, ws, 16);
  if (pl.split() && (size_t)pl.SPL * N * pl.TR > c.split_ws64_elems) pl = plan_split(F64_BM, F64_BN, M, N, K, 0, false);
  hipLaunchKernelGGL((k_gemm_f64<F64_BM, F64_BN, F64_WM, F64_WN, TA, TB, Epi>), dim3(pl.grid), dim3(NT), 0, s, A,
                     lda, Bt, ldb, M, N, K, pl.F, pl.SPL, c.split_ws64, pl.TR, epi);
  HMC_HIP(hipGetLastError());
  if (pl.split()) {
    if constexpr (Epi::kReduce) {
      hipLaunchKernelGGL((k_split_reduce_f64<Epi>), dim3((N + 3) / 4, (unsigned)(pl.TR / 128)), dim3(256), 0, s,
                         c.split_ws64, pl.SPL, pl.TR, N, pl.mt0, epi);
    } else {
      const int64_t work = (int64_t)N * (pl.TR / 4);
      const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
      hipLaunchKernelGGL((k_split_epilogue_f64<Epi>), dim3(blocks), dim3(256), 0, s, c.split_ws64, pl.SPL, pl.TR, N,
                         pl.mt0, epi);
    }
    HMC_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// Premultiplied endpoint over the upper block triangle of the symmetric H, in
// 128-row blocks a (T of them):
//   x^T H x - 2 b^T x = sum_a [ x_a^T H_aa x_a - 2 b_a^T x_a + 2 sum_{c>a} x_a^T H_ac x_c ],
// half the FP64 MFMA work of the full product.  Row blocks a and T-1-a are
// paired (T + 1 k-blocks of work per pair) and each pair's concatenated
// k-block list is cut into two halves, so the 2 ceil(T/2) items per chain
// tile carry equal work.  An item's rows are reduced against x right after
// each k-range (weight 1 on the diagonal block, which also takes the b term,
// 2 beyond it), and item i writes partial[i][chain]; the sum over items is
// the quadratic form.  The k loop is the one of k_gemm_f64.
// ---------------------------------------------------------------------------
template <typename TB>
__global__ __launch_bounds__((F64_BM / F64_WM) * (F64_BN / F64_WN) * 64) void k_quad_sym(
    const double* __restrict__ H, int64_t ldh, const TB* __restrict__ X, int64_t ldx,
    const double* __restrict__ hb, int d, int N, double* __restrict__ partial, int64_t ldp) {
  constexpr int BM = F64_BM, BN = F64_BN, WM = F64_WM, WN = F64_WN, BK = 16, LDS = BK + 2;
  constexpr int NT = (BM / WM) * (BN / WN) * 64;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int WPR = BN / WN;
  constexpr int KB = BM / BK;   // k steps per 128-wide k block
  constexpr int A_CH = BM * 2 / NT, B_CH = BN * 2 / NT;
  __shared__ double As[BM * LDS];
  __shared__ double Bs[BN * LDS];
  __shared__ double red[BM / WM][4][BN];

  const int T = (d + BM - 1) / BM;
  const int KT = (int)(ldh / BK);   // k steps of a row of H (ldh is a multiple of 16)
  const int pr = (int)blockIdx.x >> 1, half = (int)blockIdx.x & 1;
  const int64_t n_blk = (int64_t)blockIdx.y * BN;
  const int a1 = pr, a2 = T - 1 - pr;   // a1 <= a2; equal for the middle block of an odd T
  const int n1 = T - a1, n2 = (a2 != a1) ? T - a2 : 0;
  const int U = n1 + n2, U0 = (U + 1) / 2;
  const int lo = half ? U0 : 0, hi = half ? U : U0;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w / WPR, wn = w % WPR;
  const int q = lane >> 4, r = lane & 15;
  double ra[A_CH][8], rb[B_CH][8];
  doublex4 acc[TM][TN];
  double qs[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) qs[j] = 0.0;

  // acc = H[row block][k blocks kb0..kb1) . X[chain tile][same k]^T
  auto gemm = [&](int row, int kb0, int kb1) {
    const double* A = H + (int64_t)row * BM * ldh;
    const int kt0 = kb0 * KB, kt1 = min(kb1 * KB, KT);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
    auto gload = [&](int kt) {
#pragma unroll
      for (int i = 0; i < A_CH; ++i) {
        const int c = tid + i * NT, rr = c >> 1, col = (c & 1) * 8;
        ld8(A + (int64_t)rr * ldh + (int64_t)kt * BK + col, ra[i]);
      }
#pragma unroll
      for (int i = 0; i < B_CH; ++i) {
        const int c = tid + i * NT, rr = c >> 1, col = (c & 1) * 8;
        ld8(X + (n_blk + rr) * ldx + (int64_t)kt * BK + col, rb[i]);
      }
    };
    gload(kt0);
    for (int kt = kt0; kt < kt1; ++kt) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < A_CH; ++i) {
        const int c = tid + i * NT, rr = c >> 1, col = (c & 1) * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) As[rr * LDS + col + e] = ra[i][e];
      }
#pragma unroll
      for (int i = 0; i < B_CH; ++i) {
        const int c = tid + i * NT, rr = c >> 1, col = (c & 1) * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[rr * LDS + col + e] = rb[i][e];
      }
      __syncthreads();
      if (kt + 1 < kt1) gload(kt + 1);
      double af[TM][4], bfr[TN][4];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const double* src = &As[(wm * WM + i * 16 + r) * LDS + 4 * q];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) af[i][s4] = src[s4];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const double* src = &Bs[(wn * WN + j * 16 + r) * LDS + 4 * q];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) bfr[j][s4] = src[s4];
      }
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][s4], bfr[j][s4], acc[i][j], 0, 0, 0);
    }
  };
  // qs[j] += sum over the lane's rows m of x[n][m] (wt acc - 2 b_m [diagonal block])
  auto flush = [&](int row, double wt, bool diag) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int64_t m0 = (int64_t)row * BM + wm * WM + i * 16 + q;
        const int64_t n = n_blk + wn * WN + j * 16 + r;
        if (n >= N) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t m = m0 + 4 * v;
          if (m >= d) continue;
          const double xv = (double)X[n * ldx + m];
          qs[j] += xv * (wt * acc[i][j][v] - (diag ? 2.0 * hb[m] : 0.0));
        }
      }
  };
  for (int sg = 0; sg < 2; ++sg) {
    const int row = sg ? a2 : a1, base = sg ? n1 : 0, cnt = sg ? n2 : n1;
    const int s0 = max(lo, base) - base, s1 = min(hi, base + cnt) - base;
    if (s0 >= s1) continue;
    int kb0 = row + s0;
    const int kb1 = row + s1;
    if (kb0 == row) {
      gemm(row, kb0, kb0 + 1);
      flush(row, 1.0, true);
      ++kb0;
    }
    if (kb0 < kb1) {
      gemm(row, kb0, kb1);
      flush(row, 2.0, false);
    }
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) red[wm][q][wn * WN + j * 16 + r] = qs[j];
  __syncthreads();
  for (int col = tid; col < BN; col += NT) {
    double sum = 0.0;
#pragma unroll
    for (int a = 0; a < BM / WM; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) sum += red[a][b][col];
    if (n_blk + col < N) partial[(int64_t)blockIdx.x * ldp + n_blk + col] = sum;
  }
}

static size_t split_elems_f64(int M, int N, int K, int n_cu) {
  const SplitPlan pl = plan_split(F64_BM, F64_BN, M, N, K, n_cu, true, 16);
  return pl.split() ? (size_t)pl.SPL * N * pl.TR : 0;
}

// Workspace for the split tail: float32 partial planes, or null (no split);
// min_piece / max_spl as in plan_split.
struct SplitWs {
  float* ptr = nullptr;
  size_t elems = 0;
  int n_cu = 0;
  int min_piece = 8, max_spl = 8;
  // per-kernel event timing (when tc is set and its timing is on): the GEMM
  // kernel into kind, the split-tail epilogue into tail_kind
  Chains* tc = nullptr;
  int kind = -1, tail_kind = -1;
  // tile counters of the two-k-range fix-up ([0] error flag, [1 + tile]), or
  // null: the split epilogue kernel sums the partials
  int* fix = nullptr;
};

template <int BM, int BN, int WM, int WN, int NSPLIT, int S, class Epi>
static void launch_bf16_cfg(hipStream_t s, const uint16_t* A, int64_t lda, const uint16_t* B0,
                            const uint16_t* B1, int64_t ldb, int M, int N, int K, int order, const SplitWs& ws,
                            const Epi& epi) {
  constexpr int NT = (BM / WM) * (BN / WN) * 64;
  SplitPlan pl = plan_split(BM, BN, M, N, K, ws.n_cu, ws.ptr != nullptr, 64, ws.min_piece, ws.max_spl);
  if (pl.split() && (size_t)pl.SPL * N * pl.TR > ws.elems) pl = plan_split(BM, BN, M, N, K, 0, false);
  const int gn = (N + BN - 1) / BN, T = ((M + BM - 1) / BM) * gn;
  int* fix = (pl.split() && pl.SPL == 2 && ws.fix && T - pl.F <= kFixTiles) ? ws.fix : nullptr;
  const bool timed = ws.tc != nullptr && ws.kind >= 0;
  hipEvent_t ev = nullptr;
  if (timed) time_begin(*ws.tc, ws.kind, ev);
  hipLaunchKernelGGL((k_gemm_bf16<BM, BN, WM, WN, NSPLIT, S, Epi>), dim3(pl.grid), dim3(NT), 0, s, A, lda, B0, B1,
                     ldb, M, N, K, order, pl.F, pl.SPL, ws.ptr, pl.TR, fix, epi);
  HMC_HIP(hipGetLastError());
  if (timed) time_end(*ws.tc, ws.kind, ev);
  if (pl.split() && !fix) {
    const int64_t work = (int64_t)N * (pl.TR / 4);
    const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
    if (timed) time_begin(*ws.tc, ws.tail_kind, ev);
    hipLaunchKernelGGL((k_split_epilogue<Epi>), dim3(blocks), dim3(256), 0, s, ws.ptr, pl.SPL, pl.TR, N, pl.mt0,
                       epi);
    HMC_HIP(hipGetLastError());
    if (timed) time_end(*ws.tc, ws.tail_kind, ev);
  }
}

template <int BM, int BN>
static size_t split_elems(int M, int N, int K, int n_cu, int min_piece = 8, int max_spl = 8) {
  const SplitPlan pl = plan_split(BM, BN, M, N, K, n_cu, true, 64, min_piece, max_spl);
  return pl.split() ? (size_t)pl.SPL * N * pl.TR : 0;
}

// Tile configurations: 0 = 128x128 block, 4 waves of 64x64;
// 1 = 128x128, 8 waves of 64x32 (two per SIMD);
// 2 = 256x128, 4 waves of 128x64, 2 stages;
// 3 = 256x128, 8 waves of 64x64, 3 stages (one-plane operands; two-plane
//     operands use the config 1 tile): a third less G and chain traffic from
//     L2 per MFMA than 128x128 tiles.
// 4 = 256x128, 8 waves of 128x32, 3 stages (one-plane operands; two-plane
//     operands use the config 1 tile): each wave runs 4 MFMAs per k-step on
//     four A fragments and one B fragment; at the premultiplied stage shape
//     (M = K = 8192, N = 512) its main loop takes 0.87x the time of config 3.
// One-plane operands need 32 KB per stage of a 128x128 tile: a 4-stage ring
// keeps three k-tiles in flight.
template <int NSPLIT, class Epi>
static void launch_bf16_split(int cfg, hipStream_t st, const uint16_t* A, int64_t lda, const uint16_t* B0,
                              const uint16_t* B1, int64_t ldb, int M, int N, int K, int order, const SplitWs& ws,
                              const Epi& epi) {
  constexpr int S = NSPLIT == 1 ? 4 : BF_STAGES;
  if (cfg == 0)
    launch_bf16_cfg<BF_BM, BF_BN, 64, 64, NSPLIT, S>(st, A, lda, B0, B1, ldb, M, N, K, order, ws, epi);
  else if (cfg == 2)
    launch_bf16_cfg<256, 128, 128, 64, NSPLIT, 2>(st, A, lda, B0, B1, ldb, M, N, K, order, ws, epi);
  else if (cfg == 3 && NSPLIT == 1)
    launch_bf16_cfg<256, 128, 64, 64, 1, 3>(st, A, lda, B0, B1, ldb, M, N, K, order, ws, epi);
  else if (cfg == 4 && NSPLIT == 1)
    launch_bf16_cfg<256, 128, 128, 32, 1, 3>(st, A, lda, B0, B1, ldb, M, N, K, order, ws, epi);
  else
    launch_bf16_cfg<BF_BM, BF_BN, 64, 32, NSPLIT, S>(st, A, lda, B0, B1, ldb, M, N, K, order, ws, epi);
}

// partial-plane elements a launch of (cfg, planes) may need
static size_t split_elems_cfg(int cfg, int nsplit, int M, int N, int K, int n_cu) {
  if (cfg == 2) return split_elems<256, 128>(M, N, K, n_cu);
  if ((cfg == 3 || cfg == 4) && nsplit == 1) return split_elems<256, 128>(M, N, K, n_cu);
  return split_elems<BF_BM, BF_BN>(M, N, K, n_cu);
}

// kind / tail_kind >= 0: time the GEMM kernel and its split-tail epilogue in
// those timing groups of c; fixup: tiles cut into two k-ranges sum their
// partials inside the GEMM kernel (problem option split_fixup)
template <class Epi>
static void launch_bf16(Chains& c, int cfg, int split, const uint16_t* A, int64_t lda, const uint16_t* B0,
                        const uint16_t* B1, int64_t ldb, int M, int N, int K, const Epi& epi, int kind = -1,
                        int tail_kind = -1, bool fixup = false) {
  const Problem& p = *c.p;
  SplitWs ws;
  if (p.split_tail && c.split_ws) ws = SplitWs{c.split_ws, c.split_ws_elems, p.n_cu};
  if (fixup && p.split_fixup && c.fix && ws.ptr) {
    ws.fix = c.fix;
    c.fix_used = true;
  }
  if (kind >= 0) {
    ws.tc = &c;
    ws.kind = kind;
    ws.tail_kind = tail_kind;
  }
  if (split == 2) launch_bf16_split<2>(cfg, p.stream, A, lda, B0, B1, ldb, M, N, K, p.tile_order, ws, epi);
  else launch_bf16_split<1>(cfg, p.stream, A, lda, B0, B1, ldb, M, N, K, p.tile_order, ws, epi);
}

size_t gemm_split_workspace_f64(const Problem& p, int64_t B) {
  if (p.op != 1 || !p.split_tail) return 0;
  const int n = (int)p.n, d = (int)p.d;
  size_t e = split_elems_f64(n, (int)B, (int)p.ldg, p.n_cu);        // endpoint, FP64 residual
  e = std::max(e, split_elems_f64(d, (int)B, (int)round_up(d, 64), p.n_cu));   // premultiplied endpoint
  if (p.precision == HMC_F64) e = std::max(e, split_elems_f64(d, (int)B, (int)p.ldgt, p.n_cu));
  return e;
}

size_t gemm_split_workspace(const Problem& p, int64_t B) {
  if (p.precision != HMC_BF16 || p.op != 1 || !p.split_tail) return 0;
  const int n = (int)p.n, d = (int)p.d;
  size_t e = 0;
  for (int planes = 1; planes <= 2; ++planes) {
    e = std::max(e, split_elems_cfg(p.gemm_cfg, planes, n, (int)B, (int)p.ldg, p.n_cu));
    e = std::max(e, split_elems_cfg(p.gemm_cfg_grad, planes, d, (int)B, (int)p.ldgt, p.n_cu));
    e = std::max(e, split_elems_cfg(p.gemm_cfg_grad, planes, d, (int)B, (int)p.ldg, p.n_cu));   // H x
  }
  return e;
}

int64_t gemm_row_pad() { return 256; }   // largest M-tile of the GEMM configs
int64_t gemm_k_pad() { return 64; }      // k-tile multiple of every GEMM config
int64_t gemm_chain_pad() { return 128; } // N-tile of every GEMM config

static void build_premult(Chains& c);

// bf16 plane of the FP64 H (same [d_pad][ldh] layout)
__global__ void k_split_planes(const double* __restrict__ H, int64_t count, uint16_t* __restrict__ hh) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
    hh[k] = f32_to_bf16_bits((float)H[k]);
}

// Premultiplied trajectories (LinearMatrix.py:184-207 with premultiplication:
// gradient = H x - b): the gradient of a stage is one GEMM with K = d instead of
// the residual and gradient GEMMs with K = d and K = n.
bool traj_premult(const Problem& p) {
  return p.premult_traj != 0 && endpoint_premult(p) && (p.precision == HMC_BF16 || p.precision == HMC_F64);
}

bool gemm_anchored(const Problem& p) {
  return p.op == 1 && p.precision == HMC_BF16 && p.resolved_engine() == HMC_ENGINE_GEMM &&
         (traj_premult(p) || p.resid_update);
}

// xa[j] = mean over chains of X[c][j] (fixed order, FP64 sums)
template <typename T>
__global__ void k_chain_mean(const T* __restrict__ X, int64_t ldx, int64_t d, int64_t B, float* __restrict__ xa) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= d) return;
  double s = 0.0;
  for (int64_t c = 0; c < B; ++c) s += (double)X[c * ldx + j];
  xa[j] = (float)(s / (double)B);
}

// out[m] = (float)(sum_k A[m][k] x[k] - sub[m]) for m < M: one wavefront per
// row, lanes over k, fixed butterfly (FP64)
template <typename TA>
__global__ __launch_bounds__(256) void k_anchor_matvec(const TA* __restrict__ A, int64_t lda, int64_t M, int64_t K,
                                                       const float* __restrict__ x, const double* __restrict__ sub,
                                                       float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M) return;
  double s = 0.0;
  for (int64_t k = lane; k < K; k += 64) {
    double a;
    if constexpr (sizeof(TA) == 2) a = (double)bf16_bits_to_f32((uint16_t)A[m * lda + k]);
    else a = (double)A[m * lda + k];
    s += a * (double)x[k];
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[m] = (float)(s - sub[m]);
}

void gemm_set_anchor(Chains& c, const double* x_ref) {
  Problem& p = *c.p;
  if (!gemm_anchored(p)) return;
  HMC_REQUIRE(c.xa && c.yref, HMC_ERR_STATE, "anchor buffers missing");
  const int64_t d = p.d;
  if (x_ref) {
    std::vector<float> v((size_t)c.ldx, 0.0f);
    for (int64_t j = 0; j < d; ++j) {
      HMC_REQUIRE(std::isfinite(x_ref[j]), HMC_ERR_ARG, "anchor must be finite");
      v[j] = (float)x_ref[j];
    }
    HMC_HIP(hipMemcpyAsync(c.xa, v.data(), (size_t)c.ldx * 4, hipMemcpyHostToDevice, p.stream));
    HMC_HIP(hipStreamSynchronize(p.stream));
  } else {
    hipLaunchKernelGGL(k_chain_mean<float>, dim3((unsigned)((d + 255) / 256)), dim3(256), 0, p.stream,
                       (const float*)c.x0, c.ldx, d, c.B, c.xa);
    HMC_HIP(hipGetLastError());
  }
  if (traj_premult(p)) {
    if (!p.H) build_premult(c);
    hipLaunchKernelGGL(k_anchor_matvec<double>, dim3((unsigned)((d + 3) / 4)), dim3(256), 0, p.stream,
                       (const double*)p.H, p.ldh, d, d, c.xa, p.hb, c.yref);
  } else {
    // the trajectory's G (bf16 values) at the anchor: r_ref = G x_ref - d
    hipLaunchKernelGGL(k_anchor_matvec<uint16_t>, dim3((unsigned)((p.n + 3) / 4)), dim3(256), 0, p.stream,
                       (const uint16_t*)p.G, p.ldg, p.n, d, c.xa, p.dobs, c.yref);
  }
  HMC_HIP(hipGetLastError());
  c.anchor_ready = true;
}

static void ensure_h_plane(Problem& p) {
  if (p.Hh) return;
  const int64_t cnt = p.d_pad * p.ldh;
  p.Hh = (uint16_t*)dev_alloc(p.owned, (size_t)cnt * 2);
  hipLaunchKernelGGL(k_split_planes, dim3((unsigned)std::min<int64_t>((cnt + 255) / 256, 8192)), dim3(256), 0,
                     p.stream, p.H, cnt, p.Hh);
  HMC_HIP(hipGetLastError());
}

static void premult_kick_drift(Chains& c, double kick, double drift, double drift2) {
  Problem& p = *c.p;
  const int d = (int)p.d, B = (int)c.B;
  const DevModel md = p.model();
  if (!p.H) build_premult(c);
  if (p.precision == HMC_BF16) ensure_h_plane(p);
  if (p.precision == HMC_BF16) {
    HMC_REQUIRE(c.anchor_ready, HMC_ERR_STATE, "anchored trajectory without an anchor");
    // g = (H x_ref - b) + H_hi bf16(x - x_ref): one plane of the offset from
    // the anchor, written by the previous stage (or the momentum refresh).
    // The GEMM kernel is timed in group 4, its split-tail epilogue in group 7.
    const int cur = c.dx_cur;
    EpiLeap<float> e{(float*)c.xt, (float*)c.pt, c.dxp[1 - cur], nullptr, c.ldx, c.eps, c.beta, kick, drift,
                     drift2, md, d, B, c.xa, c.yref, nullptr};
    launch_bf16(c, p.gemm_cfg_grad, 1, p.Hh, p.ldh, c.dxp[cur], nullptr, c.ldx, d, B, (int)p.ldh, e, 4, 7, true);
    c.dx_cur = 1 - cur;
  } else {
    // FP64: g = H x - b from a copy of the positions (the epilogue updates xt)
    hipEvent_t ev;
    time_begin(c, 4, ev);
    HMC_HIP(hipMemcpyAsync(c.xs, c.xt, (size_t)c.B_pad * c.ldx * 8, hipMemcpyDeviceToDevice, p.stream));
    EpiLeap<double> e{(double*)c.xt, (double*)c.pt, nullptr, nullptr, c.ldx, c.eps, c.beta, kick, drift, drift2,
                      md, d, B, nullptr, nullptr, p.hb};
    launch_f64(c, (const double*)p.H, p.ldh, (const double*)c.xs, c.ldx, d, B, (int)p.ldh, e);
    time_end(c, 4, ev);
  }
}

// Trajectory: residual GEMM then gradient GEMM with kick/drift epilogue.
void gemm_kick_drift(Chains& c, double kick, double drift, double drift2, bool first) {
  Problem& p = *c.p;
  const int n = (int)p.n, d = (int)p.d, B = (int)c.B;
  const DevModel md = p.model();
  (void)first;
  hipEvent_t ev1, ev2;
  if (traj_premult(p)) {
    premult_kick_drift(c, kick, drift, drift2);
    return;
  }
  if (p.precision == HMC_BF16) {
    const bool anch = gemm_anchored(p);
    time_begin(c, 0, ev1);
    if (anch) {
      HMC_REQUIRE(c.anchor_ready, HMC_ERR_STATE, "anchored trajectory without an anchor");
      // r = (G x_ref - d) + G_hi bf16(x - x_ref) with one plane of the offset
      EpiResidAnchor e1{c.yref, c.rh, p.bf16_split_grad == 2 ? c.rl : nullptr, c.ldr, p.var, n, B};
      launch_bf16(c, p.gemm_cfg, 1, (const uint16_t*)p.G, p.ldg, c.xh, nullptr, c.ldx, n, B, (int)p.ldg, e1);
    } else {
      EpiResid<float> e1{nullptr, c.rh, c.rl, c.ldr, p.dobs, p.var, n, B};
      launch_bf16(c, p.gemm_cfg, p.bf16_split, (const uint16_t*)p.G, p.ldg, c.xh, c.xl, c.ldx, n, B, (int)p.ldg,
                  e1);
    }
    time_end(c, 0, ev1);
    // anchored: the plane carries x - x_ref for the next residual GEMM (no lo plane)
    EpiLeap<float> e2{(float*)c.xt, (float*)c.pt, c.xh, anch ? nullptr : c.xl, c.ldx, c.eps, c.beta, kick, drift,
                      drift2, md, d, B, anch ? c.xa : nullptr, nullptr, nullptr};
    time_begin(c, 3, ev2);
    launch_bf16(c, p.gemm_cfg_grad, p.bf16_split_grad, (const uint16_t*)p.Gt, p.ldgt, c.rh, c.rl, c.ldr, d, B, (int)p.ldgt, e2);
    time_end(c, 3, ev2);
  } else if (p.precision == HMC_F64) {
    EpiResid<double> e1{(double*)c.r, nullptr, nullptr, c.ldr, p.dobs, p.var, n, B};
    time_begin(c, 0, ev1);
    launch_f64(c, (const double*)p.G, p.ldg, (const double*)c.xt, c.ldx, n, B, (int)p.ldg, e1);
    time_end(c, 0, ev1);
    EpiLeap<double> e2{(double*)c.xt, (double*)c.pt, nullptr, nullptr, c.ldx, c.eps, c.beta, kick, drift, drift2,
                       md, d, B};
    time_begin(c, 3, ev2);
    launch_f64(c, (const double*)p.Gt, p.ldgt, (const double*)c.r, c.ldr, d, B, (int)p.ldgt, e2);
    time_end(c, 3, ev2);
  } else {
    throw Error{HMC_ERR_UNSUPPORTED, "GEMM engine supports float64 and bfloat16 trajectories"};
  }
}

// ---------------------------------------------------------------------------
// Premultiplied endpoint (LinearMatrix.py:164-191, the reference's algebra
// when n > d): chi_data = 0.5 (x^T H x - 2 b^T x + d^T S^-1 d) with
// H = G^T S^-1 G and b = G^T S^-1 d built once in FP64 from the target G.
// The endpoint GEMM then has K = d instead of n.
// ---------------------------------------------------------------------------
bool endpoint_premult(const Problem& p) {
  if (p.op != 1 || p.n <= 0 || p.resolved_engine() != HMC_ENGINE_GEMM) return false;
  if (p.premult >= 0) return p.premult > 0;
  // automatic: the reference's rule n > d, while H (d^2 doubles) stays modest
  const double h_bytes = 8.0 * (double)round_up(p.d, 256) * (double)round_up(p.d, 64);
  return p.n > p.d && h_bytes <= 16.0 * 1024 * 1024 * 1024;
}

int64_t endpoint_blocks(const Chains& c) {
  const Problem& p = *c.p;
  if (!endpoint_premult(p)) return (p.n + F64_BM - 1) / F64_BM;
  const int64_t T = (p.d + F64_BM - 1) / F64_BM;
  return p.premult_sym ? 2 * ((T + 1) / 2) : T;   // k_quad_sym items, or row blocks of H
}

double endpoint_const(const Problem& p) { return p.misfit_const + (endpoint_premult(p) ? 0.5 * p.dtd : 0.0); }

// W[j][k] = Gt[j][k] / var[k] (k < n), 0 in the padding
template <typename TG>
__global__ void k_scale_cols(const TG* __restrict__ Gt, int64_t ldgt, int64_t rows, int64_t n,
                             const double* __restrict__ var, double* __restrict__ W) {
  for (int64_t j = blockIdx.y; j < rows; j += gridDim.y)
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ldgt; k += (int64_t)gridDim.x * blockDim.x) {
      double g;
      if constexpr (sizeof(TG) == 2) g = (double)bf16_bits_to_f32((uint16_t)Gt[j * ldgt + k]);
      else g = (double)Gt[j * ldgt + k];
      W[j * ldgt + k] = k < n ? g / var[k] : 0.0;
    }
}

// v[k] = dobs[k] / var[k] (k < n); rows 1..127 of the 128-row operand stay 0
__global__ void k_scaled_data(const double* __restrict__ dobs, const double* __restrict__ var, int64_t n,
                              double* __restrict__ v) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    v[k] = dobs[k] / var[k];
}

template <typename TG>
static void build_premult_t(Chains& c) {
  Problem& p = *c.p;
  const int d = (int)p.d;
  const TG* Gt = (const TG*)p.Gt_tgt;
  std::vector<void*> tmp;
  try {
    double* W = (double*)dev_alloc(tmp, (size_t)p.d_pad * p.ldgt * 8);
    double* v = (double*)dev_alloc(tmp, (size_t)F64_BN * p.ldgt * 8);
    const dim3 grid((unsigned)std::min<int64_t>((p.ldgt + 255) / 256, 64), (unsigned)std::min<int64_t>(p.d_pad, 65535));
    hipLaunchKernelGGL(k_scale_cols<TG>, grid, dim3(256), 0, p.stream, Gt, p.ldgt, p.d_pad, p.n, p.var, W);
    HMC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scaled_data, dim3((unsigned)std::min<int64_t>((p.n + 255) / 256, 1024)), dim3(256), 0,
                       p.stream, p.dobs, p.var, p.n, v);
    HMC_HIP(hipGetLastError());
    // H[j2][j1] = sum_k Gt[j1][k] W[j2][k]; b[j] = sum_k Gt[j][k] v[k]
    launch_f64(c, Gt, p.ldgt, (const double*)W, p.ldgt, d, d, (int)p.ldgt, EpiStoreF64{p.H, p.ldh, d, d});
    launch_f64(c, Gt, p.ldgt, (const double*)v, p.ldgt, d, 1, (int)p.ldgt, EpiStoreF64{p.hb, p.ldh, d, 1});
    HMC_HIP(hipStreamSynchronize(p.stream));
  } catch (...) {
    for (void* q : tmp) (void)hipFree(q);
    throw;
  }
  for (void* q : tmp) (void)hipFree(q);
}

void premult_build(Chains& c) {
  if (!c.p->H) build_premult(c);
}

static void build_premult(Chains& c) {
  Problem& p = *c.p;
  p.ldh = round_up(p.d, 64);
  p.H = (double*)dev_alloc(p.owned, (size_t)p.d_pad * p.ldh * 8);
  p.hb = (double*)dev_alloc(p.owned, (size_t)p.d_pad * 8);
  if (p.tgt_dtype == HMC_BF16) build_premult_t<uint16_t>(c);
  else if (p.tgt_dtype == HMC_F32) build_premult_t<float>(c);
  else build_premult_t<double>(c);
}

// Endpoint data misfit in FP64: partial sums per row block of (G x - d)^2 / var
// (residual form), or of x (H x - 2 b) (premultiplied form).
void gemm_endpoint(Chains& c, const void* X, int x_dtype) {
  Problem& p = *c.p;
  const int n = (int)p.n, B = (int)c.B;
  HMC_REQUIRE(endpoint_blocks(c) <= c.n_mblocks, HMC_ERR_STATE, "endpoint partial buffer too small");
  if (endpoint_premult(p)) {
    if (!p.H) build_premult(c);
    const int d = (int)p.d;
    hipEvent_t ev;
    time_begin(c, 1, ev);
    if (p.premult_sym) {
      constexpr int NT = (F64_BM / F64_WM) * (F64_BN / F64_WN) * 64;
      const int64_t T = (p.d + F64_BM - 1) / F64_BM;
      const dim3 grid((unsigned)(2 * ((T + 1) / 2)), (unsigned)((B + F64_BN - 1) / F64_BN));
      if (x_dtype == HMC_F32)
        hipLaunchKernelGGL(k_quad_sym<float>, grid, dim3(NT), 0, p.stream, (const double*)p.H, p.ldh,
                           (const float*)X, c.ldx, p.hb, d, B, c.partial, c.B_pad);
      else
        hipLaunchKernelGGL(k_quad_sym<double>, grid, dim3(NT), 0, p.stream, (const double*)p.H, p.ldh,
                           (const double*)X, c.ldx, p.hb, d, B, c.partial, c.B_pad);
      HMC_HIP(hipGetLastError());
    } else if (x_dtype == HMC_F32) {
      launch_f64(c, (const double*)p.H, p.ldh, (const float*)X, c.ldx, d, B, (int)p.ldh,
                 EpiQuad<float>{(const float*)X, c.ldx, p.hb, c.partial, c.B_pad, d, B});
    } else {
      launch_f64(c, (const double*)p.H, p.ldh, (const double*)X, c.ldx, d, B, (int)p.ldh,
                 EpiQuad<double>{(const double*)X, c.ldx, p.hb, c.partial, c.B_pad, d, B});
    }
    time_end(c, 1, ev);
    return;
  }
  EpiMisfit e{p.dobs, p.var, c.partial, c.B_pad, n, B};
  hipEvent_t ev;
  time_begin(c, 1, ev);
  if (p.tgt_dtype == HMC_BF16) {
    if (x_dtype == HMC_F32)
      launch_f64(c, (const uint16_t*)p.G_tgt, p.ldg, (const float*)X, c.ldx, n, B, (int)p.ldg, e);
    else
      launch_f64(c, (const uint16_t*)p.G_tgt, p.ldg, (const double*)X, c.ldx, n, B, (int)p.ldg, e);
  } else if (p.tgt_dtype == HMC_F32) {
    if (x_dtype == HMC_F32)
      launch_f64(c, (const float*)p.G_tgt, p.ldg, (const float*)X, c.ldx, n, B, (int)p.ldg, e);
    else
      launch_f64(c, (const float*)p.G_tgt, p.ldg, (const double*)X, c.ldx, n, B, (int)p.ldg, e);
  } else {
    if (x_dtype == HMC_F32)
      launch_f64(c, (const double*)p.G_tgt, p.ldg, (const float*)X, c.ldx, n, B, (int)p.ldg, e);
    else
      launch_f64(c, (const double*)p.G_tgt, p.ldg, (const double*)X, c.ldx, n, B, (int)p.ldg, e);
  }
  time_end(c, 1, ev);
}

// Gradient of the target at model rows X (state dtype) into out (f64, stride ldx).
void gemm_gradient(Chains& c, const void* X, double* out) {
  Problem& p = *c.p;
  const int n = (int)p.n, d = (int)p.d, B = (int)c.B;
  const DevModel md = p.model();
  elem_oob_flags(c, X);
  if (p.precision == HMC_BF16) {
    // hi/lo planes of X were prepared by the caller into c.xh / c.xl
    EpiResid<float> e1{nullptr, c.rh, c.rl, c.ldr, p.dobs, p.var, n, B};
    launch_bf16(c, p.gemm_cfg, p.bf16_split, (const uint16_t*)p.G, p.ldg, c.xh, c.xl, c.ldx, n, B, (int)p.ldg, e1);
    EpiGrad<float> e2{(const float*)X, out, c.ldx, c.oob, md, d, B};
    launch_bf16(c, p.gemm_cfg_grad, p.bf16_split_grad, (const uint16_t*)p.Gt, p.ldgt, c.rh, c.rl, c.ldr, d, B, (int)p.ldgt, e2);
  } else if (p.precision == HMC_F64) {
    EpiResid<double> e1{(double*)c.r, nullptr, nullptr, c.ldr, p.dobs, p.var, n, B};
    launch_f64(c, (const double*)p.G, p.ldg, (const double*)X, c.ldx, n, B, (int)p.ldg, e1);
    EpiGrad<double> e2{(const double*)X, out, c.ldx, c.oob, md, d, B};
    launch_f64(c, (const double*)p.Gt, p.ldgt, (const double*)c.r, c.ldr, d, B, (int)p.ldgt, e2);
  } else {
    throw Error{HMC_ERR_UNSUPPORTED, "GEMM engine supports float64 and bfloat16 trajectories"};
  }
}

// ---------------------------------------------------------------------------
// Learned surrogate stage (bf16 MFMA, direct-to-LDS pipeline): 128 x 128 tiles
// (config 1); the two K = Dp / K = H products with few output tiles split K
// into float32 partials when that fills more compute units.
// ---------------------------------------------------------------------------
// the surrogate's GEMMs have few output tiles (H x B): their K splits into up
// to 16 pieces of at least one k-tile
constexpr int SUR_MIN_PIECE = 1, SUR_MAX_SPL = 16;

size_t surrogate_split_workspace(int64_t D, int64_t H, int64_t B, int n_cu) {
  auto el = [&](int64_t M, int64_t K) {
    return split_elems<BF_BM, BF_BN>((int)M, (int)B, (int)K, n_cu, SUR_MIN_PIECE, SUR_MAX_SPL);
  };
  return std::max({el(H, round_up(D, 64)), el(H, H), el(D, H)});
}

void surrogate_stage_bf16(Chains& c, const SurGemm& g, double kick, double drift, double drift2) {
  Problem& p = *c.p;
  HMC_REQUIRE(c.ldx == g.Dp, HMC_ERR_STATE, "surrogate plane stride differs from the chains' rows");
  const int B = (int)c.B, H = (int)g.H, D = (int)g.D, Dp = (int)g.Dp;
  const SplitWs ws{g.ws, g.ws_elems, p.n_cu, SUR_MIN_PIECE, SUR_MAX_SPL};
  const hipStream_t st = p.stream;
  // Z1 = W1 xn + b1 -> A1 = silu(Z1) (bf16), S1 = silu'(Z1)
  launch_bf16_split<1>(1, st, g.w1h, Dp, g.xh, nullptr, Dp, H, B, Dp, 0, ws,
                       EpiSurAct{g.b1, nullptr, g.a1h, g.s1, H, 0, H, B});
  // U2 = silu'(W2 A1 + b2) w3 (bf16)
  launch_bf16_split<1>(1, st, g.w2h, H, g.a1h, nullptr, H, H, B, H, 0, ws,
                       EpiSurAct{g.b2, g.w3, g.u2h, nullptr, H, 1, H, B});
  // U1 = S1 (W2^T U2) (bf16)
  launch_bf16_split<1>(1, st, g.w2t, H, g.u2h, nullptr, H, H, B, H, 0, ws, EpiSurMul{g.s1, g.u1h, H, H, B});
  // g = gscale (W1^T U1) + prior'(x) -> kick / drift; the epilogue also writes
  // the next stage's normalised plane (x - x_mean) / x_scale (row stride ldx = Dp)
  const DevModel md = p.model();
  uint16_t* xo = const_cast<uint16_t*>(g.xh);
  if (p.state_dtype == HMC_F64) {
    EpiLeap<double> e{(double*)c.xt, (double*)c.pt, xo, nullptr, c.ldx, c.eps, c.beta, kick, drift, drift2,
                      md, D, B, g.xm, nullptr, nullptr, g.gscale, g.xsinv};
    launch_bf16_split<1>(1, st, g.w1t, H, g.u1h, nullptr, H, D, B, H, 0, ws, e);
  } else {
    EpiLeap<float> e{(float*)c.xt, (float*)c.pt, xo, nullptr, c.ldx, c.eps, c.beta, kick, drift, drift2,
                     md, D, B, g.xm, nullptr, nullptr, g.gscale, g.xsinv};
    launch_bf16_split<1>(1, st, g.w1t, H, g.u1h, nullptr, H, D, B, H, 0, ws, e);
  }
}

// plane[c][j] = bf16(X[c][j] - xa[j]) for the chains' rows
__global__ void k_anchor_plane(const float* __restrict__ X, int64_t ldx, int64_t d, const float* __restrict__ xa,
                               uint16_t* __restrict__ plane) {
  const int64_t c = blockIdx.y;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < d; j += (int64_t)gridDim.x * blockDim.x)
    plane[c * ldx + j] = f32_to_bf16_bits(X[c * ldx + j] - xa[j]);
}

// The gradient field a trajectory stage applies, at model rows X (state
// dtype, chain rows of the set): the anchored bf16 field, the premultiplied
// FP64 field, or the residual form otherwise (gemm_gradient).
void gemm_traj_gradient(Chains& c, const void* X, double* out) {
  Problem& p = *c.p;
  const int n = (int)p.n, d = (int)p.d, B = (int)c.B;
  const DevModel md = p.model();
  if (p.precision == HMC_BF16 && gemm_anchored(p)) {
    if (!c.anchor_ready) gemm_set_anchor(c, nullptr);
    elem_oob_flags(c, X);
    uint16_t* plane = traj_premult(p) ? c.dxp[0] : c.xh;
    const dim3 grid((unsigned)std::min<int64_t>((p.d + 255) / 256, 64), (unsigned)B);
    hipLaunchKernelGGL(k_anchor_plane, grid, dim3(256), 0, p.stream, (const float*)X, c.ldx, p.d, c.xa, plane);
    HMC_HIP(hipGetLastError());
    if (traj_premult(p)) {
      ensure_h_plane(p);
      EpiGrad<float> e{(const float*)X, out, c.ldx, c.oob, md, d, B, c.yref, nullptr};
      launch_bf16(c, p.gemm_cfg_grad, 1, p.Hh, p.ldh, plane, nullptr, c.ldx, d, B, (int)p.ldh, e);
    } else {
      EpiResidAnchor e1{c.yref, c.rh, p.bf16_split_grad == 2 ? c.rl : nullptr, c.ldr, p.var, n, B};
      launch_bf16(c, p.gemm_cfg, 1, (const uint16_t*)p.G, p.ldg, plane, nullptr, c.ldx, n, B, (int)p.ldg, e1);
      EpiGrad<float> e2{(const float*)X, out, c.ldx, c.oob, md, d, B};
      launch_bf16(c, p.gemm_cfg_grad, p.bf16_split_grad, (const uint16_t*)p.Gt, p.ldgt, c.rh, c.rl, c.ldr, d, B,
                  (int)p.ldgt, e2);
    }
    return;
  }
  if (p.precision == HMC_F64 && traj_premult(p)) {
    if (!p.H) build_premult(c);
    elem_oob_flags(c, X);
    EpiGrad<double> e{(const double*)X, out, c.ldx, c.oob, md, d, B, nullptr, p.hb};
    launch_f64(c, (const double*)p.H, p.ldh, (const double*)X, c.ldx, d, B, (int)p.ldh, e);
    return;
  }
  gemm_gradient(c, X, out);
}

}  // namespace hmc

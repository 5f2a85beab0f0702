// Learned surrogate potential for the trajectory: an MLP
//   f(x) = w3 . silu(W2 silu(W1 xn + b1) + b2) + b3,   xn = (x - x_mean) / x_scale
// modelling the likelihood misfit (chi_lik ~ f_mean + f_scale * f).  Its input
// gradient drives the leapfrog; the Metropolis test keeps the exact FP64
// Hamiltonian, so the chain targets the exact posterior.  FP32 master
// weights; GEMMs on bf16 MFMA (or FP32 SIMT for checks); training minimises
//   alpha (f - y)^2 + beta/D ||grad f - t||^2
// with hand-written double backpropagation and Adam (Kingma & Ba 2015).
#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "hmc_internal.hpp"

struct hmc_surrogate;

namespace hmc {

// ---------------------------------------------------------------------------
// generic GEMM: C[m][n] = alpha * sum_k A(m,k) B(n,k) + beta * C[m][n]
//   A(m,k) = TA ? A[k*lda + m] : A[m*lda + k];  B(n,k) = TB ? B[k*ldb + n] : B[n*ldb + k]
// 64x64x32 tiles, 256 threads; dimensions must be multiples of the tile (the
// surrogate pads every dimension to 64).
// ---------------------------------------------------------------------------
constexpr int SG_T = 64, SG_K = 32;

template <bool TA, bool TB, bool MFMA>
__global__ __launch_bounds__(256) void k_sgemm(const float* __restrict__ A, int64_t lda,
                                               const float* __restrict__ B, int64_t ldb,
                                               float* __restrict__ C, int64_t ldc, int K,
                                               float alpha, float beta) {
  __shared__ __attribute__((aligned(16))) float As[SG_T][SG_K + 1];
  __shared__ __attribute__((aligned(16))) float Bs[SG_T][SG_K + 1];
  const int tid = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.y * SG_T, n0 = (int64_t)blockIdx.x * SG_T;
  // staging: 64 x 32 elements per operand, 8 per thread
  auto stage = [&](const float* __restrict__ P, int64_t ld, bool trans, int64_t r0, int k0,
                   float (*S)[SG_K + 1]) {
    if (!trans) {
      const int r = tid >> 2, kk = (tid & 3) * 8;
      const float* src = P + (r0 + r) * ld + k0 + kk;
#pragma unroll
      for (int e = 0; e < 8; ++e) S[r][kk + e] = src[e];
    } else {
      const int kk = tid >> 3, r = (tid & 7) * 8;
      const float* src = P + (int64_t)(k0 + kk) * ld + r0 + r;
#pragma unroll
      for (int e = 0; e < 8; ++e) S[r + e][kk] = src[e];
    }
  };
  if constexpr (MFMA) {
    // 4 waves in 2 x 2, each one 32x32 bf16 MFMA tile
    const int lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
    const int q = lane >> 5, r = lane & 31;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k0 = 0; k0 < K; k0 += SG_K) {
      __syncthreads();
      stage(A, lda, TA, m0, k0, As);
      stage(B, ldb, TB, n0, k0, Bs);
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          a[e] = (__bf16)As[wm * 32 + r][16 * s + 8 * q + e];
          b[e] = (__bf16)Bs[wn * 32 + r][16 * s + 8 * q + e];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t m = m0 + wm * 32 + 8 * (v >> 2) + 4 * q + (v & 3);
      const int64_t n = n0 + wn * 32 + r;
      float* cp = C + m * ldc + n;
      *cp = alpha * acc[v] + (beta != 0.f ? beta * *cp : 0.f);
    }
  } else {
    // SIMT: each thread a 4 x 4 micro tile
    const int tm = (tid >> 4) * 4, tn = (tid & 15) * 4;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += SG_K) {
      __syncthreads();
      stage(A, lda, TA, m0, k0, As);
      stage(B, ldb, TB, n0, k0, Bs);
      __syncthreads();
#pragma unroll 8
      for (int k = 0; k < SG_K; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] += As[tm + i][k] * Bs[tn + j][k];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float* cp = C + (m0 + tm + i) * ldc + n0 + tn + j;
        *cp = alpha * acc[i][j] + (beta != 0.f ? beta * *cp : 0.f);
      }
  }
}

// ---------------------------------------------------------------------------
// element-wise pieces
// ---------------------------------------------------------------------------
__device__ inline float sig(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ inline float silu_f(float z) { return z * sig(z); }
__device__ inline float dsilu_f(float z) {
  const float s = sig(z);
  return s * (1.f + z * (1.f - s));
}
__device__ inline float d2silu_f(float z) {
  const float s = sig(z);
  return s * (1.f - s) * (2.f + z * (1.f - 2.f * s));
}

// Z += bias; A = silu(Z); S = silu'(Z)   over [rows][H]
__global__ void k_act(float* Z, float* Aout, float* S, const float* bias, int64_t rows, int H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * H) return;
  const float z = Z[i] + bias[i % H];
  Z[i] = z;
  Aout[i] = silu_f(z);
  S[i] = dsilu_f(z);
}

// f[b] = A2[b].w3 + b3 ; U2 = S2 * w3      (one block of 256 per row)
__global__ void k_head(const float* A2, const float* S2, const float* w3, const float* b3, float* U2,
                       float* f, int H) {
  __shared__ float red[256];
  const int64_t b = blockIdx.x;
  float acc = 0.f;
  for (int i = threadIdx.x; i < H; i += 256) {
    acc += A2[b * H + i] * w3[i];
    U2[b * H + i] = S2[b * H + i] * w3[i];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) f[b] = red[0] + b3[0];
}

__global__ void k_mul(float* out, const float* a, const float* b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i];
}

// normalise chain rows (state dtype, stride ldx) into Xn [rows][Dp] (zero padded)
template <typename T>
__global__ void k_norm(const T* X, int64_t ldx, float* Xn, int64_t Dp, int64_t D, const float* xm,
                       const float* xs, int64_t rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * Dp) return;
  const int64_t b = i / Dp, j = i % Dp;
  Xn[i] = (j < D) ? ((float)X[b * ldx + j] - xm[j]) / xs[j] : 0.f;
}

// surrogate kick + drift: g = G[b][j] * gscale_j + prior'(x); then as EpiLeap
template <typename T>
__global__ void k_sur_leap(T* x, T* p, int64_t ldx, const float* G, int64_t Dp, int64_t D,
                           const float* gscale, const double* eps, const double* beta, double kick,
                           double drift, double drift2, int64_t rows, DevModel md) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * D) return;
  const int64_t b = i / D, j = i % D;
  const double e = eps[b];
  const double ke = beta ? kick * e * beta[b] : kick * e;
  const int64_t o = b * ldx + j;
  T xv = x[o], pv = p[o];
  T g = (T)(G[b * Dp + j] * gscale[j]) + prior_grad<T>(md, j, xv);
  if (out_of_bounds(md, j, (double)xv)) g = (T)INFINITY;
  pv -= (T)ke * g;
  xv += (T)(drift * e) * kgrad<T>(md, j, pv);
  corrector<T>(md, j, xv, pv);
  if (drift2 != 0.0) {
    xv += (T)(drift2 * e) * kgrad<T>(md, j, pv);
    corrector<T>(md, j, xv, pv);
  }
  x[o] = xv;
  p[o] = pv;
}

// training-batch targets: r = f - y ; df = 2 alpha r / Bt ; E = G - T ; dG = 2 beta/D E / Bt;
// loss partial per row
__global__ void k_residuals(const float* f, const float* y, const float* G, const float* Tg,
                            float* df, float* dG, float* lrow, int64_t Dp, int64_t D, float alpha,
                            float beta, int Bt) {
  __shared__ float red[256];
  const int64_t b = blockIdx.x;
  float acc = 0.f;
  for (int64_t j = threadIdx.x; j < Dp; j += 256) {
    const float e = (j < D) ? G[b * Dp + j] - Tg[b * Dp + j] : 0.f;
    dG[b * Dp + j] = 2.f * beta / (float)D * e / (float)Bt;
    acc += e * e;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float r = f[b] - y[b];
    df[b] = 2.f * alpha * r / (float)Bt;
    lrow[b] = alpha * r * r + beta / (float)D * red[0];
  }
}

// dS = dU * V ; dV = dU * S   (in place: dU -> dV, out dS)
__global__ void k_split(const float* dU, const float* V, const float* S, float* dS, float* dV, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dS[i] = dU[i] * V[i];
  dV[i] = dU[i] * S[i];
}

// dZ2 = dU2*w3 * silu''(Z2) + df (x) w3 * S2 ; also dw3 partial pieces handled by k_colsum
__global__ void k_dz2(const float* dU2, const float* w3, const float* Z2, const float* S2,
                      const float* df, float* dZ2, int64_t rows, int H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * H) return;
  const int64_t b = i / H;
  const int h = (int)(i % H);
  dZ2[i] = dU2[i] * w3[h] * d2silu_f(Z2[i]) + df[b] * w3[h] * S2[i];
}

// dZ1 = dS1 * silu''(Z1) + dA1 * S1
__global__ void k_dz1(const float* dS1, const float* Z1, const float* dA1, const float* S1, float* dZ1,
                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dZ1[i] = dS1[i] * d2silu_f(Z1[i]) + dA1[i] * S1[i];
}

// out[h] = sum_b X[b][h] * (Y ? Y[b][h] : 1) (+ sum_b w[b] * W[b][h])   fixed order
__global__ void k_colsum(const float* X, const float* Y, const float* wv, const float* Wm, float* out,
                         int64_t rows, int H) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= H) return;
  float acc = 0.f;
  for (int64_t b = 0; b < rows; ++b) {
    acc += X[b * H + h] * (Y ? Y[b * H + h] : 1.f);
    if (wv) acc += wv[b] * Wm[b * H + h];
  }
  out[h] = acc;
}

__global__ void k_sum(const float* v, int64_t n, float* out) {
  __shared__ float red[256];
  float acc = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += 256) acc += v[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather + normalise training rows: Xn[b] = (X[idx[b]] - xm)/xs ; Tn[b] = T[idx[b]] * xs / fs
__global__ void k_gather(const float* Xall, const float* Tall, const float* yall, const int32_t* idx,
                         float* Xn, float* Tn, float* yb, int64_t Dp, int64_t rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * Dp) return;
  const int64_t b = i / Dp, j = i % Dp;
  const int64_t src = (int64_t)idx[b] * Dp + j;
  Xn[i] = Xall[src];
  Tn[i] = Tall[src];
  if (j == 0) yb[b] = yall[idx[b]];
}

// ---------------------------------------------------------------------------
// host object
// ---------------------------------------------------------------------------
struct Surrogate {
  int64_t D = 0, Dp = 0, H = 0;
  int mode = 0;                  // 0 bf16 MFMA, 1 fp32 SIMT
  hipStream_t stream = nullptr;
  std::vector<void*> owned;
  // parameters (packed: W1 | b1 | W2 | b2 | w3 | b3) and Adam moments
  float* P = nullptr;
  float* Pm = nullptr;
  float* Pv = nullptr;
  float* dP = nullptr;
  int64_t nP = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, ow3 = 0, ob3 = 0;
  int64_t adam_t = 0;
  // normalisation
  float* xm = nullptr;
  float* xs = nullptr;
  float* gscale = nullptr;      // f_scale / x_scale
  float* xsinv = nullptr;       // 1 / x_scale
  double f_mean = 0.0, f_scale = 1.0;
  // work buffers for up to Bmax rows
  int64_t Bmax = 0;
  float *Xn = nullptr, *Z1 = nullptr, *A1 = nullptr, *S1 = nullptr, *Z2 = nullptr, *A2 = nullptr,
        *S2 = nullptr, *U2 = nullptr, *V1 = nullptr, *U1 = nullptr, *G = nullptr, *f = nullptr;
  float *Tn = nullptr, *yb = nullptr, *df = nullptr, *dG = nullptr, *dU1 = nullptr, *dS1 = nullptr,
        *dV1 = nullptr, *dU2 = nullptr, *dZ2 = nullptr, *dA1 = nullptr, *dZ1 = nullptr, *lrow = nullptr,
        *scal = nullptr;
  int32_t* idx = nullptr;
  // bf16 MFMA trajectory stage: weight planes (rebuilt when the parameters
  // change) and per-stage operands for up to Bh rows
  uint16_t *w1h = nullptr, *w1t = nullptr, *w2h = nullptr, *w2t = nullptr;
  bool planes_ok = false;
  int64_t Bh = 0;
  uint16_t *xh = nullptr, *a1h = nullptr, *u2h = nullptr, *u1h = nullptr;
  float* s1h = nullptr;
  const void* planes_for = nullptr;   // chain set whose positions xh holds
  float* ws = nullptr;
  size_t ws_elems = 0;

  float* W1() { return P + oW1; }
  float* b1() { return P + ob1; }
  float* W2() { return P + oW2; }
  float* b2() { return P + ob2; }
  float* w3() { return P + ow3; }
  float* b3() { return P + ob3; }
};

static void ensure_rows(Surrogate& s, int64_t rows) {
  rows = round_up(rows, 64);
  if (rows <= s.Bmax) return;
  auto F = [&](size_t n) { return (float*)dev_alloc(s.owned, n * sizeof(float)); };
  const int64_t RD = rows * s.Dp, RH = rows * s.H;
  s.Xn = F(RD); s.G = F(RD); s.Tn = F(RD); s.dG = F(RD);
  s.Z1 = F(RH); s.A1 = F(RH); s.S1 = F(RH); s.Z2 = F(RH); s.A2 = F(RH); s.S2 = F(RH);
  s.U2 = F(RH); s.V1 = F(RH); s.U1 = F(RH); s.dU1 = F(RH); s.dS1 = F(RH); s.dV1 = F(RH);
  s.dU2 = F(RH); s.dZ2 = F(RH); s.dA1 = F(RH); s.dZ1 = F(RH);
  s.f = F(rows); s.yb = F(rows); s.df = F(rows); s.lrow = F(rows); s.scal = F(64);
  s.idx = (int32_t*)dev_alloc(s.owned, rows * sizeof(int32_t));
  s.Bmax = rows;
}

template <bool TA, bool TB>
static void gemm(Surrogate& s, const float* A, int64_t lda, const float* B, int64_t ldb, float* C,
                 int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta) {
  HMC_REQUIRE(M % SG_T == 0 && N % SG_T == 0 && K % SG_K == 0, HMC_ERR_STATE, "surrogate GEMM padding");
  const dim3 grid((unsigned)(N / SG_T), (unsigned)(M / SG_T));
  if (s.mode == 0)
    hipLaunchKernelGGL((k_sgemm<TA, TB, true>), grid, dim3(256), 0, s.stream, A, lda, B, ldb, C, ldc,
                       (int)K, alpha, beta);
  else
    hipLaunchKernelGGL((k_sgemm<TA, TB, false>), grid, dim3(256), 0, s.stream, A, lda, B, ldb, C, ldc,
                       (int)K, alpha, beta);
  HMC_HIP(hipGetLastError());
}

static inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// dst[r][c] = bf16(src[r][c]) (trans 0) or dst[c][r] = bf16(src[r][c]) (trans 1)
__global__ void k_weight_plane(const float* __restrict__ src, int64_t lds, int64_t rows, int64_t cols,
                               uint16_t* __restrict__ dst, int64_t ldd, int trans) {
  const int64_t n = rows * cols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i % cols;
    const uint16_t v = f32_to_bf16_bits(src[r * lds + c]);
    if (trans) dst[c * ldd + r] = v;
    else dst[r * ldd + c] = v;
  }
}

// bf16 normalised rows Xh[b][j] = bf16((x - xm_j) / xs_j) for b < B, zero in the
// padding (rows up to `rows`, columns up to Dp)
// one workgroup row per chain row (blockIdx.y), 4 columns per thread
template <typename T>
__global__ void k_norm_bf16(const T* __restrict__ X, int64_t ldx, uint16_t* __restrict__ Xh, int Dp, int D,
                            const float* __restrict__ xm, const float* __restrict__ xsinv, int B) {
  const int b = blockIdx.y;
  for (int j0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x); j0 < Dp; j0 += 4 * gridDim.x * blockDim.x) {
    uint32_t h[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int j = j0 + e;
      h[e] = (b < B && j < D) ? f32_to_bf16_bits(((float)X[(int64_t)b * ldx + j] - xm[j]) * xsinv[j]) : 0u;
    }
    *reinterpret_cast<uint2*>(&Xh[(int64_t)b * Dp + j0]) = uint2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
  }
}

// (launched on `st`, the stream of the GEMMs that read the planes)
static void ensure_planes(Surrogate& s, hipStream_t st) {
  if (s.planes_ok) return;
  // parameters written on the surrogate's own stream (training) are complete
  HMC_HIP(hipStreamSynchronize(s.stream));
  const int64_t H = s.H, Dp = s.Dp, Hp = round_up(H, 256), Dpp = round_up(Dp, 256);
  if (!s.w1h) {
    s.w1h = (uint16_t*)dev_alloc(s.owned, (size_t)Hp * Dp * 2);
    s.w1t = (uint16_t*)dev_alloc(s.owned, (size_t)Dpp * H * 2);
    s.w2h = (uint16_t*)dev_alloc(s.owned, (size_t)Hp * H * 2);
    s.w2t = (uint16_t*)dev_alloc(s.owned, (size_t)Hp * H * 2);
  }
  const dim3 grid(4096);
  hipLaunchKernelGGL(k_weight_plane, grid, dim3(256), 0, st, s.W1(), Dp, H, Dp, s.w1h, Dp, 0);
  hipLaunchKernelGGL(k_weight_plane, grid, dim3(256), 0, st, s.W1(), Dp, H, Dp, s.w1t, H, 1);
  hipLaunchKernelGGL(k_weight_plane, grid, dim3(256), 0, st, s.W2(), H, H, H, s.w2h, H, 0);
  hipLaunchKernelGGL(k_weight_plane, grid, dim3(256), 0, st, s.W2(), H, H, H, s.w2t, H, 1);
  HMC_HIP(hipGetLastError());
  s.planes_ok = true;
}

static void ensure_bf16_rows(Surrogate& s, int64_t B, int n_cu) {
  const int64_t rows = round_up(B, 256);
  if (rows > s.Bh) {
    s.xh = (uint16_t*)dev_alloc(s.owned, (size_t)rows * s.Dp * 2);
    s.a1h = (uint16_t*)dev_alloc(s.owned, (size_t)rows * s.H * 2);
    s.u2h = (uint16_t*)dev_alloc(s.owned, (size_t)rows * s.H * 2);
    s.u1h = (uint16_t*)dev_alloc(s.owned, (size_t)rows * s.H * 2);
    s.s1h = (float*)dev_alloc(s.owned, (size_t)rows * s.H * 4);
    s.Bh = rows;
  }
  const size_t e = surrogate_split_workspace(s.D, s.H, B, n_cu);
  if (e > s.ws_elems) {
    s.ws = (float*)dev_alloc(s.owned, e * 4);
    s.ws_elems = e;
  }
}

// forward + input gradient for `rows` normalised rows in s.Xn (rows multiple of 64)
static void forward_grad(Surrogate& s, int64_t rows) {
  const int64_t H = s.H, Dp = s.Dp;
  gemm<false, false>(s, s.Xn, Dp, s.W1(), Dp, s.Z1, H, rows, H, Dp, 1.f, 0.f);      // Z1 = Xn W1^T
  hipLaunchKernelGGL(k_act, g1(rows * H), dim3(256), 0, s.stream, s.Z1, s.A1, s.S1, s.b1(), rows, (int)H);
  gemm<false, false>(s, s.A1, H, s.W2(), H, s.Z2, H, rows, H, H, 1.f, 0.f);         // Z2 = A1 W2^T
  hipLaunchKernelGGL(k_act, g1(rows * H), dim3(256), 0, s.stream, s.Z2, s.A2, s.S2, s.b2(), rows, (int)H);
  hipLaunchKernelGGL(k_head, dim3((unsigned)rows), dim3(256), 0, s.stream, s.A2, s.S2, s.w3(), s.b3(),
                     s.U2, s.f, (int)H);
  gemm<false, true>(s, s.U2, H, s.W2(), H, s.V1, H, rows, H, H, 1.f, 0.f);           // V1 = U2 W2
  hipLaunchKernelGGL(k_mul, g1(rows * H), dim3(256), 0, s.stream, s.U1, s.S1, s.V1, rows * H);
  gemm<false, true>(s, s.U1, H, s.W1(), Dp, s.G, Dp, rows, Dp, H, 1.f, 0.f);          // G = U1 W1
  HMC_HIP(hipGetLastError());
}

// loss and parameter gradients for one normalised batch (Xn, Tn, yb) of `rows`
static void backward(Surrogate& s, int64_t rows, float alpha, float beta) {
  const int64_t H = s.H, Dp = s.Dp;
  hipLaunchKernelGGL(k_residuals, dim3((unsigned)rows), dim3(256), 0, s.stream, s.f, s.yb, s.G, s.Tn, s.df,
                     s.dG, s.lrow, Dp, s.D, alpha, beta, (int)rows);
  float* gW1 = s.dP + s.oW1;
  float* gb1 = s.dP + s.ob1;
  float* gW2 = s.dP + s.oW2;
  float* gb2 = s.dP + s.ob2;
  float* gw3 = s.dP + s.ow3;
  float* gb3 = s.dP + s.ob3;
  gemm<true, true>(s, s.U1, H, s.dG, Dp, gW1, Dp, H, Dp, rows, 1.f, 0.f);            // dW1 = U1^T dG
  gemm<false, false>(s, s.dG, Dp, s.W1(), Dp, s.dU1, H, rows, H, Dp, 1.f, 0.f);       // dU1 = dG W1^T
  hipLaunchKernelGGL(k_split, g1(rows * H), dim3(256), 0, s.stream, s.dU1, s.V1, s.S1, s.dS1, s.dV1, rows * H);
  gemm<true, true>(s, s.U2, H, s.dV1, H, gW2, H, H, H, rows, 1.f, 0.f);              // dW2 = U2^T dV1
  gemm<false, false>(s, s.dV1, H, s.W2(), H, s.dU2, H, rows, H, H, 1.f, 0.f);         // dU2 = dV1 W2^T
  // dw3 = sum_b dU2*S2 + sum_b df*A2 ; db3 = sum df
  hipLaunchKernelGGL(k_colsum, g1(H), dim3(256), 0, s.stream, s.dU2, s.S2, s.df, s.A2, gw3, rows, (int)H);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s.stream, s.df, rows, gb3);
  hipLaunchKernelGGL(k_dz2, g1(rows * H), dim3(256), 0, s.stream, s.dU2, s.w3(), s.Z2, s.S2, s.df, s.dZ2,
                     rows, (int)H);
  gemm<true, true>(s, s.dZ2, H, s.A1, H, gW2, H, H, H, rows, 1.f, 1.f);              // dW2 += dZ2^T A1
  hipLaunchKernelGGL(k_colsum, g1(H), dim3(256), 0, s.stream, s.dZ2, nullptr, nullptr, nullptr, gb2, rows,
                     (int)H);
  gemm<false, true>(s, s.dZ2, H, s.W2(), H, s.dA1, H, rows, H, H, 1.f, 0.f);          // dA1 = dZ2 W2
  hipLaunchKernelGGL(k_dz1, g1(rows * H), dim3(256), 0, s.stream, s.dS1, s.Z1, s.dA1, s.S1, s.dZ1, rows * H);
  gemm<true, true>(s, s.dZ1, H, s.Xn, Dp, gW1, Dp, H, Dp, rows, 1.f, 1.f);           // dW1 += dZ1^T Xn
  hipLaunchKernelGGL(k_colsum, g1(H), dim3(256), 0, s.stream, s.dZ1, nullptr, nullptr, nullptr, gb1, rows,
                     (int)H);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s.stream, s.lrow, rows, s.scal);
  HMC_HIP(hipGetLastError());
}

static void adam(Surrogate& s, float lr) {
  s.adam_t += 1;
  s.planes_ok = false;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)s.adam_t), c2 = 1.f - std::pow(b2, (float)s.adam_t);
  hipLaunchKernelGGL(k_adam, g1(s.nP), dim3(256), 0, s.stream, s.P, s.dP, s.Pm, s.Pv, s.nP, lr, b1, b2, c1, c2,
                     1e-8f);
  HMC_HIP(hipGetLastError());
}

// gradient of the surrogate at the chains' trajectory positions + kick/drift
void surrogate_kick_drift(Chains& c, double kick, double drift, double drift2, bool first) {
  Surrogate& s = *reinterpret_cast<Surrogate*>(c.surrogate);
  Problem& p = *c.p;
  if (s.mode == 0) {
    // bf16 MFMA: normalised rows, then four pipelined GEMMs whose last
    // epilogue applies the kick and drift
    ensure_planes(s, p.stream);
    ensure_bf16_rows(s, c.B, p.n_cu);
    hipEvent_t ev;
    time_begin(c, 5, ev);
    // the first stage of a trajectory normalises the rows (up to the GEMM's
    // 128-row chain tile: padding rows become zeros); later stages read the
    // plane the previous stage's kick/drift epilogue wrote
    if (first || s.planes_for != &c) {
      HMC_REQUIRE(s.Dp % 4 == 0 && s.Bh < 65536, HMC_ERR_STATE, "surrogate row layout");
      const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((s.Dp / 4 + 255) / 256, 8)),
                      (unsigned)round_up(c.B, 128));
      if (p.state_dtype == HMC_F64)
        hipLaunchKernelGGL(k_norm_bf16<double>, grid, dim3(256), 0, p.stream, (const double*)c.xt, c.ldx, s.xh,
                           (int)s.Dp, (int)s.D, s.xm, s.xsinv, (int)c.B);
      else
        hipLaunchKernelGGL(k_norm_bf16<float>, grid, dim3(256), 0, p.stream, (const float*)c.xt, c.ldx, s.xh,
                           (int)s.Dp, (int)s.D, s.xm, s.xsinv, (int)c.B);
      HMC_HIP(hipGetLastError());
      s.planes_for = &c;
    }
    SurGemm g;
    g.D = s.D;
    g.Dp = s.Dp;
    g.H = s.H;
    g.xh = s.xh;
    g.w1h = s.w1h;
    g.w1t = s.w1t;
    g.w2h = s.w2h;
    g.w2t = s.w2t;
    g.b1 = s.b1();
    g.b2 = s.b2();
    g.w3 = s.w3();
    g.gscale = s.gscale;
    g.xm = s.xm;
    g.xsinv = s.xsinv;
    g.a1h = s.a1h;
    g.u2h = s.u2h;
    g.u1h = s.u1h;
    g.s1 = s.s1h;
    g.ws = s.ws;
    g.ws_elems = s.ws_elems;
    surrogate_stage_bf16(c, g, kick, drift, drift2);
    time_end(c, 5, ev);
    return;
  }
  const int64_t rows = round_up(c.B, 64);
  ensure_rows(s, rows);
  // launches go to the problem's stream for the duration of the call
  struct StreamSwap {
    Surrogate& s;
    hipStream_t saved;
    ~StreamSwap() { s.stream = saved; }
  } swap{s, s.stream};
  s.stream = p.stream;
  hipEvent_t ev;
  time_begin(c, 5, ev);
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_norm<double>, g1(rows * s.Dp), dim3(256), 0, s.stream, (const double*)c.xt, c.ldx,
                       s.Xn, s.Dp, s.D, s.xm, s.xs, rows);
  else
    hipLaunchKernelGGL(k_norm<float>, g1(rows * s.Dp), dim3(256), 0, s.stream, (const float*)c.xt, c.ldx,
                       s.Xn, s.Dp, s.D, s.xm, s.xs, rows);
  forward_grad(s, rows);
  time_end(c, 5, ev);
  hipEvent_t ev2;
  time_begin(c, 6, ev2);
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_sur_leap<double>, g1(c.B * p.d), dim3(256), 0, s.stream, (double*)c.xt, (double*)c.pt,
                       c.ldx, s.G, s.Dp, p.d, s.gscale, c.eps, (const double*)c.beta, kick, drift, drift2, c.B,
                       p.model());
  else
    hipLaunchKernelGGL(k_sur_leap<float>, g1(c.B * p.d), dim3(256), 0, s.stream, (float*)c.xt, (float*)c.pt,
                       c.ldx, s.G, s.Dp, p.d, s.gscale, c.eps, (const double*)c.beta, kick, drift, drift2, c.B,
                       p.model());
  HMC_HIP(hipGetLastError());
  time_end(c, 6, ev2);
}

}  // namespace hmc

struct hmc_surrogate {
  hmc::Surrogate s;
};

using namespace hmc;

extern "C" {

int hmc_surrogate_create(int64_t dimensions, int64_t hidden, uint64_t seed, hmc_surrogate** out) {
  return guarded("hmc_surrogate_create", [&] {
    HMC_REQUIRE(out && dimensions > 0 && hidden > 0 && hidden % 64 == 0, HMC_ERR_ARG,
                "dimensions > 0 and hidden a positive multiple of 64 are required");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      throw Error{HMC_ERR_HIP, "no HIP device is visible"};
    }
    auto* h = new hmc_surrogate();
    Surrogate& s = h->s;
    try {
      HMC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
      s.D = dimensions;
      s.Dp = round_up(dimensions, 64);
      s.H = hidden;
      s.oW1 = 0;
      s.ob1 = s.oW1 + s.H * s.Dp;
      s.oW2 = s.ob1 + s.H;
      s.ob2 = s.oW2 + s.H * s.H;
      s.ow3 = s.ob2 + s.H;
      s.ob3 = s.ow3 + s.H;
      s.nP = s.ob3 + 1;
      s.P = (float*)dev_alloc(s.owned, s.nP * 4);
      s.Pm = (float*)dev_alloc(s.owned, s.nP * 4);
      s.Pv = (float*)dev_alloc(s.owned, s.nP * 4);
      s.dP = (float*)dev_alloc(s.owned, s.nP * 4);
      // Glorot-uniform initialisation, zero biases
      std::vector<float> P0(s.nP, 0.f);
      std::mt19937_64 g(seed);
      auto fill = [&](int64_t off, int64_t rows, int64_t cols, int64_t ld, int64_t fan_in, int64_t fan_out) {
        std::uniform_real_distribution<float> U(-1.f, 1.f);
        const float a = std::sqrt(6.f / (float)(fan_in + fan_out));
        for (int64_t r = 0; r < rows; ++r)
          for (int64_t c = 0; c < cols; ++c) P0[off + r * ld + c] = a * U(g);
      };
      fill(s.oW1, s.H, s.D, s.Dp, s.D, s.H);
      fill(s.oW2, s.H, s.H, s.H, s.H, s.H);
      fill(s.ow3, 1, s.H, s.H, s.H, 1);
      h2d(s.P, P0.data(), s.nP * 4);
      std::vector<float> ones(s.Dp, 1.f), zeros(s.Dp, 0.f);
      s.xm = (float*)dev_alloc(s.owned, s.Dp * 4);
      s.xs = (float*)dev_alloc(s.owned, s.Dp * 4);
      s.gscale = (float*)dev_alloc(s.owned, s.Dp * 4);
      s.xsinv = (float*)dev_alloc(s.owned, s.Dp * 4);
      h2d(s.xs, ones.data(), s.Dp * 4);
      h2d(s.gscale, ones.data(), s.Dp * 4);
      h2d(s.xsinv, ones.data(), s.Dp * 4);
    } catch (...) {
      for (void* p : s.owned) (void)hipFree(p);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      delete h;
      throw;
    }
    *out = h;
  });
}

int hmc_surrogate_destroy(hmc_surrogate* h) {
  return guarded("hmc_surrogate_destroy", [&] {
    if (!h) return;
    (void)hipStreamSynchronize(h->s.stream);
    for (void* p : h->s.owned) (void)hipFree(p);
    (void)hipStreamDestroy(h->s.stream);
    delete h;
  });
}

int hmc_surrogate_set_mode(hmc_surrogate* h, int mode) {
  return guarded("hmc_surrogate_set_mode", [&] {
    HMC_REQUIRE(h && (mode == 0 || mode == 1), HMC_ERR_ARG, "mode must be 0 (bf16 MFMA) or 1 (fp32)");
    h->s.mode = mode;
  });
}

int hmc_surrogate_set_normalization(hmc_surrogate* h, const double* x_mean, const double* x_scale,
                                    double f_mean, double f_scale) {
  return guarded("hmc_surrogate_set_normalization", [&] {
    HMC_REQUIRE(h && x_mean && x_scale && f_scale > 0.0, HMC_ERR_ARG, "invalid normalisation");
    Surrogate& s = h->s;
    std::vector<float> m(s.Dp, 0.f), sc(s.Dp, 1.f), gs(s.Dp, 0.f), si(s.Dp, 0.f);
    for (int64_t j = 0; j < s.D; ++j) {
      HMC_REQUIRE(x_scale[j] > 0.0, HMC_ERR_ARG, "x_scale must be positive");
      m[j] = (float)x_mean[j];
      sc[j] = (float)x_scale[j];
      gs[j] = (float)(f_scale / x_scale[j]);
      si[j] = (float)(1.0 / x_scale[j]);
    }
    h2d(s.xm, m.data(), s.Dp * 4);
    h2d(s.xs, sc.data(), s.Dp * 4);
    h2d(s.gscale, gs.data(), s.Dp * 4);
    h2d(s.xsinv, si.data(), s.Dp * 4);
    s.planes_for = nullptr;
    s.f_mean = f_mean;
    s.f_scale = f_scale;
  });
}

// parameters as one packed float32 vector: W1 (H x D) | b1 | W2 (H x H) | b2 | w3 | b3
int hmc_surrogate_param_count(hmc_surrogate* h, int64_t* n) {
  return guarded("hmc_surrogate_param_count", [&] {
    HMC_REQUIRE(h && n, HMC_ERR_ARG, "NULL argument");
    const Surrogate& s = h->s;
    *n = s.H * s.D + s.H + s.H * s.H + s.H + s.H + 1;
  });
}

static void pack_unpack(Surrogate& s, float* host, bool to_device, float* dev) {
  // host layout is unpadded (W1 H x D); device rows are padded to Dp
  std::vector<float> buf(s.nP, 0.f);
  if (!to_device) d2h(buf.data(), dev, s.nP * 4);
  int64_t k = 0;
  for (int64_t r = 0; r < s.H; ++r)
    for (int64_t c = 0; c < s.D; ++c, ++k) {
      if (to_device) buf[s.oW1 + r * s.Dp + c] = host[k];
      else host[k] = buf[s.oW1 + r * s.Dp + c];
    }
  const int64_t tail = s.nP - s.ob1;
  for (int64_t i = 0; i < tail; ++i, ++k) {
    if (to_device) buf[s.ob1 + i] = host[k];
    else host[k] = buf[s.ob1 + i];
  }
  if (to_device) h2d(dev, buf.data(), s.nP * 4);
}

int hmc_surrogate_get_params(hmc_surrogate* h, float* params) {
  return guarded("hmc_surrogate_get_params", [&] {
    HMC_REQUIRE(h && params, HMC_ERR_ARG, "NULL argument");
    HMC_HIP(hipStreamSynchronize(h->s.stream));
    pack_unpack(h->s, params, false, h->s.P);
  });
}

int hmc_surrogate_set_params(hmc_surrogate* h, const float* params) {
  return guarded("hmc_surrogate_set_params", [&] {
    HMC_REQUIRE(h && params, HMC_ERR_ARG, "NULL argument");
    HMC_HIP(hipStreamSynchronize(h->s.stream));
    pack_unpack(h->s, const_cast<float*>(params), true, h->s.P);
    h->s.planes_ok = false;
    h->s.adam_t = 0;
    dzero(h->s.Pm, h->s.nP * 4);
    dzero(h->s.Pv, h->s.nP * 4);
  });
}

// upload rows X (B x D, f64) normalised into s.Xn; returns padded row count
static int64_t upload_norm(Surrogate& s, const double* X, int64_t B) {
  const int64_t rows = round_up(B, 64);
  ensure_rows(s, rows);
  std::vector<float> hx(rows * s.Dp, 0.f), xm(s.Dp), xs(s.Dp);
  d2h(xm.data(), s.xm, s.Dp * 4);
  d2h(xs.data(), s.xs, s.Dp * 4);
  for (int64_t b = 0; b < B; ++b)
    for (int64_t j = 0; j < s.D; ++j) hx[b * s.Dp + j] = ((float)X[b * s.D + j] - xm[j]) / xs[j];
  h2d(s.Xn, hx.data(), hx.size() * 4);
  return rows;
}

int hmc_surrogate_eval(hmc_surrogate* h, const double* X, int64_t B, double* f, double* g) {
  return guarded("hmc_surrogate_eval", [&] {
    HMC_REQUIRE(h && X && B > 0, HMC_ERR_ARG, "invalid argument");
    Surrogate& s = h->s;
    const int64_t rows = upload_norm(s, X, B);
    forward_grad(s, rows);
    std::vector<float> hf(rows), hg(rows * s.Dp), gs(s.Dp);
    HMC_HIP(hipMemcpyAsync(hf.data(), s.f, rows * 4, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipMemcpyAsync(hg.data(), s.G, hg.size() * 4, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipMemcpyAsync(gs.data(), s.gscale, s.Dp * 4, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
    for (int64_t b = 0; b < B; ++b) {
      if (f) f[b] = s.f_mean + s.f_scale * (double)hf[b];
      if (g)
        for (int64_t j = 0; j < s.D; ++j) g[b * s.D + j] = (double)hg[b * s.Dp + j] * (double)gs[j];
    }
  });
}

// Loss and packed parameter gradient of one batch, in normalised units:
// y (B) and t (B x D) are targets of f and of grad_xn f.
int hmc_surrogate_loss_grad(hmc_surrogate* h, const double* Xn, const double* y, const double* t, int64_t B,
                            double alpha, double beta, double* loss, float* grad) {
  return guarded("hmc_surrogate_loss_grad", [&] {
    HMC_REQUIRE(h && Xn && y && t && B > 0 && B % 64 == 0, HMC_ERR_ARG, "batch must be a positive multiple of 64");
    Surrogate& s = h->s;
    ensure_rows(s, B);
    std::vector<float> hx(B * s.Dp, 0.f), ht(B * s.Dp, 0.f), hy(B);
    for (int64_t b = 0; b < B; ++b) {
      hy[b] = (float)y[b];
      for (int64_t j = 0; j < s.D; ++j) {
        hx[b * s.Dp + j] = (float)Xn[b * s.D + j];
        ht[b * s.Dp + j] = (float)t[b * s.D + j];
      }
    }
    h2d(s.Xn, hx.data(), hx.size() * 4);
    h2d(s.Tn, ht.data(), ht.size() * 4);
    h2d(s.yb, hy.data(), B * 4);
    forward_grad(s, B);
    backward(s, B, (float)alpha, (float)beta);
    float l = 0.f;
    HMC_HIP(hipMemcpyAsync(&l, s.scal, 4, hipMemcpyDeviceToHost, s.stream));
    HMC_HIP(hipStreamSynchronize(s.stream));
    if (loss) *loss = (double)l / (double)B;
    if (grad) pack_unpack(s, grad, false, s.dP);
  });
}

// Train on N rows (X: N x D f64 raw positions, chi: N raw misfits, grad: N x D raw gradients)
// for `epochs` passes of shuffled batches (batch a multiple of 64); normalisation must be set.
int hmc_surrogate_train(hmc_surrogate* h, const double* X, const double* chi, const double* grad, int64_t N,
                        int64_t batch, int64_t epochs, double lr, double alpha, double beta, uint64_t seed,
                        double* loss_history) {
  return guarded("hmc_surrogate_train", [&] {
    HMC_REQUIRE(h && X && chi && grad && N >= batch && batch > 0 && batch % 64 == 0 && epochs > 0, HMC_ERR_ARG,
                "invalid training arguments");
    Surrogate& s = h->s;
    ensure_rows(s, batch);
    // normalised training set on the device
    std::vector<float> xm(s.Dp), xs(s.Dp);
    d2h(xm.data(), s.xm, s.Dp * 4);
    d2h(xs.data(), s.xs, s.Dp * 4);
    std::vector<void*> tmp;
    float* dX = (float*)dev_alloc(tmp, (size_t)N * s.Dp * 4);
    float* dT = (float*)dev_alloc(tmp, (size_t)N * s.Dp * 4);
    float* dy = (float*)dev_alloc(tmp, (size_t)N * 4);
    try {
      std::vector<float> hx((size_t)N * s.Dp, 0.f), ht((size_t)N * s.Dp, 0.f), hy(N);
      for (int64_t i = 0; i < N; ++i) {
        for (int64_t j = 0; j < s.D; ++j) {
          hx[i * s.Dp + j] = (float)((X[i * s.D + j] - xm[j]) / xs[j]);
          ht[i * s.Dp + j] = (float)(grad[i * s.D + j] * xs[j] / s.f_scale);   // d f_n / d x_n
        }
        hy[i] = (float)((chi[i] - s.f_mean) / s.f_scale);
      }
      h2d(dX, hx.data(), hx.size() * 4);
      h2d(dT, ht.data(), ht.size() * 4);
      h2d(dy, hy.data(), N * 4);
      std::vector<int32_t> perm(N);
      std::iota(perm.begin(), perm.end(), 0);
      std::mt19937_64 g(seed);
      const int64_t nb = N / batch;
      for (int64_t e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), g);
        double acc = 0.0;
        for (int64_t b = 0; b < nb; ++b) {
          HMC_HIP(hipMemcpyAsync(s.idx, perm.data() + b * batch, batch * 4, hipMemcpyHostToDevice, s.stream));
          hipLaunchKernelGGL(k_gather, g1(batch * s.Dp), dim3(256), 0, s.stream, dX, dT, dy, s.idx, s.Xn, s.Tn,
                             s.yb, s.Dp, batch);
          forward_grad(s, batch);
          backward(s, batch, (float)alpha, (float)beta);
          adam(s, (float)lr);
          float l = 0.f;
          HMC_HIP(hipMemcpyAsync(&l, s.scal, 4, hipMemcpyDeviceToHost, s.stream));
          HMC_HIP(hipStreamSynchronize(s.stream));
          acc += (double)l / (double)batch;
        }
        if (loss_history) loss_history[e] = acc / (double)nb;
      }
    } catch (...) {
      for (void* p : tmp) (void)hipFree(p);
      throw;
    }
    for (void* p : tmp) (void)hipFree(p);
  });
}

int hmc_chains_set_surrogate(hmc_chains* hc, hmc_surrogate* hs) {
  return guarded("hmc_chains_set_surrogate", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    if (hs) {
      HMC_REQUIRE(hs->s.D == c.p->d, HMC_ERR_ARG, "surrogate dimension does not match the problem");
      HMC_REQUIRE(c.p->resolved_engine() == HMC_ENGINE_GEMM, HMC_ERR_UNSUPPORTED,
                  "surrogate trajectories run on the GEMM engine");
    }
    c.surrogate = hs ? &hs->s : nullptr;
  });
}

}  // extern "C"

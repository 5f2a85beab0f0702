// Shared device/host definitions for the hmclab_amd HIP engine (gfx950).
//
// Element-wise target terms restate the reference's NumPy arithmetic:
//   Normal (diagonal)  hmclab/Distributions/base.py:539-574
//   Laplace            hmclab/Distributions/base.py:689-710
//   bounds/corrector   hmclab/Distributions/base.py:239-270, 361-374
//   Unit / Diagonal    hmclab/MassMatrices.py:100-142, 185-227
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/hmc_engine.h"

namespace hmc {

// ----------------------------------------------------------------------------
// Error handling: every C entry point returns an int status; the message of
// the last failure on the calling thread is kept for hmc_last_error().
// ----------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
  int code;
  std::string msg;
};

#define HMC_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw ::hmc::Error{HMC_ERR_HIP, std::string(#call) + ": " +            \
                                          hipGetErrorString(e_)};            \
  } while (0)

#define HMC_REQUIRE(cond, code, msg)                                         \
  do {                                                                       \
    if (!(cond)) throw ::hmc::Error{(code), (msg)};                          \
  } while (0)

__host__ __device__ inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// ----------------------------------------------------------------------------
// bf16 helpers (round-to-nearest-even, as __float2bfloat16)
// ----------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f32_to_bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) {           // inf / nan
    return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  }
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

__device__ __host__ inline float bf16_bits_to_f32(uint16_t b) {
  uint32_t u = ((uint32_t)b) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// ----------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC11).  Key = 64-bit seed, counter =
// (element pair, chain, proposal lo, stream<<24 | proposal hi).
// ----------------------------------------------------------------------------
struct U32x4 {
  uint32_t v[4];
};

__device__ __host__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__device__ __host__ inline U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, c.v[0], hi0, lo0);
    mulhilo32(M1, c.v[2], hi1, lo1);
    U32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words.
__device__ __host__ inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

enum : uint32_t { STREAM_MOMENTUM = 0, STREAM_SCALARS = 1, STREAM_EXCHANGE = 2 };

// Two standard normals (Box-Muller) for element pair `pair` of `chain` at
// proposal `prop`.
__device__ inline void philox_normal2(uint64_t seed, uint32_t chain, uint64_t prop, uint32_t pair,
                                      double& z0, double& z1) {
  U32x4 c = {{pair, chain, (uint32_t)prop, (STREAM_MOMENTUM << 24) | ((uint32_t)(prop >> 32) & 0xffffffu)}};
  U32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  double u1 = 1.0 - u53(r.v[0], r.v[1]);  // (0, 1]
  double u2 = u53(r.v[2], r.v[3]);
  double rad = sqrt(-2.0 * log(u1));
  double s, co;
  sincos(6.283185307179586476925286766559 * u2, &s, &co);
  z0 = rad * co;
  z1 = rad * s;
}

// Step-jitter uniform in [0.5, 1.5) and accept uniform in [0, 1).
__device__ inline void philox_scalars(uint64_t seed, uint32_t chain, uint64_t prop, double& u_eps,
                                      double& u_acc) {
  U32x4 c = {{0u, chain, (uint32_t)prop, (STREAM_SCALARS << 24) | ((uint32_t)(prop >> 32) & 0xffffffu)}};
  U32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  u_eps = 0.5 + u53(r.v[0], r.v[1]);
  u_acc = u53(r.v[2], r.v[3]);
}

// Accept uniform in [0, 1) of a replica exchange decided by `chain` after proposal `prop`.
__device__ inline double philox_exchange(uint64_t seed, uint32_t chain, uint64_t prop) {
  U32x4 c = {{0u, chain, (uint32_t)prop, (STREAM_EXCHANGE << 24) | ((uint32_t)(prop >> 32) & 0xffffffu)}};
  U32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return u53(r.v[0], r.v[1]);
}

// ----------------------------------------------------------------------------
// Element-wise part of a target: prior terms, collapsed bounds, mass diagonal.
// Passed by value to kernels.
// ----------------------------------------------------------------------------
constexpr int kMaxPriors = 4;
constexpr int kTimingKinds = 8;   // kernel timing groups of a chain set
constexpr int kFixTiles = 4096;   // tail tiles the bf16 GEMM's two-k-range fix-up counters cover
constexpr int kSellGroup = 8;     // entries per lane group of the fused engine's SELL-64 layout
constexpr int kSellMaxDim = 65472;  // rows/columns of a SELL operator (16-bit positions < 65472)
// position-ordered model arrays of the fused sparse engine: prior a and b per
// term, inverse mass, mass square root, lower and upper bound
constexpr int kPosModelArrays = 2 * kMaxPriors + 4;

struct DevPrior {
  int kind;           // HMC_PRIOR_NORMAL (a = means, b = inverse variances)
                      // HMC_PRIOR_LAPLACE (a = means, b = inverse dispersions)
  const double* a;
  const double* b;
};

struct DevModel {
  int n_priors;
  DevPrior pr[kMaxPriors];
  const double* lo;    // collapsed lower bounds or nullptr
  const double* hi;    // collapsed upper bounds or nullptr
  const double* minv;  // inverse mass diagonal or nullptr (unit mass)
  const double* msqrt; // sqrt of mass diagonal or nullptr
};

// Gradient of the prior terms at one coordinate; bounds handled by caller.
template <typename T>
__device__ inline T prior_grad(const DevModel& md, int64_t j, T x) {
  T g = T(0);
#pragma unroll
  for (int k = 0; k < kMaxPriors; ++k) {
    if (k >= md.n_priors) break;
    const DevPrior& p = md.pr[k];
    T a = (T)p.a[j], b = (T)p.b[j];
    if (p.kind == HMC_PRIOR_NORMAL) {
      g += -b * (a - x);                      // base.py:568-570
    } else {
      T s = (x > a) ? T(1) : ((x < a) ? T(-1) : T(0));
      g += s * b;                             // base.py:707-710
    }
  }
  return g;
}

// Misfit of the prior terms at one coordinate (FP64 always).
__device__ inline double prior_misfit(const DevModel& md, int64_t j, double x) {
  double m = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxPriors; ++k) {
    if (k >= md.n_priors) break;
    const DevPrior& p = md.pr[k];
    double a = p.a[j], b = p.b[j];
    if (p.kind == HMC_PRIOR_NORMAL) {
      double r = a - x;
      m += 0.5 * (r * (b * r));               // base.py:543-551
    } else {
      m += fabs(x - a) * b;                   // base.py:692-700
    }
  }
  return m;
}

__device__ inline bool out_of_bounds(const DevModel& md, int64_t j, double x) {
  return (md.lo && x < md.lo[j]) || (md.hi && x > md.hi[j]);
}

// Reflection at the collapsed bounds, lower first then upper (base.py:1129-1142).
template <typename T>
__device__ inline void corrector(const DevModel& md, int64_t j, T& x, T& p) {
  if (md.lo) {
    T lo = (T)md.lo[j];
    if (x < lo) {
      x += T(2) * (lo - x);
      p *= T(-1);
    }
  }
  if (md.hi) {
    T hi = (T)md.hi[j];
    if (x > hi) {
      x += T(2) * (hi - x);
      p *= T(-1);
    }
  }
}

template <typename T>
__device__ inline T minv_of(const DevModel& md, int64_t j) {
  return md.minv ? (T)md.minv[j] : T(1);
}

// dK/dp (MassMatrices.py:133, 218)
template <typename T>
__device__ inline T kgrad(const DevModel& md, int64_t j, T p) {
  return md.minv ? (T)md.minv[j] * p : p;
}

// One term of the kinetic energy sum (FP64): p_j^2 / M_jj.
__device__ inline double kterm(const DevModel& md, int64_t j, double p) {
  return md.minv ? p * (md.minv[j] * p) : p * p;
}

// ----------------------------------------------------------------------------
// Block-wide deterministic reduction of one double per thread.
// ----------------------------------------------------------------------------
template <int NT>
__device__ inline double block_sum(double v, double* scratch) {
  // wave-level tree (fixed order), then across waves in thread 0
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; ++i) s += scratch[i];
    scratch[NT / 64] = s;
  }
  __syncthreads();
  return scratch[NT / 64];
}

}  // namespace hmc

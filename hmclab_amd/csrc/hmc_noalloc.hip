// Batched misfit without device allocation: hmc_misfit_batch_noalloc.
//
// Callers that cannot call hipMalloc (code under graph capture, tight
// serving loops) pass device-resident models and outputs; every buffer the
// evaluation needs beyond them (the models converted to the engine's padded
// state layout and the endpoint's per-row-block partial sums) lives in one
// file-scope device array of kScratchBytes per GPU.  Its address is looked up
// once per device with hipGetSymbolAddress; the array's name in host code is
// only a stand-in and is never passed to a kernel.  Calls that overlap on one
// GPU share the array and race.
#include <algorithm>
#include <array>
#include <mutex>
#include <string>

#include "hmc_internal.hpp"

namespace hmc {

constexpr size_t kScratchBytes = (size_t)32 << 20;   // 32 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::array<void*, kMaxDevices> g_scratch{};
static std::mutex g_scratch_mu;

// device address of this GPU's scratch array (looked up on the first call on
// the device, with that device current)
static unsigned char* scratch_ptr() {
  int dev = 0;
  HMC_HIP(hipGetDevice(&dev));
  HMC_REQUIRE(dev >= 0 && dev < kMaxDevices, HMC_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    HMC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    HMC_REQUIRE(ptr != nullptr, HMC_ERR_HIP, "scratch symbol has no device address");
    g_scratch[dev] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch[dev]);
}

// models M[c][j] (float64, row stride d, c < B) -> X[c][j] in the state type,
// row stride ldx, over the whole [B_pad][ldx] block: padding rows and columns
// are written as zeros (the GEMM reads them; zero K padding keeps 0 * x out of
// the sums)
template <typename T>
__global__ void k_rows_padded(const double* __restrict__ M, int64_t d, int64_t B, T* __restrict__ X, int64_t ldx,
                              int64_t total) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = i / ldx, j = i % ldx;
    X[i] = (c < B && j < d) ? (T)M[c * d + j] : (T)0;
  }
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace hmc

using namespace hmc;

extern "C" int hmc_misfit_batch_noalloc(hmc_problem* hp, const double* M_dev, int64_t B, double* out_dev) {
  return guarded("hmc_misfit_batch_noalloc", [&] {
    HMC_REQUIRE(hp && M_dev && out_dev && B > 0, HMC_ERR_ARG, "invalid argument");
    Problem& P = hp->p;
    HMC_REQUIRE(B <= (int64_t)1 << 24, HMC_ERR_ARG, "number of models out of range");
    const bool gemm = P.resolved_engine() == HMC_ENGINE_GEMM;
    if (gemm && endpoint_premult(P))
      HMC_REQUIRE(P.H != nullptr, HMC_ERR_STATE,
                  "the premultiplied operator H is built by the first misfit evaluation; call "
                  "hmc_misfit_batch (or create chains and set their state) once before this entry point");

    // a chain set over the scratch array: state rows and endpoint partials
    hmc_chains view;
    Chains& c = view.c;
    c.p = &P;
    c.B = B;
    c.B_pad = round_up(B, 128);
    c.ldx = round_up(P.d, 64);
    c.ldr = P.n > 0 ? round_up(P.n, 64) : 64;
    const size_t es = dtype_size(P.state_dtype);
    const size_t x_bytes = align256((size_t)c.B_pad * c.ldx * es);
    size_t part_bytes = 0;
    if (P.op == 1) {
      const int64_t T = (P.d + 127) / 128;
      c.n_mblocks = std::max({P.n_pad / 128, P.d_pad / 128, 2 * ((T + 1) / 2)});
      part_bytes = align256((size_t)c.n_mblocks * c.B_pad * 8);
    }
    HMC_REQUIRE(x_bytes + part_bytes <= kScratchBytes, HMC_ERR_ARG,
                "batch needs " + std::to_string(x_bytes + part_bytes) + " bytes of scratch; at most " +
                    std::to_string(kScratchBytes) + " bytes (32 MiB) are available: use fewer models per call");
    unsigned char* base = scratch_ptr();
    void* X = base;
    if (part_bytes) c.partial = reinterpret_cast<double*>(base + x_bytes);
    // no split-tail workspace: the endpoint GEMM runs its tiles over all of K

    const int64_t total = c.B_pad * c.ldx;
    const dim3 grid((unsigned)std::min<int64_t>((total + 255) / 256, 8192));
    if (P.state_dtype == HMC_F64)
      hipLaunchKernelGGL(k_rows_padded<double>, grid, dim3(256), 0, P.stream, M_dev, P.d, B, (double*)X, c.ldx,
                         total);
    else
      hipLaunchKernelGGL(k_rows_padded<float>, grid, dim3(256), 0, P.stream, M_dev, P.d, B, (float*)X, c.ldx,
                         total);
    HMC_HIP(hipGetLastError());
    if (gemm) {
      gemm_endpoint(c, X, P.state_dtype);
      elem_finish_misfit(c, X, true, out_dev);
    } else {
      fused_misfit(c, X, out_dev);
    }
    HMC_HIP(hipStreamSynchronize(P.stream));
  });
}

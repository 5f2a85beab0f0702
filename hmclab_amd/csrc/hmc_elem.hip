// Element-wise and per-chain kernels shared by the engines:
//   momentum refresh + first drift  (MassMatrices.py:142,227; Samplers.py:1533-1545)
//   endpoint Hamiltonian, Metropolis test, accept/reject, autotune
//                                    (Samplers.py:1471-1522)
//   row conversions between host float64 rows and device state rows.
// One workgroup per chain; per-chain sums use a fixed-order tree reduction so
// results are reproducible run to run.
#include "hmc_internal.hpp"

namespace hmc {

constexpr int EW_NT = 256;

// bf16 planes for the GEMM engine: hi/lo of v, or (anchored, l null) one
// plane of v - a
template <typename T>
__device__ inline void put_planes(uint16_t* h, uint16_t* l, int64_t o, T v, float a = 0.0f) {
  if (h) {
    const float f = (float)v - a;
    const uint16_t hb = f32_to_bf16_bits(f);
    h[o] = hb;
    if (l) l[o] = f32_to_bf16_bits(f - bf16_bits_to_f32(hb));
  }
}

// Momentum draw and first drift of the trajectory.
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_refresh(T* __restrict__ xt, T* __restrict__ pt,
                                                   uint16_t* xh, uint16_t* xl,
                                                   const T* __restrict__ x0,
                                                   const double* __restrict__ p_host, int64_t ldx,
                                                   int64_t d, double* k0, double* eps, double* u,
                                                   const double* step, double drift0, uint64_t seed,
                                                   int64_t prop, int host_p, int device_scalars,
                                                   int randomize, const float* __restrict__ xa, DevModel md) {
  __shared__ double scratch[EW_NT / 64 + 1];
  const int64_t c = blockIdx.x;
  double e;
  if (device_scalars) {
    double ue, ua;
    philox_scalars(seed, (uint32_t)c, (uint64_t)prop, ue, ua);
    e = randomize ? ue * step[c] : step[c];   // rng.uniform(0.5,1.5) * stepsize
    if (threadIdx.x == 0) {
      eps[c] = e;
      u[c] = ua;
    }
  } else {
    e = eps[c];
  }
  const T dd = (T)(drift0 * e);
  double ks = 0.0;
  const int64_t npairs = (d + 1) / 2;
  for (int64_t qp = threadIdx.x; qp < npairs; qp += EW_NT) {
    double z[2] = {0.0, 0.0};
    if (!host_p) philox_normal2(seed, (uint32_t)c, (uint64_t)prop, (uint32_t)qp, z[0], z[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t j = 2 * qp + h;
      if (j >= d) break;
      const int64_t o = c * ldx + j;
      T pv = host_p ? (T)p_host[o] : (T)(md.msqrt ? md.msqrt[j] * z[h] : z[h]);
      ks += kterm(md, j, (double)pv);
      T xv = x0[o];
      xv += dd * kgrad<T>(md, j, pv);
      corrector<T>(md, j, xv, pv);
      xt[o] = xv;
      pt[o] = pv;
      put_planes<T>(xh, xl, o, xv, xa ? xa[j] : 0.0f);
    }
  }
  const double s = block_sum<EW_NT>(ks, scratch);
  if (threadIdx.x == 0) k0[c] = 0.5 * s;
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_drift(T* __restrict__ xt, T* __restrict__ pt,
                                                 uint16_t* xh, uint16_t* xl, int64_t ldx, int64_t d,
                                                 const double* eps, double drift, DevModel md) {
  const int64_t c = blockIdx.x;
  const T dd = (T)(drift * eps[c]);
  for (int64_t j = threadIdx.x; j < d; j += EW_NT) {
    const int64_t o = c * ldx + j;
    T xv = xt[o], pv = pt[o];
    xv += dd * kgrad<T>(md, j, pv);
    corrector<T>(md, j, xv, pv);
    xt[o] = xv;
    pt[o] = pv;
    put_planes<T>(xh, xl, o, xv);
  }
}

// chi of a chain row: 0.5 * sum(partials) + prior + const (+inf out of bounds)
template <typename T>
__device__ inline double chain_misfit(const T* x, int64_t d, const double* partial, int64_t nmb,
                                      int64_t ldp, int64_t c, double mconst, const DevModel& md,
                                      double* scratch, double* kin, const T* p) {
  double pm = 0.0, kk = 0.0, ob = 0.0;
  for (int64_t j = threadIdx.x; j < d; j += EW_NT) {
    const double xv = (double)x[j];
    pm += prior_misfit(md, j, xv);
    if (out_of_bounds(md, j, xv)) ob = 1.0;
    if (p) kk += kterm(md, j, (double)p[j]);
  }
  pm = block_sum<EW_NT>(pm, scratch);
  ob = block_sum<EW_NT>(ob, scratch);
  if (p) kk = block_sum<EW_NT>(kk, scratch);
  double lik = 0.0;
  for (int64_t mb = 0; mb < nmb; ++mb) lik += partial[mb * ldp + c];
  double chi = 0.5 * lik + (pm + mconst);
  if (ob > 0.0) chi += INFINITY;
  if (kin) *kin = 0.5 * kk;
  return chi;
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_metropolis(
    T* __restrict__ x0, double* chi0, const T* __restrict__ xt, const T* __restrict__ pt,
    int64_t ldx, int64_t d, const double* partial, int64_t nmb, int64_t ldp, double mconst,
    const double* k0, const double* u, double* h0o, double* h1o, int32_t* acc, int64_t* acc_count,
    double* step, int autotune, double target, double lr, int64_t prop_index, T* tr_x,
    double* tr_chi, int64_t slot, int64_t B, const double* __restrict__ beta, double* hist_rate,
    double* hist_step, int64_t hist_slot, DevModel md) {
  __shared__ double scratch[EW_NT / 64 + 1];
  __shared__ int accepted;
  __shared__ double chi_new;
  const int64_t c = blockIdx.x;
  double k1;
  const double chi1 = chain_misfit<T>(xt + c * ldx, d, partial, nmb, ldp, c, mconst, md, scratch,
                                      &k1, pt + c * ldx);
  const double bt = beta ? beta[c] : 1.0;   // tempered target exp(-beta chi)
  if (threadIdx.x == 0) {
    const double H0 = bt * chi0[c] + k0[c];
    const double H1 = bt * chi1 + k1;
    const double a = exp(H0 - H1);
    if (autotune) {
      // Samplers.py:1496-1522: record the rate and the step before the update
      if (hist_slot >= 0) {
        hist_rate[hist_slot * B + c] = a;
        hist_step[hist_slot * B + c] = step[c];
      }
      const double wgt = pow((double)(prop_index + 1), -lr);
      const double ar = isnan(a) ? 0.0 : a;
      double s = step[c] - wgt * (target - fmin(ar, 1.0));
      if (s <= 0.0) s = fmax(s, 1e-18);
      step[c] = s;
    }
    const int ok = (a > u[c]) ? 1 : 0;   // strict, Samplers.py:1486
    accepted = ok;
    chi_new = ok ? chi1 : chi0[c];
    if (h0o) h0o[c] = H0;
    if (h1o) h1o[c] = H1;
    if (acc) acc[c] = ok;
    if (acc_count) acc_count[c] += ok;
    chi0[c] = chi_new;
  }
  __syncthreads();
  if (accepted) {
    for (int64_t j = threadIdx.x; j < d; j += EW_NT) x0[c * ldx + j] = xt[c * ldx + j];
  }
  if (slot >= 0) {
    const T* src = accepted ? xt : x0;
    for (int64_t j = threadIdx.x; j < d; j += EW_NT) tr_x[(slot * B + c) * d + j] = src[c * ldx + j];
    if (threadIdx.x == 0) tr_chi[slot * B + c] = bt * chi_new;
  }
}

// Replica exchange (ParallelSampleSMP, Samplers.py:589-669): pair k = (a, b),
// b decides; the swap is accepted when
//   exp(beta_b (chi_b - chi(x_a)) + beta_a (chi_a - chi(x_b))) > u,
// i.e. exp((beta_a - beta_b)(chi_a - chi_b)) > u for one shared target; the
// positions and misfits of a and b are then exchanged.  One workgroup per pair.
template <typename T>
__global__ __launch_bounds__(EW_NT) void k_exchange(T* __restrict__ x0, double* __restrict__ chi0,
                                                   int64_t ldx, int64_t d, const int32_t* __restrict__ pairs,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ beta, int64_t* ex_count,
                                                   int32_t* swapped, uint64_t seed, int64_t prop, T* tr_x,
                                                   double* tr_chi, int64_t slot, int64_t B) {
  __shared__ int swap;
  const int64_t k = blockIdx.x;
  const int a = pairs[2 * k], b = pairs[2 * k + 1];
  if (threadIdx.x == 0) {
    const double ba = beta ? beta[a] : 1.0, bb = beta ? beta[b] : 1.0;
    const double ua = u ? u[k] : philox_exchange(seed, (uint32_t)b, (uint64_t)prop);
    const double improvement = bb * (chi0[b] - chi0[a]) + ba * (chi0[a] - chi0[b]);
    swap = exp(improvement) > ua ? 1 : 0;
    if (swapped) swapped[k] = swap;
  }
  __syncthreads();
  if (!swap) return;
  for (int64_t j = threadIdx.x; j < d; j += EW_NT) {
    const T t = x0[(int64_t)a * ldx + j];
    x0[(int64_t)a * ldx + j] = x0[(int64_t)b * ldx + j];
    x0[(int64_t)b * ldx + j] = t;
  }
  if (slot >= 0) {   // the sample recorded for this proposal is the post-exchange state
    for (int64_t j = threadIdx.x; j < d; j += EW_NT) {
      tr_x[(slot * B + a) * d + j] = x0[(int64_t)a * ldx + j];
      tr_x[(slot * B + b) * d + j] = x0[(int64_t)b * ldx + j];
    }
  }
  if (threadIdx.x == 0) {
    const double t = chi0[a];
    chi0[a] = chi0[b];
    chi0[b] = t;
    if (ex_count) {
      ex_count[a] += 1;
      ex_count[b] += 1;
    }
    if (slot >= 0) {
      tr_chi[slot * B + a] = (beta ? beta[a] : 1.0) * chi0[a];
      tr_chi[slot * B + b] = (beta ? beta[b] : 1.0) * chi0[b];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_finish_misfit(const T* __restrict__ X, int64_t ldx,
                                                         int64_t d, const double* partial,
                                                         int64_t nmb, int64_t ldp, double mconst,
                                                         double* out, DevModel md) {
  __shared__ double scratch[EW_NT / 64 + 1];
  const int64_t c = blockIdx.x;
  const double chi = chain_misfit<T>(X + c * ldx, d, partial, nmb, ldp, c, mconst, md, scratch,
                                     nullptr, (const T*)nullptr);
  if (threadIdx.x == 0) out[c] = chi;
}

template <typename T>
__global__ void k_from_f64(const double* __restrict__ src, T* __restrict__ dst, uint16_t* h,
                           uint16_t* l, int64_t ldx, int64_t d) {
  const int64_t c = blockIdx.y;
  for (int64_t j = blockIdx.x * blockDim.x + threadIdx.x; j < d; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = c * ldx + j;
    const T v = (T)src[o];
    dst[o] = v;
    put_planes<T>(h, l, o, v);
  }
}

template <typename T>
__global__ void k_to_f64(const T* __restrict__ src, int64_t lds, double* __restrict__ dst,
                         int64_t ldd, int64_t d) {
  const int64_t c = blockIdx.y;
  for (int64_t j = blockIdx.x * blockDim.x + threadIdx.x; j < d; j += (int64_t)gridDim.x * blockDim.x)
    dst[c * ldd + j] = (double)src[c * lds + j];
}

template <typename T>
__global__ __launch_bounds__(EW_NT) void k_oob(const T* __restrict__ X, int64_t ldx, int64_t d,
                                               int32_t* oob, DevModel md) {
  __shared__ double scratch[EW_NT / 64 + 1];
  const int64_t c = blockIdx.x;
  double ob = 0.0;
  for (int64_t j = threadIdx.x; j < d; j += EW_NT)
    if (out_of_bounds(md, j, (double)X[c * ldx + j])) ob = 1.0;
  ob = block_sum<EW_NT>(ob, scratch);
  if (threadIdx.x == 0) oob[c] = ob > 0.0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
void elem_refresh(Chains& c, bool host_p, double drift0, uint64_t seed, int64_t prop,
                  bool device_scalars, int randomize) {
  Problem& p = *c.p;
  // planes for the first GEMM of the trajectory: hi/lo of x, or one plane of
  // x - x_ref for anchored trajectories (premultiplied: the stage buffer 0)
  uint16_t* ph = c.xh;
  uint16_t* pl = c.xl;
  const float* xa = nullptr;
  if (c.surrogate == nullptr && gemm_anchored(p)) {
    if (!c.anchor_ready) gemm_set_anchor(c, nullptr);
    xa = c.xa;
    pl = nullptr;
    if (traj_premult(p)) {
      ph = c.dxp[0];
      c.dx_cur = 0;
    }
  }
  hipEvent_t ev;
  time_begin(c, 2, ev);
  if (p.state_dtype == HMC_F64) {
    hipLaunchKernelGGL(k_refresh<double>, dim3(c.B), dim3(EW_NT), 0, p.stream, (double*)c.xt,
                       (double*)c.pt, ph, pl, (const double*)c.x0, c.p_host, c.ldx, p.d, c.k0,
                       c.eps, c.u, c.step, drift0, seed, prop, host_p ? 1 : 0,
                       device_scalars ? 1 : 0, randomize, xa, p.model());
  } else {
    hipLaunchKernelGGL(k_refresh<float>, dim3(c.B), dim3(EW_NT), 0, p.stream, (float*)c.xt,
                       (float*)c.pt, ph, pl, (const float*)c.x0, c.p_host, c.ldx, p.d, c.k0,
                       c.eps, c.u, c.step, drift0, seed, prop, host_p ? 1 : 0,
                       device_scalars ? 1 : 0, randomize, xa, p.model());
  }
  HMC_HIP(hipGetLastError());
  time_end(c, 2, ev);
}

void elem_drift(Chains& c, double drift) {
  Problem& p = *c.p;
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_drift<double>, dim3(c.B), dim3(EW_NT), 0, p.stream, (double*)c.xt,
                       (double*)c.pt, c.xh, c.xl, c.ldx, p.d, c.eps, drift, p.model());
  else
    hipLaunchKernelGGL(k_drift<float>, dim3(c.B), dim3(EW_NT), 0, p.stream, (float*)c.xt,
                       (float*)c.pt, c.xh, c.xl, c.ldx, p.d, c.eps, drift, p.model());
  HMC_HIP(hipGetLastError());
}

void elem_metropolis(Chains& c, bool have_partials, int autotune, double target,
                     double learning_rate, int64_t prop_index, int64_t trace_slot, int64_t hist_slot) {
  Problem& p = *c.p;
  const int64_t nmb = have_partials ? endpoint_blocks(c) : 0;
  const double mconst = have_partials ? endpoint_const(p) : p.misfit_const;
  hipEvent_t ev;
  time_begin(c, 2, ev);
  if (trace_slot >= 0) HMC_REQUIRE(trace_slot < c.tr_cap, HMC_ERR_STATE, "trace capacity exceeded");
  if (hist_slot >= 0) HMC_REQUIRE(hist_slot < c.hist_cap, HMC_ERR_STATE, "tuning history capacity exceeded");
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_metropolis<double>, dim3(c.B), dim3(EW_NT), 0, p.stream, (double*)c.x0,
                       c.chi0, (const double*)c.xt, (const double*)c.pt, c.ldx, p.d, c.partial, nmb,
                       c.B_pad, mconst, c.k0, c.u, c.h0, c.h1, c.acc, c.acc_count, c.step,
                       autotune, target, learning_rate, prop_index, (double*)c.tr_x, c.tr_chi,
                       trace_slot, c.B, (const double*)c.beta, c.hist_rate, c.hist_step, hist_slot,
                       p.model());
  else
    hipLaunchKernelGGL(k_metropolis<float>, dim3(c.B), dim3(EW_NT), 0, p.stream, (float*)c.x0,
                       c.chi0, (const float*)c.xt, (const float*)c.pt, c.ldx, p.d, c.partial, nmb,
                       c.B_pad, mconst, c.k0, c.u, c.h0, c.h1, c.acc, c.acc_count, c.step,
                       autotune, target, learning_rate, prop_index, (float*)c.tr_x, c.tr_chi,
                       trace_slot, c.B, (const double*)c.beta, c.hist_rate, c.hist_step, hist_slot,
                       p.model());
  HMC_HIP(hipGetLastError());
  time_end(c, 2, ev);
}

void elem_exchange(Chains& c, const int32_t* pairs, int64_t npairs, const double* u, uint64_t seed,
                   int64_t prop, int32_t* swapped, int64_t slot) {
  if (npairs <= 0) return;
  Problem& p = *c.p;
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_exchange<double>, dim3((unsigned)npairs), dim3(EW_NT), 0, p.stream, (double*)c.x0,
                       c.chi0, c.ldx, p.d, pairs, u, (const double*)c.beta, c.ex_count, swapped, seed, prop,
                       (double*)c.tr_x, c.tr_chi, slot, c.B);
  else
    hipLaunchKernelGGL(k_exchange<float>, dim3((unsigned)npairs), dim3(EW_NT), 0, p.stream, (float*)c.x0,
                       c.chi0, c.ldx, p.d, pairs, u, (const double*)c.beta, c.ex_count, swapped, seed, prop,
                       (float*)c.tr_x, c.tr_chi, slot, c.B);
  HMC_HIP(hipGetLastError());
}

void elem_finish_misfit(Chains& c, const void* X, bool have_partials, double* out) {
  Problem& p = *c.p;
  const int64_t nmb = have_partials ? endpoint_blocks(c) : 0;
  const double mconst = have_partials ? endpoint_const(p) : p.misfit_const;
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_finish_misfit<double>, dim3(c.B), dim3(EW_NT), 0, p.stream,
                       (const double*)X, c.ldx, p.d, c.partial, nmb, c.B_pad, mconst, out,
                       p.model());
  else
    hipLaunchKernelGGL(k_finish_misfit<float>, dim3(c.B), dim3(EW_NT), 0, p.stream,
                       (const float*)X, c.ldx, p.d, c.partial, nmb, c.B_pad, mconst, out,
                       p.model());
  HMC_HIP(hipGetLastError());
}

void elem_load_rows(Chains& c, const double* src, void* dst, uint16_t* hi, uint16_t* lo) {
  Problem& p = *c.p;
  const dim3 grid((unsigned)std::min<int64_t>((p.d + 255) / 256, 64), (unsigned)c.B);
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_from_f64<double>, grid, dim3(256), 0, p.stream, src, (double*)dst, hi, lo,
                       c.ldx, p.d);
  else
    hipLaunchKernelGGL(k_from_f64<float>, grid, dim3(256), 0, p.stream, src, (float*)dst, hi, lo,
                       c.ldx, p.d);
  HMC_HIP(hipGetLastError());
}

void elem_store_rows(Chains& c, const void* src, int src_dtype, int64_t ld, double* dst) {
  Problem& p = *c.p;
  const dim3 grid((unsigned)std::min<int64_t>((p.d + 255) / 256, 64), (unsigned)c.B);
  if (src_dtype == HMC_F64)
    hipLaunchKernelGGL(k_to_f64<double>, grid, dim3(256), 0, p.stream, (const double*)src, ld, dst,
                       c.ldx, p.d);
  else
    hipLaunchKernelGGL(k_to_f64<float>, grid, dim3(256), 0, p.stream, (const float*)src, ld, dst,
                       c.ldx, p.d);
  HMC_HIP(hipGetLastError());
}

void elem_oob_flags(Chains& c, const void* X) {
  Problem& p = *c.p;
  if (p.state_dtype == HMC_F64)
    hipLaunchKernelGGL(k_oob<double>, dim3(c.B), dim3(EW_NT), 0, p.stream, (const double*)X, c.ldx,
                       p.d, c.oob, p.model());
  else
    hipLaunchKernelGGL(k_oob<float>, dim3(c.B), dim3(EW_NT), 0, p.stream, (const float*)X, c.ldx,
                       p.d, c.oob, p.model());
  HMC_HIP(hipGetLastError());
}

void integrator_schedule(int integrator, int L, double& first_drift, std::vector<Stage>& st) {
  st.clear();
  if (integrator == HMC_INT_LEAPFROG) {
    // Samplers.py:1539-1581
    first_drift = 0.5;
    for (int s = 0; s < L - 1; ++s) st.push_back({1.0, 1.0, 0.0});
    st.push_back({1.0, 0.5, 0.0});
  } else if (integrator == HMC_INT_3STAGE) {
    // Samplers.py:1663-1726, schema (a1,b1,a2,b2,a2,b1,a1)
    const double a1 = 0.11888010966548, a2 = 1.0 / 2.0 - a1;
    const double b1 = 0.29619504261126, b2 = 1.0 - 2.0 * b1;
    first_drift = a1;
    for (int s = 0; s < L; ++s) {
      st.push_back({b1, a2, 0.0});
      st.push_back({b2, a2, 0.0});
      st.push_back({b1, a1, s + 1 < L ? a1 : 0.0});
    }
  } else if (integrator == HMC_INT_4STAGE) {
    // Samplers.py:1586-1661, schema (a1,b1,a2,b2,a3,b2,a2,b1,a1)
    const double a1 = 0.071353913450279725904, a2 = 0.268548791161230105820;
    const double a3 = 1.0 - 2.0 * a1 - 2.0 * a2;
    const double b1 = 0.191667800000000000000, b2 = 1.0 / 2.0 - b1;
    first_drift = a1;
    for (int s = 0; s < L; ++s) {
      st.push_back({b1, a2, 0.0});
      st.push_back({b2, a3, 0.0});
      st.push_back({b2, a2, 0.0});
      st.push_back({b1, a1, s + 1 < L ? a1 : 0.0});
    }
  } else {
    throw Error{HMC_ERR_ARG, "unknown integrator"};
  }
}

}  // namespace hmc

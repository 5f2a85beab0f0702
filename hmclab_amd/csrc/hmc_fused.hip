// Fused trajectory engine: one workgroup owns a tile of CT chains for a whole
// proposal.  Positions live in LDS as Xs[d][CT] (parameter-major inside the
// tile), residuals as Rs[n][CT], momenta in registers.  Each leapfrog step is
//   forward  : Rs[i][c] = (sum_k G[i][k] Xs[k][c] - d_i) / var_i
//   backward : g[j][c] = sum_i G[i][j] Rs[i][c] + prior'(x); kick; drift;
//              corrector (Samplers.py:1560-1581, LinearMatrix.py:208,426)
// with one barrier after each pass; the FP64 endpoint misfit
// (LinearMatrix.py:193-202,405-415), kinetic energy and the Metropolis test
// (Samplers.py:1471-1492) finish the proposal inside the same launch.
//
// Operators: SELL-64 (rows sorted by length, 64-row slices stored lane-
// contiguous, one wavefront per slice) for sparse ray matrices, and a
// k-major dense layout (element (row, k) at base[k*ld + row]) for small dense
// G.  The forward pass uses the operator of G, the backward pass the
// operator of G^T; thread (wave w, lane r) owns the parameters of the G^T
// slices w, w+NW, ... for the whole proposal.
#include "hmc_internal.hpp"

namespace hmc {

// Threads per workgroup: 1024 (16 waves) for float32 sparse operators, whose
// 4-chain tiles fill a CU's LDS with one workgroup (16 waves overlap the
// matrix-load -> gather -> FMA chain of each row; 0.85 -> 0.79 ms per
// proposal on the 64x64 tomography), 512 otherwise (float64 and the dense
// operators would spill at the 128 VGPRs of 16 waves)
template <typename T, bool SPARSE>
constexpr int fu_nt() { return (SPARSE && sizeof(T) == 4) ? 1024 : 512; }
// waves per workgroup of the fused engine for a problem (host side)
inline int fu_nw(const Problem& p) { return (p.op == 2 && p.state_dtype == HMC_F32) ? 16 : 8; }

struct DevSell {
  int64_t rows;       // operator rows (positions >= rows are padding)
  const int32_t* base;
  const int32_t* width;
  const int32_t* perm;
  const int32_t* inv;      // position of each original row
  const uint16_t* col;
  int64_t nslices;
  int64_t total;      // stored entries (slices x 64 x width)
};

struct DevDense {
  int64_t ld;
  int64_t nrows, ncols;
};

template <int TP>
struct TypeOf;
template <>
struct TypeOf<HMC_F64> { using T = double; };
template <>
struct TypeOf<HMC_F32> { using T = float; };

// One SELL row of a lane: acc[c] += sum_k val[k] * in[col[k]][c], k ascending.
// Slices store each lane's entries in groups of SELL_G consecutive k: entry k
// of lane r sits at base + (k / SELL_G) * 64 * SELL_G + r * SELL_G + k % SELL_G,
// and slice widths are multiples of SELL_G, so one group is one 16-byte column
// load and 16-byte value loads (three vector-memory instructions per 8 entries
// for float32 values instead of sixteen).  The sum order is k ascending.
constexpr int SELL_G = kSellGroup;

// Buffer resource over n bytes of a matrix array (gfx9 raw buffer, stride 0:
// loads at byte offsets >= n return 0 and never fault).
__device__ inline __amdgpu_buffer_rsrc_t sell_rsrc(const void* p, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)n, 0x00020000);
}

// 16 bytes at byte offset voff + soff of a buffer.  The builtin's vector
// result is only ever bit-cast whole to a HIP vector type: subscripting it
// (q[i]) compiles to wrong code with this toolchain.
__device__ inline uint4 buf_load16(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// SELL_G consecutive values of one lane from byte offset voff + soff
template <typename Acc, typename V>
__device__ inline void load_group(__amdgpu_buffer_rsrc_t r, int voff, int soff, Acc out[SELL_G]) {
#pragma unroll
  for (int h = 0; h < SELL_G * (int)sizeof(V) / 16; ++h) {
    const uint4 q = buf_load16(r, voff + 16 * h, soff);
    if constexpr (sizeof(V) == 4) {
      const float4 f = __builtin_bit_cast(float4, q);
      out[4 * h] = (Acc)f.x;
      out[4 * h + 1] = (Acc)f.y;
      out[4 * h + 2] = (Acc)f.z;
      out[4 * h + 3] = (Acc)f.w;
    } else {
      const double2 f = __builtin_bit_cast(double2, q);
      out[2 * h] = (Acc)f.x;
      out[2 * h + 1] = (Acc)f.y;
    }
  }
}

// LDS byte address base + SCALE * (16-bit half HI of w): one v_mad_u32_u16 with
// op_sel choosing the half (no separate extract and shift of the column)
template <int HI, int SCALE>
__device__ inline uint32_t lds_slot(uint32_t w, uint32_t base) {
  uint32_t r;
  if constexpr (HI) asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(w), "I"(SCALE), "v"(base));
  else asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(w), "I"(SCALE), "v"(base));
  return r;
}

// The CT values of one LDS row as one vector read
template <typename Tin, int CT>
struct LdsRow {
  typedef Tin type __attribute__((ext_vector_type(CT)));
};
template <typename Tin>
struct LdsRow<Tin, 1> {
  typedef Tin type;
};

template <typename Tin, int CT>
__device__ inline typename LdsRow<Tin, CT>::type lds_row(uint32_t ad) {
  typedef typename LdsRow<Tin, CT>::type R;
  return *(const __attribute__((address_space(3))) R*)(uintptr_t)ad;
}

template <typename Acc, typename Tin, int CT>
__device__ inline Acc row_elem(const typename LdsRow<Tin, CT>::type& t, int c) {
  if constexpr (CT == 1) return (Acc)t;
  else return (Acc)t[c];
}

// One SELL row of a lane.  The loads of group k + 1 are issued before the
// gathers of group k, so the matrix-load latency overlaps the LDS gathers and
// FMAs of the group before.  Matrix loads are buffer loads: a per-lane byte
// offset plus a wave-uniform slice/group offset that advances on the scalar
// unit.  The prefetch after the last group reads the next slice's first group
// (or zeros past the array end) and is never used.
template <typename Acc, typename V, typename Tin, int CT>
__device__ inline void sell_row(const DevSell& S, const V* __restrict__ valp, int64_t base, int32_t width,
                                int lane, const Tin* __restrict__ in, Acc acc[CT]) {
  if (width <= 0) return;
  const __amdgpu_buffer_rsrc_t rc = sell_rsrc(S.col, S.total * 2);
  const __amdgpu_buffer_rsrc_t rv = sell_rsrc(valp, S.total * (int64_t)sizeof(V));
  const int vc = lane * SELL_G * 2, vv = lane * SELL_G * (int)sizeof(V);
  const uint32_t inb = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)in;
  constexpr int RB = CT * (int)sizeof(Tin);   // bytes of one LDS row
  int sc = (int)base * 2, sv = (int)base * (int)sizeof(V);
  // one group: all addresses, then all LDS reads, then the FMAs (k ascending per chain)
  auto group = [&](const uint4& cw, const Acc* vl) {
    const uint32_t wd[4] = {cw.x, cw.y, cw.z, cw.w};
    uint32_t ad[SELL_G];
#pragma unroll
    for (int h = 0; h < SELL_G / 2; ++h) {
      ad[2 * h] = lds_slot<0, RB>(wd[h], inb);
      ad[2 * h + 1] = lds_slot<1, RB>(wd[h], inb);
    }
    typename LdsRow<Tin, CT>::type t[SELL_G];
#pragma unroll
    for (int u = 0; u < SELL_G; ++u) t[u] = lds_row<Tin, CT>(ad[u]);
#pragma unroll
    for (int u = 0; u < SELL_G; ++u)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[c] += vl[u] * row_elem<Acc, Tin, CT>(t[u], c);
  };
  // two register sets in turn (no copies between groups)
  uint4 ca = buf_load16(rc, vc, sc), cb;
  Acc va[SELL_G], vb[SELL_G];
  load_group<Acc, V>(rv, vv, sv, va);
  for (int32_t k = 0;;) {
    sc += 64 * SELL_G * 2;
    sv += 64 * SELL_G * (int)sizeof(V);
    cb = buf_load16(rc, vc, sc);
    load_group<Acc, V>(rv, vv, sv, vb);
    group(ca, va);
    k += SELL_G;
    if (k >= width) break;
    sc += 64 * SELL_G * 2;
    sv += 64 * SELL_G * (int)sizeof(V);
    ca = buf_load16(rc, vc, sc);
    load_group<Acc, V>(rv, vv, sv, va);
    group(cb, vb);
    k += SELL_G;
    if (k >= width) break;
  }
}

// Forward passes over the wave's slices.  Output position pos = slice * 64 +
// lane (-1 for padding lanes): the sparse operator's rows in SELL position
// order, the dense operator's rows in order.  pre(pos) runs before the row, so
// the loads it issues overlap the row's; then f(pos, pre(pos), acc[CT]).
template <typename Acc, typename V, typename Tin, int CT, int NW, class Pre, class F>
__device__ inline void sell_pass(const DevSell& S, const V* __restrict__ val,
                                 const Tin* __restrict__ in, int w, int lane, Pre&& pre, F&& f) {
  for (int64_t s = w; s < S.nslices; s += NW) {
    const int32_t base = S.base[s], width = S.width[s];
    const int64_t pos = (s * 64 + lane < S.rows) ? s * 64 + lane : -1;
    const auto pv = pre(pos);
    Acc acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = Acc(0);
    sell_row<Acc, V, Tin, CT>(S, val, base, width, lane, in, acc);
    f(pos, pv, acc);
  }
}

// One dense row of a lane over the k-major layout (element (row, k) at
// mp[k * ld]): acc[c] += sum_k mp[k * ld] * in[k][c], k ascending, with the
// matrix loads of SELL_U consecutive k issued before their use.
constexpr int SELL_U = 8;
template <typename Acc, typename V, typename Tin, int CT>
__device__ inline void dense_row(const V* __restrict__ mp, int64_t ld, int64_t ncols,
                                 const Tin* __restrict__ in, Acc acc[CT]) {
  int64_t k = 0;
  for (; k + SELL_U <= ncols; k += SELL_U) {
    Acc vl[SELL_U];
#pragma unroll
    for (int u = 0; u < SELL_U; ++u) vl[u] = (Acc)mp[(k + u) * ld];
#pragma unroll
    for (int u = 0; u < SELL_U; ++u)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[c] += vl[u] * (Acc)in[(k + u) * CT + c];
  }
  for (; k < ncols; ++k) {
    const Acc v = (Acc)mp[k * ld];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] += v * (Acc)in[k * CT + c];
  }
}

template <typename Acc, typename V, typename Tin, int CT, int NW, class Pre, class F>
__device__ inline void dense_pass(const DevDense& D, const V* __restrict__ M,
                                  const Tin* __restrict__ in, int w, int lane, Pre&& pre, F&& f) {
  const int64_t nsl = (D.nrows + 63) / 64;
  for (int64_t s = w; s < nsl; s += NW) {
    const int64_t row = s * 64 + lane;
    const int64_t pos = row < D.nrows ? row : -1;
    const auto pv = pre(pos);
    Acc acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = Acc(0);
    if (row < D.nrows) dense_row<Acc, V, Tin, CT>(M + row, D.ld, D.ncols, in, acc);
    f(pos, pv, acc);
  }
}

// Operator bundle: forward = G (rows = data), backward = G^T (rows = params).
template <typename T, typename VT, bool SPARSE>
struct Ops {
  // sparse
  DevSell fwd_s, bwd_s;
  // dense (k-major views): forward base = Gt (ld = ldgt), backward base = G (ld = ldg)
  DevDense fwd_d, bwd_d;
  const T* fwd_val;     // trajectory precision values
  const T* bwd_val;
  const VT* fwd_tgt;    // target precision values of the forward operator
  int64_t n, d;
  // premultiplied trajectories (dense): H = G^T C^-1 G in trajectory precision,
  // row-major [d_pad][ldh] (symmetric, so also its k-major view), and b; or null
  const T* h_val;
  int64_t ldh;
  const double* hb;

  template <typename Acc, typename V, typename Tin, int CT, int NW, class Pre, class F>
  __device__ void forward(const V* val, const Tin* in, int w, int lane, Pre&& pre, F&& f) const {
    if constexpr (SPARSE) sell_pass<Acc, V, Tin, CT, NW>(fwd_s, val, in, w, lane, pre, f);
    else dense_pass<Acc, V, Tin, CT, NW>(fwd_d, val, in, w, lane, pre, f);
  }
  // Parameter slots: the LDS row (Xs[x][CT]) and model-array index of a
  // parameter.  Sparse: the parameter's backward-slice position q (the model
  // arrays passed to the kernel are the position-ordered copies); dense: the
  // parameter index itself.
  // slot owned by (wave slice s, lane), -1 for padding lanes
  __device__ int64_t slot_of(int64_t s, int lane) const {
    const int64_t x = s * 64 + lane;
    return x < d ? x : -1;
  }
  // parameter index of slot x (0 <= x < d)
  __device__ int64_t param_at(int64_t x) const {
    if constexpr (SPARSE) return bwd_s.perm[x];
    else return x;
  }
  // slot of parameter j (0 <= j < d)
  __device__ int64_t slot_at(int64_t j) const {
    if constexpr (SPARSE) return bwd_s.inv[j];
    else return j;
  }
  __device__ int64_t param_slices() const {
    if constexpr (SPARSE) return bwd_s.nslices;
    else return (d + 63) / 64;
  }
};

template <typename T>
struct T2 {
  T a, b;
};

// Model constants of one parameter slot, loaded before the slot's backward
// row so that their latency overlaps the row's streaming.  The helpers below
// apply the same conversions and operations as prior_grad, out_of_bounds,
// kgrad and corrector (absent bounds become -inf/+inf and unit mass 1, which
// leave every comparison and product unchanged).
template <typename T>
struct ParamCst {
  T a[kMaxPriors], b[kMaxPriors];
  T minv;
  double lo, hi;
};

template <typename T>
__device__ inline ParamCst<T> param_cst(const DevModel& md, int64_t x) {
  ParamCst<T> c;
#pragma unroll
  for (int k = 0; k < kMaxPriors; ++k) {
    const bool on = k < md.n_priors && x >= 0;
    c.a[k] = on ? (T)md.pr[k].a[x] : T(0);
    c.b[k] = on ? (T)md.pr[k].b[x] : T(0);
  }
  c.minv = (md.minv && x >= 0) ? (T)md.minv[x] : T(1);
  c.lo = (md.lo && x >= 0) ? md.lo[x] : -INFINITY;
  c.hi = (md.hi && x >= 0) ? md.hi[x] : INFINITY;
  return c;
}

template <typename T>
__device__ inline T prior_grad_c(const DevModel& md, const ParamCst<T>& pc, T x) {
  T g = T(0);
#pragma unroll
  for (int k = 0; k < kMaxPriors; ++k) {
    if (k >= md.n_priors) break;
    if (md.pr[k].kind == HMC_PRIOR_NORMAL) {
      g += -pc.b[k] * (pc.a[k] - x);
    } else {
      T sg = (x > pc.a[k]) ? T(1) : ((x < pc.a[k]) ? T(-1) : T(0));
      g += sg * pc.b[k];
    }
  }
  return g;
}

template <typename T>
__device__ inline void corrector_c(const ParamCst<T>& pc, T& x, T& p) {
  const T lo = (T)pc.lo, hi = (T)pc.hi;
  if (x < lo) {
    x += T(2) * (lo - x);
    p *= T(-1);
  }
  if (x > hi) {
    x += T(2) * (hi - x);
    p *= T(-1);
  }
}

struct FusedArgs {
  void* x0;            // [B][ldx] state dtype
  double* chi0;
  void* x_out;         // trajectory end positions or nullptr
  void* p_out;         // trajectory end momenta or nullptr
  const double* p_host;// host-supplied momenta [B][ldx] or nullptr
  double* eps;
  double* u;
  double* step;
  double* h0;
  double* h1;
  int32_t* acc;
  int64_t* acc_count;
  void* tr_x;
  double* tr_chi;
  int64_t slot;
  double* hist_rate;   // autotuning history [hist][B] (Samplers.py:1496-1497) or null
  double* hist_step;
  int64_t hist_slot;
  int64_t ldx, B;
  const double* dobs;
  const double* var;
  double mconst;
  uint64_t seed;
  int64_t prop;
  int device_scalars, randomize, autotune;
  const double* beta;   // per-chain inverse temperature (null: 1)
  double target, lr;
  // schedule (units of the local step): L steps of P stages; every step but
  // the last uses pat, the last uses pat_last; each stage = (kick, drift, drift2)
  double first_drift;
  int L, P;
  double pat[4][3];
  double pat_last[4][3];
};

template <typename T, typename VT, bool SPARSE, int CT, int KMAX>
__global__ __launch_bounds__((fu_nt<T, SPARSE>())) void k_fused_propose(Ops<T, VT, SPARSE> op, FusedArgs a,
                                                         DevModel md) {
  constexpr int NT = fu_nt<T, SPARSE>(), NW = NT / 64;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* Xs = reinterpret_cast<T*>(smem_raw);                         // [d][CT]
  T* Rs = Xs + round_up(op.d, 2) * CT;                            // [n][CT]
  __shared__ double scratch[NW + 1];
  __shared__ double s_k0[CT], s_eps[CT], s_chi1[CT], s_k1[CT], s_beta[CT];
  __shared__ int s_ok[CT];

  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t c0 = (int64_t)blockIdx.x * CT;
  const int64_t nps = op.param_slices();
  T* x0 = reinterpret_cast<T*>(a.x0);

  // ---- per-chain scalars
  if (tid < CT) {
    const int64_t c = c0 + tid;
    double e = 0.0;
    if (c < a.B) {
      if (a.device_scalars) {
        double ue, ua;
        philox_scalars(a.seed, (uint32_t)c, (uint64_t)a.prop, ue, ua);
        e = a.randomize ? ue * a.step[c] : a.step[c];
        a.eps[c] = e;
        a.u[c] = ua;
      } else {
        e = a.eps[c];
      }
    }
    s_eps[tid] = e;
    s_beta[tid] = (a.beta && c < a.B) ? a.beta[c] : 1.0;
  }
  // ---- load positions of the tile (parameter j into slot slot_at(j); the
  // global reads run in parameter order)
  for (int64_t j = tid; j < op.d; j += NT) {
    const int64_t x = op.slot_at(j);
#pragma unroll
    for (int c = 0; c < CT; ++c)
      Xs[x * CT + c] = (c0 + c < a.B) ? x0[(c0 + c) * a.ldx + j] : T(0);
  }
  __syncthreads();

  // ---- momentum draw (owned parameters), K0, first drift
  T preg[KMAX][CT];
  double kloc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) kloc[c] = 0.0;
#pragma unroll
  for (int ks = 0; ks < KMAX; ++ks) {
    const int64_t s = w + (int64_t)ks * NW;
    const int64_t x = (s < nps) ? op.slot_of(s, lane) : -1;
    const int64_t j = x >= 0 ? op.param_at(x) : -1;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      T pv = T(0);
      if (x >= 0 && c0 + c < a.B) {
        if (a.p_host) {
          pv = (T)a.p_host[(c0 + c) * a.ldx + j];
        } else {
          double z0, z1;
          philox_normal2(a.seed, (uint32_t)(c0 + c), (uint64_t)a.prop, (uint32_t)(j >> 1), z0, z1);
          const double z = (j & 1) ? z1 : z0;
          pv = (T)(md.msqrt ? md.msqrt[x] * z : z);
        }
        kloc[c] += kterm(md, x, (double)pv);
        T xv = Xs[x * CT + c];
        xv += (T)(a.first_drift * s_eps[c]) * kgrad<T>(md, x, pv);
        corrector<T>(md, x, xv, pv);
        Xs[x * CT + c] = xv;
      }
      preg[ks][c] = pv;
    }
  }
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const double s = block_sum<NT>(kloc[c], scratch);
    if (tid == 0) s_k0[c] = 0.5 * s;
  }
  __syncthreads();

  // ---- stages: gradient, kick, drift (+ drift2)
  for (int st = 0; st < a.L * a.P; ++st) {
    // the lane index as an opaque value per stage: without it every per-slot
    // address and model constant (invariant across stages) is hoisted out of
    // the loop and held in registers, which costs occupancy
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int step = st / a.P, k = st - step * a.P;
    const double* sc = (step + 1 < a.L) ? a.pat[k] : a.pat_last[k];
    const double kick = sc[0], drift = sc[1], drift2 = sc[2];
    // kick (gradient g without the prior), drift, reflect for owned slice ks
    // at slot x, with the slot's model constants pc
    auto update = [&](int ks, int64_t x, const T* g, const ParamCst<T>& pc) {
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const double e = s_eps[c];
        T xv = Xs[x * CT + c], pv = preg[ks][c];
        T gg = g[c] + prior_grad_c<T>(md, pc, xv);
        if ((double)xv < pc.lo || (double)xv > pc.hi) gg = (T)INFINITY;
        pv -= (T)(a.beta ? kick * e * s_beta[c] : kick * e) * gg;
        xv += (T)(drift * e) * (pc.minv * pv);
        corrector_c<T>(pc, xv, pv);
        if (drift2 != 0.0) {
          xv += (T)(drift2 * e) * (pc.minv * pv);
          corrector_c<T>(pc, xv, pv);
        }
        Xs[x * CT + c] = xv;
        preg[ks][c] = pv;
      }
    };
    if constexpr (!SPARSE && KMAX <= 2) {   // the gradients of more slices would spill
      if (op.h_val) {
        // premultiplied: g = H x - b for every owned parameter first (all
        // threads read all of x), then the updates
        T gk[KMAX][CT];
#pragma unroll
        for (int ks = 0; ks < KMAX; ++ks) {
          const int64_t s = w + (int64_t)ks * NW;
#pragma unroll
          for (int c = 0; c < CT; ++c) gk[ks][c] = T(0);
          const int64_t row = s * 64 + ln;
          if (s < nps && row < op.d) {
            dense_row<T, T, T, CT>(op.h_val + row, op.ldh, op.d, Xs, gk[ks]);
            const T bj = (T)op.hb[row];
#pragma unroll
            for (int c = 0; c < CT; ++c) gk[ks][c] -= bj;
          }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KMAX; ++ks) {
          const int64_t s = w + (int64_t)ks * NW;
          if (s < nps) {
            const int64_t x = op.slot_of(s, ln);
            if (x >= 0) update(ks, x, gk[ks], param_cst<T>(md, x));
          }
        }
        __syncthreads();
        continue;
      }
    }
    // data and variance of output position i (position-ordered arrays)
    auto data_at = [&](int64_t i) {
      return i >= 0 ? T2<T>{(T)a.dobs[i], (T)a.var[i]} : T2<T>{T(0), T(1)};
    };
    op.template forward<T, T, T, CT, NW>(op.fwd_val, Xs, w, ln, data_at, [&](int64_t i, T2<T> dv, T* acc) {
      if (i >= 0) {
#pragma unroll
        for (int c = 0; c < CT; ++c) Rs[i * CT + c] = (acc[c] - dv.a) / dv.b;
      }
    });
    __syncthreads();
    // backward over owned slices (unrolled so momenta stay in registers)
#pragma unroll
    for (int ks = 0; ks < KMAX; ++ks) {
      const int64_t s = w + (int64_t)ks * NW;
      if (s < nps) {
        const int64_t x = op.slot_of(s, ln);
        const ParamCst<T> pc = param_cst<T>(md, x);
        T g[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) g[c] = T(0);
        if constexpr (SPARSE) {
          sell_row<T, T, T, CT>(op.bwd_s, op.bwd_val, op.bwd_s.base[s], op.bwd_s.width[s], ln, Rs, g);
        } else {
          const int64_t row = s * 64 + ln;
          if (row < op.d) dense_row<T, T, T, CT>(op.bwd_val + row, op.bwd_d.ld, op.n, Rs, g);
        }
        if (x >= 0) update(ks, x, g, pc);
      }
    }
    __syncthreads();
  }

  // ---- endpoint: data misfit in FP64
  double dm[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) dm[c] = 0.0;
  if (op.n > 0) {
    auto data_at = [&](int64_t i) {
      return i >= 0 ? T2<double>{a.dobs[i], a.var[i]} : T2<double>{0.0, 1.0};
    };
    op.template forward<double, VT, T, CT, NW>(op.fwd_tgt, Xs, w, lane, data_at, [&](int64_t i, T2<double> dv,
                                                                                  double* acc) {
      if (i >= 0) {
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          const double rr = acc[c] - dv.a;
          dm[c] += rr * rr / dv.b;
        }
      }
    });
  }
  // prior misfit, bounds, K1 over owned parameters
  double pm[CT], kk[CT], ob[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) pm[c] = kk[c] = ob[c] = 0.0;
#pragma unroll
  for (int ks = 0; ks < KMAX; ++ks) {
    const int64_t s = w + (int64_t)ks * NW;
    const int64_t x = (s < nps) ? op.slot_of(s, lane) : -1;
    if (x >= 0) {
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const double xv = (double)Xs[x * CT + c];
        pm[c] += prior_misfit(md, x, xv);
        if (out_of_bounds(md, x, xv)) ob[c] = 1.0;
        kk[c] += kterm(md, x, (double)preg[ks][c]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const double sd = block_sum<NT>(dm[c], scratch);
    const double sp = block_sum<NT>(pm[c], scratch);
    const double sb = block_sum<NT>(ob[c], scratch);
    const double sk = block_sum<NT>(kk[c], scratch);
    if (tid == 0) {
      double chi = 0.5 * sd + (sp + a.mconst);
      if (sb > 0.0) chi += INFINITY;
      s_chi1[c] = chi;
      s_k1[c] = 0.5 * sk;
    }
  }
  __syncthreads();
  if (tid < CT) {
    const int c = tid;
    const int64_t ch = c0 + c;
    int ok = 0;
    if (ch < a.B) {
      const double H0 = s_beta[c] * a.chi0[ch] + s_k0[c];
      const double H1 = s_beta[c] * s_chi1[c] + s_k1[c];
      const double av = exp(H0 - H1);
      if (a.autotune) {
        if (a.hist_slot >= 0) {
          a.hist_rate[a.hist_slot * a.B + ch] = av;
          a.hist_step[a.hist_slot * a.B + ch] = a.step[ch];
        }
        const double wgt = pow((double)(a.prop + 1), -a.lr);
        const double ar = isnan(av) ? 0.0 : av;
        double sv = a.step[ch] - wgt * (a.target - fmin(ar, 1.0));
        if (sv <= 0.0) sv = fmax(sv, 1e-18);
        a.step[ch] = sv;
      }
      ok = (av > a.u[ch]) ? 1 : 0;
      if (a.h0) a.h0[ch] = H0;
      if (a.h1) a.h1[ch] = H1;
      if (a.acc) a.acc[ch] = ok;
      if (a.acc_count) a.acc_count[ch] += ok;
      if (ok) a.chi0[ch] = s_chi1[c];
      if (a.slot >= 0) a.tr_chi[a.slot * a.B + ch] = s_beta[c] * a.chi0[ch];
    }
    s_ok[c] = ok;
  }
  __syncthreads();
  // ---- write back accepted positions, trace and trajectory end positions in
  // parameter order (coalesced global writes), trajectory end momenta from
  // the owning slots' registers
  T* xo = reinterpret_cast<T*>(a.x_out);
  T* po = reinterpret_cast<T*>(a.p_out);
  T* tr = reinterpret_cast<T*>(a.tr_x);
  for (int64_t j = tid; j < op.d; j += NT) {
    const int64_t x = op.slot_at(j);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int64_t ch = c0 + c;
      if (ch < a.B) {
        const int64_t o = ch * a.ldx + j;
        const T xv = Xs[x * CT + c];
        if (xo) xo[o] = xv;
        if (s_ok[c]) x0[o] = xv;
        if (a.slot >= 0) tr[(a.slot * a.B + ch) * op.d + j] = s_ok[c] ? xv : x0[o];
      }
    }
  }
  if (po) {
#pragma unroll
    for (int ks = 0; ks < KMAX; ++ks) {
      const int64_t s = w + (int64_t)ks * NW;
      const int64_t x = (s < nps) ? op.slot_of(s, lane) : -1;
      if (x >= 0) {
        const int64_t j = op.param_at(x);
#pragma unroll
        for (int c = 0; c < CT; ++c)
          if (c0 + c < a.B) po[(c0 + c) * a.ldx + j] = preg[ks][c];
      }
    }
  }
}

// misfit (FP64) and gradient (trajectory precision) of model rows
template <typename T, typename VT, bool SPARSE, int CT, int KMAX>
__global__ __launch_bounds__((fu_nt<T, SPARSE>())) void k_fused_eval(Ops<T, VT, SPARSE> op, const T* X, int64_t ldx,
                                                      int64_t B, const double* dobs, const double* var,
                                                      double mconst, double* chi_out, double* g_out,
                                                      DevModel md) {
  constexpr int NT = fu_nt<T, SPARSE>(), NW = NT / 64;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  T* Xs = reinterpret_cast<T*>(smem_raw);
  T* Rs = Xs + round_up(op.d, 2) * CT;
  __shared__ double scratch[NW + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t c0 = (int64_t)blockIdx.x * CT;
  const int64_t nps = op.param_slices();
  for (int64_t j = tid; j < op.d; j += NT) {
    const int64_t x = op.slot_at(j);
#pragma unroll
    for (int c = 0; c < CT; ++c) Xs[x * CT + c] = (c0 + c < B) ? X[(c0 + c) * ldx + j] : T(0);
  }
  __syncthreads();
  // chain out-of-bounds flags (whole-vector semantics of misfit_bounds)
  double obl[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) obl[c] = 0.0;
  for (int64_t x = tid; x < op.d; x += NT)
#pragma unroll
    for (int c = 0; c < CT; ++c)
      if (out_of_bounds(md, x, (double)Xs[x * CT + c])) obl[c] = 1.0;
  double obc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) obc[c] = block_sum<NT>(obl[c], scratch);

  if (chi_out) {
    double dm[CT], pm[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) dm[c] = pm[c] = 0.0;
    if (op.n > 0) {
      auto data_at = [&](int64_t i) { return i >= 0 ? T2<double>{dobs[i], var[i]} : T2<double>{0.0, 1.0}; };
      op.template forward<double, VT, T, CT, NW>(op.fwd_tgt, Xs, w, lane, data_at, [&](int64_t i, T2<double> dv,
                                                                                    double* acc) {
        if (i >= 0) {
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const double rr = acc[c] - dv.a;
            dm[c] += rr * rr / dv.b;
          }
        }
      });
    }
    for (int64_t x = tid; x < op.d; x += NT)
#pragma unroll
      for (int c = 0; c < CT; ++c) pm[c] += prior_misfit(md, x, (double)Xs[x * CT + c]);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const double sd = block_sum<NT>(dm[c], scratch);
      const double sp = block_sum<NT>(pm[c], scratch);
      if (tid == 0 && c0 + c < B) {
        double chi = 0.5 * sd + (sp + mconst);
        if (obc[c] > 0.0) chi += INFINITY;
        chi_out[c0 + c] = chi;
      }
    }
  }
  if (g_out) {
    if (op.n > 0) {
      auto data_at = [&](int64_t i) { return i >= 0 ? T2<T>{(T)dobs[i], (T)var[i]} : T2<T>{T(0), T(1)}; };
      op.template forward<T, T, T, CT, NW>(op.fwd_val, Xs, w, lane, data_at, [&](int64_t i, T2<T> dv, T* acc) {
        if (i >= 0) {
#pragma unroll
          for (int c = 0; c < CT; ++c) Rs[i * CT + c] = (acc[c] - dv.a) / dv.b;
        }
      });
      __syncthreads();
    }
    for (int64_t s = w; s < nps; s += NW) {
      T g[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) g[c] = T(0);
      if (op.n > 0) {
        if constexpr (SPARSE) {
          sell_row<T, T, T, CT>(op.bwd_s, op.bwd_val, op.bwd_s.base[s], op.bwd_s.width[s], lane, Rs, g);
        } else {
          const int64_t row = s * 64 + lane;
          if (row < op.d) dense_row<T, T, T, CT>(op.bwd_val + row, op.bwd_d.ld, op.n, Rs, g);
        }
      }
      const int64_t x = op.slot_of(s, lane);
      if (x >= 0) {
        const int64_t j = op.param_at(x);
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          if (c0 + c < B) {
            double gg = (double)(g[c] + prior_grad<T>(md, x, Xs[x * CT + c]));
            if (obc[c] > 0.0) gg += INFINITY;
            g_out[(c0 + c) * ldx + j] = gg;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int fused_ct(const Problem& p, int64_t B, size_t& lds_bytes) {
  const size_t es = dtype_size(p.state_dtype);
  int dev = 0;
  HMC_HIP(hipGetDevice(&dev));
  int maxlds = 65536;
  HMC_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t budget = (size_t)maxlds - 2048;
  // float32 sparse operators prefer 4 chains per workgroup once there are
  // enough chains for a workgroup per CU: each matrix byte streamed from L2
  // serves 4 chains (64x64 tomography, 2000 rays, 1024 chains: 0.93 ms per
  // proposal against 1.34 ms with 2); otherwise, and for dense operators,
  // 2 chains (half the LDS per workgroup, two workgroups per CU)
  int n_cu = 256;
  HMC_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  const int cands_sparse_f32[] = {4, 2, 1};
  const int cands_f32[] = {2, 4, 1};
  const int cands_f64[] = {2, 1};
  const int* cands = p.state_dtype == HMC_F64 ? cands_f64
                     : (p.op == 2 && B >= 4 * (int64_t)n_cu) ? cands_sparse_f32 : cands_f32;
  const int ncand = p.state_dtype == HMC_F64 ? 2 : 3;
  for (int i = 0; i < ncand; ++i) {
    const int ct = cands[i];
    if (p.fused_ct > 0 && ct != p.fused_ct) continue;
    const size_t bytes = (size_t)(round_up(p.d, 2) + round_up(p.n, 2)) * ct * es;
    if (bytes > budget) continue;
    if (ct > 1 && B < ct) continue;
    lds_bytes = bytes;
    return ct;
  }
  if (p.fused_ct > 0)
    throw Error{HMC_ERR_UNSUPPORTED, "fused_tile_chains does not fit the fused engine's LDS tile"};
  throw Error{HMC_ERR_UNSUPPORTED, "problem too large for the fused engine's LDS tile"};
}

static int fused_kmax(const Problem& p) {
  const int64_t nslices = (p.d + 63) / 64;
  const int nw = fu_nw(p);
  const int64_t per_wave = (nslices + nw - 1) / nw;
  if (per_wave <= 2) return 2;
  if (nw == 16 && per_wave <= 4) return 4;   // instantiated for the 16-wave (float32 sparse) kernels only
  if (per_wave <= 8) return 8;
  if (per_wave <= 16) return 16;
  throw Error{HMC_ERR_UNSUPPORTED, "fused engine supports up to 8192 parameters"};
}

bool fused_supported(const Problem& p) {
  if (p.precision == HMC_BF16) return false;
  if ((p.d + 63) / 64 > (int64_t)16 * 8) return false;   // 16 slices per wave of 8 (8192 parameters)
  return true;
}

template <typename T, typename VT, bool SPARSE>
static Ops<T, VT, SPARSE> make_ops(const Problem& p) {
  Ops<T, VT, SPARSE> op{};
  op.n = p.n;
  op.d = p.d;
  if constexpr (SPARSE) {
    op.fwd_s = DevSell{p.sg.rows, p.sg.base, p.sg.width, p.sg.perm, p.sg.inv, p.sg.col, p.sg.nslices, p.sg.total};
    op.bwd_s = DevSell{p.sgt.rows, p.sgt.base, p.sgt.width, p.sgt.perm, p.sgt.inv, p.sgt.col, p.sgt.nslices, p.sgt.total};
    op.fwd_val = (const T*)p.sg.val;
    op.bwd_val = (const T*)p.sgt.val;
    op.fwd_tgt = (const VT*)p.sg.val_tgt;
  } else {
    // forward over rows i: element (i, k=j) = G[i][j] = Gt[j][i] -> base Gt, ld = ldgt
    op.fwd_d = DevDense{p.ldgt, p.n, p.d};
    op.bwd_d = DevDense{p.ldg, p.d, p.n};
    op.fwd_val = (const T*)p.Gt;
    op.bwd_val = (const T*)p.G;
    op.fwd_tgt = (const VT*)p.Gt_tgt;
    if (fused_premult(p) && p.H) {
      op.h_val = sizeof(T) == 8 ? (const T*)p.H : (const T*)p.Hf;
      op.ldh = p.ldh;
      op.hb = p.hb;
    }
  }
  return op;
}

// Premultiplied fused trajectories (LinearMatrix.py:203-207, g = H x - b): the
// dense operator when premult_trajectory is 1, or -1 with n > d (the
// reference's rule); the endpoint keeps the FP64 residual form.
bool fused_premult(const Problem& p) {
  return p.op == 1 && p.n > 0 && p.premult_traj != 0 && (p.premult_traj > 0 || p.n > p.d) &&
         (p.d + 63) / 64 <= 2 * 8;   // at most 2 slices per wave of 8 (KMAX = 2)
}

__global__ void k_h_to_f32(const double* __restrict__ H, int64_t count, float* __restrict__ out) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < count; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = (float)H[k];
}

static void fused_premult_build(Chains& c) {
  Problem& p = *c.p;
  if (!fused_premult(p)) return;
  if (!p.H) premult_build(c);
  if (p.state_dtype != HMC_F64 && !p.Hf) {
    const int64_t cnt = p.d_pad * p.ldh;
    p.Hf = (float*)dev_alloc(p.owned, (size_t)cnt * 4);
    hipLaunchKernelGGL(k_h_to_f32, dim3((unsigned)std::min<int64_t>((cnt + 255) / 256, 4096)), dim3(256), 0,
                       p.stream, p.H, cnt, p.Hf);
    HMC_HIP(hipGetLastError());
  }
}

// Position-ordered model arrays of a sparse problem: slot q holds parameter
// perm[q] (padding slots: prior weights 0, inverse mass 0, mass root 0, bounds
// -inf/+inf, which leave their never-read LDS rows at 0)
__global__ void k_gather_model(const int32_t* __restrict__ perm, int64_t npos, DevModel src,
                               double* __restrict__ dst) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npos; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = perm[q];
    for (int k = 0; k < src.n_priors; ++k) {
      dst[(2 * k) * npos + q] = j >= 0 ? src.pr[k].a[j] : 0.0;
      dst[(2 * k + 1) * npos + q] = j >= 0 ? src.pr[k].b[j] : 0.0;
    }
    const int m = 2 * kMaxPriors;
    if (src.minv) dst[m * npos + q] = j >= 0 ? src.minv[j] : 0.0;
    if (src.msqrt) dst[(m + 1) * npos + q] = j >= 0 ? src.msqrt[j] : 0.0;
    if (src.lo) dst[(m + 2) * npos + q] = j >= 0 ? src.lo[j] : -INFINITY;
    if (src.hi) dst[(m + 3) * npos + q] = j >= 0 ? src.hi[j] : INFINITY;
  }
}

// The element-wise model in the kernels' slot order: the problem's own arrays
// for dense operators, the position-ordered copies (gathered again after any
// change of priors, bounds or mass) for sparse ones.
static DevModel slot_model(Problem& p) {
  DevModel m = p.model();
  if (p.op != 2) return m;
  const int64_t npos = p.sgt.nslices * 64;
  if (p.pos_model_version != p.model_version) {
    const int blocks = (int)std::min<int64_t>((npos + 255) / 256, 1024);
    hipLaunchKernelGGL(k_gather_model, dim3(blocks), dim3(256), 0, p.stream, (const int32_t*)p.sgt.perm, npos, m,
                       p.pos_model);
    HMC_HIP(hipGetLastError());
    p.pos_model_version = p.model_version;
  }
  DevModel q = m;
  for (int k = 0; k < m.n_priors; ++k) {
    q.pr[k].a = p.pos_model + (2 * k) * npos;
    q.pr[k].b = p.pos_model + (2 * k + 1) * npos;
  }
  const int mi = 2 * kMaxPriors;
  q.minv = m.minv ? p.pos_model + mi * npos : nullptr;
  q.msqrt = m.msqrt ? p.pos_model + (mi + 1) * npos : nullptr;
  q.lo = m.lo ? p.pos_model + (mi + 2) * npos : nullptr;
  q.hi = m.hi ? p.pos_model + (mi + 3) * npos : nullptr;
  return q;
}

template <typename T, typename VT, bool SPARSE, int CT, int KMAX>
static void launch_propose(Chains& c, const FusedArgs& a, size_t lds) {
  Problem& p = *c.p;
  auto kern = k_fused_propose<T, VT, SPARSE, CT, KMAX>;
  if (lds > 65536) HMC_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = (int)((c.B + CT - 1) / CT);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(fu_nt<T, SPARSE>()), lds, p.stream, make_ops<T, VT, SPARSE>(p), a,
                     slot_model(p));
  HMC_HIP(hipGetLastError());
}

template <typename T, typename VT, bool SPARSE, int CT, int KMAX>
static void launch_eval(Chains& c, const void* X, double* chi, double* g, size_t lds) {
  Problem& p = *c.p;
  auto kern = k_fused_eval<T, VT, SPARSE, CT, KMAX>;
  if (lds > 65536) HMC_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = (int)((c.B + CT - 1) / CT);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(fu_nt<T, SPARSE>()), lds, p.stream, make_ops<T, VT, SPARSE>(p),
                     (const T*)X, c.ldx, c.B, SPARSE ? p.pos_dobs : p.dobs, SPARSE ? p.pos_var : p.var,
                     p.misfit_const, chi, g, slot_model(p));
  HMC_HIP(hipGetLastError());
}

// dispatch over (state dtype, target dtype, sparse, CT, KMAX)
template <class F>
static void dispatch(const Problem& p, int ct, int kmax, F&& f) {
  auto by_k = [&](auto t_tag, auto v_tag, auto s_tag, auto ct_tag) {
    using T = typename decltype(t_tag)::type;
    if constexpr (decltype(s_tag)::value && std::is_same<T, float>::value) {
      if (kmax == 4) {
        f(t_tag, v_tag, s_tag, ct_tag, std::integral_constant<int, 4>{});
        return;
      }
    }
    if (kmax == 2) f(t_tag, v_tag, s_tag, ct_tag, std::integral_constant<int, 2>{});
    else if (kmax == 8) f(t_tag, v_tag, s_tag, ct_tag, std::integral_constant<int, 8>{});
    else f(t_tag, v_tag, s_tag, ct_tag, std::integral_constant<int, 16>{});
  };
  auto by_ct = [&](auto t_tag, auto v_tag, auto s_tag) {
    using T = typename decltype(t_tag)::type;
    if constexpr (std::is_same<T, double>::value) {
      if (ct == 2) by_k(t_tag, v_tag, s_tag, std::integral_constant<int, 2>{});
      else by_k(t_tag, v_tag, s_tag, std::integral_constant<int, 1>{});
    } else {
      if (ct == 4) by_k(t_tag, v_tag, s_tag, std::integral_constant<int, 4>{});
      else if (ct == 2) by_k(t_tag, v_tag, s_tag, std::integral_constant<int, 2>{});
      else by_k(t_tag, v_tag, s_tag, std::integral_constant<int, 1>{});
    }
  };
  auto by_s = [&](auto t_tag, auto v_tag) {
    if (p.op == 2) by_ct(t_tag, v_tag, std::true_type{});
    else by_ct(t_tag, v_tag, std::false_type{});
  };
  using DT = std::common_type<double>;
  using FT = std::common_type<float>;
  if (p.state_dtype == HMC_F64) {
    by_s(DT{}, DT{});
  } else {
    if (p.tgt_dtype == HMC_F64) by_s(FT{}, DT{});
    else by_s(FT{}, FT{});
  }
}

void fused_propose(Chains& c, bool host_p, int L, int integrator, uint64_t seed, int64_t prop,
                   int randomize, int autotune, double target, double learning_rate,
                   int64_t trace_slot, int64_t hist_slot) {
  Problem& p = *c.p;
  size_t lds = 0;
  const int ct = fused_ct(p, c.B, lds);
  const int kmax = fused_kmax(p);
  double first = 0.0;
  std::vector<Stage> st;
  integrator_schedule(integrator, L, first, st);
  const int P = (int)st.size() / L;
  if (trace_slot >= 0) HMC_REQUIRE(trace_slot < c.tr_cap, HMC_ERR_STATE, "trace capacity exceeded");
  if (hist_slot >= 0) HMC_REQUIRE(hist_slot < c.hist_cap, HMC_ERR_STATE, "tuning history capacity exceeded");
  fused_premult_build(c);

  FusedArgs a{};
  a.x0 = c.x0;
  a.chi0 = c.chi0;
  a.x_out = host_p ? c.xt : nullptr;   // trajectory end states: only a single propose returns them
  a.p_out = host_p ? c.pt : nullptr;
  a.p_host = host_p ? c.p_host : nullptr;
  a.eps = c.eps;
  a.u = c.u;
  a.step = c.step;
  a.beta = c.beta;
  a.h0 = c.h0;
  a.h1 = c.h1;
  a.acc = c.acc;
  a.acc_count = c.acc_count;
  a.tr_x = c.tr_x;
  a.tr_chi = c.tr_chi;
  a.slot = trace_slot;
  a.hist_rate = c.hist_rate;
  a.hist_step = c.hist_step;
  a.hist_slot = hist_slot;
  a.ldx = c.ldx;
  a.B = c.B;
  a.dobs = p.op == 2 ? p.pos_dobs : p.dobs;   // the sparse forward pass writes SELL positions
  a.var = p.op == 2 ? p.pos_var : p.var;
  a.mconst = p.misfit_const;
  a.seed = seed;
  a.prop = prop;
  a.device_scalars = host_p ? 0 : 1;
  a.randomize = randomize;
  a.autotune = autotune;
  a.target = target;
  a.lr = learning_rate;
  a.first_drift = first;
  a.L = L;
  a.P = P;
  for (int k = 0; k < P; ++k) {
    const Stage& s0 = st[k];
    const Stage& s1 = st[(size_t)(L - 1) * P + k];
    a.pat[k][0] = s0.kick; a.pat[k][1] = s0.drift; a.pat[k][2] = s0.drift2;
    a.pat_last[k][0] = s1.kick; a.pat_last[k][1] = s1.drift; a.pat_last[k][2] = s1.drift2;
  }
  hipEvent_t ev;
  time_begin(c, 0, ev);
  dispatch(p, ct, kmax, [&](auto t_tag, auto v_tag, auto s_tag, auto ct_tag, auto k_tag) {
    using T = typename decltype(t_tag)::type;
    using VT = typename decltype(v_tag)::type;
    launch_propose<T, VT, decltype(s_tag)::value, decltype(ct_tag)::value, decltype(k_tag)::value>(c, a, lds);
  });
  time_end(c, 0, ev);
}

void fused_misfit(Chains& c, const void* X, double* out) {
  Problem& p = *c.p;
  size_t lds = 0;
  const int ct = fused_ct(p, c.B, lds);
  const int kmax = fused_kmax(p);
  dispatch(p, ct, kmax, [&](auto t_tag, auto v_tag, auto s_tag, auto ct_tag, auto k_tag) {
    using T = typename decltype(t_tag)::type;
    using VT = typename decltype(v_tag)::type;
    launch_eval<T, VT, decltype(s_tag)::value, decltype(ct_tag)::value, decltype(k_tag)::value>(c, X, out, nullptr, lds);
  });
}

void fused_gradient(Chains& c, const void* X, double* out) {
  Problem& p = *c.p;
  size_t lds = 0;
  const int ct = fused_ct(p, c.B, lds);
  const int kmax = fused_kmax(p);
  dispatch(p, ct, kmax, [&](auto t_tag, auto v_tag, auto s_tag, auto ct_tag, auto k_tag) {
    using T = typename decltype(t_tag)::type;
    using VT = typename decltype(v_tag)::type;
    launch_eval<T, VT, decltype(s_tag)::value, decltype(ct_tag)::value, decltype(k_tag)::value>(c, X, nullptr, out, lds);
  });
}

}  // namespace hmc

// C-ABI entry points of libhmc.so.  Every function catches engine errors and
// returns a status code; hmc_last_error() returns the thread's last message.
#include <algorithm>
#include <array>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <new>
#include <vector>

#include "hmc_internal.hpp"

namespace hmc {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

size_t dtype_size(int dt) { return dt == HMC_F64 ? 8 : (dt == HMC_F32 ? 4 : 2); }

void* dev_alloc(std::vector<void*>& owned, size_t bytes) {
  void* ptr = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error{HMC_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                   " bytes failed: " + hipGetErrorString(e)};
  }
  const hipError_t z = hipMemset(ptr, 0, bytes);
  const hipError_t w = (z == hipSuccess) ? hipStreamSynchronize(nullptr) : z;
  if (w != hipSuccess) {
    (void)hipFree(ptr);
    HMC_HIP(w);
  }
  owned.push_back(ptr);
  return ptr;
}

void h2d(void* dst, const void* src, size_t bytes) {
  HMC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  HMC_HIP(hipStreamSynchronize(nullptr));
}

void d2h(void* dst, const void* src, size_t bytes) {
  HMC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  HMC_HIP(hipStreamSynchronize(nullptr));
}

void dzero(void* dst, size_t bytes) {
  HMC_HIP(hipMemset(dst, 0, bytes));
  HMC_HIP(hipStreamSynchronize(nullptr));
}

static void* upload(std::vector<void*>& owned, const void* host, size_t bytes, hipStream_t s) {
  void* d = dev_alloc(owned, bytes);
  HMC_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s));
  HMC_HIP(hipStreamSynchronize(s));
  return d;
}

// Fault injection for host-side tests of error paths: HMC_FAULT_INJECT=<name>
// makes the named entry point fail with HMC_ERR_INJECTED before any work.
void maybe_inject(const char* name) {
  const char* f = std::getenv("HMC_FAULT_INJECT");
  if (f && std::strcmp(f, name) == 0)
    throw Error{HMC_ERR_INJECTED, std::string("fault injected in ") + name};
}

// ---- value conversion ------------------------------------------------------
static double load_elem(const void* src, int dt, size_t i) {
  if (dt == HMC_F64) return static_cast<const double*>(src)[i];
  if (dt == HMC_F32) return static_cast<const float*>(src)[i];
  return bf16_bits_to_f32(static_cast<const uint16_t*>(src)[i]);
}
static void store_elem(void* dst, int dt, size_t i, double v) {
  if (dt == HMC_F64) static_cast<double*>(dst)[i] = v;
  else if (dt == HMC_F32) static_cast<float*>(dst)[i] = (float)v;
  else static_cast<uint16_t*>(dst)[i] = f32_to_bf16_bits((float)v);
}
static int prec_rank(int dt) { return dt == HMC_F64 ? 2 : (dt == HMC_F32 ? 1 : 0); }

// storage type that represents the input values exactly for the endpoint
static int target_store(int precision, int g_dtype) {
  return prec_rank(g_dtype) <= prec_rank(precision) ? precision : g_dtype;
}

int Problem::resolved_engine() const {
  if (op != 1) return HMC_ENGINE_FUSED;  // sparse and prior-only targets
  if (engine != HMC_ENGINE_AUTO) return engine;
  if (precision == HMC_BF16) return HMC_ENGINE_GEMM;
  const double bytes = (double)n * (double)d * (double)dtype_size(precision);
  if (bytes <= 4.0 * 1024 * 1024 && fused_supported(*this)) return HMC_ENGINE_FUSED;
  return HMC_ENGINE_GEMM;
}

// ---- SELL builder ------------------------------------------------------------
// Row order of a SELL-64 operator: rows sorted by decreasing length (stable),
// 64 consecutive rows per slice; position p = slice * 64 + lane holds row
// order[p].
static std::vector<int32_t> sell_order(const std::vector<int64_t>& ptr, int64_t rows) {
  std::vector<int32_t> order(rows);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return (ptr[a + 1] - ptr[a]) > (ptr[b + 1] - ptr[b]);
  });
  return order;
}

// Entry order of one slice.  At every step k each 16-lane group of the fused
// engine gathers 16 LDS rows; when their positions differ modulo 16, 16-byte
// rows (4 float32 or 2 float64 chains) fall on distinct LDS bank groups.  Per
// step and group, Kuhn's augmenting-path matching assigns residues to lanes
// from their remaining entries (residues with more entries left tried first);
// a lane left unmatched takes a padding entry on a free residue if it has one,
// else its next entry on its fullest residue.  seq[r][k] = entry of lane r at
// step k (index into its row), or -(1 + residue) for a padding entry.
static void sell_slice_order(const std::vector<std::vector<int32_t>>& cols, int W,
                             std::vector<std::vector<int32_t>>& seq, bool reorder) {
  constexpr int NL = 64, NG = 16;
  seq.assign(NL, std::vector<int32_t>(W, 0));
  if (!reorder) {   // CSR order, padding (column 0) after each row's entries
    for (int r = 0; r < NL; ++r)
      for (int k = 0; k < W; ++k) seq[r][k] = k < (int)cols[r].size() ? k : -1;
    return;
  }
  std::vector<std::array<std::vector<int32_t>, NG>> rem(NL);
  std::vector<std::array<int, NG>> head(NL);
  std::vector<int> pads(NL);
  for (int r = 0; r < NL; ++r) {
    head[r].fill(0);
    for (int e = 0; e < (int)cols[r].size(); ++e) rem[r][cols[r][e] & (NG - 1)].push_back(e);
    pads[r] = W - (int)cols[r].size();
  }
  auto left = [&](int r, int q) { return (int)rem[r][q].size() - head[r][q]; };
  for (int k = 0; k < W; ++k) {
    for (int g = 0; g < NL / NG; ++g) {
      int by_res[NG], by_lane[NG];
      std::array<std::array<int, NG>, NG> pref;   // residues of each lane, most entries left first
      for (int i = 0; i < NG; ++i) {
        by_res[i] = by_lane[i] = -1;
        std::iota(pref[i].begin(), pref[i].end(), 0);
        const int r = g * NG + i;
        std::stable_sort(pref[i].begin(), pref[i].end(), [&](int a, int b) { return left(r, a) > left(r, b); });
      }
      std::array<bool, NG> seen;
      std::function<bool(int)> augment = [&](int i) -> bool {
        const int r = g * NG + i;
        for (int q : pref[i]) {
          if (left(r, q) <= 0 || seen[q]) continue;
          seen[q] = true;
          if (by_res[q] < 0 || augment(by_res[q])) {
            by_res[q] = i;
            by_lane[i] = q;
            return true;
          }
        }
        return false;
      };
      for (int i = 0; i < NG; ++i) {
        seen.fill(false);
        augment(i);
      }
      for (int i = 0; i < NG; ++i) {
        const int r = g * NG + i;
        int q = by_lane[i];
        if (q >= 0 && by_res[q] == i) {
          seq[r][k] = rem[r][q][head[r][q]++];
        } else if (pads[r] > 0) {
          int fr = 0;
          while (fr < NG - 1 && by_res[fr] >= 0) ++fr;
          if (by_res[fr] < 0) by_res[fr] = i;
          --pads[r];
          seq[r][k] = -(1 + fr);
        } else {
          q = pref[i][0];
          for (int t = 0; t < NG; ++t)
            if (left(r, t) > left(r, q)) q = t;
          seq[r][k] = rem[r][q][head[r][q]++];
        }
      }
    }
  }
}

// Column indices are stored as colmap[column] (the fused engine keeps both
// operands in LDS in SELL position order); each slice's entries are ordered
// by sell_slice_order.
static void build_sell(Problem& P, Sell& S, int64_t rows, int64_t cols,
                       const std::vector<int64_t>& ptr, const std::vector<int32_t>& idx,
                       const std::vector<double>& val, const std::vector<int32_t>& order,
                       const std::vector<int32_t>& colmap, int traj_dt, int tgt_dt) {
  S.rows = rows;
  S.cols = cols;
  // 16-bit column positions: a sparse target runs on the fused engine, whose
  // LDS tile holds (rows + cols) values per chain, far below 65536 of either
  HMC_REQUIRE(cols <= kSellMaxDim && rows <= kSellMaxDim, HMC_ERR_UNSUPPORTED,
              "sparse operator with more than 65472 rows or columns exceeds the fused engine's LDS tile");
  S.nslices = (rows + 63) / 64;
  // Each lane's entries are stored in groups of kSellGroup consecutive k
  // (entry k of lane r at base + (k / G) * 64 * G + r * G + k % G), so slice
  // widths are rounded up to a multiple of the group; padding entries hold
  // column 0 and value 0.
  std::vector<int32_t> base(S.nslices), width(S.nslices), perm(S.nslices * 64, -1);
  int64_t total = 0;
  int maxw = 0;
  for (int64_t s = 0; s < S.nslices; ++s) {
    int w = 0;
    for (int r = 0; r < 64; ++r) {
      const int64_t k = s * 64 + r;
      if (k < rows) {
        perm[k] = order[k];
        w = std::max<int>(w, (int)(ptr[order[k] + 1] - ptr[order[k]]));
      }
    }
    w = (w + kSellGroup - 1) / kSellGroup * kSellGroup;
    HMC_REQUIRE(total <= INT32_MAX, HMC_ERR_UNSUPPORTED, "sparse matrix too large for int32 SELL offsets");
    base[s] = (int32_t)total;
    width[s] = w;
    total += (int64_t)w * 64;
    maxw = std::max(maxw, w);
  }
  // the fused engine reads the arrays with buffer loads at 32-bit byte offsets
  HMC_REQUIRE(total * 8 + 8192 <= INT32_MAX, HMC_ERR_UNSUPPORTED,
              "sparse matrix too large for the fused engine's 32-bit buffer offsets");
  S.total = total;
  S.max_width = maxw;
  std::vector<uint16_t> col(std::max<int64_t>(total, 1), 0);
  std::vector<unsigned char> vtraj(std::max<int64_t>(total, 1) * dtype_size(traj_dt), 0);
  std::vector<unsigned char> vtgt;
  const bool alias = (traj_dt == tgt_dt);
  if (!alias) vtgt.assign(std::max<int64_t>(total, 1) * dtype_size(tgt_dt), 0);
  // the ordering costs about 8e3 operations per 64 stored entries: beyond 2^24
  // entries (seconds of setup) the slices keep CSR order
  const bool reorder = total <= ((int64_t)1 << 24);
  std::vector<std::vector<int32_t>> lane_cols(64), seq;
  for (int64_t s = 0; s < S.nslices; ++s) {
    for (int r = 0; r < 64; ++r) {
      const int32_t row = perm[s * 64 + r];
      lane_cols[r].clear();
      if (row >= 0)
        for (int64_t k = ptr[row]; k < ptr[row + 1]; ++k) lane_cols[r].push_back(colmap[idx[k]]);
    }
    sell_slice_order(lane_cols, width[s], seq, reorder);
    for (int r = 0; r < 64; ++r) {
      const int32_t row = perm[s * 64 + r];
      for (int k = 0; k < width[s]; ++k) {
        const int64_t o = base[s] + (k / kSellGroup) * 64 * kSellGroup + r * kSellGroup + k % kSellGroup;
        const int32_t e = seq[r][k];
        if (e >= 0) {
          col[o] = (uint16_t)lane_cols[r][e];
          store_elem(vtraj.data(), traj_dt, o, val[ptr[row] + e]);
          if (!alias) store_elem(vtgt.data(), tgt_dt, o, val[ptr[row] + e]);
        } else {
          const int32_t q = -(e + 1);               // padding: value 0 on a free residue
          col[o] = (uint16_t)(q < cols ? q : 0);
        }
      }
    }
  }
  S.base = (int32_t*)upload(P.owned, base.data(), base.size() * 4, P.stream);
  S.width = (int32_t*)upload(P.owned, width.data(), width.size() * 4, P.stream);
  S.perm = (int32_t*)upload(P.owned, perm.data(), perm.size() * 4, P.stream);
  std::vector<int32_t> inv(std::max<int64_t>(rows, 1), 0);
  for (int64_t p = 0; p < rows; ++p) inv[order[p]] = (int32_t)p;
  S.inv = (int32_t*)upload(P.owned, inv.data(), inv.size() * 4, P.stream);
  S.col = (uint16_t*)upload(P.owned, col.data(), col.size() * 2, P.stream);
  S.val = upload(P.owned, vtraj.data(), vtraj.size(), P.stream);
  S.val_tgt = alias ? S.val : upload(P.owned, vtgt.data(), vtgt.size(), P.stream);
}

static void set_data(Problem& P, int64_t n, const double* d_obs, const double* var, int scalar) {
  HMC_REQUIRE(d_obs != nullptr && var != nullptr, HMC_ERR_ARG, "data and variance are required");
  std::vector<double> v(n);
  for (int64_t i = 0; i < n; ++i) {
    v[i] = scalar ? var[0] : var[i];
    HMC_REQUIRE(v[i] > 0.0, HMC_ERR_ARG, "data variance must be positive");
  }
  P.dobs = (double*)upload(P.owned, d_obs, n * 8, P.stream);
  P.var = (double*)upload(P.owned, v.data(), n * 8, P.stream);
  P.n = n;
  P.dtd = 0.0;
  for (int64_t i = 0; i < n; ++i) P.dtd += d_obs[i] * (d_obs[i] / v[i]);
}

// ---- timing ------------------------------------------------------------------
void time_begin(Chains& c, int kind, hipEvent_t& a) {
  (void)kind;
  a = nullptr;
  if (!c.timing || c.timing_skip) return;
  if (c.ev_next >= c.ev_pool.size()) {
    hipEvent_t e;
    HMC_HIP(hipEventCreate(&e));
    c.ev_pool.push_back(e);
  }
  a = c.ev_pool[c.ev_next++];
  HMC_HIP(hipEventRecord(a, c.p->stream));
}
void time_end(Chains& c, int kind, hipEvent_t a) {
  if (!c.timing || a == nullptr) return;
  if (c.ev_next >= c.ev_pool.size()) {
    hipEvent_t e;
    HMC_HIP(hipEventCreate(&e));
    c.ev_pool.push_back(e);
  }
  hipEvent_t b = c.ev_pool[c.ev_next++];
  HMC_HIP(hipEventRecord(b, c.p->stream));
  c.ev_pending.push_back({kind, {a, b}});
  if (c.ev_pending.size() > 4096) time_collect(c);
}
void time_collect(Chains& c) {
  if (c.ev_pending.empty()) {
    c.ev_next = 0;
    return;
  }
  HMC_HIP(hipStreamSynchronize(c.p->stream));
  for (auto& pr : c.ev_pending) {
    float ms = 0.f;
    HMC_HIP(hipEventElapsedTime(&ms, pr.second.first, pr.second.second));
    c.t_ms[pr.first] += ms;
    c.t_n[pr.first] += 1;
  }
  c.ev_pending.clear();
  c.ev_next = 0;
}

// After a synchronised call: a tile of the two-k-range fix-up whose partner
// did not publish its partial within the kernel's bounded wait sets the flag;
// that call's trajectories are then invalid.
static void check_fixup(Chains& c) {
  if (!c.fix_used) return;
  int flag = 0;
  d2h(&flag, c.fix, 4);
  if (flag) {
    const int zero = 0;
    h2d(c.fix, &zero, 4);
    throw Error{HMC_ERR_HIP, "bf16 GEMM split fix-up: a tile's partial was not published within the bounded "
                             "wait; the results of this call are invalid"};
  }
}

// ---- proposal drivers ----------------------------------------------------------
static void gemm_proposal(Chains& c, bool host_p, int L, int integ, uint64_t seed, int64_t prop,
                          int randomize, int autotune, double target, double lr, int64_t slot,
                          int64_t hist_slot = -1) {
  double first = 0.0;
  std::vector<Stage> st;
  integrator_schedule(integ, L, first, st);
  elem_refresh(c, host_p, first, seed, prop, !host_p, randomize);
  for (size_t k = 0; k < st.size(); ++k) {
    const Stage& s = st[k];
    if (c.surrogate) surrogate_kick_drift(c, s.kick, s.drift, s.drift2, k == 0);
    else gemm_kick_drift(c, s.kick, s.drift, s.drift2, k == 0);
  }
  gemm_endpoint(c, c.xt, c.p->state_dtype);
  elem_metropolis(c, true, autotune, target, lr, prop, slot, hist_slot);
}

static void evaluate_misfit(Chains& c, const void* X, double* out_dev) {
  Problem& P = *c.p;
  if (P.resolved_engine() == HMC_ENGINE_GEMM) {
    gemm_endpoint(c, X, P.state_dtype);
    elem_finish_misfit(c, X, true, out_dev);
  } else {
    fused_misfit(c, X, out_dev);
  }
}

// rows of d doubles between host (packed) and the staging buffer (row stride
// ldx); one contiguous copy when the rows are not padded
static void copy_rows_to_device(Chains& c, const double* host, int64_t rows) {
  Problem& P = *c.p;
  if (c.ldx == P.d)
    HMC_HIP(hipMemcpyAsync(c.p_host, host, (size_t)rows * P.d * 8, hipMemcpyHostToDevice, P.stream));
  else
    HMC_HIP(hipMemcpy2DAsync(c.p_host, c.ldx * 8, host, P.d * 8, P.d * 8, rows, hipMemcpyHostToDevice, P.stream));
}
static void copy_rows_to_host(Chains& c, double* host, int64_t rows) {
  Problem& P = *c.p;
  if (c.ldx == P.d)
    HMC_HIP(hipMemcpyAsync(host, c.p_host, (size_t)rows * P.d * 8, hipMemcpyDeviceToHost, P.stream));
  else
    HMC_HIP(hipMemcpy2DAsync(host, P.d * 8, c.p_host, c.ldx * 8, P.d * 8, rows, hipMemcpyDeviceToHost, P.stream));
  HMC_HIP(hipStreamSynchronize(P.stream));
}

}  // namespace hmc

using namespace hmc;

extern "C" {

int hmc_version(void) { return 100; }

const char* hmc_last_error(void) { return g_last_error.c_str(); }

int hmc_device_count(int* count) {
  return guarded("hmc_device_count", [&] {
    HMC_REQUIRE(count, HMC_ERR_ARG, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
      (void)hipGetLastError();
      n = 0;
    }
    *count = n;
  });
}

int hmc_problem_create(int64_t dimensions, int precision, hmc_problem** out) {
  return guarded("hmc_problem_create", [&] {
    HMC_REQUIRE(out, HMC_ERR_ARG, "out is NULL");
    HMC_REQUIRE(dimensions > 0, HMC_ERR_ARG, "dimensions must be positive");
    HMC_REQUIRE(precision == HMC_F64 || precision == HMC_F32 || precision == HMC_BF16, HMC_ERR_ARG,
                "precision must be HMC_F64, HMC_F32 or HMC_BF16");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      throw Error{HMC_ERR_HIP, "no HIP device is visible"};
    }
    auto* h = new hmc_problem();
    Problem& P = h->p;
    P.d = dimensions;
    P.precision = precision;
    P.state_dtype = precision == HMC_F64 ? HMC_F64 : HMC_F32;
    P.tgt_dtype = P.state_dtype;
    try {
      HMC_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
      int dev = 0, cus = 0;
      HMC_HIP(hipGetDevice(&dev));
      HMC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      P.n_cu = cus;
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int hmc_problem_set_dense(hmc_problem* hp, const void* G, int g_dtype, int64_t n_data,
                          const double* d_obs, const double* var, int scalar) {
  return guarded("hmc_problem_set_dense", [&] {
    HMC_REQUIRE(hp && G, HMC_ERR_ARG, "NULL argument");
    Problem& P = hp->p;
    HMC_REQUIRE(P.op == 0, HMC_ERR_STATE, "likelihood already set");
    HMC_REQUIRE(n_data > 0, HMC_ERR_ARG, "n_data must be positive");
    HMC_REQUIRE(g_dtype == HMC_F64 || g_dtype == HMC_F32 || g_dtype == HMC_BF16, HMC_ERR_ARG, "bad G dtype");
    const int64_t n = n_data, d = P.d;
    P.n_pad = round_up(n, gemm_row_pad());
    P.d_pad = round_up(d, gemm_row_pad());
    P.ldg = round_up(d, 64);
    P.ldgt = round_up(n, 64);
    const int traj = P.precision;
    const int tgt = target_store(traj, g_dtype);
    P.tgt_dtype = tgt;
    const size_t et = dtype_size(traj), eg = dtype_size(tgt);
    // host staging of the padded layouts
    std::vector<unsigned char> g_traj((size_t)P.n_pad * P.ldg * et, 0);
    std::vector<unsigned char> gt_traj((size_t)P.d_pad * P.ldgt * et, 0);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < d; ++j) {
        const double v = load_elem(G, g_dtype, (size_t)i * d + j);
        store_elem(g_traj.data(), traj, (size_t)i * P.ldg + j, v);
        store_elem(gt_traj.data(), traj, (size_t)j * P.ldgt + i, v);
      }
    P.G = upload(P.owned, g_traj.data(), g_traj.size(), P.stream);
    P.Gt = upload(P.owned, gt_traj.data(), gt_traj.size(), P.stream);
    if (tgt == traj) {
      P.G_tgt = P.G;
      P.Gt_tgt = P.Gt;
    } else {
      std::vector<unsigned char>().swap(g_traj);
      std::vector<unsigned char>().swap(gt_traj);
      std::vector<unsigned char> g_t((size_t)P.n_pad * P.ldg * eg, 0);
      std::vector<unsigned char> gt_t((size_t)P.d_pad * P.ldgt * eg, 0);
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < d; ++j) {
          const double v = load_elem(G, g_dtype, (size_t)i * d + j);
          store_elem(g_t.data(), tgt, (size_t)i * P.ldg + j, v);
          store_elem(gt_t.data(), tgt, (size_t)j * P.ldgt + i, v);
        }
      P.G_tgt = upload(P.owned, g_t.data(), g_t.size(), P.stream);
      P.Gt_tgt = upload(P.owned, gt_t.data(), gt_t.size(), P.stream);
    }
    set_data(P, n, d_obs, var, scalar);
    P.op = 1;
  });
}

int hmc_problem_set_csr(hmc_problem* hp, const int32_t* indptr, const int32_t* indices,
                        const void* values, int dtype, int64_t n_data, int64_t nnz,
                        const double* d_obs, const double* var, int scalar) {
  return guarded("hmc_problem_set_csr", [&] {
    HMC_REQUIRE(hp && indptr && (nnz == 0 || (indices && values)), HMC_ERR_ARG, "NULL argument");
    Problem& P = hp->p;
    HMC_REQUIRE(P.op == 0, HMC_ERR_STATE, "likelihood already set");
    HMC_REQUIRE(P.precision != HMC_BF16, HMC_ERR_UNSUPPORTED, "sparse forward models run in float64 or float32");
    HMC_REQUIRE(n_data > 0, HMC_ERR_ARG, "n_data must be positive");
    HMC_REQUIRE(dtype == HMC_F64 || dtype == HMC_F32, HMC_ERR_ARG, "CSR values must be float64 or float32");
    const int64_t n = n_data, d = P.d;
    HMC_REQUIRE(indptr[0] == 0 && indptr[n] == nnz, HMC_ERR_ARG, "indptr does not match nnz");
    std::vector<int64_t> ptr(n + 1);
    for (int64_t i = 0; i <= n; ++i) {
      ptr[i] = indptr[i];
      if (i > 0) HMC_REQUIRE(ptr[i] >= ptr[i - 1], HMC_ERR_ARG, "indptr must be non-decreasing");
    }
    std::vector<int32_t> idx(indices, indices + nnz);
    std::vector<double> val(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
      HMC_REQUIRE(idx[k] >= 0 && idx[k] < d, HMC_ERR_ARG, "column index out of range");
      val[k] = load_elem(values, dtype, k);
    }
    const int traj = P.precision;
    const int tgt = target_store(traj, dtype);
    P.tgt_dtype = tgt;
    // transpose (counting sort by column, rows kept in increasing order)
    std::vector<int64_t> tptr(d + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) tptr[idx[k] + 1]++;
    for (int64_t j = 0; j < d; ++j) tptr[j + 1] += tptr[j];
    std::vector<int32_t> tidx(nnz);
    std::vector<double> tval(nnz);
    std::vector<int64_t> fill(tptr.begin(), tptr.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int64_t o = fill[idx[k]]++;
        tidx[o] = (int32_t)i;
        tval[o] = val[k];
      }
    // positions: data row i sits at forward-slice position ppos[i], parameter
    // j at backward-slice position qpos[j]; each operator's columns are stored
    // as positions of the other
    const std::vector<int32_t> ford = sell_order(ptr, n), bord = sell_order(tptr, d);
    std::vector<int32_t> ppos(n), qpos(d);
    for (int64_t p = 0; p < n; ++p) ppos[ford[p]] = (int32_t)p;
    for (int64_t q = 0; q < d; ++q) qpos[bord[q]] = (int32_t)q;
    build_sell(P, P.sg, n, d, ptr, idx, val, ford, qpos, traj, tgt);
    build_sell(P, P.sgt, d, n, tptr, tidx, tval, bord, ppos, traj, traj);
    set_data(P, n, d_obs, var, scalar);
    {
      const int64_t np = P.sg.nslices * 64;
      std::vector<double> pd(np, 0.0), pv(np, 1.0);
      for (int64_t p = 0; p < n; ++p) {
        pd[p] = d_obs[ford[p]];
        pv[p] = scalar ? var[0] : var[ford[p]];
      }
      P.pos_dobs = (double*)upload(P.owned, pd.data(), np * 8, P.stream);
      P.pos_var = (double*)upload(P.owned, pv.data(), np * 8, P.stream);
      P.pos_model = (double*)dev_alloc(P.owned, (size_t)kPosModelArrays * P.sgt.nslices * 64 * 8);
      P.pos_model_version = -1;
    }
    P.op = 2;
  });
}

int hmc_problem_add_prior(hmc_problem* hp, int kind, const double* a, const double* b, double constant) {
  return guarded("hmc_problem_add_prior", [&] {
    HMC_REQUIRE(hp && a && b, HMC_ERR_ARG, "NULL argument");
    Problem& P = hp->p;
    HMC_REQUIRE(kind == HMC_PRIOR_NORMAL || kind == HMC_PRIOR_LAPLACE, HMC_ERR_ARG, "unknown prior kind");
    HMC_REQUIRE(P.n_priors < kMaxPriors, HMC_ERR_UNSUPPORTED, "at most 4 element-wise prior terms");
    const int k = P.n_priors;
    P.prior_kind[k] = kind;
    P.prior_a[k] = (double*)upload(P.owned, a, P.d * 8, P.stream);
    P.prior_b[k] = (double*)upload(P.owned, b, P.d * 8, P.stream);
    P.misfit_const += constant;
    P.n_priors++;
    P.model_version++;
  });
}

int hmc_problem_set_bounds(hmc_problem* hp, const double* lower, const double* upper) {
  return guarded("hmc_problem_set_bounds", [&] {
    HMC_REQUIRE(hp, HMC_ERR_ARG, "NULL problem");
    Problem& P = hp->p;
    if (lower && upper)
      for (int64_t j = 0; j < P.d; ++j)
        HMC_REQUIRE(upper[j] > lower[j], HMC_ERR_ARG, "Bounds vectors are incompatible.");
    P.lo = lower ? (double*)upload(P.owned, lower, P.d * 8, P.stream) : nullptr;
    P.hi = upper ? (double*)upload(P.owned, upper, P.d * 8, P.stream) : nullptr;
    P.model_version++;
  });
}

int hmc_problem_set_mass(hmc_problem* hp, int kind, const double* diagonal) {
  return guarded("hmc_problem_set_mass", [&] {
    HMC_REQUIRE(hp, HMC_ERR_ARG, "NULL problem");
    Problem& P = hp->p;
    if (kind == HMC_MASS_UNIT) {
      P.minv = P.msqrt = nullptr;
    } else if (kind == HMC_MASS_DIAGONAL) {
      HMC_REQUIRE(diagonal, HMC_ERR_ARG, "diagonal mass needs the diagonal");
      std::vector<double> inv(P.d), sq(P.d);
      for (int64_t j = 0; j < P.d; ++j) {
        HMC_REQUIRE(diagonal[j] > 0.0, HMC_ERR_ARG, "mass diagonal must be positive");
        inv[j] = 1.0 / diagonal[j];       // MassMatrices.py:180
        sq[j] = std::sqrt(diagonal[j]);   // MassMatrices.py:227
      }
      P.minv = (double*)upload(P.owned, inv.data(), P.d * 8, P.stream);
      P.msqrt = (double*)upload(P.owned, sq.data(), P.d * 8, P.stream);
    } else {
      throw Error{HMC_ERR_ARG, "unknown mass matrix kind"};
    }
    P.mass_kind = kind;
    P.model_version++;
  });
}

int hmc_problem_set_engine(hmc_problem* hp, int engine) {
  return guarded("hmc_problem_set_engine", [&] {
    HMC_REQUIRE(hp, HMC_ERR_ARG, "NULL problem");
    HMC_REQUIRE(engine >= HMC_ENGINE_AUTO && engine <= HMC_ENGINE_FUSED, HMC_ERR_ARG, "unknown engine");
    Problem& P = hp->p;
    if (engine == HMC_ENGINE_FUSED)
      HMC_REQUIRE(fused_supported(P), HMC_ERR_UNSUPPORTED, "fused engine does not support this precision/size");
    P.engine = engine;
  });
}

int hmc_problem_set_option(hmc_problem* hp, const char* name, double value) {
  return guarded("hmc_problem_set_option", [&] {
    HMC_REQUIRE(hp && name, HMC_ERR_ARG, "NULL argument");
    Problem& P = hp->p;
    const std::string key(name);
    if (key == "bf16_split") {
      HMC_REQUIRE(value == 1.0 || value == 2.0, HMC_ERR_ARG, "bf16_split must be 1 or 2");
      P.bf16_split = (int)value;
      P.bf16_split_grad = (int)value;
    } else if (key == "bf16_split_grad") {
      HMC_REQUIRE(value == 1.0 || value == 2.0, HMC_ERR_ARG, "bf16_split_grad must be 1 or 2");
      P.bf16_split_grad = (int)value;
    } else if (key == "tile_order") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "tile_order must be 0 or 1");
      P.tile_order = (int)value;
    } else if (key == "incremental_residual") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "incremental_residual must be 0 or 1");
      P.resid_update = (int)value;
    } else if (key == "gemm_config") {
      HMC_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0 || value == 3.0 || value == 4.0, HMC_ERR_ARG,
                  "gemm_config must be 0, 1, 2, 3 or 4");
      P.gemm_cfg = (int)value;
      P.gemm_cfg_grad = (int)value;
    } else if (key == "gemm_config_grad") {
      HMC_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0 || value == 3.0 || value == 4.0, HMC_ERR_ARG,
                  "gemm_config_grad must be 0, 1, 2, 3 or 4");
      P.gemm_cfg_grad = (int)value;
    } else if (key == "split_fixup") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "split_fixup must be 0 or 1");
      P.split_fixup = (int)value;
    } else if (key == "gemm_split_tail") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "gemm_split_tail must be 0 or 1");
      P.split_tail = (int)value;
    } else if (key == "endpoint_premult") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "endpoint_premult must be 0 or 1");
      P.premult = (int)value;
    } else if (key == "endpoint_symmetric") {
      HMC_REQUIRE(value == 0.0 || value == 1.0, HMC_ERR_ARG, "endpoint_symmetric must be 0 or 1");
      P.premult_sym = (int)value;
    } else if (key == "premult_trajectory") {
      HMC_REQUIRE(value == -1.0 || value == 0.0 || value == 1.0, HMC_ERR_ARG, "premult_trajectory must be -1, 0 or 1");
      P.premult_traj = (int)value;
    } else if (key == "fused_tile_chains") {
      HMC_REQUIRE(value == 0.0 || value == 1.0 || value == 2.0 || value == 4.0, HMC_ERR_ARG,
                  "fused_tile_chains must be 0, 1, 2 or 4");
      P.fused_ct = (int)value;
    } else if (key == "gemm_split_cus") {
      HMC_REQUIRE(value >= 0.0 && value <= 65536.0 && value == (double)(int)value, HMC_ERR_ARG,
                  "gemm_split_cus must be a non-negative integer");
      if (value > 0.0) {
        P.n_cu = (int)value;
      } else {
        int dev = 0, cus = 0;
        HMC_HIP(hipGetDevice(&dev));
        HMC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        P.n_cu = cus;
      }
    } else {
      throw Error{HMC_ERR_ARG, "unknown option " + key};
    }
  });
}

int hmc_problem_destroy(hmc_problem* hp) {
  return guarded("hmc_problem_destroy", [&] {
    if (!hp) return;
    Problem& P = hp->p;
    if (P.stream) (void)hipStreamSynchronize(P.stream);
    for (void* ptr : P.owned) (void)hipFree(ptr);
    if (P.stream) (void)hipStreamDestroy(P.stream);
    delete hp;
  });
}

int hmc_chains_create(hmc_problem* hp, int64_t B, hmc_chains** out) {
  return guarded("hmc_chains_create", [&] {
    HMC_REQUIRE(hp && out, HMC_ERR_ARG, "NULL argument");
    HMC_REQUIRE(B > 0 && B <= (int64_t)1 << 24, HMC_ERR_ARG, "number of chains out of range");
    Problem& P = hp->p;
    auto* h = new hmc_chains();
    Chains& c = h->c;
    try {
      c.p = &P;
      c.B = B;
      c.B_pad = round_up(B, 128);
      c.ldx = round_up(P.d, 64);
      c.ldr = P.n > 0 ? round_up(P.n, 64) : 64;
      const size_t es = dtype_size(P.state_dtype);
      const size_t rows = (size_t)c.B_pad * c.ldx;
      c.x0 = dev_alloc(c.owned, rows * es);
      c.xt = dev_alloc(c.owned, rows * es);
      c.pt = dev_alloc(c.owned, rows * es);
      c.p_host = (double*)dev_alloc(c.owned, rows * 8);
      c.chi0 = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.k0 = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.eps = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.u = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.h0 = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.h1 = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.step = (double*)dev_alloc(c.owned, c.B_pad * 8);
      c.acc = (int32_t*)dev_alloc(c.owned, c.B_pad * 4);
      c.oob = (int32_t*)dev_alloc(c.owned, c.B_pad * 4);
      c.acc_count = (int64_t*)dev_alloc(c.owned, c.B_pad * 8);
      c.ex_count = (int64_t*)dev_alloc(c.owned, c.B_pad * 8);
      c.ex_tmp = (int32_t*)dev_alloc(c.owned, (size_t)c.B_pad * 2 * 4);
      c.ex_u = (double*)dev_alloc(c.owned, c.B_pad * 8);
      if (P.op == 1) {
        // endpoint partial rows: residual form (n_pad / 128 row blocks), premultiplied
        // over all of H (d_pad / 128) or over its upper block triangle (2 ceil(T/2) items)
        const int64_t T = (P.d + 127) / 128;
        c.n_mblocks = std::max({P.n_pad / 128, P.d_pad / 128, 2 * ((T + 1) / 2)});
        c.partial = (double*)dev_alloc(c.owned, (size_t)c.n_mblocks * c.B_pad * 8);
        c.split_ws64_elems = gemm_split_workspace_f64(P, c.B);
        if (c.split_ws64_elems) c.split_ws64 = (double*)dev_alloc(c.owned, c.split_ws64_elems * 8);
        HMC_REQUIRE(P.premult_traj <= 0 || traj_premult(P) || fused_premult(P), HMC_ERR_ARG,
                    "premult_trajectory needs the GEMM engine with the premultiplied endpoint "
                    "(n > d or endpoint_premult = 1)");
        if (traj_premult(P) && P.precision == HMC_BF16) {
          c.dxp[0] = (uint16_t*)dev_alloc(c.owned, rows * 2);
          c.dxp[1] = (uint16_t*)dev_alloc(c.owned, rows * 2);
        } else if (traj_premult(P)) {
          c.xs = dev_alloc(c.owned, rows * es);
        }
        if (gemm_anchored(P)) {
          c.xa = (float*)dev_alloc(c.owned, (size_t)c.ldx * 4);
          c.yref = (float*)dev_alloc(c.owned, (size_t)std::max(c.ldx, c.ldr) * 4);
        }
        if (P.precision == HMC_BF16) {
          c.xh = (uint16_t*)dev_alloc(c.owned, rows * 2);
          c.xl = (uint16_t*)dev_alloc(c.owned, rows * 2);
          c.rh = (uint16_t*)dev_alloc(c.owned, (size_t)c.B_pad * c.ldr * 2);
          c.split_ws_elems = gemm_split_workspace(P, c.B);
          if (c.split_ws_elems) c.split_ws = (float*)dev_alloc(c.owned, c.split_ws_elems * 4);
          if (c.split_ws_elems) c.fix = (int*)dev_alloc(c.owned, (size_t)(1 + kFixTiles) * 4);   // zeroed
          c.rl = (uint16_t*)dev_alloc(c.owned, (size_t)c.B_pad * c.ldr * 2);
        } else if (P.resolved_engine() == HMC_ENGINE_GEMM) {
          c.r = dev_alloc(c.owned, (size_t)c.B_pad * c.ldr * es);
        }
      }
    } catch (...) {
      for (void* ptr : c.owned) (void)hipFree(ptr);
      delete h;
      throw;
    }
    *out = h;
  });
}

int hmc_chains_destroy(hmc_chains* hc) {
  return guarded("hmc_chains_destroy", [&] {
    if (!hc) return;
    Chains& c = hc->c;
    (void)hipStreamSynchronize(c.p->stream);
    for (hipEvent_t e : c.ev_pool) (void)hipEventDestroy(e);
    for (void* ptr : c.owned) (void)hipFree(ptr);
    delete hc;
  });
}

int hmc_chains_set_state(hmc_chains* hc, const double* X) {
  return guarded("hmc_chains_set_state", [&] {
    HMC_REQUIRE(hc && X, HMC_ERR_ARG, "NULL argument");
    Chains& c = hc->c;
    Problem& P = *c.p;
    copy_rows_to_device(c, X, c.B);
    elem_load_rows(c, c.p_host, c.x0, nullptr, nullptr);
    evaluate_misfit(c, c.x0, c.chi0);
    std::vector<double> chi(c.B);
    HMC_HIP(hipMemcpyAsync(chi.data(), c.chi0, c.B * 8, hipMemcpyDeviceToHost, P.stream));
    HMC_HIP(hipStreamSynchronize(P.stream));
    for (int64_t i = 0; i < c.B; ++i) {
      HMC_REQUIRE(!std::isnan(chi[i]), HMC_ERR_ARG, "Initial position in model space gives NaN probability");
      HMC_REQUIRE(!std::isinf(chi[i]), HMC_ERR_ARG, "Initial position in model space gives inf/-inf probability");
    }
  });
}

int hmc_chains_get_state(hmc_chains* hc, double* X, double* chi) {
  return guarded("hmc_chains_get_state", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    Problem& P = *c.p;
    if (X) {
      elem_store_rows(c, c.x0, P.state_dtype, c.ldx, c.p_host);
      copy_rows_to_host(c, X, c.B);
    }
    if (chi) {
      HMC_HIP(hipMemcpyAsync(chi, c.chi0, c.B * 8, hipMemcpyDeviceToHost, P.stream));
      HMC_HIP(hipStreamSynchronize(P.stream));
      if (c.beta) {
        std::vector<double> bt(c.B);
        HMC_HIP(hipMemcpy(bt.data(), c.beta, c.B * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < c.B; ++i) chi[i] *= bt[i];
      }
    }
  });
}

int hmc_chains_set_temperatures(hmc_chains* hc, const double* temperatures) {
  return guarded("hmc_chains_set_temperatures", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    HMC_HIP(hipStreamSynchronize(c.p->stream));
    if (!temperatures) {
      if (c.beta) {
        (void)hipFree(c.beta);
        c.owned.erase(std::remove(c.owned.begin(), c.owned.end(), (void*)c.beta), c.owned.end());
        c.beta = nullptr;
      }
      return;
    }
    std::vector<double> bt(c.B_pad, 1.0);
    for (int64_t i = 0; i < c.B; ++i) {
      HMC_REQUIRE(temperatures[i] > 0.0 && std::isfinite(temperatures[i]), HMC_ERR_ARG,
                  "temperatures must be positive and finite");
      bt[i] = 1.0 / temperatures[i];
    }
    if (!c.beta) c.beta = (double*)dev_alloc(c.owned, c.B_pad * 8);
    h2d(c.beta, bt.data(), c.B_pad * 8);
  });
}

int hmc_chains_exchange(hmc_chains* hc, const int32_t* pairs, int64_t npairs, const double* u,
                        int64_t proposal, int32_t* swapped) {
  return guarded("hmc_chains_exchange", [&] {
    HMC_REQUIRE(hc && pairs && u && npairs >= 0, HMC_ERR_ARG, "invalid argument");
    Chains& c = hc->c;
    HMC_REQUIRE(2 * npairs <= c.B, HMC_ERR_ARG, "more pairs than chains");
    std::vector<char> used(c.B, 0);
    for (int64_t k = 0; k < 2 * npairs; ++k) {
      HMC_REQUIRE(pairs[k] >= 0 && pairs[k] < c.B, HMC_ERR_ARG, "pair index out of range");
      HMC_REQUIRE(!used[pairs[k]], HMC_ERR_ARG, "a chain appears twice in one exchange event");
      used[pairs[k]] = 1;
    }
    if (npairs == 0) return;
    Problem& P = *c.p;
    HMC_HIP(hipMemcpyAsync(c.ex_tmp, pairs, npairs * 2 * 4, hipMemcpyHostToDevice, P.stream));
    HMC_HIP(hipMemcpyAsync(c.ex_u, u, npairs * 8, hipMemcpyHostToDevice, P.stream));
    int32_t* flags = c.ex_tmp + c.B_pad;
    elem_exchange(c, c.ex_tmp, npairs, c.ex_u, 0, proposal, flags, -1);
    if (swapped) HMC_HIP(hipMemcpyAsync(swapped, flags, npairs * 4, hipMemcpyDeviceToHost, P.stream));
    HMC_HIP(hipStreamSynchronize(P.stream));
  });
}

int hmc_chains_set_exchange(hmc_chains* hc, int64_t interval, const int32_t* schedule, int64_t events,
                            int64_t npairs) {
  return guarded("hmc_chains_set_exchange", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    HMC_HIP(hipStreamSynchronize(c.p->stream));
    if (c.ex_pairs) {
      (void)hipFree(c.ex_pairs);
      c.owned.erase(std::remove(c.owned.begin(), c.owned.end(), (void*)c.ex_pairs), c.owned.end());
      c.ex_pairs = nullptr;
    }
    c.ex_events = c.ex_npairs = c.ex_interval = 0;
    if (!schedule || events <= 0 || npairs <= 0) return;
    HMC_REQUIRE(interval > 0 && 2 * npairs <= c.B, HMC_ERR_ARG, "invalid exchange schedule");
    for (int64_t e = 0; e < events; ++e) {
      std::vector<char> used(c.B, 0);
      for (int64_t k = 0; k < 2 * npairs; ++k) {
        const int32_t v = schedule[e * 2 * npairs + k];
        HMC_REQUIRE(v >= 0 && v < c.B && !used[v], HMC_ERR_ARG,
                    "exchange schedule rows must hold distinct chain indices");
        used[v] = 1;
      }
    }
    c.ex_pairs = (int32_t*)dev_alloc(c.owned, (size_t)events * 2 * npairs * 4);
    h2d(c.ex_pairs, schedule, (size_t)events * 2 * npairs * 4);
    c.ex_events = events;
    c.ex_npairs = npairs;
    c.ex_interval = interval;
  });
}

int hmc_chains_exchange_counts(hmc_chains* hc, int64_t* counts) {
  return guarded("hmc_chains_exchange_counts", [&] {
    HMC_REQUIRE(hc && counts, HMC_ERR_ARG, "NULL argument");
    Chains& c = hc->c;
    HMC_HIP(hipMemcpyAsync(counts, c.ex_count, c.B * 8, hipMemcpyDeviceToHost, c.p->stream));
    HMC_HIP(hipStreamSynchronize(c.p->stream));
  });
}

int hmc_misfit_batch(hmc_problem* hp, const double* M, int64_t B, double* out) {
  return guarded("hmc_misfit_batch", [&] {
    HMC_REQUIRE(hp && M && out && B > 0, HMC_ERR_ARG, "invalid argument");
    hmc_chains* hc = nullptr;
    int rc = hmc_chains_create(hp, B, &hc);
    if (rc) throw Error{rc, g_last_error};
    Chains& c = hc->c;
    try {
      copy_rows_to_device(c, M, B);
      elem_load_rows(c, c.p_host, c.xt, nullptr, nullptr);
      evaluate_misfit(c, c.xt, c.h1);
      HMC_HIP(hipMemcpyAsync(out, c.h1, B * 8, hipMemcpyDeviceToHost, hp->p.stream));
      HMC_HIP(hipStreamSynchronize(hp->p.stream));
    } catch (...) {
      hmc_chains_destroy(hc);
      throw;
    }
    hmc_chains_destroy(hc);
  });
}

int hmc_gradient_batch(hmc_problem* hp, const double* M, int64_t B, double* out) {
  return guarded("hmc_gradient_batch", [&] {
    HMC_REQUIRE(hp && M && out && B > 0, HMC_ERR_ARG, "invalid argument");
    hmc_chains* hc = nullptr;
    int rc = hmc_chains_create(hp, B, &hc);
    if (rc) throw Error{rc, g_last_error};
    Chains& c = hc->c;
    try {
      Problem& P = hp->p;
      copy_rows_to_device(c, M, B);
      elem_load_rows(c, c.p_host, c.xt, c.xh, c.xl);
      if (P.resolved_engine() == HMC_ENGINE_GEMM) gemm_gradient(c, c.xt, c.p_host);
      else fused_gradient(c, c.xt, c.p_host);
      copy_rows_to_host(c, out, B);
    } catch (...) {
      hmc_chains_destroy(hc);
      throw;
    }
    hmc_chains_destroy(hc);
  });
}

int hmc_chains_set_anchor(hmc_chains* hc, const double* x_ref) {
  return guarded("hmc_chains_set_anchor", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    if (!gemm_anchored(*c.p)) return;   // no anchored trajectories for this problem
    gemm_set_anchor(c, x_ref);
    HMC_HIP(hipStreamSynchronize(c.p->stream));
  });
}

int hmc_chains_trajectory_gradient(hmc_chains* hc, const double* X, double* out) {
  return guarded("hmc_chains_trajectory_gradient", [&] {
    HMC_REQUIRE(hc && X && out, HMC_ERR_ARG, "NULL argument");
    Chains& c = hc->c;
    Problem& P = *c.p;
    copy_rows_to_device(c, X, c.B);
    elem_load_rows(c, c.p_host, c.xt, c.xh, c.xl);
    if (P.resolved_engine() == HMC_ENGINE_GEMM) {
      HMC_REQUIRE(c.surrogate == nullptr, HMC_ERR_STATE, "a surrogate drives these chains' trajectories");
      gemm_traj_gradient(c, c.xt, c.p_host);
    } else {
      fused_gradient(c, c.xt, c.p_host);
    }
    copy_rows_to_host(c, out, c.B);
  });
}

int hmc_chains_propose(hmc_chains* hc, const double* P0, const double* eps, const double* u, int L,
                       int integrator, int32_t* accepted, double* H0, double* H1, double* X1,
                       double* P1) {
  return guarded("hmc_chains_propose", [&] {
    HMC_REQUIRE(hc && P0 && eps && u, HMC_ERR_ARG, "NULL argument");
    HMC_REQUIRE(L > 0, HMC_ERR_ARG, "amount_of_steps must be positive");
    Chains& c = hc->c;
    Problem& P = *c.p;
    copy_rows_to_device(c, P0, c.B);
    HMC_HIP(hipMemcpyAsync(c.eps, eps, c.B * 8, hipMemcpyHostToDevice, P.stream));
    HMC_HIP(hipMemcpyAsync(c.u, u, c.B * 8, hipMemcpyHostToDevice, P.stream));
    if (P.resolved_engine() == HMC_ENGINE_GEMM)
      gemm_proposal(c, true, L, integrator, 0, 0, 0, 0, 0.0, 0.0, -1);
    else
      fused_propose(c, true, L, integrator, 0, 0, 0, 0, 0.0, 0.0, -1);
    if (accepted) HMC_HIP(hipMemcpyAsync(accepted, c.acc, c.B * 4, hipMemcpyDeviceToHost, P.stream));
    if (H0) HMC_HIP(hipMemcpyAsync(H0, c.h0, c.B * 8, hipMemcpyDeviceToHost, P.stream));
    if (H1) HMC_HIP(hipMemcpyAsync(H1, c.h1, c.B * 8, hipMemcpyDeviceToHost, P.stream));
    if (X1) {
      elem_store_rows(c, c.xt, P.state_dtype, c.ldx, c.p_host);
      copy_rows_to_host(c, X1, c.B);
    }
    if (P1) {
      elem_store_rows(c, c.pt, P.state_dtype, c.ldx, c.p_host);
      copy_rows_to_host(c, P1, c.B);
    }
    HMC_HIP(hipStreamSynchronize(P.stream));
    time_collect(c);
    check_fixup(c);
  });
}

int hmc_chains_run(hmc_chains* hc, int64_t proposals, int L, int integrator, double* stepsize,
                   int randomize, int autotune, double target, double learning_rate, uint64_t seed,
                   int64_t first_proposal, int64_t thinning, int64_t* accepted_total) {
  return guarded("hmc_chains_run", [&] {
    HMC_REQUIRE(hc && stepsize, HMC_ERR_ARG, "NULL argument");
    HMC_REQUIRE(proposals >= 0 && L > 0 && thinning > 0, HMC_ERR_ARG, "invalid run arguments");
    Chains& c = hc->c;
    Problem& P = *c.p;
    HMC_HIP(hipMemcpyAsync(c.step, stepsize, c.B * 8, hipMemcpyHostToDevice, P.stream));
    HMC_HIP(hipMemsetAsync(c.acc_count, 0, c.B * 8, P.stream));
    const bool gemm = P.resolved_engine() == HMC_ENGINE_GEMM;
    for (int64_t i = 0; i < proposals; ++i) {
      const int64_t prop = first_proposal + i;
      int64_t slot = -1;
      if (c.tr_cap > 0 && prop % thinning == 0) {
        HMC_REQUIRE(c.tr_n < c.tr_cap, HMC_ERR_STATE, "trace capacity exceeded");
        slot = c.tr_n++;
      }
      int64_t hslot = -1;
      if (autotune && c.hist_cap > 0) {
        HMC_REQUIRE(c.hist_n < c.hist_cap, HMC_ERR_STATE, "tuning history capacity exceeded");
        hslot = c.hist_n++;
      }
      c.timing_skip = c.timing && (prop % c.timing_every) != 0;
      if (gemm)
        gemm_proposal(c, false, L, integrator, seed, prop, randomize, autotune, target, learning_rate, slot,
                      hslot);
      else
        fused_propose(c, false, L, integrator, seed, prop, randomize, autotune, target, learning_rate, slot,
                      hslot);
      // replica exchange after the proposal, as in the reference's loop (Samplers.py:589-595)
      if (c.ex_pairs && prop % c.ex_interval == 0 && prop / c.ex_interval < c.ex_events)
        elem_exchange(c, c.ex_pairs + (prop / c.ex_interval) * 2 * c.ex_npairs, c.ex_npairs, nullptr, seed, prop,
                      nullptr, slot);
      if (c.timing && c.ev_pending.size() > 2048) time_collect(c);
    }
    c.timing_skip = false;
    HMC_HIP(hipMemcpyAsync(stepsize, c.step, c.B * 8, hipMemcpyDeviceToHost, P.stream));
    if (accepted_total)
      HMC_HIP(hipMemcpyAsync(accepted_total, c.acc_count, c.B * 8, hipMemcpyDeviceToHost, P.stream));
    HMC_HIP(hipStreamSynchronize(P.stream));
    time_collect(c);
    check_fixup(c);
  });
}

int hmc_chains_set_tuning_history(hmc_chains* hc, int64_t capacity) {
  return guarded("hmc_chains_set_tuning_history", [&] {
    HMC_REQUIRE(hc && capacity >= 0, HMC_ERR_ARG, "invalid argument");
    Chains& c = hc->c;
    HMC_HIP(hipStreamSynchronize(c.p->stream));
    for (double* ptr : {c.hist_rate, c.hist_step}) {
      if (!ptr) continue;
      (void)hipFree(ptr);
      c.owned.erase(std::remove(c.owned.begin(), c.owned.end(), (void*)ptr), c.owned.end());
    }
    c.hist_rate = c.hist_step = nullptr;
    c.hist_cap = c.hist_n = 0;
    if (capacity > 0) {
      c.hist_rate = (double*)dev_alloc(c.owned, (size_t)capacity * c.B * 8);
      c.hist_step = (double*)dev_alloc(c.owned, (size_t)capacity * c.B * 8);
      c.hist_cap = capacity;
    }
  });
}

int hmc_chains_tuning_history_count(hmc_chains* hc, int64_t* count) {
  return guarded("hmc_chains_tuning_history_count", [&] {
    HMC_REQUIRE(hc && count, HMC_ERR_ARG, "NULL argument");
    *count = hc->c.hist_n;
  });
}

int hmc_chains_read_tuning_history(hmc_chains* hc, int64_t first, int64_t count, double* rates, double* steps) {
  return guarded("hmc_chains_read_tuning_history", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    HMC_REQUIRE(first >= 0 && count >= 0 && first + count <= c.hist_n, HMC_ERR_ARG,
                "tuning history range out of bounds");
    if (count == 0) return;
    const size_t off = (size_t)first * c.B, len = (size_t)count * c.B * 8;
    if (rates) HMC_HIP(hipMemcpyAsync(rates, c.hist_rate + off, len, hipMemcpyDeviceToHost, c.p->stream));
    if (steps) HMC_HIP(hipMemcpyAsync(steps, c.hist_step + off, len, hipMemcpyDeviceToHost, c.p->stream));
    HMC_HIP(hipStreamSynchronize(c.p->stream));
  });
}

int hmc_integrator_schedule(int integrator, int L, double* first_drift, double* stages, int64_t capacity,
                            int64_t* count) {
  return guarded("hmc_integrator_schedule", [&] {
    HMC_REQUIRE(first_drift && count && L > 0, HMC_ERR_ARG, "invalid argument");
    std::vector<Stage> st;
    integrator_schedule(integrator, L, *first_drift, st);
    *count = (int64_t)st.size();
    HMC_REQUIRE(stages == nullptr || (int64_t)st.size() <= capacity, HMC_ERR_ARG, "stage buffer too small");
    if (stages)
      for (size_t k = 0; k < st.size(); ++k) {
        stages[3 * k] = st[k].kick;
        stages[3 * k + 1] = st[k].drift;
        stages[3 * k + 2] = st[k].drift2;
      }
  });
}

int hmc_philox4x32_10(const uint32_t* counter, const uint32_t* key, uint32_t* out) {
  return guarded("hmc_philox4x32_10", [&] {
    HMC_REQUIRE(counter && key && out, HMC_ERR_ARG, "NULL argument");
    U32x4 c{{counter[0], counter[1], counter[2], counter[3]}};
    const U32x4 r = philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
  });
}

int hmc_chains_set_trace(hmc_chains* hc, int64_t capacity, int64_t param_stride) {
  return guarded("hmc_chains_set_trace", [&] {
    HMC_REQUIRE(hc && capacity >= 0, HMC_ERR_ARG, "invalid argument");
    (void)param_stride;
    Chains& c = hc->c;
    Problem& P = *c.p;
    HMC_HIP(hipStreamSynchronize(P.stream));
    auto drop = [&](void* ptr) {
      if (!ptr) return;
      (void)hipFree(ptr);
      c.owned.erase(std::remove(c.owned.begin(), c.owned.end(), ptr), c.owned.end());
    };
    drop(c.tr_x);
    drop(c.tr_chi);
    c.tr_x = nullptr;
    c.tr_chi = nullptr;
    c.tr_cap = c.tr_n = 0;
    if (capacity > 0) {
      c.tr_x = dev_alloc(c.owned, (size_t)capacity * c.B * P.d * dtype_size(P.state_dtype));
      c.tr_chi = (double*)dev_alloc(c.owned, (size_t)capacity * c.B * 8);
      c.tr_cap = capacity;
    }
  });
}

int hmc_chains_trace_count(hmc_chains* hc, int64_t* slots) {
  return guarded("hmc_chains_trace_count", [&] {
    HMC_REQUIRE(hc && slots, HMC_ERR_ARG, "NULL argument");
    *slots = hc->c.tr_n;
  });
}

int hmc_chains_reset_trace(hmc_chains* hc) {
  return guarded("hmc_chains_reset_trace", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL argument");
    hc->c.tr_n = 0;
  });
}

int hmc_chains_read_trace(hmc_chains* hc, int64_t first, int64_t count, double* X, double* chi) {
  return guarded("hmc_chains_read_trace", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    Chains& c = hc->c;
    Problem& P = *c.p;
    HMC_REQUIRE(first >= 0 && count >= 0 && first + count <= c.tr_n, HMC_ERR_ARG, "trace range out of bounds");
    if (count == 0) return;
    const size_t per = (size_t)c.B * P.d;
    if (X) {
      if (P.state_dtype == HMC_F64) {
        HMC_HIP(hipMemcpyAsync(X, (double*)c.tr_x + first * per, count * per * 8, hipMemcpyDeviceToHost, P.stream));
      } else {
        std::vector<float> tmp(count * per);
        HMC_HIP(hipMemcpyAsync(tmp.data(), (float*)c.tr_x + first * per, count * per * 4, hipMemcpyDeviceToHost, P.stream));
        HMC_HIP(hipStreamSynchronize(P.stream));
        for (size_t i = 0; i < tmp.size(); ++i) X[i] = tmp[i];
      }
    }
    if (chi)
      HMC_HIP(hipMemcpyAsync(chi, c.tr_chi + first * c.B, count * c.B * 8, hipMemcpyDeviceToHost, P.stream));
    HMC_HIP(hipStreamSynchronize(P.stream));
  });
}

int hmc_chains_trace_device(hmc_chains* hc, void** x_ptr, double** chi_ptr, int* dtype) {
  return guarded("hmc_chains_trace_device", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    if (x_ptr) *x_ptr = hc->c.tr_x;
    if (chi_ptr) *chi_ptr = hc->c.tr_chi;
    if (dtype) *dtype = hc->c.p->state_dtype;
  });
}

int hmc_chains_enable_timing(hmc_chains* hc, int on) {
  return guarded("hmc_chains_enable_timing", [&] {
    HMC_REQUIRE(hc && on >= 0, HMC_ERR_ARG, "invalid argument");
    Chains& c = hc->c;
    time_collect(c);
    c.timing = on != 0;
    c.timing_every = on > 1 ? on : 1;
    for (int k = 0; k < kTimingKinds; ++k) {
      c.t_ms[k] = 0.0;
      c.t_n[k] = 0;
    }
  });
}

int hmc_chains_timing(hmc_chains* hc, int kind, double* total_ms, int64_t* launches) {
  return guarded("hmc_chains_timing", [&] {
    HMC_REQUIRE(hc && kind >= 0 && kind < kTimingKinds, HMC_ERR_ARG, "invalid argument");
    Chains& c = hc->c;
    time_collect(c);
    if (total_ms) *total_ms = c.t_ms[kind];
    if (launches) *launches = c.t_n[kind];
  });
}

int hmc_chains_sync(hmc_chains* hc) {
  return guarded("hmc_chains_sync", [&] {
    HMC_REQUIRE(hc, HMC_ERR_ARG, "NULL chains");
    HMC_HIP(hipStreamSynchronize(hc->c.p->stream));
  });
}

}  // extern "C"

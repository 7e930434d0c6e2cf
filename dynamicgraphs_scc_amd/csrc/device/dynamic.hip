// Exact dynamic SCC maintenance under batches of edge inserts and deletes.
//
// Replaces the reference's incremental (src/scc_incremental.cc:7-27 with the
// condensation built by src/common_main.h:180-314) and decremental
// (src/scc_decremental.cc:7-24 with src/common_main.h:343-393) paths with one
// exact call.  Let G' = G - D + I and C(v) the canonical SCC labels of G.
//   X_d = union of old SCCs that lost an internal edge (only these can split;
//         the reference's read_file2 marks the same SCCs, :361-370);
//   S/T = heads/tails of inserted edges u->v with C(u) != C(v) or C(u) in X_d;
//   X   = X_d  u  (FW_G'(S) n BW_G'(T)).
// Every SCC of G' outside X is an SCC of G (a cycle through it must use an
// inserted S/T edge), and X is a union of SCCs of G', so recomputing the SCCs
// of G'[X] with the static pipeline and keeping C elsewhere gives the SCCs of G'.
// BW(T) is searched only inside FW(S) (a path from w in FW(S) to T stays in
// FW(S)), the same restriction the reference's backward search uses.
//
// The device CSR and reverse CSR are rebuilt per batch with rows kept sorted:
// every surviving old entry moves to newbegin[u] + (i - begin[u]) -
// #deleted-before + #inserted-before, every insert to its merged rank; all
// positions come from binary searches, so no row is merged sequentially.
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cstring>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

struct Dyn {
  int64_t cap_b = 0;  // capacity of each batch list
  unsigned long long *k_in = nullptr, *k_tmp = nullptr;
  unsigned long long *del_f = nullptr, *del_r = nullptr, *ins_f = nullptr, *ins_r = nullptr;
  unsigned long long *ins_sorted = nullptr;
  uint8_t* flags = nullptr;
  int32_t *src = nullptr, *dst = nullptr;  // staging of one list
  int32_t *rdel_f = nullptr, *rins_f = nullptr, *rdel_r = nullptr, *rins_r = nullptr;  // per row
  int32_t *nb_f = nullptr, *nb_r = nullptr;  // new offsets (n+1)
  int32_t *deg = nullptr;                    // n+1 scan input
  int32_t* aff = nullptr;                    // per old-SCC flag
  int32_t *adj2 = nullptr, *radj2 = nullptr;
  int64_t cap2 = 0;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int32_t n = -1;
  int32_t* h_stage = nullptr;  // pinned staging for batch upload (4 * cap_b ints)
  int64_t n_del = 0, n_ins = 0;  // final counts (forward lists)
};

constexpr unsigned long long BAD_KEY = ~0ull;

__device__ __forceinline__ int32_t key_hi(unsigned long long k) { return (int32_t)(k >> 32); }
__device__ __forceinline__ int32_t key_lo(unsigned long long k) { return (int32_t)(k & 0xFFFFFFFFull); }

// first index in a[lo, hi) with a[i] >= x (int32 rows)
__device__ __forceinline__ int32_t lb_i32(const int32_t* __restrict__ a, int32_t lo, int32_t hi, int32_t x) {
  while (lo < hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int64_t lb_u64(const unsigned long long* __restrict__ a, int64_t lo, int64_t hi,
                                          unsigned long long x) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(BLOCK) void k_make_keys(const int32_t* __restrict__ s, const int32_t* __restrict__ d,
                                                     int64_t k, bool drop_loops,
                                                     unsigned long long* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = s[i], v = d[i];
    keys[i] = (drop_loops && u == v) ? BAD_KEY
                                     : (((unsigned long long)(uint32_t)u << 32) | (uint32_t)v);
  }
}

// Flag sorted delete keys that are present in G and not re-inserted.
__global__ __launch_bounds__(BLOCK) void k_flag_del(const unsigned long long* __restrict__ keys, int64_t k,
                                                    const unsigned long long* __restrict__ ins, int64_t ni,
                                                    const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ adj,
                                                    uint8_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = keys[i];
    bool keep = key != BAD_KEY && (i == 0 || keys[i - 1] != key);
    if (keep) {
      const int32_t u = key_hi(key), v = key_lo(key);
      const int32_t s = begin[u], e = begin[u + 1];
      const int32_t p = lb_i32(adj, s, e, v);
      keep = p < e && adj[p] == v;
      if (keep && ni > 0) {
        const int64_t q = lb_u64(ins, 0, ni, key);
        keep = !(q < ni && ins[q] == key);
      }
    }
    flags[i] = keep ? 1 : 0;
  }
}

// Flag sorted insert keys that are absent from G (self-loops already BAD_KEY).
__global__ __launch_bounds__(BLOCK) void k_flag_ins(const unsigned long long* __restrict__ keys, int64_t k,
                                                    const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ adj,
                                                    uint8_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = keys[i];
    bool keep = key != BAD_KEY && (i == 0 || keys[i - 1] != key);
    if (keep) {
      const int32_t u = key_hi(key), v = key_lo(key);
      const int32_t s = begin[u], e = begin[u + 1];
      const int32_t p = lb_i32(adj, s, e, v);
      keep = !(p < e && adj[p] == v);
    }
    flags[i] = keep ? 1 : 0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_swap_keys(const unsigned long long* __restrict__ a, int64_t k,
                                                     unsigned long long* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long x = a[i];
    b[i] = ((x & 0xFFFFFFFFull) << 32) | (x >> 32);
  }
}

__global__ __launch_bounds__(BLOCK) void k_row_count(const unsigned long long* __restrict__ keys, int64_t k,
                                                     int32_t* __restrict__ cnt, int32_t val) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK)
    atomicAdd(&cnt[key_hi(keys[i])], val);
}

__global__ __launch_bounds__(BLOCK) void k_row_zero(const unsigned long long* __restrict__ keys, int64_t k,
                                                    int32_t* __restrict__ a, int32_t* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = key_hi(keys[i]);
    a[u] = 0;
    b[u] = 0;
  }
}

// deg[u] = old degree - deleted + inserted (deg[n] = 0 for the scan's total)
__global__ __launch_bounds__(BLOCK) void k_new_deg(int32_t n, const int32_t* __restrict__ begin,
                                                   const int32_t* __restrict__ rdel,
                                                   const int32_t* __restrict__ rins,
                                                   int32_t* __restrict__ deg) {
  for (int64_t u = (int64_t)blockIdx.x * BLOCK + threadIdx.x; u <= n; u += (int64_t)gridDim.x * BLOCK)
    deg[u] = (u < n) ? (begin[u + 1] - begin[u] - rdel[u] + rins[u]) : 0;
}

// Move surviving old entries.  A workgroup owns 256 consecutive rows and walks
// their entries block-strided; the row of an entry is found by binary search
// over the group's offsets staged in LDS.
constexpr int RB_ROWS = 256;
__global__ __launch_bounds__(BLOCK) void k_rebuild_old(int32_t n, const int32_t* __restrict__ begin,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ nbeg,
                                                       int32_t* __restrict__ adj2,
                                                       const int32_t* __restrict__ rdel,
                                                       const int32_t* __restrict__ rins,
                                                       const unsigned long long* __restrict__ del, int64_t nd,
                                                       const unsigned long long* __restrict__ ins, int64_t ni) {
  __shared__ int32_t s_off[RB_ROWS + 1];
  __shared__ int32_t s_nb[RB_ROWS];
  __shared__ int32_t s_d0[RB_ROWS], s_d1[RB_ROWS], s_i0[RB_ROWS], s_i1[RB_ROWS];
  __shared__ uint8_t s_ch[RB_ROWS];
  for (int64_t r0 = (int64_t)blockIdx.x * RB_ROWS; r0 < n; r0 += (int64_t)gridDim.x * RB_ROWS) {
    const int32_t rows = (int32_t)min<int64_t>(RB_ROWS, n - r0);
    for (int t = threadIdx.x; t <= rows; t += BLOCK) s_off[t] = begin[r0 + t];
    for (int t = threadIdx.x; t < rows; t += BLOCK) {
      const int32_t u = (int32_t)(r0 + t);
      s_nb[t] = nbeg[u];
      const bool ch = rdel[u] != 0 || rins[u] != 0;
      s_ch[t] = ch;
      if (ch) {
        const unsigned long long lo = (unsigned long long)(uint32_t)u << 32;
        const unsigned long long hi = (unsigned long long)(uint32_t)(u + 1) << 32;
        s_d0[t] = (int32_t)lb_u64(del, 0, nd, lo);
        s_d1[t] = (int32_t)lb_u64(del, 0, nd, hi);
        s_i0[t] = (int32_t)lb_u64(ins, 0, ni, lo);
        s_i1[t] = (int32_t)lb_u64(ins, 0, ni, hi);
      }
    }
    __syncthreads();
    const int32_t e0 = s_off[0], e1 = s_off[rows];
    for (int32_t e = e0 + threadIdx.x; e < e1; e += BLOCK) {
      // row r with s_off[r] <= e < s_off[r+1]
      int32_t lo = 0, hi = rows;
      while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid;
        else hi = mid;
      }
      const int32_t r = lo;
      const int32_t k = adj[e];
      int32_t pos = s_nb[r] + (e - s_off[r]);
      if (s_ch[r]) {
        const unsigned long long key = ((unsigned long long)(uint32_t)(r0 + r) << 32) | (uint32_t)k;
        const int64_t dp = lb_u64(del, s_d0[r], s_d1[r], key);
        if (dp < s_d1[r] && del[dp] == key) continue;  // deleted
        pos += -(int32_t)(dp - s_d0[r]) + (int32_t)(lb_u64(ins, s_i0[r], s_i1[r], key) - s_i0[r]);
      }
      adj2[pos] = k;
    }
    __syncthreads();
  }
}

// Place inserted entries at their merged rank.
__global__ __launch_bounds__(BLOCK) void k_rebuild_ins(const int32_t* __restrict__ begin,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ nbeg,
                                                       int32_t* __restrict__ adj2,
                                                       const unsigned long long* __restrict__ del, int64_t nd,
                                                       const unsigned long long* __restrict__ ins, int64_t ni) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < ni; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = ins[i];
    const int32_t u = key_hi(key), x = key_lo(key);
    const unsigned long long lo = (unsigned long long)(uint32_t)u << 32;
    const int64_t t = i - lb_u64(ins, 0, i, lo);
    const int64_t d0 = lb_u64(del, 0, nd, lo);
    const int64_t dl = lb_u64(del, d0, nd, key) - d0;
    const int32_t ol = lb_i32(adj, begin[u], begin[u + 1], x) - begin[u];
    adj2[nbeg[u] + t + ol - dl] = x;
  }
}

// Old SCCs that lost an internal edge.
__global__ __launch_bounds__(BLOCK) void k_mark_aff(const unsigned long long* __restrict__ del, int64_t nd,
                                                    const int32_t* __restrict__ canon,
                                                    int32_t* __restrict__ aff) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nd; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = key_hi(del[i]), v = key_lo(del[i]);
    const int32_t cu = canon[u];
    if (cu == canon[v]) aff[cu] = 1;
  }
}

// Seeds S (heads) and T (tails) of inserts that can merge SCCs.
__global__ __launch_bounds__(BLOCK) void k_ins_seeds(const unsigned long long* __restrict__ ins, int64_t ni,
                                                     const int32_t* __restrict__ canon,
                                                     const int32_t* __restrict__ aff,
                                                     int32_t* __restrict__ S, int32_t* __restrict__ T,
                                                     unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < ni; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool use = false;
    int32_t u = 0, v = 0;
    if (i < ni) {
      u = key_hi(ins[i]);
      v = key_lo(ins[i]);
      const int32_t cu = canon[u];
      use = cu != canon[v] || aff[cu] != 0;
    }
    wave_append(use, v, S, &ctr[C_Q_OUT]);
    wave_append(use, u, T, &ctr[C_AUX_T]);
  }
}

// colour = 0 inside FW(S), -2 outside; the live list becomes FW(S) in live[sel].
__global__ __launch_bounds__(BLOCK) void k_restrict_fw(DevCtl* __restrict__ dc, int32_t n,
                                                       const uint32_t* __restrict__ vf,
                                                       int32_t* __restrict__ color, int32_t* __restrict__ live,
                                                       int32_t sel) {
  using Out = LdsOut<0>;
  Out::init();
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(live, &dc->ctr[C_LIST_OUT], BLOCK);
    const int64_t v = base + threadIdx.x;
    bool in = false;
    if (v < n) {
      in = bit_test(vf, (int32_t)v);
      color[v] = in ? 0 : DONE;
    }
    Out::push(in, (int32_t)v, live, &dc->ctr[C_LIST_OUT]);
  }
  Out::flush(live, &dc->ctr[C_LIST_OUT]);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    dc->live.n = ld_acq(&dc->ctr[C_LIST_OUT]);
    dc->live.sel = sel;
    dc->ctr[C_LIST_OUT] = 0;
    dc->blocks_done = 0;
  }
}

// Region X: recompute state for members, keep the canonical label elsewhere;
// the live list becomes X in live[sel].
__global__ __launch_bounds__(BLOCK) void k_region(DevCtl* __restrict__ dc, int32_t n,
                                                  const int32_t* __restrict__ canon,
                                                  const int32_t* __restrict__ aff,
                                                  const uint32_t* __restrict__ vf,
                                                  const uint32_t* __restrict__ vb, bool have_fb,
                                                  int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                  int32_t* __restrict__ live, int32_t sel) {
  using Out = LdsOut<0>;
  Out::init();
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(live, &dc->ctr[C_LIST_OUT], BLOCK);
    const int64_t v = base + threadIdx.x;
    bool in = false;
    if (v < n) {
      const int32_t cv = canon[v];
      in = aff[cv] != 0 || (have_fb && bit_test(vf, (int32_t)v) && bit_test(vb, (int32_t)v));
      color[v] = in ? UNSET : DONE;
      label[v] = in ? -1 : cv;
    }
    Out::push(in, (int32_t)v, live, &dc->ctr[C_LIST_OUT]);
  }
  Out::flush(live, &dc->ctr[C_LIST_OUT]);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    dc->live.n = ld_acq(&dc->ctr[C_LIST_OUT]);
    dc->live.sel = sel;
    dc->ctr[C_LIST_OUT] = 0;
    dc->blocks_done = 0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_clear_aff(const unsigned long long* __restrict__ del, int64_t nd,
                                                     const int32_t* __restrict__ canon,
                                                     int32_t* __restrict__ aff) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nd; i += (int64_t)gridDim.x * BLOCK)
    aff[canon[key_hi(del[i])]] = 0;
}

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

Dyn* dyn_of(scc_ctx* c) { return (Dyn*)c->dyn; }

int ensure_dyn(scc_ctx* c, int64_t batch) {
  Dyn* d = dyn_of(c);
  if (!d) {
    d = new Dyn();
    c->dyn = d;
  }
  int rc;
  if (d->n != c->n) {
    // per-row arrays
    int32_t** rows[] = {&d->rdel_f, &d->rins_f, &d->rdel_r, &d->rins_r, &d->nb_f, &d->nb_r, &d->deg,
                        &d->aff};
    for (auto p : rows) {
      dev_free(c, *p);
      if ((rc = dev_alloc(c, (void**)p, 4ull * ((uint64_t)c->n + 1), "dynamic row arrays"))) return rc;
      launch_fill_i32(c->stream, *p, 0, c->n + 1);
    }
    d->n = c->n;
  }
  if (batch > d->cap_b) {
    const int64_t cap = std::max<int64_t>(batch, 1024);
    unsigned long long** ks[] = {&d->k_in, &d->k_tmp, &d->del_f, &d->del_r, &d->ins_f, &d->ins_r,
                                 &d->ins_sorted};
    for (auto p : ks) {
      dev_free(c, *p);
      if ((rc = dev_alloc(c, (void**)p, 8ull * cap, "batch keys"))) return rc;
    }
    dev_free(c, d->flags);
    dev_free(c, d->src);
    dev_free(c, d->dst);
    if ((rc = dev_alloc(c, (void**)&d->flags, cap, "batch flags")) ||
        (rc = dev_alloc(c, (void**)&d->src, 4ull * cap, "batch src")) ||
        (rc = dev_alloc(c, (void**)&d->dst, 4ull * cap, "batch dst")))
      return rc;
    if (d->h_stage) (void)hipHostFree(d->h_stage);
    d->h_stage = nullptr;
    if ((rc = hip_check(c, hipHostMalloc((void**)&d->h_stage, 8ull * cap, 0), "pinned staging")))
      return rc;
    // temp storage for sort / select / scan at this size
    size_t a = 0, b = 0, s = 0, m = 0;
    hipcub::DeviceRadixSort::SortKeys(nullptr, a, d->k_in, d->k_tmp, (int)cap);
    hipcub::DeviceSelect::Flagged(nullptr, b, d->k_in, d->flags, d->del_f, (int32_t*)nullptr, (int)cap);
    hipcub::DeviceScan::ExclusiveSum(nullptr, s, d->deg, d->nb_f, c->n + 1);
    m = std::max(std::max(a, b), s) + 256;
    if (m > d->tmp_bytes) {
      dev_free(c, d->tmp);
      if ((rc = dev_alloc(c, &d->tmp, m, "cub temp"))) return rc;
      d->tmp_bytes = m;
    }
    d->cap_b = cap;
  } else {
    size_t s = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, s, d->deg, d->nb_f, c->n + 1);
    if (s + 256 > d->tmp_bytes) {
      dev_free(c, d->tmp);
      if ((rc = dev_alloc(c, &d->tmp, s + 256, "cub temp"))) return rc;
      d->tmp_bytes = s + 256;
    }
  }
  return SCC_OK;
}

int sort_keys(scc_ctx* c, Dyn* d, unsigned long long* in, unsigned long long* out, int64_t k) {
  if (k == 0) return SCC_OK;
  size_t t = d->tmp_bytes;
  return hip_check(c, hipcub::DeviceRadixSort::SortKeys(d->tmp, t, in, out, (int)k, 0, 64, c->stream),
                   "radix sort");
}

int select_flagged(scc_ctx* c, Dyn* d, unsigned long long* in, unsigned long long* out, int64_t k,
                   int64_t* nsel) {
  *nsel = 0;
  if (k == 0) return SCC_OK;
  size_t t = d->tmp_bytes;
  int32_t* dn = (int32_t*)(c->ctr + C_AUX_SEL);  // DevCtl::ctr slot on the device
  int rc = hip_check(c, hipcub::DeviceSelect::Flagged(d->tmp, t, in, d->flags, out, dn, (int)k, c->stream),
                     "select");
  if (rc) return rc;
  int32_t h = 0;
  if ((rc = hip_check(c, hipMemcpyAsync(&h, dn, 4, hipMemcpyDeviceToHost, c->stream), "select count")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "select sync")))
    return rc;
  *nsel = h;
  return SCC_OK;
}

// Rebuild one direction: off/nb -> noff/nb2 using del/ins key lists sorted by row.
int rebuild_dir(scc_ctx* c, Dyn* d, const int32_t* off, const int32_t* nb, int32_t* rdel, int32_t* rins,
                int32_t* noff, int32_t* nb2, const unsigned long long* del, int64_t nd,
                const unsigned long long* ins, int64_t ni) {
  const int32_t n = c->n;
  if (nd) LAUNCH(k_row_count, grid_for(nd), c->stream, del, nd, rdel, 1);
  if (ni) LAUNCH(k_row_count, grid_for(ni), c->stream, ins, ni, rins, 1);
  LAUNCH(k_new_deg, grid_for((int64_t)n + 1, 4096), c->stream, n, off, rdel, rins, d->deg);
  size_t t = d->tmp_bytes;
  int rc = hip_check(c, hipcub::DeviceScan::ExclusiveSum(d->tmp, t, d->deg, noff, n + 1, c->stream), "scan");
  if (rc) return rc;
  LAUNCH(k_rebuild_old, grid_for(n, 4096), c->stream, n, off, nb, noff, nb2, rdel, rins, del, nd, ins, ni);
  if (ni) LAUNCH(k_rebuild_ins, grid_for(ni), c->stream, off, nb, noff, nb2, del, nd, ins, ni);
  if (nd) LAUNCH(k_row_zero, grid_for(nd), c->stream, del, nd, rdel, rins);
  if (ni) LAUNCH(k_row_zero, grid_for(ni), c->stream, ins, ni, rdel, rins);
  return hip_check(c, hipGetLastError(), "rebuild");
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

void destroy_sched(scc_ctx* c) { (void)c; }

void destroy_dynamic(scc_ctx* c) {
  Dyn* d = dyn_of(c);
  if (!d) return;
  void* ptrs[] = {d->k_in, d->k_tmp, d->del_f, d->del_r, d->ins_f, d->ins_r, d->ins_sorted,
                  d->flags, d->src, d->dst, d->rdel_f, d->rins_f, d->rdel_r, d->rins_r,
                  d->nb_f, d->nb_r, d->deg, d->aff, d->adj2, d->radj2, d->tmp};
  for (void* p : ptrs) dev_free(c, p);
  if (d->h_stage) (void)hipHostFree(d->h_stage);
  delete d;
  c->dyn = nullptr;
}

}  // namespace sccmi

using namespace sccmi;

extern "C" int scc_update(scc_ctx* c, const int32_t* ins_src, const int32_t* ins_dst, int64_t n_ins,
                          const int32_t* del_src, const int32_t* del_dst, int64_t n_del,
                          scc_update_stats* stats) {
  if (!c) return SCC_ERR_ARG;
  if (!c->begin) return set_error(c, SCC_ERR_STATE, "no graph uploaded");
  (void)hipSetDevice(c->device);
  if (!c->labels_ready) {
    // phase-by-phase callers: labels are a full decomposition once no vertex is live
    int r0 = do_trim_compact(c);
    if (r0) return r0;
    if (c->n_live != 0)
      return set_error(c, SCC_ERR_STATE, "scc_update needs a finished decomposition (run first)");
    if ((r0 = do_canon(c, nullptr))) return r0;
  }
  if (n_ins < 0 || n_del < 0 || (n_ins && (!ins_src || !ins_dst)) || (n_del && (!del_src || !del_dst)))
    return set_error(c, SCC_ERR_ARG, "bad batch");
  for (int64_t i = 0; i < n_ins; i++)
    if ((uint32_t)ins_src[i] >= (uint32_t)c->n || (uint32_t)ins_dst[i] >= (uint32_t)c->n)
      return set_error(c, SCC_ERR_ARG, "insert endpoint out of range at " + std::to_string(i));
  for (int64_t i = 0; i < n_del; i++)
    if ((uint32_t)del_src[i] >= (uint32_t)c->n || (uint32_t)del_dst[i] >= (uint32_t)c->n)
      return set_error(c, SCC_ERR_ARG, "delete endpoint out of range at " + std::to_string(i));
  if (c->m + n_ins >= INT32_MAX) return set_error(c, SCC_ERR_LIMIT, "edge count would exceed int32");
  (void)hipSetDevice(c->device);
  auto t0 = std::chrono::steady_clock::now();
  scc_update_stats st;
  std::memset(&st, 0, sizeof(st));
  int rc;
  if ((rc = ensure_dyn(c, std::max(n_ins, n_del)))) return rc;
  Dyn* d = dyn_of(c);
  hipStream_t s = c->stream;

  // ---- 1. sorted, filtered batch lists ------------------------------------
  auto upload_keys = [&](const int32_t* a, const int32_t* b, int64_t k, bool drop_loops) -> int {
    if (k == 0) return SCC_OK;
    std::memcpy(d->h_stage, a, 4ull * k);
    std::memcpy(d->h_stage + k, b, 4ull * k);
    int r;
    if ((r = hip_check(c, hipMemcpyAsync(d->src, d->h_stage, 4ull * k, hipMemcpyHostToDevice, s), "batch upload")) ||
        (r = hip_check(c, hipMemcpyAsync(d->dst, d->h_stage + k, 4ull * k, hipMemcpyHostToDevice, s), "batch upload")))
      return r;
    LAUNCH(k_make_keys, grid_for(k), s, d->src, d->dst, k, drop_loops, d->k_in);
    // the staging buffer is reused right away: wait for the copies
    return hip_check(c, hipStreamSynchronize(s), "batch upload sync");
  };
  // inserts: sorted keys (all, for the delete filter) then the absent ones
  int64_t ni = 0, nd = 0;
  if ((rc = upload_keys(ins_src, ins_dst, n_ins, true))) return rc;
  if ((rc = sort_keys(c, d, d->k_in, d->ins_sorted, n_ins))) return rc;
  if (n_ins) {
    LAUNCH(k_flag_ins, grid_for(n_ins), s, d->ins_sorted, n_ins, c->begin, c->adj, d->flags);
    if ((rc = select_flagged(c, d, d->ins_sorted, d->ins_f, n_ins, &ni))) return rc;
  }
  if ((rc = upload_keys(del_src, del_dst, n_del, false))) return rc;
  if ((rc = sort_keys(c, d, d->k_in, d->k_tmp, n_del))) return rc;
  if (n_del) {
    LAUNCH(k_flag_del, grid_for(n_del), s, d->k_tmp, n_del, d->ins_sorted, n_ins, c->begin, c->adj, d->flags);
    if ((rc = select_flagged(c, d, d->k_tmp, d->del_f, n_del, &nd))) return rc;
  }
  // reverse-direction lists
  if (ni) {
    LAUNCH(k_swap_keys, grid_for(ni), s, d->ins_f, ni, d->k_in);
    if ((rc = sort_keys(c, d, d->k_in, d->ins_r, ni))) return rc;
  }
  if (nd) {
    LAUNCH(k_swap_keys, grid_for(nd), s, d->del_f, nd, d->k_in);
    if ((rc = sort_keys(c, d, d->k_in, d->del_r, nd))) return rc;
  }
  st.inserted = ni;
  st.deleted = nd;
  if (ni == 0 && nd == 0) {
    st.num_sccs = -1;
    int64_t cnt = 0;
    if ((rc = do_canon(c, &cnt))) return rc;
    st.num_sccs = cnt;
    st.ms_total = ms_since(t0);
    if (stats) *stats = st;
    return SCC_OK;
  }

  // ---- 2. rebuild CSR and reverse CSR ---------------------------------------
  auto ta = std::chrono::steady_clock::now();
  const int64_t m_new = c->m + ni - nd;
  if (m_new > d->cap2 || !d->adj2) {
    dev_free(c, d->adj2);
    dev_free(c, d->radj2);
    d->adj2 = d->radj2 = nullptr;
    const int64_t cap = m_new + m_new / 32 + 4096;
    if ((rc = dev_alloc(c, (void**)&d->adj2, 4ull * cap, "spare adj")) ||
        (rc = dev_alloc(c, (void**)&d->radj2, 4ull * cap, "spare radj")))
      return rc;
    d->cap2 = cap;
  }
  if ((rc = rebuild_dir(c, d, c->begin, c->adj, d->rdel_f, d->rins_f, d->nb_f, d->adj2, d->del_f, nd,
                        d->ins_f, ni)) ||
      (rc = rebuild_dir(c, d, c->rbegin, c->radj, d->rdel_r, d->rins_r, d->nb_r, d->radj2, d->del_r, nd,
                        d->ins_r, ni)))
    return rc;
  std::swap(c->begin, d->nb_f);
  std::swap(c->rbegin, d->nb_r);
  std::swap(c->adj, d->adj2);
  std::swap(c->radj, d->radj2);
  std::swap(c->m_cap, d->cap2);
  c->m = m_new;
  if ((rc = hip_check(c, hipStreamSynchronize(s), "rebuild sync"))) return rc;
  st.ms_apply = ms_since(ta);

  // ---- 3. affected region -----------------------------------------------------
  auto tr = std::chrono::steady_clock::now();
  if (nd) LAUNCH(k_mark_aff, grid_for(nd), s, d->del_f, nd, c->canon, d->aff);
  if ((rc = zero_ctr(c))) return rc;
  // S (heads) -> scratch_i32, T (tails) -> W; both idle until the recompute
  int32_t* S = c->scratch_i32;
  int32_t* T = c->W;
  if (ni) LAUNCH(k_ins_seeds, grid_for(ni), s, d->ins_f, ni, c->canon, d->aff, S, T, c->ctr);
  if ((rc = pull_ctl(c))) return rc;
  const int64_t ns = (int64_t)c->h_ctr[C_Q_OUT];
  bool have_fb = false;
  const int64_t words = ((int64_t)c->n + 31) / 32;
  if (ns > 0) {
    // FW(S) over all of G'
    launch_fill_i32(s, c->color, UNSET, c->n);
    launch_iota(s, c->live[0], c->n);
    if ((rc = set_live(c, 0, c->n))) return rc;
    launch_fill_u32(s, c->vf, 0u, words);
    launch_fill_u32(s, c->vb, 0u, words);
    if ((rc = traverse_from(c, true, false, S, c->ctr + C_Q_OUT, nullptr))) return rc;
    // BW(T) inside FW(S)
    LAUNCH(k_restrict_fw, grid_for(c->n), s, c->dc, c->n, c->vf, c->color, c->live[1], 1);
    if ((rc = pull_ctl(c))) return rc;
    if ((rc = traverse_from(c, false, false, T, c->ctr + C_AUX_T, nullptr))) return rc;
    have_fb = true;
  }
  LAUNCH(k_region, grid_for(c->n), s, c->dc, c->n, c->canon, d->aff, c->vf, c->vb, have_fb, c->color,
         c->label, c->live[0], 0);
  // clear the old-SCC flags (indexed by old canonical labels, before canon changes)
  if (nd) LAUNCH(k_clear_aff, grid_for(nd), s, d->del_f, nd, c->canon, d->aff);
  if ((rc = pull_ctl(c))) return rc;
  st.affected = c->n_live;
  c->colors_split = false;
  st.ms_region = ms_since(tr);

  // ---- 4. recompute inside the region ------------------------------------------
  auto tc = std::chrono::steady_clock::now();
  c->labels_ready = false;
  if (c->n_live > 0) {
    scc_run_opts o;
    scc_run_opts_default(&o);
    o.pivot_policy = c->update_policy;
    scc_run_stats rs;
    if ((rc = scc_run_method2(c, &o, &rs))) return rc;
    st.num_sccs = rs.num_sccs;
  } else {
    int64_t cnt = 0;
    if ((rc = do_canon(c, &cnt))) return rc;
    st.num_sccs = cnt;
  }
  st.ms_recompute = ms_since(tc);
  st.ms_total = ms_since(t0);
  if (stats) *stats = st;
  return SCC_OK;
}

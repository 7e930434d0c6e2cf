// Vertex layout (scc_ctx_set_layout): with SCC_LAYOUT_DEGREE (or AUTO on a
// skewed graph) the uploaded
// graph is relabelled inside the context so that internal id i is the i-th
// vertex by decreasing degree class (the bit length of in + out degree: 0, 1,
// 2-3, 4-7, ...), caller id order within a class.  On skewed graphs the hubs'
// per-vertex state (colours, reach bits, WCC words) then sits in a few cache
// lines that every search level and sweep reuses, instead of being spread over
// the whole arrays.  The rows are rebuilt sorted by internal id, which is also
// hub-first order, so no traversal copies are needed.  Keeping caller order
// inside a class (rather than sorting by exact degree) keeps the writes of
// caller-indexed results (canonical labels at perm[i]) close together: R-MAT-22
// decomposes 6-7 us faster per step, R-MAT-23 5-10 us (classes of two bits per
// octave or of every fourth power of two measured slower).
//
// Every entry point keeps the caller's ids: perm[i] = caller id of internal
// vertex i, iperm[v] = internal id of caller vertex v.  Labels and colours are
// member ids, so both their index and their value are mapped on the way out;
// canonical labels are the minimum caller id of each component.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

// Degree class: the bit length of the total degree (0 for an isolated vertex).
__device__ __forceinline__ uint32_t degree_class(uint32_t d) { return 32u - (uint32_t)__clz(d); }

// ~degree class: an ascending (stable) sort puts the largest classes first,
// caller ids ascending within a class
__global__ __launch_bounds__(BLOCK) void k_neg_degree(int32_t n, const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ rbegin,
                                                      uint32_t* __restrict__ key, int32_t* __restrict__ id) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    key[v] = ~degree_class((uint32_t)(begin[v + 1] - begin[v]) + (uint32_t)(rbegin[v + 1] - rbegin[v]));
    id[v] = (int32_t)v;
  }
}

__global__ __launch_bounds__(BLOCK) void k_invert(int32_t n, const int32_t* __restrict__ perm,
                                                  int32_t* __restrict__ iperm) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    iperm[perm[i]] = (int32_t)i;
}

// row lengths of the relabelled rows: len[i] = length of caller row perm[i];
// len[n] = 0 so that an exclusive scan over n + 1 entries ends with m
__global__ __launch_bounds__(BLOCK) void k_new_lengths(int32_t n, const int32_t* __restrict__ begin,
                                                       const int32_t* __restrict__ perm,
                                                       int32_t* __restrict__ len) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i <= n; i += (int64_t)gridDim.x * BLOCK) {
    if (i == n) {
      len[i] = 0;
    } else {
      const int32_t v = perm[i];
      len[i] = begin[v + 1] - begin[v];
    }
  }
}

// key = internal row << 32 | internal neighbour; a wave per caller row
// (measured faster than a lane per row with the wave taking the long rows:
// 1.59 vs 1.84 ms per direction at R-MAT-22)
__global__ __launch_bounds__(BLOCK) void k_edge_keys(int32_t n, const int32_t* __restrict__ begin,
                                                     const int32_t* __restrict__ adj,
                                                     const int32_t* __restrict__ iperm,
                                                     unsigned long long* __restrict__ keys) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * BLOCK) >> 6;
  for (int64_t v = wave; v < n; v += nw) {
    const int32_t s = begin[v], e = begin[v + 1];
    const unsigned long long row = (unsigned long long)(uint32_t)iperm[v] << 32;
    for (int32_t i = s + lane; i < e; i += 64) keys[i] = row | (uint32_t)iperm[adj[i]];
  }
}

__global__ __launch_bounds__(BLOCK) void k_low_words(int64_t m, const unsigned long long* __restrict__ keys,
                                                     int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < m; i += (int64_t)gridDim.x * BLOCK)
    out[i] = (int32_t)(uint32_t)keys[i];
}

// out[perm[i]] = in[i], values that are vertex ids (>= 0) mapped through perm
// when map_values (labels, colours); negative codes pass unchanged
__global__ __launch_bounds__(BLOCK) void k_to_caller(int32_t n, const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ in, int map_values,
                                                     int32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t x = in[i];
    out[perm[i]] = (map_values && x >= 0) ? perm[x] : x;
  }
}

// out[iperm[v]] = in[v], values mapped through iperm (caller-set colours)
__global__ __launch_bounds__(BLOCK) void k_to_internal(int32_t n, const int32_t* __restrict__ iperm,
                                                       const int32_t* __restrict__ in, int32_t* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t x = in[v];
    out[iperm[v]] = x >= 0 ? iperm[x] : x;
  }
}

// Segmented order for batched graphs: key = graph << 32 | ~degree class of
// internal vertex i, value = its caller id (graph g owns caller ids
// [boff[g], boff[g + 1])).
__global__ __launch_bounds__(BLOCK) void k_segment_keys(int32_t n, const int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ boff, int32_t k,
                                                        const int32_t* __restrict__ begin,
                                                        const int32_t* __restrict__ rbegin,
                                                        unsigned long long* __restrict__ key,
                                                        int32_t* __restrict__ val) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = perm[i];
    int32_t lo = 0, hi = k;  // boff[lo] <= v < boff[hi]
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (boff[mid] <= v) lo = mid;
      else hi = mid;
    }
    const uint32_t deg = (uint32_t)(begin[i + 1] - begin[i]) + (uint32_t)(rbegin[i + 1] - rbegin[i]);
    key[i] = ((unsigned long long)(uint32_t)lo << 32) | (uint32_t)~degree_class(deg);
    val[i] = v;
  }
}

// rel[j] = cur_iperm[q[j]]: the current internal id of the vertex that the new
// order puts at j
__global__ __launch_bounds__(BLOCK) void k_compose(int32_t n, const int32_t* __restrict__ q,
                                                   const int32_t* __restrict__ cur_iperm,
                                                   int32_t* __restrict__ rel) {
  for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < n; j += (int64_t)gridDim.x * BLOCK)
    rel[j] = cur_iperm[q[j]];
}

// One direction: rows of (begin, adj) relabelled in place (begin, adj hold the
// caller's rows on entry and the internal rows, sorted by id, on exit).
int relabel_dir(scc_ctx* c, int32_t* begin, int32_t* adj) {
  const int64_t m = c->m;
  const int32_t n = c->n;
  int rc;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t* len = nullptr;
  void* tmp = nullptr;
  auto done = [&](int r) {
    dev_free(c, k0);
    dev_free(c, k1);
    dev_free(c, len);
    dev_free(c, tmp);
    return r;
  };
  if ((rc = dev_alloc(c, (void**)&len, 4ull * ((uint64_t)n + 1), "layout row lengths"))) return done(rc);
  LAUNCH(k_new_lengths, grid_for((int64_t)n + 1, 4096), c->stream, n, begin, c->perm, len);
  if (m > 0) {
    if ((rc = dev_alloc(c, (void**)&k0, 8ull * m, "layout keys")) ||
        (rc = dev_alloc(c, (void**)&k1, 8ull * m, "layout keys")))
      return done(rc);
    LAUNCH(k_edge_keys, grid_for((int64_t)n * 64, 4096), c->stream, n, begin, adj, c->iperm, k0);
  }
  int bits = 32;
  while (bits < 64 && (1ull << (bits - 32)) < (unsigned long long)n) bits++;
  hipcub::DoubleBuffer<unsigned long long> keys(k0, k1);
  size_t tb_sort = 0, tb_scan = 0;
  if (m > 0) (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tb_sort, keys, (int)m, 0, bits, c->stream);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb_scan, len, begin, n + 1, c->stream);
  if ((rc = dev_alloc(c, &tmp, std::max(tb_sort, tb_scan) + 256, "layout temp"))) return done(rc);
  if (m > 0) {
    if ((rc = hip_check(c, hipcub::DeviceRadixSort::SortKeys(tmp, tb_sort, keys, (int)m, 0, bits, c->stream),
                        "layout sort")))
      return done(rc);
    LAUNCH(k_low_words, grid_for(m, 4096), c->stream, m, keys.Current(), adj);
  }
  // the lengths were read from the caller's offsets above; now overwrite them
  if ((rc = hip_check(c, hipcub::DeviceScan::ExclusiveSum(tmp, tb_scan, len, begin, n + 1, c->stream),
                      "layout scan")))
    return done(rc);
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "layout sync"))) return done(rc);
  return done(SCC_OK);
}

}  // namespace

void drop_layout(scc_ctx* c) {
  dev_free(c, c->perm);
  dev_free(c, c->iperm);
  dev_free(c, c->map_a);
  dev_free(c, c->map_b);
  dev_free(c, c->canon_caller);
  dev_free(c, c->cwin);
  c->cwin = nullptr;
  c->cwin_runs = c->cwin_nwin = 0;
  c->perm = c->iperm = c->map_a = c->map_b = c->canon_caller = nullptr;
  c->h_perm.clear();
  c->h_iperm.clear();
  c->relabeled = false;
}

// The window table of k_canon_window from the host copy of perm: runs are the
// maximal internal ranges over which perm increases (the degree classes); for
// each window boundary B = w * CANON_WIN, a run's entry is its first internal
// id whose caller id is >= B.  More than CANON_RUNS runs (a batch's per-graph
// order): no table, and the canonical labels take the general path.
static int build_canon_windows(scc_ctx* c) {
  dev_free(c, c->cwin);
  c->cwin = nullptr;
  c->cwin_runs = c->cwin_nwin = 0;
  const int32_t n = c->n;
  if (n <= 0 || (int64_t)c->h_perm.size() != n) return SCC_OK;
  const int32_t* p = c->h_perm.data();
  std::vector<int32_t> start{0};
  for (int32_t v = 1; v < n; v++)
    if (p[v] < p[v - 1]) {
      if ((int)start.size() >= CANON_RUNS) return SCC_OK;
      start.push_back(v);
    }
  const int32_t runs = (int32_t)start.size();
  start.push_back(n);
  const int64_t nwin = ((int64_t)n + CANON_WIN - 1) / CANON_WIN;
  std::vector<int32_t> T((size_t)(nwin + 1) * runs);
  for (int32_t r = 0; r < runs; r++) {
    int32_t pos = start[r];
    const int32_t e = start[r + 1];
    for (int64_t w = 0; w <= nwin; w++) {
      const int64_t bound = std::min<int64_t>(n, w * CANON_WIN);
      while (pos < e && p[pos] < bound) pos++;
      T[(size_t)w * runs + r] = pos;
    }
  }
  int rc;
  if ((rc = dev_alloc(c, (void**)&c->cwin, 4ull * T.size(), "label windows"))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(c->cwin, T.data(), 4ull * T.size(), hipMemcpyHostToDevice, c->stream),
                      "label windows")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "label windows sync"))) {
    dev_free(c, c->cwin);
    c->cwin = nullptr;
    return rc;
  }
  c->cwin_runs = runs;
  c->cwin_nwin = (int32_t)nwin;
  return SCC_OK;
}

int build_layout(scc_ctx* c) {
  drop_layout(c);
  if (c->n <= 0) return SCC_OK;
  if (c->layout_req == SCC_LAYOUT_AUTO) {
    // skewed graphs only: hubs are what the relabelling packs together
    const double mean = std::max(1.0, (double)c->m / (double)c->n);
    if ((double)c->maxdeg <= 32.0 * mean) return SCC_OK;
  } else if (c->layout_req != SCC_LAYOUT_DEGREE) {
    return SCC_OK;
  }
  const int32_t n = c->n;
  int rc;
  uint32_t *key0 = nullptr, *key1 = nullptr;
  int32_t* ids = nullptr;
  void* tmp = nullptr;
  auto done = [&](int r) {
    dev_free(c, key0);
    dev_free(c, key1);
    dev_free(c, ids);
    dev_free(c, tmp);
    if (r) drop_layout(c);
    return r;
  };
  if ((rc = dev_alloc(c, (void**)&c->perm, 4ull * n, "layout perm")) ||
      (rc = dev_alloc(c, (void**)&c->iperm, 4ull * n, "layout iperm")) ||
      (rc = dev_alloc(c, (void**)&c->map_a, 4ull * ((uint64_t)n + 1), "layout map")) ||
      (rc = dev_alloc(c, (void**)&c->map_b, 4ull * ((uint64_t)n + 1), "layout map")) ||
      (rc = dev_alloc(c, (void**)&c->canon_caller, 4ull * ((uint64_t)n + 1), "caller labels")) ||
      (rc = dev_alloc(c, (void**)&key0, 4ull * n, "layout degree keys")) ||
      (rc = dev_alloc(c, (void**)&key1, 4ull * n, "layout degree keys")) ||
      (rc = dev_alloc(c, (void**)&ids, 4ull * n, "layout ids")))
    return done(rc);
  LAUNCH(k_neg_degree, grid_for(n, 4096), c->stream, n, c->begin, c->rbegin, key0, ids);
  // radix sorts are stable: equal degrees keep the caller's id order
  size_t tb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key0, key1, ids, c->perm, n, 0, 32, c->stream);
  if ((rc = dev_alloc(c, &tmp, tb + 256, "layout temp"))) return done(rc);
  if ((rc = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, key0, key1, ids, c->perm, n, 0, 32, c->stream),
                      "layout degree sort")))
    return done(rc);
  LAUNCH(k_invert, grid_for(n, 4096), c->stream, n, c->perm, c->iperm);
  dev_free(c, tmp);
  tmp = nullptr;
  if ((rc = relabel_dir(c, c->begin, c->adj)) || (rc = relabel_dir(c, c->rbegin, c->radj))) return done(rc);
  c->h_perm.resize(n);
  c->h_iperm.resize(n);
  if ((rc = hip_check(c, hipMemcpyAsync(c->h_perm.data(), c->perm, 4ull * n, hipMemcpyDeviceToHost, c->stream),
                      "layout download")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->h_iperm.data(), c->iperm, 4ull * n, hipMemcpyDeviceToHost, c->stream),
                      "layout download")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "layout sync")))
    return done(rc);
  if ((rc = build_canon_windows(c))) return done(rc);
  c->relabeled = true;
  return done(SCC_OK);
}

int layout_to_caller(scc_ctx* c, const int32_t* in, bool map_values, int32_t* out_dev) {
  LAUNCH(k_to_caller, grid_for(c->n, 4096), c->stream, c->n, c->perm, in, map_values ? 1 : 0, out_dev);
  return hip_check(c, hipGetLastError(), "layout map");
}

int layout_segment(scc_ctx* c, int32_t k) {
  if (!c->relabeled) return SCC_OK;
  if (c->end) return set_error(c, SCC_ERR_STATE, "batched graphs in the degree layout need an unchanged upload");
  const int32_t n = c->n;
  int rc;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *q = nullptr;
  void* tmp = nullptr;
  auto done = [&](int r) {
    dev_free(c, k0);
    dev_free(c, k1);
    dev_free(c, v0);
    dev_free(c, q);
    dev_free(c, tmp);
    return r;
  };
  if ((rc = dev_alloc(c, (void**)&k0, 8ull * n, "segment keys")) ||
      (rc = dev_alloc(c, (void**)&k1, 8ull * n, "segment keys")) ||
      (rc = dev_alloc(c, (void**)&v0, 4ull * n, "segment ids")) ||
      (rc = dev_alloc(c, (void**)&q, 4ull * n, "segment order")))
    return done(rc);
  LAUNCH(k_segment_keys, grid_for(n, 4096), c->stream, n, c->perm, c->d_boff, k, c->begin, c->rbegin, k0, v0);
  int bits = 32;
  while (bits < 64 && (1ull << (bits - 32)) < (unsigned long long)k) bits++;
  size_t tb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, q, n, 0, bits, c->stream);
  if ((rc = dev_alloc(c, &tmp, tb + 256, "segment temp"))) return done(rc);
  // stable: equal (graph, degree) keep the current order, i.e. caller id order
  if ((rc = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, q, n, 0, bits, c->stream),
                      "segment sort")))
    return done(rc);
  // rows move from the current internal order to the new one: relabel_dir with
  // the relative permutation (map_a: new -> current, map_b: current -> new)
  LAUNCH(k_compose, grid_for(n, 4096), c->stream, n, q, c->iperm, c->map_a);
  LAUNCH(k_invert, grid_for(n, 4096), c->stream, n, c->map_a, c->map_b);
  std::swap(c->perm, c->map_a);
  std::swap(c->iperm, c->map_b);
  rc = relabel_dir(c, c->begin, c->adj);
  if (!rc) rc = relabel_dir(c, c->rbegin, c->radj);
  std::swap(c->perm, c->map_a);
  std::swap(c->iperm, c->map_b);
  if (rc) {  // the rows are half moved: the graph is unusable
    c->relabeled = false;
    return done(rc);
  }
  // the new order itself: perm = q, iperm = its inverse
  if ((rc = hip_check(c, hipMemcpyAsync(c->perm, q, 4ull * n, hipMemcpyDeviceToDevice, c->stream), "segment perm")))
    return done(rc);
  LAUNCH(k_invert, grid_for(n, 4096), c->stream, n, c->perm, c->iperm);
  // the rows moved: their heads follow
  if ((rc = build_heads(c))) return done(rc);
  if ((rc = hip_check(c, hipMemcpyAsync(c->h_perm.data(), c->perm, 4ull * n, hipMemcpyDeviceToHost, c->stream),
                      "layout download")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->h_iperm.data(), c->iperm, 4ull * n, hipMemcpyDeviceToHost, c->stream),
                      "layout download")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "segment sync")))
    return done(rc);
  // (per graph and class: usually more runs than the table holds)
  return done(build_canon_windows(c));
}

int layout_colors_to_internal(scc_ctx* c, const int32_t* colors_host) {
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(c->map_b, colors_host, 4ull * c->n, hipMemcpyHostToDevice, c->stream),
                      "colours upload")))
    return rc;
  LAUNCH(k_to_internal, grid_for(c->n, 4096), c->stream, c->n, c->iperm, c->map_b, c->color);
  return hip_check(c, hipStreamSynchronize(c->stream), "colours sync");
}

// Rows of internal CSR (ib, ia) as caller rows (ob, oa), each sorted by id.
void layout_rows_to_caller(const scc_ctx* c, const std::vector<int32_t>& ib, const std::vector<int32_t>& ia,
                           int32_t* ob, int32_t* oa) {
  const int32_t n = c->n;
  std::vector<int32_t> off((size_t)n + 1, 0);
  for (int32_t v = 0; v < n; v++) {
    const int32_t i = c->h_iperm[v];
    off[v + 1] = off[v] + (ib[i + 1] - ib[i]);
  }
  if (ob) std::copy(off.begin(), off.end(), ob);
  if (!oa) return;
  for (int32_t v = 0; v < n; v++) {
    const int32_t i = c->h_iperm[v];
    int32_t* row = oa + off[v];
    const int32_t len = ib[i + 1] - ib[i];
    for (int32_t j = 0; j < len; j++) row[j] = c->h_perm[ia[ib[i] + j]];
    std::sort(row, row + len);
  }
}

}  // namespace sccmi

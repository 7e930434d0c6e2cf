// Traversal copies of the adjacency: every row reordered by decreasing degree
// (in + out) of the neighbour, ties by id.  The searches and trims only ask
// "is there a neighbour with property X" and stop at the first one; on
// skewed graphs the neighbours with the property (reached, still live) are
// mostly the hubs, so meeting them first shortens the scans.  The id-sorted
// rows stay the reference layout (downloads, update membership searches); the
// copies are built at upload (one stable radix sort per direction on
// (row, ~degree) keys, so equal degrees keep the id order) and dropped once
// the graph is updated.  With them (and in the degree layout, whose id-sorted
// rows are already in that order) go the row heads: the first entry of every
// traversal row in a dense array per direction.  A bottom-up search level
// tests a candidate's first opposite neighbour before anything else and most
// candidates stop there; reading it from the heads is a sequential 4-byte
// stream instead of one row-start cache line per candidate.
#include <hipcub/hipcub.hpp>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

__global__ __launch_bounds__(BLOCK) void k_row_heads(int32_t n, const int32_t* __restrict__ begin,
                                                     const int32_t* __restrict__ end,
                                                     const int32_t* __restrict__ adj, int32_t* __restrict__ head) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t s = begin[v];
    head[v] = s < end[v] ? adj[s] : -1;
  }
}

__global__ __launch_bounds__(BLOCK) void k_degsum(int32_t n, const int32_t* __restrict__ begin,
                                                  const int32_t* __restrict__ rbegin, uint32_t* __restrict__ deg) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK)
    deg[v] = (uint32_t)(begin[v + 1] - begin[v]) + (uint32_t)(rbegin[v + 1] - rbegin[v]);
}

// key = row << 32 | ~deg(neighbour); a wave per row keeps the writes coalesced
__global__ __launch_bounds__(BLOCK) void k_order_keys(int32_t n, const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ adj,
                                                      const uint32_t* __restrict__ deg,
                                                      unsigned long long* __restrict__ keys) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * BLOCK) >> 6;
  for (int64_t v = wave; v < n; v += nw) {
    const int32_t s = begin[v], e = begin[v + 1];
    for (int32_t i = s + lane; i < e; i += 64)
      keys[i] = ((unsigned long long)v << 32) | (uint32_t)~deg[adj[i]];
  }
}

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

int order_dir(scc_ctx* c, const int32_t* begin, const int32_t* adj, const uint32_t* deg, int32_t** out) {
  const int64_t m = c->m;
  int rc;
  unsigned long long *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *v1 = nullptr;
  void* tmp = nullptr;
  auto fail = [&](int r) {
    dev_free(c, k0);
    dev_free(c, k1);
    dev_free(c, v0);
    dev_free(c, v1);
    dev_free(c, tmp);
    return r;
  };
  if ((rc = dev_alloc(c, (void**)&k0, 8ull * m, "order keys")) ||
      (rc = dev_alloc(c, (void**)&k1, 8ull * m, "order keys")) ||
      (rc = dev_alloc(c, (void**)&v0, 4ull * m, "traversal rows")) ||
      (rc = dev_alloc(c, (void**)&v1, 4ull * m, "traversal rows")))
    return fail(rc);
  if ((rc = hip_check(c, hipMemcpyAsync(v0, adj, 4ull * m, hipMemcpyDeviceToDevice, c->stream), "order copy")))
    return fail(rc);
  LAUNCH(k_order_keys, grid_for((int64_t)c->n * 64, 4096), c->stream, c->n, begin, adj, deg, k0);
  int bits = 32;
  while (bits < 64 && (1ull << (bits - 32)) < (unsigned long long)c->n) bits++;
  hipcub::DoubleBuffer<unsigned long long> keys(k0, k1);
  hipcub::DoubleBuffer<int32_t> vals(v0, v1);
  size_t tb = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, vals, (int)m, 0, bits, c->stream);
  if ((rc = dev_alloc(c, &tmp, tb + 256, "order temp"))) return fail(rc);
  if ((rc = hip_check(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, vals, (int)m, 0, bits, c->stream),
                      "order sort")))
    return fail(rc);
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "order sync"))) return fail(rc);
  int32_t* res = vals.Current();
  if (res == v0) v0 = nullptr;
  else v1 = nullptr;
  *out = res;
  return fail(SCC_OK);
}

}  // namespace

void drop_order(scc_ctx* c) {
  dev_free(c, c->tadj);
  dev_free(c, c->tradj);
  dev_free(c, c->head_out);
  dev_free(c, c->head_in);
  c->tadj = c->tradj = nullptr;
  c->head_out = c->head_in = nullptr;
}

int build_heads(scc_ctx* c) {
  if (c->n <= 0 || c->end || c->no_alloc) return SCC_OK;
  dev_free(c, c->head_out);
  dev_free(c, c->head_in);
  c->head_out = c->head_in = nullptr;
  int rc;
  if ((rc = dev_alloc(c, (void**)&c->head_out, 4ull * c->n, "row heads")) ||
      (rc = dev_alloc(c, (void**)&c->head_in, 4ull * c->n, "row heads"))) {
    drop_order(c);
    return rc;
  }
  LAUNCH(k_row_heads, grid_for(c->n, 4096), c->stream, c->n, c->begin, row_end(c), tr_adj(c), c->head_out);
  LAUNCH(k_row_heads, grid_for(c->n, 4096), c->stream, c->n, c->rbegin, rrow_end(c), tr_radj(c), c->head_in);
  return hip_check(c, hipGetLastError(), "row heads");
}

int build_order(scc_ctx* c) {
  drop_order(c);
  if (c->m <= 0 || c->n <= 0 || c->end) return SCC_OK;
  int rc;
  uint32_t* deg = nullptr;
  if ((rc = dev_alloc(c, (void**)&deg, 4ull * c->n, "degrees"))) return rc;
  LAUNCH(k_degsum, grid_for(c->n, 4096), c->stream, c->n, c->begin, c->rbegin, deg);
  rc = order_dir(c, c->begin, c->adj, deg, &c->tadj);
  if (!rc) rc = order_dir(c, c->rbegin, c->radj, deg, &c->tradj);
  dev_free(c, deg);
  if (rc) drop_order(c);
  return rc;
}

}  // namespace sccmi

// Batched graphs and launch-shape autotuning.
//
// Batched: many independent graphs are decomposed by the same launches.  The
// caller uploads their disjoint union (graph g owns vertex ids
// [off[g], off[g+1]) and no edge leaves its graph); every graph starts as its
// own partition (colour = its first vertex id), so the first forward-backward
// step takes one pivot per graph and every later phase -- trims, WCC
// colouring, partition rounds -- serves all graphs in each launch.  SCCs never
// cross graphs, so the union's canonical labels are the graphs' canonical
// labels shifted by off[g].  (No reference counterpart: the reference
// decomposes one graph per process.)
//
// Autotuning: coordinate descent over the launch-shape knobs (lane/wave split
// of row scans, split threshold of huge rows, workgroups per sweep, the two
// direction-switch ratios of the reachability searches), timing the whole
// decomposition of the uploaded graph (or batch) for each candidate.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

__device__ __forceinline__ int32_t graph_of(const int32_t* __restrict__ off, int32_t k, int32_t v) {
  int32_t lo = 0, hi = k;  // off[lo] <= v < off[hi]
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (off[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Edges whose endpoints lie in different graphs (must be none).
__global__ __launch_bounds__(BLOCK) void k_cross_edges(int32_t n, const int32_t* __restrict__ begin,
                                                       const int32_t* __restrict__ end,
                                                       const int32_t* __restrict__ adj,
                                                       const int32_t* __restrict__ off, int32_t k,
                                                       unsigned long long* __restrict__ bad) {
  unsigned long long cnt = 0;
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t g = graph_of(off, k, (int32_t)v);
    const int32_t lo = off[g], hi = off[g + 1];
    for (int32_t i = begin[v]; i < end[v]; i++) {
      const int32_t u = adj[i];
      cnt += (u < lo || u >= hi);
    }
  }
  block_count(cnt, bad);
}

// Colour of every vertex = first vertex of its graph.
__global__ __launch_bounds__(BLOCK) void k_batch_colors(int32_t n, const int32_t* __restrict__ off, int32_t k,
                                                        int32_t* __restrict__ color) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK)
    color[v] = off[graph_of(off, k, (int32_t)v)];
}

// SCCs per graph: canonical roots counted into the graph's slot.
__global__ __launch_bounds__(BLOCK) void k_batch_counts(int32_t n, const int32_t* __restrict__ canon,
                                                        const int32_t* __restrict__ off, int32_t k,
                                                        unsigned long long* __restrict__ cnt) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK)
    if (canon[v] == (int32_t)v) atomicAdd(&cnt[graph_of(off, k, (int32_t)v)], 1ull);
}

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)


int push_tuning(scc_ctx* c) {
  int rc = hip_check(c, hipMemcpyAsync(&c->dc->tune, &c->tune, sizeof(Tune), hipMemcpyHostToDevice, c->stream),
                     "tuning upload");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "tuning sync");
}

bool tuning_ok(const scc_tuning* t) {
  const int l = t->light_row;
  return (l == 4 || l == 8 || l == 16 || l == 32 || l == 64) && t->huge_row >= 256 && t->huge_row <= (1 << 30) &&
         t->grid_cap >= 16 && t->grid_cap <= 65536 && t->bu_alpha > 0.f && t->bu_alpha <= 1e6f &&
         t->td_beta > 0.f && t->td_beta <= 1e6f && t->wcc_dense >= 0 && t->wcc_dense <= 64;
}

void to_api(const Tune& a, scc_tuning* t) {
  std::memset(t, 0, sizeof(*t));
  t->light_row = a.light;
  t->huge_row = a.huge_row;
  t->grid_cap = a.grid_cap;
  t->bu_alpha = a.bu_alpha;
  t->td_beta = a.td_beta;
  t->wcc_dense = a.wcc_dense;
}

Tune from_api(const scc_tuning* t) {
  Tune a;
  a.light = t->light_row;
  a.huge_row = t->huge_row;
  a.grid_cap = t->grid_cap;
  a.bu_alpha = t->bu_alpha;
  a.td_beta = t->td_beta;
  a.wcc_dense = t->wcc_dense > 0 ? t->wcc_dense : WCC_DENSE;
  return a;
}

}  // namespace

// Fresh state with the batch colours (or the plain reset without a batch).
int start_run(scc_ctx* c) {
  int rc;
  if ((rc = reset_state(c))) return rc;
  if (c->batch_k <= 0) return SCC_OK;
  LAUNCH(k_batch_colors, grid_for(c->n, 4096), c->stream, c->n, c->d_boff, c->batch_k, c->color);
  c->multi_colors = true;
  c->colors_split = true;
  c->fresh = false;
  // every vertex is live and no edge crosses graphs (scc_batch_set checks),
  // so the first trim sweep needs only the row lengths
  c->rows_decide = true;
  return hip_check(c, hipGetLastError(), "batch colours");
}

}  // namespace sccmi

using namespace sccmi;

extern "C" {

int scc_batch_set(scc_ctx* c, int32_t num_graphs, const int32_t* offsets) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  dev_free(c, c->d_boff);
  c->d_boff = nullptr;
  c->batch_k = 0;
  c->h_boff.clear();
  if (num_graphs == 0) return SCC_OK;  // back to one graph
  if (num_graphs < 0 || !offsets) return set_error(c, SCC_ERR_ARG, "bad batch");
  if (offsets[0] != 0 || offsets[num_graphs] != c->n)
    return set_error(c, SCC_ERR_ARG, "batch offsets must run from 0 to n");
  for (int32_t g = 0; g < num_graphs; g++)
    if (offsets[g + 1] <= offsets[g])
      return set_error(c, SCC_ERR_ARG, "batch offsets must increase (empty graph at " + std::to_string(g) + ")");
  int rc;
  if ((rc = dev_alloc(c, (void**)&c->d_boff, 4ull * ((uint64_t)num_graphs + 1), "batch offsets"))) return rc;
  if ((rc = hip_check(c, hipMemcpyAsync(c->d_boff, offsets, 4ull * ((uint64_t)num_graphs + 1),
                                        hipMemcpyHostToDevice, c->stream), "batch offsets")))
    return rc;
  // the degree layout becomes a per-graph degree order, so that every graph
  // keeps its id range inside the context as well
  if (c->relabeled && (rc = layout_segment(c, num_graphs))) {
    dev_free(c, c->d_boff);
    c->d_boff = nullptr;
    if (!c->relabeled) free_graph(c);  // a failed move leaves no usable graph
    return rc;
  }
  // no edge may leave its graph (either direction)
  if ((rc = zero_ctr(c))) return rc;
  LAUNCH(k_cross_edges, grid_for(c->n, 4096), c->stream, c->n, c->begin, row_end(c), c->adj, c->d_boff,
         num_graphs, c->ctr + C_CHANGED);
  LAUNCH(k_cross_edges, grid_for(c->n, 4096), c->stream, c->n, c->rbegin, rrow_end(c), c->radj, c->d_boff,
         num_graphs, c->ctr + C_CHANGED);
  if ((rc = pull_ctl(c))) return rc;
  if (c->h_ctr[C_CHANGED] != 0) {
    dev_free(c, c->d_boff);
    c->d_boff = nullptr;
    return set_error(c, SCC_ERR_ARG, std::to_string(c->h_ctr[C_CHANGED]) + " edges cross between batch graphs");
  }
  c->batch_k = num_graphs;
  c->h_boff.assign(offsets, offsets + num_graphs + 1);
  return SCC_OK;
}

int scc_batch_run(scc_ctx* c, const scc_run_opts* opts, scc_run_stats* st) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  int rc;
  if ((rc = start_run(c))) return rc;
  return scc_run_method2(c, opts, st);
}

int scc_batch_counts(scc_ctx* c, int64_t* counts) {
  if (!c || !counts) return SCC_ERR_ARG;
  if (c->batch_k <= 0) return set_error(c, SCC_ERR_STATE, "no batch set");
  (void)hipSetDevice(c->device);
  int rc;
  const int32_t* d_canon = nullptr;
  if ((rc = scc_labels_device(c, &d_canon))) return rc;
  unsigned long long* d_cnt = nullptr;
  if ((rc = dev_alloc(c, (void**)&d_cnt, 8ull * c->batch_k, "batch counts"))) return rc;
  std::vector<unsigned long long> h(c->batch_k);
  rc = hip_check(c, hipMemsetAsync(d_cnt, 0, 8ull * c->batch_k, c->stream), "batch counts");
  if (!rc) {
    LAUNCH(k_batch_counts, grid_for(c->n, 4096), c->stream, c->n, d_canon, c->d_boff, c->batch_k, d_cnt);
    rc = hip_check(c, hipMemcpyAsync(h.data(), d_cnt, 8ull * c->batch_k, hipMemcpyDeviceToHost, c->stream),
                   "batch counts");
  }
  if (!rc) rc = hip_check(c, hipStreamSynchronize(c->stream), "batch counts sync");
  dev_free(c, d_cnt);
  if (rc) return rc;
  for (int32_t g = 0; g < c->batch_k; g++) counts[g] = (int64_t)h[g];
  return SCC_OK;
}

int scc_ctx_get_tuning(scc_ctx* c, scc_tuning* t) {
  if (!c || !t) return SCC_ERR_ARG;
  to_api(c->tune, t);
  return SCC_OK;
}

int scc_ctx_set_tuning(scc_ctx* c, const scc_tuning* t) {
  if (!c) return SCC_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (!t) {
    c->tune = Tune{LEVEL_LIGHT, HUGE_ROW, 24.f, 14.f, GRID_CAP, WCC_DENSE};
    return push_tuning(c);
  }
  if (!tuning_ok(t)) return set_error(c, SCC_ERR_ARG, "tuning value out of range");
  c->tune = from_api(t);
  return push_tuning(c);
}

int scc_autotune(scc_ctx* c, const scc_run_opts* opts, int32_t reps, scc_tuning* best_out, double* best_ms,
                 double* default_ms) {
  if (!c || reps < 1) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  int rc;
  auto run_ms = [&](double* ms) -> int {
    std::vector<double> ts;
    for (int r = 0; r <= reps; r++) {  // first run warms up
      int e;
      if ((e = start_run(c))) return e;
      if ((e = hip_check(c, hipStreamSynchronize(c->stream), "tune sync"))) return e;
      auto t0 = std::chrono::steady_clock::now();
      if ((e = scc_run_method2(c, opts, nullptr))) return e;
      if ((e = hip_check(c, hipStreamSynchronize(c->stream), "tune sync"))) return e;
      if (r > 0) ts.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(ts.begin(), ts.end());
    *ms = ts[ts.size() / 2];
    return SCC_OK;
  };
  // (from the context's defaults)
  Tune best{LEVEL_LIGHT, HUGE_ROW, 24.f, 14.f, GRID_CAP, WCC_DENSE};
  c->tune = best;
  if ((rc = push_tuning(c))) return rc;
  double bms = 0;
  if ((rc = run_ms(&bms))) return rc;
  if (default_ms) *default_ms = bms;
  const int lights[] = {4, 8, 16, 32, 64};
  const int huges[] = {2048, 4096, 8192, 16384, 65536};
  const int grids[] = {512, 1024, 2048, 4096};
  const float alphas[] = {4.f, 8.f, 14.f, 24.f, 48.f};
  const float betas[] = {8.f, 16.f, 24.f, 48.f, 96.f};
  const int denses[] = {2, 3, 4};
  for (int pass = 0; pass < 2; pass++) {
    bool improved = false;
    for (int knob = 0; knob < 6; knob++) {
      const int nc = knob == 2 ? 4 : (knob == 5 ? 3 : 5);
      for (int i = 0; i < nc; i++) {
        Tune t = best;
        switch (knob) {
          case 0: t.light = lights[i]; break;
          case 1: t.huge_row = huges[i]; break;
          case 2: t.grid_cap = grids[i]; break;
          case 3: t.bu_alpha = alphas[i]; break;
          case 4: t.td_beta = betas[i]; break;
          default: t.wcc_dense = denses[i]; break;
        }
        if (std::memcmp(&t, &best, sizeof(Tune)) == 0) continue;
        c->tune = t;
        if ((rc = push_tuning(c))) return rc;
        double ms = 0;
        if ((rc = run_ms(&ms))) return rc;
        if (ms < 0.98 * bms) {  // keep a change only when it gains over noise
          bms = ms;
          best = t;
          improved = true;
        }
      }
    }
    if (!improved) break;
  }
  c->tune = best;
  if ((rc = push_tuning(c))) return rc;
  if (best_out) to_api(best, best_out);
  if (best_ms) *best_ms = bms;
  // leave the labels of a final run in place
  if ((rc = start_run(c))) return rc;
  return scc_run_method2(c, opts, nullptr);
}

}  // extern "C"

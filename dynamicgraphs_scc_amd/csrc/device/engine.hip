// libsccmi engine: context, device memory, graph upload, phase drivers and the
// method-2 pipeline behind the C ABI of include/sccmi.h.
//
// Data layout in HBM (per context, n vertices, m edges, all int32):
//   begin[n+1] adj[m] rbegin[n+1] radj[m]            graph (CSR + reverse CSR)
//   tadj[m] tradj[m]                                 traversal copies, hubs first (order.hip)
//   end[n] rend[n]                                   row ends once updates made rows gapped
//   color[n] label[n] canon[n] W[n]                  per-vertex state
//   cw[n] (u64)                                      WCC words: colour << 32 | label
//   live[2][n] q[2][n] cand[2][n] huge[2][n]         ping-pong vertex lists
//   bq[2][n] bcand[2][n] bhuge[2][n]                 the backward search's lists
//   vf[n/32] vb[n/32]                                forward / backward reach bitmaps (adjacent)
//   pkey[n+1] (u64), rep[3(n+1)], scratch[n+1]       per-partition slots
//   DevCtl                                           device-resident loop control
// A partition is identified by its colour; colours are -1 (the initial
// partition, slot n) or the id of a member vertex (slot = that id), so the
// per-partition arrays need no allocator.
//
// Iterative phases (trim sweeps, search levels, WCC iterations) are queued in
// batches of launches; each launch's last workgroup decides the next step on
// the device, and the host synchronises once per batch to see whether the
// phase has finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <string>

#include "common.h"
#include "engine.h"
#include "kernels.h"

using namespace sccmi;

namespace sccmi {

int set_error(scc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_check(scc_ctx* c, hipError_t e, const char* what) {
  if (c->fault_after == 0) {
    c->lost = true;
    return set_error(c, SCC_ERR_DEVICE_LOST, std::string("device lost (injected) at ") + what);
  }
  if (c->fault_after > 0) c->fault_after--;
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
      return set_error(c, SCC_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e));
    return set_error(c, SCC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  }
  return SCC_OK;
}

int dev_alloc(scc_ctx* c, void** p, uint64_t bytes, const char* what) {
  *p = nullptr;
  if (c->no_alloc)
    return set_error(c, SCC_ERR_STATE, std::string("allocation of ") + what + " in a call that allocates nothing");
  if (bytes == 0) bytes = 256;
  if (c->budget && c->bytes + bytes > c->budget)
    return set_error(c, SCC_ERR_BUDGET,
                     std::string("HBM budget exceeded allocating ") + what + " (" +
                         std::to_string(bytes) + " B, held " + std::to_string(c->bytes) +
                         " B, budget " + std::to_string(c->budget) + " B)");
  int rc = hip_check(c, hipMalloc(p, bytes), what);
  if (rc) {
    *p = nullptr;
    return rc;
  }
  c->bytes += bytes;
  c->allocs.push_back(*p);
  c->alloc_sizes.push_back(bytes);
  return SCC_OK;
}

void dev_free(scc_ctx* c, void* p) {
  if (!p) return;
  for (size_t i = 0; i < c->allocs.size(); i++) {
    if (c->allocs[i] == p) {
      (void)hipFree(p);
      c->bytes -= c->alloc_sizes[i];
      c->allocs.erase(c->allocs.begin() + i);
      c->alloc_sizes.erase(c->alloc_sizes.begin() + i);
      return;
    }
  }
}

GraphPtrs gp(const scc_ctx* c) {
  return GraphPtrs{c->begin, row_end(c), tr_adj(c), c->rbegin, rrow_end(c), tr_radj(c), c->n, c->maxdeg,
                   c->relabeled ? c->perm : nullptr, c->ckey};
}
CanonKeys ck(const scc_ctx* c) { return CanonKeys{c->relabeled ? c->perm : nullptr, c->ckey}; }

// The rows the partition steps follow, per direction: the cut rows once a
// fixpoint has cut that direction, else the graph's (for searches, traversal
// copies where they exist).
struct PartRows {
  const int32_t *b, *e, *a, *rb, *re, *ra;
};
static PartRows part_rows(const scc_ctx* c, bool traversal) {
  PartRows R;
  R.b = c->begin;
  R.rb = c->rbegin;
  R.e = c->cut_out ? c->fend : row_end(c);
  R.a = c->cut_out ? c->fadj : (traversal ? tr_adj(c) : c->adj);
  R.re = c->cut_in ? c->frend : rrow_end(c);
  R.ra = c->cut_in ? c->fradj : (traversal ? tr_radj(c) : c->radj);
  return R;
}

static void drop_cut_rows(scc_ctx* c) {
  int32_t** arrs[] = {&c->fadj, &c->fradj, &c->fend, &c->frend};
  for (auto p : arrs) {
    dev_free(c, *p);
    *p = nullptr;
  }
  c->f_cap = 0;
  c->f_n = 0;
  c->cut_out = c->cut_in = false;
}

// The cut rows' buffers (8 (m + n) bytes), allocated on first use and kept
// while they are large enough.  False without room in the budget or on the
// device, and in a call that allocates nothing: the steps then follow the
// graph's rows.  Also false while fewer than 1/8 of the vertices are live (an
// update's affected region): the first cut of a direction starts by clearing
// every row end, which would cost more than the region's sweeps save.
static bool cut_buffers(scc_ctx* c) {
  if (c->no_alloc || 8 * c->n_live < (int64_t)c->n) return false;
  if (c->fadj && c->f_cap >= c->m_cap && c->f_n >= c->n) return true;
  drop_cut_rows(c);
  const uint64_t cap = (uint64_t)std::max<int64_t>(c->m_cap, 1);
  if (c->budget && c->bytes + 8ull * cap + 8ull * (uint64_t)c->n + 1024 > c->budget) return false;
  if (dev_alloc(c, (void**)&c->fadj, 4ull * cap, "cut rows") ||
      dev_alloc(c, (void**)&c->fradj, 4ull * cap, "cut rows") ||
      dev_alloc(c, (void**)&c->fend, 4ull * c->n, "cut row ends") ||
      dev_alloc(c, (void**)&c->frend, 4ull * c->n, "cut row ends")) {
    drop_cut_rows(c);
    (void)hipGetLastError();
    c->err.clear();
    return false;
  }
  c->f_cap = c->m_cap;
  c->f_n = c->n;
  return true;
}

// Copy the control block to the host and refresh the host copies of the live list.
// pull_ctl in two halves, so that launches can be queued behind the copy
// before the host waits for it (launches that only read the control block's
// end state, or return at once once the loop it records is done).  A context
// without a readback event (the no-alloc decomposition creates none) waits for
// the whole stream instead, which includes the copy.
int ctl_readback_start(scc_ctx* c) {
  if (int rc = hip_check(c, hipMemcpyAsync(c->h_dc, c->dc, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream),
                         "control readback"))
    return rc;
  if (!c->ev_rb) return SCC_OK;
  return hip_check(c, hipEventRecord(c->ev_rb, c->stream), "control readback event");
}
int ctl_readback_wait(scc_ctx* c) {
  int rc;
  if ((rc = hip_check(c, c->ev_rb ? hipEventSynchronize(c->ev_rb) : hipStreamSynchronize(c->stream),
                      "control readback wait")))
    return rc;
  if ((rc = hip_check(c, hipGetLastError(), "kernel launch"))) return rc;
  c->cur = c->h_dc->live.sel;
  c->n_live = (int64_t)c->h_dc->live.n;
  return SCC_OK;
}

int pull_ctl(scc_ctx* c) {
  int rc = hip_check(c, hipMemcpyAsync(c->h_dc, c->dc, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream),
                     "control readback");
  if (rc) return rc;
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
  if ((rc = hip_check(c, hipGetLastError(), "kernel launch"))) return rc;
  c->cur = c->h_dc->live.sel;
  c->n_live = (int64_t)c->h_dc->live.n;
  return SCC_OK;
}


// One phase-trace record: construct at phase entry, finish() after the
// phase's final synchronisation.
struct PhaseRec {
  scc_ctx* c;
  int kind;
  int64_t before;
  std::chrono::steady_clock::time_point t0;
  PhaseRec(scc_ctx* cc, int k) : c(cc), kind(k), before(cc->n_live), t0(std::chrono::steady_clock::now()) {}
  void finish(int64_t aux) {
    if (c->no_alloc || c->trace.size() >= 65536) return;
    scc_phase_rec r;
    r.kind = kind;
    r.round = c->round_idx;
    r.live_before = before;
    r.live_after = c->n_live;
    r.aux = aux;
    r.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->trace.push_back(r);
  }
};

int zero_ctr(scc_ctx* c) {
  launch_ctl_init(c->stream, c->dc, 1, 0, 0, 0, 0.0);
  return hip_check(c, hipGetLastError(), "counter reset");
}

int set_live(scc_ctx* c, int sel, int64_t n) {
  launch_ctl_init(c->stream, c->dc, 2, sel, (unsigned long long)n, 0, 0.0);
  c->cur = sel;
  c->n_live = n;
  return hip_check(c, hipGetLastError(), "live set");
}

int check_cancel(scc_ctx* c) {
  if (c->cancel.load()) {
    c->cancel.store(0);
    return set_error(c, SCC_ERR_CANCELLED, "cancelled");
  }
  if (c->lost) return set_error(c, SCC_ERR_DEVICE_LOST, "device lost");
  return SCC_OK;
}

uint64_t state_bytes(int32_t n) {
  const uint64_t N = (uint64_t)n + 1;
  const uint64_t words = ((uint64_t)n + 31) / 32 + 1;
  return 4 * N * 4            // color label canon W
         + 8 * N * 4          // live[2] q[2] cand[2] huge[2]
         + 6 * N * 4          // bq[2] bcand[2] bhuge[2]
         + 2 * words * 4      // vf vb
         + 8 * N + 12 * N     // pkey rep
         + 8 * N              // ckey
         + 4 * N              // scratch
         + 8 * N;             // packed WCC words
}
// Traversal copies (8 m) and, while they are built, sort keys and temp (~20 m).
uint64_t order_bytes(int64_t m) {
  return 28ull * (uint64_t)(m > 0 ? m : 1) + (1ull << 20);
}

int free_graph(scc_ctx* c) {
  dynamic_graph_replaced(c);
  dev_free(c, c->cw);
  c->cw = nullptr;
  drop_order(c);
  drop_layout(c);
  dev_free(c, c->d_boff);
  c->d_boff = nullptr;
  c->batch_k = 0;
  c->h_boff.clear();
  int32_t** arrs[] = {&c->begin, &c->adj, &c->rbegin, &c->radj, &c->end, &c->rend, &c->color, &c->label, &c->canon,
                      &c->W, &c->live[0], &c->live[1], &c->q[0], &c->q[1], &c->cand[0],
                      &c->cand[1], &c->huge, &c->huge2, &c->bq[0], &c->bq[1], &c->bcand[0], &c->bcand[1],
                      &c->bhuge[0], &c->bhuge[1], &c->rep, &c->scratch_i32};
  for (auto p : arrs) {
    dev_free(c, *p);
    *p = nullptr;
  }
  drop_cut_rows(c);
  dev_free(c, c->vf);  // vb shares this allocation
  dev_free(c, c->pkey);
  dev_free(c, c->ckey);
  c->vf = c->vb = nullptr;
  c->pkey = nullptr;
  c->ckey = nullptr;
  c->n = 0;
  c->m = 0;
  c->m_cap = 0;
  c->n_live = 0;
  c->labels_ready = false;
  c->graph_broken = false;
  return SCC_OK;
}

int alloc_graph(scc_ctx* c, int32_t n, int64_t m_cap) {
  free_graph(c);
  const uint64_t N = (uint64_t)n + 1;
  const uint64_t words = ((uint64_t)n + 31) / 32 + 1;
  int rc = 0;
#define ALLOC(ptr, bytes) \
  if ((rc = dev_alloc(c, (void**)&(ptr), (bytes), #ptr))) return (free_graph(c), rc)
  ALLOC(c->begin, 4 * N);
  ALLOC(c->rbegin, 4 * N);
  ALLOC(c->adj, 4 * (uint64_t)std::max<int64_t>(m_cap, 1));
  ALLOC(c->radj, 4 * (uint64_t)std::max<int64_t>(m_cap, 1));
  ALLOC(c->color, 4 * N);
  ALLOC(c->label, 4 * N);
  ALLOC(c->canon, 4 * N);
  ALLOC(c->W, 4 * N);
  ALLOC(c->live[0], 4 * N);
  ALLOC(c->live[1], 4 * N);
  ALLOC(c->q[0], 4 * N);
  ALLOC(c->q[1], 4 * N);
  ALLOC(c->cand[0], 4 * N);
  ALLOC(c->cand[1], 4 * N);
  ALLOC(c->huge, 4 * N);
  ALLOC(c->huge2, 4 * N);
  ALLOC(c->bq[0], 4 * N);
  ALLOC(c->bq[1], 4 * N);
  ALLOC(c->bcand[0], 4 * N);
  ALLOC(c->bcand[1], 4 * N);
  ALLOC(c->bhuge[0], 4 * N);
  ALLOC(c->bhuge[1], 4 * N);
  ALLOC(c->vf, 8 * words);  // vf then vb, adjacent so that one fill clears both
  c->vb = c->vf + ((uint64_t)n + 31) / 32;
  ALLOC(c->pkey, 8 * N);
  ALLOC(c->rep, 12 * N);
  ALLOC(c->scratch_i32, 4 * N);
  ALLOC(c->cw, 8 * N);
  ALLOC(c->ckey, 8 * N);
#undef ALLOC
  c->n = n;
  c->m_cap = m_cap;
  return SCC_OK;
}

int require_graph(scc_ctx* c, bool materialize) {
  if (!c->begin) return set_error(c, SCC_ERR_STATE, "no graph uploaded");
  if (c->graph_broken)
    return set_error(c, SCC_ERR_STATE, "an update failed part-way; the graph must be uploaded again");
  (void)hipSetDevice(c->device);
  c->pivot_ready = false;
  if (materialize && c->reset_pending) {
    c->reset_pending = false;
    launch_reset(c->stream, c->dc, c->n, c->color, c->label, c->live[0]);
    return hip_check(c, hipGetLastError(), "reset");
  }
  return SCC_OK;
}

// The host side of a reset; the launch is left to the next call.
static void reset_host_state(scc_ctx* c) {
  c->labels_ready = false;
  c->multi_colors = false;
  c->colors_split = false;
  c->classes = false;
  c->trace.clear();
  c->round_idx = -1;
  c->fresh = true;
  c->rows_decide = false;
  c->cur = 0;
  c->n_live = c->n;
  c->pivot_ready = false;
  c->canon_stale = false;
  c->last_color_round = false;
  c->w_fresh = false;
  c->cut_out = c->cut_in = false;
  c->fix_idx = 0;
  c->search_idx = 0;
}

int reset_state(scc_ctx* c) {
  c->reset_pending = false;
  launch_reset(c->stream, c->dc, c->n, c->color, c->label, c->live[0]);
  reset_host_state(c);
  return hip_check(c, hipGetLastError(), "reset");
}

// ---- phases -----------------------------------------------------------------------

int do_trim_compact(scc_ctx* c) {
  c->fresh = false;
  c->rows_decide = false;
  if (c->n_live == 0) return SCC_OK;
  int rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_compact(c->stream, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->dc, c->color, c->live[0], c->live[1]);
  return pull_ctl(c);
}

int force_pivot(scc_ctx* c, int32_t pivot) {
  launch_force_pivot(c->stream, pivot, c->n, c->color, c->pkey);
  return hip_check(c, hipGetLastError(), "force pivot");
}

// repeat_global_trim1: sweeps until one trims <= stop.
int do_trim1(scc_ctx* c, int64_t stop, int64_t* trimmed) {
  if (trimmed) *trimmed = 0;
  c->w_fresh = false;
  c->pivot_ready = false;
  if (c->n_live == 0) return SCC_OK;
  PhaseRec pr(c, SCC_PHASE_TRIM1);
  const int32_t epoch = ++c->trim_epoch;
  const unsigned grid = grid_for(c->n_live, (unsigned)c->tune.grid_cap);
  // a deferred reset goes into the first sweep, with the global pivot key
  const bool fused = c->reset_pending;
  c->reset_pending = false;
  const int64_t words = ((int64_t)c->n + 31) / 32;
  // queue as many sweeps as the last loop of this kind needed (a sweep that
  // finds the loop ended returns at once)
  int32_t& hint = c->trim_hint[stop < 0 ? 1 : 0];
  int batch = std::max(1, hint), rc;
  for (int guard = 0;; guard++) {
    if ((rc = check_cancel(c))) return rc;
    for (int i = 0; i < batch; i++) {
      const int fresh = c->fresh ? (fused && guard == 0 && i == 0 ? 2 : 1) : (c->rows_decide ? 1 : 0);
      launch_trim1_step(c->stream, grid, c->dc, gp(c), c->color, c->label, c->live[0], c->live[1], fresh,
                        (long long)stop, epoch, c->pkey, c->vf, 2 * words);
      c->fresh = false;
      c->rows_decide = false;
    }
    if ((rc = pull_ctl(c))) return rc;
    if (c->h_dc->loop.done) break;
    batch = 4;
    if (guard > (int64_t)c->n + 8) return set_error(c, SCC_ERR_LIMIT, "trim did not converge");
  }
  hint = std::min(64, c->h_dc->loop.iters);
  if (trimmed) *trimmed = (int64_t)c->h_dc->loop.total0;
  // the fused sweep's pivot key stands when that sweep ended the loop
  c->pivot_ready = fused && c->h_dc->loop.iters == 1;
  pr.finish(c->h_dc->loop.iters);
  return SCC_OK;
}

int do_trim2(scc_ctx* c, int64_t* removed) {
  c->fresh = false;
  c->rows_decide = false;
  c->w_fresh = false;
  if (removed) *removed = 0;
  if (c->n_live == 0) return SCC_OK;
  PhaseRec pr(c, SCC_PHASE_TRIM2);
  int rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_trim2(c->stream, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->dc, gp(c), c->color, c->label, c->live[0], c->live[1]);
  if ((rc = pull_ctl(c))) return rc;
  if (removed) *removed = (int64_t)c->h_dc->ctr[C_TRIMMED];
  pr.finish((int64_t)c->h_dc->ctr[C_TRIMMED]);
  return SCC_OK;
}

// The arrays of the forward (fwd) or backward search; with own_lists the
// backward search uses its own lists so that it can run alongside the forward one.
static LevelDir level_dir(scc_ctx* c, bool fwd, bool own_lists) {
  LevelDir d;
  const PartRows R = part_rows(c, true);
  d.off = fwd ? R.b : R.rb;
  d.offe = fwd ? R.e : R.re;
  d.nb = fwd ? R.a : R.ra;
  d.roff = fwd ? R.rb : R.b;
  d.roffe = fwd ? R.re : R.e;
  d.rnb = fwd ? R.ra : R.a;
  d.rhead = fwd ? (c->cut_in ? nullptr : c->head_in) : (c->cut_out ? nullptr : c->head_out);
  d.V = fwd ? c->vf : c->vb;
  const bool b = !fwd && own_lists;
  d.q0 = b ? c->bq[0] : c->q[0];
  d.q1 = b ? c->bq[1] : c->q[1];
  d.c0 = b ? c->bcand[0] : c->cand[0];
  d.c1 = b ? c->bcand[1] : c->cand[1];
  d.h0 = b ? c->bhuge[0] : c->huge;
  d.h1 = b ? c->bhuge[1] : c->huge2;
  return d;
}

// Queue level launches in batches until the search (both == false: the one of
// dc->lv; both: the two of dc->lv and dc->lv2) has finished.  hint holds the
// launches the last such search needed.  post, when given, queues kernels
// behind every batch that do nothing until the searches have ended, so the
// batch that ends them also runs what follows them before the host reads back.
static int level_loop(scc_ctx* c, bool both, bool fwd, bool multi, int64_t* levels, bool timed,
                      const std::function<void()>& post = nullptr) {
  const unsigned grid = grid_for(c->n_live, (unsigned)c->tune.grid_cap);
  const int32_t max_chunks = std::max<int32_t>(1, (c->maxdeg + 1023) / 1024);
  const LevelDir F = level_dir(c, fwd || both, both), B = level_dir(c, false, true);
  int& hint = c->lv_hint[both ? 2 : (fwd ? 0 : 1)];
  // batches sized by the search of the same ordinal last time, else by the
  // last search of this kind; outside kernel timing, a few launches are queued
  // behind each control-block copy (they return at once when the search has
  // ended), so the device is not idle while the host reads it back
  const int slot = std::min(c->search_idx++, (int32_t)(sizeof(c->lv_hints) / sizeof(c->lv_hints[0])) - 1);
  int batch = c->lv_hints[slot] > 0 ? c->lv_hints[slot] : std::max(2, hint), rc;
  auto levels_q = [&](int k) {
    for (int i = 0; i < k; i++) {
      if (both)
        launch_level2(c->stream, 2 * grid, c->dc, multi ? 1 : 0, F, B, c->color, c->live[0], c->live[1],
                      max_chunks, c->tune.light);
      else
        launch_level(c->stream, grid, c->dc, multi ? 1 : 0, F, c->color, c->live[0], c->live[1], max_chunks,
                     c->tune.light);
    }
  };
  for (int guard = 0;; guard++) {
    if ((rc = check_cancel(c))) return rc;
    if (timed) (void)hipEventRecord(c->ev[0], c->stream);
    levels_q(batch);
    if (timed) (void)hipEventRecord(c->ev[1], c->stream);
    if (post) post();
    if (timed) {
      if ((rc = pull_ctl(c))) return rc;
    } else {
      if ((rc = ctl_readback_start(c))) return rc;
      levels_q(std::min(batch, 2));
      if ((rc = ctl_readback_wait(c))) return rc;
    }
    if (timed) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
      c->kernel_ms += ms;
      c->kernel_launches += batch;
    }
    if (c->h_dc->lv.done && (!both || c->h_dc->lv2.done)) break;
    batch = std::max(4, batch);
    if (guard > 4 * (int64_t)c->n + 64) return set_error(c, SCC_ERR_LIMIT, "search did not converge");
  }
  const int used = both ? std::max(c->h_dc->lv.levels, c->h_dc->lv2.levels) : c->h_dc->lv.levels;
  if (levels) *levels += c->h_dc->lv.levels + (both ? c->h_dc->lv2.levels : 0);
  c->last_dirs = both ? 2 : (fwd ? 1 : -1);
  // next time, queue one more launch than this search needed (its last level
  // admits nothing)
  hint = std::min(64, used + 1);
  c->lv_hints[slot] = hint;
  return SCC_OK;
}

// Run the level launches of one search whose frontier has been seeded.
int run_levels(scc_ctx* c, bool fwd, bool multi, int64_t* levels, bool timed) {
  return level_loop(c, false, fwd, multi, levels, timed);
}

// The forward and the backward search from the partitions' pivots, advancing
// together (one launch runs a level of each); post as in level_loop.
int traverse_both(scc_ctx* c, bool multi, int64_t* levels, bool timed, const std::function<void()>& post) {
  // (both search controls were set by the step's set-up launch)
  // (a streaming pass with no completion count: more workgroups hide its latency)
  if (multi)
    launch_seed_pivots2(c->stream, grid_for(c->n_live, 2048u), c->dc, c->n, 1, level_dir(c, true, true),
                        level_dir(c, false, true), c->color, c->pkey, c->live[0], c->live[1], ck(c));
  else  // one partition: its pivot is pkey[n]
    launch_seed_single(c->stream, c->dc, c->n, level_dir(c, true, true), level_dir(c, false, true), c->pkey,
                       ck(c));
  return level_loop(c, true, true, multi, levels, timed, post);
}

// Bottom-up entry threshold of a search: frontier edges per candidate (the
// multi-partition scale on searches that serve many partitions).
static double bu_ratio(const scc_ctx* c, bool multi) {
  const double avg = c->n ? (double)c->m / (double)c->n : 0.0;
  return avg / ((double)c->tune.bu_alpha * (multi ? MULTI_BU_SCALE : 1.0));
}

int traverse_from(scc_ctx* c, bool fwd, bool multi, const int32_t* seeds,
                  const unsigned long long* d_count, int64_t* levels) {
  const int32_t* off = fwd ? c->begin : c->rbegin;
  uint32_t* V = fwd ? c->vf : c->vb;
  launch_ctl_init(c->stream, c->dc, 8, 0, 0, 0, bu_ratio(c, multi));
  launch_seed_list(c->stream, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->dc, off, fwd ? row_end(c) : rrow_end(c), c->color, seeds, d_count, V, c->q[0]);
  return run_levels(c, fwd, multi, levels, false);
}

// One forward-backward step for every live partition at once (multi) or for
// the single partition of colour -1 (the global step).
int do_fb_round(scc_ctx* c, int policy, bool multi, int32_t pivot_override, int64_t* scc,
                int32_t* pivot_out, int64_t* levels, bool timed, int64_t* sets) {
  c->fresh = false;
  c->rows_decide = false;
  c->w_fresh = false;
  int rc;
  if (scc) *scc = 0;
  if (c->n_live == 0) return SCC_OK;
  PhaseRec pr(c, multi ? SCC_PHASE_FB_ROUND : SCC_PHASE_FB_GLOBAL);
  int64_t lv = 0;
  if (!levels) levels = &lv;
  const int64_t lv0 = *levels;
  const unsigned grid = grid_for(c->n_live, (unsigned)c->tune.grid_cap);
  // learned pivots need the network (else the plain rule); sched_pivots
  // writes this graph's scores for every live vertex first
  if (policy == SCC_PIVOT_LEARNED && !sched_ready(c)) policy = SCC_PIVOT_DEGREE_PRODUCT;
  if (pivot_override < 0 && policy == SCC_PIVOT_LEARNED && (rc = sched_pivots(c))) return rc;
  // counters, both search controls and (one partition) its slot; several
  // partitions reset their slots from the live list
  // the fused first trim sweep left the global pivot key and clear bitmaps
  const bool ready = !multi && c->pivot_ready && pivot_override < 0 && policy == SCC_PIVOT_DEGREE_PRODUCT;
  c->pivot_ready = false;
  launch_fb_setup(c->stream, c->dc, c->n, c->pkey, c->rep, multi ? 0 : (ready ? 2 : 1), bu_ratio(c, multi));
  if (multi)
    launch_reset_slots(c->stream, grid, c->dc, c->n, c->color, c->pkey, c->rep, c->live[0], c->live[1], 0);
  const int64_t words = ((int64_t)c->n + 31) / 32;
  if (ready) {
    // nothing to launch
  } else if (pivot_override >= 0) {
    if ((rc = force_pivot(c, pivot_override))) return rc;
    launch_fill_u32(c->stream, c->vf, 0u, 2 * words);
  } else {
    // (also clears vf and vb)
    // (a streaming pass with no completion count: more workgroups hide its latency)
    launch_pivot_key(c->stream, grid_for(c->n_live, 2048u), c->dc, gp(c), c->color, policy,
                     policy == SCC_PIVOT_LEARNED ? c->sched_scores : nullptr, c->pkey, c->live[0], c->live[1],
                     c->vf, 2 * words);
  }
  // reach-set sizes (when asked), FW n BW -> done, the rest -> new partitions:
  // queued behind the level batches, they run once both searches have ended
  auto post = [&]() {
    if (sets) launch_fb_stats(c->stream, grid, c->dc, gp(c), c->vf, c->vb, c->live[0], c->live[1]);
    launch_classify(c->stream, grid, c->dc, c->n, c->color, c->label, c->vf, c->vb, c->pkey, c->rep,
                    c->live[0], c->live[1], ck(c), multi ? 0 : 1);
    launch_recolor(c->stream, grid, c->dc, c->n, c->color, c->vf, c->vb, c->rep, c->live[0], c->live[1]);
  };
  if ((rc = traverse_both(c, multi, levels, timed, post))) return rc;
  const unsigned long long* h = c->h_dc->ctr;
  if (scc) *scc = (int64_t)h[C_SCC];
  if (pivot_out) *pivot_out = multi ? -1 : (int32_t)h[C_PIVOT];
  if (sets) {
    sets[0] = (int64_t)h[C_F];
    sets[1] = (int64_t)h[C_B];
    sets[2] = (int64_t)h[C_EF];
    sets[3] = (int64_t)h[C_EB];
    sets[4] = (int64_t)h[C_FB];
  }
  c->colors_split = true;
  if (multi) c->classes = true;
  c->last_color_round = false;
  pr.finish(*levels - lv0);
  return SCC_OK;
}

// The rest of a global step whose searches a queued chain left open (they
// continue from the device state): more level launches, then the
// classification behind them.
static int fb_continue(scc_ctx* c, int64_t* scc, int64_t* levels) {
  c->fresh = false;
  c->rows_decide = false;
  PhaseRec pr(c, SCC_PHASE_FB_GLOBAL);
  const unsigned grid = grid_for(c->n_live, (unsigned)c->tune.grid_cap);
  auto post = [&]() {
    launch_classify(c->stream, grid, c->dc, c->n, c->color, c->label, c->vf, c->vb, c->pkey, c->rep, c->live[0],
                    c->live[1], ck(c), 1);
    launch_recolor(c->stream, grid, c->dc, c->n, c->color, c->vf, c->vb, c->rep, c->live[0], c->live[1]);
  };
  int64_t lv = 0;
  const int rc = level_loop(c, true, true, false, &lv, false, post);
  if (rc) return rc;
  *scc = (int64_t)c->h_dc->ctr[C_SCC];
  *levels = lv;
  c->colors_split = true;
  pr.finish(lv);
  return SCC_OK;
}

// WCC colouring of every live partition (W to its fixpoint, colour = root).
// With seed, the roots also seed the forward search of a colouring round.
// in_dir: the minimum over each vertex's same-colour backward reach instead
// (rows and reverse rows swap roles).
static int wcc_fixpoint(scc_ctx* c, bool seed, bool in_dir = false) {
  c->fresh = false;
  c->rows_decide = false;
  int rc;
  const unsigned grid = grid_for(c->n_live, (unsigned)c->tune.grid_cap);
  const int64_t words = ((int64_t)c->n + 31) / 32;
  // the first sweep pulls over the rows of the fixpoint's direction and writes
  // them cut (from the graph's rows the first time in a decomposition, in
  // place after that); the later iterations follow the cut rows
  int32_t *fe = nullptr, *fout = nullptr;
  bool& cut = in_dir ? c->cut_in : c->cut_out;
  if (cut_buffers(c)) {
    fe = in_dir ? c->frend : c->fend;
    fout = in_dir ? c->fradj : c->fadj;
    // (rows of vertices resolved before the first cut stay empty)
    if (!cut && (rc = hip_check(c, hipMemcpyAsync(fe, in_dir ? c->rbegin : c->begin, 4ull * c->n,
                                                  hipMemcpyDeviceToDevice, c->stream),
                                "cut row ends")))
      return rc;
  }
  const PartRows R0 = part_rows(c, false);
  launch_ctl_init(c->stream, c->dc, 1 | 4 | 8, 0, 0, 0, 0.0);
  launch_fill_u32(c->stream, c->vf, 0u, 2 * words);
  // (tags from the graph's row lengths)
  launch_cw_init(c->stream, c->n, in_dir ? c->rbegin : c->begin, in_dir ? rrow_end(c) : row_end(c), c->color, c->cw);
  launch_wcc_pull(c->stream, grid, c->dc, in_dir ? R0.rb : R0.b, in_dir ? R0.re : R0.e, in_dir ? R0.ra : R0.a,
                  c->cw, c->q[0], c->live[0], c->live[1], fe, fout);
  if (fout) cut = true;
  const PartRows R = part_rows(c, false);
  const int32_t* pb = in_dir ? R.rb : R.b;
  const int32_t* pe = in_dir ? R.re : R.e;
  const int32_t* pa = in_dir ? R.ra : R.a;
  const int32_t* qb = in_dir ? R.b : R.rb;
  const int32_t* qe = in_dir ? R.e : R.re;
  const int32_t* qa = in_dir ? R.a : R.ra;
  // Push launches in batches, sized by what the fixpoint of the same ordinal
  // needed last time (repeated decompositions of a graph repeat their
  // fixpoints).  The control block is copied after each batch, and a few more
  // launches are queued behind the copy before the host waits for it, so the
  // device is not idle while the host reads it back; launches that find the
  // fixpoint done return at once.
  const int slot = std::min(c->fix_idx++, (int32_t)(sizeof(c->it_hints) / sizeof(c->it_hints[0])) - 1);
  int batch = c->it_hints[slot] > 0 ? c->it_hints[slot] : std::max(4, c->it_hint);
  auto pushes = [&](int k) {
    for (int i = 0; i < k; i++)
      launch_wcc_push(c->stream, grid, c->dc, pb, pe, pa, qb, qe, qa, c->cw, c->q[0], c->q[1], c->vf, c->vb,
                      c->live[0], c->live[1]);
  };
  pushes(batch);
  for (int guard = 0;; guard++) {
    if ((rc = check_cancel(c))) return rc;
    if ((rc = ctl_readback_start(c))) return rc;
    pushes(std::min(batch, 4));
    if ((rc = ctl_readback_wait(c))) return rc;
    if (c->h_dc->loop.done) break;
    batch = std::min(2 * batch, 64);
    pushes(batch);
    if (guard > 4 * (int64_t)c->n + 64) return set_error(c, SCC_ERR_LIMIT, "WCC did not converge");
  }
  c->it_hint = std::min(64, c->h_dc->loop.iters + 1);
  c->it_hints[slot] = c->it_hint;
  if (seed) {
    launch_ctl_init(c->stream, c->dc, 8, 0, 0, 0, bu_ratio(c, true));
    launch_fill_u32(c->stream, c->vf, 0u, ((int64_t)c->n + 31) / 32);
  }
  launch_wcc_color(c->stream, grid, c->dc, c->color, c->cw, c->W, seed ? 1 : 0, R.b, R.e, c->vf, c->q[0],
                   c->live[0], c->live[1], ck(c));
  c->colors_split = true;
  return hip_check(c, hipGetLastError(), "wcc colour");
}

int do_wcc(scc_ctx* c, int64_t* parts) {
  if (parts) *parts = 0;
  if (c->n_live == 0) return SCC_OK;
  PhaseRec pr(c, SCC_PHASE_WCC);
  int rc;
  if ((rc = wcc_fixpoint(c, false))) return rc;
  if ((rc = pull_ctl(c))) return rc;
  if (parts) *parts = (int64_t)c->h_dc->ctr[C_ROOTS];
  c->classes = true;
  c->w_fresh = true;  // (cleared by every other phase)
  pr.finish(c->h_dc->loop.iters);
  return SCC_OK;
}

// One colouring round: WCC classes, then each root's forward reach inside its
// class is the root's SCC (see k_extract).
int do_color_round(scc_ctx* c, int64_t* scc) {
  if (scc) *scc = 0;
  if (c->n_live == 0) return SCC_OK;
  PhaseRec pr(c, SCC_PHASE_COLOR_ROUND);
  int rc;
  // Right after another colouring round, each class (what its extraction left)
  // first splits by the minimum over its backward reach inside the class:
  // members of one SCC have the same ancestors there (a path into the SCC
  // continues inside it), so no SCC splits, and this round works on much finer
  // classes (planted 8 M, colouring rounds only: 16 rounds and 49 ms before,
  // 7 rounds and 40 ms with it).  After a forward-backward round, whose
  // forward / backward / rest split already refines the classes, it costs more
  // than it saves, and is skipped.
  if (c->color_pairs && c->last_color_round && (rc = wcc_fixpoint(c, false, true))) return rc;
  // Right after the WCC phase, its labels are this round's (the same minimum
  // over the same classes): only the roots' seeding is left to do.
  if (c->w_fresh) {
    c->w_fresh = false;
    launch_ctl_init(c->stream, c->dc, 1 | 8, 0, 0, 0, bu_ratio(c, true));
    launch_fill_u32(c->stream, c->vf, 0u, ((int64_t)c->n + 31) / 32);
    const PartRows R = part_rows(c, false);
    launch_wcc_color(c->stream, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->dc, c->color, nullptr, c->W, 1,
                     R.b, R.e, c->vf, c->q[0], c->live[0], c->live[1], ck(c));
    if ((rc = hip_check(c, hipGetLastError(), "colouring seeds"))) return rc;
  } else if ((rc = wcc_fixpoint(c, true))) {
    return rc;
  }
  const int64_t iters = c->h_dc->loop.iters;
  if ((rc = run_levels(c, true, true, nullptr, false))) return rc;
  launch_extract(c->stream, grid_for(c->n_live, (unsigned)c->tune.grid_cap), c->dc, c->color, c->label, c->vf,
                 c->live[0], c->live[1], ck(c));
  if ((rc = pull_ctl(c))) return rc;
  if (scc) *scc = (int64_t)c->h_dc->ctr[C_SCC];
  c->last_color_round = true;
  c->classes = true;
  pr.finish(iters);
  return SCC_OK;
}

int do_fb_parts(scc_ctx* c, int policy, int mode, bool trim_between, int64_t* rounds) {
  int64_t r = 0;
  int rc;
  if (mode == SCC_PARTS_LEARNED && !sched_ready(c)) mode = SCC_PARTS_COLOR;
  while (c->n_live > 0) {
    if ((rc = check_cancel(c))) return rc;
    c->round_idx = (int32_t)r;
    if (trim_between && r > 0) {
      if ((rc = do_trim1(c, 0, nullptr))) return rc;
      if (c->n_live == 0) break;
    }
    if (mode == SCC_PARTS_LEARNED) {
      // one step per round, chosen by the scheduler
      int action = 0;
      if ((rc = sched_choose(c, &action))) return rc;
      if (action == 0) {
        if (!c->classes && (rc = do_wcc(c, nullptr))) return rc;
        if ((rc = do_fb_round(c, policy, true, -1, nullptr, nullptr, nullptr, false, nullptr)))
          return rc;
      } else if ((rc = do_color_round(c, nullptr))) {
        return rc;
      }
    } else if (mode == SCC_PARTS_COLOR) {
      if ((rc = do_color_round(c, nullptr))) return rc;
    } else {
      if ((rc = do_fb_round(c, policy, true, -1, nullptr, nullptr, nullptr, false, nullptr))) return rc;
      if (mode == SCC_PARTS_BOTH && c->n_live > 0) {
        if (trim_between && (rc = do_trim1(c, 0, nullptr))) return rc;
        if ((rc = do_color_round(c, nullptr))) return rc;
      }
    }
    r++;
  }
  c->round_idx = -1;
  if (rounds) *rounds = r;
  return SCC_OK;
}

int do_canon(scc_ctx* c, int64_t* count) {
  int rc;
  PhaseRec pr(c, SCC_PHASE_CANON);
  // (in the degree layout the caller's canonical labels in the same pass)
  launch_canon(c->stream, c->dc, c->n, c->label, c->ckey, c->relabeled ? c->perm : nullptr, c->canon,
               c->canon_caller);
  if ((rc = pull_ctl(c))) return rc;
  pr.finish((int64_t)c->h_dc->ctr[C_COUNT]);
  if (count) *count = (int64_t)c->h_dc->ctr[C_COUNT];
  c->labels_ready = true;
  c->canon_stale = false;
  return SCC_OK;
}

int materialize_canon(scc_ctx* c) {
  if (!c->canon_stale) return SCC_OK;
  launch_canon(c->stream, c->dc, c->n, c->label, c->ckey, nullptr, c->canon, nullptr);
  c->canon_stale = false;
  return hip_check(c, hipGetLastError(), "internal canonical labels");
}

// ---- a fresh decomposition queued as one chain -------------------------------------
// From the global step just read back: level launches for the next chain (as
// many as the longer search took: its last launch admitted nothing) and each
// launch's workgroups per direction, from the work the two searches' levels
// had (a top-down frontier or bottom-up candidate list, 256 per workgroup,
// at least 8; the full grid where a level split hub rows over the grid).
// A level that grows on another graph only runs with fewer workgroups.
static void chain_learn_search(scc_ctx* c) {
  const DevCtl& d = *c->h_dc;
  const int levels = std::max(d.lv.levels, d.lv2.levels);
  c->chain_levels = std::min(64, std::max(1, levels));
  auto need = [&](const unsigned long long* lg, int have, int i) -> int64_t {
    if (i >= have) return 0;
    const unsigned long long in = lg[1];
    if ((lg[0] & 1) == MODE_TD) {
      if ((in >> 32) > 0) return INT64_MAX;  // rows split over the grid
      return (int64_t)(in & 0xFFFFFFFFull);
    }
    return (int64_t)in;
  };
  for (int i = 0; i < 64; i++) {
    const int64_t w = std::max(need(d.lvlog[i], d.lv.levels, i), need(d.lvlog2[i], d.lv2.levels, i));
    if (w == INT64_MAX) {
      c->chain_level_grid[i] = 0;
      continue;
    }
    const int64_t g = ((w + 255) / 256 + 7) / 8 * 8;
    c->chain_level_grid[i] = (int32_t)std::max<int64_t>(8, std::min<int64_t>(g, 1 << 20));
  }
}

// Stages of method 2 in order; run_method2 runs them from a given one.
enum Stage : int { ST_TRIM_FIRST = 0, ST_FB, ST_FB_CONTINUE, ST_TRIM_SECOND, ST_TRIM2, ST_TRIM_THIRD, ST_WCC, ST_DONE };

// Queue the trims, the global step, the trims after it and the canonical labels
// behind each other with a gate after each phase (ChainCtl), read the control
// block back once, and return the stage the host continues from: ST_DONE when
// the chain ran to its end (labels final), else the stage after the gate that
// stopped it.  The chain's length and each phase's launches come from the last
// decomposition on this context; a phase that needs more launches stops the
// chain at its gate and continues with host read-backs as usual.
static int run_chain(scc_ctx* c, const scc_run_opts& o, scc_run_stats* st, int* resume) {
  const int shape = c->chain_shape;
  auto grid_at = [&](int gate) {
    const int64_t live = gate < 0 || c->chain_live[gate] < 0 ? c->n : c->chain_live[gate];
    return grid_for(std::max<int64_t>(live, 1), (unsigned)c->tune.grid_cap);
  };
  const int64_t words = ((int64_t)c->n + 31) / 32;
  const long long stop = (long long)o.trim_stop;
  auto trim_loop = [&](int gate, int k) {
    const int32_t epoch = ++c->trim_epoch;
    for (int i = 0; i < c->chain_sweeps[k]; i++) {
      const int fresh = c->fresh ? 2 : 0;
      launch_trim1_step(c->stream, grid_at(gate - 1), c->dc, gp(c), c->color, c->label, c->live[0], c->live[1], fresh,
                        stop, epoch, c->pkey, c->vf, 2 * words);
      c->fresh = false;
      c->rows_decide = false;
    }
    return epoch;
  };
  c->reset_pending = false;
  c->pivot_ready = false;
  auto t0 = std::chrono::steady_clock::now();
  // under the learned phase schedule each phase's device time goes into the
  // phase trace (the scheduler's round features read it): events between the
  // phases, read after the chain's one synchronisation.  Skipped when the last
  // chain on this context ran to its end (no partition step read the trace).
  const bool timed = o.partition_mode == SCC_PARTS_LEARNED && !c->chain_ended;
  auto mark = [&](int i) {
    if (timed) (void)hipEventRecord(c->ev[2 + i], c->stream);
  };
  mark(0);
  // the last gate queued marks the chain's end (and zeroes the counters for
  // the canonicalisation behind it)
  const int lastg = shape >= 3 && o.use_trim2 ? 4 : std::min(shape, 2);
  // gate 0: the first trim loop; its first sweep does the deferred reset and
  // leaves the global pivot key, which stands when it is the loop's only sweep
  const int32_t e0 = trim_loop(0, 0);
  mark(1);
  if (lastg == 0) {
    launch_chain_gate(c->stream, c->dc, 0, GATE_TRIM, e0, 1, 1, 1);
  } else {  // gate 0 with the global forward-backward step's set-up and seeding
    const unsigned grid = grid_at(0);
    const LevelDir F = level_dir(c, true, true), B = level_dir(c, false, true);
    launch_chain_gate_fb(c->stream, c->dc, e0, 0, c->n, c->pkey, c->rep, bu_ratio(c, false), F, B, ck(c));
    const int32_t max_chunks = std::max<int32_t>(1, (c->maxdeg + 1023) / 1024);
    for (int i = 0; i < c->chain_levels; i++) {
      const unsigned gi = c->chain_level_grid[i] > 0 ? std::min<unsigned>(grid, (unsigned)c->chain_level_grid[i]) : grid;
      launch_level2(c->stream, 2 * std::max(8u, gi), c->dc, 0, F, B, c->color, c->live[0], c->live[1], max_chunks,
                    c->tune.light);
    }
    launch_classify(c->stream, grid, c->dc, c->n, c->color, c->label, c->vf, c->vb, c->pkey, c->rep, c->live[0],
                    c->live[1], ck(c), 1);
    // gate 1 (when the chain goes on past it) runs inside the recolouring launch
    launch_recolor(c->stream, grid, c->dc, c->n, c->color, c->vf, c->vb, c->rep, c->live[0], c->live[1],
                   lastg >= 2);
    if (lastg == 1) launch_chain_gate(c->stream, c->dc, 1, GATE_SEARCH, 0, 0, 1, 1);
    mark(2);
  }
  // the last gate, when it follows a trim loop, runs inside the window
  // canonicalisation (which needs no counter reset: its count was zeroed by the
  // step's set-up and nothing since adds to it)
  const bool fold_last = c->relabeled && c->cwin && (lastg == 2 || lastg == 4);
  int32_t e_last = 0;
  if (lastg >= 2) {  // gate 2: trim1 after the step (its counters zeroed for trim2 or the labels)
    const int32_t e2 = trim_loop(2, 1);
    mark(3);
    e_last = e2;
    if (!(fold_last && lastg == 2)) launch_chain_gate(c->stream, c->dc, 2, GATE_TRIM, e2, 0, lastg == 2, 1);
  }
  if (lastg == 4) {  // gates 3 and 4: trim2, then trim1
    launch_trim2(c->stream, grid_at(2), c->dc, gp(c), c->color, c->label, c->live[0], c->live[1]);
    launch_chain_gate(c->stream, c->dc, 3, GATE_TRIM2, 0, 0, 0, 0);
    mark(4);
    const int32_t e4 = trim_loop(4, 2);
    mark(5);
    e_last = e4;
    if (!fold_last) launch_chain_gate(c->stream, c->dc, 4, GATE_TRIM, e4, 0, 1, 1);
  }
  // canonical labels: skipped by a chain that stopped (halt odd); a chain that
  // ran to its end has its halt cleared by this launch.  In the degree layout
  // only the caller-order labels (sequential writes); the internal-order ones
  // follow when an update needs them.
  if (c->relabeled && c->cwin)
    launch_canon_window(c->stream, c->dc, c->n, c->cwin_runs, c->cwin_nwin, c->cwin, c->label, c->ckey, c->perm,
                        c->pkey + c->n, c->canon_caller, fold_last ? lastg : -1, e_last);
  else if (c->relabeled)
    launch_canon_caller(c->stream, c->dc, c->n, c->label, c->ckey, c->iperm, c->canon_caller, false);
  else
    launch_canon(c->stream, c->dc, c->n, c->label, c->ckey, nullptr, c->canon, nullptr, false);
  int rc = pull_ctl(c);
  if (rc) return rc;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const ChainCtl ch = c->h_dc->chain;
  const int last = ch.gate;  // the last gate that ran
  // (the last gate queued always decides: halt 2, which the canonicalisation
  // clears, or odd at a gate that ran)
  if (last < 0 || last >= CHAIN_GATES || (ch.halt != 0 && ch.halt != 2 && (ch.halt & 1) == 0))
    return set_error(c, SCC_ERR_STATE, "queued chain ended without a decision");
  // what the gates saw
  st->trimmed_first = last >= 0 ? (int64_t)ch.count[0] : 0;
  if (last >= 1) {
    st->scc_first = (int64_t)ch.count[1];
    st->pivot_first = (int32_t)ch.pivot;
    st->fb_levels = ch.iters[1];
  }
  st->trimmed_second = (last >= 2 ? (int64_t)ch.count[2] : 0) + (last >= 3 ? (int64_t)ch.count[3] : 0) +
                       (last >= 4 ? (int64_t)ch.count[4] : 0);
  for (int g = 0; g <= last && g < CHAIN_GATES; g++) c->chain_live[g] = (int64_t)ch.live[g];
  // sweeps per trim loop: a loop that ended inside the chain needed iters[g]
  // (the first loop too when its gate failed only because it took more than one
  // sweep); one still open gets twice what it had, since the loop that
  // continues it afterwards counts only its own sweeps
  {
    const int gates[3] = {0, 2, 4};
    for (int k = 0; k < 3 && gates[k] <= last; k++) {
      const int g = gates[k], it = std::max(1, ch.iters[g]);
      const bool open = g == last && !ch.ok[g] && !(g == 0 && c->h_dc->loop.done);
      c->chain_sweeps[k] = std::min(64, open ? 2 * it : it);
    }
  }
  // the phase trace (the chain synchronises once: its wall time goes to the last record)
  static const int kinds[CHAIN_GATES] = {SCC_PHASE_TRIM1, SCC_PHASE_FB_GLOBAL, SCC_PHASE_TRIM1, SCC_PHASE_TRIM2,
                                         SCC_PHASE_TRIM1, 0};
  int64_t before = c->n;
  for (int g = 0; g <= last && g < 5 && c->trace.size() < 65536 && !c->no_alloc; g++) {
    scc_phase_rec r;
    r.kind = kinds[g];
    r.round = -1;
    r.live_before = before;
    r.live_after = (int64_t)ch.live[g];
    r.aux = g == 3 ? (int64_t)ch.count[3] : ch.iters[g];
    r.ms = g == last ? ms : 0.0;
    if (timed) {  // phase g lies between marks g and g + 1
      float e = 0.f;
      if (hipEventElapsedTime(&e, c->ev[2 + g], c->ev[3 + g]) == hipSuccess) r.ms = e;
    }
    c->trace.push_back(r);
    before = r.live_after;
  }
  if ((ch.halt & 1) == 0) {  // ran to its end: every vertex decided, labels final
    if (!c->no_alloc && c->trace.size() < 65536) {
      scc_phase_rec r{SCC_PHASE_CANON, -1, 0, 0, (int64_t)c->h_dc->ctr[C_COUNT], 0.0};
      c->trace.push_back(r);
    }
    st->num_sccs = (int64_t)c->h_dc->ctr[C_COUNT];
    c->chain_ended = true;
    c->labels_ready = true;
    c->canon_stale = c->relabeled;
    c->colors_split = last >= 1;
    // next time, stop where the vertices ran out
    c->chain_shape = last >= 3 ? 3 : last;
    if (last >= 1) chain_learn_search(c);
    *resume = ST_DONE;
    st->chain_stop = 5;
    return hip_check(c, hipGetLastError(), "chain");
  }
  // stopped at gate `last` (halt odd): continue from there
  c->chain_ended = false;
  launch_ctl_init(c->stream, c->dc, 64, 0, 0, 0, 0.0);
  st->chain_stop = last;
  const bool ok = ch.ok[last] != 0;
  const LoopCtl& l = c->h_dc->loop;
  switch (last) {
    case 0:
      // the fused sweep's pivot key stands when it alone ended the loop
      c->pivot_ready = ok;
      *resume = ok || l.done ? ST_FB : ST_TRIM_FIRST;
      if (ok) c->chain_shape = 3;
      break;
    case 1:
      c->colors_split = true;
      *resume = ok ? ST_TRIM_SECOND : ST_FB_CONTINUE;
      if (ok) c->chain_shape = 3;
      break;
    case 2:
      *resume = ok ? ST_TRIM2 : ST_TRIM_SECOND;
      c->colors_split = true;
      if (ok) c->chain_shape = 3;
      break;
    default:  // gate 4 (gate 3 never stops the chain)
      c->colors_split = true;
      *resume = ok ? ST_WCC : ST_TRIM_THIRD;
      break;
  }
  return hip_check(c, hipGetLastError(), "chain");
}

}  // namespace sccmi

// ---------------------------------------------------------------------------------
extern "C" {

int scc_ctx_create(int device, uint64_t hbm_budget_bytes, scc_ctx** out) {
  if (!out) return SCC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SCC_ERR_HIP;
  if (device < 0 || device >= ndev) return SCC_ERR_ARG;
  auto* c = new scc_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  int rc;
  if ((rc = hip_check(c, hipSetDevice(device), "hipSetDevice")) ||
      (rc = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream")) ||
      (rc = dev_alloc(c, (void**)&c->dc, sizeof(DevCtl), "control block")) ||
      (rc = hip_check(c, hipHostMalloc((void**)&c->h_dc, sizeof(DevCtl), 0), "pinned control"))) {
    scc_ctx_destroy(c);
    return rc;
  }
  c->ctr = c->dc->ctr;
  c->h_ctr = c->h_dc->ctr;
  std::memset(c->h_dc, 0, sizeof(DevCtl));
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  (void)hipEventCreateWithFlags(&c->ev_rb, hipEventDisableTiming);
  launch_ctl_init(c->stream, c->dc, 1 | 2 | 4 | 8 | 16 | 64, 0, 0, 0, 0.0);
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "init"))) {
    scc_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return SCC_OK;
}

int scc_ctx_destroy(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_dynamic(c);
  destroy_sched(c);
  free_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  if (c->h_dc) (void)hipHostFree(c->h_dc);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_rb) (void)hipEventDestroy(c->ev_rb);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SCC_OK;
}

const char* scc_last_error(const scc_ctx* c) { return c ? c->err.c_str() : "null context"; }
uint64_t scc_ctx_bytes(const scc_ctx* c) { return c ? c->bytes : 0; }
uint64_t scc_bytes_needed(int32_t n, int64_t m) {
  // graph, state, traversal copies or relabelling temporaries, scratch, the
  // degree layout's maps (perm, iperm, two staging arrays, caller-id labels),
  // row heads, control block
  return 8ull * (uint64_t)(m > 0 ? m : 1) + 8ull * ((uint64_t)n + 1) + state_bytes(n) + order_bytes(m) +
         4ull * ((uint64_t)n + 1) + 20ull * ((uint64_t)n + 1) + 8ull * ((uint64_t)n + 1) + sizeof(DevCtl);
}
int scc_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return SCC_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return SCC_ERR_HIP;
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) return SCC_ERR_HIP;
  *free_bytes = f;
  *total_bytes = t;
  return SCC_OK;
}
int scc_ctx_set_budget(scc_ctx* c, uint64_t hbm_budget_bytes) {
  if (!c) return SCC_ERR_ARG;
  if (hbm_budget_bytes && c->bytes > hbm_budget_bytes)
    return set_error(c, SCC_ERR_BUDGET, "context already holds " + std::to_string(c->bytes) +
                                            " B, more than the new budget");
  c->budget = hbm_budget_bytes;
  return SCC_OK;
}
void scc_ctx_cancel(scc_ctx* c) {
  if (c) c->cancel.store(1);
}
void scc_ctx_inject_fault(scc_ctx* c, int after) {
  if (c) {
    c->fault_after = after;
    if (after < 0) c->lost = false;
  }
}

static int validate_csr(scc_ctx* c, int32_t n, int64_t m, const int32_t* off, const int32_t* nb,
                        const char* name) {
  if (off[0] != 0 || (int64_t)off[n] != m)
    return set_error(c, SCC_ERR_ARG, std::string(name) + ": offsets must start at 0 and end at m");
  for (int32_t v = 0; v < n; v++)
    if (off[v + 1] < off[v])
      return set_error(c, SCC_ERR_ARG, std::string(name) + ": offsets decrease at " +
                                           std::to_string(v));
  for (int64_t i = 0; i < m; i++)
    if ((uint32_t)nb[i] >= (uint32_t)n)
      return set_error(c, SCC_ERR_ARG, std::string(name) + ": vertex id out of range at " +
                                           std::to_string(i));
  return SCC_OK;
}

static int32_t max_degree(int32_t n, const int32_t* a, const int32_t* b) {
  int32_t mx = 0;
  for (int32_t v = 0; v < n; v++) mx = std::max(mx, std::max(a[v + 1] - a[v], b[v + 1] - b[v]));
  return mx;
}

int scc_graph_upload(scc_ctx* c, int32_t n, int64_t m, const int32_t* begin, const int32_t* node_idx,
                     const int32_t* r_begin, const int32_t* r_node_idx) {
  if (!c) return SCC_ERR_ARG;
  if (n < 0 || m < 0 || m >= INT32_MAX || n >= WCC_MASK)
    return set_error(c, SCC_ERR_LIMIT, "graph exceeds int32 limits (n < 2^29, m < 2^31)");
  if (!begin || !r_begin || (m > 0 && (!node_idx || !r_node_idx)))
    return set_error(c, SCC_ERR_ARG, "null graph array");
  int rc;
  if ((rc = validate_csr(c, n, m, begin, node_idx, "begin/node_idx")) ||
      (rc = validate_csr(c, n, m, r_begin, r_node_idx, "r_begin/r_node_idx")))
    return rc;
  (void)hipSetDevice(c->device);
  destroy_dynamic(c);
  if ((rc = alloc_graph(c, n, m))) return rc;
  c->m = m;
  c->maxdeg = max_degree(n, begin, r_begin);
  if ((rc = hip_check(c, hipMemcpyAsync(c->begin, begin, 4ull * (n + 1), hipMemcpyHostToDevice, c->stream), "upload")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->rbegin, r_begin, 4ull * (n + 1), hipMemcpyHostToDevice, c->stream), "upload")) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->adj, node_idx, 4ull * m, hipMemcpyHostToDevice, c->stream), "upload"))) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->radj, r_node_idx, 4ull * m, hipMemcpyHostToDevice, c->stream), "upload"))))
    return rc;
  if ((rc = build_layout(c))) {  // a half-relabelled graph must not stay usable
    free_graph(c);
    return rc;
  }
  if (!c->relabeled && (rc = build_order(c))) return rc;
  if ((rc = build_heads(c))) return rc;
  // the fresh state is deferred like scc_state_reset's: the next call makes it,
  // or scc_run_method2 fuses it into its first sweep
  reset_host_state(c);
  c->reset_pending = true;
  c->ran = false;
  return hip_check(c, hipStreamSynchronize(c->stream), "upload sync");
}

int scc_graph_upload_device(scc_ctx* c, int32_t n, int64_t m, const int32_t* d_begin,
                            const int32_t* d_node_idx, const int32_t* d_r_begin,
                            const int32_t* d_r_node_idx) {
  if (!c) return SCC_ERR_ARG;
  if (n < 0 || m < 0 || m >= INT32_MAX || n >= WCC_MASK)
    return set_error(c, SCC_ERR_LIMIT, "graph exceeds int32 limits");
  int rc;
  (void)hipSetDevice(c->device);
  destroy_dynamic(c);
  if ((rc = alloc_graph(c, n, m))) return rc;
  c->m = m;
  if ((rc = hip_check(c, hipMemcpyAsync(c->begin, d_begin, 4ull * (n + 1), hipMemcpyDeviceToDevice, c->stream), "d2d")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->rbegin, d_r_begin, 4ull * (n + 1), hipMemcpyDeviceToDevice, c->stream), "d2d")) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->adj, d_node_idx, 4ull * m, hipMemcpyDeviceToDevice, c->stream), "d2d"))) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->radj, d_r_node_idx, 4ull * m, hipMemcpyDeviceToDevice, c->stream), "d2d"))))
    return rc;
  {  // degrees for the huge-row chunking
    std::vector<int32_t> b(n + 1), rb(n + 1);
    if ((rc = hip_check(c, hipMemcpyAsync(b.data(), c->begin, 4ull * (n + 1), hipMemcpyDeviceToHost, c->stream), "d2h")) ||
        (rc = hip_check(c, hipMemcpyAsync(rb.data(), c->rbegin, 4ull * (n + 1), hipMemcpyDeviceToHost, c->stream), "d2h")) ||
        (rc = hip_check(c, hipStreamSynchronize(c->stream), "sync")))
      return rc;
    c->maxdeg = max_degree(n, b.data(), rb.data());
  }
  if ((rc = build_layout(c))) {  // a half-relabelled graph must not stay usable
    free_graph(c);
    return rc;
  }
  if (!c->relabeled && (rc = build_order(c))) return rc;
  if ((rc = build_heads(c))) return rc;
  // the fresh state is deferred like scc_state_reset's: the next call makes it,
  // or scc_run_method2 fuses it into its first sweep
  reset_host_state(c);
  c->reset_pending = true;
  c->ran = false;
  return hip_check(c, hipStreamSynchronize(c->stream), "upload sync");
}

int scc_graph_info(const scc_ctx* c, int32_t* n, int64_t* m) {
  if (!c) return SCC_ERR_ARG;
  if (n) *n = c->n;
  if (m) *m = c->m;
  return c->graph_broken ? SCC_ERR_STATE : SCC_OK;
}

int scc_graph_download(scc_ctx* c, int32_t* begin, int32_t* node_idx, int32_t* r_begin,
                       int32_t* r_node_idx) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  int rc = 0;
  if (c->relabeled) {
    // internal rows to host, then each caller row rebuilt and sorted
    std::vector<int32_t> ib(c->n + 1), ia(c->m), rb(c->n + 1), ra(c->m);
    if ((rc = c->end ? download_gapped(c, ib.data(), ia.data(), rb.data(), ra.data())
                     : (hip_check(c, hipMemcpyAsync(ib.data(), c->begin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download") ||
                        hip_check(c, hipMemcpyAsync(rb.data(), c->rbegin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download") ||
                        (c->m && hip_check(c, hipMemcpyAsync(ia.data(), c->adj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download")) ||
                        (c->m && hip_check(c, hipMemcpyAsync(ra.data(), c->radj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download")) ||
                        hip_check(c, hipStreamSynchronize(c->stream), "download sync"))
                           ? SCC_ERR_HIP
                           : SCC_OK))
      return rc;
    layout_rows_to_caller(c, ib, ia, begin, node_idx);
    layout_rows_to_caller(c, rb, ra, r_begin, r_node_idx);
    return SCC_OK;
  }
  if (c->end) return download_gapped(c, begin, node_idx, r_begin, r_node_idx);
  if (begin && (rc = hip_check(c, hipMemcpyAsync(begin, c->begin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (r_begin && (rc = hip_check(c, hipMemcpyAsync(r_begin, c->rbegin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (node_idx && c->m && (rc = hip_check(c, hipMemcpyAsync(node_idx, c->adj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (r_node_idx && c->m && (rc = hip_check(c, hipMemcpyAsync(r_node_idx, c->radj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "download sync");
}

int scc_state_reset(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c, false)) return r0;
  // deferred: the next call performs it, or scc_run_method2 fuses it into its first sweep
  reset_host_state(c);
  c->reset_pending = true;
  return SCC_OK;
}

int scc_state_set_colors(scc_ctx* c, const int32_t* colors) {
  if (!c || !colors) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  for (int32_t v = 0; v < c->n; v++)
    if (colors[v] < -1 || colors[v] >= c->n)
      return set_error(c, SCC_ERR_ARG, "colour out of range at " + std::to_string(v));
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = reset_state(c))) return rc;
  c->multi_colors = false;
  for (int32_t v = 0; v < c->n && !c->multi_colors; v++) c->multi_colors = colors[v] != -1;
  c->colors_split = c->multi_colors;
  c->fresh = false;
  c->rows_decide = false;
  if (c->relabeled) return layout_colors_to_internal(c, colors);
  if ((rc = hip_check(c, hipMemcpyAsync(c->color, colors, 4ull * c->n, hipMemcpyHostToDevice, c->stream),
                      "colours upload")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "colours sync");
}

int scc_ctx_set_layout(scc_ctx* c, int layout) {
  if (!c) return SCC_ERR_ARG;
  if (layout != SCC_LAYOUT_INPUT && layout != SCC_LAYOUT_DEGREE && layout != SCC_LAYOUT_AUTO)
    return set_error(c, SCC_ERR_ARG, "bad layout");
  c->layout_req = layout;
  return SCC_OK;
}

int scc_graph_layout(const scc_ctx* c, int* layout) {
  if (!c || !layout) return SCC_ERR_ARG;
  *layout = c->relabeled ? SCC_LAYOUT_DEGREE : SCC_LAYOUT_INPUT;
  return SCC_OK;
}

#define REQUIRE_GRAPH(c)                                                    \
  do {                                                                      \
    if (!(c)) return SCC_ERR_ARG;                                           \
    if (int r0_ = require_graph(c)) return r0_;                             \
  } while (0)

int scc_trim1(scc_ctx* c, int64_t trim_stop, int64_t* trimmed) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_trim1(c, trim_stop, trimmed);
}

int scc_fb_global(scc_ctx* c, int policy, int32_t pivot_override, int64_t* scc_size,
                  int32_t* pivot_out) {
  REQUIRE_GRAPH(c);
  if (pivot_override >= c->n) return set_error(c, SCC_ERR_ARG, "pivot out of range");
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  if (c->relabeled && pivot_override >= 0) pivot_override = c->h_iperm[pivot_override];
  rc = do_fb_round(c, policy, c->colors_split, pivot_override, scc_size, pivot_out, nullptr, false, nullptr);
  if (!rc && c->relabeled && pivot_out && *pivot_out >= 0) *pivot_out = c->h_perm[*pivot_out];
  return rc;
}

int scc_trim2(scc_ctx* c, int64_t* removed) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_trim2(c, removed);
}

int scc_wcc(scc_ctx* c, int64_t* parts) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_wcc(c, parts);
}

int scc_color_round(scc_ctx* c, int64_t* resolved) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_color_round(c, resolved);
}

int scc_partitions_resolve(scc_ctx* c, int policy, int mode, int trim_between, int64_t* rounds) {
  REQUIRE_GRAPH(c);
  if (mode < SCC_PARTS_FB || mode > SCC_PARTS_LEARNED) return set_error(c, SCC_ERR_ARG, "bad partition mode");
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_fb_parts(c, policy, mode, trim_between != 0, rounds);
}

int scc_fb_partitions(scc_ctx* c, int policy, int trim_between, int64_t* rounds) {
  return scc_partitions_resolve(c, policy, SCC_PARTS_FB, trim_between, rounds);
}

void scc_run_opts_default(scc_run_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->pivot_policy = SCC_PIVOT_DEGREE_PRODUCT;
  // a sweep after the first gathers neighbour colours; it pays only when it
  // removes most of what it examines (what it would remove is otherwise left
  // to the searches, as unreached candidates, and to the trim after them)
  o->trim_stop = -70;
  o->use_trim2 = 1;
  o->trim_between = 0;
  o->use_wcc = 1;
  o->time_phases = 0;
  o->partition_mode = SCC_PARTS_COLOR;
}

int scc_run_method2(scc_ctx* c, const scc_run_opts* opts_in, scc_run_stats* st_out) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c, false)) return r0;  // a deferred reset goes into the first trim sweep
  return sccmi::run_method2(c, opts_in, st_out);
}

}  // extern "C"

namespace sccmi {

// The method-2 pipeline on a context already checked (scc_update runs it on
// its affected region while the context counts as mid-update).
int run_method2(scc_ctx* c, const scc_run_opts* opts_in, scc_run_stats* st_out) {
  scc_run_opts opts;
  if (opts_in) opts = *opts_in;
  else scc_run_opts_default(&opts);
  if (opts.partition_mode < SCC_PARTS_FB || opts.partition_mode > SCC_PARTS_LEARNED)
    return set_error(c, SCC_ERR_ARG, "bad partition mode");
  scc_run_stats st;
  std::memset(&st, 0, sizeof(st));
  st.pivot_first = -1;
  st.chain_stop = -1;
  c->labels_ready = false;
  c->last_color_round = false;
  c->w_fresh = false;
  c->cut_out = c->cut_in = false;
  c->fix_idx = 0;
  c->search_idx = 0;
  c->kernel_ms = 0;
  c->kernel_launches = 0;
  const bool tp = opts.time_phases != 0;
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t* E = c->ev;
  int rc = 0;
  auto mark = [&](int i) {
    if (tp) (void)hipEventRecord(E[2 + i], c->stream);
  };
  auto elapsed = [&](int a, int b) -> double {
    float ms = 0;
    if (tp) {
      (void)hipEventSynchronize(E[2 + b]);
      (void)hipEventElapsedTime(&ms, E[2 + a], E[2 + b]);
    }
    return ms;
  };
  // A fresh decomposition with plain global pivots queues its phases as one
  // chain (run_chain) and continues here only when the chain stopped early (the
  // learned phase schedule only decides the partition steps after it).
  int from = ST_TRIM_FIRST;
  const bool chain = !c->chain_off && !tp && !c->kernel_timing && c->reset_pending && c->fresh && c->n > 0 &&
                     !c->colors_split && !c->multi_colors && opts.pivot_policy != SCC_PIVOT_MIN_ID;
  if (chain) {
    if ((rc = check_cancel(c))) return rc;
    if ((rc = run_chain(c, opts, &st, &from))) return rc;
  }
  do {
    if (from == ST_DONE) break;
    mark(0);
    // (trim sweeps drop finished vertices from the list, so no compaction first)
    if (from <= ST_TRIM_FIRST) {
      int64_t t = 0;
      if ((rc = do_trim1(c, opts.trim_stop, &t))) break;
      st.trimmed_first += t;
    }
    mark(1);
    st.ms_trim1 = elapsed(0, 1);
    if (c->n_live == 0) break;
    if ((rc = check_cancel(c))) break;
    if (from <= ST_FB) {
      // the global step runs over colour -1; with caller-set colours it is a
      // partition round.  Learned pivots rank the candidates of partition
      // rounds; the global step keeps the plain rule (one partition, whose
      // max in x out vertex the network's degree features rank first anyway)
      int64_t sets[5] = {0, 0, 0, 0, 0};
      const int gpol = !c->colors_split && opts.pivot_policy == SCC_PIVOT_LEARNED ? SCC_PIVOT_DEGREE_PRODUCT
                                                                                 : opts.pivot_policy;
      if ((rc = do_fb_round(c, gpol, c->colors_split, -1, &st.scc_first, &st.pivot_first,
                            &st.fb_levels, c->kernel_timing, tp || c->kernel_timing ? sets : nullptr)))
        break;
      st.fb_f = sets[0];
      st.fb_b = sets[1];
      st.fb_ef = sets[2];
      st.fb_eb = sets[3];
      st.fb_fb = sets[4];
    } else if (from == ST_FB_CONTINUE) {
      // the chain's search needed more level launches than it queued
      if ((rc = fb_continue(c, &st.scc_first, &st.fb_levels))) break;
    }
    if (chain && from <= ST_FB_CONTINUE)
      chain_learn_search(c);
    mark(2);
    st.ms_fb_global = elapsed(1, 2);
    int64_t t = 0;
    if (from <= ST_TRIM_SECOND) {
      if ((rc = do_trim1(c, opts.trim_stop, &t))) break;
      st.trimmed_second += t;
    }
    if (opts.use_trim2 && c->n_live > 0 && from <= ST_TRIM_THIRD) {
      if (from <= ST_TRIM2) {
        if ((rc = do_trim2(c, &t))) break;
        st.trimmed_second += t;
      }
      if ((rc = do_trim1(c, opts.trim_stop, &t))) break;
      st.trimmed_second += t;
    }
    mark(3);
    st.ms_trim2 = elapsed(2, 3);
    if (c->n_live == 0) break;
    if ((rc = check_cancel(c))) break;
    if (opts.use_wcc) {
      if ((rc = do_wcc(c, &st.wcc_parts))) break;
    }
    mark(4);
    st.ms_wcc = elapsed(3, 4);
    if ((rc = do_fb_parts(c, opts.pivot_policy, opts.partition_mode, opts.trim_between != 0,
                          &st.fb_rounds)))
      break;
    mark(5);
    st.ms_fb_parts = elapsed(4, 5);
  } while (false);
  if (rc) return rc;
  mark(4);
  if (from != ST_DONE && (rc = do_canon(c, &st.num_sccs))) return rc;
  mark(5);
  st.ms_canon = elapsed(4, 5);
  auto t1 = std::chrono::steady_clock::now();
  st.ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
  st.ms_fb_kernels = c->kernel_ms;
  st.fb_launches = c->kernel_launches;
  if (c->relabeled && st.pivot_first >= 0) st.pivot_first = c->h_perm[st.pivot_first];
  c->ran = true;
  if (st_out) *st_out = st;
  return SCC_OK;
}

}  // namespace sccmi

extern "C" {

int scc_labels_download(scc_ctx* c, int32_t* out, int canonical) {
  REQUIRE_GRAPH(c);
  if (!out) return SCC_ERR_ARG;
  int rc;
  if (canonical && !c->labels_ready) {
    if ((rc = do_canon(c, nullptr))) return rc;
  }
  const int32_t* src = canonical ? c->canon : c->label;
  if (c->relabeled) {
    if (canonical) {
      src = c->canon_caller;  // computed with canon
    } else {
      if ((rc = layout_to_caller(c, c->label, true, c->map_a))) return rc;
      src = c->map_a;
    }
  }
  if ((rc = hip_check(c, hipMemcpyAsync(out, src, 4ull * c->n, hipMemcpyDeviceToHost, c->stream),
                      "labels download")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "labels sync");
}

int scc_colors_download(scc_ctx* c, int32_t* out) {
  REQUIRE_GRAPH(c);
  if (!out) return SCC_ERR_ARG;
  int rc;
  const int32_t* src = c->color;
  if (c->relabeled) {
    if ((rc = layout_to_caller(c, c->color, true, c->map_a))) return rc;
    src = c->map_a;
  }
  if ((rc = hip_check(c, hipMemcpyAsync(out, src, 4ull * c->n, hipMemcpyDeviceToHost, c->stream),
                      "colours download")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "colours sync");
}

int scc_count(scc_ctx* c, int64_t* count) {
  REQUIRE_GRAPH(c);
  return do_canon(c, count);
}

int scc_labels_device(scc_ctx* c, const int32_t** d_labels) {
  REQUIRE_GRAPH(c);
  if (!d_labels) return SCC_ERR_ARG;
  int rc;
  if (!c->labels_ready && (rc = do_canon(c, nullptr))) return rc;
  *d_labels = c->relabeled ? c->canon_caller : c->canon;  // caller-indexed, computed with canon
  return SCC_OK;
}

int scc_ctx_set_kernel_timing(scc_ctx* c, int enable) {
  if (!c) return SCC_ERR_ARG;
  c->kernel_timing = enable != 0;
  return SCC_OK;
}

int scc_ctx_set_chain(scc_ctx* c, int enable, int32_t shape, int32_t level_launches, int32_t sweeps) {
  if (!c) return SCC_ERR_ARG;
  if (shape > 3 || level_launches > 64 || sweeps > 64 || shape < -1 || level_launches == 0 || sweeps == 0 ||
      level_launches < -1 || sweeps < -1)
    return set_error(c, SCC_ERR_ARG, "chain: shape in [0, 3], launches and sweeps in [1, 64] (-1 keeps)");
  c->chain_off = enable == 0;
  if (shape >= 0) c->chain_shape = shape;
  if (level_launches > 0) c->chain_levels = level_launches;
  if (sweeps > 0)
    for (auto& x : c->chain_sweeps) x = sweeps;
  return SCC_OK;
}

int scc_level_log(scc_ctx* c, int64_t* out, int32_t* levels) {
  if (!c || !out || !levels) return SCC_ERR_ARG;
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = pull_ctl(c))) return rc;
  const int32_t k = std::min(64, c->h_dc->lv.levels);
  for (int32_t i = 0; i < k; i++)
    for (int j = 0; j < 4; j++) out[4 * i + j] = (int64_t)c->h_dc->lvlog[i][j];
  if (c->last_dirs == -1)
    for (int32_t i = 0; i < k; i++) out[4 * i] |= 256;  // a backward search alone
  int32_t k2 = 0;
  if (c->last_dirs == 2) {
    k2 = std::min(64 - k, c->h_dc->lv2.levels);
    for (int32_t i = 0; i < k2; i++) {
      for (int j = 0; j < 4; j++) out[4 * (k + i) + j] = (int64_t)c->h_dc->lvlog2[i][j];
      out[4 * (k + i)] |= 256;  // backward search
    }
  }
  *levels = k + k2;
  return SCC_OK;
}

int scc_ctx_synchronize(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  (void)hipSetDevice(c->device);
  return hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
}

}  // extern "C"

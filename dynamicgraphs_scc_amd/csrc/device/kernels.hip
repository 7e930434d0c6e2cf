// gfx950 kernels of the SCC engine.  All data is int32 (ids, offsets, colours,
// labels) plus 1-bit-per-vertex reach bitmaps; no floating point and no MFMA:
// every kernel here is an HBM/L2-bound sweep over a vertex list.
//
// Launch shape: 256-thread workgroups (4 wave64s), grid-stride loops over a
// vertex list whose trip count is uniform across the workgroup, so wave- and
// block-wide collectives (ballot compaction, counters) can be used in the body.
// Row scans use a lane-per-vertex phase over the first LIGHT entries and then a
// wave-cooperative phase for the rest of long rows (64 lanes x 4 loads in flight).
//
// List inputs come from the device control block (DevCtl): the current vertex
// list is live[dc->live.sel][0, dc->live.n).  Kernels that shrink the list write
// the survivors to the other buffer and their last workgroup flips the selector.
#include "common.h"
#include "kernels.h"

namespace sccmi {


// ---- row scans -----------------------------------------------------------------

// Does row [s, e) of `nb` hold an entry k with pred(k, v, c)?  (v, c) are the
// owning vertex and its colour; in the wave-cooperative phase they are taken
// from the owning lane.  Lanes with act=false join the wave-wide steps and
// return false.  Must be reached by all lanes of the wave.
template <class Pred>
__device__ __forceinline__ bool row_any(bool act, int32_t v, int32_t c, int32_t s, int32_t e,
                                        const int32_t* __restrict__ nb, Pred pred, int light) {
  bool found = false;
  const int lim = act ? min(e - s, light) : 0;
  const int mx = wave_max(lim);
  // the first entry alone (it usually decides), then four entries per step
  // with their loads all in flight before any is tested
  if (lim > 0) found = pred(nb[s], v, c);
  if (__ballot(lim > 1 && !found) != 0) {
    for (int j = 1; j < mx; j += 4) {
      int32_t k[4];
#pragma unroll
      for (int t = 0; t < 4; t++) k[t] = (!found && j + t < lim) ? nb[s + j + t] : -1;
      bool hit[4];
#pragma unroll
      for (int t = 0; t < 4; t++) hit[t] = k[t] >= 0 && pred(k[t], v, c);
      found = found || hit[0] || hit[1] || hit[2] || hit[3];
      if (__ballot((j + 4 < lim) && !found) == 0) break;
    }
  }
  uint64_t hm = __ballot(act && !found && (e - s > light));
  const int lane = lane_id();
  while (hm) {
    const int L = __ffsll((unsigned long long)hm) - 1;
    hm &= hm - 1;
    const int32_t hs = __shfl(s, L) + light, he = __shfl(e, L);
    const int32_t hv = __shfl(v, L), hc = __shfl(c, L);
    bool f = false;
    for (int32_t b = hs; b < he; b += 256) {
      bool hit = false;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int32_t idx = b + t * 64 + lane;
        if (idx < he) hit |= pred(nb[idx], hv, hc);
      }
      if (__ballot(hit) != 0) {
        f = true;
        break;
      }
    }
    if (lane == L) found = f;
  }
  return found;
}

// ---- utility kernels ---------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_fill_i32(int32_t* __restrict__ p, int32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    p[i] = v;
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    p[i] = v;
}

__global__ __launch_bounds__(BLOCK) void k_iota(int32_t* __restrict__ p, int32_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    p[i] = (int32_t)i;
}

// Control-block set-up (one thread): the counters, the live list, the loop
// control and the level control.  what: bit 0 counters, 1 live, 2 loop, 3 level.
__device__ void init_level(LevelCtl& v, double bu_ratio) {
  v.done = 0;
  v.mode = MODE_TD;
  v.q_sel = 0;
  v.c_sel = -1;
  v.h_sel = 0;
  v.levels = 0;
  v.blocks_done = 0;
  v.nq = v.nh = v.nc = v.mf = v.prev_nq = 0;
  v.out_q = v.out_c = v.out_h = v.out_edges = 0;
  v.visited = 0;
  v.bu_edges_per_cand = bu_ratio;
}

// what: bit 0 counters, 1 live list, 2 loop control, 3 and 5 the two level
// controls, 4 tuning defaults, 6 the chain state (no halt).  One thread.
__device__ void ctl_init_body(DevCtl* dc, int what, int32_t live_sel, unsigned long long live_n, long long stop,
                              double bu_ratio) {
  if (what & 1) {
    for (int i = 0; i < C_NUM; i++) dc->ctr[i] = 0;
    dc->blocks_done = 0;
    for (int i = 0; i < 9 * 16; i++) dc->done[i] = dc->done2[i] = 0;
  }
  if (what & 2) {
    dc->live.sel = live_sel;
    dc->live.n = live_n;
  }
  if (what & 4) {
    LoopCtl& l = dc->loop;
    l.done = 0;
    l.iters = 0;
    l.blocks_done = 0;
    l.stop = stop;
    l.epoch = 0;  // trim loops number themselves from 1
    l.acc0 = l.acc1 = l.total0 = 0;
  }
  if (what & 16) {
    dc->tune.light = LEVEL_LIGHT;
    dc->tune.huge_row = HUGE_ROW;
    dc->tune.td_beta = 24.f;
    dc->tune.bu_alpha = 14.f;
    dc->tune.grid_cap = GRID_CAP;
    dc->tune.wcc_dense = WCC_DENSE;
  }
  if (what & 8) init_level(dc->lv, bu_ratio);
  if (what & 32) init_level(dc->lv2, bu_ratio);
  if (what & 64) {
    ChainCtl& ch = dc->chain;
    ch.halt = 0;
    ch.gate = -1;
    for (int i = 0; i < CHAIN_GATES; i++) {
      ch.ok[i] = ch.iters[i] = 0;
      ch.live[i] = ch.count[i] = 0;
    }
    ch.pivot = 0;
  }
}

__global__ void k_ctl_init(DevCtl* dc, int what, int32_t live_sel, unsigned long long live_n,
                           long long stop, double bu_ratio) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl_init_body(dc, what, live_sel, live_n, stop, bu_ratio);
}

// Fresh decomposition state in one launch: colours -1, labels -1, the live
// list 0..n-1, zeroed counters.
__global__ __launch_bounds__(BLOCK) void k_reset(DevCtl* __restrict__ dc, int32_t n, int32_t* __restrict__ color,
                                                 int32_t* __restrict__ label, int32_t* __restrict__ live0) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    color[v] = UNSET;
    label[v] = -1;
    live0[v] = (int32_t)v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl_init_body(dc, 1 | 2 | 64, 0, (unsigned long long)n, 0, 0.0);
}

// Before a sweep of k_trim1_step with fresh == 2: the control block as the reset
// leaves it (counters zero, live list 0..n-1 in live[0], no chain halt) and the
// global pivot slot empty.
__global__ void k_fresh_init(DevCtl* __restrict__ dc, int32_t n, unsigned long long* __restrict__ pkey) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctl_init_body(dc, 1 | 2 | 64, 0, (unsigned long long)n, 0, 0.0);
  pkey[n] = 0ull;
}

// Set-up of a forward-backward step in one launch: zeroed counters, both
// search controls, and (single: every live vertex has colour -1) the one
// partition slot in use.
__device__ void fb_setup_body(DevCtl* __restrict__ dc, int32_t n, unsigned long long* __restrict__ pkey,
                              int32_t* __restrict__ rep, int single, double bu_ratio) {
  ctl_init_body(dc, 1 | 8 | 32, 0, 0, 0, bu_ratio);
  if (single) {  // single == 2: pkey[n] already holds the pivot key (k_trim1_step, fresh == 2)
    if (single == 1) pkey[n] = 0ull;
    rep[3 * (int64_t)n + 0] = rep[3 * (int64_t)n + 1] = rep[3 * (int64_t)n + 2] = INT32_MAX;
  }
}
__global__ void k_fb_setup(DevCtl* __restrict__ dc, int32_t n, unsigned long long* __restrict__ pkey,
                           int32_t* __restrict__ rep, int single, double bu_ratio) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || chain_halted(dc)) return;
  fb_setup_body(dc, n, pkey, rep, single, bu_ratio);
}

// ---- chain gates -------------------------------------------------------------------
// One thread, after a phase of a queued chain (ChainCtl): records what the
// phase did and stops the chain when it did not finish within its launches
// (a trim loop still open, or, with single, ended after more than its first
// sweep, whose pivot key then no longer stands; a search still open), or when
// this is the chain's last gate and vertices are still live.  No live vertex:
// the later phases are skipped (halt 2).  One thread; nothing when halted.
__device__ void gate_body(DevCtl* __restrict__ dc, int gate, int kind, int32_t epoch, int single, int last) {
  if (chain_halted(dc)) return;
  ChainCtl& ch = dc->chain;
  bool ok = true;
  if (kind == GATE_TRIM) {
    const LoopCtl& l = dc->loop;
    ok = l.epoch == epoch && l.done && (!single || l.iters == 1);
    ch.iters[gate] = l.iters;
    ch.count[gate] = l.total0;
  } else if (kind == GATE_SEARCH) {
    ok = dc->lv.done && dc->lv2.done;
    ch.iters[gate] = dc->lv.levels + dc->lv2.levels;
    ch.count[gate] = dc->ctr[C_SCC];
    ch.pivot = dc->ctr[C_PIVOT];
  } else {  // GATE_TRIM2: one launch
    ch.count[gate] = dc->ctr[C_TRIMMED];
  }
  const unsigned long long live = dc->live.n;
  ch.ok[gate] = ok;
  ch.live[gate] = live;
  ch.gate = gate;
  if (!ok || (last && live > 0)) ch.halt = 2 * gate + 1;
  else if (live == 0) ch.halt = 2;
}
// zero_ctr: then zero the counters, whatever the gate decided (the chain's
// canonicalisation, queued next, counts from zero)
__global__ void k_chain_gate(DevCtl* __restrict__ dc, int gate, int kind, int32_t epoch, int single, int last,
                             int zero_ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  gate_body(dc, gate, kind, epoch, single, last);
  if (zero_ctr) ctl_init_body(dc, 1, 0, 0, 0, 0.0);
}

// ---- trim1 -------------------------------------------------------------------
// src/scc_trim1.cc:38-110: v of colour c is trimmed (label v, colour -2) when it
// has no out-neighbour k != v of colour c, or no such in-neighbour.  Survivors
// are compacted into the other list (replaces create_trim1_compact, :211-275).
// Reads of neighbour colours race with trims of the same sweep; trimming is
// monotone, so the fixpoint is the same (the unique maximal set where every
// vertex keeps a same-colour in- and out-neighbour).  One launch is one sweep;
// the last workgroup ends the loop when a sweep trims <= stop
// (src/scc_trim1.cc:195-209 repeat_global_trim1).  epoch numbers the loop:
// loop state of another epoch is stale, so the first sweep of a loop needs no
// separate initialisation (the per-sweep accumulators are left zero by every
// sweep's last workgroup).
//
// fresh == 2 is the first sweep straight after a reset that was not launched:
// it also writes colour -1 / label -1 of every survivor (the reset's work),
// takes the global step's pivot key (max in-degree x out-degree over the
// survivors, into pkey[n], which the launch before zeroed) and clears the
// step's reach bitmaps (zwords words at zbits), so that a decomposition whose
// trim loop ends after this sweep needs no reset or pivot-key pass.
__global__ __launch_bounds__(BLOCK) void k_trim1_step(DevCtl* __restrict__ dc, GraphPtrs G,
                                                      int32_t* __restrict__ color,
                                                      int32_t* __restrict__ label,
                                                      int32_t* __restrict__ L0, int32_t* __restrict__ L1,
                                                      int fresh, long long stop, int32_t epoch,
                                                      unsigned long long* __restrict__ pkey,
                                                      uint32_t* __restrict__ zbits, int64_t zwords) {
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = chain_halted(dc) || (dc->loop.epoch == epoch ? dc->loop.done : 0);
  __syncthreads();
  if (s_done) return;
  const int32_t* __restrict__ list_in = LIVE_IN(dc, L0, L1);
  int32_t* __restrict__ list_out = LIVE_OUT(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  auto live = [color](int32_t k, int32_t pv, int32_t pc) { return k != pv && color[k] == pc; };
  unsigned long long trimmed = 0;
  // (room for a whole sweep's survivors of most workgroups: one flush, at the end)
  using Out = LdsOut<0, 8192>;
  Out::init();
  unsigned long long best = 0ull;  // fresh == 2: this thread's largest pivot key
  if (fresh == 2)
    for (int64_t w = (int64_t)blockIdx.x * BLOCK + threadIdx.x; w < zwords; w += (int64_t)gridDim.x * BLOCK)
      zbits[w] = 0u;
  if (fresh) {
    // fresh state: the live list is 0..n-1, none is done and every edge joins
    // two vertices of one colour (all colours -1, or batch colours that no
    // edge crosses), so only empty rows decide this sweep (a row holding nothing but
    // a self-loop survives it and falls to the next sweep, which reads
    // colours; trimming later is always safe, trimming early never happens).
    // Four vertices per thread and step (base + t * BLOCK + thread: loads
    // coalesced, survivors pushed in id order), their row bounds loaded together.
    for (int64_t base = (int64_t)blockIdx.x * 4 * BLOCK; base < n_in; base += (int64_t)gridDim.x * 4 * BLOCK) {
      Out::maybe_flush(list_out, &dc->loop.acc1, 4 * BLOCK);
      int32_t b[4], e[4], rb[4], re[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int64_t v = base + t * BLOCK + threadIdx.x;
        b[t] = e[t] = rb[t] = re[t] = 0;
        if (v < n_in) {
          b[t] = G.begin[v];
          e[t] = G.end[v];
          rb[t] = G.rbegin[v];
          re[t] = G.rend[v];
        }
      }
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int64_t v = base + t * BLOCK + threadIdx.x;
        const bool valid = v < n_in;
        const bool trim = valid && !(e[t] > b[t] && re[t] > rb[t]);
        if (trim) {
          label[v] = (int32_t)v;
          color[v] = DONE;
          G.ckey[v] = canon_key(G.perm, (int32_t)v);
          trimmed++;
        } else if (fresh == 2 && valid) {
          label[v] = -1;
          color[v] = UNSET;
          const unsigned long long p = (unsigned long long)(e[t] - b[t]) * (unsigned long long)(re[t] - rb[t]);
          const unsigned long long k = pivot_key(p > 0x1FFFFFFFFull ? 0x1FFFFFFFFull : p, (int32_t)v);
          best = k > best ? k : best;
        }
        Out::push(valid && !trim, (int32_t)v, list_out, &dc->loop.acc1);
      }
    }
  }
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; !fresh && base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(list_out, &dc->loop.acc1, BLOCK);
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n_in;
    int32_t v = 0, c = DONE;
    bool out_ok, in_ok;
    {
      v = valid ? list_in[i] : 0;
      c = valid ? color[v] : DONE;
      const bool act = valid && c != DONE;
      int32_t s = 0, e = 0;
      if (act) {
        s = G.begin[v];
        e = G.end[v];
      }
      out_ok = row_any(act, v, c, s, e, G.adj, live, LIGHT);
      const bool act2 = act && out_ok;
      int32_t rs = 0, re = 0;
      if (act2) {
        rs = G.rbegin[v];
        re = G.rend[v];
      }
      in_ok = row_any(act2, v, c, rs, re, G.radj, live, LIGHT);
    }
    const bool act = valid && c != DONE;
    const bool trim = act && !in_ok;
    if (trim) {
      label[v] = v;
      color[v] = DONE;
      G.ckey[v] = canon_key(G.perm, v);
      trimmed++;
    }
    Out::push(act && !trim, v, list_out, &dc->loop.acc1);
  }
  Out::flush(list_out, &dc->loop.acc1);
  if (fresh == 2) {  // the workgroup's largest key, one atomic (the slot is read first)
    __shared__ unsigned long long s_best[WAVES];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(best, o);
      best = y > best ? y : best;
    }
    if (lane_id() == 0) s_best[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long x = 0ull;
      for (int w = 0; w < WAVES; w++) x = s_best[w] > x ? s_best[w] : x;
      if (x > pkey[G.n]) atomicMax(&pkey[G.n], x);
    }
  }
  block_count(trimmed, &dc->loop.acc0);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    LoopCtl& l = dc->loop;
    const unsigned long long t = ld_acq(&l.acc0), out = ld_acq(&l.acc1);
    if (l.epoch != epoch) {  // the loop's first sweep
      l.epoch = epoch;
      l.total0 = 0;
      l.iters = 0;
      l.stop = stop;
    }
    l.total0 += t;
    l.iters++;
    // stop >= 0: a sweep trimming at most `stop` vertices ends the loop (the
    // reference's TRIM_STOP); stop < 0: one trimming at most -stop percent of
    // the vertices it examined does
    const bool small = stop >= 0 ? (long long)t <= stop
                                 : t * 100ull <= (unsigned long long)(-stop) * (t + out);
    l.done = small || out == 0;
    l.acc0 = l.acc1 = 0;
    l.blocks_done = 0;
    dc->live.sel ^= 1;
    dc->live.n = out;
  }
}

// Drop vertices of colour -2 from the live list (one launch).
__global__ __launch_bounds__(BLOCK) void k_compact(DevCtl* __restrict__ dc, const int32_t* __restrict__ color,
                                                   int32_t* __restrict__ L0, int32_t* __restrict__ L1) {
  const int32_t* __restrict__ list_in = LIVE_IN(dc, L0, L1);
  int32_t* __restrict__ list_out = LIVE_OUT(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  using Out = LdsOut<0>;
  Out::init();
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(list_out, &dc->ctr[C_LIST_OUT], BLOCK);
    const int64_t i = base + threadIdx.x;
    int32_t v = 0;
    bool keep = false;
    if (i < n_in) {
      v = list_in[i];
      keep = color[v] != DONE;
    }
    Out::push(keep, v, list_out, &dc->ctr[C_LIST_OUT]);
  }
  Out::flush(list_out, &dc->ctr[C_LIST_OUT]);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    dc->live.n = ld_acq(&dc->ctr[C_LIST_OUT]);
    dc->live.sel ^= 1;
    dc->ctr[C_LIST_OUT] = 0;
    dc->blocks_done = 0;
  }
}

// ---- trim2 -------------------------------------------------------------------
// src/scc_trim2_new.cc:3-39: the unique same-colour neighbour other than v
// (repeats skipped), or -1 when there are zero or several.  Rows are short for
// every vertex trim2 can remove, so a lane scans alone and gives up at two.
__device__ __forceinline__ int32_t unique_nbr(const int32_t* __restrict__ off, const int32_t* __restrict__ offe,
                                              const int32_t* __restrict__ nb,
                                              const int32_t* __restrict__ color, int32_t v, int32_t c) {
  int32_t the = -1;
  const int32_t s = off[v], e = offe[v];
  for (int32_t i = s; i < e; i++) {
    const int32_t k = nb[i];
    if (k == v || k == the || color[k] != c) continue;
    if (the >= 0) return -1;
    the = k;
  }
  return the;
}

// src/scc_trim2_new.cc:42-101: {v, k} with v < k, each the other's unique
// same-colour out-neighbour (or in-neighbour), is a 2-vertex SCC labelled v.
__global__ __launch_bounds__(BLOCK) void k_trim2(DevCtl* __restrict__ dc, GraphPtrs G,
                                                 int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                 const int32_t* __restrict__ L0,
                                                 const int32_t* __restrict__ L1) {
  if (chain_halted(dc)) return;
  const int32_t* __restrict__ list_in = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  unsigned long long removed = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list_in[i];
    const int32_t c = color[v];
    if (c == DONE) continue;
    int32_t k = unique_nbr(G.begin, G.end, G.adj, color, v, c);
    bool pair = k > v && unique_nbr(G.begin, G.end, G.adj, color, k, c) == v;
    if (!pair) {
      k = unique_nbr(G.rbegin, G.rend, G.radj, color, v, c);
      pair = k > v && unique_nbr(G.rbegin, G.rend, G.radj, color, k, c) == v;
    }
    if (pair) {
      label[v] = v;
      label[k] = v;
      color[v] = DONE;
      color[k] = DONE;
      const unsigned long long kv = canon_key(G.perm, v), kk = canon_key(G.perm, k);
      G.ckey[v] = kv < kk ? kv : kk;
      removed += 2;
    }
  }
  block_count(removed, &dc->ctr[C_TRIMMED]);
}

// ---- pivot selection -----------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_reset_slots(const DevCtl* __restrict__ dc, int32_t n,
                                                       const int32_t* __restrict__ color,
                                                       unsigned long long* __restrict__ pkey,
                                                       int32_t* __restrict__ rep,
                                                       const int32_t* __restrict__ L0,
                                                       const int32_t* __restrict__ L1, int single) {
  if (single) {  // every live vertex has colour -1: only slot n is in use
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      pkey[n] = 0ull;
      rep[3 * (int64_t)n + 0] = rep[3 * (int64_t)n + 1] = rep[3 * (int64_t)n + 2] = INT32_MAX;
    }
    return;
  }
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t c = color[list[i]];
    if (c == DONE) continue;
    const int32_t s = slot_of(c, n);
    pkey[s] = 0ull;
    rep[3 * (int64_t)s + 0] = INT32_MAX;
    rep[3 * (int64_t)s + 1] = INT32_MAX;
    rep[3 * (int64_t)s + 2] = INT32_MAX;
  }
}

// Per-partition argmax of the pivot score.  SCC_PIVOT_DEGREE_PRODUCT scores
// in-degree x out-degree (clipped to 33 bits); SCC_PIVOT_MIN_ID scores 0 so the
// id tie-break picks the smallest member (src_CUDA/scc_cuda_fb_global.cu:426-440);
// SCC_PIVOT_LEARNED takes scores[] from the scheduler.  The slot is read first
// so that most lanes of a large partition skip the atomic.
__global__ __launch_bounds__(BLOCK) void k_pivot_key(const DevCtl* __restrict__ dc, GraphPtrs G,
                                                     const int32_t* __restrict__ color, int policy,
                                                     const uint32_t* __restrict__ scores,
                                                     unsigned long long* __restrict__ pkey,
                                                     const int32_t* __restrict__ L0,
                                                     const int32_t* __restrict__ L1,
                                                     uint32_t* __restrict__ zbits, int64_t zwords) {
  // the step's reach bitmaps start empty (nothing here reads them)
  for (int64_t w = (int64_t)blockIdx.x * BLOCK + threadIdx.x; w < zwords; w += (int64_t)gridDim.x * BLOCK)
    zbits[w] = 0u;
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  // PK list entries per thread and iteration: one slot check and reduction
  // (four barriers) per PK * BLOCK entries
  constexpr int PK = 4;
  __shared__ int32_t s_slot;
  __shared__ int s_mixed;
  __shared__ unsigned long long s_best[WAVES];
  for (int64_t base = (int64_t)blockIdx.x * PK * BLOCK; base < n_in;
       base += (int64_t)gridDim.x * PK * BLOCK) {
    bool act[PK];
    int32_t slot[PK];
    unsigned long long key[PK];
    int32_t first = -1;
#pragma unroll
    for (int k = 0; k < PK; k++) {
      const int64_t i = base + (int64_t)k * BLOCK + threadIdx.x;
      act[k] = false;
      slot[k] = 0;
      key[k] = 0;
      if (i < n_in) {
        const int32_t v = list[i];
        const int32_t c = color[v];
        if (c != DONE) {
          unsigned long long score = 0;
          if (policy == SCC_PIVOT_DEGREE_PRODUCT) {
            const unsigned long long p = (unsigned long long)(G.end[v] - G.begin[v]) *
                                         (unsigned long long)(G.rend[v] - G.rbegin[v]);
            score = p > 0x1FFFFFFFFull ? 0x1FFFFFFFFull : p;
          } else if (policy == SCC_PIVOT_LEARNED && scores) {
            // the network's estimate in 256 steps, then the plain degree
            // product among equal estimates (saturated estimates tie often)
            const unsigned long long p = (unsigned long long)(G.end[v] - G.begin[v]) *
                                         (unsigned long long)(G.rend[v] - G.rbegin[v]);
            score = ((unsigned long long)(scores[v] >> 24) << 25) | (p > 0x1FFFFFFull ? 0x1FFFFFFull : p);
          }
          key[k] = pivot_key(score, v);
          slot[k] = slot_of(c, G.n);
          act[k] = true;
          if (first < 0) first = slot[k];
        }
      }
    }
    // one slot for the whole workgroup (the global step): reduce in LDS first.
    // The previous iteration ended on a barrier, so no wave still reads s_mixed.
    if (threadIdx.x == 0) s_mixed = 0;
    const uint64_t am = __ballot(first >= 0);
    const int32_t ws = am ? __shfl(first, __ffsll((unsigned long long)am) - 1) : -1;
    __syncthreads();
    if (lane_id() == 0 && am) s_slot = ws;
    __syncthreads();
    const int32_t bs = s_slot;
    bool mis = false;
#pragma unroll
    for (int k = 0; k < PK; k++) mis |= act[k] && slot[k] != bs;
    if (__ballot(mis) != 0 && lane_id() == 0) s_mixed = 1;
    __syncthreads();
    if (!s_mixed) {
      unsigned long long x = 0ull;
#pragma unroll
      for (int k = 0; k < PK; k++) x = (act[k] && key[k] > x) ? key[k] : x;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = __shfl_xor(x, o);
        x = y > x ? y : x;
      }
      if (lane_id() == 0) s_best[threadIdx.x >> 6] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long b = 0;
        for (int w = 0; w < WAVES; w++) b = s_best[w] > b ? s_best[w] : b;
        if (b && b > pkey[bs]) atomicMax(&pkey[bs], b);
      }
    } else {
#pragma unroll
      for (int k = 0; k < PK; k++) grouped_max_u64(act[k], slot[k], key[k], pkey);
    }
    // every wave has read s_mixed and s_best before the next iteration resets them
    __syncthreads();
  }
}

// Force the pivot of `p`'s partition (src/scc_fb_global.cc:140-147 override).
__global__ void k_force_pivot(int32_t p, int32_t n, const int32_t* __restrict__ color,
                              unsigned long long* __restrict__ pkey) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int32_t c = color[p];
    if (c != DONE) pkey[slot_of(c, n)] = pivot_key(0x1FFFFFFFFull, p);
  }
}

// Entries per chunk of a row cut over the grid in a top-down level
constexpr int TD_CHUNK = 1024;  // (2048 measured 7 us slower per R-MAT-22 step, 512 3 us)

// Seeds of the two searches of a step, which run together: each live vertex
// that is its partition's pivot marks itself in both visited bitmaps and enters
// both frontiers (or, for a row of at least huge_row entries, the chunked
// expansion of that direction's first level).  The summed row lengths give the
// first direction-switch decision.
__global__ __launch_bounds__(BLOCK) void k_seed_pivots2(DevCtl* __restrict__ dc, int32_t n, int multi, LevelDir F,
                                                        LevelDir B, const int32_t* __restrict__ color,
                                                        const unsigned long long* __restrict__ pkey,
                                                        const int32_t* __restrict__ L0,
                                                        const int32_t* __restrict__ L1, CanonKeys K) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  const int32_t huge = dc->tune.huge_row;
  unsigned long long ef = 0, eb = 0, cnt = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool piv = false;
    int32_t v = 0, df = 0, db = 0;
    if (i < n_in) {
      v = list[i];
      const int32_t c = color[v];
      piv = c != DONE && pivot_of_key(pkey[slot_of(c, n)]) == v;
      if (piv) {
        K.ckey[v] = canon_key(K.perm, v);  // the pivot's SCC starts its key slot
        const uint32_t bit = 1u << (v & 31);
        atomicOr(&F.V[v >> 5], bit);
        atomicOr(&B.V[v >> 5], bit);
        df = F.offe[v] - F.off[v];
        db = B.offe[v] - B.off[v];
        ef += (unsigned long long)df;
        eb += (unsigned long long)db;
        cnt++;
        // the pivot, when there is one partition (with many, thousands of
        // same-address stores would serialise for a value nobody reads)
        if (!multi) dc->ctr[C_PIVOT] = (unsigned long long)v;
      }
    }
    wave_append(piv && df < huge, v, F.q0, &dc->lv.nq);
    wave_append(piv && df >= huge, v, F.h0, &dc->lv.nh);
    wave_append(piv && db < huge, v, B.q0, &dc->lv2.nq);
    wave_append(piv && db >= huge, v, B.h0, &dc->lv2.nh);
  }
  block_count(ef, &dc->lv.mf);
  block_count(eb, &dc->lv2.mf);
  block_count(cnt, &dc->lv.visited);
  block_count(cnt, &dc->lv2.visited);
}

// The seeds of the global step's two searches when there is one partition
// (colour -1, slot n): its pivot, from pkey[n], with no pass over the live list.
__device__ void seed_single_body(DevCtl* __restrict__ dc, int32_t n, const LevelDir& F, const LevelDir& B,
                                 const unsigned long long* __restrict__ pkey, CanonKeys K) {
  const unsigned long long key = pkey[n];
  if (key == 0ull) return;  // no live vertex
  const int32_t v = pivot_of_key(key);
  K.ckey[v] = canon_key(K.perm, v);
  const uint32_t bit = 1u << (v & 31);
  atomicOr(&F.V[v >> 5], bit);
  atomicOr(&B.V[v >> 5], bit);
  const int32_t df = F.offe[v] - F.off[v], db = B.offe[v] - B.off[v];
  // the only frontier vertex: a row longer than one chunk is cut into chunks
  // over the grid (one wave would otherwise walk all of it alone)
  const int32_t huge = min(dc->tune.huge_row, (int32_t)TD_CHUNK);
  if (df < huge) F.q0[dc->lv.nq++] = v;
  else F.h0[dc->lv.nh++] = v;
  if (db < huge) B.q0[dc->lv2.nq++] = v;
  else B.h0[dc->lv2.nh++] = v;
  dc->lv.mf += (unsigned long long)df;
  dc->lv2.mf += (unsigned long long)db;
  dc->lv.visited += 1;
  dc->lv2.visited += 1;
  dc->ctr[C_PIVOT] = (unsigned long long)v;
}
__global__ void k_seed_single(DevCtl* __restrict__ dc, int32_t n, LevelDir F, LevelDir B,
                              const unsigned long long* __restrict__ pkey, CanonKeys K) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || chain_halted(dc)) return;
  seed_single_body(dc, n, F, B, pkey, K);
}

// A chain's first gate (its first trim loop, which must have ended after the
// fused first sweep) with, when it lets the chain go on, the global step's
// set-up and its pivot's seeding: k_chain_gate, k_fb_setup (single 2) and
// k_seed_single in one launch.
__global__ void k_chain_gate_fb(DevCtl* __restrict__ dc, int32_t epoch, int last, int32_t n,
                                unsigned long long* __restrict__ pkey, int32_t* __restrict__ rep, double bu_ratio,
                                LevelDir F, LevelDir B, CanonKeys K) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  gate_body(dc, 0, GATE_TRIM, epoch, 1, last);
  if (chain_halted(dc)) return;
  fb_setup_body(dc, n, pkey, rep, 2, bu_ratio);
  seed_single_body(dc, n, F, B, pkey, K);
}

// Seed a search from an explicit vertex list (count at *ns): vertices not yet
// in V and not of colour -2 enter the frontier once.
__global__ __launch_bounds__(BLOCK) void k_seed_list(DevCtl* __restrict__ dc, const int32_t* __restrict__ off, const int32_t* __restrict__ offe,
                                                     const int32_t* __restrict__ color,
                                                     const int32_t* __restrict__ seeds,
                                                     const unsigned long long* __restrict__ ns,
                                                     uint32_t* __restrict__ V, int32_t* __restrict__ q0) {
  const int64_t n_in = (int64_t)*ns;
  unsigned long long edges = 0, cnt = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool take = false;
    int32_t v = 0;
    if (i < n_in) {
      v = seeds[i];
      if (color[v] != DONE) {
        const uint32_t bit = 1u << (v & 31);
        take = !(atomicOr(&V[v >> 5], bit) & bit);
        if (take) {
          edges += (unsigned long long)(offe[v] - off[v]);
          cnt++;
        }
      }
    }
    wave_append(take, v, q0, &dc->lv.nq);
  }
  block_count(edges, &dc->lv.mf);
  block_count(cnt, &dc->lv.visited);
}

// ---- reachability (direction-optimising, partition-aware) ----------------------
// One level of the forward (rows = out-edges) or backward (rows = in-edges)
// search of src/scc_fb_global.cc:32-129 / gm_bfs_template.h, for every
// partition at once: a vertex is admitted when it is reached from a visited
// vertex of the same colour (the navigator of src/scc_fb_global.cc:64-68 and
// :123-127; during a search colours do not change, so "same colour" is the
// partition).  Visited state is a bitmap, claimed with atomicOr.

// Set bit v of V for every lane with flag, one atomicOr per distinct word in
// the wave (lanes hold nearby ids, so a wave touches few words).  All lanes.
__device__ __forceinline__ void set_bits_wave(bool flag, int32_t v, uint32_t* __restrict__ V) {
  uint64_t m = __ballot(flag);
  const int lane = lane_id();
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int32_t w = __shfl(v >> 5, leader);
    const bool mine = flag && (v >> 5) == w;
    uint32_t bits = mine ? (1u << (v & 31)) : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bits |= __shfl_xor(bits, o);
    if (lane == leader) atomicOr(&V[w], bits);
    m &= ~__ballot(mine);
  }
}

__device__ __forceinline__ bool try_claim(uint32_t* __restrict__ V, const int32_t* __restrict__ color,
                                          int32_t k, int32_t c) {
  const uint32_t bit = 1u << (k & 31);
  if (V[k >> 5] & bit) return false;
  if (color[k] != c) return false;
  return !(atomicOr(&V[k >> 5], bit) & bit);
}

// try_claim for four neighbours at once (k[t] < 0: none).  Each round of
// loads is issued for all four before any result is used: visited words, then
// the colours of the unvisited, then the atomics; a lane then waits on memory
// three times per four neighbours instead of three times per neighbour.
__device__ __forceinline__ void claim4(uint32_t* __restrict__ V, const int32_t* __restrict__ color,
                                       const int32_t (&k)[4], int32_t c, bool (&cl)[4]) {
  uint32_t w[4];
#pragma unroll
  for (int t = 0; t < 4; t++) w[t] = k[t] >= 0 ? V[k[t] >> 5] : 0xFFFFFFFFu;
  bool open[4];
  int32_t ck[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    open[t] = k[t] >= 0 && !((w[t] >> (k[t] & 31)) & 1u);
    ck[t] = open[t] ? color[k[t]] : DONE;
  }
#pragma unroll
  for (int t = 0; t < 4; t++) {
    cl[t] = false;
    if (open[t] && ck[t] == c) {
      const uint32_t bit = 1u << (k[t] & 31);
      cl[t] = !(atomicOr(&V[k[t] >> 5], bit) & bit);
    }
  }
}

// claim4 for four consecutive entries of one row held by one lane: entries
// whose ids share a visited word are claimed with one atomicOr issued by the
// first of them, and an id repeated among the four is claimed at most once.
__device__ __forceinline__ void claim4_run(uint32_t* __restrict__ V, const int32_t* __restrict__ color,
                                           const int32_t (&k)[4], int32_t c, bool (&cl)[4]) {
  uint32_t w[4];
#pragma unroll
  for (int t = 0; t < 4; t++) w[t] = k[t] >= 0 ? V[k[t] >> 5] : 0xFFFFFFFFu;
  bool want[4];
  int32_t ck[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    want[t] = k[t] >= 0 && !((w[t] >> (k[t] & 31)) & 1u);
    ck[t] = want[t] ? color[k[t]] : DONE;
  }
#pragma unroll
  for (int t = 0; t < 4; t++) want[t] = want[t] && ck[t] == c;
  // lead[t]: the group (word) of t starts at t; bits[t]: the group's bits
  int g[4];
  uint32_t bits[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    g[t] = t;
    bits[t] = 0u;
#pragma unroll
    for (int u = 0; u < t; u++)
      if (g[t] == t && want[u] && want[t] && (k[u] >> 5) == (k[t] >> 5)) g[t] = u;
  }
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const uint32_t b = want[t] ? 1u << (k[t] & 31) : 0u;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (g[t] == u) bits[u] |= b;
  }
  uint32_t old[4];
#pragma unroll
  for (int t = 0; t < 4; t++) old[t] = (want[t] && g[t] == t) ? atomicOr(&V[k[t] >> 5], bits[t]) : 0xFFFFFFFFu;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    uint32_t og = old[0];
#pragma unroll
    for (int u = 1; u < 4; u++)
      if (g[t] == u) og = old[u];
    bool dup = false;
#pragma unroll
    for (int u = 0; u < t; u++) dup = dup || (want[u] && k[u] == k[t]);
    cl[t] = want[t] && !dup && !((og >> (k[t] & 31)) & 1u);
  }
}

// Whether any of four candidates parents k[t] (k[t] < 0: none) is visited
// and, with several partitions, of colour c; the loads of all four are issued
// together.
__device__ __forceinline__ bool reached4(const uint32_t* __restrict__ V, const int32_t* __restrict__ color,
                                         int multi, const int32_t (&k)[4], int32_t c) {
  uint32_t w[4];
#pragma unroll
  for (int t = 0; t < 4; t++) w[t] = k[t] >= 0 ? V[k[t] >> 5] : 0u;
  bool hit[4];
#pragma unroll
  for (int t = 0; t < 4; t++) hit[t] = k[t] >= 0 && ((w[t] >> (k[t] & 31)) & 1u);
  if (multi) {
    int32_t ck[4];
#pragma unroll
    for (int t = 0; t < 4; t++) ck[t] = hit[t] ? color[k[t]] : DONE;
#pragma unroll
    for (int t = 0; t < 4; t++) hit[t] = hit[t] && ck[t] == c;
  }
  return hit[0] || hit[1] || hit[2] || hit[3];
}

// Lanes per frontier vertex in a top-down level: 1, or a power of two up to 16
// when the frontier (nq vertices, mf edges) leaves most of the search's
// lanes idle and its rows are long enough to feed a group (a mean row of at
// least 2G entries; a group takes up to 8G entries of a row, in two four-load
// steps).
__device__ __forceinline__ int td_group(unsigned long long nq, unsigned long long mf, unsigned long long lanes) {
  int G = 1;
  while (G < 16 && nq * (unsigned long long)(2 * G) <= lanes && mf >= nq * (unsigned long long)(4 * G)) G *= 2;
  return G;
}


// One level.  Top-down: expand the frontier (rows of at least HUGE_ROW entries
// are deferred to the next launch, which cuts them into TD_CHUNK pieces spread
// over the whole grid).  Bottom-up: every candidate not yet visited scans its
// opposite-direction row for a visited parent of its colour
// (gm_bfs_template.h:281-301, :810-829); a parent admitted earlier in the same
// launch also counts (reachability needs no levels).  The last workgroup
// applies Beamer's rule (BU when the frontier's edges exceed the candidates'
// edges / 14, back to TD when the frontier shrinks below n/24) and ends the
// search when a level admits nothing.  A switch to BU drops deferred rows:
// their neighbours find them as parents.
// One workgroup's part of a level: workgroup bid of nblk working on the
// search whose control is lv / done / lvlog and whose arrays are D.
template <int s_light>
__device__ __forceinline__ void level_part(DevCtl* __restrict__ dc, LevelCtl* __restrict__ lv,
                                           uint32_t* __restrict__ done, unsigned long long (*lvlog)[4],
                                           int multi, const LevelDir& D, const int32_t* __restrict__ color,
                                           const int32_t* __restrict__ L0, const int32_t* __restrict__ L1,
                                           int32_t max_chunks, unsigned bid, unsigned nblk) {
  const int32_t* __restrict__ off = D.off;
  const int32_t* __restrict__ offe = D.offe;
  const int32_t* __restrict__ nb = D.nb;
  const int32_t* __restrict__ roff = D.roff;
  const int32_t* __restrict__ roffe = D.roffe;
  const int32_t* __restrict__ rnb = D.rnb;
  uint32_t* __restrict__ V = D.V;
  int32_t* __restrict__ q0 = D.q0;
  int32_t* __restrict__ q1 = D.q1;
  int32_t* __restrict__ c0 = D.c0;
  int32_t* __restrict__ c1 = D.c1;
  int32_t* __restrict__ h0 = D.h0;
  int32_t* __restrict__ h1 = D.h1;
  __shared__ int s_done, s_mode, s_q, s_c, s_h, s_huge, s_group;
  __shared__ unsigned long long s_nq, s_nh, s_nc;
  if (threadIdx.x == 0) {
    s_huge = dc->tune.huge_row;
    s_done = lv->done;
    s_mode = lv->mode;
    s_q = lv->q_sel;
    s_c = lv->c_sel;
    s_h = lv->h_sel;
    s_nq = lv->nq;
    s_nh = lv->nh;
    s_nc = lv->nc;
    s_group = td_group(lv->nq, lv->mf, (unsigned long long)nblk * BLOCK);
  }
  __syncthreads();
  if (s_done) return;
  const int lane = lane_id();
  unsigned long long edges = 0;
  using QOut = LdsOut<0>;
  using COut = LdsOut<1>;
  QOut::init();
  COut::init();
  int32_t* q_out_g = s_q ? q0 : q1;
  if (s_mode == MODE_TD) {
    const int32_t* __restrict__ q_in = s_q ? q1 : q0;
    int32_t* __restrict__ q_out = s_q ? q0 : q1;
    const int32_t* __restrict__ h_in = s_h ? h1 : h0;
    int32_t* __restrict__ h_out = s_h ? h0 : h1;
    // (a) rows deferred by the previous level, chunked over the grid
    const int64_t nh = (int64_t)s_nh;
    if (nh > 0) {
      const int64_t total = nh * max_chunks;
      for (int64_t g = bid; g < total; g += nblk) {
        QOut::maybe_flush(q_out, &lv->out_q, BLOCK);
        const int32_t u = h_in[g % nh];
        const int32_t ch = (int32_t)(g / nh);
        const int32_t s = off[u] + ch * TD_CHUNK;
        const int32_t e = min(s + TD_CHUNK, offe[u]);
        if (s >= e) continue;  // uniform across the workgroup
        const int32_t c = color[u];
        // each thread takes TD_CHUNK / BLOCK consecutive entries: the lanes of
        // a wave are that many entries apart, so where the sorted row is dense
        // (hub neighbours with nearby ids) they claim in different visited
        // words instead of queueing on one
        constexpr int PER = TD_CHUNK / BLOCK;
        const int32_t tb = s + (int32_t)threadIdx.x * PER;
        for (int32_t j = 0; j < PER; j += 4) {
          int32_t k[4];
          bool cl[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const int32_t idx = tb + j + t;
            k[t] = idx < e ? nb[idx] : -1;
          }
          claim4_run(V, color, k, c, cl);
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (cl[t]) edges += (unsigned long long)(offe[k[t]] - off[k[t]]);
#pragma unroll
          for (int t = 0; t < 4; t++) QOut::push(cl[t], k[t], q_out, &lv->out_q);
        }
      }
    }
    // (b) the frontier
    const int64_t nq = (int64_t)s_nq;
    const int G = s_group;
    if (G > 1) {
      // lane groups: G lanes share a frontier vertex and take its first
      // 8G entries (two four-load steps); longer rows pass their rest to the
      // whole wave as below
      const int g = threadIdx.x & (G - 1);
      const int per_blk = BLOCK / G;
      const int limg = 8 * G;
      for (int64_t base = (int64_t)bid * per_blk; base < nq; base += (int64_t)nblk * per_blk) {
        QOut::maybe_flush(q_out, &lv->out_q, BLOCK);
        const int64_t i = base + threadIdx.x / G;
        const bool valid = i < nq;
        int32_t u = 0, s = 0, e = 0, c = DONE;
        if (valid) {
          u = q_in[i];
          s = off[u];
          e = offe[u];
          c = color[u];
        }
        const int32_t d = e - s;
        const bool hug = valid && d >= s_huge;
        wave_append(hug && g == 0, u, h_out, &lv->out_h);
        const bool act = valid && !hug;
        const int lim = act ? min(d, limg) : 0;
        const int mx = wave_max(lim);
        for (int j = 0; j < mx; j += 4 * G) {
          int32_t k[4];
          bool cl[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const int jj = j + t * G + g;
            k[t] = jj < lim ? nb[s + jj] : -1;
          }
          claim4(V, color, k, c, cl);
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (cl[t]) edges += (unsigned long long)(offe[k[t]] - off[k[t]]);
#pragma unroll
          for (int t = 0; t < 4; t++) QOut::push(cl[t], k[t], q_out, &lv->out_q);
        }
        uint64_t hm = __ballot(act && g == 0 && d > limg);
        while (hm) {
          const int L = __ffsll((unsigned long long)hm) - 1;
          hm &= hm - 1;
          const int32_t hs = __shfl(s, L) + limg, he = __shfl(e, L), hc = __shfl(c, L);
          for (int32_t b = hs; b < he; b += 256) {
            int32_t k[4];
            bool cl[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
              const int32_t idx = b + t * 64 + lane;
              k[t] = idx < he ? nb[idx] : -1;
            }
            claim4(V, color, k, hc, cl);
#pragma unroll
            for (int t = 0; t < 4; t++)
              if (cl[t]) edges += (unsigned long long)(offe[k[t]] - off[k[t]]);
#pragma unroll
            for (int t = 0; t < 4; t++) QOut::push(cl[t], k[t], q_out, &lv->out_q);
          }
        }
      }
    }
    for (int64_t base = (int64_t)bid * BLOCK; G == 1 && base < nq; base += (int64_t)nblk * BLOCK) {
      QOut::maybe_flush(q_out, &lv->out_q, BLOCK);
      const int64_t i = base + threadIdx.x;
      const bool valid = i < nq;
      int32_t u = 0, s = 0, e = 0, c = DONE;
      if (valid) {
        u = q_in[i];
        s = off[u];
        e = offe[u];
        c = color[u];
      }
      const int32_t d = e - s;
      const bool hug = valid && d >= s_huge;
      wave_append(hug, u, h_out, &lv->out_h);
      const bool act = valid && !hug;
      const int lim = act ? min(d, s_light) : 0;
      const int mx = wave_max(lim);
      for (int j = 0; j < mx; j += 4) {
        int32_t k[4];
        bool cl[4];
#pragma unroll
        for (int t = 0; t < 4; t++) k[t] = j + t < lim ? nb[s + j + t] : -1;
        claim4(V, color, k, c, cl);
#pragma unroll
        for (int t = 0; t < 4; t++)
          if (cl[t]) edges += (unsigned long long)(offe[k[t]] - off[k[t]]);
#pragma unroll
        for (int t = 0; t < 4; t++) QOut::push(cl[t], k[t], q_out, &lv->out_q);
      }
      uint64_t hm = __ballot(act && d > s_light);
      while (hm) {
        const int L = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int32_t hs = __shfl(s, L) + s_light, he = __shfl(e, L), hc = __shfl(c, L);
        for (int32_t b = hs; b < he; b += 256) {
          int32_t k[4];
          bool cl[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const int32_t idx = b + t * 64 + lane;
            k[t] = idx < he ? nb[idx] : -1;
          }
          claim4(V, color, k, hc, cl);
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (cl[t]) edges += (unsigned long long)(offe[k[t]] - off[k[t]]);
#pragma unroll
          for (int t = 0; t < 4; t++) QOut::push(cl[t], k[t], q_out, &lv->out_q);
        }
      }
    }
  } else {
    const int32_t* __restrict__ c_in = (s_c < 0) ? LIVE_IN(dc, L0, L1) : (s_c ? c1 : c0);
    const int64_t ncin = (s_c < 0) ? (int64_t)dc->live.n : (int64_t)s_nc;
    int32_t* __restrict__ c_out = (s_c == 0) ? c1 : c0;
    int32_t* __restrict__ q_out = s_q ? q0 : q1;
    __shared__ int32_t s_qs[BLOCK], s_qe[BLOCK], s_qc[BLOCK], s_qhit[BLOCK];
    __shared__ int s_bq;
    for (int64_t base = (int64_t)bid * BLOCK; base < ncin; base += (int64_t)nblk * BLOCK) {
      QOut::maybe_flush(q_out, &lv->out_q, BLOCK);
      COut::maybe_flush(c_out, &lv->out_c, BLOCK);
      const int64_t i = base + threadIdx.x;
      const bool valid = i < ncin;
      int32_t v = 0, c = DONE, s = 0, e = 0, hd = -1;
      bool act = false;
      if (valid) {
        v = c_in[i];
        // independent loads: visited bit, colour (several partitions only), row
        // bounds, the row's first entry (from the heads: no row-start line)
        const bool vis = bit_test(V, v);
        c = multi ? color[v] : UNSET;
        s = roff[v];
        e = roffe[v];
        if (D.rhead) hd = D.rhead[v];
        act = !vis && c != DONE;
      }
      // phase 1: each lane tests the first s_light entries of its row
      bool found = false;
      {
        const int lim = act ? min(e - s, s_light) : 0;
        const int mx = wave_max(lim);
        // the first entry alone (hub-first rows put a likely parent there),
        // then four per step with their loads issued together
        if (lim > 0) {
          const int32_t k0[4] = {D.rhead ? hd : rnb[s], -1, -1, -1};
          found = reached4(V, color, multi, k0, c);
        }
        if (__ballot(lim > 1 && !found) != 0) {
          for (int j = 1; j < mx; j += 4) {
            int32_t k[4];
#pragma unroll
            for (int t = 0; t < 4; t++) k[t] = (!found && j + t < lim) ? rnb[s + j + t] : -1;
            found = found || reached4(V, color, multi, k, c);
            if (__ballot((j + 4 < lim) && !found) == 0) break;
          }
        }
      }
      // phase 2: rows still open go to a workgroup queue and are scanned by
      // 16-lane groups, so one long row no longer holds up its wave's other lanes
      if (threadIdx.x == 0) s_bq = 0;
      __syncthreads();
      int slot = -1;
      if (act && !found && e - s > s_light) {
        slot = atomicAdd(&s_bq, 1);
        s_qs[slot] = s + s_light;
        s_qe[slot] = e;
        s_qc[slot] = c;
        s_qhit[slot] = 0;
      }
      __syncthreads();
      {
        const int nbq = s_bq;
        const int gl = threadIdx.x & 15, gsh = lane & 48;
        for (int j = threadIdx.x >> 4; j < nbq; j += BLOCK / 16) {
          const int32_t rs = s_qs[j], re = s_qe[j], rc = s_qc[j];
          bool hit = false;
          for (int32_t b = rs; b < re; b += 64) {
            int32_t k[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
              const int32_t idx = b + t * 16 + gl;
              k[t] = idx < re ? rnb[idx] : -1;
            }
            const bool h = reached4(V, color, multi, k, rc);
            if ((__ballot(h) >> gsh) & 0xFFFFull) {
              hit = true;
              break;
            }
          }
          if (gl == 0 && hit) s_qhit[j] = 1;
        }
      }
      __syncthreads();
      if (slot >= 0) found = s_qhit[slot] != 0;
      // (no edge count: after a bottom-up level only the admitted count decides
      // the next direction)
      set_bits_wave(found, v, V);
      QOut::push(found, v, q_out, &lv->out_q);
      COut::push(act && !found, v, c_out, &lv->out_c);
    }
    COut::flush(c_out, &lv->out_c);
  }
  QOut::flush(q_out_g, &lv->out_q);
  block_count(edges, &lv->out_edges);
  if (last_block2(done, bid, nblk) && threadIdx.x == 0) {
    const unsigned long long oq = ld_acq(&lv->out_q), oc = ld_acq(&lv->out_c);
    const unsigned long long oh = ld_acq(&lv->out_h), oe = ld_acq(&lv->out_edges);
    const unsigned long long nlive = dc->live.n;
    if (lv->levels < 64) {
      unsigned long long* lg = lvlog[lv->levels];
      lg[0] = (unsigned long long)s_mode;
      lg[1] = s_mode == MODE_TD ? (unsigned long long)s_nq + ((unsigned long long)s_nh << 32)
                                : (s_c < 0 ? nlive : (unsigned long long)s_nc);
      lg[2] = oq;
      lg[3] = oe;
    }
    lv->visited += oq;
    lv->levels++;
    if (s_mode == MODE_TD) {
      lv->done = (oq == 0 && oh == 0);
      lv->h_sel ^= 1;
      lv->nh = oh;
    } else {
      lv->done = (oq == 0);
      lv->c_sel = (s_c == 0) ? 1 : 0;
      lv->nc = oc;
      lv->nh = 0;
    }
    // one partition: every live vertex reached means nothing is left to find
    if (!multi && lv->visited >= nlive) lv->done = 1;
    lv->q_sel ^= 1;
    lv->prev_nq = lv->nq;
    lv->nq = oq;
    lv->mf = oe;
    if (s_mode == MODE_TD) {
      const double ncand = (lv->c_sel < 0) ? (double)nlive : (double)lv->nc;
      if ((double)oe > ncand * lv->bu_edges_per_cand && ncand > 0) {
        lv->mode = MODE_BU;
        lv->nh = 0;
      }
    } else if (oq < lv->prev_nq &&
               (double)oq * (multi ? dc->tune.td_beta * MULTI_TD_SCALE : dc->tune.td_beta) < (double)nlive) {
      lv->mode = MODE_TD;
    }
    lv->out_q = lv->out_c = lv->out_h = lv->out_edges = 0;
    lv->blocks_done = 0;
  }
}

template <int s_light>
__global__ __launch_bounds__(BLOCK) void k_level(DevCtl* __restrict__ dc, int multi, LevelDir D,
                                                 const int32_t* __restrict__ color,
                                                 const int32_t* __restrict__ L0, const int32_t* __restrict__ L1,
                                                 int32_t max_chunks) {
  level_part<s_light>(dc, &dc->lv, dc->done, dc->lvlog, multi, D, color, L0, L1, max_chunks, blockIdx.x,
                      gridDim.x);
}

// The forward and the backward search of a step advance one level each in the
// same launch: workgroups alternate between them in runs of eight (so each XCD
// serves both) and each search completes on its own counters.  The grid is a
// multiple of 16.  A finished search's workgroups return at once.
template <int s_light>
__global__ __launch_bounds__(BLOCK) void k_level2(DevCtl* __restrict__ dc, int multi, LevelDir F, LevelDir B,
                                                  const int32_t* __restrict__ color,
                                                  const int32_t* __restrict__ L0, const int32_t* __restrict__ L1,
                                                  int32_t max_chunks) {
  if (chain_halted(dc)) return;
  const bool bw = (blockIdx.x >> 3) & 1u;
  const unsigned bid = ((blockIdx.x >> 4) << 3) | (blockIdx.x & 7u);
  level_part<s_light>(dc, bw ? &dc->lv2 : &dc->lv, bw ? dc->done2 : dc->done, bw ? dc->lvlog2 : dc->lvlog, multi,
                      bw ? B : F, color, L0, L1, max_chunks, bid, gridDim.x >> 1);
}

// The kernels after a forward-backward step's searches are queued right behind
// each batch of level launches, before the host has seen whether the searches
// ended; they do nothing until both have (every workgroup reads the same
// flags: the level launches before them in the stream have finished).
__device__ __forceinline__ bool searches_open(const DevCtl* __restrict__ dc) {
  return chain_halted(dc) || !(dc->lv.done && dc->lv2.done);
}

// Sizes of a step's reach sets for the traffic model: |F|, |B|, E_F (out-degrees
// over F), E_B (in-degrees over B), |F n B|.
__global__ __launch_bounds__(BLOCK) void k_fb_stats(DevCtl* __restrict__ dc, GraphPtrs G,
                                                    const uint32_t* __restrict__ vf,
                                                    const uint32_t* __restrict__ vb,
                                                    const int32_t* __restrict__ L0,
                                                    const int32_t* __restrict__ L1) {
  if (searches_open(dc)) return;
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  unsigned long long f = 0, b = 0, ef = 0, eb = 0, fb = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    const bool inf = bit_test(vf, v), inb = bit_test(vb, v);
    if (inf) {
      f++;
      ef += (unsigned long long)(G.end[v] - G.begin[v]);
    }
    if (inb) {
      b++;
      eb += (unsigned long long)(G.rend[v] - G.rbegin[v]);
    }
    fb += (inf && inb) ? 1 : 0;
  }
  block_count(f, &dc->ctr[C_F]);
  block_count(b, &dc->ctr[C_B]);
  block_count(ef, &dc->ctr[C_EF]);
  block_count(eb, &dc->ctr[C_EB]);
  block_count(fb, &dc->ctr[C_FB]);
}

// ---- classification after a forward-backward step -------------------------------
// src/scc_fb_global.cc:99-113: forward and backward reached -> SCC of the pivot
// (label pivot, colour -2).  The rest split into fw-only / bw-only / neither
// (src/scc_fb_seq2.cc:245-285); each new part's colour is its minimum member,
// gathered here with atomicMin and applied by k_recolor.
__global__ __launch_bounds__(BLOCK) void k_classify(DevCtl* __restrict__ dc, int32_t n,
                                                    int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                    const uint32_t* __restrict__ vf,
                                                    const uint32_t* __restrict__ vb,
                                                    const unsigned long long* __restrict__ pkey,
                                                    int32_t* __restrict__ rep, int32_t* __restrict__ L0,
                                                    int32_t* __restrict__ L1, CanonKeys K, int single) {
  if (searches_open(dc)) return;
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  int32_t* __restrict__ list_out = LIVE_OUT(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  constexpr int PER = 8;  // entries per thread and step
  // the partition of the workgroup's first entry (single: the one partition,
  // colour -1): its three representatives and its SCC's canonical key are
  // reduced in LDS and written with one global atomic each, instead of every
  // wave hitting the same addresses
  __shared__ int32_t s_hot, s_hmin[3], s_hpiv;
  __shared__ unsigned long long s_hkey;
  if (threadIdx.x == 0) {
    if (single) {
      s_hot = n;
      s_hpiv = pivot_of_key(pkey[n]);
    } else {
      const int64_t i0 = (int64_t)blockIdx.x * PER * BLOCK;
      const int32_t c0 = i0 < n_in ? color[list[i0]] : DONE;
      s_hot = c0 != DONE ? slot_of(c0, n) : -1;
      s_hpiv = c0 != DONE ? pivot_of_key(pkey[s_hot]) : -1;
    }
    s_hkey = ~0ull;
  }
  if (threadIdx.x < 3) s_hmin[threadIdx.x] = INT32_MAX;
  __syncthreads();
  const int32_t hot = s_hot;
  unsigned long long nscc = 0;
  using Out = LdsOut<0, 8192>;  // (few flushes: each is a barrier and a waited atomic)
  Out::init();
  // eight entries per thread and step (base + t * BLOCK + thread: survivors
  // keep the list order); each round of loads (entries; then colours, both
  // reach words and caller ids) is issued for all eight before any is used
  unsigned long long hkey = ~0ull;  // this thread's least key among the hot SCC's members
  for (int64_t base = (int64_t)blockIdx.x * PER * BLOCK; base < n_in; base += (int64_t)gridDim.x * PER * BLOCK) {
    Out::maybe_flush(list_out, &dc->ctr[C_LIST_OUT], PER * BLOCK);
    int32_t v[PER], c[PER], pm[PER];
    uint32_t wf[PER], wb[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
      const int64_t i = base + t * BLOCK + threadIdx.x;
      v[t] = i < n_in ? list[i] : -1;
    }
#pragma unroll
    for (int t = 0; t < PER; t++) {
      c[t] = DONE;
      wf[t] = wb[t] = 0u;
      pm[t] = v[t];
      if (v[t] >= 0) {
        c[t] = color[v[t]];
        wf[t] = vf[v[t] >> 5];
        wb[t] = vb[v[t] >> 5];
        if (K.perm) pm[t] = K.perm[v[t]];
      }
    }
#pragma unroll
    for (int t = 0; t < PER; t++) {
      const bool f = c[t] != DONE && ((wf[t] >> (v[t] & 31)) & 1u);
      const bool b = c[t] != DONE && ((wb[t] >> (v[t] & 31)) & 1u);
      const unsigned long long kv = ((unsigned long long)(uint32_t)pm[t] << 32) | (uint32_t)v[t];
      bool keep = false, hk = false, inscc = false;
      int64_t rslot = 0;
      int32_t piv = 0;
      if (c[t] != DONE) {
        const int32_t sl = slot_of(c[t], n);
        if (f && b) {
          piv = sl == hot ? s_hpiv : pivot_of_key(pkey[sl]);
          label[v[t]] = piv;
          color[v[t]] = DONE;
          nscc++;
          if (sl == hot) hkey = kv < hkey ? kv : hkey;
          else inscc = true;
        } else {
          const int k = f ? 0 : (b ? 1 : 2);
          rslot = 3 * (int64_t)sl + k;
          keep = true;
          hk = sl == hot;
          if (hk) atomicMin(&s_hmin[k], v[t]);
        }
      }
      grouped_min_i32(keep && !hk, rslot, v[t], rep);
      if (inscc && kv < K.ckey[piv]) atomicMin(&K.ckey[piv], kv);  // small SCCs: a direct atomic, read first
      Out::push(keep, v[t], list_out, &dc->ctr[C_LIST_OUT]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(hkey, o);
    hkey = y < hkey ? y : hkey;
  }
  if (lane_id() == 0 && hkey != ~0ull) atomicMin(&s_hkey, hkey);
  Out::flush(list_out, &dc->ctr[C_LIST_OUT]);  // (begins with a barrier)
  if (threadIdx.x < 3 && hot >= 0) {
    const int32_t x = s_hmin[threadIdx.x];
    int32_t* r = &rep[3 * (int64_t)hot + threadIdx.x];
    if (x < *r) atomicMin(r, x);
  }
  if (threadIdx.x == 3 && hot >= 0 && s_hkey != ~0ull && s_hkey < K.ckey[s_hpiv]) atomicMin(&K.ckey[s_hpiv], s_hkey);
  block_count(nscc, &dc->ctr[C_SCC]);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    dc->live.n = ld_acq(&dc->ctr[C_LIST_OUT]);
    dc->live.sel ^= 1;
    dc->ctr[C_LIST_OUT] = 0;
    dc->blocks_done = 0;
  }
}

// gate1: also the chain's gate after the global step (not its last gate), by
// thread 0 of workgroup 0 first.  A workgroup that reads the halt it may set
// returns, one that read it before runs: the same either way, since the gate
// stops the chain only when the searches are still open (nothing to recolour
// then) or no vertex is live (nothing to recolour either).
__global__ __launch_bounds__(BLOCK) void k_recolor(DevCtl* __restrict__ dc, int32_t n,
                                                   int32_t* __restrict__ color,
                                                   const uint32_t* __restrict__ vf,
                                                   const uint32_t* __restrict__ vb,
                                                   const int32_t* __restrict__ rep,
                                                   const int32_t* __restrict__ L0,
                                                   const int32_t* __restrict__ L1, int gate1) {
  if (gate1 && blockIdx.x == 0 && threadIdx.x == 0) gate_body(dc, 1, GATE_SEARCH, 0, 0, 0);
  if (searches_open(dc)) return;
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    const int32_t s = slot_of(color[v], n);
    const bool f = bit_test(vf, v), b = bit_test(vb, v);
    const int t = f ? 0 : (b ? 1 : 2);
    color[v] = rep[3 * (int64_t)s + t];
  }
}

// ---- WCC colouring --------------------------------------------------------------
// ---- WCC labels, packed with the colour -------------------------------------------
// cw[v] = colour << 32 | W: one 8-byte gather per neighbour tests the colour
// and reads the label, and a 64-bit atomicMin on the packed word lowers the
// label of a same-colour neighbour.  Vertices of colour -2 never match a live
// colour (-1 or a member id).
__device__ __forceinline__ unsigned long long cw_pack(int32_t c, int32_t w) {
  return ((unsigned long long)(uint32_t)c << 32) | (uint32_t)w;
}
__device__ __forceinline__ uint32_t cw_hi(unsigned long long x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ int32_t cw_lo(unsigned long long x) { return (int32_t)(uint32_t)x; }

// src/scc_weak.cc:137-150 for every vertex: W = v, tagged when out-degree >= 50.
__global__ __launch_bounds__(BLOCK) void k_cw_init(int32_t n, const int32_t* __restrict__ begin,
                                                   const int32_t* __restrict__ bend,
                                                   const int32_t* __restrict__ color,
                                                   unsigned long long* __restrict__ cw) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t w = (bend[v] - begin[v] >= 50) ? ((int32_t)v | WCC_TAG) : (int32_t)v;
    cw[v] = cw_pack(color[v], w);
  }
}

// One pull sweep (src/scc_weak.cc:65-125, both phases fused): every live
// vertex pulls the minimum label of its same-colour out-neighbours, then jumps
// to its root's label when the root is not self-rooted.  Vertices whose label
// dropped are appended to qout (count at *qcount).  All threads of the
// workgroup; workgroup bid of nblk.
// With CUT, the sweep also writes each row cut to the neighbours it matched
// (live, same colour) at fout[begin[v]...fe[v]): colours do not change within
// a fixpoint and only split between steps, so later sweeps and steps may
// follow the cut rows.  fout may be adj itself (bend == fe): a row is
// compacted by the thread or wave that reads it, every store at or behind
// entries already loaded.
template <bool CUT>
__device__ __forceinline__ void wcc_pull_sweep(const int32_t* __restrict__ list, int64_t n_in,
                                               const int32_t* __restrict__ begin, const int32_t* bend,
                                               const int32_t* adj, unsigned long long* __restrict__ cw,
                                               int32_t* __restrict__ qout, unsigned long long* qcount,
                                               int32_t* fe, int32_t* fout) {
  const int lane = lane_id();
  using Out = LdsOut<0>;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(qout, qcount, BLOCK);
    const int64_t i = base + threadIdx.x;
    int32_t v = 0, s = 0, e = 0, m = INT32_MAX, w0 = INT32_MAX;
    uint32_t hv = (uint32_t)DONE;
    if (i < n_in) {
      v = list[i];
      const unsigned long long x = cw[v];
      hv = cw_hi(x);
      w0 = m = cw_lo(x);
      s = begin[v];
      e = bend[v];
    }
    const bool act = hv != (uint32_t)DONE;
    const int len = act ? e - s : 0;
    const int lim = len < 32 ? len : 0;
    int32_t k = s;  // (CUT) next store of v's row
    // four neighbours per step, their loads issued together
    for (int j = 0; j < lim; j += 4) {
      int32_t a[4];
#pragma unroll
      for (int t = 0; t < 4; t++) a[t] = j + t < lim ? adj[s + j + t] : -1;
      unsigned long long y[4];
#pragma unroll
      for (int t = 0; t < 4; t++) y[t] = a[t] >= 0 ? cw[a[t]] : 0ull;
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (a[t] >= 0 && cw_hi(y[t]) == hv) {
          m = min(m, cw_lo(y[t]));
          if (CUT) {
            if (fout != adj || k != s + j + t) fout[k] = a[t];  // (in place: entries not yet moved stay)
            k++;
          }
        }
    }
    if (CUT && act && len < 32) fe[v] = k;
    uint64_t hm = __ballot(len >= 32);
    while (hm) {
      const int L = __ffsll((unsigned long long)hm) - 1;
      hm &= hm - 1;
      const int32_t hs = __shfl(s, L), he = __shfl(e, L);
      const uint32_t hh = __shfl(hv, L);
      int32_t hmin = INT32_MAX;
      int32_t w = hs;  // (CUT) next store of the row, the same in every lane
      for (int32_t b = hs; b < he; b += 256) {
        int32_t a[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int32_t idx = b + t * 64 + lane;
          a[t] = idx < he ? adj[idx] : -1;
        }
        unsigned long long y[4];
#pragma unroll
        for (int t = 0; t < 4; t++) y[t] = a[t] >= 0 ? cw[a[t]] : 0ull;
        bool keep[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          keep[t] = a[t] >= 0 && cw_hi(y[t]) == hh;
          if (keep[t]) hmin = min(hmin, cw_lo(y[t]));
        }
        if (CUT) {
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const uint64_t km = __ballot(keep[t]);
            const int32_t pos = w + lanes_below(km);
            if (keep[t] && (fout != adj || pos != b + t * 64 + lane)) fout[pos] = a[t];
            w += __popcll(km);
          }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) hmin = min(hmin, __shfl_xor(hmin, o));
      if (lane == L) {
        m = min(m, hmin);
        if (CUT) fe[v] = w;
      }
    }
    bool changed = false;
    if (act) {
      const int32_t r = m & WCC_MASK;
      if (r != v) {
        const unsigned long long y = cw[r];
        if (cw_hi(y) == hv && (cw_lo(y) & WCC_MASK) != r) m = min(m, cw_lo(y));
      }
      if (m < w0) {  // only v's own thread writes cw[v] in a pull sweep
        cw[v] = cw_pack((int32_t)hv, m);
        changed = true;
      }
    }
    Out::push(changed, v, qout, qcount);
  }
  Out::flush(qout, qcount);
}

// First iteration: a pull sweep over the live list; the vertices whose label
// dropped form the frontier of the later iterations.  With fout, the sweep
// cuts the rows it reads (wcc_pull_sweep<true>).
__global__ __launch_bounds__(BLOCK) void k_wcc_pull(DevCtl* __restrict__ dc, const int32_t* __restrict__ begin,
                                                    const int32_t* bend, const int32_t* adj,
                                                    unsigned long long* __restrict__ cw,
                                                    int32_t* __restrict__ q0, const int32_t* __restrict__ L0,
                                                    const int32_t* __restrict__ L1, int32_t* fe, int32_t* fout) {
  LdsOut<0>::init();
  if (fout)
    wcc_pull_sweep<true>(LIVE_IN(dc, L0, L1), (int64_t)dc->live.n, begin, bend, adj, cw, q0, &dc->lv.out_q, fe,
                         fout);
  else
    wcc_pull_sweep<false>(LIVE_IN(dc, L0, L1), (int64_t)dc->live.n, begin, bend, adj, cw, q0, &dc->lv.out_q,
                          nullptr, nullptr);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    LevelCtl& lv = dc->lv;
    const unsigned long long nq = ld_acq(&lv.out_q);
    lv.nq = nq;
    lv.out_q = 0;
    lv.q_sel = 0;
    lv.levels = 1;
    lv.mode = 1;  // the last iteration pulled (claimed no frontier bits)
    dc->loop.iters = 1;
    dc->loop.total0 += nq;
    dc->loop.done = nq == 0;
    lv.blocks_done = 0;
  }
}

// Local rounds of a push iteration, and the capacity of each of the two LDS
// queues they alternate between.
constexpr int WCC_LOCAL_ROUNDS = 8;
constexpr int WCC_LOCAL_CAP = 2048;
struct LocalQ {
  __device__ static int32_t* buf0() {
    __shared__ int32_t b[WCC_LOCAL_CAP];
    return b;
  }
  __device__ static int32_t* buf1() {
    __shared__ int32_t b[WCC_LOCAL_CAP];
    return b;
  }
  __device__ static int* cnt() {
    __shared__ int c[4];
    return c;
  }
};

// Later iterations.  While the frontier (vertices whose label dropped) holds
// more than 1/wcc_dense of the live set (a tuning knob), an iteration is another pull sweep
// over the whole live list; below that, every frontier vertex pushes its
// label to its same-colour in-neighbours (reverse rows) with a 64-bit
// atomicMin; a neighbour whose label dropped is pushed from again by the same
// workgroup in its local rounds, or joins the next frontier once (claimed in
// the bitmap of this iteration's parity; each iteration clears the claims of
// its own input frontier, so a bitmap is clean when it is reused).  Ends when
// an iteration leaves nothing to the next frontier.  The fixpoint is the pull
// fixpoint: W[v] = min tag over v's same-colour forward-reachable set.
__global__ __launch_bounds__(BLOCK) void k_wcc_push(DevCtl* __restrict__ dc, const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ bend,
                                                    const int32_t* __restrict__ adj,
                                                    const int32_t* __restrict__ rbegin,
                                                    const int32_t* __restrict__ rend,
                                                    const int32_t* __restrict__ radj,
                                                    unsigned long long* __restrict__ cw,
                                                    int32_t* __restrict__ q0, int32_t* __restrict__ q1,
                                                    uint32_t* __restrict__ B0, uint32_t* __restrict__ B1,
                                                    const int32_t* __restrict__ L0, const int32_t* __restrict__ L1) {
  __shared__ int s_done, s_sel, s_t, s_pulled;
  __shared__ unsigned long long s_nq;
  if (threadIdx.x == 0) {
    s_done = dc->loop.done;
    s_sel = dc->lv.q_sel;
    s_t = dc->lv.levels;
    s_pulled = dc->lv.mode;
    s_nq = dc->lv.nq;
  }
  __syncthreads();
  if (s_done) return;
  const int32_t* __restrict__ qin = s_sel ? q1 : q0;
  int32_t* __restrict__ qout = s_sel ? q0 : q1;
  uint32_t* __restrict__ Bin = ((s_t - 1) & 1) ? B1 : B0;
  uint32_t* __restrict__ Bout = (s_t & 1) ? B1 : B0;
  const int64_t nq = (int64_t)s_nq;
  const int64_t n_live = (int64_t)dc->live.n;
  const bool dense = nq * dc->tune.wcc_dense > n_live;
  const int lane = lane_id();
  using Out = LdsOut<0>;
  Out::init();
  if (!s_pulled && dense) {
    // the input frontier's claims were made by the previous (push) iteration
    // (a push iteration clears them as it reads its input, below)
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nq; i += (int64_t)gridDim.x * BLOCK) {
      const int32_t k = qin[i];
      atomicAnd(&Bin[k >> 5], ~(1u << (k & 31)));
    }
  }
  if (dense) {
    wcc_pull_sweep<false>(LIVE_IN(dc, L0, L1), n_live, begin, bend, adj, cw, qout, &dc->lv.out_q, nullptr,
                          nullptr);
  } else {
    // Local rounds.  A vertex whose label a push lowered goes to this
    // workgroup's LDS queue, and the workgroup pushes from that queue itself,
    // round after round, up to WCC_LOCAL_ROUNDS rounds, before the launch
    // ends: a label then travels several hops per launch instead of one (the
    // frontier's tail is a long chain of small iterations).  What a full LDS
    // queue or the last round leaves goes to the next launch's frontier, as a
    // claim in Bout.  The fixpoint does not depend on the order of the
    // relaxations: every vertex lowered is pushed again after its last
    // lowering, here or by a later launch.
    int32_t* lq0 = LocalQ::buf0();
    int32_t* lq1 = LocalQ::buf1();
    int* lc = LocalQ::cnt();  // [0],[1]: reserved in lq0/lq1; [2],[3]: their valid limits
    if (threadIdx.x == 0) {
      lc[0] = lc[1] = 0;
      lc[2] = lc[3] = WCC_LOCAL_CAP;
    }
    __syncthreads();
    // push label x to four in-neighbours u[t] (u[t] < 0: none); low[t] when
    // the push lowered u[t]'s label.  Each round of loads or atomics is
    // issued for all four together.
    auto lower4 = [&](const int32_t (&u)[4], unsigned long long x, bool (&low)[4]) {
      unsigned long long y[4];
#pragma unroll
      for (int t = 0; t < 4; t++) y[t] = u[t] >= 0 ? cw[u[t]] : 0ull;
      bool want[4];
#pragma unroll
      for (int t = 0; t < 4; t++) want[t] = u[t] >= 0 && cw_hi(y[t]) == cw_hi(x) && cw_lo(x) < cw_lo(y[t]);
      unsigned long long old[4];
#pragma unroll
      for (int t = 0; t < 4; t++) old[t] = want[t] ? atomicMin(&cw[u[t]], x) : 0ull;
#pragma unroll
      for (int t = 0; t < 4; t++) low[t] = want[t] && old[t] > x;
    };
    // all lanes of a wave: claim the lowered vertices for the next launch's
    // frontier (once each)
    auto to_global = [&](const bool (&g)[4], const int32_t (&u)[4]) {
      uint32_t ob[4];
#pragma unroll
      for (int t = 0; t < 4; t++) ob[t] = g[t] ? atomicOr(&Bout[u[t] >> 5], 1u << (u[t] & 31)) : 0xFFFFFFFFu;
#pragma unroll
      for (int t = 0; t < 4; t++) Out::push(g[t] && !((ob[t] >> (u[t] & 31)) & 1u), u[t], qout, &dc->lv.out_q);
    };
    // all lanes of a wave: lowered vertices into the local queue `nx`, or,
    // when it is full or `last`, to the global frontier
    auto sink4 = [&](const bool (&low)[4], const int32_t (&u)[4], int nx, bool last) {
      int32_t* lq = nx ? lq1 : lq0;
      bool g[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        g[t] = low[t];
        if (last) continue;
        const uint64_t m = __ballot(low[t]);
        if (m == 0) continue;
        const int kk = __popcll(m);
        int b = 0;
        if (lane == 0) {
          b = atomicAdd(&lc[nx], kk);
          if (b + kk > WCC_LOCAL_CAP) {
            atomicMin(&lc[2 + nx], b);
            b = -1;
          }
        }
        b = __shfl(b, 0);
        if (b >= 0) {
          if (low[t]) lq[b + lanes_below(m)] = u[t];
          g[t] = false;
        }
      }
      to_global(g, u);
    };
    // all lanes of a wave: push the labels of the lane's vertex k (act) to its
    // in-neighbours
    auto push_rows = [&](bool act, int32_t k, unsigned long long x, int32_t s, int32_t e, int nx, bool last) {
      const int len = act ? e - s : 0;
      const int lim = len < 32 ? len : 0;
      const int wl = wave_max(lim);
      for (int j = 0; j < wl; j += 4) {
        int32_t u[4];
        bool low[4];
#pragma unroll
        for (int t = 0; t < 4; t++) u[t] = j + t < lim ? radj[s + j + t] : -1;
        lower4(u, x, low);
        sink4(low, u, nx, last);
      }
      uint64_t hm = __ballot(len >= 32);
      while (hm) {
        const int L = __ffsll((unsigned long long)hm) - 1;
        hm &= hm - 1;
        const int32_t hs = __shfl(s, L), he = __shfl(e, L);
        const unsigned long long hx = __shfl(x, L);
        for (int32_t b = hs; b < he; b += 256) {
          int32_t u[4];
          bool low[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const int32_t idx = b + t * 64 + lane;
            u[t] = idx < he ? radj[idx] : -1;
          }
          lower4(u, hx, low);
          sink4(low, u, nx, last);
        }
      }
    };
    for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < nq; base += (int64_t)gridDim.x * BLOCK) {
      Out::maybe_flush(qout, &dc->lv.out_q, 4 * BLOCK);
      const int64_t i = base + threadIdx.x;
      int32_t k = 0, s = 0, e = 0;
      unsigned long long x = 0;
      bool act = false;
      if (i < nq) {
        k = qin[i];
        if (!s_pulled) atomicAnd(&Bin[k >> 5], ~(1u << (k & 31)));
        x = cw[k];
        s = rbegin[k];
        e = rend[k];
        act = cw_hi(x) != (uint32_t)DONE;
      }
      push_rows(act, k, x, s, e, 0, WCC_LOCAL_ROUNDS == 0);
    }
    // rounds over the local queue: round r reads queue r & 1 and fills the
    // other one; the last round sends what it lowers to the global frontier
    for (int r = 0; r < WCC_LOCAL_ROUNDS; r++) {
      const int cu = r & 1, nx = cu ^ 1;
      __syncthreads();
      const int nc = min(lc[cu], lc[2 + cu]);
      if (nc == 0) break;  // (the same in every thread: read after the barrier)
      const bool last = r + 1 == WCC_LOCAL_ROUNDS;
      __syncthreads();
      if (threadIdx.x == 0) {
        lc[nx] = 0;
        lc[2 + nx] = WCC_LOCAL_CAP;
      }
      __syncthreads();
      const int32_t* lq = cu ? lq1 : lq0;
      for (int b0 = 0; b0 < nc; b0 += BLOCK) {
        Out::maybe_flush(qout, &dc->lv.out_q, 4 * BLOCK);
        const int ii = b0 + (int)threadIdx.x;
        int32_t k = 0, s = 0, e = 0;
        unsigned long long x = 0;
        bool act = false;
        if (ii < nc) {
          k = lq[ii];
          x = cw[k];
          s = rbegin[k];
          e = rend[k];
          act = cw_hi(x) != (uint32_t)DONE;
        }
        push_rows(act, k, x, s, e, nx, last);
      }
      if (threadIdx.x == 0) lc[cu] = 0;  // (consumed; a later break sees it empty)
    }
    Out::flush(qout, &dc->lv.out_q);
  }
  if (last_block2(dc->done) && threadIdx.x == 0) {
    LevelCtl& lv = dc->lv;
    const unsigned long long nn = ld_acq(&lv.out_q);
    lv.nq = nn;
    lv.out_q = 0;
    lv.q_sel ^= 1;
    lv.levels++;
    lv.mode = dense ? 1 : 0;
    dc->loop.iters++;
    dc->loop.total0 += nn;
    dc->loop.done = nn == 0;
    lv.blocks_done = 0;
  }
}

// src/scc_weak.cc:166-194: each class takes its root's id as colour (a unique
// colour: the root is a member of its class); counts roots.  With seed != 0 the
// roots also seed a forward search (V, q0, dc->lv) for root-SCC extraction.
// With cw, the labels come from the fixpoint's packed words and are stored to
// W on the way; without, they are read from W.
__global__ __launch_bounds__(BLOCK) void k_wcc_color(DevCtl* __restrict__ dc, int32_t* __restrict__ color,
                                                     const unsigned long long* __restrict__ cw,
                                                     int32_t* __restrict__ W, int seed,
                                                     const int32_t* __restrict__ off, const int32_t* __restrict__ offe,
                                                     uint32_t* __restrict__ V, int32_t* __restrict__ q0,
                                                     const int32_t* __restrict__ L0,
                                                     const int32_t* __restrict__ L1, CanonKeys K) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  unsigned long long roots = 0, edges = 0;
  using Out = LdsOut<0>;
  Out::init();
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    if (seed) Out::maybe_flush(q0, &dc->lv.nq, BLOCK);
    const int64_t i = base + threadIdx.x;
    bool root = false;
    int32_t v = 0;
    if (i < n_in) {
      v = list[i];
      int32_t w;
      if (cw) {
        w = cw_lo(cw[v]);
        W[v] = w;
      } else {
        w = W[v];
      }
      if (color[v] != DONE) {
        const int32_t r = w & WCC_MASK;
        color[v] = r;
        root = r == v;
        if (root) {
          roots++;
          if (seed) {
            K.ckey[v] = canon_key(K.perm, v);  // the root's SCC starts its key slot
            atomicOr(&V[v >> 5], 1u << (v & 31));
            edges += (unsigned long long)(offe[v] - off[v]);
          }
        }
      }
    }
    if (seed) Out::push(root, v, q0, &dc->lv.nq);
  }
  if (seed) Out::flush(q0, &dc->lv.nq);
  block_count(roots, &dc->ctr[C_ROOTS]);
  if (seed) {
    block_count(edges, &dc->lv.mf);
    block_count(roots, &dc->lv.visited);
  }
}

// Root-SCC extraction after a colouring: every class member reaches its root
// (W is a minimum over the forward-reachable set and the path to the root stays
// in the class), so the root's forward reach inside its class is its SCC.
// Reached vertices take the root (their colour) as label and leave the list.
__global__ __launch_bounds__(BLOCK) void k_extract(DevCtl* __restrict__ dc, int32_t* __restrict__ color,
                                                   int32_t* __restrict__ label,
                                                   const uint32_t* __restrict__ V, int32_t* __restrict__ L0,
                                                   int32_t* __restrict__ L1, CanonKeys K) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  int32_t* __restrict__ list_out = LIVE_OUT(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  unsigned long long nscc = 0;
  using Out = LdsOut<0>;
  Out::init();
  // the class of the workgroup's first entry: its SCC's key reduced in LDS
  __shared__ int32_t s_hot;
  __shared__ unsigned long long s_hkey;
  if (threadIdx.x == 0) {
    const int64_t i0 = (int64_t)blockIdx.x * BLOCK;
    s_hot = i0 < n_in ? color[list[i0]] : DONE;
    s_hkey = ~0ull;
  }
  __syncthreads();
  const int32_t hot = s_hot;
  unsigned long long hkey = ~0ull;  // this thread's least key among the hot SCC's members
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(list_out, &dc->ctr[C_LIST_OUT], BLOCK);
    const int64_t i = base + threadIdx.x;
    bool keep = false, inscc = false;
    int32_t v = 0, c = DONE;
    if (i < n_in) {
      v = list[i];
      c = color[v];
      if (c != DONE) {
        if (bit_test(V, v)) {
          label[v] = c;
          color[v] = DONE;
          nscc++;
          if (c == hot) {
            const unsigned long long kv = canon_key(K.perm, v);
            hkey = kv < hkey ? kv : hkey;
          } else {
            inscc = true;
          }
        } else {
          keep = true;
        }
      }
    }
    if (inscc) {  // small SCCs (many labels per wave): a direct atomic, read first
      const unsigned long long kv = canon_key(K.perm, v);
      if (kv < K.ckey[c]) atomicMin(&K.ckey[c], kv);
    }
    Out::push(keep, v, list_out, &dc->ctr[C_LIST_OUT]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(hkey, o);
    hkey = y < hkey ? y : hkey;
  }
  if (lane_id() == 0 && hkey != ~0ull) atomicMin(&s_hkey, hkey);
  Out::flush(list_out, &dc->ctr[C_LIST_OUT]);  // (begins with a barrier)
  if (threadIdx.x == 0 && hot != DONE && s_hkey != ~0ull && s_hkey < K.ckey[hot]) atomicMin(&K.ckey[hot], s_hkey);
  block_count(nscc, &dc->ctr[C_SCC]);
  if (last_block2(dc->done) && threadIdx.x == 0) {
    dc->live.n = ld_acq(&dc->ctr[C_LIST_OUT]);
    dc->live.sel ^= 1;
    dc->ctr[C_LIST_OUT] = 0;
    dc->blocks_done = 0;
  }
}

// ---- canonical labels -----------------------------------------------------------
// One pass: the canonical label of v is the low half of ckey[label[v]] (the
// phases that gave the labels min-reduced their members' keys there, see
// CanonKeys), its caller-id label the high half, stored at cout[perm[v]] in
// the degree vertex layout.  A label that many workgroups meet (a giant SCC's)
// is read once per workgroup.

// The most frequent label held by at least two of entries
// [blockIdx.x * BLOCK, + 64) that are not self-labelled, or -1.  All threads.
__device__ int32_t block_hot_label(int32_t n, const int32_t* __restrict__ label) {
  __shared__ int32_t s_hot;
  if (threadIdx.x < 64) {  // wave 0
    const int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int32_t l = v < n ? label[v] : -1;
    bool pend = (uint32_t)l < (uint32_t)n && l != (int32_t)v;
    int32_t best = -1;
    int bestc = 1;
    uint64_t m = __ballot(pend);
    while (m) {
      const int32_t L = __shfl(l, __ffsll((unsigned long long)m) - 1);
      const bool same = pend && l == L;
      const int cnt = __popcll(__ballot(same));
      if (cnt > bestc) {
        bestc = cnt;
        best = L;
      }
      if (same) pend = false;
      m = __ballot(pend);
    }
    if (threadIdx.x == 0) s_hot = best;
  }
  __syncthreads();
  return s_hot;
}

template <bool CALLER>
__global__ __launch_bounds__(BLOCK) void k_canon_keys(DevCtl* __restrict__ dc, int32_t n,
                                                      const int32_t* __restrict__ label,
                                                      const unsigned long long* __restrict__ ckey,
                                                      const int32_t* __restrict__ perm,
                                                      int32_t* __restrict__ out, int32_t* __restrict__ cout) {
  if (dc->chain.halt & 1) return;  // a chain stopped before its end
  // a chain that ran to its end (halt 2) leaves no halt for later launches:
  // every workgroup reads 0 or 2 here and runs either way
  if (blockIdx.x == 0 && threadIdx.x == 0 && dc->chain.halt == 2) dc->chain.halt = 0;
  const int32_t hot = block_hot_label(n, label);
  __shared__ unsigned long long s_hk;
  if (threadIdx.x == 0) s_hk = hot >= 0 ? ckey[hot] : ~0ull;
  __syncthreads();
  const unsigned long long hk = s_hk;
  // -1 (both halves) for a vertex not yet labelled
  auto key = [&](int32_t l) -> unsigned long long {
    if (l == hot) return hk;
    return (uint32_t)l < (uint32_t)n ? ckey[l] : ~0ull;
  };
  unsigned long long self = 0;
  // four labels per lane (16-byte loads and stores; the four gathers are
  // independent), then the tail
  const int64_t n4 = (int64_t)n / 4;
  const int4* __restrict__ label4 = reinterpret_cast<const int4*>(label);
  const int4* __restrict__ perm4 = reinterpret_cast<const int4*>(perm);
  int4* __restrict__ out4 = reinterpret_cast<int4*>(out);
  for (int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x; q < n4; q += (int64_t)gridDim.x * BLOCK) {
    const int4 l = label4[q];
    const unsigned long long k0 = key(l.x), k1 = key(l.y), k2 = key(l.z), k3 = key(l.w);
    const int4 c = make_int4((int32_t)k0, (int32_t)k1, (int32_t)k2, (int32_t)k3);
    out4[q] = c;
    if (CALLER) {
      const int4 pv = perm4[q];
      cout[pv.x] = (int32_t)(k0 >> 32);
      cout[pv.y] = (int32_t)(k1 >> 32);
      cout[pv.z] = (int32_t)(k2 >> 32);
      cout[pv.w] = (int32_t)(k3 >> 32);
    }
    const int32_t v = (int32_t)(4 * q);
    self += (c.x == v) + (c.y == v + 1) + (c.z == v + 2) + (c.w == v + 3);
  }
  if (blockIdx.x == 0)
    for (int64_t v = 4 * n4 + threadIdx.x; v < n; v += BLOCK) {
      const unsigned long long k = key(label[v]);
      out[v] = (int32_t)k;
      if (CALLER) cout[perm[v]] = (int32_t)(k >> 32);
      self += ((int32_t)k == (int32_t)v) ? 1 : 0;
    }
  block_count_nowait(self, &dc->ctr[C_COUNT]);
}

// Caller-order canonical labels in the degree vertex layout: cout[u] is the
// high half (the minimum caller id) of the key of the label of internal vertex
// iperm[u].  The writes are sequential; the label and key reads follow the
// degree classes, within each of which caller order is internal order.  Eight
// caller ids per thread, each stage's loads issued together.  Counts the
// caller ids that are their component's label.  In a chain: nothing when it
// stopped (halt odd), and a finished chain's halt is cleared.
__global__ __launch_bounds__(BLOCK) void k_canon_caller(DevCtl* __restrict__ dc, int32_t n,
                                                        const int32_t* __restrict__ label,
                                                        const unsigned long long* __restrict__ ckey,
                                                        const int32_t* __restrict__ iperm,
                                                        int32_t* __restrict__ cout) {
  if (dc->chain.halt & 1) return;
  if (blockIdx.x == 0 && threadIdx.x == 0 && dc->chain.halt == 2) dc->chain.halt = 0;
  auto key = [&](int32_t l) -> unsigned long long { return (uint32_t)l < (uint32_t)n ? ckey[l] : ~0ull; };
  unsigned long long self = 0;
  const int64_t n8 = (int64_t)n / 8;
  const int4* __restrict__ ip4 = reinterpret_cast<const int4*>(iperm);
  int4* __restrict__ c4 = reinterpret_cast<int4*>(cout);
  for (int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x; q < n8; q += (int64_t)gridDim.x * BLOCK) {
    const int4 a = ip4[2 * q], b = ip4[2 * q + 1];
    const int32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    int32_t l[8];
#pragma unroll
    for (int k = 0; k < 8; k++) l[k] = label[v[k]];
    int32_t o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (int32_t)(key(l[k]) >> 32);
    c4[2 * q] = make_int4(o[0], o[1], o[2], o[3]);
    c4[2 * q + 1] = make_int4(o[4], o[5], o[6], o[7]);
    const int32_t u = (int32_t)(8 * q);
#pragma unroll
    for (int k = 0; k < 8; k++) self += o[k] == u + k;
  }
  if (blockIdx.x == 0)
    for (int64_t u = 8 * n8 + threadIdx.x; u < n; u += BLOCK) {
      const int32_t o = (int32_t)(key(label[iperm[u]]) >> 32);
      cout[u] = o;
      self += o == (int32_t)u;
    }
  block_count_nowait(self, &dc->ctr[C_COUNT]);
}

// Caller-order canonical labels by windows of CANON_WIN caller ids, one
// workgroup each.  Inside a run of internal ids whose caller ids increase (a
// degree class), the members of a window are one contiguous internal range
// (the table T gives where each run enters each window), so the workgroup
// reads labels, keys and caller ids sequentially, run by run, assembles the
// window's labels in LDS and writes them out in caller order.  The label of
// the global step's SCC (its pivot, from *gkey = the step's pivot key, 0 for
// none) is the label of most members: its key is read once per workgroup
// instead of by every wave (one L2 line would serve all of them).  Counts the
// caller ids that are their component's label.  In a chain: nothing when it
// stopped (halt odd), and a finished chain's halt is cleared.
__global__ __launch_bounds__(BLOCK) void k_canon_window(DevCtl* __restrict__ dc, int32_t n, int32_t runs,
                                                        const int32_t* __restrict__ T,
                                                        const int32_t* __restrict__ label,
                                                        const unsigned long long* __restrict__ ckey,
                                                        const int32_t* __restrict__ perm,
                                                        const unsigned long long* __restrict__ gkey,
                                                        int32_t* __restrict__ cout, int gate,
                                                        int32_t epoch) {
  // gate >= 0: the chain's last gate (a trim loop of this epoch) runs here
  // first, in thread 0 of workgroup 0.  Each workgroup decides from the halt
  // its thread 0 reads; when this gate stops the chain, workgroups that read
  // before it may still write labels, which the host's continuation rewrites
  __shared__ int s_skip;
  if (threadIdx.x == 0) {
    if (gate >= 0 && blockIdx.x == 0) gate_body(dc, gate, GATE_TRIM, epoch, 0, 1);
    const int32_t h = dc->chain.halt;
    s_skip = h & 1;
    if (blockIdx.x == 0 && h == 2) dc->chain.halt = 0;
  }
  __syncthreads();
  if (s_skip) return;
  __shared__ int32_t s_buf[CANON_WIN];
  __shared__ int32_t s_pre[CANON_RUNS + 1], s_base[CANON_RUNS];
  __shared__ int32_t s_hot;
  __shared__ unsigned long long s_hk;
  const int64_t w = blockIdx.x;
  const int64_t U0 = w * CANON_WIN;
  if (threadIdx.x == 64) {  // (a thread of the second wave: the first loads the table)
    const unsigned long long g = *gkey;
    s_hot = g ? pivot_of_key(g) : -1;
    s_hk = g ? ckey[s_hot] : ~0ull;
  }
  if (threadIdx.x < runs) {
    const int32_t a = T[w * runs + threadIdx.x], b = T[(w + 1) * runs + threadIdx.x];
    s_base[threadIdx.x] = a;
    s_pre[threadIdx.x + 1] = b - a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_pre[0] = 0;
    for (int r = 0; r < runs; r++) s_pre[r + 1] += s_pre[r];
  }
  __syncthreads();
  const int32_t total = s_pre[runs];
  const int32_t hot = s_hot;
  const unsigned long long hk = s_hk;
  auto key = [&](int32_t l) -> unsigned long long {
    if (l == hot) return hk;
    return (uint32_t)l < (uint32_t)n ? ckey[l] : ~0ull;
  };
  // four entries per thread and step, each round of loads issued together
  for (int32_t j0 = threadIdx.x; j0 < total; j0 += 4 * BLOCK) {
    int32_t v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int32_t j = j0 + t * BLOCK;
      v[t] = -1;
      if (j < total) {
        int lo = 0, hi = runs;  // s_pre[lo] <= j < s_pre[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_pre[mid] <= j) lo = mid;
          else hi = mid;
        }
        v[t] = s_base[lo] + (j - s_pre[lo]);
      }
    }
    int32_t l[4], u[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      l[t] = v[t] >= 0 ? label[v[t]] : -1;
      u[t] = v[t] >= 0 ? perm[v[t]] : -1;
    }
    unsigned long long k[4];
#pragma unroll
    for (int t = 0; t < 4; t++) k[t] = key(l[t]);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int64_t pos = (int64_t)u[t] - U0;
      if (v[t] >= 0 && pos >= 0 && pos < total) s_buf[pos] = (int32_t)(k[t] >> 32);
    }
  }
  __syncthreads();
  unsigned long long self = 0;
  for (int32_t i = threadIdx.x; i < total; i += BLOCK) {
    const int32_t o = s_buf[i];
    cout[U0 + i] = o;
    self += o == (int32_t)(U0 + i);
  }
  block_count_nowait(self, &dc->ctr[C_COUNT]);
}

// ---- host launchers -------------------------------------------------------------

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

void launch_fill_i32(hipStream_t st, int32_t* p, int32_t v, int64_t n) {
  if (n > 0) LAUNCH(k_fill_i32, grid_for(n, 4096), st, p, v, n);
}
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, int64_t n) {
  if (n > 0) LAUNCH(k_fill_u32, grid_for(n, 4096), st, p, v, n);
}
void launch_iota(hipStream_t st, int32_t* p, int32_t n) {
  if (n > 0) LAUNCH(k_iota, grid_for(n, 4096), st, p, n);
}
void launch_ctl_init(hipStream_t st, DevCtl* dc, int what, int32_t live_sel, unsigned long long live_n,
                     long long stop, double bu_ratio) {
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, st, dc, what, live_sel, live_n, stop, bu_ratio);
}
void launch_trim1_step(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, int32_t* color,
                       int32_t* label, int32_t* L0, int32_t* L1, int fresh, long long stop, int32_t epoch,
                       unsigned long long* pkey, uint32_t* zbits, int64_t zwords) {
  if (fresh == 2) hipLaunchKernelGGL(k_fresh_init, dim3(1), dim3(64), 0, st, dc, G.n, pkey);
  LAUNCH(k_trim1_step, grid, st, dc, G, color, label, L0, L1, fresh, stop, epoch, pkey, zbits, zwords);
}
void launch_seed_single(hipStream_t st, DevCtl* dc, int32_t n, const LevelDir& F, const LevelDir& B,
                        const unsigned long long* pkey, CanonKeys K) {
  hipLaunchKernelGGL(k_seed_single, dim3(1), dim3(64), 0, st, dc, n, F, B, pkey, K);
}
void launch_compact(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* color, int32_t* L0,
                    int32_t* L1) {
  LAUNCH(k_compact, grid, st, dc, color, L0, L1);
}
void launch_trim2(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, int32_t* color,
                  int32_t* label, const int32_t* L0, const int32_t* L1) {
  LAUNCH(k_trim2, grid, st, dc, G, color, label, L0, L1);
}
void launch_reset_slots(hipStream_t st, unsigned grid, const DevCtl* dc, int32_t n,
                        const int32_t* color, unsigned long long* pkey, int32_t* rep,
                        const int32_t* L0, const int32_t* L1, int single) {
  LAUNCH(k_reset_slots, single ? 1u : grid, st, dc, n, color, pkey, rep, L0, L1, single);
}
void launch_pivot_key(hipStream_t st, unsigned grid, const DevCtl* dc, GraphPtrs G,
                      const int32_t* color, int policy, const uint32_t* scores,
                      unsigned long long* pkey, const int32_t* L0, const int32_t* L1, uint32_t* zbits,
                      int64_t zwords) {
  LAUNCH(k_pivot_key, grid, st, dc, G, color, policy, scores, pkey, L0, L1, zbits, zwords);
}
void launch_reset(hipStream_t st, DevCtl* dc, int32_t n, int32_t* color, int32_t* label, int32_t* live0) {
  LAUNCH(k_reset, grid_for(n > 0 ? n : 1, 4096), st, dc, n, color, label, live0);
}
void launch_fb_setup(hipStream_t st, DevCtl* dc, int32_t n, unsigned long long* pkey, int32_t* rep, int single,
                     double bu_ratio) {
  hipLaunchKernelGGL(k_fb_setup, dim3(1), dim3(64), 0, st, dc, n, pkey, rep, single, bu_ratio);
}
void launch_chain_gate(hipStream_t st, DevCtl* dc, int gate, int kind, int32_t epoch, int single, int last,
                       int zero_ctr) {
  hipLaunchKernelGGL(k_chain_gate, dim3(1), dim3(64), 0, st, dc, gate, kind, epoch, single, last, zero_ctr);
}
void launch_chain_gate_fb(hipStream_t st, DevCtl* dc, int32_t epoch, int last, int32_t n, unsigned long long* pkey,
                          int32_t* rep, double bu_ratio, const LevelDir& F, const LevelDir& B, CanonKeys K) {
  hipLaunchKernelGGL(k_chain_gate_fb, dim3(1), dim3(64), 0, st, dc, epoch, last, n, pkey, rep, bu_ratio, F, B, K);
}
void launch_force_pivot(hipStream_t st, int32_t p, int32_t n, const int32_t* color,
                        unsigned long long* pkey) {
  LAUNCH(k_force_pivot, 1, st, p, n, color, pkey);
}
void launch_seed_list(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* off, const int32_t* offe,
                      const int32_t* color, const int32_t* seeds, const unsigned long long* ns,
                      uint32_t* V, int32_t* q0) {
  LAUNCH(k_seed_list, grid, st, dc, off, offe, color, seeds, ns, V, q0);
}
void launch_seed_pivots2(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int multi, const LevelDir& F,
                         const LevelDir& B, const int32_t* color, const unsigned long long* pkey,
                         const int32_t* L0, const int32_t* L1, CanonKeys K) {
  LAUNCH(k_seed_pivots2, grid, st, dc, n, multi, F, B, color, pkey, L0, L1, K);
}
void launch_level(hipStream_t st, unsigned grid, DevCtl* dc, int multi, const LevelDir& D,
                  const int32_t* color, const int32_t* L0, const int32_t* L1, int32_t max_chunks, int light) {
#define LEVEL(L) LAUNCH(k_level<L>, grid, st, dc, multi, D, color, L0, L1, max_chunks)
  switch (light) {
    case 4: LEVEL(4); break;
    case 8: LEVEL(8); break;
    case 32: LEVEL(32); break;
    case 64: LEVEL(64); break;
    default: LEVEL(16); break;
  }
#undef LEVEL
}
void launch_level2(hipStream_t st, unsigned grid, DevCtl* dc, int multi, const LevelDir& F, const LevelDir& B,
                   const int32_t* color, const int32_t* L0, const int32_t* L1, int32_t max_chunks, int light) {
  grid = (grid + 15u) & ~15u;
#define LEVEL(L) LAUNCH(k_level2<L>, grid, st, dc, multi, F, B, color, L0, L1, max_chunks)
  switch (light) {
    case 4: LEVEL(4); break;
    case 8: LEVEL(8); break;
    case 32: LEVEL(32); break;
    case 64: LEVEL(64); break;
    default: LEVEL(16); break;
  }
#undef LEVEL
}
void launch_fb_stats(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, const uint32_t* vf,
                     const uint32_t* vb, const int32_t* L0, const int32_t* L1) {
  LAUNCH(k_fb_stats, grid, st, dc, G, vf, vb, L0, L1);
}
void launch_classify(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int32_t* color,
                     int32_t* label, const uint32_t* vf, const uint32_t* vb,
                     const unsigned long long* pkey, int32_t* rep, int32_t* L0, int32_t* L1, CanonKeys K,
                     int single) {
  LAUNCH(k_classify, grid, st, dc, n, color, label, vf, vb, pkey, rep, L0, L1, K, single);
}
void launch_recolor(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int32_t* color,
                    const uint32_t* vf, const uint32_t* vb, const int32_t* rep, const int32_t* L0,
                    const int32_t* L1, int gate1) {
  LAUNCH(k_recolor, grid, st, dc, n, color, vf, vb, rep, L0, L1, gate1);
}
void launch_cw_init(hipStream_t st, int32_t n, const int32_t* begin, const int32_t* bend, const int32_t* color,
                    unsigned long long* cw) {
  if (n > 0) LAUNCH(k_cw_init, grid_for(n, 4096), st, n, begin, bend, color, cw);
}
void launch_wcc_pull(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* begin, const int32_t* bend,
                     const int32_t* adj, unsigned long long* cw, int32_t* q0, const int32_t* L0, const int32_t* L1,
                     int32_t* fe, int32_t* fout) {
  LAUNCH(k_wcc_pull, grid, st, dc, begin, bend, adj, cw, q0, L0, L1, fe, fout);
}
void launch_wcc_push(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* begin, const int32_t* bend,
                     const int32_t* adj, const int32_t* rbegin, const int32_t* rend, const int32_t* radj,
                     unsigned long long* cw, int32_t* q0, int32_t* q1, uint32_t* B0, uint32_t* B1,
                     const int32_t* L0, const int32_t* L1) {
  LAUNCH(k_wcc_push, grid, st, dc, begin, bend, adj, rbegin, rend, radj, cw, q0, q1, B0, B1, L0, L1);
}
void launch_wcc_color(hipStream_t st, unsigned grid, DevCtl* dc, int32_t* color, const unsigned long long* cw,
                      int32_t* W, int seed, const int32_t* off, const int32_t* offe, uint32_t* V, int32_t* q0,
                      const int32_t* L0, const int32_t* L1, CanonKeys K) {
  LAUNCH(k_wcc_color, grid, st, dc, color, cw, W, seed, off, offe, V, q0, L0, L1, K);
}
void launch_extract(hipStream_t st, unsigned grid, DevCtl* dc, int32_t* color, int32_t* label,
                    const uint32_t* V, int32_t* L0, int32_t* L1, CanonKeys K) {
  LAUNCH(k_extract, grid, st, dc, color, label, V, L0, L1, K);
}
void launch_canon_window(hipStream_t st, DevCtl* dc, int32_t n, int32_t runs, int32_t nwin, const int32_t* table,
                         const int32_t* label, const unsigned long long* ckey, const int32_t* perm,
                         const unsigned long long* gkey, int32_t* cout, int gate, int32_t epoch) {
  if (n <= 0 || nwin <= 0) return;
  LAUNCH(k_canon_window, nwin, st, dc, n, runs, table, label, ckey, perm, gkey, cout, gate, epoch);
}
void launch_canon_caller(hipStream_t st, DevCtl* dc, int32_t n, const int32_t* label,
                         const unsigned long long* ckey, const int32_t* iperm, int32_t* cout, bool zero_ctr) {
  if (zero_ctr) hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, st, dc, 1, 0, 0ull, 0ll, 0.0);
  if (n <= 0) return;
  LAUNCH(k_canon_caller, grid_for(n / 8 > 0 ? n / 8 : 1, 2048), st, dc, n, label, ckey, iperm, cout);
}
void launch_canon(hipStream_t st, DevCtl* dc, int32_t n, const int32_t* label, const unsigned long long* ckey,
                  const int32_t* perm, int32_t* out, int32_t* cout, bool zero_ctr) {
  if (zero_ctr) hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, st, dc, 1, 0, 0ull, 0ll, 0.0);
  if (n <= 0) return;
  if (perm)
    LAUNCH(k_canon_keys<true>, grid_for((n + 3) / 4, 1024), st, dc, n, label, ckey, perm, out, cout);
  else
    LAUNCH(k_canon_keys<false>, grid_for((n + 3) / 4, 1024), st, dc, n, label, ckey, perm, out, cout);
}

#undef LAUNCH
}  // namespace sccmi

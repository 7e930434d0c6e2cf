// Host launchers of the SCC kernels (defined in kernels.hip).  List-processing
// kernels read the current vertex list and its length from the device control
// block, so the host only chooses the grid (an upper bound on the list length).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace sccmi {

// Row v of the forward CSR is adj[begin[v] .. end[v]); the reverse CSR alike.
struct GraphPtrs {
  const int32_t *begin, *end, *adj, *rbegin, *rend, *radj;
  int32_t n;
  int32_t maxdeg;
  // canonical-label keys (see CanonKeys)
  const int32_t* perm;
  unsigned long long* ckey;
};

// Canonical-label keys.  key(v) = caller id of v << 32 | v (the caller id is
// perm[v] in the degree vertex layout, v itself otherwise).  For every label l
// in use, ckey[l] holds the minimum key over the vertices labelled l: the phase
// that gives a label sets it (trims store their vertex's key, a step's pivots
// and a colouring's roots start their slot with their own key, the members'
// keys are min-reduced in), so the canonical label of v is the low half of
// ckey[label[v]] (internal id of the member with the smallest caller id, the
// smallest id in the input layout) and its caller-id label the high half.
struct CanonKeys {
  const int32_t* perm;
  unsigned long long* ckey;
};
__device__ __forceinline__ unsigned long long canon_key(const int32_t* __restrict__ perm, int32_t v) {
  const uint32_t cid = perm ? (uint32_t)perm[v] : (uint32_t)v;
  return ((unsigned long long)cid << 32) | (uint32_t)v;
}

// grid size for a list of n items (grid-stride beyond the cap)
inline unsigned grid_for(int64_t n, unsigned cap = 2048) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

void launch_fill_i32(hipStream_t st, int32_t* p, int32_t v, int64_t n);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, int64_t n);
void launch_iota(hipStream_t st, int32_t* p, int32_t n);
// what: bit 0 counters, 1 live list, 2 loop control, 3 level control, 4 tuning
// defaults, 5 the second level control (backward search run alongside)
void launch_ctl_init(hipStream_t st, DevCtl* dc, int what, int32_t live_sel, unsigned long long live_n,
                     long long stop, double bu_ratio);
// one sweep of trim loop number epoch (a new number starts a loop)
// fresh: 0 a sweep over the live list, 1 the first sweep after a reset (row
// lengths only), 2 the same with the reset fused in (colours / labels of every
// vertex written, the global pivot key into pkey[n], zwords words at zbits cleared)
void launch_trim1_step(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, int32_t* color,
                       int32_t* label, int32_t* L0, int32_t* L1, int fresh, long long stop, int32_t epoch,
                       unsigned long long* pkey, uint32_t* zbits, int64_t zwords);
// Seeds of the global step (one partition) from pkey[n]
void launch_seed_single(hipStream_t st, DevCtl* dc, int32_t n, const LevelDir& F, const LevelDir& B,
                        const unsigned long long* pkey, CanonKeys K);
void launch_compact(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* color, int32_t* L0,
                    int32_t* L1);
void launch_trim2(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, int32_t* color,
                  int32_t* label, const int32_t* L0, const int32_t* L1);
void launch_reset_slots(hipStream_t st, unsigned grid, const DevCtl* dc, int32_t n,
                        const int32_t* color, unsigned long long* pkey, int32_t* rep,
                        const int32_t* L0, const int32_t* L1, int single);
// per-partition pivot keys; also zeroes zbits[0, zwords) (the step's reach bitmaps)
void launch_pivot_key(hipStream_t st, unsigned grid, const DevCtl* dc, GraphPtrs G,
                      const int32_t* color, int policy, const uint32_t* scores,
                      unsigned long long* pkey, const int32_t* L0, const int32_t* L1, uint32_t* zbits,
                      int64_t zwords);
// fresh state: colours -1, labels -1, live list 0..n-1, counters zeroed, live = all
void launch_reset(hipStream_t st, DevCtl* dc, int32_t n, int32_t* color, int32_t* label, int32_t* live0);
// forward-backward step set-up: counters zeroed, both search controls, and with
// single the slot of colour -1 reset
void launch_fb_setup(hipStream_t st, DevCtl* dc, int32_t n, unsigned long long* pkey, int32_t* rep, int single,
                     double bu_ratio);
// A gate of a queued chain (ChainCtl): kind GATE_TRIM (a trim loop of the given
// epoch; single: it must have ended after one sweep), GATE_SEARCH, GATE_TRIM2.
enum : int { GATE_TRIM = 0, GATE_SEARCH = 1, GATE_TRIM2 = 2 };
// zero_ctr: then zero the counters (for the canonicalisation queued next).
void launch_chain_gate(hipStream_t st, DevCtl* dc, int gate, int kind, int32_t epoch, int single, int last,
                       int zero_ctr);
// Gate 0 with, when the chain goes on, the global step's set-up and seeding.
void launch_chain_gate_fb(hipStream_t st, DevCtl* dc, int32_t epoch, int last, int32_t n, unsigned long long* pkey,
                          int32_t* rep, double bu_ratio, const LevelDir& F, const LevelDir& B, CanonKeys K);
void launch_force_pivot(hipStream_t st, int32_t p, int32_t n, const int32_t* color,
                        unsigned long long* pkey);
void launch_seed_list(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* off, const int32_t* offe,
                      const int32_t* color, const int32_t* seeds, const unsigned long long* ns,
                      uint32_t* V, int32_t* q0);
void launch_seed_pivots2(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int multi, const LevelDir& F,
                         const LevelDir& B, const int32_t* color, const unsigned long long* pkey,
                         const int32_t* L0, const int32_t* L1, CanonKeys K);
void launch_level(hipStream_t st, unsigned grid, DevCtl* dc, int multi, const LevelDir& D,
                  const int32_t* color, const int32_t* L0, const int32_t* L1, int32_t max_chunks, int light);
// both searches of a step, one level each; grid = workgroups per search x 2,
// rounded up to a multiple of 16
void launch_level2(hipStream_t st, unsigned grid, DevCtl* dc, int multi, const LevelDir& F, const LevelDir& B,
                   const int32_t* color, const int32_t* L0, const int32_t* L1, int32_t max_chunks, int light);
void launch_fb_stats(hipStream_t st, unsigned grid, DevCtl* dc, GraphPtrs G, const uint32_t* vf,
                     const uint32_t* vb, const int32_t* L0, const int32_t* L1);
void launch_classify(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int32_t* color,
                     int32_t* label, const uint32_t* vf, const uint32_t* vb,
                     const unsigned long long* pkey, int32_t* rep, int32_t* L0, int32_t* L1, CanonKeys K,
                     int single);
void launch_recolor(hipStream_t st, unsigned grid, DevCtl* dc, int32_t n, int32_t* color,
                    const uint32_t* vf, const uint32_t* vb, const int32_t* rep, const int32_t* L0,
                    const int32_t* L1,
                    int gate1 = 0);
void launch_cw_init(hipStream_t st, int32_t n, const int32_t* begin, const int32_t* bend, const int32_t* color,
                    unsigned long long* cw);
void launch_wcc_pull(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* begin, const int32_t* bend,
                     const int32_t* adj, unsigned long long* cw, int32_t* q0, const int32_t* L0, const int32_t* L1,
                     int32_t* fe = nullptr, int32_t* fout = nullptr);
// later WCC iterations: pull sweeps over the live list while the frontier is
// large, frontier pushes over the reverse rows after
void launch_wcc_push(hipStream_t st, unsigned grid, DevCtl* dc, const int32_t* begin, const int32_t* bend,
                     const int32_t* adj, const int32_t* rbegin, const int32_t* rend, const int32_t* radj,
                     unsigned long long* cw, int32_t* q0, int32_t* q1, uint32_t* B0, uint32_t* B1,
                     const int32_t* L0, const int32_t* L1);
// colour = class root from the labels (from cw when given, storing them to W,
// else from W); seed: the roots also seed the extraction search
void launch_wcc_color(hipStream_t st, unsigned grid, DevCtl* dc, int32_t* color, const unsigned long long* cw,
                      int32_t* W, int seed, const int32_t* off, const int32_t* offe, uint32_t* V, int32_t* q0,
                      const int32_t* L0, const int32_t* L1, CanonKeys K);
void launch_extract(hipStream_t st, unsigned grid, DevCtl* dc, int32_t* color, int32_t* label,
                    const uint32_t* V, int32_t* L0, int32_t* L1, CanonKeys K);
// canonical labels into out, and (perm != NULL) the caller-id labels into
// cout[perm[v]], from the keys the phases left (CanonKeys); zeroes the
// counters first and leaves the SCC count in dc->ctr[C_COUNT]
// Caller-order canonical labels by caller-id windows (degree layout with a
// window table, scc_ctx::cwin); counts into ctr[C_COUNT] like launch_canon.
// gkey: the global step's pivot key (its SCC's label is read once per workgroup), 0 for none.
void launch_canon_window(hipStream_t st, DevCtl* dc, int32_t n, int32_t runs, int32_t nwin, const int32_t* table,
                         const int32_t* label, const unsigned long long* ckey, const int32_t* perm,
                         const unsigned long long* gkey, int32_t* cout, int gate = -1, int32_t epoch = 0);
// Caller-order canonical labels only (degree vertex layout): cout[u] from the
// label of iperm[u]; the internal-order labels are left to launch_canon.
void launch_canon_caller(hipStream_t st, DevCtl* dc, int32_t n, const int32_t* label,
                         const unsigned long long* ckey, const int32_t* iperm, int32_t* cout, bool zero_ctr);
void launch_canon(hipStream_t st, DevCtl* dc, int32_t n, const int32_t* label, const unsigned long long* ckey,
                  const int32_t* perm, int32_t* out, int32_t* cout,
                  bool zero_ctr = true);

}  // namespace sccmi

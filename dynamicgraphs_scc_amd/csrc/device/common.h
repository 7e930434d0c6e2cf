// Shared device helpers, constants and the context definition of libsccmi.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <string>
#include <vector>

#include "sccmi.h"

namespace sccmi {

constexpr int32_t DONE = -2;   // colour of a vertex whose SCC is known
constexpr int32_t UNSET = -1;  // initial colour
constexpr int BLOCK = 256;     // threads per workgroup (4 wave64s)
constexpr int WAVES = BLOCK / 64;
constexpr int LIGHT = 16;      // trim row scans: row entries a lane scans alone before the wave helps
constexpr int LEVEL_LIGHT = 32; // default of the same for search levels (tunable; planted graphs'
                               // rows of ~10-30 entries are then scanned by their lane alone)
constexpr int HUGE_ROW = 8192; // default: rows at least this long are expanded by many blocks
constexpr int GRID_CAP = 640;  // default: workgroups of a sweep launch (2.5 per CU; more only adds
                               // per-workgroup set-up and flushes to the short levels)
constexpr int WCC_DENSE = 2;   // default: label propagation pulls over the whole live set while
                               // the frontier holds more than half of it

// Direction switches of searches that serve many partitions at once (partition
// rounds, colouring-round searches, a batch's first step): bottom-up levels
// there also scan candidates of partitions whose search has ended, so they
// enter bottom-up later and leave it sooner than a single search.  The
// multi-partition thresholds are bu_alpha * MULTI_BU_SCALE and
// td_beta * MULTI_TD_SCALE (4 and 8 at the defaults 14 and 24; measured on the
// 512-graph batch: 179 -> 152 ms, planted and R-MAT graphs unchanged).
constexpr double MULTI_BU_SCALE = 4.0 / 14.0;
constexpr float MULTI_TD_SCALE = 1.0f / 3.0f;
constexpr int CANON_WIN = 4096;  // caller ids per workgroup of k_canon_window
constexpr int CANON_RUNS = 64;   // runs of increasing caller ids the window table can hold
constexpr int32_t WCC_TAG = 0x20000000, WCC_MASK = 0x1FFFFFFF;  // src/scc_weak.cc:57-62

// Counter slots of DevCtl::ctr (unsigned long long).
enum Ctr : int {
  C_LIST_OUT = 0,   // compaction output count (live / survivors)
  C_TRIMMED = 1,    // vertices removed by trim2
  C_Q_OUT = 2,      // seeds / frontier count outside the level loop
  C_CAND_OUT = 3,
  C_SCC = 4,        // SCC members found by classify / extract
  C_CHANGED = 5,
  C_ROOTS = 6,      // WCC roots
  C_AUX_T = 7,      // second seed list (dynamic updates)
  C_AUX_SEL = 8,    // selected-count slot for library selects (int32)
  C_COUNT = 9,      // SCC count from canonicalisation
  C_F = 10,         // |F|  forward-reached vertices of a step
  C_B = 11,         // |B|  backward-reached
  C_EF = 12,        // sum of out-degrees over F
  C_EB = 13,        // sum of in-degrees over B
  C_FB = 14,        // |F n B|
  C_PIVOT = 15,     // a pivot of the last seeded search (the pivot, for one partition)
  C_NUM = 16
};

// ---- device-resident control ------------------------------------------------
// Iterative phases run as a sequence of launches of one kernel.  Each launch
// reads its input sizes from this block; the last workgroup to finish a launch
// (counted with an atomic) writes the next launch's inputs and the `done` flag,
// so the host queues launches in batches and synchronises only to learn that a
// phase has finished.  Launches after `done` exit at once.

// The live list a kernel reads (LIVE_IN) and the one it compacts into.
#define LIVE_IN(dc, L0, L1) ((dc)->live.sel ? (L1) : (L0))
#define LIVE_OUT(dc, L0, L1) ((dc)->live.sel ? (L0) : (L1))

struct LiveCtl {  // the current vertex list: live[sel][0, n)
  int32_t sel;
  int32_t pad;
  unsigned long long n;
};

struct LoopCtl {  // trim sweeps and WCC iterations
  int32_t done;
  int32_t iters;
  uint32_t blocks_done;
  int32_t epoch;  // trim loop the state belongs to (a sweep of a new loop starts it afresh)
  long long stop;               // trim: stop when a sweep trims <= stop
  unsigned long long acc0, acc1;  // per-launch accumulators
  unsigned long long total0;      // sum of acc0 over launches
};

enum : int32_t { MODE_TD = 0, MODE_BU = 1 };

struct LevelCtl {  // one reachability search (all partitions at once)
  int32_t done;
  int32_t mode;     // MODE_TD / MODE_BU for the next launch
  int32_t q_sel;    // frontier q[q_sel] of length nq
  int32_t c_sel;    // bottom-up candidates: -1 = the live list, else cand[c_sel] (nc)
  int32_t h_sel;    // rows deferred by the last top-down level: huge[h_sel] (nh)
  int32_t levels;
  uint32_t blocks_done;
  int32_t pad;
  unsigned long long nq, nh, nc, mf, prev_nq;
  unsigned long long out_q, out_c, out_h, out_edges;
  unsigned long long visited;  // vertices admitted over all levels (seeds included)
  double bu_edges_per_cand;    // Beamer's switch: BU when m_f > n_cand * this
};

// Launch-shape knobs the autotuner sets (scc_tuning); kernels read them here.
struct Tune {
  int32_t light;     // lane-alone prefix of a row scan
  int32_t huge_row;  // rows split over the grid in top-down levels
  float td_beta;     // bottom-up -> top-down when the frontier < live / td_beta
  float bu_alpha;    // top-down -> bottom-up when frontier edges > candidates' edges / bu_alpha
  int32_t grid_cap;  // workgroups per sweep launch
  int32_t wcc_dense; // propagation iterations pull while the frontier > live / wcc_dense
};

// A decomposition queued as one chain of launches (no host read-back between
// its phases): a gate launch after each phase checks that the phase finished
// within the launches queued for it.  When it did not (or when the chain's
// last phase leaves vertices live), the gate stops the chain: halt = 2 * gate
// + 1, and every later guarded launch returns at once, so the state is the one
// the gate saw and the host resumes from there.  halt = 2: no vertex is live,
// so the later phases are skipped but canonicalisation runs.
constexpr int CHAIN_GATES = 6;
struct ChainCtl {
  int32_t halt;
  int32_t gate;  // last gate reached
  int32_t ok[CHAIN_GATES];         // the gate's phase had finished
  int32_t iters[CHAIN_GATES];      // trim loop: sweeps; search: levels (fw + bw)
  unsigned long long live[CHAIN_GATES];   // live vertices at the gate
  unsigned long long count[CHAIN_GATES];  // trimmed (trim loops, trim2) or SCC members (search)
  unsigned long long pivot;        // the global step's pivot
};

struct DevCtl {
  Tune tune;
  LiveCtl live;
  LoopCtl loop;
  LevelCtl lv;
  uint32_t blocks_done;  // (unused; completion counts live in done[])
  int32_t pad;
  unsigned long long ctr[C_NUM];
  uint32_t done[9 * 16];  // last_block2 slots, 64 B apart
  // per level of the last reachability search: mode, input size, admitted, their edges
  unsigned long long lvlog[64][4];
  // the backward search when both directions run in the same launches
  LevelCtl lv2;
  uint32_t done2[9 * 16];
  unsigned long long lvlog2[64][4];
  ChainCtl chain;
};

// Guard of a launch that may be queued in a chain: true when it must return.
__device__ __forceinline__ bool chain_halted(const DevCtl* __restrict__ dc) { return dc->chain.halt != 0; }

// The arrays one reachability direction reads and writes: rows to expand
// (off/offe/nb), the opposite rows a bottom-up level scans (roff/roffe/rnb), the
// visited bitmap, and ping-pong frontier, candidate and deferred-row lists.
struct LevelDir {
  const int32_t *off, *offe, *nb, *roff, *roffe, *rnb;
  const int32_t* rhead;  // first entry of each opposite row (rnb[roff[v]]; -1 for an empty row), or null
  uint32_t* V;
  int32_t *q0, *q1, *c0, *c1, *h0, *h1;
};

// ---- wave64 primitives -----------------------------------------------------

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Number of set bits of m below this lane.
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Append `value` for every lane with `flag` to out[] at a base reserved with one
// atomic per wave.  Must be reached by all lanes of the wave.
__device__ __forceinline__ void wave_append(bool flag, int32_t value, int32_t* out,
                                            unsigned long long* count) {
  uint64_t m = __ballot(flag);
  if (m == 0) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(count, (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (flag) out[base + lanes_below(m)] = value;
}

// Block-wide append with one atomic per block (order inside the block kept).
// Must be reached by all threads of the block.
__device__ __forceinline__ void block_append(bool flag, int32_t value, int32_t* out,
                                             unsigned long long* count) {
  __shared__ int s_cnt[WAVES];
  __shared__ unsigned long long s_base;
  uint64_t m = __ballot(flag);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) s_cnt[w] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) {
      int c = s_cnt[i];
      s_cnt[i] = tot;
      tot += c;
    }
    s_base = tot ? atomicAdd(count, (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  if (flag) out[s_base + s_cnt[w] + lanes_below(m)] = value;
  __syncthreads();
}

// Block-wide sum added to *count with one atomic.  All threads must call.  The
// atomic's return value is kept, so thread 0 waits for it to be performed
// before it goes on (before the workgroup's completion count in last_block2).
__device__ __forceinline__ void block_count(unsigned long long x, unsigned long long* count) {
  __shared__ unsigned long long s_sum[WAVES];
  __shared__ volatile unsigned long long s_sink;
  x = wave_sum_u64(x);
  if (lane_id() == 0) s_sum[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) t += s_sum[i];
    if (t) s_sink = atomicAdd(count, t);
  }
  __syncthreads();
}

// block_count for totals that no completion count depends on: the atomic's
// result is not waited for.
__device__ __forceinline__ void block_count_nowait(unsigned long long x, unsigned long long* count) {
  __shared__ unsigned long long s_sum[WAVES];
  x = wave_sum_u64(x);
  if (lane_id() == 0) s_sum[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) t += s_sum[i];
    if (t) atomicAdd(count, t);
  }
}

// LDS-staged list output.  Lanes push entries into a CAP-entry workgroup buffer
// (reserved with LDS atomics); the buffer goes to the global list with one
// global atomic per flush.  A wave whose reservation does not fit writes its
// entries straight to the global list, and the buffer's valid prefix ends at
// the first failed reservation.  ID separates buffers of one kernel.
template <int ID, int CAP = 2048>
struct LdsOut {
  __device__ static int32_t* buf() {
    __shared__ int32_t b[CAP];
    return b;
  }
  __device__ static int* cnt() {
    __shared__ int c[2];  // [0] reserved, [1] valid limit
    return c;
  }
  __device__ static unsigned long long* gbase() {
    __shared__ unsigned long long g;
    return &g;
  }
  // all threads, at kernel start
  __device__ static void init() {
    if (threadIdx.x == 0) {
      cnt()[0] = 0;
      cnt()[1] = CAP;
    }
    __syncthreads();
  }
  // all threads: copy the buffer out (one global atomic) and empty it
  __device__ static void flush(int32_t* out, unsigned long long* gcount) {
    __syncthreads();
    const int c = min(cnt()[0], cnt()[1]);
    if (c > 0) {
      if (threadIdx.x == 0) *gbase() = atomicAdd(gcount, (unsigned long long)c);
      __syncthreads();
      const unsigned long long b = *gbase();
      int32_t* B = buf();
      for (int i = threadIdx.x; i < c; i += BLOCK) out[b + i] = B[i];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      cnt()[0] = 0;
      cnt()[1] = CAP;
    }
    __syncthreads();
  }
  // all threads: flush when fewer than `room` free entries remain
  __device__ static void maybe_flush(int32_t* out, unsigned long long* gcount, int room) {
    __syncthreads();
    const int c = cnt()[0];
    if (c > CAP - room || cnt()[1] < CAP) flush(out, gcount);
  }
  // all lanes of a wave
  __device__ static void push(bool flag, int32_t v, int32_t* out, unsigned long long* gcount) {
    const uint64_t m = __ballot(flag);
    if (m == 0) return;
    const int k = __popcll(m);
    int base = 0;
    if (lane_id() == 0) {
      base = atomicAdd(&cnt()[0], k);
      if (base + k > CAP) {
        atomicMin(&cnt()[1], base);
        base = -1 - (int)atomicAdd(gcount, (unsigned long long)k);  // global, encoded < 0
      }
    }
    base = __shfl(base, 0);
    if (flag) {
      const int r = lanes_below(m);
      if (base >= 0) buf()[base + r] = v;
      else out[(unsigned long long)(-1 - base) + r] = v;
    }
  }
};

// Completion count of a launch in two levels: the workgroups of each XCD
// (blockIdx % 8) count on their own slot, the last of each slot on the top
// slot, so no address sees more than grid/8 atomics.  Counters reset
// themselves.  True in every thread of the workgroup that finishes last.
// bid/nb: the workgroup's index and count within its share of the grid (a
// launch that runs two searches gives each half of it).
//
// Ordering.  What the last workgroup reads of the other workgroups' work is
// only device-scope atomic counters (list lengths, sums; read with ld_acq),
// never plain stores: lists and per-vertex arrays are read by the next launch,
// after the kernel boundary.  Every such counter update is issued by a thread
// that waits for the atomic's returned value (block_count, LdsOut, the level
// kernels' totals) before the workgroup's barrier and its completion count, so
// it has been performed at the memory-side coherence point before the count is
// issued; the counts are device-scope atomics performed at the same point, and
// the last workgroup issues its counter loads after its count has returned,
// behind an agent-scope acquire fence.  The counts themselves stay relaxed: an
// agent-scope release per workgroup (the L2 writeback of every XCD's dirty
// lines, per workgroup) took an R-MAT-22 decomposition from 0.345 to 0.441 ms,
// and acquire-release counts to 0.488 ms, for data no consumer here needs.
__device__ __forceinline__ bool last_block2(uint32_t* ctr, unsigned bid, unsigned nb) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned x = bid & 7u;
    const unsigned in_slot = (nb > x) ? (nb - x + 7u) / 8u : 0u;
    int last = 0;
    const unsigned prev = atomicAdd(&ctr[(1 + x) * 16], 1u);
    if (prev == in_slot - 1) {
      atomicExch(&ctr[(1 + x) * 16], 0u);
      const unsigned nslots = nb < 8 ? nb : 8;
      const unsigned p2 = atomicAdd(&ctr[0], 1u);
      if (p2 == nslots - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        atomicExch(&ctr[0], 0u);
        last = 1;
      }
    }
    s_last = last;
  }
  __syncthreads();
  return s_last;
}
__device__ __forceinline__ bool last_block2(uint32_t* ctr) { return last_block2(ctr, blockIdx.x, gridDim.x); }

// atomicMax of `val` into target[slot] for every lane with act, after the lanes
// of the wave that share a slot have reduced among themselves (one atomic per
// distinct slot per wave; skipped when the slot already holds a larger value).
__device__ __forceinline__ void grouped_max_u64(bool act, int32_t slot, unsigned long long val,
                                                unsigned long long* target) {
  const int lane = lane_id();
  bool pend = act;
  uint64_t m = __ballot(pend);
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int32_t S = __shfl(slot, leader);
    const bool mine = pend && slot == S;
    unsigned long long x = mine ? val : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long y = __shfl_xor(x, o);
      x = y > x ? y : x;
    }
    if (lane == leader && x > target[S]) atomicMax(&target[S], x);
    if (mine) pend = false;
    m = __ballot(pend);
  }
}

// atomicMin counterpart for int32 slots.
__device__ __forceinline__ void grouped_min_i32(bool act, int64_t slot, int32_t val, int32_t* target) {
  const int lane = lane_id();
  bool pend = act;
  uint64_t m = __ballot(pend);
  while (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int64_t S = __shfl(slot, leader);
    const bool mine = pend && slot == S;
    int32_t x = mine ? val : INT32_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    if (lane == leader && x < target[S]) atomicMin(&target[S], x);
    if (mine) pend = false;
    m = __ballot(pend);
  }
}

// Agent-scope acquire load (counters other workgroups update with atomics).
__device__ __forceinline__ unsigned long long ld_acq(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool bit_test(const uint32_t* bm, int32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Slot of a partition colour in the per-partition arrays: colours are vertex
// ids (a member of the partition) or -1 for the initial partition (slot n).
__device__ __forceinline__ int32_t slot_of(int32_t c, int32_t n) { return c >= 0 ? c : n; }

// Pivot key: larger wins.  High bits carry the score, low 31 bits (2^31-1-v)
// break ties toward the smaller id.
__device__ __forceinline__ unsigned long long pivot_key(unsigned long long score, int32_t v) {
  return (score << 31) | (unsigned long long)(0x7FFFFFFF - v);
}
__device__ __forceinline__ int32_t pivot_of_key(unsigned long long k) {
  return 0x7FFFFFFF - (int32_t)(k & 0x7FFFFFFFull);
}

}  // namespace sccmi

// ---- the context -------------------------------------------------------------

struct scc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;  // 0 = none
  uint64_t bytes = 0;
  std::string err;
  std::atomic<int> cancel{0};
  int fault_after = -1;
  bool lost = false;
  bool kernel_timing = false;

  // graph (device)
  int32_t n = 0;
  int64_t m = 0;
  int64_t m_cap = 0;  // edge capacity of adj/radj
  int32_t *begin = nullptr, *adj = nullptr, *rbegin = nullptr, *radj = nullptr;
  // Row ends.  Uploaded graphs are packed (row v ends where v+1 begins) and
  // leave these NULL; the update engine keeps rows with spare room behind them
  // (see dynamic.hip) and sets them.
  int32_t *end = nullptr, *rend = nullptr;
  // Traversal copies: the same rows, each ordered by decreasing degree of the
  // neighbour (ties by id), so scans that stop at the first live / reached
  // neighbour meet hubs first.  NULL when absent (after updates).
  int32_t *tadj = nullptr, *tradj = nullptr;
  // first entry of every traversal row (out, in; -1 when empty), with the copies
  int32_t *head_out = nullptr, *head_in = nullptr;

  // Vertex layout (layout.hip): with SCC_LAYOUT_DEGREE requested at upload the
  // device graph is relabelled; perm[i] = caller id of internal vertex i,
  // iperm[v] = internal id of caller vertex v (host copies beside them), and
  // map_a / map_b are n + 1 ints of staging for mapped downloads.
  int32_t layout_req = 0;
  bool relabeled = false;
  int32_t *perm = nullptr, *iperm = nullptr, *map_a = nullptr, *map_b = nullptr;
  // canonical labels in caller ids (caller-indexed), written with canon[]
  int32_t* canon_caller = nullptr;
  std::vector<int32_t> h_perm, h_iperm;
  // caller-id windows of the canonical labels (k_canon_window): for window w
  // of CANON_WIN caller ids and run r (a maximal internal range whose caller
  // ids increase: a degree class), cwin[w * cwin_runs + r] = the first
  // internal id of run r whose caller id is in window w or later; null when
  // the order has more than CANON_RUNS runs
  int32_t* cwin = nullptr;
  int32_t cwin_runs = 0, cwin_nwin = 0;

  // per-vertex state
  int32_t *color = nullptr, *label = nullptr, *canon = nullptr, *W = nullptr;
  unsigned long long* cw = nullptr;  // WCC iterations: colour << 32 | label
  sccmi::Tune tune = {sccmi::LEVEL_LIGHT, sccmi::HUGE_ROW, 24.f, 14.f, sccmi::GRID_CAP, sccmi::WCC_DENSE};  // host copy of dc->tune

  // batched graphs: vertex ranges [boff[g], boff[g+1]) (scc_batch_set)
  int32_t batch_k = 0;
  int32_t* d_boff = nullptr;
  std::vector<int32_t> h_boff;
  int32_t* live[2] = {nullptr, nullptr};
  int cur = 0;            // live[cur] is the current live list (host copy of dc->live)
  int64_t n_live = 0;     // its length (host copy, valid at phase boundaries)
  int32_t* q[2] = {nullptr, nullptr};
  int32_t* cand[2] = {nullptr, nullptr};
  int32_t *huge = nullptr, *huge2 = nullptr;  // deferred huge-row frontier vertices (ping-pong)
  // the backward search's lists when both directions run together
  int32_t* bq[2] = {nullptr, nullptr};
  int32_t* bcand[2] = {nullptr, nullptr};
  int32_t* bhuge[2] = {nullptr, nullptr};
  int32_t last_dirs = 1;                  // last search: 1 forward, -1 backward, 2 both (level log)
  uint32_t *vf = nullptr, *vb = nullptr;  // reach bitmaps
  unsigned long long* pkey = nullptr;     // per-slot pivot key (n+1)
  unsigned long long* ckey = nullptr;     // per-label canonical key (n+1), see CanonKeys in kernels.h
  int32_t* rep = nullptr;                 // per-slot new-partition representatives 3*(n+1)
  int32_t* scratch_i32 = nullptr;         // n+1 ints for canonicalisation / misc
  sccmi::DevCtl* dc = nullptr;            // device control block
  sccmi::DevCtl* h_dc = nullptr;          // pinned host mirror
  unsigned long long* ctr = nullptr;      // == dc->ctr
  unsigned long long* h_ctr = nullptr;    // == h_dc->ctr
  int32_t maxdeg = 0;                     // max(out-degree, in-degree), for huge-row chunking
  int32_t update_policy = 0;              // pivot policy of dynamic recomputes
  bool colors_split = false;              // colours other than -1/-2 exist (after FB/WCC/set)
  bool fresh = false;                     // state just reset: all colours -1, nothing done
  bool rows_decide = false;               // just reset, caller colours that no edge crosses: the next
                                          // trim sweep may decide on row lengths alone
  int32_t lv_hint[3] = {8, 8, 8};         // level launches per search batch (fw, bw, both)
  int32_t it_hint = 4;                    // WCC iterations per batch (the last fixpoint's)
  int32_t it_hints[32] = {};              // the same per fixpoint ordinal of a decomposition (0: none yet)
  int32_t fix_idx = 0;                    // ordinal of the next fixpoint in this decomposition
  int32_t lv_hints[32] = {};              // level launches per search ordinal of a decomposition (0: none yet)
  int32_t search_idx = 0;                 // ordinal of the next search in this decomposition
  int32_t trim_hint[2] = {3, 3};          // trim sweeps per batch (absolute stop, relative stop)
  // the queued chain of a fresh decomposition (ChainCtl), sized by the last one:
  // its last gate (0 first trim loop, 1 global search, 2 second trim loop, 3
  // trim2 and the third trim loop), sweeps per trim loop, level launches, and
  // live vertices at each gate (launch grids)
  int32_t chain_shape = 3;
  int32_t chain_sweeps[3] = {1, 1, 1};
  int32_t chain_levels = 8;
  int32_t chain_level_grid[64] = {};      // workgroups per direction of each level launch (0: full)
  int64_t chain_live[sccmi::CHAIN_GATES] = {-1, -1, -1, -1, -1, -1};
  bool chain_off = false;                 // scc_ctx_set_chain(ctx, 0)
  bool chain_ended = false;               // the last chain ran to its end (no partition steps followed)
  bool color_pairs = true;                // a colouring round right after another first splits classes by backward reach
  bool last_color_round = false;          // the last partition step of this decomposition was a colouring round
  bool w_fresh = false;                   // W holds the labels of the current classes (the WCC phase just ran)
  // Rows cut to live same-colour neighbours (offsets begin / rbegin, ends fend
  // / frend): the first pull sweep of every propagation fixpoint writes the
  // rows it reads, cut, and cut_out / cut_in say which directions hold this
  // decomposition's cut rows.  Allocated on first use for an edge capacity of
  // f_cap and f_n vertices.
  int32_t *fadj = nullptr, *fradj = nullptr, *fend = nullptr, *frend = nullptr;
  int64_t f_cap = 0;
  int32_t f_n = 0;
  bool cut_out = false, cut_in = false;
  int32_t trim_epoch = 0;                 // number of the last trim loop queued
  bool labels_ready = false;              // canon[] valid for the current graph
  bool ran = false;
  // scc_state_reset defers its reset launch: the next call either performs it
  // (require_graph) or, in scc_run_method2, fuses it into the first trim sweep
  bool reset_pending = false;
  // pkey[n] holds the global step's pivot key and the reach bitmaps are clear
  // (left by that fused sweep when it ended the trim loop); cleared by every call
  bool pivot_ready = false;
  // canon (internal order) not computed: the last decomposition's chain wrote
  // the caller-order labels only (materialize_canon)
  bool canon_stale = false;
  // An update failed after it began to rewrite the rows: the directions, the
  // edge count or the labels may disagree, so graph calls are refused until
  // the next upload.
  bool graph_broken = false;

  // dynamic-update scratch (allocated on demand) and the gapped layout's room
  void* dyn = nullptr;
  scc_update_room room = {2, 1, 1, 0, 1 << 20};

  // timing
  hipEvent_t ev[8] = {};
  hipEvent_t ev_rb = nullptr;   // end of the last control-block copy (ctl_readback_start)
  std::vector<hipEvent_t> kev;  // per-launch event pairs for kernel timing
  double kernel_ms = 0;
  int64_t kernel_launches = 0;

  // learned-scheduler scores per vertex (SCC_PIVOT_LEARNED), NULL when unset
  uint32_t* sched_scores = nullptr;
  void* sched = nullptr;
  bool multi_colors = false;  // caller-set colours: the global step is a partition round
  bool classes = false;       // live colours are WCC classes or finer (after WCC / a round)

  // phase trace of the current run (scc_phase_trace)
  std::vector<scc_phase_rec> trace;
  int32_t round_idx = -1;

  std::vector<void*> allocs;
  std::vector<uint64_t> alloc_sizes;
  // buffers are pieces of a static array (scc_run_device_noalloc): dev_alloc
  // refuses, and no phase trace is kept
  bool no_alloc = false;
};

namespace sccmi {
int set_error(scc_ctx* c, int code, const std::string& msg);
// Entry check of every call that uses the device graph: one is uploaded and
// usable.  Makes the context's device current and performs a deferred reset
// (materialize = false leaves it to the caller, scc_run_method2).
int require_graph(scc_ctx* c, bool materialize = true);
// Row-end arrays of the forward and reverse CSR.
inline const int32_t* row_end(const scc_ctx* c) { return c->end ? c->end : c->begin + 1; }
inline const int32_t* rrow_end(const scc_ctx* c) { return c->rend ? c->rend : c->rbegin + 1; }
inline const int32_t* tr_adj(const scc_ctx* c) { return c->tadj ? c->tadj : c->adj; }
inline const int32_t* tr_radj(const scc_ctx* c) { return c->tradj ? c->tradj : c->radj; }
}  // namespace sccmi

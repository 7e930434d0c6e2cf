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
constexpr int LIGHT = 16;      // row entries a lane scans alone before the wave helps
constexpr int HUGE_ROW = 8192; // rows at least this long are expanded by many blocks
constexpr int32_t WCC_TAG = 0x20000000, WCC_MASK = 0x1FFFFFFF;  // src/scc_weak.cc:57-62

// Counter slots in the context's device counter array (unsigned long long).
enum Ctr : int {
  C_LIST_OUT = 0,   // compaction output count (live / survivors)
  C_TRIMMED = 1,    // vertices trimmed by a trim sweep / removed by trim2
  C_Q_OUT = 2,      // next-frontier count
  C_CAND_OUT = 3,   // bottom-up candidates left
  C_SCC = 4,        // SCC members found by classify
  C_CHANGED = 5,    // WCC iteration changed something
  C_ROOTS = 6,      // WCC roots
  C_Q_EDGES = 7,    // sum of traversal-direction degrees of the next frontier
  C_HUGE = 8,       // frontier vertices deferred to the many-block expansion
  C_COUNT = 9,      // generic counter (SCC count, ...)
  C_AUX = 10,
  C_NUM = 16
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

// Block-wide sum added to *count with one atomic.  All threads must call.
__device__ __forceinline__ void block_count(unsigned long long x, unsigned long long* count) {
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
  __syncthreads();
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

  // per-vertex state
  int32_t *color = nullptr, *label = nullptr, *canon = nullptr, *W = nullptr;
  int32_t* live[2] = {nullptr, nullptr};
  int cur = 0;            // live[cur] is the current live list
  int64_t n_live = 0;     // its length (host copy)
  int32_t* q[2] = {nullptr, nullptr};
  int32_t* cand[2] = {nullptr, nullptr};
  int32_t* huge = nullptr;          // frontier vertices with huge rows
  uint32_t *vf = nullptr, *vb = nullptr;  // reach bitmaps
  unsigned long long* pkey = nullptr;     // per-slot pivot key (n+1)
  int32_t* rep = nullptr;                 // per-slot new-partition representatives 3*(n+1)
  int32_t* scratch_i32 = nullptr;         // n+1 ints for canonicalisation / misc
  unsigned long long* ctr = nullptr;      // device counters
  unsigned long long* h_ctr = nullptr;    // pinned host mirror
  bool labels_ready = false;              // canon[] valid for the current graph
  bool ran = false;

  // dynamic-update scratch (allocated on demand)
  void* dyn = nullptr;

  // timing
  hipEvent_t ev[8] = {};
  std::vector<hipEvent_t> kev;  // per-launch event pairs for kernel timing
  double kernel_ms = 0;
  int64_t kernel_launches = 0;

  // learned-scheduler scores per vertex (SCC_PIVOT_LEARNED), NULL when unset
  uint32_t* sched_scores = nullptr;
  void* sched = nullptr;
  bool multi_colors = false;  // caller-set colours: the global step is a partition round

  std::vector<void*> allocs;
  std::vector<uint64_t> alloc_sizes;
};

// gfx950 kernels of the SCC engine.  All data is int32 (ids, offsets, colours,
// labels) plus 1-bit-per-vertex reach bitmaps; no floating point and no MFMA:
// every kernel here is an HBM/L2-bound sweep over a vertex list.
//
// Launch shape: 256-thread workgroups (4 wave64s), grid-stride loops over a
// vertex list whose trip count is uniform across the workgroup, so wave- and
// block-wide collectives (ballot compaction, counters) can be used in the body.
// Row scans use a lane-per-vertex phase over the first LIGHT entries and then a
// wave-cooperative phase for the rest of long rows (64 lanes x 4 loads in flight).
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
                                        const int32_t* __restrict__ nb, Pred pred) {
  bool found = false;
  const int lim = act ? min(e - s, LIGHT) : 0;
  const int mx = wave_max(lim);
  for (int j = 0; j < mx; j++) {
    if (j < lim && !found) found = pred(nb[s + j], v, c);
    if (__ballot((j + 1 < lim) && !found) == 0) break;
  }
  uint64_t hm = __ballot(act && !found && (e - s > LIGHT));
  const int lane = lane_id();
  while (hm) {
    const int L = __ffsll((unsigned long long)hm) - 1;
    hm &= hm - 1;
    const int32_t hs = __shfl(s, L) + LIGHT, he = __shfl(e, L);
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

// ---- trim1 -------------------------------------------------------------------
// src/scc_trim1.cc:38-110: v of colour c is trimmed (label v, colour -2) when it
// has no out-neighbour k != v of colour c, or no such in-neighbour.  Survivors
// are compacted into list_out (replaces create_trim1_compact, :211-275).  Reads
// of neighbour colours race with trims of the same sweep; trimming is monotone,
// so the fixpoint is the same (the unique maximal set where every vertex keeps a
// same-colour in- and out-neighbour).
__global__ __launch_bounds__(BLOCK) void k_trim1(const int32_t* __restrict__ begin,
                                                 const int32_t* __restrict__ adj,
                                                 const int32_t* __restrict__ rbegin,
                                                 const int32_t* __restrict__ radj,
                                                 int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                 const int32_t* __restrict__ list_in, int64_t n_in,
                                                 int32_t* __restrict__ list_out,
                                                 unsigned long long* __restrict__ ctr) {
  auto live = [color](int32_t k, int32_t pv, int32_t pc) { return k != pv && color[k] == pc; };
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n_in;
    const int32_t v = valid ? list_in[i] : 0;
    const int32_t c = valid ? color[v] : DONE;
    const bool act = valid && c != DONE;
    int32_t s = 0, e = 0;
    if (act) {
      s = begin[v];
      e = begin[v + 1];
    }
    const bool out_ok = row_any(act, v, c, s, e, adj, live);
    const bool act2 = act && out_ok;
    int32_t rs = 0, re = 0;
    if (act2) {
      rs = rbegin[v];
      re = rbegin[v + 1];
    }
    const bool in_ok = row_any(act2, v, c, rs, re, radj, live);
    const bool trim = act && !in_ok;
    if (trim) {
      label[v] = v;
      color[v] = DONE;
    }
    block_count(trim ? 1ull : 0ull, &ctr[C_TRIMMED]);
    block_append(act && !trim, v, list_out, &ctr[C_LIST_OUT]);
  }
}

// ---- trim2 -------------------------------------------------------------------
// src/scc_trim2_new.cc:3-39: the unique same-colour neighbour other than v
// (repeats skipped), or -1 when there are zero or several.  Rows are short for
// every vertex trim2 can remove, so a lane scans alone and gives up at two.
__device__ __forceinline__ int32_t unique_nbr(const int32_t* __restrict__ off,
                                              const int32_t* __restrict__ nb,
                                              const int32_t* __restrict__ color, int32_t v, int32_t c) {
  int32_t the = -1;
  const int32_t s = off[v], e = off[v + 1];
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
__global__ __launch_bounds__(BLOCK) void k_trim2(const int32_t* __restrict__ begin,
                                                 const int32_t* __restrict__ adj,
                                                 const int32_t* __restrict__ rbegin,
                                                 const int32_t* __restrict__ radj,
                                                 int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                 const int32_t* __restrict__ list_in, int64_t n_in,
                                                 unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    int removed = 0;
    if (i < n_in) {
      const int32_t v = list_in[i];
      const int32_t c = color[v];
      if (c != DONE) {
        int32_t k = unique_nbr(begin, adj, color, v, c);
        bool pair = k > v && unique_nbr(begin, adj, color, k, c) == v;
        if (!pair) {
          k = unique_nbr(rbegin, radj, color, v, c);
          pair = k > v && unique_nbr(rbegin, radj, color, k, c) == v;
        }
        if (pair) {
          label[v] = v;
          label[k] = v;
          color[v] = DONE;
          color[k] = DONE;
          removed = 2;
        }
      }
    }
    block_count((unsigned long long)removed, &ctr[C_TRIMMED]);
  }
}

// ---- pivot selection -----------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_reset_slots(const int32_t* __restrict__ list, int64_t n_in,
                                                       int32_t n, const int32_t* __restrict__ color,
                                                       unsigned long long* __restrict__ pkey,
                                                       int32_t* __restrict__ rep) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    const int32_t c = color[v];
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
__global__ __launch_bounds__(BLOCK) void k_pivot_key(const int32_t* __restrict__ list, int64_t n_in,
                                                     int32_t n, const int32_t* __restrict__ color,
                                                     const int32_t* __restrict__ begin,
                                                     const int32_t* __restrict__ rbegin, int policy,
                                                     const uint32_t* __restrict__ scores,
                                                     unsigned long long* __restrict__ pkey) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    const int32_t c = color[v];
    if (c == DONE) continue;
    unsigned long long score = 0;
    if (policy == SCC_PIVOT_DEGREE_PRODUCT) {
      const unsigned long long p = (unsigned long long)(begin[v + 1] - begin[v]) *
                                   (unsigned long long)(rbegin[v + 1] - rbegin[v]);
      score = p > 0x1FFFFFFFFull ? 0x1FFFFFFFFull : p;
    } else if (policy == SCC_PIVOT_LEARNED) {
      score = scores[v];
    }
    const unsigned long long key = pivot_key(score, v);
    unsigned long long* slot = &pkey[slot_of(c, n)];
    if (key > *slot) atomicMax(slot, key);
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

// Drop vertices of colour -2 from a list (order kept inside each workgroup).
__global__ __launch_bounds__(BLOCK) void k_compact_live(const int32_t* __restrict__ list, int64_t n_in,
                                                        const int32_t* __restrict__ color,
                                                        int32_t* __restrict__ list_out,
                                                        unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    int32_t v = 0;
    bool keep = false;
    if (i < n_in) {
      v = list[i];
      keep = color[v] != DONE;
    }
    block_append(keep, v, list_out, &ctr[C_LIST_OUT]);
  }
}

// Seed a traversal: every live vertex that is its partition's pivot marks
// itself in V and enters the frontier.
__global__ __launch_bounds__(BLOCK) void k_seed(const int32_t* __restrict__ list, int64_t n_in, int32_t n,
                                                const int32_t* __restrict__ color,
                                                const unsigned long long* __restrict__ pkey,
                                                uint32_t* __restrict__ V, int32_t* __restrict__ q_out,
                                                unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool piv = false;
    int32_t v = 0;
    if (i < n_in) {
      v = list[i];
      const int32_t c = color[v];
      piv = c != DONE && pivot_of_key(pkey[slot_of(c, n)]) == v;
      if (piv) atomicOr(&V[v >> 5], 1u << (v & 31));
    }
    wave_append(piv, v, q_out, &ctr[C_Q_OUT]);
  }
}

// ---- reachability (direction-optimising, partition-aware) ----------------------
// One level of the forward (rows = out-edges) or backward (rows = in-edges)
// search of src/scc_fb_global.cc:32-129 / gm_bfs_template.h, for every
// partition at once: a vertex is admitted when it is reached from a visited
// vertex of the same colour (the navigator of src/scc_fb_global.cc:64-68 and
// :123-127; during a step colours do not change, so "same colour" is the
// partition).  Visited state is a bitmap, claimed with atomicOr.

__device__ __forceinline__ bool try_claim(uint32_t* __restrict__ V, const int32_t* __restrict__ color,
                                          int32_t k, int32_t c) {
  const uint32_t bit = 1u << (k & 31);
  if (V[k >> 5] & bit) return false;
  if (color[k] != c) return false;
  return !(atomicOr(&V[k >> 5], bit) & bit);
}

// Top-down: expand every frontier vertex's row.  Rows of at least HUGE_ROW
// entries are deferred to k_td_huge (appended to `huge`).  `noff` gives the
// degree used to estimate the next level's work (C_Q_EDGES).
__global__ __launch_bounds__(BLOCK) void k_td(const int32_t* __restrict__ off, const int32_t* __restrict__ nb,
                                              const int32_t* __restrict__ color, uint32_t* __restrict__ V,
                                              const int32_t* __restrict__ q_in, int64_t n_qin,
                                              int32_t* __restrict__ q_out, int32_t* __restrict__ huge,
                                              unsigned long long* __restrict__ ctr) {
  const int lane = lane_id();
  unsigned long long edges = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_qin; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n_qin;
    int32_t u = 0, s = 0, e = 0, c = DONE;
    if (valid) {
      u = q_in[i];
      s = off[u];
      e = off[u + 1];
      c = color[u];
    }
    const int32_t d = e - s;
    const bool hug = valid && d >= HUGE_ROW;
    wave_append(hug, u, huge, &ctr[C_HUGE]);
    const bool act = valid && !hug;
    const int lim = act ? min(d, LIGHT) : 0;
    const int mx = wave_max(lim);
    for (int j = 0; j < mx; j++) {
      int32_t k = 0;
      bool cl = false;
      if (j < lim) {
        k = nb[s + j];
        cl = try_claim(V, color, k, c);
        if (cl) edges += (unsigned long long)(off[k + 1] - off[k]);
      }
      wave_append(cl, k, q_out, &ctr[C_Q_OUT]);
    }
    uint64_t hm = __ballot(act && d > LIGHT);
    while (hm) {
      const int L = __ffsll((unsigned long long)hm) - 1;
      hm &= hm - 1;
      const int32_t hs = __shfl(s, L) + LIGHT, he = __shfl(e, L), hc = __shfl(c, L);
      for (int32_t b = hs; b < he; b += 64) {
        const int32_t idx = b + lane;
        int32_t k = 0;
        bool cl = false;
        if (idx < he) {
          k = nb[idx];
          cl = try_claim(V, color, k, hc);
          if (cl) edges += (unsigned long long)(off[k + 1] - off[k]);
        }
        wave_append(cl, k, q_out, &ctr[C_Q_OUT]);
      }
    }
  }
  edges = wave_sum_u64(edges);
  if (lane == 0 && edges) atomicAdd(&ctr[C_Q_EDGES], edges);
}

// Top-down expansion of the deferred huge rows: the rows are cut into chunks of
// CHUNK entries and the grid's workgroups take chunks round-robin.
constexpr int TD_CHUNK = 4096;
__global__ __launch_bounds__(BLOCK) void k_td_huge(const int32_t* __restrict__ off,
                                                   const int32_t* __restrict__ nb,
                                                   const int32_t* __restrict__ color, uint32_t* __restrict__ V,
                                                   const int32_t* __restrict__ huge, int64_t n_huge,
                                                   int32_t* __restrict__ q_out,
                                                   unsigned long long* __restrict__ ctr) {
  int64_t total = 0;
  for (int64_t h = 0; h < n_huge; h++) {
    const int32_t u = huge[h];
    total += (off[u + 1] - off[u] + TD_CHUNK - 1) / TD_CHUNK;
  }
  unsigned long long edges = 0;
  for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
    int64_t acc = 0, h = 0;
    int32_t u = 0, nch = 0;
    for (;; h++) {
      u = huge[h];
      nch = (off[u + 1] - off[u] + TD_CHUNK - 1) / TD_CHUNK;
      if (g < acc + nch) break;
      acc += nch;
    }
    const int32_t s = off[u] + (int32_t)(g - acc) * TD_CHUNK;
    const int32_t e = min(s + TD_CHUNK, off[u + 1]);
    const int32_t c = color[u];
    for (int32_t b = s; b < e; b += BLOCK) {
      const int32_t idx = b + threadIdx.x;
      int32_t k = 0;
      bool cl = false;
      if (idx < e) {
        k = nb[idx];
        cl = try_claim(V, color, k, c);
        if (cl) edges += (unsigned long long)(off[k + 1] - off[k]);
      }
      wave_append(cl, k, q_out, &ctr[C_Q_OUT]);
    }
  }
  edges = wave_sum_u64(edges);
  if (lane_id() == 0 && edges) atomicAdd(&ctr[C_Q_EDGES], edges);
}

// Bottom-up: every candidate not yet visited scans its opposite-direction row
// for a visited parent of its colour (gm_bfs_template.h:281-301, :810-829).
// A parent visited earlier in the same pass also counts (reachability needs no
// levels).  Newly visited vertices form the next frontier; the rest stay
// candidates.  MULTI=false when only one partition is being searched, where a
// visited parent is in the partition by construction.
template <bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_bu(const int32_t* __restrict__ roff, const int32_t* __restrict__ rnb,
                                              const int32_t* __restrict__ off,
                                              const int32_t* __restrict__ color, uint32_t* __restrict__ V,
                                              const int32_t* __restrict__ c_in, int64_t n_cin,
                                              int32_t* __restrict__ c_out, int32_t* __restrict__ q_out,
                                              unsigned long long* __restrict__ ctr) {
  auto reached = [V, color](int32_t k, int32_t pv, int32_t pc) {
    (void)pv;
    if (!((V[k >> 5] >> (k & 31)) & 1u)) return false;
    return !MULTI || color[k] == pc;
  };
  unsigned long long edges = 0;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_cin; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n_cin;
    int32_t v = 0, c = DONE, s = 0, e = 0;
    bool act = false;
    if (valid) {
      v = c_in[i];
      if (!bit_test(V, v)) {
        c = color[v];
        act = c != DONE;
        if (act) {
          s = roff[v];
          e = roff[v + 1];
        }
      }
    }
    const bool found = row_any(act, v, c, s, e, rnb, reached);
    if (found) {
      atomicOr(&V[v >> 5], 1u << (v & 31));
      edges += (unsigned long long)(off[v + 1] - off[v]);
    }
    block_append(found, v, q_out, &ctr[C_Q_OUT]);
    block_append(act && !found, v, c_out, &ctr[C_CAND_OUT]);
  }
  block_count(edges, &ctr[C_Q_EDGES]);
}

template __global__ void k_bu<true>(const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                                    uint32_t*, const int32_t*, int64_t, int32_t*, int32_t*,
                                    unsigned long long*);
template __global__ void k_bu<false>(const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                                     uint32_t*, const int32_t*, int64_t, int32_t*, int32_t*,
                                     unsigned long long*);

// ---- classification after a forward-backward step -------------------------------
// src/scc_fb_global.cc:99-113: forward and backward reached -> SCC of the pivot
// (label pivot, colour -2).  The rest split into fw-only / bw-only / neither
// (src/scc_fb_seq2.cc:245-285); each new part's colour is its minimum member,
// gathered here with atomicMin and applied by k_recolor.
__global__ __launch_bounds__(BLOCK) void k_classify(const int32_t* __restrict__ list, int64_t n_in,
                                                    int32_t n, int32_t* __restrict__ color,
                                                    int32_t* __restrict__ label,
                                                    const uint32_t* __restrict__ vf,
                                                    const uint32_t* __restrict__ vb,
                                                    const unsigned long long* __restrict__ pkey,
                                                    int32_t* __restrict__ rep, int32_t* __restrict__ list_out,
                                                    unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool keep = false, scc = false;
    int32_t v = 0;
    if (i < n_in) {
      v = list[i];
      const int32_t c = color[v];
      if (c != DONE) {
        const int32_t s = slot_of(c, n);
        const bool f = bit_test(vf, v), b = bit_test(vb, v);
        if (f && b) {
          label[v] = pivot_of_key(pkey[s]);
          color[v] = DONE;
          scc = true;
        } else {
          const int t = f ? 0 : (b ? 1 : 2);
          atomicMin(&rep[3 * (int64_t)s + t], v);
          keep = true;
        }
      }
    }
    block_count(scc ? 1ull : 0ull, &ctr[C_SCC]);
    block_append(keep, v, list_out, &ctr[C_LIST_OUT]);
  }
}

__global__ __launch_bounds__(BLOCK) void k_recolor(const int32_t* __restrict__ list, int64_t n_in, int32_t n,
                                                   int32_t* __restrict__ color,
                                                   const uint32_t* __restrict__ vf,
                                                   const uint32_t* __restrict__ vb,
                                                   const int32_t* __restrict__ rep) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    const int32_t c = color[v];
    const int32_t s = slot_of(c, n);
    const bool f = bit_test(vf, v), b = bit_test(vb, v);
    const int t = f ? 0 : (b ? 1 : 2);
    color[v] = rep[3 * (int64_t)s + t];
  }
}

// ---- WCC colouring --------------------------------------------------------------
// src/scc_weak.cc:128-150: W[v] = v, tagged when out-degree >= 50.
__global__ __launch_bounds__(BLOCK) void k_wcc_init(const int32_t* __restrict__ list, int64_t n_in,
                                                    const int32_t* __restrict__ begin,
                                                    int32_t* __restrict__ W) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK) {
    const int32_t v = list[i];
    W[v] = (begin[v + 1] - begin[v] >= 50) ? (v | WCC_TAG) : v;
  }
}

// src/scc_weak.cc:65-125 (propagate_color), both phases fused per vertex: pull
// the minimum W over same-colour out-neighbours, then jump to the root's value
// when the root is not self-rooted.  Each vertex writes only its own W and W
// only decreases, so racing reads change the path to the fixpoint, not the
// fixpoint (W[v] = min tag over v's same-colour forward-reachable set).
__global__ __launch_bounds__(BLOCK) void k_wcc_iter(const int32_t* __restrict__ list, int64_t n_in,
                                                    const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ adj,
                                                    const int32_t* __restrict__ color,
                                                    int32_t* __restrict__ W,
                                                    unsigned long long* __restrict__ ctr) {
  const int lane = lane_id();
  bool changed_any = false;
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n_in;
    int32_t v = 0, c = DONE, s = 0, e = 0, m = INT32_MAX, w0 = INT32_MAX;
    if (valid) {
      v = list[i];
      c = color[v];
      s = begin[v];
      e = begin[v + 1];
      w0 = m = __hip_atomic_load(&W[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const bool act = valid && c != DONE;
    // lane-serial over short rows, wave-cooperative min over long ones
    const int lim = act ? min(e - s, 64) : 0;
    for (int j = 0; j < lim; j++) {
      const int32_t k = adj[s + j];
      if (color[k] == c) m = min(m, __hip_atomic_load(&W[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    uint64_t hm = __ballot(act && e - s > 64);
    while (hm) {
      const int L = __ffsll((unsigned long long)hm) - 1;
      hm &= hm - 1;
      const int32_t hs = __shfl(s, L) + 64, he = __shfl(e, L), hc = __shfl(c, L);
      int32_t hmin = INT32_MAX;
      for (int32_t b = hs; b < he; b += 64) {
        const int32_t idx = b + lane;
        if (idx < he) {
          const int32_t k = adj[idx];
          if (color[k] == hc)
            hmin = min(hmin, __hip_atomic_load(&W[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) hmin = min(hmin, __shfl_xor(hmin, o));
      if (lane == L) m = min(m, hmin);
    }
    if (act) {
      const int32_t r = m & WCC_MASK;
      if (r != v) {
        const int32_t wr = __hip_atomic_load(&W[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((wr & WCC_MASK) != r) m = min(m, wr);
      }
      if (m < w0) {
        __hip_atomic_store(&W[v], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        changed_any = true;
      }
    }
  }
  if (__ballot(changed_any) != 0 && lane == 0) atomicAdd(&ctr[C_CHANGED], 1ull);
}

// src/scc_weak.cc:166-194: each class takes its root's id as colour (a fresh,
// unique colour: the root is a member of the class); counts roots.
__global__ __launch_bounds__(BLOCK) void k_wcc_color(const int32_t* __restrict__ list, int64_t n_in,
                                                     int32_t* __restrict__ color,
                                                     const int32_t* __restrict__ W,
                                                     unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n_in; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool root = false;
    if (i < n_in) {
      const int32_t v = list[i];
      if (color[v] != DONE) {
        const int32_t r = W[v] & WCC_MASK;
        color[v] = r;
        root = r == v;
      }
    }
    block_count(root ? 1ull : 0ull, &ctr[C_ROOTS]);
  }
}

// ---- canonical labels -----------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_canon_min(int32_t n, const int32_t* __restrict__ label,
                                                     int32_t* __restrict__ mn) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    const int32_t l = label[v];
    if ((uint32_t)l < (uint32_t)n) atomicMin(&mn[l], (int32_t)v);
  }
}

__global__ __launch_bounds__(BLOCK) void k_canon_apply(int32_t n, const int32_t* __restrict__ label,
                                                       const int32_t* __restrict__ mn,
                                                       int32_t* __restrict__ out,
                                                       unsigned long long* __restrict__ ctr) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t v = base + threadIdx.x;
    bool self = false;
    if (v < n) {
      const int32_t l = label[v];
      const int32_t c = ((uint32_t)l < (uint32_t)n) ? mn[l] : -1;
      out[v] = c;
      self = c == (int32_t)v;
    }
    block_count(self ? 1ull : 0ull, &ctr[C_COUNT]);
  }
}

}  // namespace sccmi

// ---- host launchers -------------------------------------------------------------
namespace sccmi {

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
void launch_trim1(hipStream_t st, const int32_t* begin, const int32_t* adj, const int32_t* rbegin,
                  const int32_t* radj, int32_t* color, int32_t* label, const int32_t* list_in,
                  int64_t n_in, int32_t* list_out, unsigned long long* ctr) {
  if (n_in > 0)
    LAUNCH(k_trim1, grid_for(n_in), st, begin, adj, rbegin, radj, color, label, list_in, n_in,
           list_out, ctr);
}
void launch_trim2(hipStream_t st, const int32_t* begin, const int32_t* adj, const int32_t* rbegin,
                  const int32_t* radj, int32_t* color, int32_t* label, const int32_t* list_in,
                  int64_t n_in, unsigned long long* ctr) {
  if (n_in > 0)
    LAUNCH(k_trim2, grid_for(n_in), st, begin, adj, rbegin, radj, color, label, list_in, n_in, ctr);
}
void launch_reset_slots(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n,
                        const int32_t* color, unsigned long long* pkey, int32_t* rep) {
  if (n_in > 0) LAUNCH(k_reset_slots, grid_for(n_in), st, list, n_in, n, color, pkey, rep);
}
void launch_pivot_key(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n,
                      const int32_t* color, const int32_t* begin, const int32_t* rbegin, int policy,
                      const uint32_t* scores, unsigned long long* pkey) {
  if (n_in > 0)
    LAUNCH(k_pivot_key, grid_for(n_in), st, list, n_in, n, color, begin, rbegin, policy, scores,
           pkey);
}
void launch_force_pivot(hipStream_t st, int32_t p, int32_t n, const int32_t* color,
                        unsigned long long* pkey) {
  LAUNCH(k_force_pivot, 1, st, p, n, color, pkey);
}
void launch_compact_live(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* color,
                         int32_t* list_out, unsigned long long* ctr) {
  if (n_in > 0) LAUNCH(k_compact_live, grid_for(n_in), st, list, n_in, color, list_out, ctr);
}
void launch_seed(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, const int32_t* color,
                 const unsigned long long* pkey, uint32_t* V, int32_t* q_out,
                 unsigned long long* ctr) {
  if (n_in > 0) LAUNCH(k_seed, grid_for(n_in), st, list, n_in, n, color, pkey, V, q_out, ctr);
}
void launch_td(hipStream_t st, const int32_t* off, const int32_t* nb, const int32_t* color,
               uint32_t* V, const int32_t* q_in, int64_t n_qin, int32_t* q_out, int32_t* huge,
               unsigned long long* ctr) {
  if (n_qin > 0) LAUNCH(k_td, grid_for(n_qin), st, off, nb, color, V, q_in, n_qin, q_out, huge, ctr);
}
void launch_td_huge(hipStream_t st, const int32_t* off, const int32_t* nb, const int32_t* color,
                    uint32_t* V, const int32_t* huge, int64_t n_huge, int32_t* q_out,
                    unsigned long long* ctr) {
  if (n_huge > 0) LAUNCH(k_td_huge, 2048, st, off, nb, color, V, huge, n_huge, q_out, ctr);
}
void launch_bu(hipStream_t st, bool multi, const int32_t* roff, const int32_t* rnb,
               const int32_t* off, const int32_t* color, uint32_t* V, const int32_t* c_in,
               int64_t n_cin, int32_t* c_out, int32_t* q_out, unsigned long long* ctr) {
  if (n_cin <= 0) return;
  if (multi)
    LAUNCH(k_bu<true>, grid_for(n_cin), st, roff, rnb, off, color, V, c_in, n_cin, c_out, q_out, ctr);
  else
    LAUNCH(k_bu<false>, grid_for(n_cin), st, roff, rnb, off, color, V, c_in, n_cin, c_out, q_out, ctr);
}
void launch_classify(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, int32_t* color,
                     int32_t* label, const uint32_t* vf, const uint32_t* vb,
                     const unsigned long long* pkey, int32_t* rep, int32_t* list_out,
                     unsigned long long* ctr) {
  if (n_in > 0)
    LAUNCH(k_classify, grid_for(n_in), st, list, n_in, n, color, label, vf, vb, pkey, rep, list_out,
           ctr);
}
void launch_recolor(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, int32_t* color,
                    const uint32_t* vf, const uint32_t* vb, const int32_t* rep) {
  if (n_in > 0) LAUNCH(k_recolor, grid_for(n_in), st, list, n_in, n, color, vf, vb, rep);
}
void launch_wcc_init(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* begin,
                     int32_t* W) {
  if (n_in > 0) LAUNCH(k_wcc_init, grid_for(n_in), st, list, n_in, begin, W);
}
void launch_wcc_iter(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* begin,
                     const int32_t* adj, const int32_t* color, int32_t* W, unsigned long long* ctr) {
  if (n_in > 0) LAUNCH(k_wcc_iter, grid_for(n_in), st, list, n_in, begin, adj, color, W, ctr);
}
void launch_wcc_color(hipStream_t st, const int32_t* list, int64_t n_in, int32_t* color,
                      const int32_t* W, unsigned long long* ctr) {
  if (n_in > 0) LAUNCH(k_wcc_color, grid_for(n_in), st, list, n_in, color, W, ctr);
}
void launch_canon(hipStream_t st, int32_t n, const int32_t* label, int32_t* mn, int32_t* out,
                  unsigned long long* ctr) {
  if (n <= 0) return;
  LAUNCH(k_fill_i32, grid_for(n, 4096), st, mn, INT32_MAX, (int64_t)n);
  LAUNCH(k_canon_min, grid_for(n), st, n, label, mn);
  LAUNCH(k_canon_apply, grid_for(n), st, n, label, mn, out, ctr);
}

#undef LAUNCH
}  // namespace sccmi

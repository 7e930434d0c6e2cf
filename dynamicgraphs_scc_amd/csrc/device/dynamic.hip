// Exact dynamic SCC maintenance under batches of edge inserts and deletes.
//
// Replaces the reference's incremental (src/scc_incremental.cc:7-27 with the
// condensation built by src/common_main.h:180-314) and decremental
// (src/scc_decremental.cc:7-24 with src/common_main.h:343-393) paths with one
// exact call.  Let G' = G - D + I and C(v) the canonical SCC labels of G.
// First P = FW_G'(p) n BW_G'(p) for the max in x out vertex p: one
// forward-backward step over G'.  On graphs with a giant SCC, P is the new
// giant (old giant minus what split off, plus every SCC merged into it), and
// the rest of the work stays outside it.  Then, in G' - P,
//   X_d = union of old SCCs that lost an internal edge, minus P (only these can
//         split; the reference's read_file2 marks the same SCCs, :361-370);
//   S/T = heads/tails of inserted edges u->v outside P with C(u) != C(v) or
//         C(u) in X_d;
//   X   = X_d  u  (FW(S) n BW(T)), both searches confined to G' - P.
// Every SCC of G' outside P and X is an SCC of G (a cycle joining old SCCs must
// use an inserted S/T edge, and one through P would put its vertices in P), and
// X is a union of SCCs of G', so recomputing the SCCs of G'[X] with the static
// pipeline, P's label from the first step, and C elsewhere gives the SCCs of
// G'.  BW(T) is searched only inside FW(S) (a path from w in FW(S) to T stays
// in FW(S)), the same restriction the reference's backward search uses.
//
// Graph storage under updates ("gapped" CSR, per direction): row u occupies
// adj[begin[u] .. end[u]) inside its own room adj[begin[u] .. cap[u]); rooms
// are laid out in vertex order with 2 + len/8 spare entries each, followed by
// an overflow area.  The first update converts the packed upload to this
// layout.  A batch then touches only the rows it changes: their old entries
// are staged, and each row is rewritten sorted (survivors at
// i - #deleted-before + #inserted-before, inserts at their merged rank, all by
// binary search), in its own room when it fits, else in a new room taken from
// the overflow area.  When the overflow area runs out, the layout is rebuilt
// with fresh spare room.  Cost per batch is O(batch + entries of changed rows),
// not O(edges).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {

namespace {

// Device counters of the update engine (own block, apart from DevCtl).
enum DynCtr : int {
  D_ROWS = 0,      // changed rows collected
  D_BUMP_F = 1,    // next free entry of the forward overflow area
  D_BUMP_R = 2,    // same, reverse
  D_OVERFLOW = 3,  // a room could not be taken
  D_MAXLEN = 4,    // longest changed row after the batch
  D_STAGED = 5,    // entries staged from the changed rows
  D_DELENT = 6,    // row entries the delete keys of the direction remove (duplicates included)
  D_NUM = 8
};

struct Dyn {
  int64_t cap_b = 0;  // capacity of each batch list
  unsigned long long *k_in = nullptr, *k_tmp = nullptr;
  unsigned long long *del_f = nullptr, *del_r = nullptr, *ins_f = nullptr, *ins_r = nullptr;
  unsigned long long *ins_sorted = nullptr;
  uint8_t* flags = nullptr;
  int32_t *src = nullptr, *dst = nullptr;  // staging of one list
  int32_t *rdel = nullptr, *rins = nullptr;  // per-row entries removed / added in the direction being applied (n+1)
  // per delete key of the direction being applied: the row entries it removes
  // (every copy of a duplicated edge), and their exclusive prefix sum over the
  // sorted key list (cap_b + 1 each)
  int32_t *dcopies = nullptr, *dcum = nullptr;
  int32_t* rowmark = nullptr;                // per row: 1 + index in the changed-row list, else 0
  int32_t* aff = nullptr;                    // per old-SCC flag
  int32_t *cap_f = nullptr, *cap_r = nullptr;  // room ends of the gapped layout
  int64_t size_f = 0, size_r = 0;              // entries allocated for adj / radj
  int64_t cap_rows = 0;                        // capacity of the changed-row arrays
  int32_t *rows = nullptr, *tmpoff = nullptr, *newbeg = nullptr, *newcap = nullptr, *lens = nullptr;
  int4* rng = nullptr;  // per changed row: its delete range, its insert range
  int32_t* stage = nullptr;  // staged old entries of changed rows
  int64_t stage_cap = 0;
  int32_t* crow = nullptr;   // per chunk of staged entries: its first and last changed row (2 per chunk)
  unsigned long long* dctr = nullptr;    // DynCtr (device)
  unsigned long long* h_dctr = nullptr;  // pinned mirror
  void* tmp = nullptr;  // library temp storage
  size_t tmp_bytes = 0;
  int32_t n = -1;
  int32_t* h_stage = nullptr;  // pinned staging for batch upload (4 * cap_b ints: inserts, then deletes)
  int64_t rebuilds = 0;        // layout rebuilds of the current update (both directions)
};

constexpr unsigned long long BAD_KEY = ~0ull;
constexpr int32_t OUTSIDE = -3;  // region search: outside FW(S)

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
                                                     int64_t k, bool drop_loops, const int32_t* __restrict__ map,
                                                     unsigned long long* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    // (map: caller ids to internal ids in the degree layout, validated on the host)
    const int32_t u = map ? map[s[i]] : s[i], v = map ? map[d[i]] : d[i];
    keys[i] = (drop_loops && u == v) ? BAD_KEY
                                     : (((unsigned long long)(uint32_t)u << 32) | (uint32_t)v);
  }
}

// Flag sorted delete keys that are present in G and not re-inserted.
__global__ __launch_bounds__(BLOCK) void k_flag_del(const unsigned long long* __restrict__ keys, int64_t k,
                                                    const unsigned long long* __restrict__ ins, int64_t ni,
                                                    const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ end,
                                                    const int32_t* __restrict__ adj,
                                                    uint8_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = keys[i];
    bool keep = key != BAD_KEY && (i == 0 || keys[i - 1] != key);
    if (keep) {
      const int32_t u = key_hi(key), v = key_lo(key);
      const int32_t s = begin[u], e = end[u];
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
                                                    const int32_t* __restrict__ end,
                                                    const int32_t* __restrict__ adj,
                                                    uint8_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < k; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = keys[i];
    bool keep = key != BAD_KEY && (i == 0 || keys[i - 1] != key);
    if (keep) {
      const int32_t u = key_hi(key), v = key_lo(key);
      const int32_t s = begin[u], e = end[u];
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

// Per sorted, filtered delete key u->v: the entries equal to v in row u (a
// graph may hold an edge several times; a delete removes every copy, as
// src/common_main.h:379-386 drops every matching edge), added to the row's
// removed count.  copies[k] = 0 closes the list for the exclusive scan.
__global__ __launch_bounds__(BLOCK) void k_del_copies(const unsigned long long* __restrict__ keys, int64_t k,
                                                      const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ end,
                                                      const int32_t* __restrict__ adj,
                                                      int32_t* __restrict__ copies, int32_t* __restrict__ rdel) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i <= k; i += (int64_t)gridDim.x * BLOCK) {
    if (i == k) {
      copies[k] = 0;
      continue;
    }
    const int32_t u = key_hi(keys[i]), v = key_lo(keys[i]);
    const int32_t e = end[u];
    const int32_t p = lb_i32(adj, begin[u], e, v);
    const int32_t q = lb_i32(adj, p, e, v + 1);  // v < n <= INT32_MAX - 1
    copies[i] = q - p;
    atomicAdd(&rdel[u], q - p);
  }
}

__global__ void k_dctr_zero(unsigned long long* __restrict__ dctr, int mask) {
  const int i = threadIdx.x;
  if (i < D_NUM && ((mask >> i) & 1)) dctr[i] = 0;
}
__global__ void k_store_total(const int32_t* __restrict__ p, unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0) *out = (unsigned long long)(uint32_t)p[0];
}

// ---- gapped CSR ------------------------------------------------------------------
// Room of a row of len entries: len + spare + spare8 eighths of len (the
// context's scc_update_room; 2 and 1 by default).
__device__ __forceinline__ int32_t room_of(int32_t len, int2 sp) {
  return len + sp.x + (int32_t)(((int64_t)len * sp.y) >> 3);
}

// Changed rows: every row with a delete or an insert enters the list once.
__global__ __launch_bounds__(BLOCK) void k_collect_rows(const unsigned long long* __restrict__ keys, int64_t k,
                                                        int32_t* __restrict__ rowmark,
                                                        int32_t* __restrict__ rows,
                                                        unsigned long long* __restrict__ count) {
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < k; base += (int64_t)gridDim.x * BLOCK) {
    const int64_t i = base + threadIdx.x;
    bool take = false;
    int32_t u = 0;
    if (i < k) {
      u = key_hi(keys[i]);
      if (i == 0 || key_hi(keys[i - 1]) != u) take = atomicCAS(&rowmark[u], 0, -1) == 0;
    }
    wave_append(take, u, rows, count);
  }
}

// Per changed row: staged length, new room (in place, or from the overflow
// area when the row outgrows its room), longest row.
__global__ __launch_bounds__(BLOCK) void k_row_plan(const int32_t* __restrict__ rows, int64_t nr,
                                                    const int32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ end,
                                                    const int32_t* __restrict__ cap,
                                                    const int32_t* __restrict__ rdel,
                                                    const int32_t* __restrict__ rins,
                                                    int32_t* __restrict__ rowmark, int32_t* __restrict__ lens,
                                                    int32_t* __restrict__ newbeg, int32_t* __restrict__ newcap,
                                                    const unsigned long long* __restrict__ del, int64_t nd,
                                                    const unsigned long long* __restrict__ ins, int64_t ni,
                                                    int4* __restrict__ rng,
                                                    unsigned long long* __restrict__ dctr, int bump_slot,
                                                    long long size, int2 sp) {
  for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < nr; j += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = rows[j];
    // the row's ranges in the sorted delete / insert lists
    const unsigned long long klo = (unsigned long long)(uint32_t)u << 32;
    const unsigned long long khi = (unsigned long long)(uint32_t)(u + 1) << 32;
    int4 r;
    r.x = (int32_t)lb_u64(del, 0, nd, klo);
    r.y = (int32_t)lb_u64(del, r.x, nd, khi);
    r.z = (int32_t)lb_u64(ins, 0, ni, klo);
    r.w = (int32_t)lb_u64(ins, r.z, ni, khi);
    rng[j] = r;
    const int32_t b = begin[u], len = end[u] - b;
    const int32_t nlen = len - rdel[u] + rins[u];
    rowmark[u] = (int32_t)j + 1;
    lens[j] = len;
    if (j == 0) lens[nr] = 0;  // the scan's last input
    if (nlen <= cap[u] - b) {
      newbeg[j] = b;
      newcap[j] = cap[u];
    } else {
      const int32_t room = room_of(nlen, sp);
      const unsigned long long at = atomicAdd(&dctr[bump_slot], (unsigned long long)room);
      if ((long long)(at + room) > size) {
        atomicOr(&dctr[D_OVERFLOW], 1ull);
        newbeg[j] = b;
        newcap[j] = cap[u];
      } else {
        newbeg[j] = (int32_t)at;
        newcap[j] = (int32_t)(at + room);
      }
    }
    atomicMax(&dctr[D_MAXLEN], (unsigned long long)nlen);
  }
}

// Copy rows into a destination laid out by doff (an exclusive scan over the
// list): row j (vertex rows[j], or j when rows is NULL) moves
// src[sbeg[u] .. send[u]) to dst[doff[j] ..].  A workgroup takes 256 rows and
// walks their destination range; the row of a position is found by binary
// search over the group's offsets staged in LDS.  Positions past a row's
// length (spare room) are skipped.
constexpr int RB_ROWS = 256;
// A group whose destination range is longer than this is left to
// k_gather_heavy (with the degree vertex layout the first group holds the 256
// largest hubs, millions of entries that one workgroup would copy alone).
constexpr int32_t RB_HEAVY = 65536;
constexpr int32_t RB_CHUNK = 8192;

// Stage group [j0, j0 + cnt)'s offsets (relative to doff[j0]), sources and lengths in LDS.
__device__ __forceinline__ void stage_group(const int32_t* __restrict__ rows, int64_t j0, int32_t cnt,
                                            const int32_t* __restrict__ sbeg, const int32_t* __restrict__ send,
                                            const int32_t* __restrict__ doff, int32_t o0, int32_t* s_off,
                                            int32_t* s_src, int32_t* s_len) {
  for (int t = threadIdx.x; t <= cnt; t += BLOCK) s_off[t] = doff[j0 + t] - o0;
  for (int t = threadIdx.x; t < cnt; t += BLOCK) {
    const int32_t u = rows ? rows[j0 + t] : (int32_t)(j0 + t);
    s_src[t] = sbeg[u];
    s_len[t] = send[u] - sbeg[u];
  }
}

// Copy destination positions [e0, e1) of a staged group.
__device__ __forceinline__ void copy_group(int32_t cnt, int32_t o0, int32_t e0, int32_t e1,
                                           const int32_t* s_off, const int32_t* s_src, const int32_t* s_len,
                                           const int32_t* __restrict__ src, int32_t* __restrict__ dst) {
  for (int32_t e = e0 + threadIdx.x; e < e1; e += BLOCK) {
    int32_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= e) lo = mid;
      else hi = mid;
    }
    const int32_t i = e - s_off[lo];
    if (i < s_len[lo]) dst[o0 + e] = src[s_src[lo] + i];
  }
}

__global__ __launch_bounds__(BLOCK) void k_gather_rows(const int32_t* __restrict__ rows, int64_t nr,
                                                       const int32_t* __restrict__ sbeg,
                                                       const int32_t* __restrict__ send,
                                                       const int32_t* __restrict__ src,
                                                       const int32_t* __restrict__ doff,
                                                       int32_t* __restrict__ dst, int32_t* __restrict__ heavy) {
  __shared__ int32_t s_off[RB_ROWS + 1];
  __shared__ int32_t s_src[RB_ROWS], s_len[RB_ROWS];
  for (int64_t j0 = (int64_t)blockIdx.x * RB_ROWS; j0 < nr; j0 += (int64_t)gridDim.x * RB_ROWS) {
    const int32_t cnt = (int32_t)min<int64_t>(RB_ROWS, nr - j0);
    const int32_t o0 = doff[j0];
    stage_group(rows, j0, cnt, sbeg, send, doff, o0, s_off, s_src, s_len);
    __syncthreads();
    const int32_t total = s_off[cnt];
    if (total > RB_HEAVY) {  // (uniform) listed for k_gather_heavy
      if (threadIdx.x == 0) heavy[1 + atomicAdd(&heavy[0], 1)] = (int32_t)(j0 / RB_ROWS);
    } else {
      copy_group(cnt, o0, 0, total, s_off, s_src, s_len, src, dst);
    }
    __syncthreads();
  }
}

// The listed heavy groups, in RB_CHUNK-position pieces: the chunks of all
// listed groups are numbered in list order and dealt round-robin over the grid,
// so every workgroup gets about the same number whatever the groups' sizes.
// Each batch of up to BLOCK groups has its chunk counts loaded in parallel and
// prefix-summed in LDS; a workgroup stages only the groups it has chunks in.
__global__ __launch_bounds__(BLOCK) void k_gather_heavy(const int32_t* __restrict__ rows, int64_t nr,
                                                        const int32_t* __restrict__ sbeg,
                                                        const int32_t* __restrict__ send,
                                                        const int32_t* __restrict__ src,
                                                        const int32_t* __restrict__ doff,
                                                        int32_t* __restrict__ dst,
                                                        const int32_t* __restrict__ heavy) {
  __shared__ int32_t s_off[RB_ROWS + 1];
  __shared__ int32_t s_src[RB_ROWS], s_len[RB_ROWS];
  __shared__ int32_t s_incl[BLOCK], s_nch[BLOCK];
  const int32_t nh = heavy[0];
  const int64_t G = gridDim.x;
  int64_t acc = 0;  // chunks of the batches before this one (the same in every thread)
  for (int32_t h0 = 0; h0 < nh; h0 += BLOCK) {
    const int32_t nb = min(BLOCK, nh - h0);
    int32_t mine = 0;
    if ((int32_t)threadIdx.x < nb) {
      const int64_t j0 = (int64_t)heavy[1 + h0 + threadIdx.x] * RB_ROWS;
      const int32_t cnt = (int32_t)min<int64_t>(RB_ROWS, nr - j0);
      mine = (doff[j0 + cnt] - doff[j0] + RB_CHUNK - 1) / RB_CHUNK;
    }
    s_nch[threadIdx.x] = mine;
    s_incl[threadIdx.x] = mine;
    __syncthreads();
    for (int o = 1; o < BLOCK; o <<= 1) {  // inclusive scan of the chunk counts
      const int32_t add = (int)threadIdx.x >= o ? s_incl[threadIdx.x - o] : 0;
      __syncthreads();
      s_incl[threadIdx.x] += add;
      __syncthreads();
    }
    for (int32_t k = 0; k < nb; k++) {
      const int32_t nch = s_nch[k];
      const int64_t first = acc + s_incl[k] - nch;  // number of the group's first chunk
      int64_t c0 = ((int64_t)blockIdx.x - first) % G;
      if (c0 < 0) c0 += G;
      if (c0 >= nch) continue;  // uniform across the workgroup
      const int64_t j0 = (int64_t)heavy[1 + h0 + k] * RB_ROWS;
      const int32_t cnt = (int32_t)min<int64_t>(RB_ROWS, nr - j0);
      const int32_t o0 = doff[j0];
      const int32_t total = doff[j0 + cnt] - o0;
      stage_group(rows, j0, cnt, sbeg, send, doff, o0, s_off, s_src, s_len);
      __syncthreads();
      for (int64_t ch = c0; ch < nch; ch += G)
        copy_group(cnt, o0, (int32_t)ch * RB_CHUNK, min(total, ((int32_t)ch + 1) * RB_CHUNK), s_off, s_src,
                   s_len, src, dst);
      __syncthreads();  // before the next group is staged
    }
    acc += s_incl[BLOCK - 1];
    __syncthreads();  // s_incl / s_nch are rewritten by the next batch
  }
}

// Index r in [lo, hi) of the changed row holding staged entry e (tmpoff[lo] <= e).
__device__ __forceinline__ int32_t row_in(const int32_t* __restrict__ tmpoff, int32_t lo, int32_t hi, int32_t e) {
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (tmpoff[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The staged entries in chunks of STG_CHUNK per workgroup.  The changed rows
// that hold each chunk's first and last entries come from crow (k_chunk_rows:
// two binary searches over all rows per chunk, every chunk in parallel); the
// chunk's row offsets are staged in LDS, where every entry of the chunk then
// finds its row.  (An entry-parallel search over all rows costs ~17 dependent
// loads per entry, and one thread searching for its workgroup's chunk before
// each chunk serialises 34 of them: a 100k-edge batch on R-MAT-23 stages 41 M
// entries per direction, hub rows hit by deletes being rewritten whole.)  A
// chunk spread over more than STG_ROWS rows (rows with nothing staged)
// searches that range in memory instead.
constexpr int32_t STG_CHUNK = 2048;
constexpr int32_t STG_ROWS = 1024;

__global__ __launch_bounds__(BLOCK) void k_chunk_rows(const int32_t* __restrict__ tmpoff, int32_t nr,
                                                      int32_t* __restrict__ crow) {
  const int32_t total = tmpoff[nr];
  const int64_t nch = ((int64_t)total + STG_CHUNK - 1) / STG_CHUNK;
  for (int64_t ch = (int64_t)blockIdx.x * BLOCK + threadIdx.x; ch < nch; ch += (int64_t)gridDim.x * BLOCK) {
    const int32_t E0 = (int32_t)(ch * STG_CHUNK), E1 = (int32_t)min<int64_t>(total, (ch + 1) * STG_CHUNK);
    const int32_t r0 = row_in(tmpoff, 0, nr, E0);
    crow[2 * ch] = r0;
    crow[2 * ch + 1] = row_in(tmpoff, r0, nr, E1 - 1);
  }
}

// One chunk's entries of a thread: e[t] = E0 + thread + t * BLOCK (ok[t] when
// inside the chunk), r[t] its changed row and off[t] that row's first staged
// entry; r0 / nrows: the chunk's rows.
constexpr int32_t STG_PER = STG_CHUNK / BLOCK;
struct StagedBatch {
  int32_t e[STG_PER], r[STG_PER], off[STG_PER];
  bool ok[STG_PER];
  int32_t r0, nrows;
};

template <class Body>
__device__ __forceinline__ void for_staged(const int32_t* __restrict__ tmpoff, const int32_t* __restrict__ crow,
                                           int32_t total, Body&& body) {
  __shared__ int32_t s_off[STG_ROWS + 1];
  __shared__ int32_t s_r0, s_nrows;
  for (int64_t c0 = (int64_t)blockIdx.x * STG_CHUNK; c0 < total; c0 += (int64_t)gridDim.x * STG_CHUNK) {
    const int32_t E0 = (int32_t)c0, E1 = (int32_t)min<int64_t>(total, c0 + STG_CHUNK);
    if (threadIdx.x == 0) {
      const int64_t ch = c0 / STG_CHUNK;
      const int32_t r0 = crow[2 * ch];
      s_r0 = r0;
      s_nrows = crow[2 * ch + 1] - r0 + 1;
    }
    __syncthreads();
    StagedBatch B;
    B.r0 = s_r0;
    B.nrows = s_nrows;
    const bool lds = B.nrows <= STG_ROWS;
    if (lds)
      for (int32_t t = threadIdx.x; t <= B.nrows; t += BLOCK) s_off[t] = tmpoff[B.r0 + t];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < STG_PER; t++) {
      const int32_t e = E0 + (int32_t)threadIdx.x + t * BLOCK;
      B.e[t] = e;
      B.ok[t] = e < E1;
      B.r[t] = B.r0;
      B.off[t] = 0;
      if (!B.ok[t]) continue;
      if (lds) {
        int32_t lo = 0, hi = B.nrows;
        while (hi - lo > 1) {
          const int32_t mid = (lo + hi) >> 1;
          if (s_off[mid] <= e) lo = mid;
          else hi = mid;
        }
        B.r[t] = B.r0 + lo;
        B.off[t] = s_off[lo];
      } else {
        B.r[t] = row_in(tmpoff, B.r0, B.r0 + B.nrows, e);
        B.off[t] = tmpoff[B.r[t]];
      }
    }
    body(B);  // (called by every thread: it may synchronise the workgroup)
    __syncthreads();  // s_off is restaged by the next chunk
  }
}

// Stage the old entries of the changed rows (each thread's loads issued together).
__global__ __launch_bounds__(BLOCK) void k_stage_rows(const int32_t* __restrict__ rows, int32_t nr,
                                                      const int32_t* __restrict__ tmpoff,
                                                      const int32_t* __restrict__ crow,
                                                      const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ adj,
                                                      int32_t* __restrict__ stage) {
  const int32_t total = tmpoff[nr];
  for_staged(tmpoff, crow, total, [&](const StagedBatch& B) {
    int32_t x[STG_PER];
#pragma unroll
    for (int t = 0; t < STG_PER; t++) x[t] = B.ok[t] ? adj[begin[rows[B.r[t]]] + (B.e[t] - B.off[t])] : 0;
#pragma unroll
    for (int t = 0; t < STG_PER; t++)
      if (B.ok[t]) stage[B.e[t]] = x[t];
  });
}

// First index i in [0, n) with a[i] >= key (n when none), over LDS.
__device__ __forceinline__ int32_t lb_u64_lds(const unsigned long long* a, int32_t n, unsigned long long key) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Survivors of the staged rows, written sorted to their new rooms:
// old index - #entries deleted before it + #inserted before it.  The deleted
// entries before key x of row r are dcum[first delete key >= x] - dcum[row's
// first delete key] (every copy of a deleted key counts).  A chunk inside one
// row (a hub row: most staged entries) with at most MRG_KEYS delete and insert
// keys searches them in LDS.
constexpr int32_t MRG_KEYS = 512;
__global__ __launch_bounds__(BLOCK) void k_merge_rows(const int32_t* __restrict__ rows, int32_t nr,
                                                      const int32_t* __restrict__ tmpoff,
                                                      const int32_t* __restrict__ crow,
                                                      const int32_t* __restrict__ stage,
                                                      const int32_t* __restrict__ newbeg,
                                                      const int4* __restrict__ rng,
                                                      const unsigned long long* __restrict__ del,
                                                      const int32_t* __restrict__ dcum,
                                                      const unsigned long long* __restrict__ ins,
                                                      int32_t* __restrict__ adj) {
  __shared__ unsigned long long s_del[MRG_KEYS], s_ins[MRG_KEYS];
  __shared__ int32_t s_dcum[MRG_KEYS + 1];
  const int32_t total = tmpoff[nr];
  for_staged(tmpoff, crow, total, [&](const StagedBatch& B) {
    int32_t k[STG_PER];
#pragma unroll
    for (int t = 0; t < STG_PER; t++) k[t] = B.ok[t] ? stage[B.e[t]] : 0;
    const int4 g0 = rng[B.r0];
    const int32_t nd = g0.y - g0.x, ni = g0.w - g0.z;
    const bool fast = B.nrows == 1 && nd <= MRG_KEYS && ni <= MRG_KEYS;  // (uniform)
    if (fast) {
      for (int32_t i = threadIdx.x; i < nd; i += BLOCK) s_del[i] = del[g0.x + i];
      for (int32_t i = threadIdx.x; i <= nd; i += BLOCK) s_dcum[i] = dcum[g0.x + i] - dcum[g0.x];
      for (int32_t i = threadIdx.x; i < ni; i += BLOCK) s_ins[i] = ins[g0.z + i];
    }
    __syncthreads();
    if (fast) {
      const unsigned long long hi = (unsigned long long)(uint32_t)rows[B.r0] << 32;
      const int32_t nb = newbeg[B.r0];
#pragma unroll
      for (int t = 0; t < STG_PER; t++) {
        if (!B.ok[t]) continue;
        const unsigned long long key = hi | (uint32_t)k[t];
        const int32_t dp = nd ? lb_u64_lds(s_del, nd, key) : 0;
        if (dp < nd && s_del[dp] == key) continue;  // deleted (each copy)
        const int32_t ip = ni ? lb_u64_lds(s_ins, ni, key) : 0;
        adj[nb + (B.e[t] - B.off[t]) - s_dcum[dp] + ip] = k[t];
      }
    } else {
#pragma unroll
      for (int t = 0; t < STG_PER; t++) {
        if (!B.ok[t]) continue;
        const int32_t r = B.r[t];
        const int4 g = rng[r];
        const unsigned long long key = ((unsigned long long)(uint32_t)rows[r] << 32) | (uint32_t)k[t];
        int32_t dents = 0;
        if (g.y > g.x) {
          const int32_t dp = (int32_t)lb_u64(del, g.x, g.y, key);
          if (dp < g.y && del[dp] == key) continue;  // deleted (each copy)
          dents = dcum[dp] - dcum[g.x];
        }
        const int32_t ip = g.w > g.z ? (int32_t)lb_u64(ins, g.z, g.w, key) : g.z;
        adj[newbeg[r] + (B.e[t] - B.off[t]) - dents + (ip - g.z)] = k[t];
      }
    }
  });
}

// Inserted entries at their merged rank in the new room: inserts before it +
// staged entries below it - deleted entries below it (an insert key is never a
// delete key: the filters drop deletes that are re-inserted and inserts that
// are present).
__global__ __launch_bounds__(BLOCK) void k_place_ins(const unsigned long long* __restrict__ ins, int64_t ni,
                                                     const int32_t* __restrict__ rowmark,
                                                     const int32_t* __restrict__ tmpoff,
                                                     const int32_t* __restrict__ stage,
                                                     const int32_t* __restrict__ newbeg,
                                                     const int4* __restrict__ rng,
                                                     const unsigned long long* __restrict__ del,
                                                     const int32_t* __restrict__ dcum,
                                                     int32_t* __restrict__ adj) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < ni; i += (int64_t)gridDim.x * BLOCK) {
    const unsigned long long key = ins[i];
    const int32_t u = key_hi(key), x = key_lo(key);
    const int32_t j = rowmark[u] - 1;
    const int4 g = rng[j];
    const int32_t t = (int32_t)i - g.z;
    const int32_t dl = g.y > g.x ? dcum[lb_u64(del, g.x, g.y, key)] - dcum[g.x] : 0;
    const int32_t s0 = tmpoff[j], s1 = tmpoff[j + 1];
    const int32_t ol = lb_i32(stage, s0, s1, x) - s0;
    adj[newbeg[j] + t + ol - dl] = x;
  }
}

// Install the new rows and clear the per-row scratch.
__global__ __launch_bounds__(BLOCK) void k_commit_rows(const int32_t* __restrict__ rows, int64_t nr,
                                                       const int32_t* __restrict__ newbeg,
                                                       const int32_t* __restrict__ newcap,
                                                       int32_t* __restrict__ rdel, int32_t* __restrict__ rins,
                                                       int32_t* __restrict__ rowmark, int32_t* __restrict__ begin,
                                                       int32_t* __restrict__ end, int32_t* __restrict__ cap) {
  for (int64_t j = (int64_t)blockIdx.x * BLOCK + threadIdx.x; j < nr; j += (int64_t)gridDim.x * BLOCK) {
    const int32_t u = rows[j];
    const int32_t nlen = end[u] - begin[u] - rdel[u] + rins[u];
    begin[u] = newbeg[j];
    end[u] = newbeg[j] + nlen;
    cap[u] = newcap[j];
    rdel[u] = 0;
    rins[u] = 0;
    rowmark[u] = 0;
  }
}

// Room sizes of every row for a fresh layout (rooms[n] = 0 for the scan).
__global__ __launch_bounds__(BLOCK) void k_room_sizes(int32_t n, const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ end,
                                                      int32_t* __restrict__ rooms, int2 sp) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v <= n; v += (int64_t)gridDim.x * BLOCK)
    rooms[v] = v < n ? room_of(end[v] - begin[v], sp) : 0;
}

__global__ __launch_bounds__(BLOCK) void k_room_ends(int32_t n, const int32_t* __restrict__ begin,
                                                     const int32_t* __restrict__ end,
                                                     const int32_t* __restrict__ nbeg,
                                                     int32_t* __restrict__ nend, int32_t* __restrict__ ncap) {
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK) {
    nend[v] = nbeg[v] + (end[v] - begin[v]);
    ncap[v] = nbeg[v + 1];
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

// Seeds S (heads) and T (tails) of inserts that can merge SCCs outside P.
__global__ __launch_bounds__(BLOCK) void k_ins_seeds(const unsigned long long* __restrict__ ins, int64_t ni,
                                                     const int32_t* __restrict__ canon,
                                                     const int32_t* __restrict__ aff,
                                                     const int32_t* __restrict__ color,
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
      use = color[u] != DONE && color[v] != DONE && (cu != canon[v] || aff[cu] != 0);
    }
    wave_append(use, v, S, &ctr[C_Q_OUT]);
    wave_append(use, u, T, &ctr[C_AUX_T]);
  }
}

// Outside P: colour 0 inside FW(S), OUTSIDE elsewhere; the live list becomes
// FW(S) in live[sel].
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
    if (v < n && color[v] != DONE) {
      in = bit_test(vf, (int32_t)v);
      color[v] = in ? 0 : OUTSIDE;
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

// Region X outside P: recompute state for members, keep the canonical label
// elsewhere; the live list becomes X in live[sel].
__global__ __launch_bounds__(BLOCK) void k_region(DevCtl* __restrict__ dc, int32_t n,
                                                  const int32_t* __restrict__ canon,
                                                  const int32_t* __restrict__ aff,
                                                  const uint32_t* __restrict__ vf,
                                                  const uint32_t* __restrict__ vb, bool have_fb,
                                                  int32_t* __restrict__ color, int32_t* __restrict__ label,
                                                  int32_t* __restrict__ live, int32_t sel, CanonKeys K) {
  using Out = LdsOut<0>;
  Out::init();
  for (int64_t base = (int64_t)blockIdx.x * BLOCK; base < n; base += (int64_t)gridDim.x * BLOCK) {
    Out::maybe_flush(live, &dc->ctr[C_LIST_OUT], BLOCK);
    const int64_t v = base + threadIdx.x;
    bool in = false;
    if (v < n && color[v] != DONE) {  // P keeps its new label
      const int32_t cv = canon[v];
      in = aff[cv] != 0 || (have_fb && bit_test(vf, (int32_t)v) && bit_test(vb, (int32_t)v));
      color[v] = in ? UNSET : DONE;
      label[v] = in ? -1 : cv;
      // the kept SCC's label is its canonical member, whose key is the SCC's least
      if (!in && cv == (int32_t)v) K.ckey[v] = canon_key(K.perm, (int32_t)v);
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

// Live vertices (outside P) back to one domain.
__global__ __launch_bounds__(BLOCK) void k_unset_live(const DevCtl* __restrict__ dc, int32_t* __restrict__ color,
                                                      const int32_t* __restrict__ L0,
                                                      const int32_t* __restrict__ L1) {
  const int32_t* __restrict__ list = LIVE_IN(dc, L0, L1);
  const int64_t n_in = (int64_t)dc->live.n;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_in; i += (int64_t)gridDim.x * BLOCK)
    color[list[i]] = UNSET;
}

__global__ __launch_bounds__(BLOCK) void k_clear_aff(const unsigned long long* __restrict__ del, int64_t nd,
                                                     const int32_t* __restrict__ canon,
                                                     int32_t* __restrict__ aff) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nd; i += (int64_t)gridDim.x * BLOCK)
    aff[canon[key_hi(del[i])]] = 0;
}

#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, st, __VA_ARGS__)

Dyn* dyn_of(scc_ctx* c) { return (Dyn*)c->dyn; }

// Library temp storage of at least `bytes`.
int ensure_tmp(scc_ctx* c, Dyn* d, size_t bytes) {
  if (bytes <= d->tmp_bytes) return SCC_OK;
  dev_free(c, d->tmp);
  d->tmp = nullptr;
  d->tmp_bytes = 0;
  int rc = dev_alloc(c, &d->tmp, bytes, "library temp");
  if (rc) return rc;
  d->tmp_bytes = bytes;
  return SCC_OK;
}

int ensure_dyn(scc_ctx* c, int64_t batch) {
  Dyn* d = dyn_of(c);
  if (!d) {
    d = new Dyn();
    c->dyn = d;
  }
  int rc;
  if (!d->dctr) {
    if ((rc = dev_alloc(c, (void**)&d->dctr, 8ull * D_NUM, "update counters"))) return rc;
    if ((rc = hip_check(c, hipHostMalloc((void**)&d->h_dctr, 8ull * D_NUM, 0), "pinned counters")))
      return rc;
  }
  if (d->n != c->n) {
    int32_t** rows[] = {&d->rdel, &d->rins, &d->rowmark, &d->aff};
    for (auto p : rows) {
      dev_free(c, *p);
      *p = nullptr;
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
      *p = nullptr;
      if ((rc = dev_alloc(c, (void**)p, 8ull * cap, "batch keys"))) return rc;
    }
    dev_free(c, d->flags);
    dev_free(c, d->src);
    dev_free(c, d->dst);
    dev_free(c, d->dcopies);
    dev_free(c, d->dcum);
    d->flags = nullptr;
    d->src = d->dst = d->dcopies = d->dcum = nullptr;
    if ((rc = dev_alloc(c, (void**)&d->flags, cap, "batch flags")) ||
        (rc = dev_alloc(c, (void**)&d->src, 4ull * cap, "batch src")) ||
        (rc = dev_alloc(c, (void**)&d->dst, 4ull * cap, "batch dst")) ||
        (rc = dev_alloc(c, (void**)&d->dcopies, 4ull * (cap + 1), "delete copies")) ||
        (rc = dev_alloc(c, (void**)&d->dcum, 4ull * (cap + 1), "delete copy offsets")))
      return rc;
    if (d->h_stage) (void)hipHostFree(d->h_stage);
    d->h_stage = nullptr;
    if ((rc = hip_check(c, hipHostMalloc((void**)&d->h_stage, 16ull * cap, 0), "pinned staging")))
      return rc;
    // changed rows: at most one per delete plus one per insert
    const int64_t cr = 2 * cap + 1;
    int32_t** ra[] = {&d->rows, &d->tmpoff, &d->newbeg, &d->newcap, &d->lens};
    for (auto p : ra) {
      dev_free(c, *p);
      *p = nullptr;
      if ((rc = dev_alloc(c, (void**)p, 4ull * (cr + 1), "changed-row arrays"))) return rc;
    }
    dev_free(c, d->rng);
    d->rng = nullptr;
    if ((rc = dev_alloc(c, (void**)&d->rng, 16ull * (cr + 1), "changed-row ranges"))) return rc;
    d->cap_rows = cr;
    size_t a = 0, b = 0, s = 0;
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, a, d->k_in, d->k_tmp, (int)cap);
    (void)hipcub::DeviceSelect::Flagged(nullptr, b, d->k_in, d->flags, d->del_f, (int32_t*)nullptr, (int)cap);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, s, d->lens, d->tmpoff, (int)(cr + 1));
    if ((rc = ensure_tmp(c, d, std::max(std::max(a, b), s) + 256))) return rc;
    d->cap_b = cap;
  }
  return SCC_OK;
}

// What ensure_dyn, the gapped layout and its rebuilds allocate, bounded from
// above, for a graph growing to m_max edges.
uint64_t gapped_entries(int64_t n, int64_t m) {
  const uint64_t t = (uint64_t)m + 2ull * (uint64_t)n + (uint64_t)m / 8;
  return t + std::max<uint64_t>(t / 8, 1u << 20);
}
uint64_t update_bytes(int64_t n, int64_t m, int64_t m_max, int64_t max_batch) {
  const uint64_t N = (uint64_t)n + 1;
  const int64_t mm = std::max(m, m_max);
  const uint64_t cap_b = (uint64_t)std::max<int64_t>(max_batch, 1024);
  const uint64_t g = gapped_entries(n, mm) + 4 * cap_b;
  uint64_t b = 4 * 4 * N;                       // per-row counters, marks, old-SCC flags
  b += 2 * 3 * 4 * N;                           // begin / end / room ends of a rebuilt layout
  b += 2 * 2 * 4 * g;                           // adj + radj, old and new during a rebuild
  b += 4 * (uint64_t)mm + 8 * ((uint64_t)mm / 2048 + 2);  // staged rows (at most every entry), chunk rows
  b += cap_b * (7 * 8 + 1 + 4 + 4);             // keys, flags, src, dst
  b += 2 * 4 * (cap_b + 1);                     // delete copies and their offsets
  b += 5 * 4 * (2 * cap_b + 2);                 // changed-row arrays
  b += 2 * 8 * cap_b + 8 * N + (1u << 20);      // sort / select / scan temp
  return b + 4096;
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

int pull_dctr(scc_ctx* c, Dyn* d) {
  int rc = hip_check(c, hipMemcpyAsync(d->h_dctr, d->dctr, 8ull * D_NUM, hipMemcpyDeviceToHost, c->stream),
                     "update counters");
  if (rc) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "update counters sync");
}

// (Re)build one direction in the gapped layout with fresh room for every row
// and an overflow area of at least min_over entries.  Reads the current rows
// (packed when *end is NULL), writes new begin / end / room ends / adj.
int regap_dir(scc_ctx* c, Dyn* d, bool fwd, int64_t min_over) {
  const int32_t n = c->n;
  int32_t*& beg = fwd ? c->begin : c->rbegin;
  int32_t*& en = fwd ? c->end : c->rend;
  int32_t*& adj = fwd ? c->adj : c->radj;
  int32_t*& cap = fwd ? d->cap_f : d->cap_r;
  int64_t& size = fwd ? d->size_f : d->size_r;
  const int32_t* cur_end = en ? en : beg + 1;
  int32_t *nbeg = nullptr, *nend = nullptr, *ncap = nullptr, *nadj = nullptr;
  int rc;
  if ((rc = dev_alloc(c, (void**)&nbeg, 4ull * ((uint64_t)n + 1), "gapped begin")) ||
      (rc = dev_alloc(c, (void**)&nend, 4ull * ((uint64_t)n + 1), "gapped end")) ||
      (rc = dev_alloc(c, (void**)&ncap, 4ull * ((uint64_t)n + 1), "gapped rooms"))) {
    dev_free(c, nbeg);
    dev_free(c, nend);
    dev_free(c, ncap);
    return rc;
  }
  auto fail = [&](int r) {
    dev_free(c, nbeg);
    dev_free(c, nend);
    dev_free(c, ncap);
    dev_free(c, nadj);
    return r;
  };
  // room sizes -> scan (in ncap as scratch) -> total
  const int2 sp = make_int2(c->room.row_spare, c->room.row_spare_eighths);
  LAUNCH(k_room_sizes, grid_for((int64_t)n + 1, 4096), c->stream, n, beg, cur_end, ncap, sp);
  size_t t = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, t, ncap, nbeg, n + 1);
  if ((rc = ensure_tmp(c, d, t + 256))) return fail(rc);
  t = d->tmp_bytes;
  if ((rc = hip_check(c, hipcub::DeviceScan::ExclusiveSum(d->tmp, t, ncap, nbeg, n + 1, c->stream), "room scan")))
    return fail(rc);
  int32_t total = 0;
  if ((rc = hip_check(c, hipMemcpyAsync(&total, nbeg + n, 4, hipMemcpyDeviceToHost, c->stream), "room total")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "room total sync")))
    return fail(rc);
  const int64_t over = std::max<int64_t>(
      std::max<int64_t>((int64_t)total * c->room.overflow_eighths / 8, c->room.overflow_min), min_over);
  const int64_t nsize = (int64_t)total + over;
  if (nsize >= INT32_MAX) return fail(set_error(c, SCC_ERR_LIMIT, "gapped layout would exceed int32 offsets"));
  if ((rc = dev_alloc(c, (void**)&nadj, 4ull * (uint64_t)nsize, "gapped adjacency"))) return fail(rc);
  // heavy-group list: count, then up to one entry per group
  int32_t* heavy = nullptr;
  if ((rc = dev_alloc(c, (void**)&heavy, 4ull * (uint64_t)((int64_t)n / RB_ROWS + 2), "regap heavy groups")))
    return fail(rc);
  launch_fill_i32(c->stream, heavy, 0, 1);
  LAUNCH(k_gather_rows, grid_for(n, 4096), c->stream, (const int32_t*)nullptr, (int64_t)n, beg, cur_end, adj,
         nbeg, nadj, heavy);
  LAUNCH(k_gather_heavy, 2048, c->stream, (const int32_t*)nullptr, (int64_t)n, beg, cur_end, adj, nbeg, nadj,
         (const int32_t*)heavy);
  LAUNCH(k_room_ends, grid_for(n, 4096), c->stream, n, beg, cur_end, nbeg, nend, ncap);
  if ((rc = hip_check(c, hipGetLastError(), "regap"))) {
    dev_free(c, heavy);
    return fail(rc);
  }
  rc = hip_check(c, hipStreamSynchronize(c->stream), "regap sync");
  dev_free(c, heavy);
  if (rc) return fail(rc);
  dev_free(c, beg);
  dev_free(c, en);
  dev_free(c, adj);
  dev_free(c, cap);
  beg = nbeg;
  en = nend;
  cap = ncap;
  adj = nadj;
  size = nsize;
  unsigned long long b = (unsigned long long)total;
  if ((rc = hip_check(c, hipMemcpyAsync(d->dctr + (fwd ? D_BUMP_F : D_BUMP_R), &b, 8, hipMemcpyHostToDevice,
                                        c->stream), "bump reset")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "bump reset sync");
}

// Apply one direction's sorted, filtered delete / insert lists to the gapped rows.
int apply_dir(scc_ctx* c, Dyn* d, bool fwd, const unsigned long long* del, int64_t nd,
              const unsigned long long* ins, int64_t ni) {
  hipStream_t s = c->stream;
  int rc;
  // changed rows and per-row counts
  LAUNCH(k_dctr_zero, 1, s, d->dctr, (1 << D_ROWS) | (1 << D_DELENT));
  if (nd) {
    LAUNCH(k_del_copies, grid_for(nd + 1), s, del, nd, fwd ? c->begin : c->rbegin, fwd ? row_end(c) : rrow_end(c),
           fwd ? c->adj : c->radj, d->dcopies, d->rdel);
    size_t t = d->tmp_bytes;
    if ((rc = hip_check(c, hipcub::DeviceScan::ExclusiveSum(d->tmp, t, d->dcopies, d->dcum, (int)(nd + 1), s),
                        "delete copies scan")))
      return rc;
    LAUNCH(k_store_total, 1, s, d->dcum + nd, d->dctr + D_DELENT);
  }
  if (ni) LAUNCH(k_row_count, grid_for(ni), s, ins, ni, d->rins, 1);
  if (nd) LAUNCH(k_collect_rows, grid_for(nd), s, del, nd, d->rowmark, d->rows, d->dctr + D_ROWS);
  if (ni) LAUNCH(k_collect_rows, grid_for(ni), s, ins, ni, d->rowmark, d->rows, d->dctr + D_ROWS);
  if ((rc = pull_dctr(c, d))) return rc;
  const int64_t nr = (int64_t)d->h_dctr[D_ROWS];
  if (nr == 0) return SCC_OK;
  for (int attempt = 0;; attempt++) {
    LAUNCH(k_dctr_zero, 1, s, d->dctr, (1 << D_OVERFLOW) | (1 << D_MAXLEN) | (1 << D_STAGED));
    LAUNCH(k_row_plan, grid_for(nr), s, d->rows, nr, fwd ? c->begin : c->rbegin, fwd ? c->end : c->rend,
           fwd ? d->cap_f : d->cap_r, d->rdel, d->rins, d->rowmark, d->lens, d->newbeg, d->newcap, del, nd,
           ins, ni, d->rng, d->dctr,
           fwd ? (int)D_BUMP_F : (int)D_BUMP_R, (long long)(fwd ? d->size_f : d->size_r),
           make_int2(c->room.row_spare, c->room.row_spare_eighths));
    size_t t = d->tmp_bytes;
    if ((rc = hip_check(c, hipcub::DeviceScan::ExclusiveSum(d->tmp, t, d->lens, d->tmpoff, (int)(nr + 1), s),
                        "staging scan")))
      return rc;
    LAUNCH(k_store_total, 1, s, d->tmpoff + nr, d->dctr + D_STAGED);
    if ((rc = pull_dctr(c, d))) return rc;
    if (!d->h_dctr[D_OVERFLOW]) break;
    if (attempt > 0) return set_error(c, SCC_ERR_LIMIT, "update: no room after rebuilding the layout");
    // out of overflow room: rebuild this direction with fresh room, plan again
    const int64_t staged = (int64_t)d->h_dctr[D_STAGED];
    if ((rc = regap_dir(c, d, fwd, 2 * (staged + ni + 3 * nr) + 4096))) return rc;
    d->rebuilds++;
  }
  c->maxdeg = std::max<int32_t>(c->maxdeg, (int32_t)d->h_dctr[D_MAXLEN]);
  const int64_t staged = (int64_t)d->h_dctr[D_STAGED];
  if (staged > d->stage_cap) {
    dev_free(c, d->stage);
    dev_free(c, d->crow);
    d->stage = nullptr;
    d->crow = nullptr;
    const int64_t cap = staged + staged / 2 + 4096;
    if ((rc = dev_alloc(c, (void**)&d->stage, 4ull * (uint64_t)cap, "staged rows")) ||
        (rc = dev_alloc(c, (void**)&d->crow, 8ull * (uint64_t)(cap / STG_CHUNK + 2), "staged chunk rows")))
      return rc;
    d->stage_cap = cap;
  }
  int32_t* beg = fwd ? c->begin : c->rbegin;
  int32_t* en = fwd ? c->end : c->rend;
  int32_t* adj = fwd ? c->adj : c->radj;
  (void)en;
  const unsigned sg = (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (staged + STG_CHUNK - 1) / STG_CHUNK));
  if (staged) {
    LAUNCH(k_chunk_rows, grid_for((staged + STG_CHUNK - 1) / STG_CHUNK), s, d->tmpoff, (int32_t)nr, d->crow);
    LAUNCH(k_stage_rows, sg, s, d->rows, (int32_t)nr, d->tmpoff, d->crow, beg, adj, d->stage);
    LAUNCH(k_merge_rows, sg, s, d->rows, (int32_t)nr, d->tmpoff, d->crow, d->stage, d->newbeg, d->rng, del,
           (const int32_t*)d->dcum, ins, adj);
  }
  if (ni)
    LAUNCH(k_place_ins, grid_for(ni), s, ins, ni, d->rowmark, d->tmpoff, d->stage, d->newbeg, d->rng, del,
           (const int32_t*)d->dcum, adj);
  LAUNCH(k_commit_rows, grid_for(nr), s, d->rows, nr, d->newbeg, d->newcap, d->rdel, d->rins, d->rowmark, beg, en,
         fwd ? d->cap_f : d->cap_r);
  return hip_check(c, hipGetLastError(), "apply rows");
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace


void destroy_dynamic(scc_ctx* c) {
  Dyn* d = dyn_of(c);
  if (!d) return;
  void* ptrs[] = {d->k_in, d->k_tmp, d->del_f, d->del_r, d->ins_f, d->ins_r, d->ins_sorted,
                  d->flags, d->src, d->dst, d->dcopies, d->dcum, d->rdel, d->rins, d->rowmark, d->aff, d->cap_f, d->cap_r,
                  d->rows, d->tmpoff, d->newbeg, d->newcap, d->lens, d->rng, d->stage, d->crow, d->dctr, d->tmp};
  for (void* p : ptrs) dev_free(c, p);
  if (d->h_stage) (void)hipHostFree(d->h_stage);
  if (d->h_dctr) (void)hipHostFree(d->h_dctr);
  delete d;
  c->dyn = nullptr;
}

// A new upload replaces the gapped layout: drop its room ends.
void dynamic_graph_replaced(scc_ctx* c) {
  Dyn* d = dyn_of(c);
  if (!d) return;
  dev_free(c, d->cap_f);
  dev_free(c, d->cap_r);
  d->cap_f = d->cap_r = nullptr;
  d->size_f = d->size_r = 0;
}

// Copy a gapped graph out packed (rows in vertex order).
int download_gapped(scc_ctx* c, int32_t* begin, int32_t* node_idx, int32_t* r_begin, int32_t* r_node_idx) {
  Dyn* d = dyn_of(c);
  const int32_t n = c->n;
  int rc;
  for (int dir = 0; dir < 2; dir++) {
    int32_t* ob = dir ? r_begin : begin;
    int32_t* oa = dir ? r_node_idx : node_idx;
    if (!ob && !oa) continue;
    const int64_t size = dir ? d->size_r : d->size_f;
    std::vector<int32_t> hb(n + 1), he(n + 1), ha(size);
    if ((rc = hip_check(c, hipMemcpyAsync(hb.data(), dir ? c->rbegin : c->begin, 4ull * n, hipMemcpyDeviceToHost,
                                          c->stream), "download")) ||
        (rc = hip_check(c, hipMemcpyAsync(he.data(), dir ? c->rend : c->end, 4ull * n, hipMemcpyDeviceToHost,
                                          c->stream), "download")) ||
        (rc = hip_check(c, hipMemcpyAsync(ha.data(), dir ? c->radj : c->adj, 4ull * size, hipMemcpyDeviceToHost,
                                          c->stream), "download")) ||
        (rc = hip_check(c, hipStreamSynchronize(c->stream), "download sync")))
      return rc;
    int64_t o = 0;
    for (int32_t v = 0; v < n; v++) {
      if (ob) ob[v] = (int32_t)o;
      for (int32_t i = hb[v]; i < he[v]; i++) {
        if (oa) oa[o] = ha[i];
        o++;
      }
    }
    if (ob) ob[n] = (int32_t)o;
  }
  return SCC_OK;
}

}  // namespace sccmi

using namespace sccmi;

extern "C" uint64_t scc_update_bytes_needed(int32_t n, int64_t m, int64_t m_max, int64_t max_batch) {
  return update_bytes(n, m, m_max, max_batch);
}

extern "C" int scc_update(scc_ctx* c, const int32_t* ins_src, const int32_t* ins_dst, int64_t n_ins,
                          const int32_t* del_src, const int32_t* del_dst, int64_t n_del,
                          scc_update_stats* stats) {
  if (!c) return SCC_ERR_ARG;
  if (int r0 = require_graph(c)) return r0;
  c->cut_out = c->cut_in = false;  // (the rows change)
  if (!c->labels_ready) {
    // phase-by-phase callers: labels are a full decomposition once no vertex is live
    int r0 = do_trim_compact(c);
    if (r0) return r0;
    if (c->n_live != 0)
      return set_error(c, SCC_ERR_STATE, "scc_update needs a finished decomposition (run first)");
    if ((r0 = do_canon(c, nullptr))) return r0;
  }
  if (int r1 = materialize_canon(c)) return r1;
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
  drop_order(c);  // the traversal copies describe the old rows
  Dyn* d = dyn_of(c);
  hipStream_t s = c->stream;

  // ---- 1. sorted, filtered batch lists ------------------------------------
  // (each list has its own pinned region, so no copy waits for the other; the
  // caller's ids become internal ids on the device, through iperm)
  auto upload_keys = [&](const int32_t* a, const int32_t* b, int64_t k, bool drop_loops, int region) -> int {
    if (k == 0) return SCC_OK;
    int32_t* h = d->h_stage + 2 * d->cap_b * region;
    std::memcpy(h, a, 4ull * k);
    std::memcpy(h + k, b, 4ull * k);
    int r;
    if ((r = hip_check(c, hipMemcpyAsync(d->src, h, 4ull * k, hipMemcpyHostToDevice, s), "batch upload")) ||
        (r = hip_check(c, hipMemcpyAsync(d->dst, h + k, 4ull * k, hipMemcpyHostToDevice, s), "batch upload")))
      return r;
    LAUNCH(k_make_keys, grid_for(k), s, d->src, d->dst, k, drop_loops, c->relabeled ? (const int32_t*)c->iperm : nullptr,
           d->k_in);
    return hip_check(c, hipGetLastError(), "batch keys");
  };
  // inserts: sorted keys (all, for the delete filter) then the absent ones
  int64_t ni = 0, nd = 0;
  if ((rc = upload_keys(ins_src, ins_dst, n_ins, true, 0))) return rc;
  if ((rc = sort_keys(c, d, d->k_in, d->ins_sorted, n_ins))) return rc;
  if (n_ins) {
    LAUNCH(k_flag_ins, grid_for(n_ins), s, d->ins_sorted, n_ins, c->begin, row_end(c), c->adj, d->flags);
    if ((rc = select_flagged(c, d, d->ins_sorted, d->ins_f, n_ins, &ni))) return rc;
  }
  if ((rc = upload_keys(del_src, del_dst, n_del, false, 1))) return rc;
  if ((rc = sort_keys(c, d, d->k_in, d->k_tmp, n_del))) return rc;
  if (n_del) {
    LAUNCH(k_flag_del, grid_for(n_del), s, d->k_tmp, n_del, d->ins_sorted, n_ins, c->begin, row_end(c), c->adj,
           d->flags);
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

  // ---- 2. apply to the rows (gapped layout, converted on first use) ----------
  auto ta = std::chrono::steady_clock::now();
  d->rebuilds = 0;
  if (!c->end) {
    if ((rc = regap_dir(c, d, true, 4 * d->cap_b)) || (rc = regap_dir(c, d, false, 4 * d->cap_b))) return rc;
  }
  // From here on a failure leaves the two directions, the edge count or the
  // labels out of step: the context refuses further graph calls until the
  // graph is uploaded again (cleared at the end of a successful update).
  c->graph_broken = true;
  c->labels_ready = false;
  if ((rc = apply_dir(c, d, true, d->del_f, nd, d->ins_f, ni))) return rc;
  const int64_t del_entries = nd ? (int64_t)d->h_dctr[D_DELENT] : 0;  // every copy of a deleted edge
  if ((rc = apply_dir(c, d, false, d->del_r, nd, d->ins_r, ni))) return rc;
  st.deleted = del_entries;
  c->m = c->m + ni - del_entries;
  c->m_cap = d->size_f;
  if ((rc = hip_check(c, hipStreamSynchronize(s), "apply sync"))) return rc;
  st.ms_apply = ms_since(ta);
  st.rebuilds = d->rebuilds;

  // ---- 3. affected region ------------------------------------------------------
  auto tr = std::chrono::steady_clock::now();
  if (nd) LAUNCH(k_mark_aff, grid_for(nd), s, d->del_f, nd, c->canon, d->aff);
  // P = SCC of the max in x out vertex of G', by one forward-backward step over
  // all of G'.  On graphs with a giant SCC this is the new giant, merges into it
  // included, and everything left to do lies outside it.
  // From the fresh state over G' (every vertex live, no colours), one trim
  // sweep on row lengths first: a vertex with an empty row in either direction
  // is a singleton SCC of G' (labelled itself, outside every search below, and
  // no inserted edge at it can close a cycle).  The sweep also leaves the
  // step's pivot key and clear reach bitmaps.
  c->colors_split = false;
  c->multi_colors = false;
  c->classes = false;
  c->fresh = false;
  c->rows_decide = false;
  {
    const int64_t words = ((int64_t)c->n + 31) / 32;
    launch_trim1_step(s, grid_for(c->n, (unsigned)c->tune.grid_cap), c->dc, gp(c), c->color, c->label, c->live[0],
                      c->live[1], 2, 0LL, ++c->trim_epoch, c->pkey, c->vf, 2 * words);
    if ((rc = pull_ctl(c))) return rc;
    c->pivot_ready = true;
  }
  int64_t psize = 0;
  if ((rc = do_fb_round(c, SCC_PIVOT_DEGREE_PRODUCT, false, -1, &psize, nullptr, nullptr, false, nullptr)))
    return rc;
  c->pivot_ready = false;  // (the step returns early, key unused, when the sweep left nothing live)
  st.giant = psize;
  LAUNCH(k_unset_live, grid_for(std::max<int64_t>(1, c->n_live)), s, c->dc, c->color, c->live[0], c->live[1]);
  if ((rc = zero_ctr(c))) return rc;
  // S (heads) -> scratch_i32, T (tails) -> W; both idle until the recompute
  int32_t* S = c->scratch_i32;
  int32_t* T = c->W;
  if (ni) LAUNCH(k_ins_seeds, grid_for(ni), s, d->ins_f, ni, c->canon, d->aff, c->color, S, T, c->ctr);
  if ((rc = pull_ctl(c))) return rc;
  const int64_t ns = (int64_t)c->h_ctr[C_Q_OUT];
  bool have_fb = false;
  const int64_t words = ((int64_t)c->n + 31) / 32;
  if (ns > 0 && c->n_live > 0) {
    // FW(S) within G' - P
    launch_fill_u32(s, c->vf, 0u, 2 * words);  // vf and vb (adjacent)
    if ((rc = traverse_from(c, true, false, S, c->ctr + C_Q_OUT, nullptr))) return rc;
    // BW(T) inside FW(S)
    LAUNCH(k_restrict_fw, grid_for(c->n), s, c->dc, c->n, c->vf, c->color, c->live[1], 1);
    if ((rc = pull_ctl(c))) return rc;
    if ((rc = traverse_from(c, false, false, T, c->ctr + C_AUX_T, nullptr))) return rc;
    have_fb = true;
  }
  LAUNCH(k_region, grid_for(c->n), s, c->dc, c->n, c->canon, d->aff, c->vf, c->vb, have_fb, c->color,
         c->label, c->live[0], 0, ck(c));
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
    if ((rc = run_method2(c, &o, &rs))) return rc;
    st.num_sccs = rs.num_sccs;
  } else {
    int64_t cnt = 0;
    if ((rc = do_canon(c, &cnt))) return rc;
    st.num_sccs = cnt;
  }
  st.ms_recompute = ms_since(tc);
  c->graph_broken = false;
  st.ms_total = ms_since(t0);
  if (stats) *stats = st;
  return SCC_OK;
}

extern "C" int scc_ctx_set_update_room(scc_ctx* c, const scc_update_room* r) {
  if (!c) return SCC_ERR_ARG;
  if (!r) {
    scc_update_room_default(&c->room);
    return SCC_OK;
  }
  if (r->row_spare < 0 || r->row_spare > (1 << 20) || r->row_spare_eighths < 0 || r->row_spare_eighths > 64 ||
      r->overflow_eighths < 0 || r->overflow_eighths > 64 || r->overflow_min < 0 ||
      r->overflow_min > (int64_t)1 << 30)
    return set_error(c, SCC_ERR_ARG, "bad update room");
  c->room = *r;
  return SCC_OK;
}

extern "C" void scc_update_room_default(scc_update_room* r) {
  if (!r) return;
  std::memset(r, 0, sizeof(*r));
  r->row_spare = 2;
  r->row_spare_eighths = 1;
  r->overflow_eighths = 1;
  r->overflow_min = 1 << 20;
}

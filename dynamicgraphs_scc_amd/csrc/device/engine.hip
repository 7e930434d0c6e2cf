// libsccmi engine: context, device memory, graph upload, phase drivers and the
// method-2 pipeline behind the C ABI of include/sccmi.h.
//
// Data layout in HBM (per context, n vertices, m edges, all int32):
//   begin[n+1] adj[m] rbegin[n+1] radj[m]            graph (CSR + reverse CSR)
//   color[n] label[n] canon[n] W[n]                  per-vertex state
//   live[2][n] q[2][n] cand[2][n] huge[n]            ping-pong vertex lists
//   vf[n/32] vb[n/32]                                forward / backward reach bitmaps
//   pkey[n+1] (u64), rep[3(n+1)], scratch[n+1]       per-partition slots
// A partition is identified by its colour; colours are -1 (the initial
// partition, slot n) or the id of a member vertex (slot = that id), so the
// per-partition arrays need no allocator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
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

int sync_ctr(scc_ctx* c) {
  int rc = hip_check(c, hipMemcpyAsync(c->h_ctr, c->ctr, sizeof(unsigned long long) * C_NUM,
                                       hipMemcpyDeviceToHost, c->stream),
                     "counter readback");
  if (rc) return rc;
  rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync");
  if (rc) return rc;
  return hip_check(c, hipGetLastError(), "kernel launch");
}

int zero_ctr(scc_ctx* c) {
  return hip_check(c, hipMemsetAsync(c->ctr, 0, sizeof(unsigned long long) * C_NUM, c->stream),
                   "counter reset");
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
         + 7 * N * 4          // live[2] q[2] cand[2] huge
         + 2 * words * 4      // vf vb
         + 8 * N + 12 * N     // pkey rep
         + 4 * N;             // scratch
}

int free_graph(scc_ctx* c) {
  int32_t** arrs[] = {&c->begin, &c->adj, &c->rbegin, &c->radj, &c->color, &c->label, &c->canon,
                      &c->W, &c->live[0], &c->live[1], &c->q[0], &c->q[1], &c->cand[0],
                      &c->cand[1], &c->huge, &c->rep, &c->scratch_i32};
  for (auto p : arrs) {
    dev_free(c, *p);
    *p = nullptr;
  }
  dev_free(c, c->vf);
  dev_free(c, c->vb);
  dev_free(c, c->pkey);
  c->vf = c->vb = nullptr;
  c->pkey = nullptr;
  c->n = 0;
  c->m = 0;
  c->m_cap = 0;
  c->n_live = 0;
  c->labels_ready = false;
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
  ALLOC(c->vf, 4 * words);
  ALLOC(c->vb, 4 * words);
  ALLOC(c->pkey, 8 * N);
  ALLOC(c->rep, 12 * N);
  ALLOC(c->scratch_i32, 4 * N);
#undef ALLOC
  c->n = n;
  c->m_cap = m_cap;
  return SCC_OK;
}

int reset_state(scc_ctx* c) {
  launch_fill_i32(c->stream, c->color, UNSET, c->n);
  launch_fill_i32(c->stream, c->label, -1, c->n);
  launch_iota(c->stream, c->live[0], c->n);
  c->cur = 0;
  c->n_live = c->n;
  c->labels_ready = false;
  c->multi_colors = false;
  return hip_check(c, hipGetLastError(), "state reset");
}

// ---- phases -----------------------------------------------------------------------

int do_trim_compact(scc_ctx* c) {
  int rc;
  if (c->n_live == 0) return SCC_OK;
  if ((rc = zero_ctr(c))) return rc;
  launch_compact_live(c->stream, c->live[c->cur], c->n_live, c->color, c->live[c->cur ^ 1], c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  c->cur ^= 1;
  c->n_live = (int64_t)c->h_ctr[C_LIST_OUT];
  return SCC_OK;
}

int force_pivot(scc_ctx* c, int32_t pivot) {
  launch_force_pivot(c->stream, pivot, c->n, c->color, c->pkey);
  return hip_check(c, hipGetLastError(), "force pivot");
}

int do_trim1(scc_ctx* c, int64_t stop, int64_t* trimmed) {
  int64_t total = 0;
  int rc;
  while (c->n_live > 0) {
    if ((rc = zero_ctr(c))) return rc;
    launch_trim1(c->stream, c->begin, c->adj, c->rbegin, c->radj, c->color, c->label,
                 c->live[c->cur], c->n_live, c->live[c->cur ^ 1], c->ctr);
    if ((rc = sync_ctr(c))) return rc;
    const int64_t t = (int64_t)c->h_ctr[C_TRIMMED];
    c->n_live = (int64_t)c->h_ctr[C_LIST_OUT];
    c->cur ^= 1;
    total += t;
    if (t <= stop) break;
  }
  if (trimmed) *trimmed = total;
  return SCC_OK;
}

int do_trim2(scc_ctx* c, int64_t* removed) {
  int rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_trim2(c->stream, c->begin, c->adj, c->rbegin, c->radj, c->color, c->label,
               c->live[c->cur], c->n_live, c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  if (removed) *removed = (int64_t)c->h_ctr[C_TRIMMED];
  return SCC_OK;
}

// Forward (fwd) or backward reachability from every partition's pivot,
// alternating top-down and bottom-up levels (Beamer's switch rule with the
// edge estimate m_f against m_u / 14 and back when the frontier shrinks below
// n_live / 24).
int traverse(scc_ctx* c, bool fwd, bool multi, int64_t* levels, bool timed) {
  const int32_t *off = fwd ? c->begin : c->rbegin, *nb = fwd ? c->adj : c->radj;
  const int32_t *roff = fwd ? c->rbegin : c->begin, *rnb = fwd ? c->radj : c->adj;
  uint32_t* V = fwd ? c->vf : c->vb;
  int rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_seed(c->stream, c->live[c->cur], c->n_live, c->n, c->color, c->pkey, V, c->q[0], c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  int64_t nq = (int64_t)c->h_ctr[C_Q_OUT];
  int qc = 0;
  const int32_t* cand_in = c->live[c->cur];
  int64_t ncand = c->n_live;
  int cc = 0;
  const double avg_deg = c->n ? (double)c->m / (double)c->n : 0.0;
  int64_t mf = 0;
  bool bu = false;
  int64_t prev_nq = nq;
  while (nq > 0) {
    if ((rc = check_cancel(c))) return rc;
    if (!bu) {
      if ((double)mf > (double)ncand * avg_deg / 14.0 && ncand > 0) bu = true;
    } else if (nq < prev_nq && nq * 24 < c->n_live) {
      bu = false;
    }
    if ((rc = zero_ctr(c))) return rc;
    if (timed) hipEventRecord(c->ev[0], c->stream);
    if (!bu) {
      launch_td(c->stream, off, nb, c->color, V, c->q[qc], nq, c->q[qc ^ 1], c->huge, c->ctr);
      if ((rc = sync_ctr(c))) return rc;
      const int64_t nh = (int64_t)c->h_ctr[C_HUGE];
      if (nh > 0) {
        launch_td_huge(c->stream, off, nb, c->color, V, c->huge, nh, c->q[qc ^ 1], c->ctr);
        if ((rc = sync_ctr(c))) return rc;
      }
      c->kernel_launches += nh > 0 ? 2 : 1;
    } else {
      launch_bu(c->stream, multi, roff, rnb, off, c->color, V, cand_in, ncand, c->cand[cc],
                c->q[qc ^ 1], c->ctr);
      if ((rc = sync_ctr(c))) return rc;
      cand_in = c->cand[cc];
      ncand = (int64_t)c->h_ctr[C_CAND_OUT];
      cc ^= 1;
      c->kernel_launches += 1;
    }
    if (timed) {
      hipEventRecord(c->ev[1], c->stream);
      hipEventSynchronize(c->ev[1]);
      float ms = 0;
      hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
      c->kernel_ms += ms;
    }
    prev_nq = nq;
    nq = (int64_t)c->h_ctr[C_Q_OUT];
    mf = (int64_t)c->h_ctr[C_Q_EDGES];
    qc ^= 1;
    if (levels) (*levels)++;
  }
  return SCC_OK;
}

// One forward-backward step for every live partition at once (multi) or for
// the single partition of colour -1 (the global step).
int do_fb_round(scc_ctx* c, int policy, bool multi, int32_t pivot_override, int64_t* scc,
                int32_t* pivot_out, int64_t* levels, bool timed) {
  int rc;
  if (c->n_live == 0) {
    if (scc) *scc = 0;
    return SCC_OK;
  }
  launch_reset_slots(c->stream, c->live[c->cur], c->n_live, c->n, c->color, c->pkey, c->rep);
  if (pivot_override >= 0) {
    if ((rc = force_pivot(c, pivot_override))) return rc;
  } else {
    launch_pivot_key(c->stream, c->live[c->cur], c->n_live, c->n, c->color, c->begin, c->rbegin,
                     policy, c->sched_scores, c->pkey);
  }
  const int64_t words = ((int64_t)c->n + 31) / 32;
  launch_fill_u32(c->stream, c->vf, 0u, words);
  launch_fill_u32(c->stream, c->vb, 0u, words);
  if (pivot_out) {
    unsigned long long k = 0;
    if ((rc = hip_check(c, hipMemcpyAsync(&k, c->pkey + c->n, sizeof(k), hipMemcpyDeviceToHost,
                                          c->stream), "pivot readback")))
      return rc;
    if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "stream sync"))) return rc;
    *pivot_out = k ? (int32_t)(0x7FFFFFFF - (int32_t)(k & 0x7FFFFFFFull)) : -1;
  }
  if ((rc = traverse(c, true, multi, levels, timed))) return rc;
  if ((rc = traverse(c, false, multi, levels, timed))) return rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_classify(c->stream, c->live[c->cur], c->n_live, c->n, c->color, c->label, c->vf, c->vb,
                  c->pkey, c->rep, c->live[c->cur ^ 1], c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  if (scc) *scc = (int64_t)c->h_ctr[C_SCC];
  c->cur ^= 1;
  c->n_live = (int64_t)c->h_ctr[C_LIST_OUT];
  launch_recolor(c->stream, c->live[c->cur], c->n_live, c->n, c->color, c->vf, c->vb, c->rep);
  return hip_check(c, hipGetLastError(), "recolor");
}

int do_wcc(scc_ctx* c, int64_t* parts) {
  int rc;
  launch_wcc_init(c->stream, c->live[c->cur], c->n_live, c->begin, c->W);
  for (int it = 0;; it++) {
    if ((rc = check_cancel(c))) return rc;
    if ((rc = zero_ctr(c))) return rc;
    launch_wcc_iter(c->stream, c->live[c->cur], c->n_live, c->begin, c->adj, c->color, c->W, c->ctr);
    if ((rc = sync_ctr(c))) return rc;
    if (c->h_ctr[C_CHANGED] == 0) break;
    if (it > 4 * (int64_t)c->n + 64) return set_error(c, SCC_ERR_LIMIT, "WCC did not converge");
  }
  if ((rc = zero_ctr(c))) return rc;
  launch_wcc_color(c->stream, c->live[c->cur], c->n_live, c->color, c->W, c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  if (parts) *parts = (int64_t)c->h_ctr[C_ROOTS];
  return SCC_OK;
}

int do_fb_parts(scc_ctx* c, int policy, bool trim_between, int64_t* rounds) {
  int64_t r = 0;
  int rc;
  while (c->n_live > 0) {
    if ((rc = check_cancel(c))) return rc;
    if (trim_between && r > 0) {
      if ((rc = do_trim1(c, 0, nullptr))) return rc;
      if (c->n_live == 0) break;
    }
    if ((rc = do_fb_round(c, policy, true, -1, nullptr, nullptr, nullptr, false))) return rc;
    r++;
  }
  if (rounds) *rounds = r;
  return SCC_OK;
}

int do_canon(scc_ctx* c, int64_t* count) {
  int rc;
  if ((rc = zero_ctr(c))) return rc;
  launch_canon(c->stream, c->n, c->label, c->scratch_i32, c->canon, c->ctr);
  if ((rc = sync_ctr(c))) return rc;
  if (count) *count = (int64_t)c->h_ctr[C_COUNT];
  c->labels_ready = true;
  return SCC_OK;
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
      (rc = dev_alloc(c, (void**)&c->ctr, sizeof(unsigned long long) * C_NUM, "counters")) ||
      (rc = hip_check(c, hipHostMalloc((void**)&c->h_ctr, sizeof(unsigned long long) * C_NUM, 0),
                      "pinned counters"))) {
    scc_ctx_destroy(c);
    return rc;
  }
  for (auto& e : c->ev) hipEventCreate(&e);
  *out = c;
  return SCC_OK;
}

int scc_ctx_destroy(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  destroy_dynamic(c);
  destroy_sched(c);
  free_graph(c);
  for (void* p : c->allocs) hipFree(p);
  c->allocs.clear();
  if (c->h_ctr) hipHostFree(c->h_ctr);
  for (auto& e : c->ev)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SCC_OK;
}

const char* scc_last_error(const scc_ctx* c) { return c ? c->err.c_str() : "null context"; }
uint64_t scc_ctx_bytes(const scc_ctx* c) { return c ? c->bytes : 0; }
uint64_t scc_bytes_needed(int32_t n, int64_t m) {
  return 8ull * (uint64_t)(m > 0 ? m : 1) + 8ull * ((uint64_t)n + 1) + state_bytes(n) +
         sizeof(unsigned long long) * C_NUM;
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
  hipSetDevice(c->device);
  destroy_dynamic(c);
  if ((rc = alloc_graph(c, n, m))) return rc;
  c->m = m;
  if ((rc = hip_check(c, hipMemcpyAsync(c->begin, begin, 4ull * (n + 1), hipMemcpyHostToDevice, c->stream), "upload")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->rbegin, r_begin, 4ull * (n + 1), hipMemcpyHostToDevice, c->stream), "upload")) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->adj, node_idx, 4ull * m, hipMemcpyHostToDevice, c->stream), "upload"))) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->radj, r_node_idx, 4ull * m, hipMemcpyHostToDevice, c->stream), "upload"))))
    return rc;
  if ((rc = reset_state(c))) return rc;
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
  hipSetDevice(c->device);
  destroy_dynamic(c);
  if ((rc = alloc_graph(c, n, m))) return rc;
  c->m = m;
  if ((rc = hip_check(c, hipMemcpyAsync(c->begin, d_begin, 4ull * (n + 1), hipMemcpyDeviceToDevice, c->stream), "d2d")) ||
      (rc = hip_check(c, hipMemcpyAsync(c->rbegin, d_r_begin, 4ull * (n + 1), hipMemcpyDeviceToDevice, c->stream), "d2d")) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->adj, d_node_idx, 4ull * m, hipMemcpyDeviceToDevice, c->stream), "d2d"))) ||
      (m > 0 && (rc = hip_check(c, hipMemcpyAsync(c->radj, d_r_node_idx, 4ull * m, hipMemcpyDeviceToDevice, c->stream), "d2d"))))
    return rc;
  if ((rc = reset_state(c))) return rc;
  c->ran = false;
  return hip_check(c, hipStreamSynchronize(c->stream), "upload sync");
}

int scc_graph_info(const scc_ctx* c, int32_t* n, int64_t* m) {
  if (!c) return SCC_ERR_ARG;
  if (n) *n = c->n;
  if (m) *m = c->m;
  return SCC_OK;
}

int scc_graph_download(scc_ctx* c, int32_t* begin, int32_t* node_idx, int32_t* r_begin,
                       int32_t* r_node_idx) {
  if (!c) return SCC_ERR_ARG;
  if (!c->begin) return set_error(c, SCC_ERR_STATE, "no graph");
  hipSetDevice(c->device);
  int rc = 0;
  if (begin && (rc = hip_check(c, hipMemcpyAsync(begin, c->begin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (r_begin && (rc = hip_check(c, hipMemcpyAsync(r_begin, c->rbegin, 4ull * (c->n + 1), hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (node_idx && c->m && (rc = hip_check(c, hipMemcpyAsync(node_idx, c->adj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  if (r_node_idx && c->m && (rc = hip_check(c, hipMemcpyAsync(r_node_idx, c->radj, 4ull * c->m, hipMemcpyDeviceToHost, c->stream), "download"))) return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "download sync");
}

int scc_state_reset(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  if (!c->begin) return set_error(c, SCC_ERR_STATE, "no graph uploaded");
  hipSetDevice(c->device);
  return reset_state(c);
}

int scc_state_set_colors(scc_ctx* c, const int32_t* colors) {
  if (!c || !colors) return SCC_ERR_ARG;
  if (!c->begin) return set_error(c, SCC_ERR_STATE, "no graph uploaded");
  for (int32_t v = 0; v < c->n; v++)
    if (colors[v] < -1 || colors[v] >= c->n)
      return set_error(c, SCC_ERR_ARG, "colour out of range at " + std::to_string(v));
  hipSetDevice(c->device);
  int rc;
  if ((rc = reset_state(c))) return rc;
  c->multi_colors = false;
  for (int32_t v = 0; v < c->n && !c->multi_colors; v++) c->multi_colors = colors[v] != -1;
  return hip_check(c, hipMemcpyAsync(c->color, colors, 4ull * c->n, hipMemcpyHostToDevice, c->stream),
                   "colours upload");
}

#define REQUIRE_GRAPH(c)                                                    \
  do {                                                                      \
    if (!(c)) return SCC_ERR_ARG;                                           \
    if (!(c)->begin) return set_error((c), SCC_ERR_STATE, "no graph uploaded"); \
    hipSetDevice((c)->device);                                              \
  } while (0)

int scc_trim1(scc_ctx* c, int64_t trim_stop, int64_t* trimmed) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  return do_trim1(c, trim_stop, trimmed);
}

int scc_fb_global(scc_ctx* c, int policy, int32_t pivot_override, int64_t* scc_size,
                  int32_t* pivot_out) {
  REQUIRE_GRAPH(c);
  if (pivot_override >= c->n) return set_error(c, SCC_ERR_ARG, "pivot out of range");
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_fb_round(c, policy, false, pivot_override, scc_size, pivot_out, nullptr, false);
}

int scc_trim2(scc_ctx* c, int64_t* removed) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  return do_trim2(c, removed);
}

int scc_wcc(scc_ctx* c, int64_t* parts) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_wcc(c, parts);
}

int scc_fb_partitions(scc_ctx* c, int policy, int trim_between, int64_t* rounds) {
  REQUIRE_GRAPH(c);
  c->labels_ready = false;
  int rc = do_trim_compact(c);
  if (rc) return rc;
  return do_fb_parts(c, policy, trim_between != 0, rounds);
}

void scc_run_opts_default(scc_run_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->pivot_policy = SCC_PIVOT_DEGREE_PRODUCT;
  o->trim_stop = 0;
  o->use_trim2 = 1;
  o->trim_between = 1;
  o->use_wcc = 1;
  o->time_phases = 0;
}

int scc_run_method2(scc_ctx* c, const scc_run_opts* opts_in, scc_run_stats* st_out) {
  REQUIRE_GRAPH(c);
  scc_run_opts opts;
  if (opts_in) opts = *opts_in;
  else scc_run_opts_default(&opts);
  scc_run_stats st;
  std::memset(&st, 0, sizeof(st));
  st.pivot_first = -1;
  c->labels_ready = false;
  c->kernel_ms = 0;
  c->kernel_launches = 0;
  const bool tp = opts.time_phases != 0;
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t* E = c->ev;
  int rc = 0;
  auto mark = [&](int i) {
    if (tp) hipEventRecord(E[2 + i], c->stream);
  };
  auto elapsed = [&](int a, int b) -> double {
    float ms = 0;
    if (tp) {
      hipEventSynchronize(E[2 + b]);
      hipEventElapsedTime(&ms, E[2 + a], E[2 + b]);
    }
    return ms;
  };
  do {
    mark(0);
    if ((rc = do_trim_compact(c))) break;
    if ((rc = do_trim1(c, opts.trim_stop, &st.trimmed_first))) break;
    mark(1);
    st.ms_trim1 = elapsed(0, 1);
    if (c->n_live == 0) break;
    if ((rc = check_cancel(c))) break;
    // the global step runs over colour -1; with caller-set colours it is a
    // partition round
    const bool multi = c->multi_colors;
    if ((rc = do_fb_round(c, opts.pivot_policy, multi, -1, &st.scc_first, &st.pivot_first,
                          &st.fb_levels, c->kernel_timing)))
      break;
    mark(2);
    st.ms_fb_global = elapsed(1, 2);
    int64_t t = 0;
    if ((rc = do_trim1(c, opts.trim_stop, &t))) break;
    st.trimmed_second = t;
    if (opts.use_trim2) {
      if ((rc = do_trim2(c, &t))) break;
      st.trimmed_second += t;
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
    if ((rc = do_fb_parts(c, opts.pivot_policy, opts.trim_between != 0, &st.fb_rounds))) break;
    mark(5);
    st.ms_fb_parts = elapsed(4, 5);
  } while (false);
  if (rc) return rc;
  mark(4);
  if ((rc = do_canon(c, &st.num_sccs))) return rc;
  mark(5);
  st.ms_canon = elapsed(4, 5);
  auto t1 = std::chrono::steady_clock::now();
  st.ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
  st.ms_fb_kernels = c->kernel_ms;
  st.fb_launches = c->kernel_launches;
  c->ran = true;
  if (st_out) *st_out = st;
  return SCC_OK;
}

int scc_labels_download(scc_ctx* c, int32_t* out, int canonical) {
  REQUIRE_GRAPH(c);
  if (!out) return SCC_ERR_ARG;
  int rc;
  if (canonical && !c->labels_ready) {
    if ((rc = do_canon(c, nullptr))) return rc;
  }
  if ((rc = hip_check(c, hipMemcpyAsync(out, canonical ? c->canon : c->label, 4ull * c->n,
                                        hipMemcpyDeviceToHost, c->stream), "labels download")))
    return rc;
  return hip_check(c, hipStreamSynchronize(c->stream), "labels sync");
}

int scc_colors_download(scc_ctx* c, int32_t* out) {
  REQUIRE_GRAPH(c);
  if (!out) return SCC_ERR_ARG;
  int rc;
  if ((rc = hip_check(c, hipMemcpyAsync(out, c->color, 4ull * c->n, hipMemcpyDeviceToHost, c->stream),
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
  *d_labels = c->canon;
  return SCC_OK;
}

int scc_ctx_set_kernel_timing(scc_ctx* c, int enable) {
  if (!c) return SCC_ERR_ARG;
  c->kernel_timing = enable != 0;
  return SCC_OK;
}

int scc_ctx_synchronize(scc_ctx* c) {
  if (!c) return SCC_ERR_ARG;
  hipSetDevice(c->device);
  return hip_check(c, hipStreamSynchronize(c->stream), "synchronize");
}

}  // extern "C"

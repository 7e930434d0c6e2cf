/*
 * sccserve.h — request loop over device-resident graphs (libsccserve.so).
 *
 * No reference counterpart: the reference runs one graph per process and
 * applies one update file per run (src/common_main.h:519-725).  Here a server
 * keeps many graphs resident in HBM and takes a stream of requests:
 *
 *   LOAD    upload a graph and decompose it (method 2)
 *   UPDATE  apply an insert/delete batch exactly (scc_update)
 *   LABELS  copy a graph's canonical labels out
 *   UNLOAD  free a graph
 *
 * Continuous batching: `slots` worker threads each run one request at a time;
 * every graph has its own context and HIP stream, so requests on different
 * graphs overlap on the device.  Requests on one graph run in submission
 * order.  A request is admitted when a slot is free, its graph is idle and its
 * HBM reserve fits the server budget: a LOAD reserves the graph's bytes plus
 * its update headroom (scc_update_bytes_needed), an UPDATE that grows a graph
 * beyond its headroom reserves more; memory freed by UNLOAD (or by failed
 * loads) admits waiting requests.  LOADs are admitted in arrival order.
 *
 * Coalescing: when an UPDATE starts, the UPDATEs queued right behind it on the
 * same graph (no other request between them) are merged into the same device
 * update, up to `coalesce` requests and max_batch edges per list.  Per edge
 * the last request's intent wins (within one request an insert beats a
 * delete, as in scc_update; inserted self-loops carry none), so the graph and
 * labels after the group equal those after applying its requests one by one.
 * A group ends before a request with an out-of-range vertex id (it runs alone
 * and fails alone), before one whose edges would exceed max_batch or the HBM
 * budget, and before one that inserts an edge an earlier member deletes (a
 * duplicated edge then keeps one copy one by one, all copies in one batch).
 * Every member reports its own latency; members share start and end times.
 *
 * Latency of a request = arrival (the submit time, or the arrival time the
 * caller passes, on the server clock scc_server_now_ms) to labels final
 * (the update's results are on the device and synchronised).
 *
 * Threading: every function may be called from any host thread.
 * Errors: negative scc_status codes (sccmi.h) and scc_server_last_error.
 */
#ifndef SCCSERVE_H
#define SCCSERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scc_server scc_server;

typedef enum scc_request_kind {
  SCC_REQ_LOAD = 1,
  SCC_REQ_UPDATE = 2,
  SCC_REQ_LABELS = 3,
  SCC_REQ_UNLOAD = 4
} scc_request_kind;

/* request states reported by scc_server_wait (besides negative errors) */
enum { SCC_REQ_DONE = 0, SCC_REQ_QUEUED = 1, SCC_REQ_RUNNING = 2 };

typedef struct scc_server_config {
  int32_t device;            /* HIP device */
  int32_t slots;             /* requests executing at once (worker threads) */
  int32_t queue_limit;       /* queued requests beyond which submits fail with SCC_ERR_LIMIT */
  int32_t pivot_policy;      /* scc_pivot_policy for loads and updates */
  uint64_t hbm_budget_bytes; /* sum of the graphs' reserves; 0 = 90% of the device's free memory */
  int64_t max_batch;         /* largest update list (inserts or deletes) accepted */
  double growth;             /* edge headroom reserved per graph, as a fraction of its edges */
  int32_t layout;            /* scc_layout of every loaded graph (SCC_LAYOUT_INPUT by default) */
  int32_t coalesce;          /* most UPDATE requests merged into one device update (<= 1: none; default 32) */
} scc_server_config;

typedef struct scc_request_info {
  int32_t kind;      /* scc_request_kind */
  int32_t status;    /* SCC_REQ_DONE / QUEUED / RUNNING or a negative scc_status */
  int32_t graph;
  int32_t group;     /* UPDATE: requests applied by the same device update (1 when alone) */
  int64_t num_sccs;  /* LOAD / UPDATE: SCCs after the request */
  int64_t inserted, deleted, affected;  /* UPDATE: effective changes, region size (of the whole
                                           group when group > 1) */
  double arrival_ms, start_ms, end_ms;  /* server clock */
  double ms_latency;                    /* end - arrival */
} scc_request_info;

typedef struct scc_server_stats {
  int64_t submitted, completed, failed, queued, running;
  int64_t resident_graphs;
  uint64_t reserved_bytes, budget_bytes;
  int64_t updates_done;          /* UPDATE requests completed since the last reset */
  int64_t update_edges;          /* inserted + deleted edges they applied */
  double update_p50_ms, update_p99_ms, update_max_ms, update_mean_ms;  /* latency */
  double busy_ms;                /* first arrival to last completion of those updates */
  double updates_per_s;          /* update_edges / busy_ms */
  double max_concurrency;        /* most requests seen executing at once */
} scc_server_stats;

void scc_server_config_default(scc_server_config* cfg);
int scc_server_create(const scc_server_config* cfg, scc_server** out);
/* Fails queued requests with SCC_ERR_CANCELLED, waits for running ones, frees
 * every graph. */
int scc_server_destroy(scc_server* s);
const char* scc_server_last_error(const scc_server* s);
double scc_server_now_ms(const scc_server* s);

/* The host arrays are copied at submit.  *graph receives the new graph's id
 * (usable for further submits at once); *ticket identifies the request. */
int scc_server_load(scc_server* s, int32_t n, int64_t m, const int32_t* begin, const int32_t* node_idx,
                    const int32_t* r_begin, const int32_t* r_node_idx, int32_t* graph, int64_t* ticket);
/* Lists are copied at submit.  arrival_ms < 0 means now. */
int scc_server_update(scc_server* s, int32_t graph, const int32_t* ins_src, const int32_t* ins_dst,
                      int64_t n_ins, const int32_t* del_src, const int32_t* del_dst, int64_t n_del,
                      double arrival_ms, int64_t* ticket);
/* out (n int32) must stay valid until the request ends. */
int scc_server_labels(scc_server* s, int32_t graph, int32_t* out, int64_t* ticket);
int scc_server_unload(scc_server* s, int32_t graph, int64_t* ticket);

/* Wait up to timeout_ms (< 0: forever) for a request; returns its status
 * (SCC_REQ_DONE, QUEUED or RUNNING on timeout, or the request's error) and
 * fills *info when given.  Finished requests stay queryable. */
int scc_server_wait(scc_server* s, int64_t ticket, double timeout_ms, scc_request_info* info);
/* Wait until no request is queued or running (timeout_ms < 0: forever).
 * Returns SCC_OK or SCC_REQ_QUEUED on timeout. */
int scc_server_drain(scc_server* s, double timeout_ms);
/* Cancel a queued request (SCC_ERR_STATE when it already started or ended). */
int scc_server_cancel(scc_server* s, int64_t ticket);

int scc_server_stats_get(scc_server* s, scc_server_stats* out);
int scc_server_stats_reset(scc_server* s);

/* Test hook: the graph's context reports a lost device after `after` more
 * HIP calls (scc_ctx_inject_fault).  Its later requests fail; other graphs
 * are unaffected. */
int scc_server_inject_fault(scc_server* s, int32_t graph, int after);

#ifdef __cplusplus
}
#endif
#endif /* SCCSERVE_H */

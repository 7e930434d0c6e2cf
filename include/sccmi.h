/*
 * sccmi.h — C ABI of the MI355X (gfx950) SCC engine, libsccmi.so.
 *
 * One context owns one device-resident graph (int32 CSR + reverse CSR), the
 * per-vertex colour and label arrays, scratch, and one HIP stream.  Every entry
 * point is called from one host thread per context; several contexts may run
 * concurrently (each on its own stream).  Calls never exit the process: they
 * return SCC_OK (0) or a negative scc_status and record a message readable with
 * scc_last_error().
 *
 * Conventions kept from the reference (src/scc.h:11-19, src/scc_color.cc:15-31):
 *   colour -1 = not yet assigned, -2 = SCC found, any other value = the
 *   partition a vertex currently belongs to; label = a member id of the
 *   vertex's SCC (pivot, self, or the smaller vertex of a 2-cycle), -1 before
 *   it is known.  scc_labels_download(..., canonical=1) returns min-id
 *   canonical labels, the form every parity check compares.
 *
 * Each phase entry point replaces one function of the reference's scc.h and
 * cites it; scc_run_method2 replaces the whole method-2 sequence.
 */
#ifndef SCCMI_H
#define SCCMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scc_ctx scc_ctx;

typedef enum scc_status {
  SCC_OK = 0,
  SCC_ERR_ARG = -1,         /* bad argument (sizes, null pointers, ids out of range) */
  SCC_ERR_HIP = -2,         /* a HIP runtime call failed */
  SCC_ERR_OOM = -3,         /* device allocation failed */
  SCC_ERR_BUDGET = -4,      /* request exceeds the context's HBM budget */
  SCC_ERR_STATE = -5,       /* call out of order (e.g. no graph uploaded) */
  SCC_ERR_DEVICE_LOST = -6, /* device unusable (also forced by the fault hook) */
  SCC_ERR_CANCELLED = -7,   /* stopped by scc_ctx_cancel or a deadline */
  SCC_ERR_LIMIT = -8        /* a size limit of the engine is exceeded */
} scc_status;

/* Pivot policies for forward-backward steps. */
typedef enum scc_pivot_policy {
  SCC_PIVOT_DEGREE_PRODUCT = 0, /* max(in-degree x out-degree), ties -> min id (plain heuristic) */
  SCC_PIVOT_MIN_ID = 1,         /* min id of the partition (src_CUDA/scc_cuda_fb_global.cu:426-440) */
  SCC_PIVOT_LEARNED = 2         /* scores from the learned scheduler (scc_sched_* below), the
                                   degree product among equal scores; scc_run_method2 applies it
                                   to the partition rounds and keeps the plain rule for its
                                   global step (scc_fb_global applies it as asked) */
} scc_pivot_policy;

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device`.  hbm_budget_bytes = 0 means no
 * budget; otherwise allocations that would exceed it fail with SCC_ERR_BUDGET.
 * (No reference counterpart: the reference state is process-global,
 *  src/scc.h:11-13.) */
int scc_ctx_create(int device, uint64_t hbm_budget_bytes, scc_ctx** out);
int scc_ctx_destroy(scc_ctx* ctx);
const char* scc_last_error(const scc_ctx* ctx);
/* Device bytes currently held by the context. */
uint64_t scc_ctx_bytes(const scc_ctx* ctx);
/* Device bytes a graph of n vertices and m edges needs in a context.  With room
 * beyond it (no budget, or a budget above it), the partition steps also keep a
 * copy of the rows cut to each class (8 (m + n) bytes, made at the first label
 * propagation over at least 1/8 of the vertices); without that room they follow
 * the graph's rows. */
uint64_t scc_bytes_needed(int32_t n, int64_t m);
/* Free and total device memory of a HIP device (hipMemGetInfo). */
int scc_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);
/* Change the context's HBM budget (0 = none).  Fails with SCC_ERR_BUDGET when
 * the context already holds more than the new budget. */
int scc_ctx_set_budget(scc_ctx* ctx, uint64_t hbm_budget_bytes);
/* Ask a running call on this context to stop at its next phase boundary
 * (thread-safe).  The call returns SCC_ERR_CANCELLED. */
void scc_ctx_cancel(scc_ctx* ctx);
/* Test hook: make the next `after` HIP calls succeed, then report
 * SCC_ERR_DEVICE_LOST (after < 0 disables).  For host-side failure tests. */
void scc_ctx_inject_fault(scc_ctx* ctx, int after);

/* ---- graph -------------------------------------------------------------- */

/* Copy an int32 CSR + reverse CSR to the device (rows need not be sorted).
 * Replaces the H2D load of src_CUDA/scc_cuda_graph.cu:4-25 and takes the arrays
 * gm_graph exposes (gm_graph/inc/gm_graph.h:154-160).  Validates offsets and
 * ids; resets the state (colours -1, labels -1). */
int scc_graph_upload(scc_ctx* ctx, int32_t n, int64_t m, const int32_t* begin,
                     const int32_t* node_idx, const int32_t* r_begin, const int32_t* r_node_idx);
/* Same, from device pointers already in HBM (copied device-to-device). */
int scc_graph_upload_device(scc_ctx* ctx, int32_t n, int64_t m, const int32_t* d_begin,
                            const int32_t* d_node_idx, const int32_t* d_r_begin,
                            const int32_t* d_r_node_idx);
/* Returns SCC_ERR_STATE (with n and m still filled in) when an earlier
 * scc_update failed after it began to rewrite the rows: every call that uses
 * the graph is then refused with SCC_ERR_STATE until the next upload. */
int scc_graph_info(const scc_ctx* ctx, int32_t* n, int64_t* m);
/* Vertex layout inside the context (no counterpart in the reference: a device
 * memory layout choice).  SCC_LAYOUT_DEGREE relabels the next uploaded graph
 * by decreasing degree class (bit length of in + out degree), caller id order
 * within a class, so that hubs' per-vertex state shares cache lines; it pays
 * on skewed (R-MAT) graphs.  Every entry point still takes and returns the
 * caller's vertex ids, labels and colours, and canonical labels stay the
 * minimum caller id of each component.  With a batch (scc_batch_set) the order
 * becomes per graph (graph, then decreasing degree class), so every graph
 * keeps its id range; that needs an upload without updates since. */
typedef enum {
  SCC_LAYOUT_INPUT = 0,  /* device ids are the caller's ids (default) */
  SCC_LAYOUT_DEGREE = 1, /* device ids by decreasing total degree class */
  SCC_LAYOUT_AUTO = 2    /* SCC_LAYOUT_DEGREE when the largest in- or out-degree exceeds 32 times
                            the mean degree (skewed graphs), else SCC_LAYOUT_INPUT */
} scc_layout;
/* Takes effect at the next scc_graph_upload / scc_graph_upload_device. */
int scc_ctx_set_layout(scc_ctx* ctx, int layout);
/* The layout the uploaded graph is held in (SCC_LAYOUT_INPUT or SCC_LAYOUT_DEGREE). */
int scc_graph_layout(const scc_ctx* ctx, int* layout);

/* Download the current device CSR (after updates).  Any pointer may be NULL. */
int scc_graph_download(scc_ctx* ctx, int32_t* begin, int32_t* node_idx, int32_t* r_begin,
                       int32_t* r_node_idx);

/* Reset colours to -1 and labels to -1 (src/scc_main.cc:96-118 prepare,
 * src/scc_color.cc:15-31 initialize_color). */
int scc_state_reset(scc_ctx* ctx);
/* Set the initial colour of every vertex (partitioned runs: vertices of
 * different colours are never in one SCC).  Colours must be -1 or in [0, n). */
int scc_state_set_colors(scc_ctx* ctx, const int32_t* colors);

/* ---- phases (each replaces one function of src/scc.h) -------------------- */

/* repeat_global_trim1 (src/scc.h:43, src/scc_trim1.cc:195-209): sweep until a
 * sweep trims <= trim_stop vertices (0 = fixpoint).  *trimmed = total. */
int scc_trim1(scc_ctx* ctx, int64_t trim_stop, int64_t* trimmed);
/* do_fw_bw_global_main (src/scc.h:81, src/scc_fb_global.cc:133-219): one
 * forward-backward step from one pivot over all live vertices of colour -1.
 * pivot_override >= 0 forces the pivot (src/scc_fb_global.cc:140-147). */
int scc_fb_global(scc_ctx* ctx, int policy, int32_t pivot_override, int64_t* scc_size,
                  int32_t* pivot_out);
/* do_global_trim2_new (src/scc.h:57, src/scc_trim2_new.cc:80-101): one pass. */
int scc_trim2(scc_ctx* ctx, int64_t* removed);
/* do_global_wcc + create_work_items_from_wcc (src/scc.h:64-65,
 * src/scc_weak.cc:128-271): recolour live vertices by WCC root. *parts = #roots. */
int scc_wcc(scc_ctx* ctx, int64_t* parts);
/* start_workers_fw_bw_dfs (src/scc.h:77, src/scc_fb_seq2.cc:292-351): resolve
 * every remaining partition by rounds of forward-backward from one pivot per
 * partition (all partitions at once), with trim1 between rounds when
 * trim_between != 0.  *rounds = rounds used. */
int scc_fb_partitions(scc_ctx* ctx, int policy, int trim_between, int64_t* rounds);
/* One colouring step over every live partition (no counterpart in the
 * reference's CPU path; the step of SCC_PARTS_COLOR): W = minimum tag over each
 * vertex's same-colour forward reach, each class recoloured by its root, and
 * each root's SCC (its forward reach inside its class) resolved.
 * *resolved = vertices resolved. */
int scc_color_round(scc_ctx* ctx, int64_t* resolved);
/* The partition phase with an explicit scc_partition_mode. */
int scc_partitions_resolve(scc_ctx* ctx, int policy, int mode, int trim_between, int64_t* rounds);

/* ---- whole pipeline ------------------------------------------------------ */

typedef struct scc_run_opts {
  int32_t pivot_policy;    /* scc_pivot_policy for every FB step */
  int32_t trim_stop;       /* trim1 ends after a sweep trimming <= trim_stop vertices (0 = fixpoint;
                              the reference CPU uses 100), or, when negative, <= -trim_stop
                              percent of the vertices the sweep examined */
  int32_t use_trim2;       /* reference USE_TRIM2 (src/scc_main.cc:46) */
  int32_t trim_between;    /* trim1 between partition FB rounds */
  int32_t use_wcc;         /* WCC colouring before the partition rounds */
  int32_t time_phases;     /* record per-phase HIP-event timings in stats */
  int32_t partition_mode;  /* scc_partition_mode: how the partition phase resolves what is left */
} scc_run_opts;

/* Partition phase (after the global step and WCC).  Every round resolves all
 * remaining partitions at once. */
typedef enum scc_partition_mode {
  SCC_PARTS_FB = 0,     /* one forward-backward step per partition per round (src/scc_fb_seq2.cc) */
  SCC_PARTS_COLOR = 1,  /* (default) WCC colouring, then each class's root SCC by a forward search;
                           the first step reuses the WCC phase's labels, and a step right after
                           another first splits the classes by the minimum over backward reach */
  SCC_PARTS_BOTH = 2,   /* a forward-backward step, then a colouring step, per round */
  SCC_PARTS_LEARNED = 3 /* each step chosen by the learned scheduler (scc_sched_*): a
                           forward-backward step or a colouring step (the plain rule unless the
                           network's vote is far from even).  The forward-backward steps take
                           pivots by pivot_policy (SCC_PIVOT_LEARNED for the network's
                           scores).  Without weights: SCC_PARTS_COLOR */
} scc_partition_mode;

typedef struct scc_run_stats {
  int64_t trimmed_first;   /* repeat_global_trim1 count */
  int64_t scc_first;       /* size of the global FB's SCC */
  int32_t pivot_first;     /* pivot of the global FB */
  int32_t chain_stop;      /* queued chain (scc_ctx_set_chain): -1 none, 0-4 the gate that stopped
                              it (the run went on phase by phase), 5 it ran to its end */
  int64_t trimmed_second;  /* trim1 + trim2 + trim1 after the global FB */
  int64_t wcc_parts;       /* partitions after WCC */
  int64_t fb_rounds;       /* partition FB rounds */
  int64_t num_sccs;        /* total SCC count (src/scc_main.cc:147-161) */
  int64_t fb_levels;       /* BFS levels of the global FB (fw + bw) */
  double ms_total;         /* host wall time of the run, ms */
  double ms_trim1, ms_fb_global, ms_trim2, ms_wcc, ms_fb_parts, ms_canon; /* HIP events */
  double ms_fb_kernels;    /* summed device time of the global FB traversal kernels */
  int64_t fb_launches;     /* traversal kernel launches in the global FB */
  /* reach sets of the global FB (filled when time_phases or kernel timing is on):
   * |F|, |B|, out-degrees summed over F, in-degrees summed over B, |F n B| */
  int64_t fb_f, fb_b, fb_ef, fb_eb, fb_fb;
} scc_run_stats;

/* Default options: degree-product pivot, trim_stop -70 (70 %), trim2 on, no trim between
 * partition rounds (measured faster on planted graphs), WCC on, no phase
 * timing, SCC_PARTS_COLOR. */
void scc_run_opts_default(scc_run_opts* o);

/* do_baseline_global_wcc_fb (src/scc_main.cc:272-351): trim1 -> global FB ->
 * trim1/trim2/trim1 -> WCC -> partition FB, then canonicalises labels on the
 * device.  Starts from the current state (call scc_state_reset first for a
 * from-scratch run).  stats may be NULL. */
int scc_run_method2(scc_ctx* ctx, const scc_run_opts* opts, scc_run_stats* stats);

/* ---- phase trace ------------------------------------------------------------
 * Every phase call appends one record (host wall time, live-set sizes), the
 * counts src/scc_main.cc:198-344 prints per phase.  Cleared by scc_state_reset
 * (and by scc_graph_upload).  The learned scheduler's features come from it. */
typedef enum scc_phase_kind {
  SCC_PHASE_TRIM1 = 1,
  SCC_PHASE_FB_GLOBAL = 2,
  SCC_PHASE_TRIM2 = 3,
  SCC_PHASE_WCC = 4,
  SCC_PHASE_FB_ROUND = 5,
  SCC_PHASE_COLOR_ROUND = 6,
  SCC_PHASE_CANON = 7
} scc_phase_kind;

typedef struct scc_phase_rec {
  int32_t kind;         /* scc_phase_kind */
  int32_t round;        /* partition round index, -1 before the rounds */
  int64_t live_before;  /* undecided vertices before the phase */
  int64_t live_after;   /* and after */
  int64_t aux;          /* trim sweeps, search levels, WCC iterations or SCCs found */
  double ms;            /* host wall time of the phase (it ends synchronised) */
} scc_phase_rec;

/* Copy up to cap records of the current trace; *n = records available. */
int scc_phase_trace(scc_ctx* ctx, scc_phase_rec* out, int64_t cap, int64_t* n);

/* ---- learned scheduler -------------------------------------------------------
 * (No reference counterpart: the reference's pivot is the first vertex of a
 * colour, src/scc_color.cc:64-93, and its phase order is fixed,
 * src/scc_main.cc:272-351.)  A 16 -> 64 -> 64 -> 2 MLP, BF16 matrix-core
 * arithmetic with FP32 accumulation, evaluated per live vertex: head 0 scores
 * the vertex as a pivot (estimated log-size of its SCC relative to its
 * partition), head 1 votes for a colouring step over a forward-backward step.
 * Parameters (FP32, row-major): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2].
 * Features (per sample, 16 floats): log2(1+outdeg)/16, log2(1+indeg)/16,
 * log2(1+outdeg*indeg)/32, log2(1+|partition|)/24, |partition|/live,
 * (outdeg-indeg)/(outdeg+indeg+1), log2(1+live)/24, live/n, classes flag,
 * last trim / forward-backward / colouring resolved fractions, log2(1+ns per
 * live vertex)/16 of the last forward-backward and colouring steps, the kind
 * of the last partition step (0 forward-backward, 1 colouring, 0.5 none yet),
 * (m/n)/64. */
#define SCC_SCHED_FEATURES 16
#define SCC_SCHED_HIDDEN 64
#define SCC_SCHED_HEADS 2
#define SCC_SCHED_NPARAMS 5378

/* Install weights (NULL disables the scheduler: SCC_PIVOT_LEARNED and
 * SCC_PARTS_LEARNED then fall back to the plain rules). */
int scc_sched_set_weights(scc_ctx* ctx, const float* params);
int scc_sched_get_weights(scc_ctx* ctx, float* params);
/* out[n][2] = network(X[n][16]) on the device. */
int scc_sched_forward(scc_ctx* ctx, const float* X, int64_t n, float* out);
/* Mean loss over n samples and its gradient w.r.t. every parameter.  Loss per
 * sample: M0 (out0 - Y0)^2 + M1 BCE(sigmoid(out1), Y1). */
int scc_sched_gradients(scc_ctx* ctx, const float* X, const float* Y, const float* M, int64_t n,
                        float* loss, float* grads);
/* Adam training on the device from the installed weights; *last_loss = loss of
 * the final mini-batch. */
int scc_sched_train(scc_ctx* ctx, const float* X, const float* Y, const float* M, int64_t n,
                    int32_t epochs, int32_t batch, float lr, uint64_t seed, float* last_loss);

typedef struct scc_rollout_stats {
  int64_t rounds;     /* partition rounds */
  int64_t agree;      /* rounds where the installed network chose the measured better step */
  int64_t num_sccs;
  double ms_total;    /* wall time of the rollout (both continuations measured each round) */
  double ms_best;     /* partition time from the first round, the faster of the two alternations */
} scc_rollout_stats;

/* Phase traces for training: a from-scratch decomposition in which, at every
 * partition round, each of the two steps is taken from the same state and
 * followed by the continuation policy until nothing is live (timed; the
 * installed network's choices, else alternating steps), the state is
 * restored, up to samples_per_round sampled live vertices are recorded
 * (features in X, head-1 target Y1 = 1 when the colouring-first continuation
 * finished sooner), and the run goes on with the better step (or the installed
 * network's choice), taking the other one with probability `explore`.  After
 * the run each sample's head-0 target is log2(1+|SCC|)/log2(1+|partition|).
 * Labels are exact. */
int scc_sched_rollout(scc_ctx* ctx, const scc_run_opts* opts, uint64_t seed, float explore,
                      int32_t samples_per_round, float* X, float* Y, float* M, int64_t cap,
                      int64_t* n, scc_rollout_stats* stats);

/* ---- results -------------------------------------------------------------- */

/* Labels (G_SCC, src/scc.h:12).  canonical != 0 returns min-id labels. */
int scc_labels_download(scc_ctx* ctx, int32_t* out, int canonical);
/* Colours (G_Color, src/scc.h:11). */
int scc_colors_download(scc_ctx* ctx, int32_t* out);
/* Number of SCCs (vertices whose label is themselves; src/scc_main.cc:147-161). */
int scc_count(scc_ctx* ctx, int64_t* count);
/* Device pointers of the canonical label / colour arrays (valid until the next
 * call on the context), for callers that keep results in HBM. */
int scc_labels_device(scc_ctx* ctx, const int32_t** d_labels);

/* ---- dynamic updates ------------------------------------------------------ */

typedef struct scc_update_stats {
  int64_t inserted;        /* edges actually added (absent before) */
  int64_t deleted;         /* edge entries actually removed (present before; every copy of a
                              duplicated edge counts) */
  int64_t affected;        /* vertices whose SCC was recomputed */
  int64_t num_sccs;        /* SCC count after the batch */
  double ms_total;         /* wall time of the call */
  double ms_apply;         /* CSR rebuild */
  double ms_region;        /* affected-region search */
  double ms_recompute;     /* SCC recompute inside the region */
  int64_t giant;           /* size of P, the SCC of the max in x out vertex, found first */
  int64_t rebuilds;        /* row-layout rebuilds the batch needed (per direction; the first
                              update's conversion to the gapped layout not counted) */
} scc_update_stats;

/* Apply one batch of edge inserts and deletes to the device graph and bring
 * the canonical labels up to date, recomputing only SCCs the batch can change.
 * A delete removes every copy of a duplicated edge (stats->deleted counts
 * the entries removed).  A failure after the rows began to change (an
 * allocation beyond the budget, a HIP error, a cancel) leaves the context
 * refusing graph calls with SCC_ERR_STATE until the graph is uploaded again
 * (scc_graph_info reports it); failures before that (bad ids, no labels yet)
 * leave the graph and labels as they were.
 * Replaces the reference's insert path (src/scc_incremental.cc:7-27 with
 * src/common_main.h:180-314, methods 5/6/7/11) and delete path
 * (src/scc_decremental.cc:7-24 with src/common_main.h:343-393, methods 8/9/10)
 * as one exact call.  Requires a completed run (labels known).  Inserting a
 * present edge or deleting an absent one is a no-op; self-loops are ignored. */
int scc_update(scc_ctx* ctx, const int32_t* ins_src, const int32_t* ins_dst, int64_t n_ins,
               const int32_t* del_src, const int32_t* del_dst, int64_t n_del,
               scc_update_stats* stats);

/* Room of the gapped row layout scc_update keeps (no reference counterpart:
 * the reference rebuilds graphs on the host).  Row u of length L gets
 * L + row_spare + row_spare_eighths * L / 8 entries; a row outgrowing its room
 * moves to an overflow area of max(overflow_eighths / 8 of all rooms,
 * overflow_min) entries; when that runs out the layout is rebuilt with fresh
 * room.  Defaults 2, 1, 1, 1 << 20.  Less room saves HBM and rebuilds more
 * often.  Takes effect at the next layout build (the first update after an
 * upload, or a rebuild); scc_update_bytes_needed assumes the defaults. */
typedef struct scc_update_room {
  int32_t row_spare;
  int32_t row_spare_eighths;
  int32_t overflow_eighths;
  int32_t pad;
  int64_t overflow_min;
} scc_update_room;
void scc_update_room_default(scc_update_room* r);
/* NULL restores the defaults. */
int scc_ctx_set_update_room(scc_ctx* ctx, const scc_update_room* r);

/* Upper bound of the device bytes scc_update adds to a context that holds a
 * graph of n vertices and m edges (scc_bytes_needed(n, m)), while the graph
 * grows to at most m_max edges through batches of at most max_batch edges per
 * direction: spare adjacency arrays, per-row counters, batch buffers.  (No
 * reference counterpart: the reference rebuilds graphs on the host.) */
uint64_t scc_update_bytes_needed(int32_t n, int64_t m, int64_t m_max, int64_t max_batch);

/* ---- batched graphs -----------------------------------------------------------
 * (No reference counterpart: the reference decomposes one graph per process.)
 * Many independent graphs decomposed by the same launches: upload their
 * disjoint union (graph g owns vertex ids [offsets[g], offsets[g+1]), no edge
 * leaves its graph), name the boundaries, then run.  Each graph starts as its
 * own partition, so the first forward-backward step takes one pivot per graph.
 * Canonical labels are union ids (graph-local id = label - offsets[g]). */
/* Set (num_graphs > 0) or clear (0) the batch; fails with SCC_ERR_ARG when an
 * edge crosses between graphs.  A new upload clears it. */
int scc_batch_set(scc_ctx* ctx, int32_t num_graphs, const int32_t* offsets);
/* Fresh decomposition of every graph of the batch (method 2 from per-graph partitions). */
int scc_batch_run(scc_ctx* ctx, const scc_run_opts* opts, scc_run_stats* stats);
/* SCC count of each graph after a run (num_graphs entries). */
int scc_batch_counts(scc_ctx* ctx, int64_t* counts);

/* ---- launch-shape tuning --------------------------------------------------------
 * (No reference counterpart: generalises the fixed switch points of
 * gm_graph/inc/gm_bfs_template.h:424, :841-842.) */
typedef struct scc_tuning {
  int32_t light_row; /* search levels: a lane scans the first light_row entries of a row alone, the
                        wave the rest; one of 4, 8, 16, 32, 64 (32) */
  int32_t huge_row;  /* top-down frontier rows at least this long are split over the grid (8192) */
  int32_t grid_cap;  /* workgroups per sweep launch (640) */
  int32_t wcc_dense; /* label propagation iterations are pull sweeps over the whole live set while
                        the frontier holds more than 1/wcc_dense of it, else frontier pushes (2;
                        0 = the default) */
  float bu_alpha;    /* top-down -> bottom-up when frontier edges > candidates' edges / bu_alpha (14);
                        searches serving many partitions at once use bu_alpha * 2/7 */
  float td_beta;     /* bottom-up -> top-down when frontier < live / td_beta and shrinking (24);
                        searches serving many partitions at once use td_beta / 3 */
} scc_tuning;
int scc_ctx_get_tuning(scc_ctx* ctx, scc_tuning* t);
/* NULL restores the defaults. */
int scc_ctx_set_tuning(scc_ctx* ctx, const scc_tuning* t);
/* Coordinate descent over the knobs, timing full decompositions of the uploaded
 * graph (or batch): median of reps runs per candidate; a change is kept when it
 * gains more than 2 %.  Installs the best tuning, leaves the labels of a run
 * with it, and reports the best and the default median wall times (ms). */
int scc_autotune(scc_ctx* ctx, const scc_run_opts* opts, int32_t reps, scc_tuning* best,
                 double* best_ms, double* default_ms);

/* Per-level record of the last reachability search the context ran (up to 64
 * levels; a forward-backward step runs both searches together and records the
 * forward levels, then the backward ones): out[4*i..] = {mode (bit 0: 0
 * top-down, 1 bottom-up; bit 8: backward search), input size (top-down:
 * frontier + (split rows << 32); bottom-up: candidates), vertices admitted,
 * row-length sum of the admitted}.  *levels = levels recorded.  For analysis. */
int scc_level_log(scc_ctx* ctx, int64_t* out, int32_t* levels);

/* ---- timing helpers (HIP events on the context's stream) ------------------ */

/* Record per-launch durations of the global-FB traversal kernels during the
 * next runs (enable != 0).  Sum and count are returned in scc_run_stats. */
int scc_ctx_set_kernel_timing(scc_ctx* ctx, int enable);
/* Block until the context's stream is idle. */
int scc_ctx_synchronize(scc_ctx* ctx);

/* Queued chain (no reference counterpart: the reference returns to the host
 * after every phase).  scc_run_method2 from a fresh state on the plain
 * schedule (degree-product pivots, no phase or kernel timing, partition mode
 * not learned) queues the first trim loop, the global step, the trims after it
 * and the canonical labels as one chain of launches and reads the control
 * block back once.  A gate launch after each phase stops the chain when the
 * phase needs more launches than were queued for it, or when vertices are
 * still live after the chain's last phase; the call then continues phase by
 * phase with read-backs, as without the chain.  The chain's length (shape: its
 * last gate, 0 = first trim loop, 1 = global step, 2 = second trim loop, 3 =
 * trim2 and the third trim loop), level launches and sweeps per trim loop
 * follow the last decomposition on the context; values < 0 keep them.
 * Results do not depend on any of this.  enable = 0 turns the chain off. */
int scc_ctx_set_chain(scc_ctx* ctx, int enable, int32_t shape, int32_t level_launches, int32_t sweeps);

/* ---- decomposition that allocates nothing ---------------------------------------
 * (No counterpart in the reference.  For callers that cannot allocate device
 * memory: code preparing HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.)  The canonical SCC labels of a graph already in HBM, as
 * scc_graph_upload_device + scc_run_method2 + scc_labels_download(canonical)
 * give them on a context, computed with every scratch buffer (per-vertex
 * state, lists, bitmaps, control block) in one static device array of
 * SCC_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends.  No hipMalloc, hipFree, pinned or library allocation is made.
 * The caller passes the int32 CSR + reverse CSR (device pointers, a valid CSR
 * as scc_graph_upload_device takes) and d_labels (n int32, device); work runs
 * on `stream` (a hipStream_t, NULL for the null stream) and the call returns
 * once the labels are written.  n above SCC_NOALLOC_MAX_N fails with
 * SCC_ERR_LIMIT.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race and must be serialised by the caller.  A message for
 * a failure is written to err (err_len bytes, may be NULL). */
#define SCC_NOALLOC_MAX_N 262144
#define SCC_NOALLOC_SCRATCH_BYTES (32u << 20)
int scc_run_device_noalloc(int device, void* stream, int32_t n, int64_t m, const int32_t* d_begin,
                           const int32_t* d_node_idx, const int32_t* d_r_begin, const int32_t* d_r_node_idx,
                           int32_t* d_labels, int64_t* num_sccs, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* SCCMI_H */

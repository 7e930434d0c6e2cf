// Internal interface shared by the engine's translation units.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace sccmi {

int set_error(scc_ctx* c, int code, const std::string& msg);
int hip_check(scc_ctx* c, hipError_t e, const char* what);
int dev_alloc(scc_ctx* c, void** p, uint64_t bytes, const char* what);
void dev_free(scc_ctx* c, void* p);
GraphPtrs gp(const scc_ctx* c);
CanonKeys ck(const scc_ctx* c);  // canonical-key pointers of the context
int pull_ctl(scc_ctx* c);   // control block -> host (syncs the stream)
// pull_ctl in two halves: queue the copy (and an event), later wait for it
int ctl_readback_start(scc_ctx* c);
int ctl_readback_wait(scc_ctx* c);
int zero_ctr(scc_ctx* c);   // zero DevCtl::ctr (queued)
int set_live(scc_ctx* c, int sel, int64_t n);
int check_cancel(scc_ctx* c);
int reset_state(scc_ctx* c);
int start_run(scc_ctx* c);  // reset_state, then the batch's per-graph colours when a batch is set
int alloc_graph(scc_ctx* c, int32_t n, int64_t m_cap);
int free_graph(scc_ctx* c);
int build_order(scc_ctx* c);  // traversal copies of the rows, hubs first (order.hip)
void drop_order(scc_ctx* c);

// vertex layout (layout.hip): internal relabelling and the mappings at the boundary
int build_layout(scc_ctx* c);  // relabel the uploaded rows when SCC_LAYOUT_DEGREE is requested
void drop_layout(scc_ctx* c);
int layout_to_caller(scc_ctx* c, const int32_t* in, bool map_values, int32_t* out_dev);
int layout_colors_to_internal(scc_ctx* c, const int32_t* colors_host);
// re-relabel by (batch graph, decreasing degree) once c->d_boff holds k graphs' caller ranges
int layout_segment(scc_ctx* c, int32_t k);
void layout_rows_to_caller(const scc_ctx* c, const std::vector<int32_t>& ib, const std::vector<int32_t>& ia,
                           int32_t* ob, int32_t* oa);

// phases (operate on the live list of the control block)
int do_trim_compact(scc_ctx* c);  // drop colour -2 entries from the live list
int do_trim1(scc_ctx* c, int64_t stop, int64_t* trimmed);
int do_trim2(scc_ctx* c, int64_t* removed);
int force_pivot(scc_ctx* c, int32_t pivot);
int run_levels(scc_ctx* c, bool fwd, bool multi, int64_t* levels, bool timed);
int traverse_both(scc_ctx* c, bool multi, int64_t* levels, bool timed, const std::function<void()>& post);
int traverse_from(scc_ctx* c, bool fwd, bool multi, const int32_t* seeds,
                  const unsigned long long* d_count, int64_t* levels);
int do_fb_round(scc_ctx* c, int policy, bool multi, int32_t pivot_override, int64_t* scc,
                int32_t* pivot_out, int64_t* levels, bool timed, int64_t* sets);
int do_wcc(scc_ctx* c, int64_t* parts);
int do_color_round(scc_ctx* c, int64_t* scc);
int do_fb_parts(scc_ctx* c, int policy, int mode, bool trim_between, int64_t* rounds);
int do_canon(scc_ctx* c, int64_t* count);
// The internal-order canonical labels (c->canon) when the last decomposition
// left only the caller-order ones (a chain in the degree layout).
int materialize_canon(scc_ctx* c);
// Row heads of the traversal rows (order.hip), after build_order / the layout.
int build_heads(scc_ctx* c);
int run_method2(scc_ctx* c, const scc_run_opts* opts, scc_run_stats* st);  // scc_run_method2 unchecked

// optional subsystems (dynamic.hip, sched.hip)
void destroy_dynamic(scc_ctx* c);
void destroy_sched(scc_ctx* c);
void dynamic_graph_replaced(scc_ctx* c);
// learned scheduler (sched.hip)
bool sched_ready(const scc_ctx* c);
int sched_choose(scc_ctx* c, int* action);  // 0: forward-backward step, 1: colouring step
int sched_pivots(scc_ctx* c);               // per-vertex pivot scores into c->sched_scores
int download_gapped(scc_ctx* c, int32_t* begin, int32_t* node_idx, int32_t* r_begin, int32_t* r_node_idx);

}  // namespace sccmi

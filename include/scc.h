/*
 * scc.h — the reference's C++ SCC entry points (src/scc.h:11-92, with the work
 * item and queue of src/my_work_queue.h:9-26 it includes), served by the
 * MI355X engine.  Every name the reference's scc.h declares is declared here
 * with the reference's signature, so a driver written against the reference's
 * call sequence compiles against this header and links against libsccshim.so.
 * The graph container is sccmi::gm_graph (same public CSR members as the
 * reference's gm_graph, gm_graph/inc/gm_graph.h:154-160), and every phase runs
 * on the GPU through the C ABI in sccmi.h.
 *
 * State: one process-global context bound to the graph the first phase call
 * sees (src/scc.h:11-13 keeps G_Color / G_SCC / G_num_nodes as globals).
 * G_Color and G_SCC are host mirrors refreshed after every phase call (turn
 * that off with scc_shim_set_mirror(false) and call scc_shim_sync() when the
 * arrays are needed); the shim owns them (a driver's own G_SCC allocation is
 * replaced at the first phase call).  Writes into the mirrors are not sent to
 * the device.  Errors print a message and exit(EXIT_FAILURE), as the reference
 * does (src/common_main.h:459-473).
 *
 * Where the engine differs from the reference (each entry below says which):
 *   - labels are member ids of the component, as in the reference, but not
 *     always the same member (the reference's pivots depend on thread timing);
 *     the SCC partition, and so the min-id canonical labels, are identical;
 *   - colours are partition ids chosen as a member vertex, so a step resolves
 *     every live partition at once: calls that the reference applies to one
 *     partition or one work item act on all of them (their results are exact;
 *     more vertices may be resolved than the reference would at that point);
 *   - Tarjan is served by the device decomposition: do_tarjan_all and its
 *     variants give the reference's SCC partition, not its DFS-root labels.
 */
#ifndef SCCMI_SCC_H
#define SCCMI_SCC_H

#include <stdint.h>

#include <unordered_set>
#include <utility>
#include <vector>

#include "sccmi_gm_graph.h"

typedef sccmi::node_t node_t;
typedef sccmi::edge_t edge_t;
using sccmi::gm_graph;

#define MAX_THREADS 1024  // src/scc.h:14

extern int32_t* G_Color;  // src/scc.h:11
extern node_t* G_SCC;     // src/scc.h:12
extern int32_t G_num_nodes;  // src/scc.h:13

// ---- work items and their queue (src/my_work_queue.h:9-26) ----
// A host-side queue of partition descriptions.  The device resolves every live
// partition in each step, so start_workers_fw_bw(_dfs) empty the queue (and
// delete its items) whatever it holds.
typedef std::unordered_set<node_t> NODE_SET;
class my_work {
 public:
  int color;           // colour of the partition
  int count;           // its vertex count
  NODE_SET* color_set; // its vertices, or NULL
  int depth;
};
bool is_work_q_empty_from_seq_context();
my_work* work_q_fetch(int thread_id);  // NULL when empty
void work_q_fetch_N(int thread_id, int N, std::vector<my_work*>& works);
void work_q_put(int thread_id, my_work* work);
void work_q_put_all(int thread_id, std::vector<my_work*>& work);
void work_q_init(int num_threads);
void work_q_print_max_depth();
int work_q_size();

// ---- colour (src/scc.h:18-26) ----
void initialize_color();  // G_Color = -1 and G_SCC = -1 on the device (src/scc_color.cc:15-31)
void finalize_color();
// Fresh colour ids from one process-wide counter (src/scc_color.cc:38-54 hands
// them out in per-thread chunks); the first is 0.  Thread-safe.
int get_new_color();
// The counter's current value: -1 before any get_new_color (src/scc_color.cc:56-61).
int get_curr_color();
// First vertex of `color` in get_compact_trim_targets() order, i.e. the minimum
// id, or NIL_NODE (src/scc_color.cc:64-93, whose first trim target depends on
// thread timing).  Reads the device colours.
node_t choose_pivot_from_color(gm_graph& G, int color);
// The vertices of `color` (src/scc_color.cc:95-113); the caller owns the set.
NODE_SET* generate_compact_set(gm_graph& G, int color);

// ---- Tarjan (src/scc.h:28-34), served by the device decomposition ----
void initialize_tarjan();
void finalize_tarjan();
// Every vertex not yet done gets its SCC (src/scc_tarjan.cc:137-149); labels are
// member ids, canonical partition equal to the reference's.
void do_tarjan_all(gm_graph& G);
// The reference resolves the SCCs reachable from `root` inside `root_color`
// (src/scc_tarjan.cc:99-135); here every undecided vertex is resolved.
void do_tarjan(gm_graph& G, int root_color, node_t root);
// src/scc_tarjan.cc:151-180: every undecided vertex resolved (G_WCC unused).
void do_tarjan_parallel_color(gm_graph& G, node_t* G_WCC);

// ---- trim1 (src/scc.h:36-48) ----
void initialize_trim1();
// One trim1 sweep over the undecided vertices; returns the number trimmed
// (src/scc_trim1.cc:112-135; the sweep from the fresh state decides on row
// lengths alone, so it can trim fewer than the reference's first sweep).
int do_global_trim1(gm_graph& G);
int do_global_trim1_compact(gm_graph& G);  // the same sweep (src/scc_trim1.cc:137-162)
// Sweeps until one trims <= TRIM_STOP; returns the total (src/scc_trim1.cc:195-209).
int repeat_global_trim1(gm_graph& G, int TRIM_STOP = 100);
int repeat_global_trim1_compact(gm_graph& G, int TRIM_STOP = 100);
// The undecided vertices (colour != -2) in id order, read from the device
// (src/scc_trim1.cc:32 returns the compacted trim targets).
std::vector<node_t>& get_compact_trim_targets();
// Trims to the fixpoint (over every partition) and returns how many of w's
// partition were trimmed; w->count and w->color_set are updated
// (src/scc_trim1.cc:292-303).
int repeat_local_trim1(gm_graph& G, my_work* w);

// ---- trim2 (src/scc.h:50-58) ----
void initialize_trim2();
void finalize_trim2();
// One 2-cycle pass (src/scc_trim2.cc:114-152 and src/scc_trim2_new.cc:80-101 are
// two variants of it; both are served by the pass of scc_trim2_new.cc:3-77).
int do_global_trim2(gm_graph& G);
int repeat_global_trim2(gm_graph& G, int exit_count);  // passes until one removes <= exit_count
int do_global_trim2_new(gm_graph& G);
int repeat_global_trim2_new(gm_graph& G, int exit_count);

// ---- WCC (src/scc.h:60-66) ----
void initialize_WCC();
void finalize_WCC();
void do_global_wcc(gm_graph& G);
// Queues one work item per WCC class (src/scc_weak.cc:198-271), so that
// work_q_size() reports the class count as the reference's driver prints it.
void create_work_items_from_wcc(gm_graph& G);
// Per vertex: its WCC class (a member id) after do_global_wcc, NIL_NODE for
// vertices already done (src/scc_weak.cc:13).  Valid until the next call.
node_t* get_WCC();

// ---- FW-BW (src/scc.h:68-79) ----
// One forward-backward step from one pivot per live partition (the reference
// takes `work`'s partition, src/scc_fb_seq.cc:182-300); queues one new item per
// live partition left afterwards into new_works and returns the SCC vertices
// found.
int do_fw_bw_single_thread(gm_graph& G, my_work* work, std::vector<my_work*>& new_works);
void start_workers_fw_bw(gm_graph& G, int local_items);
void start_workers_fw_bw_dfs(gm_graph& G, int local_items);
void initialize_global_fb();
int do_fw_bw_global_main(gm_graph& G, int curr_color, int count, bool create_work_items);
void create_works_after_bfs_trim(gm_graph& G);

// ---- post analysis (src/scc.h:81-88) ----
void initialize_analyze();
void finalize_analyze();
// src/scc_post_analyze.cc:84-131: "[root members...]" for SCCs of that size
// (of at least that size); roots are the canonical (minimum) members.
void print_scc_of_size(int size);
void print_scc_of_nontrivial_size(int min_size);
void create_histogram_and_print();
void output_scc_list();  // writes scc_list.txt (src/scc_post_analyze.cc:31-53)

// ---- graph building for the update methods (src/scc.h:90-92) ----
// As in src/scc_incremental.cc:7-27 and src/scc_decremental.cc:7-24 these add
// edges to G before it is frozen for a from-scratch run: insert_idea1 the
// original and inserted edges, insert_idea2 the condensation edges (a pair
// summing to 0 is skipped), delete_idea1 every kept edge of every SCC group
// (pairs with -1 skipped; the reference reads only group 5).
void insert_idea1(gm_graph& G, std::vector<std::pair<int, int>> orig_edges,
                  std::vector<std::pair<int, int>> insert_edges);
void insert_idea2(gm_graph& G, std::vector<std::pair<int, int>> scc_edges);
void delete_idea1(gm_graph& G, std::vector<std::vector<std::pair<int, int>>> rem_edges);

// ---- exact device updates (no reference counterpart: the reference rebuilds
// and reruns; these apply the batch to the resident graph and update labels) ----
void insert_edges(gm_graph& G, const std::vector<std::pair<int, int>>& insert_edges);
void delete_edges(gm_graph& G, const std::vector<std::pair<int, int>>& delete_edges);

// ---- whole pipeline and shim controls (no reference counterpart) ----
// Method 2 (src/scc_main.cc:272-351) in one call; returns the SCC count.
long long run_method2(gm_graph& G, int pivot_policy = 0);
void scc_shim_bind(gm_graph& G);    // upload G to the device context now (no phase runs)
void scc_shim_set_mirror(bool on);  // refresh G_Color/G_SCC after each call (default on)
void scc_shim_sync();               // refresh them now
void scc_shim_reset();              // G_Color = G_SCC = -1 on the device (prepare())
void scc_shim_release();            // drop the device context

#endif  // SCCMI_SCC_H

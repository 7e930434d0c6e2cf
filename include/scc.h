/*
 * scc.h — the reference's C++ SCC entry points (src/scc.h:11-93), served by the
 * MI355X engine.  A driver written against the reference keeps its calls; the
 * graph container is sccmi::gm_graph (same public CSR members as the
 * reference's gm_graph, gm_graph/inc/gm_graph.h:154-160), and every phase runs
 * on the GPU through the C ABI in sccmi.h.
 *
 * State: one process-global context bound to the graph the first phase call
 * sees (src/scc.h:11-13 keeps G_Color / G_SCC / G_num_nodes as globals).
 * G_Color and G_SCC are host mirrors refreshed after every phase call (turn
 * that off with scc_shim_set_mirror(false) and call scc_shim_sync() when the
 * arrays are needed).  Errors print a message and exit(EXIT_FAILURE), as the
 * reference does (src/common_main.h:459-473).
 */
#ifndef SCCMI_SCC_H
#define SCCMI_SCC_H

#include <stdint.h>

#include <vector>

#include "sccmi_gm_graph.h"

typedef sccmi::node_t node_t;
typedef sccmi::edge_t edge_t;
using sccmi::gm_graph;

extern int32_t* G_Color;  // src/scc.h:11
extern node_t* G_SCC;     // src/scc.h:12
extern int32_t G_num_nodes;  // src/scc.h:13

// ---- colour (src/scc.h:18-26) ----
void initialize_color();  // G_Color = -1 everywhere
void finalize_color();

// ---- trim1 (src/scc.h:39-48) ----
void initialize_trim1();
int repeat_global_trim1(gm_graph& G, int TRIM_STOP = 100);
int repeat_global_trim1_compact(gm_graph& G, int TRIM_STOP = 100);

// ---- trim2 (src/scc.h:50-58) ----
void initialize_trim2();
void finalize_trim2();
int do_global_trim2_new(gm_graph& G);

// ---- WCC (src/scc.h:60-66) ----
void initialize_WCC();
void finalize_WCC();
void do_global_wcc(gm_graph& G);
void create_work_items_from_wcc(gm_graph& G);

// ---- FW-BW (src/scc.h:68-79) ----
void initialize_global_fb();
void start_workers_fw_bw(gm_graph& G, int local_items);
void start_workers_fw_bw_dfs(gm_graph& G, int local_items);
int do_fw_bw_global_main(gm_graph& G, int curr_color, int count, bool create_work_items);
void create_works_after_bfs_trim(gm_graph& G);

// ---- post analysis (src/scc.h:81-88) ----
void initialize_analyze();
void finalize_analyze();
void create_histogram_and_print();
void output_scc_list();  // writes scc_list.txt (src/scc_post_analyze.cc:31-53)

// ---- dynamic updates (src/scc.h:90-92), exact on the device ----
// insert_idea1/2 and delete_idea1 built a new graph for a from-scratch rerun;
// here they apply the edges to the device graph and update the labels.
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

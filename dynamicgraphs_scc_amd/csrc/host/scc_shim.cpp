// The reference's scc.h entry points (src/scc.h:11-93) over one process-global
// device context.  Each call makes sure the context holds the caller's graph
// (uploading it when the gm_graph changed), runs the phase on the GPU through
// the C ABI, and refreshes the host mirrors G_Color / G_SCC.
#include "scc.h"

#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <vector>

#include "sccmi.h"

int32_t* G_Color = nullptr;
node_t* G_SCC = nullptr;
int32_t G_num_nodes = 0;

namespace {

scc_ctx* g_ctx = nullptr;
const gm_graph* g_bound = nullptr;
uint64_t g_version = 0;
bool g_mirror = true;
std::vector<int32_t> g_color_buf, g_label_buf;

[[noreturn]] void die(const char* what) {
  fprintf(stderr, "scc: %s: %s\n", what, g_ctx ? scc_last_error(g_ctx) : "no context");
  exit(EXIT_FAILURE);
}

void check(int rc, const char* what) {
  if (rc != SCC_OK) die(what);
}

void bind(gm_graph& G) {
  if (!G.is_frozen() && !G.make_reverse_edges()) die("graph freeze");
  if (!g_ctx) {
    if (scc_ctx_create(0, 0, &g_ctx) != SCC_OK) die("no HIP device");
    // skewed graphs are relabelled by degree inside the context; ids stay the caller's
    scc_ctx_set_layout(g_ctx, SCC_LAYOUT_AUTO);
  }
  if (g_bound == &G && g_version == G.version()) return;
  const int32_t n = G.num_nodes();
  check(scc_graph_upload(g_ctx, n, G.num_edges(), G.begin, G.node_idx, G.r_begin, G.r_node_idx),
        "graph upload");
  g_bound = &G;
  g_version = G.version();
  G_num_nodes = n;
  g_color_buf.assign(n, -1);
  g_label_buf.assign(n, -1);
  G_Color = g_color_buf.data();
  G_SCC = g_label_buf.data();
}

void mirror() {
  if (g_mirror) scc_shim_sync();
}

}  // namespace

void scc_shim_sync() {
  if (!g_ctx || !g_bound) return;
  check(scc_colors_download(g_ctx, G_Color), "colour download");
  check(scc_labels_download(g_ctx, G_SCC, 0), "label download");
}
void scc_shim_set_mirror(bool on) { g_mirror = on; }
void scc_shim_bind(gm_graph& G) { bind(G); }
void scc_shim_reset() {
  if (!g_ctx || !g_bound) return;
  check(scc_state_reset(g_ctx), "state reset");
  mirror();
}
void scc_shim_release() {
  if (g_ctx) scc_ctx_destroy(g_ctx);
  g_ctx = nullptr;
  g_bound = nullptr;
}

// src/scc_color.cc:15-31 (initialize_color): every colour -1; labels -1 with it
// (src/scc_main.cc:104-108).
void initialize_color() { scc_shim_reset(); }
void finalize_color() {}
void initialize_trim1() {}
void initialize_trim2() {}
void finalize_trim2() {}
void initialize_WCC() {}
void finalize_WCC() {}
void initialize_global_fb() {}
void initialize_analyze() {}
void finalize_analyze() {}

int repeat_global_trim1(gm_graph& G, int TRIM_STOP) {
  bind(G);
  int64_t t = 0;
  check(scc_trim1(g_ctx, TRIM_STOP, &t), "trim1");
  mirror();
  return (int)t;
}

int repeat_global_trim1_compact(gm_graph& G, int TRIM_STOP) { return repeat_global_trim1(G, TRIM_STOP); }

int do_global_trim2_new(gm_graph& G) {
  bind(G);
  int64_t t = 0;
  check(scc_trim2(g_ctx, &t), "trim2");
  mirror();
  return (int)t;
}

// curr_color and count are tracked by the engine; work items are the
// partitions' colours, so create_work_items needs no separate step.
int do_fw_bw_global_main(gm_graph& G, int curr_color, int count, bool create_work_items) {
  (void)curr_color;
  (void)count;
  (void)create_work_items;
  bind(G);
  int64_t size = 0;
  int32_t pivot = -1;
  check(scc_fb_global(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, -1, &size, &pivot), "global FW-BW");
  mirror();
  return (int)size;
}

void create_works_after_bfs_trim(gm_graph& G) { bind(G); }

void do_global_wcc(gm_graph& G) {
  bind(G);
  int64_t parts = 0;
  check(scc_wcc(g_ctx, &parts), "WCC");
  mirror();
}

void create_work_items_from_wcc(gm_graph& G) { bind(G); }

void start_workers_fw_bw_dfs(gm_graph& G, int local_items) {
  (void)local_items;
  bind(G);
  int64_t rounds = 0;
  check(scc_partitions_resolve(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, SCC_PARTS_BOTH, 1, &rounds),
        "partition FW-BW");
  mirror();
}

void start_workers_fw_bw(gm_graph& G, int local_items) {
  (void)local_items;
  bind(G);
  int64_t rounds = 0;
  check(scc_partitions_resolve(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, SCC_PARTS_FB, 1, &rounds),
        "partition FW-BW");
  mirror();
}

long long run_method2(gm_graph& G, int pivot_policy) {
  bind(G);
  check(scc_state_reset(g_ctx), "state reset");
  scc_run_opts o;
  scc_run_opts_default(&o);
  o.pivot_policy = pivot_policy;
  scc_run_stats st;
  check(scc_run_method2(g_ctx, &o, &st), "method 2");
  mirror();
  return st.num_sccs;
}

void insert_edges(gm_graph& G, const std::vector<std::pair<int, int>>& ins) {
  bind(G);
  std::vector<int32_t> s(ins.size()), d(ins.size());
  for (size_t i = 0; i < ins.size(); i++) {
    s[i] = ins[i].first;
    d[i] = ins[i].second;
  }
  scc_update_stats st;
  check(scc_update(g_ctx, s.data(), d.data(), (int64_t)s.size(), nullptr, nullptr, 0, &st),
        "insert");
  mirror();
}

void delete_edges(gm_graph& G, const std::vector<std::pair<int, int>>& del) {
  bind(G);
  std::vector<int32_t> s(del.size()), d(del.size());
  for (size_t i = 0; i < del.size(); i++) {
    s[i] = del[i].first;
    d[i] = del[i].second;
  }
  scc_update_stats st;
  check(scc_update(g_ctx, nullptr, nullptr, 0, s.data(), d.data(), (int64_t)s.size(), &st),
        "delete");
  mirror();
}

// src/scc_post_analyze.cc:55-81: "size => count" for every SCC size, ascending.
void create_histogram_and_print() {
  if (!g_ctx || !g_bound) return;
  std::vector<int32_t> canon(G_num_nodes);
  check(scc_labels_download(g_ctx, canon.data(), 1), "label download");
  std::vector<int32_t> size(G_num_nodes, 0);
  for (int32_t v = 0; v < G_num_nodes; v++) size[canon[v]]++;
  std::map<int, int> hist;
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (canon[v] == v) hist[size[v]]++;
  for (auto& kv : hist) printf("%d => %d\n", kv.first, kv.second);
  printf("\n");
}

// src/scc_post_analyze.cc:31-53: "vertex dense_id" lines, dense ids in order of
// first appearance.  Any member can label an SCC; the first member met in id
// order is the minimum either way, so the numbering matches the reference's.
void output_scc_list() {
  if (!g_ctx || !g_bound) return;
  FILE* f = fopen("scc_list.txt", "w");
  if (!f) {
    printf("Error: Cannot open scc_list output file");
    exit(1);
  }
  std::vector<int32_t> canon(G_num_nodes), dense(G_num_nodes, -1);
  check(scc_labels_download(g_ctx, canon.data(), 1), "label download");
  int32_t next = 0;
  for (int32_t v = 0; v < G_num_nodes; v++) {
    int32_t& d = dense[canon[v]];
    if (d < 0) d = next++;
    fprintf(f, "%d %d\n", v, d);
  }
  fclose(f);
}

// The reference's scc.h entry points (src/scc.h:11-93) over one process-global
// device context.  Each call makes sure the context holds the caller's graph
// (uploading it when the gm_graph changed), runs the phase on the GPU through
// the C ABI, and refreshes the host mirrors G_Color / G_SCC.
#include "scc.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>

#include <atomic>
#include <climits>
#include <map>
#include <mutex>
#include <unordered_map>
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

// G_Color from the device (the calls below read colours whatever the mirror setting)
void pull_colors() { check(scc_colors_download(g_ctx, G_Color), "colour download"); }

// one work item per live colour, with its vertex count
void items_per_colour(std::vector<my_work*>& out) {
  pull_colors();
  std::unordered_map<int32_t, int> cnt;
  std::vector<int32_t> order;
  for (int32_t v = 0; v < G_num_nodes; v++) {
    const int32_t c = G_Color[v];
    if (c == -2) continue;
    auto it = cnt.find(c);
    if (it == cnt.end()) {
      cnt.emplace(c, 1);
      order.push_back(c);
    } else {
      it->second++;
    }
  }
  for (int32_t c : order) {
    my_work* w = new my_work();
    w->color = c;
    w->count = cnt[c];
    w->color_set = nullptr;
    w->depth = 0;
    out.push_back(w);
  }
}

// the host-side work queue (src/my_work_queue.cc keeps a LIFO vector under a lock)
std::mutex g_qmu;
std::vector<my_work*> g_queue;
int g_qmax_depth = 0;

void drain_queue() {
  std::lock_guard<std::mutex> lk(g_qmu);
  for (my_work* w : g_queue) {
    delete w->color_set;
    delete w;
  }
  g_queue.clear();
}

std::atomic<int> g_color_ctr{-1};
std::vector<node_t> g_targets, g_wcc;

}  // namespace

// ---- gm runtime (gm_graph/inc/gm_runtime.h:46-50) ----
void gm_rt_initialize() {}
int gm_rt_get_num_threads() { return omp_get_max_threads(); }
void gm_rt_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int gm_rt_thread_id() { return omp_get_thread_num(); }

// ---- work queue (src/my_work_queue.h:18-25) ----
bool is_work_q_empty_from_seq_context() {
  std::lock_guard<std::mutex> lk(g_qmu);
  return g_queue.empty();
}
my_work* work_q_fetch(int thread_id) {
  (void)thread_id;
  std::lock_guard<std::mutex> lk(g_qmu);
  if (g_queue.empty()) return nullptr;
  my_work* w = g_queue.back();
  g_queue.pop_back();
  return w;
}
void work_q_fetch_N(int thread_id, int N, std::vector<my_work*>& works) {
  (void)thread_id;
  std::lock_guard<std::mutex> lk(g_qmu);
  for (int i = 0; i < N && !g_queue.empty(); i++) {
    works.push_back(g_queue.back());
    g_queue.pop_back();
  }
}
void work_q_put(int thread_id, my_work* work) {
  (void)thread_id;
  std::lock_guard<std::mutex> lk(g_qmu);
  g_queue.push_back(work);
  if (work && work->depth > g_qmax_depth) g_qmax_depth = work->depth;
}
void work_q_put_all(int thread_id, std::vector<my_work*>& work) {
  for (my_work* w : work) work_q_put(thread_id, w);
}
void work_q_init(int num_threads) {
  (void)num_threads;
  drain_queue();
  g_qmax_depth = 0;
}
void work_q_print_max_depth() { printf("max depth = %d\n", g_qmax_depth); }
int work_q_size() {
  std::lock_guard<std::mutex> lk(g_qmu);
  return (int)g_queue.size();
}

// ---- colours (src/scc_color.cc:38-113) ----
int get_new_color() { return ++g_color_ctr; }
int get_curr_color() { return g_color_ctr.load(); }

node_t choose_pivot_from_color(gm_graph& G, int color) {
  bind(G);
  pull_colors();
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (G_Color[v] == color) return v;
  return gm_graph::NIL_NODE;
}

NODE_SET* generate_compact_set(gm_graph& G, int color) {
  bind(G);
  pull_colors();
  NODE_SET* S = new NODE_SET();
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (G_Color[v] == color) S->insert(v);
  return S;
}

std::vector<node_t>& get_compact_trim_targets() {
  g_targets.clear();
  if (!g_ctx || !g_bound) return g_targets;
  pull_colors();
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (G_Color[v] != -2) g_targets.push_back(v);
  return g_targets;
}

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
void initialize_color() {
  g_color_ctr = -1;
  scc_shim_reset();
}
void finalize_color() {}
void initialize_tarjan() {}
void finalize_tarjan() {}
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

// one sweep: the loop ends after any sweep trimming at most INT64_MAX vertices
int do_global_trim1(gm_graph& G) {
  bind(G);
  int64_t t = 0;
  check(scc_trim1(g_ctx, INT64_MAX, &t), "trim1 sweep");
  mirror();
  return (int)t;
}

int do_global_trim1_compact(gm_graph& G) { return do_global_trim1(G); }

int repeat_local_trim1(gm_graph& G, my_work* w) {
  bind(G);
  if (!w) return 0;
  auto members = [&]() {
    pull_colors();
    int k = 0;
    if (w->color_set) {
      for (node_t v : *w->color_set) k += (v >= 0 && v < G_num_nodes && G_Color[v] == w->color);
    } else {
      for (int32_t v = 0; v < G_num_nodes; v++) k += G_Color[v] == w->color;
    }
    return k;
  };
  const int before = members();
  int64_t t = 0;
  check(scc_trim1(g_ctx, 0, &t), "trim1");
  const int after = members();
  if (w->color_set)
    for (auto it = w->color_set->begin(); it != w->color_set->end();)
      it = (*it >= 0 && *it < G_num_nodes && G_Color[*it] == w->color) ? std::next(it) : w->color_set->erase(it);
  w->count -= before - after;
  mirror();
  return before - after;
}

int do_global_trim2_new(gm_graph& G) {
  bind(G);
  int64_t t = 0;
  check(scc_trim2(g_ctx, &t), "trim2");
  mirror();
  return (int)t;
}

int do_global_trim2(gm_graph& G) { return do_global_trim2_new(G); }

// src/scc_trim2.cc:155-167 (and src/scc_trim2_new.cc:104-116): passes until one
// removes at most exit_count vertices, printing each pass's count
int repeat_global_trim2(gm_graph& G, int exit_count) {
  int total = 0, c = 0;
  do {
    c = do_global_trim2_new(G);
    total += c;
    printf("trim2 = %d\n", c);
  } while (c > exit_count);
  return total;
}

int repeat_global_trim2_new(gm_graph& G, int exit_count) { return repeat_global_trim2(G, exit_count); }

// Every undecided vertex resolved from the current state (trims, forward-backward
// steps, colouring steps), then canonical labels: the SCC partition of Tarjan.
void do_tarjan_all(gm_graph& G) {
  bind(G);
  scc_run_opts o;
  scc_run_opts_default(&o);
  scc_run_stats st;
  check(scc_run_method2(g_ctx, &o, &st), "decomposition");
  mirror();
}

void do_tarjan(gm_graph& G, int root_color, node_t root) {
  (void)root_color;
  (void)root;
  do_tarjan_all(G);
}

void do_tarjan_parallel_color(gm_graph& G, node_t* wcc) {
  (void)wcc;
  do_tarjan_all(G);
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

void create_work_items_from_wcc(gm_graph& G) {
  bind(G);
  std::vector<my_work*> items;
  items_per_colour(items);
  work_q_put_all(0, items);
}

node_t* get_WCC() {
  if (!g_ctx || !g_bound) return nullptr;
  pull_colors();
  g_wcc.assign(G_num_nodes, gm_graph::NIL_NODE);
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (G_Color[v] != -2) g_wcc[v] = G_Color[v];
  return g_wcc.data();
}

int do_fw_bw_single_thread(gm_graph& G, my_work* work, std::vector<my_work*>& new_works) {
  (void)work;
  bind(G);
  int64_t size = 0;
  int32_t pivot = -1;
  check(scc_fb_global(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, -1, &size, &pivot), "FW-BW step");
  items_per_colour(new_works);
  mirror();
  return (int)size;
}

void start_workers_fw_bw_dfs(gm_graph& G, int local_items) {
  (void)local_items;
  bind(G);
  int64_t rounds = 0;
  check(scc_partitions_resolve(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, SCC_PARTS_BOTH, 1, &rounds),
        "partition FW-BW");
  drain_queue();
  mirror();
}

void start_workers_fw_bw(gm_graph& G, int local_items) {
  (void)local_items;
  bind(G);
  int64_t rounds = 0;
  check(scc_partitions_resolve(g_ctx, SCC_PIVOT_DEGREE_PRODUCT, SCC_PARTS_FB, 1, &rounds),
        "partition FW-BW");
  drain_queue();
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

// src/scc_post_analyze.cc:84-131: "[root member ...]" with roots the canonical
// (minimum) members; sizes counted over canonical labels.
namespace {
void print_sccs(int lo, int hi, int wrap, bool newline) {
  if (!g_ctx || !g_bound) return;
  std::vector<int32_t> canon(G_num_nodes), size(G_num_nodes, 0);
  check(scc_labels_download(g_ctx, canon.data(), 1), "label download");
  for (int32_t v = 0; v < G_num_nodes; v++) size[canon[v]]++;
  std::vector<std::vector<int32_t>> members(G_num_nodes);
  for (int32_t v = 0; v < G_num_nodes; v++)
    if (canon[v] != v && size[canon[v]] >= lo && size[canon[v]] <= hi) members[canon[v]].push_back(v);
  for (int32_t i = 0; i < G_num_nodes; i++) {
    if (canon[i] != i || size[i] < lo || size[i] > hi) continue;
    printf("[%d ", i);
    int cnt = 0;
    for (int32_t j : members[i]) {
      printf("%d ", j);
      cnt++;
      if ((cnt % wrap) == wrap - 1) printf("\n");
    }
    printf(newline ? "]\n" : "]");
  }
}
}  // namespace

void print_scc_of_size(int size) { print_sccs(size, size, 16, false); }
void print_scc_of_nontrivial_size(int min_size) { print_sccs(min_size, INT_MAX, 32, true); }

// src/scc_incremental.cc:7-17
void insert_idea1(gm_graph& G, std::vector<std::pair<int, int>> orig_edges,
                  std::vector<std::pair<int, int>> insert_edges) {
  for (auto& e : orig_edges) G.add_edge(e.first, e.second);
  for (auto& e : insert_edges) G.add_edge(e.first, e.second);
}

// src/scc_incremental.cc:19-27
void insert_idea2(gm_graph& G, std::vector<std::pair<int, int>> scc_edges) {
  for (auto& e : scc_edges)
    if (e.first + e.second != 0) G.add_edge(e.first, e.second);
}

// src/scc_decremental.cc:7-24, over every SCC group
void delete_idea1(gm_graph& G, std::vector<std::vector<std::pair<int, int>>> rem_edges) {
  for (auto& grp : rem_edges)
    for (auto& e : grp)
      if (e.first != -1 && e.second != -1) G.add_edge(e.first, e.second);
  printf("num_new_edges:%d\n", (int)G.num_edges());
}

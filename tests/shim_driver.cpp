// A driver written against the reference's scc.h call sequence: the phase
// calls of src/scc_main.cc:96-178 (prepare / run / post_process) for methods
// 0-4 and the graph-building update calls of src/common_main.h:502-725, made
// through include/scc.h.  Tests build it against libsccshim.so (a missing or
// mis-typed entry point fails the build) and run it on the GPU, comparing its
// SCC count and scc_list.txt with the oracle.
//
// usage: shim_driver <edge list> <0-4 | ins | del> [<edge file>]
//   0-4: decompose the 1-indexed edge list with that method's call sequence
//   ins: insert_idea1(G, edge list, edge file), then method 2
//   del: delete_idea1(G, groups of the edges not in the edge file), then method 2
#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <set>
#include <utility>
#include <vector>

#include "scc.h"

static std::vector<std::pair<int, int>> read_pairs(const char* path, int* max_id) {
  std::vector<std::pair<int, int>> e;
  FILE* f = fopen(path, "r");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", path);
    exit(1);
  }
  int a, b;
  while (fscanf(f, "%d %d", &a, &b) == 2) {
    e.emplace_back(a - 1, b - 1);
    if (max_id && a > *max_id) *max_id = a;
    if (max_id && b > *max_id) *max_id = b;
  }
  fclose(f);
  return e;
}

static void prepare(gm_graph& G) {
  gm_rt_initialize();
  G.make_reverse_edges();
  G_num_nodes = G.num_nodes();
  work_q_init(gm_rt_get_num_threads());
  initialize_color();
  initialize_trim1();
  initialize_trim2();
  initialize_tarjan();
  initialize_analyze();
  initialize_global_fb();
  initialize_WCC();
}

static void method0(gm_graph& G) {
  int trimmed = repeat_global_trim1(G);
  if (G_num_nodes - trimmed == 0) return;
  my_work* w = new my_work();
  w->color = get_curr_color();
  assert(w->color == -1);
  w->color_set = NULL;
  w->count = G_num_nodes - trimmed;
  w->depth = 0;
  work_q_put(0, w);
  start_workers_fw_bw(G, 1);
}

static void method1(gm_graph& G) {
  int trimmed = repeat_global_trim1(G);
  int left = G_num_nodes - trimmed;
  if (left == 0) return;
  int scc = do_fw_bw_global_main(G, get_curr_color(), left, false);
  printf("First_SCC_size = %d\n", scc);
  repeat_global_trim1_compact(G);
  create_works_after_bfs_trim(G);
  start_workers_fw_bw_dfs(G, 1);
}

static void method2(gm_graph& G) {
  int trimmed = repeat_global_trim1(G);
  int left = G_num_nodes - trimmed;
  if (left == 0) return;
  int scc = do_fw_bw_global_main(G, get_curr_color(), left, false);
  printf("First_SCC_size = %d\n", scc);
  trimmed = repeat_global_trim1_compact(G);
  trimmed += do_global_trim2_new(G);
  trimmed += repeat_global_trim1_compact(G, 100);
  // the calls a driver may make between the phases
  std::vector<node_t>& targets = get_compact_trim_targets();
  if (!targets.empty()) {
    node_t p = choose_pivot_from_color(G, G_Color[targets[0]]);
    assert(p == targets[0] || G_Color[p] == G_Color[targets[0]]);
    NODE_SET* s = generate_compact_set(G, G_Color[targets[0]]);
    assert(s->count(targets[0]) == 1);
    delete s;
  }
  do_global_wcc(G);
  node_t* wcc = get_WCC();
  assert(wcc != NULL);
  create_work_items_from_wcc(G);
  printf("#queue_size = %d\n", work_q_size());
  start_workers_fw_bw_dfs(G, 40);
  assert(work_q_size() == 0);
}

static void method4(gm_graph& G) {
  int trimmed = repeat_global_trim1(G);
  printf("Trimmed = %d\n", trimmed);
  do_tarjan_all(G);
}

static int post_process(gm_graph& G) {
  int count = 0;
  for (int i = 0; i < G.num_nodes(); i++)
    if (G_SCC[i] == i) count++;
  printf("Total # SCCs = %d\n", count);
  output_scc_list();
  finalize_color();
  finalize_tarjan();
  finalize_trim2();
  finalize_WCC();
  finalize_analyze();
  return count;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <edge list> <0-4 | ins | del> [<edge file>]\n", argv[0]);
    return 2;
  }
  int n = 0;
  std::vector<std::pair<int, int>> edges = read_pairs(argv[1], &n);
  gm_graph G;
  for (int i = 0; i < n; i++) G.add_node();
  const char* mode = argv[2];
  if (!strcmp(mode, "ins") || !strcmp(mode, "del")) {
    if (argc < 4) return 2;
    std::vector<std::pair<int, int>> batch = read_pairs(argv[3], nullptr);
    if (!strcmp(mode, "ins")) {
      insert_idea1(G, edges, batch);
    } else {
      // the kept edges, grouped by source vertex mod 7 (any grouping: every group is taken)
      std::set<std::pair<int, int>> gone(batch.begin(), batch.end());
      std::vector<std::vector<std::pair<int, int>>> groups(7);
      for (auto& e : edges)
        if (!gone.count(e)) groups[e.first % 7].push_back(e);
      groups[3].emplace_back(-1, -1);  // skipped, as the reference skips them
      delete_idea1(G, groups);
    }
    mode = "2";
  } else {
    for (auto& e : edges) G.add_edge(e.first, e.second);
  }
  prepare(G);
  switch (atoi(mode)) {
    case 0: method0(G); break;
    case 1: method1(G); break;
    case 2: method2(G); break;
    case 3: do_tarjan_all(G); break;
    case 4: method4(G); break;
    default: return 2;
  }
  scc_shim_sync();
  post_process(G);
  scc_shim_release();
  return 0;
}

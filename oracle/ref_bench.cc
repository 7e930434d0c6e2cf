// Timing harness for the reference's own method 2 (test infrastructure: the
// CPU baseline of bench.py).  Built by `make -C oracle ref` against the
// reference's phase sources compiled where they lie (src/*.cc except the
// driver src/scc_main.cc, and gm_graph/src/*.cc); nothing here reimplements a
// phase.  It loads the graph once with the reference's gm binary loader
// (gm_graph/src/gm_graph_binary_loader.cc), then runs `reps` decompositions,
// each one the reference driver's prepare() (src/scc_main.cc:96-118, untimed)
// followed by the phase calls of do_baseline_global_wcc_fb()
// (src/scc_main.cc:272-351) timed as src/common_main.h:801-807 times run().
// The globals the phase sources expect are defined here with the values
// src/scc_main.cc:16-43 gives them for method 2.  OpenMP thread i is pinned,
// once, to the i-th CPU of the process's affinity set (src/common_main.h:48-59
// pins it to CPU i, which need not be in the set, once per run of a process).
//
// usage: ref_bench <graph.gm> <threads> <reps> [labels_out]
// prints one line per run: "run <i> ms <t> sccs <count>", and writes the last
// run's G_SCC (int32, native order) to labels_out when given.
#include <omp.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <vector>

#include "gm.h"
#include "my_work_queue.h"
#include "scc.h"

node_t* G_SCC;
int32_t G_num_nodes;
std::vector<int> scc_list, vec_scc_count, level_ver, affect_level, new_edge_nodes;
int met_algo = 2;
int good_init_pivot = 0, yes_yes = 0, flag11 = 0, count_trim_spec = 0;
int max_affec_scc = 0, flag10 = 0, max_scc = -1, min_scc = 0, flag_del = 0;

static void pin_threads() {
  cpu_set_t allowed;
  CPU_ZERO(&allowed);
  if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return;
  std::vector<int> cpus;
  for (int c = 0; c < CPU_SETSIZE; c++)
    if (CPU_ISSET(c, &allowed)) cpus.push_back(c);
  if (cpus.empty()) return;
#pragma omp parallel
  {
    cpu_set_t one;
    CPU_ZERO(&one);
    CPU_SET(cpus[omp_get_thread_num() % cpus.size()], &one);
    pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
  }
}

static void prepare(gm_graph& G) {
  gm_rt_initialize();
  G_num_nodes = G.num_nodes();
  G_SCC = new node_t[G.num_nodes()];
#pragma omp parallel for
  for (int i = 0; i < G_num_nodes; i++) G_SCC[i] = gm_graph::NIL_NODE;
  work_q_init(gm_rt_get_num_threads());
  initialize_color();
  initialize_trim1();
  initialize_trim2();
  initialize_tarjan();
  initialize_analyze();
  initialize_global_fb();
  initialize_WCC();
}

static void method2(gm_graph& G) {
  int trimmed = repeat_global_trim1(G);
  int curr_color = get_curr_color();
  int curr_count = G_num_nodes - trimmed;
  if (curr_count == 0) return;
  int scc_size = do_fw_bw_global_main(G, curr_color, curr_count, false);
  trimmed = repeat_global_trim1_compact(G);
  trimmed += do_global_trim2_new(G);
  trimmed += repeat_global_trim1_compact(G, 100);
  curr_count = curr_count - trimmed - scc_size;
  if (curr_count == 0) return;
  do_global_wcc(G);
  create_work_items_from_wcc(G);
  start_workers_fw_bw_dfs(G, 40);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <graph.gm> <threads> <reps> [labels_out]\n", argv[0]);
    return 2;
  }
  const int threads = atoi(argv[2]), reps = atoi(argv[3]);
  gm_rt_set_num_threads(threads > 0 ? threads : 1);
  gm_graph G;
  if (!G.load_binary(argv[1])) {
    fprintf(stderr, "cannot load %s\n", argv[1]);
    return 1;
  }
  G.make_reverse_edges();
  printf("loaded n %d m %d threads %d\n", (int)G.num_nodes(), (int)G.num_edges(), gm_rt_get_num_threads());
  fflush(stdout);
  pin_threads();  // once: pinning the team again before every run slowed later runs ~100x here
  for (int r = 0; r < reps; r++) {
    prepare(G);
    struct timeval t1, t2;
    gettimeofday(&t1, NULL);
    method2(G);
    gettimeofday(&t2, NULL);
    const double ms = (t2.tv_sec - t1.tv_sec) * 1000.0 + (t2.tv_usec - t1.tv_usec) * 0.001;
    int count = 0;
    for (int i = 0; i < G.num_nodes(); i++) count += G_SCC[i] == i;
    printf("run %d ms %.4f sccs %d\n", r, ms, count);
    fflush(stdout);
    if (r == reps - 1 && argc > 4) {
      FILE* f = fopen(argv[4], "wb");
      if (!f || fwrite(G_SCC, sizeof(node_t), G.num_nodes(), f) != (size_t)G.num_nodes()) return 1;
      fclose(f);
    }
    finalize_color();
    finalize_tarjan();
    finalize_trim2();
    finalize_WCC();
    finalize_analyze();
    delete[] G_SCC;
  }
  return 0;
}

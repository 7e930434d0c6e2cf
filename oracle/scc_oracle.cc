// ORACLE — test infrastructure only.
//
// CPU restatement of the reference's OpenMP SCC path ("method 2": trim1 ->
// global forward-backward -> trim1/trim2/trim1 -> WCC colouring -> per-partition
// forward-backward through a work queue), an iterative Tarjan, and the
// phase-level functions the GPU parity tests compare against.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the checker or the timed CPU baseline, never as the product path.
//
// Every function cites the reference file:line it restates.  Conventions kept
// from the reference (src/scc.h:19, src/scc_color.cc:15-31): colour -1 =
// unassigned, -2 = SCC found, fresh colours from an atomic counter; label
// G_SCC[v] = a member of v's SCC (pivot, self, or the smaller of a 2-cycle).
//
// Parity pinning of each function is described in the header of oracle/__init__.py.
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

namespace {

struct G {
  int32_t n;
  const int32_t *begin, *adj, *rbegin, *radj;
};

constexpr int32_t DONE = -2;   // colour of vertices whose SCC is known
constexpr int32_t NONE = -1;   // initial colour

struct State {
  const G& g;
  int32_t* color;
  int32_t* label;
  std::atomic<int32_t> next_color{-1};
  std::vector<int32_t> targets;  // compacted trim targets (src/scc_trim1.cc:5, :211-275)
  explicit State(const G& gg, int32_t* c, int32_t* l) : g(gg), color(c), label(l) {}
  // src/scc_color.cc:38-54 hands out fresh colours from a shared counter.
  int32_t new_color() { return next_color.fetch_add(1) + 1; }
};

// ---- trim1 ----------------------------------------------------------------
// src/scc_trim1.cc:38-110 (trim_once_node): v of colour c is trimmed when it has
// no out-neighbour k != v of colour c, or (else) no such in-neighbour.
inline bool trim_once(State& s, int32_t v) {
  int32_t c = s.color[v];
  if (c == DONE) return false;
  const G& g = s.g;
  bool live = false;
  for (int32_t i = g.begin[v]; i < g.begin[v + 1]; i++) {
    int32_t k = g.adj[i];
    if (k != v && s.color[k] == c) { live = true; break; }
  }
  if (live) {
    live = false;
    for (int32_t i = g.rbegin[v]; i < g.rbegin[v + 1]; i++) {
      int32_t k = g.radj[i];
      if (k != v && s.color[k] == c) { live = true; break; }
    }
  }
  if (live) return false;
  s.label[v] = v;
  s.color[v] = DONE;
  return true;
}

// src/scc_trim1.cc:112-135 (do_global_trim1, dynamic schedule of 512 over N)
int64_t trim1_all(State& s) {
  int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : cnt)
  for (int32_t v = 0; v < s.g.n; v++) cnt += trim_once(s, v) ? 1 : 0;
  return cnt;
}

// src/scc_trim1.cc:137-162 (do_global_trim1_compact)
int64_t trim1_targets(State& s) {
  int64_t cnt = 0;
  const int64_t T = (int64_t)s.targets.size();
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : cnt)
  for (int64_t i = 0; i < T; i++) cnt += trim_once(s, s.targets[i]) ? 1 : 0;
  return cnt;
}

// src/scc_trim1.cc:211-275 (create_trim1_compact): the target list becomes the
// not-yet-done vertices, drawn from all of 0..N-1 the first time and from the
// previous list afterwards.  Kept in increasing order here.
void compact_targets(State& s) {
  std::vector<int32_t> out;
  if (s.targets.empty()) {
    for (int32_t v = 0; v < s.g.n; v++)
      if (s.color[v] != DONE) out.push_back(v);
  } else {
    for (int32_t v : s.targets)
      if (s.color[v] != DONE) out.push_back(v);
  }
  s.targets.swap(out);
}

// src/scc_trim1.cc:277-290
int64_t repeat_trim1_compact(State& s, int64_t stop) {
  compact_targets(s);
  int64_t total = 0, c;
  do {
    c = trim1_targets(s);
    total += c;
  } while (c > stop);
  return total;
}

// src/scc_trim1.cc:195-209: sweeps over all vertices until a sweep trims at most
// `stop`, switching to the compact list once 10 % of N has been trimmed.
int64_t repeat_trim1(State& s, int64_t stop) {
  int64_t total = 0, c;
  do {
    c = trim1_all(s);
    total += c;
    if ((double)total >= s.g.n * 0.1) return total + repeat_trim1_compact(s, stop);
  } while (c > stop);
  return total;
}

// ---- trim2 ----------------------------------------------------------------
// src/scc_trim2_new.cc:3-39: true when n has exactly one distinct neighbour of
// colour c other than itself (consecutive repeats skipped, as in the reference).
inline bool unique_nbr(const State& s, const int32_t* off, const int32_t* nb, int32_t v,
                       int32_t c, int32_t& out) {
  int32_t the = -1, cnt = 0;
  for (int32_t i = off[v]; i < off[v + 1]; i++) {
    int32_t k = nb[i];
    if (k == v || k == the || s.color[k] != c) continue;
    if (++cnt == 2) return false;
    the = k;
  }
  out = the;
  return cnt == 1;
}

// src/scc_trim2_new.cc:42-77 (trim_2nd_new_main)
inline int trim2_once(State& s, int32_t v) {
  int32_t c = s.color[v];
  if (c == DONE) return 0;
  const G& g = s.g;
  int32_t k, kk;
  if (unique_nbr(s, g.begin, g.adj, v, c, k) && v < k &&
      unique_nbr(s, g.begin, g.adj, k, c, kk) && kk == v) {
    s.color[v] = s.color[k] = DONE;
    s.label[v] = s.label[k] = v;
    return 2;
  }
  if (unique_nbr(s, g.rbegin, g.radj, v, c, k) && v < k &&
      unique_nbr(s, g.rbegin, g.radj, k, c, kk) && kk == v) {
    s.color[v] = s.color[k] = DONE;
    s.label[v] = s.label[k] = v;
    return 2;
  }
  return 0;
}

// src/scc_trim2_new.cc:80-101 (do_global_trim2_new, one pass over the targets)
int64_t trim2_targets(State& s) {
  int64_t cnt = 0;
  const int64_t T = (int64_t)s.targets.size();
#pragma omp parallel for schedule(static) reduction(+ : cnt)
  for (int64_t i = 0; i < T; i++) cnt += trim2_once(s, s.targets[i]);
  return cnt;
}

// ---- global forward-backward ---------------------------------------------
// src/scc_color.cc:64-93 (choose_pivot_from_color): first vertex of colour c in
// target-list order, or in id order when the list is empty.
int32_t choose_pivot(const State& s, int32_t c) {
  if (s.targets.empty()) {
    for (int32_t v = 0; v < s.g.n; v++)
      if (s.color[v] == c) return v;
  } else {
    for (int32_t v : s.targets)
      if (s.color[v] == c) return v;
  }
  return -1;
}

// Level-synchronous parallel reachability standing in for gm_bfs_template
// (gm_graph/inc/gm_bfs_template.h:79-337, thresholds :424, :834-846): top-down
// while the frontier is small, bottom-up (scan the opposite-direction rows for
// an already-reached parent) once it reaches 5 % of N, back to top-down as it
// shrinks.  `claim(k)` tries to admit k (navigator + visit in one atomic step);
// `reached(k)` tells whether k has been admitted.
template <class Claim, class Reached, class Eligible>
void bfs(const State& s, int32_t root, const int32_t* off, const int32_t* nb,
         const int32_t* roff, const int32_t* rnb, Claim claim, Reached reached,
         Eligible eligible) {
  const int32_t n = s.g.n;
  std::vector<int32_t> frontier{root};
  const int nt = omp_get_max_threads();
  std::vector<std::vector<int32_t>> local(nt);
  bool bottom_up = false;
  while (!frontier.empty()) {
    size_t prev = frontier.size();
    if (!bottom_up && (double)frontier.size() >= 0.05 * n) bottom_up = true;
#pragma omp parallel
    {
      auto& L = local[omp_get_thread_num()];
      L.clear();
      if (!bottom_up) {
#pragma omp for schedule(dynamic, 64) nowait
        for (size_t i = 0; i < frontier.size(); i++) {
          int32_t u = frontier[i];
          for (int32_t e = off[u]; e < off[u + 1]; e++) {
            int32_t k = nb[e];
            if (claim(k)) L.push_back(k);
          }
        }
      } else {
        const std::vector<int32_t>& cand = s.targets;
        const int64_t T = cand.empty() ? n : (int64_t)cand.size();
#pragma omp for schedule(dynamic, 512) nowait
        for (int64_t i = 0; i < T; i++) {
          int32_t v = cand.empty() ? (int32_t)i : cand[i];
          if (!eligible(v)) continue;
          for (int32_t e = roff[v]; e < roff[v + 1]; e++) {
            if (reached(rnb[e])) {
              if (claim(v)) L.push_back(v);
              break;
            }
          }
        }
      }
    }
    frontier.clear();
    for (auto& L : local) frontier.insert(frontier.end(), L.begin(), L.end());
    if (bottom_up && frontier.size() <= 2 * prev && (double)frontier.size() < 0.05 * n)
      bottom_up = false;
  }
}

// src/scc_fb_global.cc:133-219 (do_fw_bw_global_main) with the visitors of
// fw_trim_global :32-69 and bw_trim_global :71-129.  Returns the SCC size.
int64_t fw_bw_global(State& s, int32_t base, int64_t count, int32_t pivot_override,
                     int32_t* pivot_out) {
  int32_t pivot = pivot_override >= 0 ? pivot_override : choose_pivot(s, base);
  if (pivot_out) *pivot_out = pivot;
  if (pivot < 0) return 0;
  if (count == 1) {
    s.color[pivot] = DONE;
    s.label[pivot] = pivot;
    return 1;
  }
  const int32_t fw = s.new_color(), bw = s.new_color();
  int32_t* color = s.color;
  const G& g = s.g;
  // forward: navigator colour == base; visit sets fw colour (pivot included)
  color[pivot] = fw;
  bfs(s, pivot, g.begin, g.adj, g.rbegin, g.radj,
      [&](int32_t k) { return __sync_bool_compare_and_swap(&color[k], base, fw); },
      [&](int32_t k) { return color[k] == fw; },
      [&](int32_t k) { return color[k] == base; });
  // backward: navigator colour in {fw, base}; fw -> SCC (label pivot), base -> bw
  std::atomic<int64_t> scc{0};
  int32_t* label = s.label;
  auto claim_bw = [&](int32_t k) {
    int32_t c = color[k];
    if (c == fw) {
      if (__sync_bool_compare_and_swap(&color[k], fw, DONE)) {
        label[k] = pivot;
        scc.fetch_add(1, std::memory_order_relaxed);
        return true;
      }
    } else if (c == base) {
      return __sync_bool_compare_and_swap(&color[k], base, bw);
    }
    return false;
  };
  // the pivot is the first vertex the backward search visits
  claim_bw(pivot);
  // "reached by the backward search": colour bw, or DONE with this pivot's label
  bfs(s, pivot, g.rbegin, g.radj, g.begin, g.adj, claim_bw,
      [&](int32_t k) {
        int32_t c = color[k];
        return c == bw || (c == DONE && label[k] == pivot);
      },
      [&](int32_t k) {
        int32_t c = color[k];
        return c == fw || c == base;
      });
  return scc.load();
}

// ---- WCC colouring ----------------------------------------------------------
// src/scc_weak.cc:57-62 tags high-degree roots so they lose min comparisons.
constexpr int32_t WCC_TAG = 0x20000000, WCC_MASK = 0x1FFFFFFF;

// src/scc_weak.cc:65-125 (propagate_color) after the init of :128-150.
void wcc_fixpoint(State& s, const std::vector<int32_t>& cand, int32_t* W) {
  const G& g = s.g;
  const int64_t T = (int64_t)cand.size();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < T; i++) {
    int32_t v = cand[i];
    if (s.color[v] == DONE) continue;
    W[v] = (g.begin[v + 1] - g.begin[v] >= 50) ? (v | WCC_TAG) : v;
  }
  bool changed;
  do {
    changed = false;
    bool ch1 = false, ch2 = false;
#pragma omp parallel for schedule(dynamic, 32) reduction(|| : ch1)
    for (int64_t i = 0; i < T; i++) {
      int32_t v = cand[i];
      int32_t c = s.color[v];
      if (c == DONE) continue;
      int32_t m = W[v];
      for (int32_t e = g.begin[v]; e < g.begin[v + 1]; e++) {
        int32_t k = g.adj[e];
        if (s.color[k] != c) continue;
        int32_t wk = __atomic_load_n(&W[k], __ATOMIC_RELAXED);
        if (wk < m) { m = wk; ch1 = true; }
      }
      if (m != W[v]) __atomic_store_n(&W[v], m, __ATOMIC_RELAXED);
    }
#pragma omp parallel for schedule(static) reduction(|| : ch2)
    for (int64_t i = 0; i < T; i++) {
      int32_t v = cand[i];
      if (s.color[v] == DONE) continue;
      int32_t r = __atomic_load_n(&W[v], __ATOMIC_RELAXED) & WCC_MASK;
      if (r != v) {
        int32_t wr = __atomic_load_n(&W[r], __ATOMIC_RELAXED);
        if ((wr & WCC_MASK) != r) {
          __atomic_store_n(&W[v], wr, __ATOMIC_RELAXED);
          ch2 = true;
        }
      }
    }
    changed = ch1 || ch2;
  } while (changed);
}

// ---- per-partition forward-backward through a work queue --------------------
struct Work {
  int32_t color;
  std::vector<int32_t> members;
};

// src/scc_fb_seq2.cc:180-289 (do_fw_bw_dfs) with the DFS visitors of :10-177:
// pivot = first member; forward DFS inside the base colour; backward DFS through
// {fw, base}; the intersection is the pivot's SCC; fw-only, bw-only and the rest
// become new work items.
int64_t fw_bw_item(State& s, Work& w, std::vector<Work>& out, std::vector<int32_t>& stack) {
  if (w.members.empty()) return 0;
  const G& g = s.g;
  int32_t* color = s.color;
  const int32_t base = w.color;
  const int32_t pivot = w.members[0];
  if (w.members.size() == 1) {
    color[pivot] = DONE;
    s.label[pivot] = pivot;
    return 1;
  }
  const int32_t fw = s.new_color(), bw = s.new_color();
  stack.clear();
  color[pivot] = fw;
  stack.push_back(pivot);
  while (!stack.empty()) {
    int32_t u = stack.back();
    stack.pop_back();
    for (int32_t e = g.begin[u]; e < g.begin[u + 1]; e++) {
      int32_t k = g.adj[e];
      if (color[k] == base) {
        color[k] = fw;
        stack.push_back(k);
      }
    }
  }
  int64_t scc = 0;
  color[pivot] = DONE;
  s.label[pivot] = pivot;
  scc++;
  stack.push_back(pivot);
  while (!stack.empty()) {
    int32_t u = stack.back();
    stack.pop_back();
    for (int32_t e = g.rbegin[u]; e < g.rbegin[u + 1]; e++) {
      int32_t k = g.radj[e];
      int32_t c = color[k];
      if (c == fw) {
        color[k] = DONE;
        s.label[k] = pivot;
        scc++;
        stack.push_back(k);
      } else if (c == base) {
        color[k] = bw;
        stack.push_back(k);
      }
    }
  }
  Work a{fw, {}}, b{bw, {}}, r{base, {}};
  for (int32_t v : w.members) {
    int32_t c = color[v];
    if (c == fw) a.members.push_back(v);
    else if (c == bw) b.members.push_back(v);
    else if (c == base) r.members.push_back(v);
  }
  if (!a.members.empty()) out.push_back(std::move(a));
  if (!b.members.empty()) out.push_back(std::move(b));
  if (!r.members.empty()) out.push_back(std::move(r));
  return scc;
}

// src/scc_fb_seq2.cc:292-351 with the LIFO queue of src/my_work_queue.cc:8-193:
// every thread fetches up to max(N/2,1) items, processes them, keeps up to N
// new items locally and returns the rest to the shared queue; the run ends when
// the queue is empty and no thread holds work.
void run_work_queue(State& s, std::vector<Work>& init, int local_items) {
  std::mutex mu;
  std::vector<Work> q = std::move(init);
  std::atomic<int> busy{0};
  const int fetch_n = std::max(local_items / 2, 1);
#pragma omp parallel
  {
    std::vector<Work> mine, made;
    std::vector<int32_t> stack;
    while (true) {
      if (mine.empty()) {
        std::unique_lock<std::mutex> lk(mu);
        int take = std::min<int>(fetch_n, (int)q.size());
        for (int i = 0; i < take; i++) {
          mine.push_back(std::move(q.back()));
          q.pop_back();
        }
        if (take > 0) busy.fetch_add(1);
        else if (busy.load() == 0) break;
      }
      if (mine.empty()) {
        std::this_thread::yield();
        continue;
      }
      while (!mine.empty()) {
        Work w = std::move(mine.back());
        mine.pop_back();
        fw_bw_item(s, w, made, stack);
        while ((int)mine.size() < local_items && !made.empty()) {
          mine.push_back(std::move(made.back()));
          made.pop_back();
        }
        if (!made.empty()) {
          std::lock_guard<std::mutex> lk(mu);
          for (auto& x : made) q.push_back(std::move(x));
          made.clear();
        }
      }
      busy.fetch_sub(1);
    }
  }
}

// src/scc_weak.cc:128-196 (do_global_wcc colour assignment) and :198-271
// (create_work_items_from_wcc): each root gets a fresh colour, members copy it,
// and each root's member set becomes one work item.
void wcc_work_items(State& s, std::vector<Work>& items, int32_t* W) {
  const std::vector<int32_t>& cand = s.targets;
  std::vector<int32_t> root_slot(s.g.n, -1);
  for (int32_t v : cand) {
    if (s.color[v] == DONE) continue;
    int32_t r = W[v] & WCC_MASK;
    if (r == v) {
      root_slot[v] = (int32_t)items.size();
      items.push_back(Work{s.new_color(), {}});
    }
  }
  for (int32_t v : cand) {
    if (s.color[v] == DONE) continue;
    int32_t r = W[v] & WCC_MASK;
    Work& w = items[root_slot[r]];
    s.color[v] = w.color;
    w.members.push_back(v);
  }
}

struct Graph_ {
  G g;
};

}  // namespace

extern "C" {

// Full method-2 run (src/scc_main.cc:272-351), timing only the run() window as
// the reference does (src/common_main.h:801-807).  label_out receives G_SCC.
// stats (may be NULL) receives {trim1, scc_first, trim_second, trim2_pairs,
// wcc_items, pivot, sweeps_unused, 0}.
int oracle_method2(int32_t n, const int32_t* begin, const int32_t* adj, const int32_t* rbegin,
                   const int32_t* radj, int32_t* label_out, int32_t* color_out, int nthreads,
                   int64_t trim_stop, double* run_ms, int64_t* stats) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  G g{n, begin, adj, rbegin, radj};
  std::vector<int32_t> color(n, NONE);
  std::vector<int32_t> W(n, -1);
  for (int32_t v = 0; v < n; v++) label_out[v] = -1;
  State s(g, color.data(), label_out);
  int64_t st[8] = {0, 0, 0, 0, 0, -1, 0, 0};
  auto t0 = std::chrono::steady_clock::now();
  do {
    int64_t trimmed = repeat_trim1(s, trim_stop);
    st[0] = trimmed;
    int64_t curr = n - trimmed;
    if (curr == 0) break;
    int32_t piv = -1;
    int64_t scc = fw_bw_global(s, NONE, curr, -1, &piv);
    st[1] = scc;
    st[5] = piv;
    trimmed = repeat_trim1_compact(s, trim_stop);
    int64_t t2 = trim2_targets(s);
    st[3] = t2;
    trimmed += t2 + repeat_trim1_compact(s, 100);
    st[2] = trimmed;
    curr = curr - trimmed - scc;
    if (curr == 0) break;
    wcc_fixpoint(s, s.targets, W.data());
    std::vector<Work> items;
    wcc_work_items(s, items, W.data());
    st[4] = (int64_t)items.size();
    run_work_queue(s, items, 40);
  } while (false);
  auto t1 = std::chrono::steady_clock::now();
  if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (color_out) memcpy(color_out, color.data(), sizeof(int32_t) * n);
  if (stats) memcpy(stats, st, sizeof(st));
  for (int32_t v = 0; v < n; v++)
    if (label_out[v] < 0) return -1;  // every vertex must be labelled
  return 0;
}

// Iterative restatement of src/scc_tarjan.cc:99-149 (scc_tarjan_recurse /
// do_tarjan_all): label of each vertex = the root of its SCC in DFS order.
int oracle_tarjan(int32_t n, const int32_t* begin, const int32_t* adj, int32_t* label_out) {
  std::vector<int32_t> index(n, -1), low(n, 0), stk;
  std::vector<uint8_t> on(n, 0);
  std::vector<std::pair<int32_t, int32_t>> call;  // (vertex, next edge)
  int32_t counter = 0;
  for (int32_t r = 0; r < n; r++) {
    if (index[r] >= 0) continue;
    call.push_back({r, begin[r]});
    index[r] = low[r] = counter++;
    stk.push_back(r);
    on[r] = 1;
    while (!call.empty()) {
      int32_t v = call.back().first;
      int32_t& e = call.back().second;
      if (e < begin[v + 1]) {
        int32_t w = adj[e++];
        if (index[w] < 0) {
          index[w] = low[w] = counter++;
          stk.push_back(w);
          on[w] = 1;
          call.push_back({w, begin[w]});
        } else if (on[w]) {
          low[v] = std::min(low[v], index[w]);
        }
        continue;
      }
      if (low[v] == index[v]) {
        while (true) {
          int32_t w = stk.back();
          stk.pop_back();
          on[w] = 0;
          label_out[w] = v;
          if (w == v) break;
        }
      }
      call.pop_back();
      if (!call.empty()) {
        int32_t p = call.back().first;
        low[p] = std::min(low[p], low[v]);
      }
    }
  }
  return 0;
}

// Min-id canonical labels: out[v] = smallest vertex sharing v's label.
int oracle_canonicalize(int32_t n, const int32_t* label, int32_t* out) {
  std::vector<int32_t> mn(n, INT32_MAX);
  for (int32_t v = 0; v < n; v++) {
    if ((uint32_t)label[v] >= (uint32_t)n) return -1;
    mn[label[v]] = std::min(mn[label[v]], v);
  }
  for (int32_t v = 0; v < n; v++) out[v] = mn[label[v]];
  return 0;
}

// Phase-level: trim1 sweeps over all vertices until a sweep trims nothing
// (TRIM_STOP=0 as the reference GPU path uses, src_CUDA/scc_cuda_main.cpp:421-422).
// color/label are in-out.  Returns the number trimmed.
int64_t oracle_trim1_fixpoint(int32_t n, const int32_t* begin, const int32_t* adj,
                              const int32_t* rbegin, const int32_t* radj, int32_t* color,
                              int32_t* label) {
  G g{n, begin, adj, rbegin, radj};
  State s(g, color, label);
  int64_t total = 0, c;
  do {
    c = trim1_all(s);
    total += c;
  } while (c > 0);
  return total;
}

// Phase-level: one trim2 pass over every vertex in id order (sequential, so
// the outcome is deterministic).  Returns the number of vertices removed.
int64_t oracle_trim2_pass(int32_t n, const int32_t* begin, const int32_t* adj,
                          const int32_t* rbegin, const int32_t* radj, int32_t* color,
                          int32_t* label) {
  G g{n, begin, adj, rbegin, radj};
  State s(g, color, label);
  int64_t cnt = 0;
  for (int32_t v = 0; v < n; v++) cnt += trim2_once(s, v);
  return cnt;
}

// Phase-level: the pairs trim_2nd_new_main (src/scc_trim2_new.cc:42-77)
// accepts when every colour is read as it stands before the pass.  out[v] = 1
// for both members of each such pair.  A pass removes only pairs, so a pair
// that qualifies before the pass still qualifies whenever it is examined (its
// members' other same-colour neighbours can only leave): every trim2 pass, in
// any order or in parallel, removes at least these.  Returns the pair count.
int64_t oracle_trim2_snapshot(int32_t n, const int32_t* begin, const int32_t* adj,
                              const int32_t* rbegin, const int32_t* radj, const int32_t* color,
                              uint8_t* out) {
  G g{n, begin, adj, rbegin, radj};
  State s(g, const_cast<int32_t*>(color), nullptr);
  memset(out, 0, n);
  int64_t pairs = 0;
  for (int32_t v = 0; v < n; v++) {
    const int32_t c = color[v];
    if (c == DONE) continue;
    int32_t k, kk;
    const bool fw = unique_nbr(s, g.begin, g.adj, v, c, k) && v < k && unique_nbr(s, g.begin, g.adj, k, c, kk) &&
                    kk == v;
    if (!fw && !(unique_nbr(s, g.rbegin, g.radj, v, c, k) && v < k &&
                 unique_nbr(s, g.rbegin, g.radj, k, c, kk) && kk == v))
      continue;
    out[v] = out[k] = 1;
    pairs++;
  }
  return pairs;
}

// Phase-level: the WCC fixpoint W over all vertices whose colour is not -2
// (others untouched).
void oracle_wcc(int32_t n, const int32_t* begin, const int32_t* adj, int32_t* color, int32_t* W) {
  G g{n, begin, adj, begin, adj};
  State s(g, color, nullptr);
  std::vector<int32_t> cand;
  for (int32_t v = 0; v < n; v++)
    if (color[v] != DONE) cand.push_back(v);
  wcc_fixpoint(s, cand, W);
}

// Phase-level: vertices reachable from `pivot` along rows (off, nb) through
// vertices of the pivot's colour; out[v] = 1 when reached.  Sequential DFS.
void oracle_reach(int32_t n, const int32_t* off, const int32_t* nb, const int32_t* color,
                  int32_t pivot, uint8_t* out) {
  memset(out, 0, n);
  const int32_t c = color[pivot];
  std::vector<int32_t> st{pivot};
  out[pivot] = 1;
  while (!st.empty()) {
    int32_t u = st.back();
    st.pop_back();
    for (int32_t e = off[u]; e < off[u + 1]; e++) {
      int32_t k = nb[e];
      if (!out[k] && color[k] == c) {
        out[k] = 1;
        st.push_back(k);
      }
    }
  }
}

// Phase-level: the global forward-backward step of src/scc_fb_global.cc:133-219
// from a given pivot over vertices of colour `base` (color/label in-out).
// Returns the SCC size.
int64_t oracle_fb_global(int32_t n, const int32_t* begin, const int32_t* adj,
                         const int32_t* rbegin, const int32_t* radj, int32_t* color,
                         int32_t* label, int32_t base, int32_t pivot) {
  G g{n, begin, adj, rbegin, radj};
  State s(g, color, label);
  int32_t mx = -1;
  for (int32_t v = 0; v < n; v++) mx = std::max(mx, color[v]);
  s.next_color.store(mx);
  int64_t cnt = 0;
  for (int32_t v = 0; v < n; v++) cnt += color[v] == base;
  return fw_bw_global(s, base, cnt, pivot, nullptr);
}

}  // extern "C"

// Request loop over device-resident graphs: continuous batching across worker
// slots, per-graph ordering, admission by HBM budget, latency statistics.
// Host code over the C ABI in sccmi.h; every graph owns one scc_ctx (one HIP
// stream), so requests on different graphs overlap on the device.
#include "sccserve.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sccmi.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t ticket = 0;
  int kind = 0;
  int32_t graph = -1;
  int32_t n = 0;
  int64_t m = 0;
  std::vector<int32_t> a, b, c, d;  // LOAD: CSR arrays; UPDATE: ins src/dst, del src/dst
  int32_t* out = nullptr;           // LABELS target
  int status = SCC_REQ_QUEUED;
  scc_request_info info{};
};

struct Graph {
  int32_t id = -1;
  scc_ctx* ctx = nullptr;
  int32_t n = 0;
  int64_t m = 0, m_max = 0;
  uint64_t reserve = 0;
  bool busy = false, failed = false, closing = false;
  int pending_fault = INT_MIN;
  std::deque<Request*> q;
};

thread_local std::string t_err;

}  // namespace

struct scc_server {
  scc_server_config cfg{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::map<int32_t, Graph> graphs;
  std::deque<Request*> loads;  // LOADs not yet admitted, in arrival order
  int64_t next_ticket = 1;
  int32_t next_graph = 0;
  uint64_t reserved = 0, budget = 0;
  int64_t queued = 0, running = 0, submitted = 0, completed = 0, failed = 0;
  int64_t max_running = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  // update statistics since the last reset
  std::vector<double> upd_lat;
  int64_t upd_edges = 0;
  double upd_first = -1, upd_last = -1;
  std::string err;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  int fail(int rc, const std::string& msg) {
    std::lock_guard<std::mutex> g(errmu);
    err = msg;
    return rc;
  }
  std::mutex errmu;

  uint64_t graph_reserve(int32_t n, int64_t m, int64_t m_max) const {
    return scc_bytes_needed(n, std::max(m, m_max)) +
           scc_update_bytes_needed(n, m, m_max, cfg.max_batch) + (1u << 20);
  }

  // ---- scheduling (called with mu held) ---------------------------------------
  // Extra bytes an update of n_ins inserts needs before it can run (0 when
  // within headroom).
  uint64_t growth_need(const Graph& g, int64_t n_ins, int64_t* new_max) const {
    *new_max = g.m_max;
    if (g.m + n_ins <= g.m_max) return 0;
    const int64_t want = g.m + n_ins;
    *new_max = want + (int64_t)(cfg.growth * (double)want);
    const uint64_t r2 = graph_reserve(g.n, g.m, *new_max);
    return r2 > g.reserve ? r2 - g.reserve : 0;
  }

  Request* pick() {
    Request* best = nullptr;
    for (auto& kv : graphs) {
      Graph& g = kv.second;
      if (g.busy || g.q.empty()) continue;
      Request* r = g.q.front();
      if (r->kind == SCC_REQ_LOAD) {
        if (loads.empty() || loads.front() != r) continue;
        const uint64_t need = graph_reserve(r->n, r->m, r->m + (int64_t)(cfg.growth * (double)r->m));
        if (reserved + need > budget) continue;
      } else if (r->kind == SCC_REQ_UPDATE && !g.failed && g.ctx) {
        int64_t nm;
        const uint64_t extra = growth_need(g, (int64_t)r->a.size(), &nm);
        // an update that could never fit runs and fails; one that fits later waits
        if (extra && reserved + extra > budget && g.reserve + extra <= budget) continue;
      }
      if (!best || r->info.arrival_ms < best->info.arrival_ms ||
          (r->info.arrival_ms == best->info.arrival_ms && r->ticket < best->ticket))
        best = r;
    }
    return best;
  }

  void worker() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      Request* r = nullptr;
      cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
      if (stop) return;
      Graph& g = graphs[r->graph];
      g.busy = true;
      g.q.pop_front();
      std::vector<Request*> group{r};
      std::unique_ptr<Request> merged;
      if (r->kind == SCC_REQ_UPDATE && !g.failed && g.ctx && cfg.coalesce > 1) merged = coalesce(g, group);
      Request* run = merged ? merged.get() : r;
      if (r->kind == SCC_REQ_LOAD) {
        loads.pop_front();
        g.m_max = r->m + (int64_t)(cfg.growth * (double)r->m);
        g.reserve = graph_reserve(r->n, r->m, g.m_max);
        reserved += g.reserve;
      }
      int64_t grow_to = g.m_max;
      uint64_t extra = 0;
      if (r->kind == SCC_REQ_UPDATE && !g.failed && g.ctx) {
        extra = growth_need(g, (int64_t)run->a.size(), &grow_to);
        if (extra && reserved + extra <= budget) {
          reserved += extra;
          g.reserve += extra;
          g.m_max = grow_to;
        } else if (extra) {
          extra = UINT64_MAX;  // cannot fit even alone
        }
      }
      const double t_start = now();
      for (Request* x : group) {
        x->status = SCC_REQ_RUNNING;
        x->info.start_ms = t_start;
        x->info.group = r->kind == SCC_REQ_UPDATE ? (int32_t)group.size() : 0;
        queued--;
      }
      running++;  // one device request, whatever the group size
      max_running = std::max(max_running, running);
      const int fault = g.pending_fault;
      g.pending_fault = INT_MIN;
      lk.unlock();

      const int rc = execute(g, *run, fault, extra == UINT64_MAX);

      lk.lock();
      const double t_end = now();
      running--;
      for (Request* x : group) {
        if (x != run) {
          x->info.num_sccs = run->info.num_sccs;
          x->info.inserted = run->info.inserted;
          x->info.deleted = run->info.deleted;
          x->info.affected = run->info.affected;
        }
        x->info.end_ms = t_end;
        x->info.ms_latency = t_end - x->info.arrival_ms;
        x->info.status = x->status = rc;
        if (rc == SCC_OK) completed++;
        else failed++;
        if (x->kind == SCC_REQ_UPDATE && rc == SCC_OK) {
          upd_lat.push_back(x->info.ms_latency);
          if (upd_first < 0 || x->info.arrival_ms < upd_first) upd_first = x->info.arrival_ms;
        }
      }
      if (r->kind == SCC_REQ_UPDATE && rc == SCC_OK) {
        upd_edges += run->info.inserted + run->info.deleted;  // a group's edges once
        upd_last = std::max(upd_last, t_end);
      }
      if (r->kind == SCC_REQ_LOAD) {
        if (rc == SCC_OK) g.m = r->m;
        else {
          g.failed = true;
          reserved -= g.reserve;
          g.reserve = 0;
        }
      }
      // an update that failed after it began to change the rows leaves the
      // context refusing graph calls (scc_graph_info reports it): the graph is
      // failed, so later requests are refused instead of served stale labels
      if (r->kind == SCC_REQ_UPDATE && rc != SCC_OK &&
          (rc == SCC_ERR_DEVICE_LOST || (g.ctx && scc_graph_info(g.ctx, nullptr, nullptr) != SCC_OK)))
        g.failed = true;
      if (r->kind == SCC_REQ_UNLOAD && rc == SCC_OK) {
        reserved -= g.reserve;
        g.reserve = 0;
      }
      // drop bulk inputs; the requests stay queryable
      for (Request* x : group) {
        std::vector<int32_t>().swap(x->a);
        std::vector<int32_t>().swap(x->b);
        std::vector<int32_t>().swap(x->c);
        std::vector<int32_t>().swap(x->d);
      }
      merged.reset();
      g.busy = false;
      if (r->kind == SCC_REQ_UNLOAD && g.q.empty()) graphs.erase(r->graph);
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  // ---- coalescing (mu held) ------------------------------------------------------
  static bool ids_in_range(const Request& x, int32_t n) {
    for (const std::vector<int32_t>* v : {&x.a, &x.b, &x.c, &x.d})
      for (int32_t u : *v)
        if (u < 0 || u >= n) return false;
    return true;
  }

  // Moves the UPDATEs queued right behind group[0] on g into the group and
  // returns the one update that applies them all, or nullptr when group[0]
  // stays alone.  Per edge key the last member's intent wins; within a member
  // an insert beats a delete (scc_update drops a delete its own batch
  // inserts); inserted self-loops carry no intent (scc_update ignores them).
  // A member that inserts a key an earlier member deletes is left for the
  // next group: one by one, a duplicated edge deleted and inserted again keeps
  // one copy, while one batch holding the insert would keep all copies.
  std::unique_ptr<Request> coalesce(Graph& g, std::vector<Request*>& group) {
    Request* r = group[0];
    if (g.q.empty() || g.q.front()->kind != SCC_REQ_UPDATE || !ids_in_range(*r, g.n)) return nullptr;
    const int64_t n = g.n;
    enum : uint8_t { DEL = 1, INS = 2 };
    std::unordered_map<int64_t, uint8_t> last;  // key -> last intent
    int64_t n_ins = 0, n_del = 0;  // upper bounds of the merged lists
    auto admit = [&](const Request& x, bool check) {
      if (check) {
        if (n_ins + (int64_t)x.a.size() > cfg.max_batch || n_del + (int64_t)x.c.size() > cfg.max_batch)
          return false;
        if (!ids_in_range(x, g.n)) return false;
        for (size_t i = 0; i < x.a.size(); i++) {
          if (x.a[i] == x.b[i]) continue;
          auto it = last.find((int64_t)x.a[i] * n + x.b[i]);
          if (it != last.end() && it->second == DEL) return false;
        }
        int64_t nm;
        const uint64_t extra = growth_need(g, n_ins + (int64_t)x.a.size(), &nm);
        if (extra && reserved + extra > budget) return false;
      }
      for (size_t i = 0; i < x.c.size(); i++) last[(int64_t)x.c[i] * n + x.d[i]] = DEL;
      for (size_t i = 0; i < x.a.size(); i++)
        if (x.a[i] != x.b[i]) last[(int64_t)x.a[i] * n + x.b[i]] = INS;
      n_ins += (int64_t)x.a.size();
      n_del += (int64_t)x.c.size();
      return true;
    };
    admit(*r, false);
    while ((int)group.size() < cfg.coalesce && !g.q.empty() && g.q.front()->kind == SCC_REQ_UPDATE &&
           admit(*g.q.front(), true)) {
      group.push_back(g.q.front());
      g.q.pop_front();
    }
    if (group.size() == 1) return nullptr;
    auto m = std::make_unique<Request>();
    m->kind = SCC_REQ_UPDATE;
    m->graph = r->graph;
    m->ticket = r->ticket;
    for (const auto& kv : last) {
      const int32_t u = (int32_t)(kv.first / n), v = (int32_t)(kv.first % n);
      if (kv.second == INS) {
        m->a.push_back(u);
        m->b.push_back(v);
      } else {
        m->c.push_back(u);
        m->d.push_back(v);
      }
    }
    return m;
  }

  // ---- execution (mu not held; the graph is marked busy) ------------------------
  int execute(Graph& g, Request& r, int fault, bool no_room) {
    if (fault != INT_MIN && g.ctx) scc_ctx_inject_fault(g.ctx, fault);
    switch (r.kind) {
      case SCC_REQ_LOAD: {
        int rc = scc_ctx_create(cfg.device, g.reserve, &g.ctx);
        if (rc) {
          g.ctx = nullptr;
          return fail(rc, "load: no context on the device");
        }
        if (fault != INT_MIN) scc_ctx_inject_fault(g.ctx, fault);
        if ((rc = scc_ctx_set_layout(g.ctx, cfg.layout)) == SCC_OK &&
            (rc = scc_graph_upload(g.ctx, r.n, r.m, r.a.data(), r.b.data(), r.c.data(), r.d.data())) == SCC_OK) {
          scc_run_opts o;
          scc_run_opts_default(&o);
          o.pivot_policy = cfg.pivot_policy;
          scc_run_stats st;
          rc = scc_run_method2(g.ctx, &o, &st);
          if (rc == SCC_OK) r.info.num_sccs = st.num_sccs;
        }
        if (rc) {
          std::string e = std::string("load: ") + scc_last_error(g.ctx);
          scc_ctx_destroy(g.ctx);
          g.ctx = nullptr;
          return fail(rc, e);
        }
        g.n = r.n;
        return SCC_OK;
      }
      case SCC_REQ_UPDATE: {
        if (g.failed || !g.ctx) return fail(SCC_ERR_STATE, "update: the graph failed earlier");
        if (no_room) return fail(SCC_ERR_BUDGET, "update: growing the graph exceeds the server budget");
        int rc = scc_ctx_set_budget(g.ctx, g.reserve);
        scc_update_stats st;
        if (rc == SCC_OK)
          rc = scc_update(g.ctx, r.a.data(), r.b.data(), (int64_t)r.a.size(), r.c.data(), r.d.data(),
                          (int64_t)r.c.size(), &st);
        if (rc) return fail(rc, std::string("update: ") + scc_last_error(g.ctx));
        r.info.num_sccs = st.num_sccs;
        r.info.inserted = st.inserted;
        r.info.deleted = st.deleted;
        r.info.affected = st.affected;
        g.m += st.inserted - st.deleted;
        return SCC_OK;
      }
      case SCC_REQ_LABELS: {
        if (g.failed || !g.ctx) return fail(SCC_ERR_STATE, "labels: the graph failed earlier");
        int rc = scc_labels_download(g.ctx, r.out, 1);
        if (rc) return fail(rc, std::string("labels: ") + scc_last_error(g.ctx));
        return SCC_OK;
      }
      case SCC_REQ_UNLOAD: {
        if (g.ctx) scc_ctx_destroy(g.ctx);
        g.ctx = nullptr;
        return SCC_OK;
      }
    }
    return fail(SCC_ERR_ARG, "unknown request kind");
  }

  // ---- submission -------------------------------------------------------------
  int submit(std::unique_ptr<Request> r, double arrival, int64_t* ticket) {
    std::lock_guard<std::mutex> lk(mu);
    if (stop) return fail(SCC_ERR_STATE, "server is shutting down");
    if (queued >= cfg.queue_limit) return fail(SCC_ERR_LIMIT, "request queue is full");
    auto it = graphs.find(r->graph);
    if (r->kind != SCC_REQ_LOAD && (it == graphs.end() || it->second.closing))
      return fail(SCC_ERR_ARG, "unknown or unloaded graph " + std::to_string(r->graph));
    if (r->kind == SCC_REQ_LOAD) {
      Graph g;
      g.id = r->graph = next_graph++;
      it = graphs.emplace(g.id, std::move(g)).first;
      loads.push_back(r.get());
    }
    if (r->kind == SCC_REQ_UNLOAD) it->second.closing = true;
    r->ticket = next_ticket++;
    r->info.kind = r->kind;
    r->info.graph = r->graph;
    r->info.status = SCC_REQ_QUEUED;
    r->info.arrival_ms = arrival < 0 ? now() : arrival;
    r->info.start_ms = r->info.end_ms = -1;
    *ticket = r->ticket;
    it->second.q.push_back(r.get());
    reqs.emplace(r->ticket, std::move(r));
    queued++;
    submitted++;
    cv_work.notify_one();
    return SCC_OK;
  }
};

extern "C" {

void scc_server_config_default(scc_server_config* c) {
  if (!c) return;
  c->device = 0;
  c->slots = 4;
  c->queue_limit = 1 << 16;
  c->pivot_policy = SCC_PIVOT_DEGREE_PRODUCT;
  c->hbm_budget_bytes = 0;
  c->max_batch = 100000;
  c->growth = 0.25;
  c->layout = SCC_LAYOUT_INPUT;
  c->coalesce = 32;
}

int scc_server_create(const scc_server_config* cfg, scc_server** out) {
  if (!out) return SCC_ERR_ARG;
  *out = nullptr;
  scc_server_config c;
  if (cfg) c = *cfg;
  else scc_server_config_default(&c);
  if (c.slots < 1 || c.slots > 64 || c.queue_limit < 1 || c.max_batch < 1 || c.growth < 0 ||
      c.growth > 16 || c.layout < SCC_LAYOUT_INPUT || c.layout > SCC_LAYOUT_AUTO || c.coalesce > (1 << 20)) {
    t_err = "bad server configuration";
    return SCC_ERR_ARG;
  }
  uint64_t fr = 0, tot = 0;
  if (scc_device_memory(c.device, &fr, &tot) != SCC_OK) {
    t_err = "no HIP device " + std::to_string(c.device);
    return SCC_ERR_HIP;
  }
  auto* s = new scc_server();
  s->cfg = c;
  s->t0 = Clock::now();
  s->budget = c.hbm_budget_bytes ? c.hbm_budget_bytes : (uint64_t)(0.9 * (double)fr);
  for (int i = 0; i < c.slots; i++) s->workers.emplace_back([s] { s->worker(); });
  *out = s;
  return SCC_OK;
}

int scc_server_destroy(scc_server* s) {
  if (!s) return SCC_ERR_ARG;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // fail what has not started, let running requests end
    for (auto& kv : s->graphs) {
      for (Request* r : kv.second.q) {
        r->status = r->info.status = SCC_ERR_CANCELLED;
        s->queued--;
      }
      kv.second.q.clear();
    }
    s->loads.clear();
    s->cv_done.wait(lk, [&] { return s->running == 0; });
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->graphs)
    if (kv.second.ctx) scc_ctx_destroy(kv.second.ctx);
  delete s;
  return SCC_OK;
}

const char* scc_server_last_error(const scc_server* s) {
  if (!s) return t_err.c_str();
  auto* m = const_cast<scc_server*>(s);
  std::lock_guard<std::mutex> g(m->errmu);
  t_err = m->err;
  return t_err.c_str();
}

double scc_server_now_ms(const scc_server* s) { return s ? s->now() : 0.0; }

int scc_server_load(scc_server* s, int32_t n, int64_t m, const int32_t* begin, const int32_t* node_idx,
                    const int32_t* r_begin, const int32_t* r_node_idx, int32_t* graph, int64_t* ticket) {
  if (!s || !graph || !ticket) return SCC_ERR_ARG;
  if (n < 0 || m < 0 || !begin || !r_begin || (m > 0 && (!node_idx || !r_node_idx)))
    return s->fail(SCC_ERR_ARG, "load: bad graph arrays");
  const uint64_t need = s->graph_reserve(n, m, m + (int64_t)(s->cfg.growth * (double)m));
  if (need > s->budget)
    return s->fail(SCC_ERR_BUDGET, "load: the graph needs " + std::to_string(need) +
                                       " B, more than the server budget " + std::to_string(s->budget));
  auto r = std::make_unique<Request>();
  r->kind = SCC_REQ_LOAD;
  r->n = n;
  r->m = m;
  r->a.assign(begin, begin + n + 1);
  r->b.assign(node_idx, node_idx + m);
  r->c.assign(r_begin, r_begin + n + 1);
  r->d.assign(r_node_idx, r_node_idx + m);
  Request* raw = r.get();
  int rc = s->submit(std::move(r), -1, ticket);
  if (rc == SCC_OK) *graph = raw->graph;
  return rc;
}

int scc_server_update(scc_server* s, int32_t graph, const int32_t* ins_src, const int32_t* ins_dst,
                      int64_t n_ins, const int32_t* del_src, const int32_t* del_dst, int64_t n_del,
                      double arrival_ms, int64_t* ticket) {
  if (!s || !ticket) return SCC_ERR_ARG;
  if (n_ins < 0 || n_del < 0 || (n_ins && (!ins_src || !ins_dst)) || (n_del && (!del_src || !del_dst)))
    return s->fail(SCC_ERR_ARG, "update: bad batch");
  if (n_ins > s->cfg.max_batch || n_del > s->cfg.max_batch)
    return s->fail(SCC_ERR_LIMIT, "update: batch larger than max_batch");
  auto r = std::make_unique<Request>();
  r->kind = SCC_REQ_UPDATE;
  r->graph = graph;
  r->a.assign(ins_src, ins_src + n_ins);
  r->b.assign(ins_dst, ins_dst + n_ins);
  r->c.assign(del_src, del_src + n_del);
  r->d.assign(del_dst, del_dst + n_del);
  return s->submit(std::move(r), arrival_ms, ticket);
}

int scc_server_labels(scc_server* s, int32_t graph, int32_t* out, int64_t* ticket) {
  if (!s || !ticket || !out) return SCC_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->kind = SCC_REQ_LABELS;
  r->graph = graph;
  r->out = out;
  return s->submit(std::move(r), -1, ticket);
}

int scc_server_unload(scc_server* s, int32_t graph, int64_t* ticket) {
  if (!s || !ticket) return SCC_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->kind = SCC_REQ_UNLOAD;
  r->graph = graph;
  return s->submit(std::move(r), -1, ticket);
}

int scc_server_wait(scc_server* s, int64_t ticket, double timeout_ms, scc_request_info* info) {
  if (!s) return SCC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return s->fail(SCC_ERR_ARG, "unknown ticket");
  Request* r = it->second.get();
  auto done = [&] { return r->status != SCC_REQ_QUEUED && r->status != SCC_REQ_RUNNING; };
  if (timeout_ms < 0) s->cv_done.wait(lk, done);
  else s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done);
  if (info) {
    *info = r->info;
    info->status = r->status;
  }
  return r->status;
}

int scc_server_drain(scc_server* s, double timeout_ms) {
  if (!s) return SCC_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] { return s->queued == 0 && s->running == 0; };
  if (timeout_ms < 0) s->cv_done.wait(lk, idle);
  else s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), idle);
  return idle() ? (int)SCC_OK : (int)SCC_REQ_QUEUED;
}

int scc_server_cancel(scc_server* s, int64_t ticket) {
  if (!s) return SCC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return s->fail(SCC_ERR_ARG, "unknown ticket");
  Request* r = it->second.get();
  if (r->status != SCC_REQ_QUEUED) return s->fail(SCC_ERR_STATE, "request already started");
  auto gi = s->graphs.find(r->graph);
  if (gi != s->graphs.end()) {
    auto& q = gi->second.q;
    q.erase(std::remove(q.begin(), q.end(), r), q.end());
    if (r->kind == SCC_REQ_UNLOAD) gi->second.closing = false;
    if (r->kind == SCC_REQ_LOAD) {
      // the graph never existed: fail what was queued behind it
      s->loads.erase(std::remove(s->loads.begin(), s->loads.end(), r), s->loads.end());
      for (Request* x : q) {
        x->status = x->info.status = SCC_ERR_CANCELLED;
        s->queued--;
      }
      q.clear();
      s->graphs.erase(gi);
    }
  }
  r->status = r->info.status = SCC_ERR_CANCELLED;
  s->queued--;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
  return SCC_OK;
}

int scc_server_stats_get(scc_server* s, scc_server_stats* o) {
  if (!s || !o) return SCC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *o = scc_server_stats{};
  o->submitted = s->submitted;
  o->completed = s->completed;
  o->failed = s->failed;
  o->queued = s->queued;
  o->running = s->running;
  for (auto& kv : s->graphs)
    if (kv.second.ctx) o->resident_graphs++;
  o->reserved_bytes = s->reserved;
  o->budget_bytes = s->budget;
  o->max_concurrency = (double)s->max_running;
  o->updates_done = (int64_t)s->upd_lat.size();
  o->update_edges = s->upd_edges;
  if (!s->upd_lat.empty()) {
    std::vector<double> v = s->upd_lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      // nearest rank
      size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
      return v[std::min(v.size() - 1, k > 0 ? k - 1 : 0)];
    };
    o->update_p50_ms = pct(50);
    o->update_p99_ms = pct(99);
    o->update_max_ms = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    o->update_mean_ms = sum / (double)v.size();
    o->busy_ms = s->upd_last - s->upd_first;
    o->updates_per_s = o->busy_ms > 0 ? (double)s->upd_edges / (o->busy_ms * 1e-3) : 0.0;
  }
  return SCC_OK;
}

int scc_server_stats_reset(scc_server* s) {
  if (!s) return SCC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->upd_lat.clear();
  s->upd_edges = 0;
  s->upd_first = s->upd_last = -1;
  s->max_running = s->running;
  return SCC_OK;
}

int scc_server_inject_fault(scc_server* s, int32_t graph, int after) {
  if (!s) return SCC_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->graphs.find(graph);
  if (it == s->graphs.end()) return s->fail(SCC_ERR_ARG, "unknown graph");
  it->second.pending_fault = after;
  return SCC_OK;
}

}  // extern "C"

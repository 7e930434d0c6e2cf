// libsccgen: host graph container, CSR/reverse-CSR builder, seeded generators.
//
// The CSR this produces is the layout the reference's gm_graph exposes
// (gm_graph/inc/gm_graph.h:154-160): begin/node_idx for out-edges with rows
// sorted by destination (gm_graph.cc:99-100, do_semi_sort :385-419) and
// r_begin/r_node_idx for in-edges with rows sorted by source
// (make_reverse_edges :193-287, semi-sorted at :276).  The reference builds the
// transpose with an atomic in-degree count, a sequential prefix sum and a
// scatter; here every step (count, scan, scatter, per-row sort) is parallel.
#include "sccgen.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

struct sccgen_graph {
  int32_t n = 0;
  int64_t m = 0;
  std::vector<int32_t> begin, adj, rbegin, radj;
};

namespace {

thread_local std::string g_err;

void set_err(const std::string& s) { g_err = s; }

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Stream key for (seed, stream): distinct streams give unrelated sequences.
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed * 0x9E3779B97F4A7C15ull + mix64(stream + 0x632BE59BD9B4E019ull));
}

inline uint64_t rng_at(uint64_t key, uint64_t i) {
  return mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// Uniform double in [0,1) from 53 random bits.
inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Uniform integer in [0, bound) (bound > 0) by 128-bit multiply-shift.
inline uint64_t bounded(uint64_t r, uint64_t bound) {
  return (uint64_t)(((unsigned __int128)r * bound) >> 64);
}

// Exclusive prefix sum of cnt[0..n) into off[0..n]; returns the total.
int64_t exclusive_scan(const uint32_t* cnt, int64_t* off, int64_t n) {
  int nt = omp_get_max_threads();
  std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    int64_t lo = n * t / nt, hi = n * (t + 1) / nt, s = 0;
    for (int64_t i = lo; i < hi; i++) s += cnt[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < nt; i++) part[i + 1] += part[i];
    int64_t run = part[t];
    for (int64_t i = lo; i < hi; i++) {
      off[i] = run;
      run += cnt[i];
    }
  }
  off[n] = part[nt];
  return part[nt];
}

// Group (key -> value) pairs into CSR rows: rows sorted ascending, optional
// removal of self-loops (key == value) and duplicates.  Output offsets int32.
bool group_rows(int32_t n, int64_t m, const int32_t* key, const int32_t* val,
                bool dedup, std::vector<int32_t>& off_out,
                std::vector<int32_t>& adj_out) {
  std::vector<uint32_t> cnt(n, 0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    if (dedup && key[e] == val[e]) continue;
    __atomic_fetch_add(&cnt[key[e]], 1u, __ATOMIC_RELAXED);
  }
  std::vector<int64_t> off(n + 1);
  int64_t total = exclusive_scan(cnt.data(), off.data(), n);
  if (total >= (int64_t)INT32_MAX) {
    set_err("edge count exceeds int32 offsets");
    return false;
  }
  std::vector<int32_t> tmp(total);
  {
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; e++) {
      if (dedup && key[e] == val[e]) continue;
      int64_t p = __atomic_fetch_add(&cur[key[e]], (int64_t)1, __ATOMIC_RELAXED);
      tmp[p] = val[e];
    }
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; v++) std::sort(tmp.begin() + off[v], tmp.begin() + off[v + 1]);

  off_out.assign(n + 1, 0);
  if (!dedup) {
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v <= n; v++) off_out[v] = (int32_t)off[v];
    adj_out.swap(tmp);
    return true;
  }
  // unique count per row
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; v++) {
    uint32_t u = 0;
    for (int64_t i = off[v]; i < off[v + 1]; i++)
      if (i == off[v] || tmp[i] != tmp[i - 1]) u++;
    cnt[v] = u;
  }
  std::vector<int64_t> off2(n + 1);
  int64_t total2 = exclusive_scan(cnt.data(), off2.data(), n);
  adj_out.assign(total2, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; v++) {
    int64_t w = off2[v];
    for (int64_t i = off[v]; i < off[v + 1]; i++)
      if (i == off[v] || tmp[i] != tmp[i - 1]) adj_out[w++] = tmp[i];
  }
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v <= n; v++) off_out[v] = (int32_t)off2[v];
  return true;
}

sccgen_graph* build_graph(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                          bool dedup) {
  if (n < 0 || m < 0) {
    set_err("negative size");
    return nullptr;
  }
  for (int64_t e = 0; e < m; e++) {
    if ((uint32_t)src[e] >= (uint32_t)n || (uint32_t)dst[e] >= (uint32_t)n) {
      set_err("edge endpoint out of range at index " + std::to_string(e));
      return nullptr;
    }
  }
  sccgen_graph* g = new (std::nothrow) sccgen_graph();
  if (!g) {
    set_err("out of memory");
    return nullptr;
  }
  g->n = n;
  if (!group_rows(n, m, src, dst, dedup, g->begin, g->adj)) {
    delete g;
    return nullptr;
  }
  g->m = (int64_t)g->adj.size();
  // reverse CSR from the (possibly deduplicated) forward rows
  std::vector<int32_t> fs(g->m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; v++)
    for (int32_t i = g->begin[v]; i < g->begin[v + 1]; i++) fs[i] = v;
  if (!group_rows(n, g->m, g->adj.data(), fs.data(), false, g->rbegin, g->radj)) {
    delete g;
    return nullptr;
  }
  return g;
}

// Seeded Fisher-Yates permutation of [0, n).
void seeded_permutation(uint64_t key, int64_t n, std::vector<int32_t>& perm) {
  perm.resize(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  for (int64_t i = n - 1; i > 0; i--) {
    int64_t j = (int64_t)bounded(rng_at(key, (uint64_t)i), (uint64_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
}

enum Streams : uint64_t {
  S_RMAT_EDGE = 1,
  S_PERM = 2,
  S_SIZES = 3,
  S_CYCLE = 4,
  S_INTRA = 5,
  S_RANK = 6,
  S_INTER = 7,
  S_UPD_SIZE = 8,
  S_UPD_INS = 9,
  S_UPD_DEL = 10,
};

}  // namespace

extern "C" {

const char* sccgen_last_error(void) { return g_err.c_str(); }

uint64_t sccgen_rng(uint64_t key, uint64_t i) { return rng_at(key, i); }

sccgen_graph* sccgen_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                                int dedup) {
  return build_graph(n, m, src, dst, dedup != 0);
}

sccgen_graph* sccgen_read_edgelist(const char* path) {
  std::ifstream in(path);
  if (!in) {
    set_err(std::string("cannot open ") + path);
    return nullptr;
  }
  std::vector<int32_t> s, d;
  std::string line;
  int64_t maxid = 0;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    // reference splits on a single space and takes the first two tokens
    size_t sp = line.find(' ');
    if (sp == std::string::npos) {
      set_err("malformed line: " + line);
      return nullptr;
    }
    long a = std::strtol(line.c_str(), nullptr, 10);
    long b = std::strtol(line.c_str() + sp + 1, nullptr, 10);
    if (a < 1 || b < 1 || a > INT32_MAX || b > INT32_MAX) {
      set_err("vertex id out of range: " + line);
      return nullptr;
    }
    s.push_back((int32_t)(a - 1));
    d.push_back((int32_t)(b - 1));
    maxid = std::max<int64_t>(maxid, std::max(a, b));
  }
  return build_graph((int32_t)maxid, (int64_t)s.size(), s.data(), d.data(), false);
}

namespace {
inline char* put_uint(char* p, uint32_t x) {
  char tmp[12];
  int k = 0;
  do {
    tmp[k++] = (char)('0' + x % 10);
    x /= 10;
  } while (x);
  while (k) *p++ = tmp[--k];
  return p;
}
}  // namespace

int sccgen_write_edgelist(const sccgen_graph* g, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_err(std::string("cannot open ") + path);
    return -1;
  }
  // rows are formatted in parallel chunks, written in order
  const int32_t n = g->n;
  const int32_t chunk = 1 << 16;
  const int32_t nchunks = (n + chunk - 1) / chunk;
  int rc = 0;
  for (int32_t c0 = 0; c0 < nchunks && rc == 0; c0 += 64) {
    const int32_t c1 = std::min(nchunks, c0 + 64);
    std::vector<std::string> out(c1 - c0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t c = c0; c < c1; c++) {
      const int32_t v0 = c * chunk, v1 = std::min(n, v0 + chunk);
      std::string& s = out[c - c0];
      s.resize((size_t)(g->begin[v1] - g->begin[v0]) * 24);
      char* p = &s[0];
      for (int32_t v = v0; v < v1; v++)
        for (int32_t i = g->begin[v]; i < g->begin[v + 1]; i++) {
          p = put_uint(p, (uint32_t)v + 1);
          *p++ = ' ';
          p = put_uint(p, (uint32_t)g->adj[i] + 1);
          *p++ = '\n';
        }
      s.resize(p - &s[0]);
    }
    for (auto& s : out)
      if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) rc = -1;
  }
  std::fclose(f);
  if (rc) set_err("write failed");
  return rc;
}

namespace {
constexpr uint32_t GM_MAGIC = 0x03939999u;
inline uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }
}  // namespace

sccgen_graph* sccgen_read_gm_binary(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_err(std::string("cannot open ") + path);
    return nullptr;
  }
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (size < 20) {
    std::fclose(f);
    set_err("gm binary: file too short for its header");
    return nullptr;
  }
  std::vector<uint32_t> w((size_t)size / 4);
  const size_t got = std::fread(w.data(), 4, w.size(), f);
  std::fclose(f);
  if (got != w.size()) {
    set_err("gm binary: read failed");
    return nullptr;
  }
  // header words are big-endian; an old writer left them in host order
  bool flip;
  if (bswap32(w[0]) == GM_MAGIC) flip = true;
  else if (w[0] == GM_MAGIC) flip = false;
  else {
    set_err("gm binary: magic word mismatch");
    return nullptr;
  }
  auto word = [&](size_t i) -> uint32_t { return flip ? bswap32(w[i]) : w[i]; };
  if (word(1) != 4 || word(2) != 4) {
    set_err("gm binary: only 4-byte node and edge ids are supported");
    return nullptr;
  }
  const int64_t n = (int32_t)word(3), m = (int32_t)word(4);
  if (n < 0 || m < 0 || (uint64_t)(5 + (n + 1) + m) * 4 > (uint64_t)size) {
    set_err("gm binary: sizes in the header do not fit the file");
    return nullptr;
  }
  const size_t b0 = 5, a0 = 5 + (size_t)n + 1;
  std::vector<int32_t> src(m), dst(m);
  std::atomic<int> bad{0};
  if ((int32_t)word(b0) != 0 || (int64_t)(int32_t)word(b0 + n) != m) bad = 1;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < n; v++) {
    const int64_t lo = (int32_t)word(b0 + v), hi = (int32_t)word(b0 + v + 1);
    if (lo < 0 || hi < lo || hi > m) {
      bad = 1;
      continue;
    }
    for (int64_t i = lo; i < hi; i++) {
      const int32_t k = (int32_t)word(a0 + i);
      if (k < 0 || k >= n) bad = 1;
      src[i] = (int32_t)v;
      dst[i] = k;
    }
  }
  if (bad) {
    set_err("gm binary: malformed CSR (offsets or ids out of range)");
    return nullptr;
  }
  return build_graph((int32_t)n, m, src.data(), dst.data(), false);
}

int sccgen_write_gm_binary(const sccgen_graph* g, const char* path) {
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_err(std::string("cannot open ") + path);
    return -1;
  }
  const int64_t n = g->n, m = g->begin[g->n];
  std::vector<uint32_t> w(5 + (size_t)n + 1 + (size_t)m);
  w[0] = bswap32(GM_MAGIC);
  w[1] = bswap32(4);
  w[2] = bswap32(4);
  w[3] = bswap32((uint32_t)n);
  w[4] = bswap32((uint32_t)m);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v <= n; v++) w[5 + v] = bswap32((uint32_t)g->begin[v]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) w[6 + n + i] = bswap32((uint32_t)g->adj[i]);
  const bool ok = std::fwrite(w.data(), 4, w.size(), f) == w.size();
  std::fclose(f);
  if (!ok) {
    set_err("write failed");
    return -1;
  }
  return 0;
}

sccgen_graph* sccgen_rmat(int scale, double edge_factor, double a, double b, double c,
                          uint64_t seed, int permute, int dedup, int64_t n_keep) {
  if (scale < 1 || scale > 30 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) {
    set_err("bad R-MAT parameters");
    return nullptr;
  }
  int64_t nfull = (int64_t)1 << scale;
  int64_t n = (n_keep > 0 && n_keep <= nfull) ? n_keep : nfull;
  int64_t m = (int64_t)std::llround(edge_factor * (double)n);
  if (m >= (int64_t)INT32_MAX) {
    set_err("too many edges for int32 offsets");
    return nullptr;
  }
  std::vector<int32_t> perm;
  if (permute) seeded_permutation(stream_key(seed, S_PERM), nfull, perm);
  // thresholds in 32-bit fixed point
  const double ab = a + b, abc = a + b + c;
  const uint32_t ta = (uint32_t)std::min(4294967295.0, a * 4294967296.0);
  const uint32_t tab = (uint32_t)std::min(4294967295.0, ab * 4294967296.0);
  const uint32_t tabc = (uint32_t)std::min(4294967295.0, abc * 4294967296.0);
  const uint64_t key = stream_key(seed, S_RMAT_EDGE);
  std::vector<int32_t> src(m), dst(m);
  std::vector<uint8_t> keep(n_keep > 0 ? m : 0, 1);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    uint64_t u = 0, v = 0;
    for (int lvl = 0; lvl < scale; lvl += 2) {
      uint64_t r = rng_at(key, (uint64_t)e * 16 + (uint64_t)(lvl >> 1));
      for (int h = 0; h < 2 && lvl + h < scale; h++) {
        uint32_t x = (uint32_t)(r >> (32 * h));
        int bu, bv;
        if (x < ta) { bu = 0; bv = 0; }
        else if (x < tab) { bu = 0; bv = 1; }
        else if (x < tabc) { bu = 1; bv = 0; }
        else { bu = 1; bv = 1; }
        u = (u << 1) | (uint64_t)bu;
        v = (v << 1) | (uint64_t)bv;
      }
    }
    int64_t pu = permute ? perm[u] : (int64_t)u, pv = permute ? perm[v] : (int64_t)v;
    if (n_keep > 0 && (pu >= n || pv >= n)) {
      keep[e] = 0;
      pu = pv = 0;
    }
    src[e] = (int32_t)pu;
    dst[e] = (int32_t)pv;
  }
  if (n_keep > 0) {
    int64_t w = 0;
    for (int64_t e = 0; e < m; e++)
      if (keep[e]) {
        src[w] = src[e];
        dst[w] = dst[e];
        w++;
      }
    src.resize(w);
    dst.resize(w);
    m = w;
  }
  return build_graph((int32_t)n, m, src.data(), dst.data(), dedup != 0);
}

sccgen_graph* sccgen_planted(int32_t n, double giant_frac, int32_t n_medium, int32_t size_min,
                             int32_t size_max, double alpha, int32_t intra_out,
                             double inter_per_vertex, double pair_frac, uint64_t seed,
                             int32_t* comp_out) {
  if (n < 1 || giant_frac < 0 || giant_frac > 1 || size_min < 1 || size_max < size_min ||
      alpha <= 1.0 || intra_out < 0 || inter_per_vertex < 0 || pair_frac < 0 || pair_frac > 1) {
    set_err("bad planted-SCC parameters");
    return nullptr;
  }
  // ---- component sizes (logical, before the id permutation)
  std::vector<int64_t> comp_start;  // start of each component in logical ids
  std::vector<int32_t> comp_size;
  int64_t giant = (int64_t)std::llround(giant_frac * n);
  if (giant > 0) {
    comp_start.push_back(0);
    comp_size.push_back((int32_t)giant);
  }
  int64_t used = giant;
  {
    const uint64_t key = stream_key(seed, S_SIZES);
    const double e = 1.0 - alpha;
    const double lo = std::pow((double)size_min, e), hi = std::pow((double)size_max + 1.0, e);
    for (int32_t i = 0; i < n_medium; i++) {
      double uu = u01(rng_at(key, (uint64_t)i));
      int64_t s = (int64_t)std::floor(std::pow(lo + uu * (hi - lo), 1.0 / e));
      s = std::max<int64_t>(size_min, std::min<int64_t>(size_max, s));
      if (used + s > n) break;
      comp_start.push_back(used);
      comp_size.push_back((int32_t)s);
      used += s;
    }
  }
  int64_t rest = n - used;
  int64_t pairs = (int64_t)std::floor(rest * pair_frac / 2.0);
  for (int64_t i = 0; i < pairs; i++) {
    comp_start.push_back(used);
    comp_size.push_back(2);
    used += 2;
  }
  while (used < n) {
    comp_start.push_back(used);
    comp_size.push_back(1);
    used += 1;
  }
  const int64_t ncomp = (int64_t)comp_size.size();
  std::vector<int32_t> comp_of(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t c = 0; c < ncomp; c++)
    for (int64_t i = 0; i < comp_size[c]; i++) comp_of[comp_start[c] + i] = (int32_t)c;

  // ---- edges (logical ids); count first for exact allocation
  std::vector<int64_t> ebase(ncomp + 1, 0);
  for (int64_t c = 0; c < ncomp; c++) {
    int64_t s = comp_size[c];
    int64_t k = (s >= 3) ? s * (1 + (int64_t)intra_out) : (s == 2 ? 2 : 0);
    ebase[c + 1] = ebase[c] + k;
  }
  const int64_t m_intra = ebase[ncomp];
  const int64_t draws_per_vertex_floor = (int64_t)std::floor(inter_per_vertex);
  const double draws_frac = inter_per_vertex - (double)draws_per_vertex_floor;
  const int64_t max_draws = (int64_t)n * (draws_per_vertex_floor + 1);
  if (m_intra + max_draws >= (int64_t)INT32_MAX) {
    set_err("too many edges for int32 offsets");
    return nullptr;
  }
  std::vector<int32_t> src(m_intra + max_draws), dst(m_intra + max_draws);
  std::vector<uint8_t> keep(max_draws, 0);
  const uint64_t kcyc = stream_key(seed, S_CYCLE), kin = stream_key(seed, S_INTRA);
#pragma omp parallel
  {
    std::vector<int32_t> order;
#pragma omp for schedule(dynamic, 16)
    for (int64_t c = 0; c < ncomp; c++) {
      int64_t s = comp_size[c], st = comp_start[c], w = ebase[c];
      if (s == 2) {
        src[w] = (int32_t)st; dst[w] = (int32_t)(st + 1);
        src[w + 1] = (int32_t)(st + 1); dst[w + 1] = (int32_t)st;
        continue;
      }
      if (s < 3) continue;
      // random Hamiltonian cycle through the component's members
      order.resize(s);
      for (int64_t i = 0; i < s; i++) order[i] = (int32_t)(st + i);
      uint64_t kc = kcyc ^ mix64((uint64_t)c);
      for (int64_t i = s - 1; i > 0; i--) {
        int64_t j = (int64_t)bounded(rng_at(kc, (uint64_t)i), (uint64_t)(i + 1));
        std::swap(order[i], order[j]);
      }
      for (int64_t i = 0; i < s; i++) {
        src[w] = order[i];
        dst[w] = order[(i + 1) % s];
        w++;
      }
      // intra_out random intra-component out-edges per vertex
      for (int64_t i = 0; i < s; i++) {
        for (int32_t k = 0; k < intra_out; k++) {
          uint64_t r = rng_at(kin, (uint64_t)(st + i) * 64 + (uint64_t)k);
          src[w] = (int32_t)(st + i);
          dst[w] = (int32_t)(st + (int64_t)bounded(r, (uint64_t)s));
          w++;
        }
      }
    }
  }
  // topological rank of components: a uniform random permutation
  std::vector<int32_t> rank;
  seeded_permutation(stream_key(seed, S_RANK), ncomp, rank);
  const uint64_t kint = stream_key(seed, S_INTER);
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < n; u++) {
    for (int64_t k = 0; k <= draws_per_vertex_floor; k++) {
      uint64_t r0 = rng_at(kint, (uint64_t)u * 64 + (uint64_t)k * 2);
      if (k == draws_per_vertex_floor && u01(r0) >= draws_frac) continue;
      uint64_t r1 = rng_at(kint, (uint64_t)u * 64 + (uint64_t)k * 2 + 1);
      int64_t v = (int64_t)bounded(r1, (uint64_t)n);
      int32_t cu = comp_of[u], cv = comp_of[v];
      int64_t slot = u * (draws_per_vertex_floor + 1) + k;
      if (cu == cv) continue;
      bool fwd = rank[cu] < rank[cv];
      src[m_intra + slot] = (int32_t)(fwd ? u : v);
      dst[m_intra + slot] = (int32_t)(fwd ? v : u);
      keep[slot] = 1;
    }
  }
  // compact dropped inter draws
  int64_t w = m_intra;
  for (int64_t s = 0; s < max_draws; s++)
    if (keep[s]) {
      src[w] = src[m_intra + s];
      dst[w] = dst[m_intra + s];
      w++;
    }
  src.resize(w);
  dst.resize(w);
  // relabel by a seeded permutation
  std::vector<int32_t> perm;
  seeded_permutation(stream_key(seed, S_PERM), n, perm);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < w; e++) {
    src[e] = perm[src[e]];
    dst[e] = perm[dst[e]];
  }
  if (comp_out) {
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; v++) comp_out[perm[v]] = comp_of[v];
  }
  return build_graph(n, w, src.data(), dst.data(), true);
}

int32_t sccgen_num_nodes(const sccgen_graph* g) { return g ? g->n : -1; }
int64_t sccgen_num_edges(const sccgen_graph* g) { return g ? g->m : -1; }

void sccgen_copy(const sccgen_graph* g, int32_t* begin, int32_t* node_idx, int32_t* r_begin,
                 int32_t* r_node_idx) {
  if (!g) return;
  if (begin) std::memcpy(begin, g->begin.data(), sizeof(int32_t) * (g->n + 1));
  if (node_idx && g->m) std::memcpy(node_idx, g->adj.data(), sizeof(int32_t) * g->m);
  if (r_begin) std::memcpy(r_begin, g->rbegin.data(), sizeof(int32_t) * (g->n + 1));
  if (r_node_idx && g->m) std::memcpy(r_node_idx, g->radj.data(), sizeof(int32_t) * g->m);
}

const int32_t* sccgen_array(const sccgen_graph* g, int which) {
  if (!g) return nullptr;
  switch (which) {
    case 0: return g->begin.data();
    case 1: return g->adj.data();
    case 2: return g->rbegin.data();
    case 3: return g->radj.data();
  }
  return nullptr;
}

void sccgen_free(sccgen_graph* g) { delete g; }

int sccgen_canonicalize(int32_t n, const int32_t* labels, int32_t* out) {
  std::vector<int32_t> mn(n, INT32_MAX);
  for (int32_t v = 0; v < n; v++) {
    int32_t l = labels[v];
    if ((uint32_t)l >= (uint32_t)n) {
      set_err("label out of range at vertex " + std::to_string(v));
      return -1;
    }
    if (v < mn[l]) mn[l] = v;
  }
  for (int32_t v = 0; v < n; v++) out[v] = mn[labels[v]];
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Dynamic update stream.  The tracked edge set is the base CSR (with a
// deleted flag per base edge) plus an append-only list of inserted edges
// (with their own deleted flags) and a hash set for membership.  Deletes are
// drawn uniformly over present edges by rejection on [0, M_base + n_inserted).
struct sccgen_updates {
  int32_t n = 0;
  int64_t m_base = 0;
  std::vector<int32_t> begin, adj;  // base CSR (sorted rows)
  std::vector<int32_t> base_src;    // source of each base edge
  std::vector<uint8_t> base_dead;
  std::vector<int32_t> ins_s, ins_d;
  std::vector<uint8_t> ins_dead;
  std::unordered_map<uint64_t, int64_t> ins_live;  // key (u<<32|v) -> slot of live inserted edge
  int64_t live = 0;
  uint64_t seed = 0;
  int64_t lo = 1000, hi = 100000;
  int64_t batch = 0;

  bool base_has(int32_t u, int32_t v, int64_t* idx) const {
    auto b = adj.begin() + begin[u], e = adj.begin() + begin[u + 1];
    auto it = std::lower_bound(b, e, v);
    if (it != e && *it == v) {
      if (idx) *idx = it - adj.begin();
      return true;
    }
    return false;
  }
  bool present(int32_t u, int32_t v) const {
    uint64_t k = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
    int64_t idx;
    if (base_has(u, v, &idx)) return !base_dead[idx];
    return ins_live.count(k) != 0;
  }
  int64_t next_size() const {
    uint64_t r = rng_at(stream_key(seed, S_UPD_SIZE), (uint64_t)batch);
    double x = std::log((double)lo) + u01(r) * (std::log((double)hi) - std::log((double)lo));
    int64_t s = (int64_t)std::floor(std::exp(x));
    return std::max<int64_t>(1, std::min<int64_t>(hi, s));
  }
};

extern "C" {

sccgen_updates* sccgen_updates_create(const sccgen_graph* base, uint64_t seed, int64_t size_lo,
                                      int64_t size_hi) {
  if (!base || size_lo < 1 || size_hi < size_lo) {
    set_err("bad update-stream parameters");
    return nullptr;
  }
  auto* u = new sccgen_updates();
  u->n = base->n;
  u->m_base = base->m;
  u->begin = base->begin;
  u->adj = base->adj;
  u->base_src.resize(base->m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < base->n; v++)
    for (int32_t i = base->begin[v]; i < base->begin[v + 1]; i++) u->base_src[i] = v;
  u->base_dead.assign(base->m, 0);
  u->live = base->m;
  u->seed = seed;
  u->lo = size_lo;
  u->hi = size_hi;
  return u;
}

int64_t sccgen_updates_peek(const sccgen_updates* u) { return u ? u->next_size() : -1; }

int sccgen_updates_set_sizes(sccgen_updates* u, int64_t size_lo, int64_t size_hi) {
  if (!u || size_lo < 1 || size_hi < size_lo) {
    set_err("bad batch sizes");
    return -1;
  }
  u->lo = size_lo;
  u->hi = size_hi;
  return 0;
}

int sccgen_updates_next(sccgen_updates* u, int64_t cap, int32_t* ins_src, int32_t* ins_dst,
                        int64_t* n_ins, int32_t* del_src, int32_t* del_dst, int64_t* n_del) {
  if (!u) return -1;
  int64_t s = u->next_size();
  int64_t ni = s / 2, nd = s - ni;
  if (ni > cap || nd > cap) {
    set_err("batch capacity too small");
    return -2;
  }
  const uint64_t kins = stream_key(u->seed ^ (uint64_t)u->batch * 0x2545F4914F6CDD1Dull, S_UPD_INS);
  const uint64_t kdel = stream_key(u->seed ^ (uint64_t)u->batch * 0x2545F4914F6CDD1Dull, S_UPD_DEL);
  uint64_t ctr = 0;
  // deletes first: uniform over the edges present before this batch
  int64_t got_d = 0;
  std::unordered_set<uint64_t> chosen;
  while (got_d < nd && u->live > 0) {
    uint64_t r = rng_at(kdel, ctr++);
    uint64_t pool = (uint64_t)(u->m_base + (int64_t)u->ins_s.size());
    uint64_t i = bounded(r, pool);
    int32_t a, b;
    if ((int64_t)i < u->m_base) {
      if (u->base_dead[i]) continue;
      a = u->base_src[i];
      b = u->adj[i];
    } else {
      int64_t j = (int64_t)i - u->m_base;
      if (u->ins_dead[j]) continue;
      a = u->ins_s[j];
      b = u->ins_d[j];
    }
    uint64_t k = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    if (!chosen.insert(k).second) continue;
    del_src[got_d] = a;
    del_dst[got_d] = b;
    got_d++;
  }
  // apply deletes to the tracked set
  for (int64_t i = 0; i < got_d; i++) {
    int32_t a = del_src[i], b = del_dst[i];
    int64_t idx;
    if (u->base_has(a, b, &idx) && !u->base_dead[idx]) {
      u->base_dead[idx] = 1;
    } else {
      uint64_t k = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
      auto it = u->ins_live.find(k);
      if (it != u->ins_live.end()) {
        u->ins_dead[it->second] = 1;
        u->ins_live.erase(it);
      }
    }
    u->live--;
  }
  // inserts: uniform random pairs u != v that are not present
  int64_t got_i = 0;
  ctr = 0;
  while (got_i < ni) {
    uint64_t r = rng_at(kins, ctr++);
    int32_t a = (int32_t)bounded(r, (uint64_t)u->n);
    int32_t b = (int32_t)bounded(rng_at(kins ^ 0x5bd1e995ull, ctr), (uint64_t)u->n);
    if (a == b || u->present(a, b)) continue;
    uint64_t k = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    int64_t idx;
    if (u->base_has(a, b, &idx)) {
      u->base_dead[idx] = 0;  // re-insert of a deleted base edge
    } else {
      u->ins_s.push_back(a);
      u->ins_d.push_back(b);
      u->ins_dead.push_back(0);
      u->ins_live.emplace(k, (int64_t)u->ins_s.size() - 1);
    }
    u->live++;
    ins_src[got_i] = a;
    ins_dst[got_i] = b;
    got_i++;
  }
  *n_ins = got_i;
  *n_del = got_d;
  u->batch++;
  return 0;
}

int64_t sccgen_updates_num_edges(const sccgen_updates* u) { return u ? u->live : -1; }

sccgen_graph* sccgen_updates_graph(const sccgen_updates* u) {
  if (!u) return nullptr;
  // live edges, compacted in parallel chunks (order kept)
  const int64_t mb = u->m_base, mi = (int64_t)u->ins_s.size(), tot = mb + mi;
  const int64_t chunk = 1 << 20, nch = (tot + chunk - 1) / chunk;
  std::vector<int64_t> cnt(nch + 1, 0);
  auto dead = [&](int64_t i) { return i < mb ? u->base_dead[i] != 0 : u->ins_dead[i - mb] != 0; };
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nch; c++) {
    int64_t k = 0;
    for (int64_t i = c * chunk; i < std::min(tot, (c + 1) * chunk); i++) k += !dead(i);
    cnt[c + 1] = k;
  }
  for (int64_t c = 0; c < nch; c++) cnt[c + 1] += cnt[c];
  std::vector<int32_t> s(cnt[nch]), d(cnt[nch]);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t c = 0; c < nch; c++) {
    int64_t o = cnt[c];
    for (int64_t i = c * chunk; i < std::min(tot, (c + 1) * chunk); i++) {
      if (dead(i)) continue;
      s[o] = i < mb ? u->base_src[i] : u->ins_s[i - mb];
      d[o] = i < mb ? u->adj[i] : u->ins_d[i - mb];
      o++;
    }
  }
  return build_graph(u->n, (int64_t)s.size(), s.data(), d.data(), true);
}

void sccgen_updates_free(sccgen_updates* u) { delete u; }

}  // extern "C"

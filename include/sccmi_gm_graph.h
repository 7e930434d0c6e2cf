/*
 * sccmi_gm_graph.h — a graph container with the public interface of the
 * reference's gm_graph that SCC code touches (gm_graph/inc/gm_graph.h:139-484):
 * a flexible edge list built with add_node / add_edge, frozen into a CSR with
 * rows sorted by destination (freeze, gm_graph.cc:60-101, semi-sorted :99-100)
 * and a reverse CSR (make_reverse_edges, gm_graph.cc:193-287).  The arrays
 * begin / node_idx / r_begin / r_node_idx are public as in the reference.
 * Self-loops and duplicate edges are kept, as the reference keeps them.
 * The CSR is built by libsccgen (parallel count / scan / scatter / row sort).
 */
#ifndef SCCMI_GM_GRAPH_H
#define SCCMI_GM_GRAPH_H

#include <stdint.h>

#include <utility>
#include <vector>

#include "sccgen.h"

namespace sccmi {

typedef int32_t node_t;  // gm_graph/inc/gm_graph_typedef.h:17-18
typedef int32_t edge_t;

class gm_graph {
 public:
  static constexpr node_t NIL_NODE = -1;
  static constexpr edge_t NIL_EDGE = -1;

  edge_t* begin = nullptr;
  node_t* node_idx = nullptr;
  edge_t* r_begin = nullptr;
  node_t* r_node_idx = nullptr;

  gm_graph() = default;
  gm_graph(const gm_graph&) = delete;
  gm_graph& operator=(const gm_graph&) = delete;
  ~gm_graph() { clear(); }

  node_t num_nodes() const { return frozen_ ? sccgen_num_nodes(h_) : n_; }
  edge_t num_edges() const { return frozen_ ? (edge_t)sccgen_num_edges(h_) : (edge_t)src_.size(); }
  bool is_frozen() const { return frozen_; }
  bool has_reverse_edge() const { return frozen_; }

  node_t add_node() {
    thaw();
    return n_++;
  }
  // returns the edge's index in insertion order
  edge_t add_edge(node_t n, node_t m) {
    thaw();
    src_.push_back(n);
    dst_.push_back(m);
    return (edge_t)src_.size() - 1;
  }
  // CSR + reverse CSR with sorted rows; returns false on bad ids
  bool freeze() {
    if (frozen_) return true;
    sccgen_graph* h = sccgen_from_edges(n_, (int64_t)src_.size(), src_.data(), dst_.data(), 0);
    if (!h) return false;
    h_ = h;
    begin = const_cast<edge_t*>(sccgen_array(h_, 0));
    node_idx = const_cast<node_t*>(sccgen_array(h_, 1));
    r_begin = const_cast<edge_t*>(sccgen_array(h_, 2));
    r_node_idx = const_cast<node_t*>(sccgen_array(h_, 3));
    frozen_ = true;
    ++version_;
    return true;
  }
  bool make_reverse_edges() { return freeze(); }
  // adopt a built graph (takes ownership)
  void adopt(sccgen_graph* h) {
    clear();
    h_ = h;
    n_ = sccgen_num_nodes(h);
    begin = const_cast<edge_t*>(sccgen_array(h_, 0));
    node_idx = const_cast<node_t*>(sccgen_array(h_, 1));
    r_begin = const_cast<edge_t*>(sccgen_array(h_, 2));
    r_node_idx = const_cast<node_t*>(sccgen_array(h_, 3));
    frozen_ = true;
    ++version_;
  }
  // gm binary CSR files (gm_graph/src/gm_graph_binary_loader.cc:48-248)
  bool load_binary(const char* filename) {
    sccgen_graph* h = sccgen_read_gm_binary(filename);
    if (!h) return false;
    adopt(h);
    return true;
  }
  bool store_binary(const char* filename) {
    if (!freeze()) return false;
    return sccgen_write_gm_binary(h_, filename) == 0;
  }
  void clear() {
    if (h_) sccgen_free(h_);
    h_ = nullptr;
    begin = r_begin = nullptr;
    node_idx = r_node_idx = nullptr;
    frozen_ = false;
  }
  // bumps whenever the CSR changes (the engine re-uploads on change)
  uint64_t version() const { return version_; }

 private:
  // back to the flexible form: keep the frozen edges as the edge list
  void thaw() {
    if (!frozen_) return;
    const int32_t n = sccgen_num_nodes(h_);
    src_.clear();
    dst_.clear();
    for (int32_t v = 0; v < n; v++)
      for (int32_t i = begin[v]; i < begin[v + 1]; i++) {
        src_.push_back(v);
        dst_.push_back(node_idx[i]);
      }
    n_ = n;
    clear();
  }

  sccgen_graph* h_ = nullptr;
  bool frozen_ = false;
  node_t n_ = 0;
  std::vector<node_t> src_, dst_;
  uint64_t version_ = 0;
};

}  // namespace sccmi

// The gm runtime calls the reference's driver makes (gm_graph/inc/gm_runtime.h:46-50):
// OpenMP thread count and thread id of the host-side work.  Defined in libsccshim.so.
void gm_rt_initialize();
int gm_rt_get_num_threads();
void gm_rt_set_num_threads(int n);
int gm_rt_thread_id();

#endif  // SCCMI_GM_GRAPH_H

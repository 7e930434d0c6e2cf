"""Host graph container and seeded synthetic workloads (ctypes over libsccgen.so).

``Graph`` holds the int32 CSR + reverse CSR the reference's ``gm_graph`` exposes
(gm_graph/inc/gm_graph.h:154-160): ``begin``/``node_idx`` out-rows sorted by
destination and ``r_begin``/``r_node_idx`` in-rows sorted by source
(gm_graph/src/gm_graph.cc:99-100, :276).

Workloads (all seeded, counter-based, independent of thread count):

* ``rmat(scale, ...)``: directed R-MAT, (a,b,c,d) = (0.57,0.19,0.19,0.05),
  self-loops/duplicates removed, ids permuted by a seeded Fisher-Yates.
* ``planted(n, ...)``: planted-SCC graph whose true SCC partition is known.
* ``UpdateStream``: insert/delete batches against a base graph.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libsccgen.so")
_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"{_LIB_PATH} is not built; run __graft_entry__.build() or "
            "make -C dynamicgraphs_scc_amd/csrc")
    lib = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    lib.sccgen_last_error.restype = C.c_char_p
    lib.sccgen_from_edges.restype = vp
    lib.sccgen_from_edges.argtypes = [C.c_int32, C.c_int64, i32p, i32p, C.c_int]
    lib.sccgen_read_edgelist.restype = vp
    lib.sccgen_read_edgelist.argtypes = [C.c_char_p]
    lib.sccgen_write_edgelist.restype = C.c_int
    lib.sccgen_write_edgelist.argtypes = [vp, C.c_char_p]
    lib.sccgen_read_gm_binary.restype = vp
    lib.sccgen_read_gm_binary.argtypes = [C.c_char_p]
    lib.sccgen_write_gm_binary.restype = C.c_int
    lib.sccgen_write_gm_binary.argtypes = [vp, C.c_char_p]
    lib.sccgen_rmat.restype = vp
    lib.sccgen_rmat.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_uint64, C.c_int, C.c_int, C.c_int64]
    lib.sccgen_planted.restype = vp
    lib.sccgen_planted.argtypes = [C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_double, C.c_int32, C.c_double, C.c_double, C.c_uint64,
                                   vp]
    lib.sccgen_num_nodes.restype = C.c_int32
    lib.sccgen_num_nodes.argtypes = [vp]
    lib.sccgen_num_edges.restype = C.c_int64
    lib.sccgen_num_edges.argtypes = [vp]
    lib.sccgen_copy.restype = None
    lib.sccgen_copy.argtypes = [vp, vp, vp, vp, vp]
    lib.sccgen_free.restype = None
    lib.sccgen_free.argtypes = [vp]
    lib.sccgen_updates_create.restype = vp
    lib.sccgen_updates_create.argtypes = [vp, C.c_uint64, C.c_int64, C.c_int64]
    lib.sccgen_updates_set_sizes.restype = C.c_int
    lib.sccgen_updates_set_sizes.argtypes = [vp, C.c_int64, C.c_int64]
    lib.sccgen_updates_peek.restype = C.c_int64
    lib.sccgen_updates_peek.argtypes = [vp]
    lib.sccgen_updates_next.restype = C.c_int
    lib.sccgen_updates_next.argtypes = [vp, C.c_int64, i32p, i32p, C.POINTER(C.c_int64),
                                        i32p, i32p, C.POINTER(C.c_int64)]
    lib.sccgen_updates_num_edges.restype = C.c_int64
    lib.sccgen_updates_num_edges.argtypes = [vp]
    lib.sccgen_updates_graph.restype = vp
    lib.sccgen_updates_graph.argtypes = [vp]
    lib.sccgen_updates_free.restype = None
    lib.sccgen_updates_free.argtypes = [vp]
    lib.sccgen_canonicalize.restype = C.c_int
    lib.sccgen_canonicalize.argtypes = [C.c_int32, i32p, i32p]
    _lib = lib
    return lib


def _err(lib) -> str:
    return lib.sccgen_last_error().decode()


@dataclass
class Graph:
    """Int32 CSR + reverse CSR (rows sorted)."""

    begin: np.ndarray
    node_idx: np.ndarray
    r_begin: np.ndarray
    r_node_idx: np.ndarray

    @property
    def num_nodes(self) -> int:
        return int(self.begin.shape[0] - 1)

    @property
    def num_edges(self) -> int:
        return int(self.node_idx.shape[0])

    def out_degree(self) -> np.ndarray:
        return np.diff(self.begin)

    def in_degree(self) -> np.ndarray:
        return np.diff(self.r_begin)

    def edges(self):
        """(src, dst) arrays in CSR order."""
        src = np.repeat(np.arange(self.num_nodes, dtype=np.int32), self.out_degree())
        return src, self.node_idx.copy()

    def arrays(self):
        return self.begin, self.node_idx, self.r_begin, self.r_node_idx

    def write_edgelist(self, path: str) -> None:
        """Write the reference's 1-indexed text edge list (src/common_main.h:61-83)."""
        lib = _load()
        h = _to_handle(self)
        try:
            if lib.sccgen_write_edgelist(h, path.encode()) != 0:
                raise RuntimeError(_err(lib))
        finally:
            lib.sccgen_free(h)

    def write_gm_binary(self, path: str) -> None:
        """Write the forward CSR as a gm binary graph (big-endian,
        gm_graph/src/gm_graph_binary_loader.cc:207-248)."""
        lib = _load()
        h = _to_handle(self)
        try:
            if lib.sccgen_write_gm_binary(h, path.encode()) != 0:
                raise RuntimeError(_err(lib))
        finally:
            lib.sccgen_free(h)


def _from_handle(h) -> Graph:
    lib = _load()
    if not h:
        raise RuntimeError(_err(lib))
    try:
        n = lib.sccgen_num_nodes(h)
        m = lib.sccgen_num_edges(h)
        b = np.empty(n + 1, np.int32)
        a = np.empty(m, np.int32)
        rb = np.empty(n + 1, np.int32)
        ra = np.empty(m, np.int32)
        lib.sccgen_copy(h, b.ctypes.data, a.ctypes.data, rb.ctypes.data, ra.ctypes.data)
    finally:
        lib.sccgen_free(h)
    return Graph(b, a, rb, ra)


def _to_handle(g: Graph):
    lib = _load()
    src, dst = g.edges()
    h = lib.sccgen_from_edges(g.num_nodes, g.num_edges, np.ascontiguousarray(src, np.int32),
                              np.ascontiguousarray(dst, np.int32), 0)
    if not h:
        raise RuntimeError(_err(lib))
    return h


def from_edges(n: int, src, dst, dedup: bool = False) -> Graph:
    """CSR + reverse CSR from an edge list.  ``dedup=False`` keeps self-loops and
    duplicates exactly as the reference loader does (gm_graph.cc:459-481)."""
    lib = _load()
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    if src.shape != dst.shape:
        raise ValueError("src and dst differ in length")
    return _from_handle(lib.sccgen_from_edges(int(n), int(src.shape[0]), src, dst, int(dedup)))


def read_edgelist(path: str) -> Graph:
    """Load the reference's text edge list (src/common_main.h:61-83): 1-indexed,
    single-space separated, N = max id."""
    lib = _load()
    return _from_handle(lib.sccgen_read_edgelist(path.encode()))


def read_gm_binary(path: str) -> Graph:
    """Load a gm binary graph (gm_graph/src/gm_graph_binary_loader.cc:48-205):
    rows sorted and the reverse CSR built, as load_binary does."""
    lib = _load()
    return _from_handle(lib.sccgen_read_gm_binary(path.encode()))


RMAT_ABCD = (0.57, 0.19, 0.19, 0.05)


def rmat(scale: int, edge_factor: float = 16, seed: int = 1, abc=RMAT_ABCD[:3],
         permute: bool = True, dedup: bool = True, n_keep: int = 0) -> Graph:
    lib = _load()
    a, b, c = abc
    return _from_handle(lib.sccgen_rmat(int(scale), float(edge_factor), a, b, c,
                                        int(seed) & (2**64 - 1), int(permute), int(dedup),
                                        int(n_keep)))


def planted(n: int, seed: int = 3, giant_frac: float = 0.4, n_medium: int = 200_000,
            size_min: int = 3, size_max: int = 1000, alpha: float = 2.1, intra_out: int = 8,
            inter_per_vertex: float = 4.0, pair_frac: float = 0.5, with_components=False):
    """Planted-SCC graph.  With ``with_components`` also returns the planted
    component index of every vertex (the true SCC partition)."""
    lib = _load()
    comp = np.empty(n, np.int32) if with_components else None
    h = lib.sccgen_planted(int(n), giant_frac, int(n_medium), int(size_min), int(size_max),
                           alpha, int(intra_out), inter_per_vertex, pair_frac,
                           int(seed) & (2**64 - 1),
                           comp.ctypes.data if comp is not None else None)
    g = _from_handle(h)
    return (g, comp) if with_components else g


def disjoint_union(gs) -> tuple:
    """(union Graph, vertex offsets) of independent graphs: graph g keeps its
    rows, its ids shifted by offsets[g]; no edge joins two graphs."""
    ns = np.array([g.num_nodes for g in gs], np.int64)
    ms = np.array([g.num_edges for g in gs], np.int64)
    voff = np.concatenate([[0], np.cumsum(ns)])
    eoff = np.concatenate([[0], np.cumsum(ms)])
    if voff[-1] >= 2**31 or eoff[-1] >= 2**31:
        raise ValueError("union exceeds int32 ids or offsets")
    begin = np.empty(voff[-1] + 1, np.int32)
    rbegin = np.empty(voff[-1] + 1, np.int32)
    adj = np.empty(eoff[-1], np.int32)
    radj = np.empty(eoff[-1], np.int32)
    for i, g in enumerate(gs):
        v0, v1, e0, e1 = voff[i], voff[i + 1], eoff[i], eoff[i + 1]
        begin[v0:v1] = g.begin[:-1] + e0
        rbegin[v0:v1] = g.r_begin[:-1] + e0
        adj[e0:e1] = g.node_idx + v0
        radj[e0:e1] = g.r_node_idx + v0
    begin[-1] = rbegin[-1] = eoff[-1]
    return Graph(begin, adj, rbegin, radj), voff.astype(np.int32)


def batched_config(k: int = 512, seed: int = 4000, n_lo: int = 1_000, n_hi: int = 1_000_000):
    """The batched workload: k graphs with vertex counts log-uniform in
    [n_lo, n_hi]; even g is R-MAT (edge factor uniform in [8, 32], (a,b,c,d) =
    (0.57, 0.19, 0.19, 0.05), 2^ceil(log2 N) ids generated and only edges
    between the first N kept, ids permuted, duplicates and self-loops removed),
    odd g is planted-SCC (40 % giant, 0.024 N medium SCCs of power-law sizes in
    [3, 1000], singletons and 2-cycles, 4 inter-component edges per vertex).
    Graph g is seeded with seed + g.  Returns the list of graphs."""
    out = []
    for g in range(k):
        rng = np.random.default_rng(seed + g)
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        if g % 2 == 0:
            ef = float(rng.uniform(8, 32))
            scale = max(1, int(np.ceil(np.log2(n))))
            out.append(rmat(scale, ef, seed=seed + g, n_keep=n))
        else:
            out.append(planted(n, seed=seed + g, n_medium=max(1, int(round(0.024 * n)))))
    return out


def canonicalize(labels) -> np.ndarray:
    """Min-id canonical labels."""
    lib = _load()
    lab = np.ascontiguousarray(labels, np.int32)
    out = np.empty_like(lab)
    if lib.sccgen_canonicalize(lab.shape[0], lab, out) != 0:
        raise ValueError(_err(lib))
    return out


class UpdateStream:
    """Seeded insert/delete batches against a base graph: sizes log-uniform in
    [lo, hi], half inserts of uniform random absent pairs, half deletes drawn
    uniformly among the edges present at that point of the stream."""

    def __init__(self, base: Graph, seed: int = 5, lo: int = 1000, hi: int = 100_000):
        lib = _load()
        self._lib = lib
        hb = _to_handle(base) if not isinstance(base, int) else base
        self._h = lib.sccgen_updates_create(hb, int(seed) & (2**64 - 1), int(lo), int(hi))
        lib.sccgen_free(hb)
        if not self._h:
            raise RuntimeError(_err(lib))
        self.hi = int(hi)

    def next(self):
        """Returns (ins_src, ins_dst, del_src, del_dst) int32 arrays."""
        cap = self.hi
        isrc = np.empty(cap, np.int32)
        idst = np.empty(cap, np.int32)
        dsrc = np.empty(cap, np.int32)
        ddst = np.empty(cap, np.int32)
        ni = C.c_int64()
        nd = C.c_int64()
        rc = self._lib.sccgen_updates_next(self._h, cap, isrc, idst, C.byref(ni), dsrc, ddst,
                                           C.byref(nd))
        if rc != 0:
            raise RuntimeError(_err(self._lib))
        return (isrc[:ni.value].copy(), idst[:ni.value].copy(), dsrc[:nd.value].copy(),
                ddst[:nd.value].copy())

    def set_sizes(self, lo: int, hi: int) -> None:
        """Batch sizes of the following batches: log-uniform in [lo, hi]."""
        if self._lib.sccgen_updates_set_sizes(self._h, int(lo), int(hi)) != 0:
            raise RuntimeError(_err(self._lib))
        self.hi = int(hi)

    def num_edges(self) -> int:
        return int(self._lib.sccgen_updates_num_edges(self._h))

    def graph(self) -> Graph:
        return _from_handle(self._lib.sccgen_updates_graph(self._h))

    def close(self):
        if self._h:
            self._lib.sccgen_updates_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Shared test helpers: fixture loading, independent checkers (scipy, brute force)."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from dynamicgraphs_scc_amd import graphs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden_cases():
    return sorted(manifest().keys())


def load_golden(name):
    """(graph as the reference loads it, canonical labels, scc_count)."""
    info = manifest()[name]
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    n = int(z["n"])
    if "src" in z:
        src, dst = z["src"], z["dst"]
    else:
        p = info["params"]
        if p["kind"] == "rmat":
            g0 = graphs.rmat(p["scale"], p["edge_factor"], seed=p["seed"])
        elif p["kind"] == "planted":
            g0 = graphs.planted(p["n"], seed=p["seed"], n_medium=p["n_medium"])
        else:
            raise ValueError(p)
        src, dst = g0.edges()
        sha = hashlib.sha256(src.astype("<i4").tobytes() + dst.astype("<i4").tobytes()).hexdigest()
        assert sha == info["edges_sha256"], f"{name}: regenerated graph differs from the fixture"
    g = graphs.from_edges(n, src, dst, dedup=False)
    return g, z["canon"].astype(np.int32), int(z["scc_count"])


def scipy_canonical(g) -> np.ndarray:
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    n = g.num_nodes
    A = csr_matrix((np.ones(g.num_edges, np.int8), g.node_idx, g.begin), shape=(n, n))
    _, lab = connected_components(A, directed=True, connection="strong")
    return graphs.canonicalize(lab.astype(np.int32))


def brute_canonical(g) -> np.ndarray:
    """Mutual reachability by transitive closure (tiny graphs only)."""
    n = g.num_nodes
    R = np.eye(n, dtype=bool)
    src, dst = g.edges()
    R[src, dst] = True
    for k in range(n):  # Floyd-Warshall closure
        R |= np.outer(R[:, k], R[k, :])
    M = R & R.T
    return np.array([int(np.argmax(M[v])) for v in range(n)], np.int32)


def random_graph(n, m, seed, self_loops=True, dups=True):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    g = graphs.from_edges(n, src, dst, dedup=not dups)
    return g


def same_partition(a, b) -> bool:
    """True when two labelings induce the same partition."""
    return np.array_equal(graphs.canonicalize(np.unique(a, return_inverse=True)[1].astype(np.int32)),
                          graphs.canonicalize(np.unique(b, return_inverse=True)[1].astype(np.int32)))


def apply_update_host(n, src, dst, ins, dels):
    """The reference's update semantics on a multigraph edge list: a delete drops
    every copy of its edge (src/common_main.h:379-386), an insert adds an edge
    that is absent (once, whatever the batch repeats), inserted self-loops are
    ignored, and a delete whose edge the same batch inserts is dropped."""
    keys = src.astype(np.int64) * n + dst
    ik = np.unique(ins[0].astype(np.int64) * n + ins[1])
    ik = ik[(ik // n) != (ik % n)]
    dk = np.setdiff1d(np.unique(dels[0].astype(np.int64) * n + dels[1]), ik)
    present = np.isin(ik, keys)
    ik = ik[~present]
    keep = ~np.isin(keys, dk)
    removed = int((~keep).sum())
    k2 = np.concatenate([keys[keep], ik])
    return (k2 // n).astype(np.int32), (k2 % n).astype(np.int32), removed

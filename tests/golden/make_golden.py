"""Write the golden fixtures in this folder from the reference's own binary.

Each case is a graph (hand-made edge lists, or seeded R-MAT / planted graphs
from dynamicgraphs_scc_amd.graphs) written as the reference's 1-indexed text
edge list, run through oracle/_ref/scc_ref (the reference's OpenMP method 2,
compiled from its sources by ``make -C oracle ref``) with ``-p``, and stored as:

* ``<case>.npz``: ``src``, ``dst`` (0-indexed input edges, reference order;
  omitted for the large seeded graphs, which tests regenerate from the
  manifest's params and check against ``edges_sha256``),
  ``n`` (vertex count as the reference loads it: max 1-indexed id),
  ``canon`` (min-id canonical label per vertex, from scc_list.txt),
  ``scc_count`` (the ``Total # SCCs`` line).
* ``manifest.json``: per case the parameters, n, m, scc_count, runtimes and the
  SHA-256 of ``canon`` (little-endian int32).

Run in a container that has /root/reference:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from dynamicgraphs_scc_amd import graphs  # noqa: E402
import oracle  # noqa: E402


def hand_cases():
    """Small graphs covering the edge cases the path must handle."""
    cases = {}
    # a single directed cycle 0->1->...->9->0 plus a tail 10->0, 9->11
    e = [(i, (i + 1) % 10) for i in range(10)] + [(10, 0), (9, 11)]
    cases["cycle_tail"] = (12, e)
    # isolated 2-cycles, a 3-cycle, self-loops and duplicate edges
    e = [(0, 1), (1, 0), (2, 3), (3, 2), (3, 2), (4, 4), (4, 5), (5, 6), (6, 4), (5, 5),
         (7, 8), (8, 7), (8, 7), (9, 9)]
    cases["pairs_loops_dups"] = (10, e)
    # a DAG (every vertex its own SCC)
    e = [(i, j) for i in range(12) for j in range(i + 1, min(12, i + 4))]
    cases["dag"] = (12, e)
    # two SCCs joined one way, with an isolated vertex in the middle (id 5)
    e = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 3), (6, 7), (7, 6), (4, 6)]
    cases["chain_of_sccs"] = (8, e)
    # a 2-cycle whose members also have other same-colour neighbours (not trim2-able)
    e = [(0, 1), (1, 0), (1, 2), (2, 3), (3, 1), (0, 3)]
    cases["pair_not_isolated"] = (4, e)
    # nested cycles (figure-eight) and a long path feeding back
    e = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 2)] + [(5 + i, 6 + i) for i in range(20)] + [(25, 5), (25, 0)]
    cases["figure_eight_long_cycle"] = (26, e)
    return cases


def synthetic_cases():
    cases = {}
    for scale, seed in [(8, 11), (10, 12), (12, 13)]:
        g = graphs.rmat(scale, 16, seed=seed)
        cases[f"rmat{scale}_s{seed}"] = (g, {"kind": "rmat", "scale": scale, "edge_factor": 16,
                                             "seed": seed})
    g = graphs.rmat(16, 16, seed=1)
    cases["rmat16_s1"] = (g, {"kind": "rmat", "scale": 16, "edge_factor": 16, "seed": 1})
    g = graphs.planted(20000, seed=3, n_medium=500)
    cases["planted20k_s3"] = (g, {"kind": "planted", "n": 20000, "n_medium": 500, "seed": 3})
    # sparse R-MAT leaves non-trivial components for WCC + per-partition FB
    g = graphs.rmat(12, 3, seed=14)
    cases["rmat12_ef3_s14"] = (g, {"kind": "rmat", "scale": 12, "edge_factor": 3, "seed": 14})
    return cases


def run_case(name, n, src, dst, params, manifest, threads_list=(1, 4)):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "edges.txt")
        with open(path, "w") as f:
            for a, b in zip(src.tolist(), dst.tolist()):
                f.write(f"{a + 1} {b + 1}\n")
        n_ref = int(max(src.max(), dst.max()) + 1) if len(src) else 0
        canon = None
        times = {}
        count = None
        for t in threads_list:
            res = oracle.run_ref(path, threads=t, method=2, flag="-p", cwd=td)
            c = oracle.dense_to_canonical(res["dense"])
            if canon is None:
                canon, count = c, res["scc_count"]
            elif not np.array_equal(canon, c) or count != res["scc_count"]:
                raise SystemExit(f"{name}: reference disagrees with itself across threads")
            times[str(t)] = res.get("runtime_ms")
    assert canon.shape[0] == n_ref
    edges_sha = hashlib.sha256(src.astype("<i4").tobytes() + dst.astype("<i4").tobytes()).hexdigest()
    arrays = dict(n=np.int64(n_ref), canon=canon.astype(np.int32), scc_count=np.int64(count))
    if len(src) <= 200_000:  # large seeded graphs are regenerated from params + edges_sha256
        arrays.update(src=src.astype(np.int32), dst=dst.astype(np.int32))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
    manifest[name] = {
        "params": params, "n": n_ref, "n_generated": int(n), "m": int(len(src)),
        "scc_count": int(count), "ref_runtime_ms": times,
        "edges_sha256": edges_sha,
        "canon_sha256": hashlib.sha256(canon.astype("<i4").tobytes()).hexdigest(),
    }
    print(f"{name}: n={n_ref} m={len(src)} sccs={count}")


def main():
    if not oracle.have_ref():
        oracle.build(with_ref=True)
    manifest = {}
    for name, (n, e) in hand_cases().items():
        src = np.array([a for a, _ in e], np.int32)
        dst = np.array([b for _, b in e], np.int32)
        run_case(name, n, src, dst, {"kind": "hand"}, manifest)
    for name, (g, params) in synthetic_cases().items():
        src, dst = g.edges()
        run_case(name, g.num_nodes, src, dst, params, manifest)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

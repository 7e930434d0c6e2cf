"""The dynamic serving workload at its full size, checked against the oracle.

Base graph: directed R-MAT scale 23 (8,388,608 vertices), edge factor 16,
seed 23, resident in the degree vertex layout as it is served.  Batches come
from the serving stream (seed 5): sizes log-uniform in [1k, 100k] edges, half
inserts of uniform random pairs, half deletes of present edges, plus one batch
of exactly 100k edges.  The gapped row layout is given no spare room per row
and a small overflow area, so rows that gain edges move and the overflow runs
out: the batches go through the layout rebuild.  After every batch the
canonical labels equal the oracle's from-scratch decomposition of the
host-tracked edge set (src/scc_incremental.cc:7-27, src/scc_decremental.cc:7-24
and src/common_main.h:343-393 define the update semantics); the device CSR
equals the host's after the last batch."""
import time

import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import LAYOUT_DEGREE, SccEngine

pytestmark = pytest.mark.gpu


def test_rmat23_stream_with_rebuilds_matches_oracle():
    t0 = time.time()
    base = graphs.rmat(23, 16, seed=23)
    n = base.num_nodes
    print(f"  base n={n} m={base.num_edges} ({time.time() - t0:.1f}s)", flush=True)
    us = graphs.UpdateStream(base, seed=5, lo=1000, hi=100_000)
    rebuilds = 0
    sizes = []
    with SccEngine() as e:
        e.set_layout(LAYOUT_DEGREE)
        e.upload(base)
        del base
        e.set_update_room(row_spare=0, row_spare_eighths=0, overflow_eighths=0, overflow_min=1 << 16)
        e.run_method2()
        for b in range(6):
            if b == 3:
                us.set_sizes(100_000, 100_000)
            elif b == 4:
                us.set_sizes(1000, 100_000)
            isrc, idst, dsrc, ddst = us.next()
            sizes.append(len(isrc) + len(dsrc))
            st = e.update(isrc, idst, dsrc, ddst)
            rebuilds += st["rebuilds"]
            assert st["inserted"] == len(isrc) and st["deleted"] == len(dsrc)
            cur = us.graph()
            want = oracle.scc_canonical(cur, 16)
            got = e.labels()
            print(f"  batch {b}: +{len(isrc)} -{len(dsrc)} rebuilds {st['rebuilds']} "
                  f"affected {st['affected']} sccs {st['num_sccs']} ({time.time() - t0:.1f}s)", flush=True)
            assert np.array_equal(got, want), f"batch {b}: {(got != want).sum()} labels differ"
            assert st["num_sccs"] == len(np.unique(want))
        g = e.download_graph()
        assert np.array_equal(g.begin, cur.begin) and np.array_equal(g.node_idx, cur.node_idx)
        assert np.array_equal(g.r_begin, cur.r_begin) and np.array_equal(g.r_node_idx, cur.r_node_idx)
    us.close()
    assert 100_000 in sizes
    assert rebuilds > 0
    print(f"  sizes {sizes}, rebuilds {rebuilds} ({time.time() - t0:.1f}s)", flush=True)

"""GPU behaviour of the queued chain (scc_ctx_set_chain): a fresh decomposition
queued as one chain of launches gives the labels and run statistics of the
phase-by-phase run; a chain stopped at any of its gates (too few level launches
or sweeps queued, or vertices still live after its last phase) continues
phase by phase to the same labels; later calls on the context see no halted
state.  Labels are compared bit-exactly with the oracle."""
import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import LAYOUT_DEGREE, LAYOUT_INPUT, SccEngine
from helpers import load_golden, random_graph

pytestmark = pytest.mark.gpu


def _path(n):
    """A path 0 -> 1 -> ... -> n-1 plus one 3-cycle at its end: the first trim
    loop peels the path over several sweeps."""
    src = np.concatenate([np.arange(n - 1), [n - 1, n - 2]]).astype(np.int32)
    dst = np.concatenate([np.arange(1, n), [n - 3, n - 1]]).astype(np.int32)
    return graphs.from_edges(n, src, dst)


def _dag(n, m, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keep = lo != hi
    return graphs.from_edges(n, lo[keep].astype(np.int32), hi[keep].astype(np.int32))


def _cases():
    return {
        "rmat14": graphs.rmat(14, 8, seed=5),
        "rmat12_ef16": graphs.rmat(12, 16, seed=9),
        "planted30k": graphs.planted(30000, seed=11, n_medium=500),
        "random_sparse": random_graph(20000, 24000, seed=3),
        "dag": _dag(5000, 20000, seed=4),
        "path": _path(3000),
        "golden_rmat12": load_golden("rmat12_s13")[0],
    }


# (deterministic ones: how far a relative-stop trim loop gets before the WCC
# phase depends on the order of its sweeps' colour reads, so wcc_parts is not)
STATS = ("num_sccs", "scc_first", "pivot_first")


@pytest.mark.parametrize("layout", [LAYOUT_INPUT, LAYOUT_DEGREE])
def test_chain_equals_phase_by_phase(layout):
    cases = _cases()
    with SccEngine() as on, SccEngine() as off:
        off.set_chain(False)
        for e in (on, off):
            e.set_layout(layout)
        for name, g in cases.items():
            want = oracle.scc_canonical(g, 8)
            off.upload(g)
            so = off.run_method2()
            assert so["chain_stop"] == -1
            assert np.array_equal(off.labels(), want), name
            on.upload(g)
            stops = []
            for rep in range(3):  # the chain resizes itself from the last run
                if rep:
                    on.reset()
                s = on.run_method2()
                stops.append(s["chain_stop"])
                assert np.array_equal(on.labels(), want), (name, rep)
                for k in STATS:
                    assert s[k] == so[k], (name, rep, k)
            print(f"  {name}: chain stops {stops}", flush=True)
            assert stops[0] >= 0
        # a graph that every vertex leaves within the chain runs it to its end
        on.upload(cases["rmat14"])
        on.run_method2()
        on.reset()
        assert on.run_method2()["chain_stop"] == 5


@pytest.mark.parametrize("layout", [LAYOUT_INPUT, LAYOUT_DEGREE])
@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("levels,sweeps", [(1, 1), (2, 4), (8, 1), (64, 64)])
def test_chain_stopped_at_each_gate(layout, shape, levels, sweeps):
    """Chains of every length with too few (or many) level launches and sweeps:
    wherever a gate stops them, the run ends with the oracle's labels.  In the
    degree layout the last gate runs inside the canonical-label launch and the
    gate after the global step inside the recolouring."""
    with SccEngine() as e:
        e.set_layout(layout)
        for name, g in _cases().items():
            if name == "golden_rmat12":
                continue
            want = oracle.scc_canonical(g, 8)
            e.upload(g)
            e.set_chain(True, shape, levels, sweeps)
            s = e.run_method2()
            last_gate = {0: 0, 1: 1, 2: 2, 3: 4}[shape]
            assert s["chain_stop"] == 5 or 0 <= s["chain_stop"] <= last_gate, s["chain_stop"]
            assert np.array_equal(e.labels(), want), (name, shape, levels, sweeps, s["chain_stop"])
            assert s["num_sccs"] == len(np.unique(want))


def test_phase_calls_after_a_chain():
    """After chained runs (stopped and complete) the phase-by-phase calls and
    updates work on the same context: no launch is left halted."""
    g = graphs.planted(30000, seed=11, n_medium=500)
    want = oracle.scc_canonical(g, 8)
    with SccEngine() as e:
        e.upload(g)
        e.set_chain(True, 1, 1, 1)  # stops at the global step's gate
        e.run_method2()
        assert np.array_equal(e.labels(), want)
        e.reset()
        e.repeat_global_trim1(0)
        e.do_fw_bw_global_main()
        e.repeat_global_trim1(0)
        e.do_global_trim2_new()
        e.do_global_wcc()
        e.start_workers_fw_bw_dfs()
        assert np.array_equal(e.labels(), want)
        r = graphs.rmat(13, 8, seed=2)
        e.upload(r)
        assert e.run_method2()["chain_stop"] >= 0
        e.reset()
        e.repeat_global_trim1(0)
        e.do_fw_bw_global_main()
        e.repeat_global_trim1(0)
        e.do_global_trim2_new()
        e.do_global_wcc()
        e.start_workers_fw_bw_dfs()
        assert np.array_equal(e.labels(), oracle.scc_canonical(r, 8))
        # an update right after a chained run
        e.reset()
        e.run_method2()
        src = np.array([0, 5, 9], np.int32)
        dst = np.array([7, 1, 2], np.int32)
        e.update(src, dst, src[:0], dst[:0])
        s, d = r.edges()
        g2 = graphs.from_edges(r.num_nodes, np.concatenate([s, src]), np.concatenate([d, dst]))
        assert np.array_equal(e.labels(), oracle.scc_canonical(g2, 8))


def test_chain_phase_trace():
    g = graphs.rmat(14, 8, seed=5)
    with SccEngine() as e:
        e.upload(g)
        s = e.run_method2()
        tr = e.phase_trace()
    kinds = [r["phase"] for r in tr]
    print("  trace", kinds, "stop", s["chain_stop"], flush=True)
    assert kinds[0] == "trim1" and kinds[1] == "fb_global" and kinds[-1] == "canon"
    assert tr[0]["live_before"] == g.num_nodes
    assert all(r["live_after"] <= r["live_before"] for r in tr if r["phase"] != "canon")

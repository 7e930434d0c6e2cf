"""Batched graphs through the same launches and launch-shape tuning: labels
equal the oracle's on the union of the graphs, per-graph counts, rejected
inputs, and exact labels under every tuning the autotuner can pick."""
import time

import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import SccEngine, SccError
from helpers import golden_cases, load_golden

pytestmark = pytest.mark.gpu


def _check_batch(e, gs):
    u, off = graphs.disjoint_union(gs)
    print(f"  batch of {len(gs)}: n={u.num_nodes} m={u.num_edges}", flush=True)
    e.upload(u)
    e.batch_set(off)
    e.batch_run()
    got = e.labels()
    want = oracle.scc_canonical(u, 16)
    assert np.array_equal(got, want)
    counts = e.batch_counts()
    for i in range(len(gs)):
        seg = want[off[i]:off[i + 1]]
        assert counts[i] == int(np.count_nonzero(seg == np.arange(off[i], off[i + 1])))
    return u, off


def test_batch_of_golden_and_synthetic_graphs():
    gs = [load_golden(n)[0] for n in golden_cases() if load_golden(n)[0].num_nodes < 100_000]
    gs += [graphs.rmat(12, 8, seed=s) for s in range(6)] + [graphs.planted(5000, seed=s, n_medium=100)
                                                           for s in range(6)]
    with SccEngine() as e:
        _check_batch(e, gs)


def test_batched_config_reduced_sizes():
    gs = graphs.batched_config(k=64, n_lo=1_000, n_hi=100_000)
    with SccEngine() as e:
        _check_batch(e, gs)


def test_batched_config_full():
    """512 graphs, vertex counts log-uniform in [1e3, 1e6]."""
    t = time.time()
    gs = graphs.batched_config()
    print(f"  generated in {time.time() - t:.1f}s", flush=True)
    with SccEngine() as e:
        _check_batch(e, gs)


def test_batch_rejects_cross_edges_and_bad_offsets():
    a, b = graphs.rmat(8, 4, seed=1), graphs.rmat(8, 4, seed=2)
    u, off = graphs.disjoint_union([a, b])
    src, dst = u.edges()
    bad = graphs.from_edges(u.num_nodes, np.append(src, 0), np.append(dst, u.num_nodes - 1))
    with SccEngine() as e:
        e.upload(bad)
        with pytest.raises(SccError) as ei:
            e.batch_set(off)
        assert ei.value.code == -1
        e.upload(u)
        for o in ([0, 300, u.num_nodes], [0, 0, u.num_nodes], [1, 256, u.num_nodes]):
            with pytest.raises(SccError):
                e.batch_set(o)
        e.batch_set(off)
        e.batch_run()
        e.batch_set(None)
        e.reset()
        e.run_method2()
        assert np.array_equal(e.labels(), oracle.scc_canonical(u, 8))


@pytest.mark.parametrize("knobs", [dict(light_row=4, huge_row=256, grid_cap=16, bu_alpha=0.5, td_beta=1.0),
                                   dict(light_row=64, huge_row=65536, grid_cap=4096, bu_alpha=1000.0,
                                        td_beta=1000.0),
                                   dict(light_row=4, huge_row=2048, grid_cap=512, bu_alpha=4.0, td_beta=8.0)])
def test_labels_exact_under_tuning(knobs):
    g = graphs.rmat(15, 12, seed=5)
    p = graphs.planted(40000, seed=6, n_medium=800)
    with SccEngine() as e:
        e.set_tuning(**knobs)
        assert e.tuning()["light_row"] == knobs["light_row"]
        for h in (g, p):
            e.upload(h)
            e.run_method2()
            assert np.array_equal(e.labels(), oracle.scc_canonical(h, 8))
        e.set_tuning()
        assert e.tuning()["light_row"] == 32 and e.tuning()["grid_cap"] == 640


def test_autotune_picks_a_valid_tuning():
    g = graphs.rmat(17, 16, seed=7)
    with SccEngine() as e:
        e.upload(g)
        r = e.autotune(reps=3)
        print("  autotune", r, flush=True)
        assert r["best_ms"] <= r["default_ms"] * 1.05
        assert np.array_equal(e.labels(), oracle.scc_canonical(g, 8))
        with pytest.raises(SccError):
            e.set_tuning(light_row=0)

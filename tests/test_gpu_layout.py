"""GPU parity with the degree vertex layout (SCC_LAYOUT_DEGREE): the context
relabels the graph by decreasing degree inside HBM, and every result seen
through the C ABI (canonical and raw labels, colours, pivots, the downloaded
CSR, updates) must be exactly what the input layout gives, in caller ids."""
import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import (LAYOUT_DEGREE, PIVOT_DEGREE_PRODUCT, PIVOT_MIN_ID,
                                          SCC_ERR_STATE, SccEngine, SccError)
from helpers import golden_cases, load_golden, random_graph, same_partition, scipy_canonical

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deng():
    e = SccEngine()
    e.set_layout(LAYOUT_DEGREE)
    yield e
    e.close()


def run(e, g, **kw):
    print(f"  degree layout n={g.num_nodes} m={g.num_edges} {kw}", flush=True)
    e.upload(g)
    st = e.run_method2(**kw)
    return e.labels(canonical=True), st


def _same_graph(a, b):
    for x, y in ((a.begin, b.begin), (a.node_idx, b.node_idx), (a.r_begin, b.r_begin),
                 (a.r_node_idx, b.r_node_idx)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name", golden_cases())
def test_golden_fixture(deng, name):
    g, canon, count = load_golden(name)
    lab, st = run(deng, g)
    assert np.array_equal(lab, canon)
    assert st["num_sccs"] == count


@pytest.mark.parametrize("policy", [PIVOT_DEGREE_PRODUCT, PIVOT_MIN_ID])
@pytest.mark.parametrize("flags", [dict(), dict(use_wcc=0), dict(trim_between=1), dict(trim_stop=0)])
def test_pipeline_options(deng, policy, flags):
    g = graphs.rmat(13, 6, seed=31)
    lab, _ = run(deng, g, pivot_policy=policy, **flags)
    assert np.array_equal(lab, oracle.scc_canonical(g, 8))


@pytest.mark.parametrize("seed,n,ef", [(1, 1, 0), (2, 2, 1), (3, 50, 0.5), (5, 3000, 1.3),
                                       (7, 20000, 5.0)])
def test_random_graphs(deng, seed, n, ef):
    g = random_graph(n, int(ef * n), seed)
    lab, _ = run(deng, g)
    assert np.array_equal(lab, scipy_canonical(g))


def test_raw_labels_colours_and_download(deng):
    g = graphs.rmat(14, 8, seed=5)
    lab, st = run(deng, g)
    raw = deng.labels(canonical=False)
    # a raw label is a member of the vertex's own component, in caller ids
    assert ((raw >= 0) & (raw < g.num_nodes)).all()
    assert np.array_equal(lab[raw], lab)
    assert (deng.colors() == -2).all()
    _same_graph(deng.download_graph(), g)
    # the reported first pivot is a caller id inside the first SCC found
    p = st["pivot_first"]
    assert 0 <= p < g.num_nodes and np.count_nonzero(lab == lab[p]) == st["scc_first"]


def test_phase_fb_global_given_pivot(deng):
    g = graphs.rmat(14, 8, seed=12)
    canon = scipy_canonical(g)
    big = int(np.bincount(canon).argmax())
    pivot = int(np.flatnonzero(canon == big)[-1])
    deng.upload(g)
    size, p = deng.do_fw_bw_global_main(pivot=pivot)
    assert p == pivot
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    assert size == oracle.fb_global(g, col, lab, -1, pivot)
    gcol = deng.colors()
    assert np.array_equal(gcol == -2, col == -2)
    assert np.array_equal(deng.labels(canonical=False)[col == -2], lab[col == -2])
    live = col != -2
    assert same_partition(gcol[live], col[live])
    # mid-run canonical labels: -1 until resolved, the component minimum after
    mid = deng.labels(canonical=True)
    assert (mid[~live] == canon[~live]).all() and (mid[live] == -1).all()


def test_phase_trim1_with_colours(deng):
    g = random_graph(20000, 30000, 9)
    colors = np.random.default_rng(9).integers(-1, 3, g.num_nodes).astype(np.int32)
    deng.upload(g)
    deng.set_colors(colors)
    assert np.array_equal(deng.colors(), colors)
    deng.repeat_global_trim1(0)
    col = colors.copy()
    lab = np.full(g.num_nodes, -1, np.int32)
    oracle.trim1_fixpoint(g, col, lab)
    assert np.array_equal(deng.colors() == -2, col == -2)
    assert np.array_equal(deng.labels(canonical=False), lab)


def test_update_stream(deng):
    base = graphs.rmat(12, 6, seed=2)
    deng.upload(base)
    deng.run_method2()
    us = graphs.UpdateStream(base, seed=12, lo=20, hi=3000)
    for b in range(6):
        isrc, idst, dsrc, ddst = us.next()
        print(f"  batch {b}: +{len(isrc)} -{len(dsrc)}", flush=True)
        st = deng.update(isrc, idst, dsrc, ddst)
        cur = us.graph()
        want = oracle.scc_canonical(cur, 8)
        assert np.array_equal(deng.labels(), want), f"batch {b}"
        assert st["num_sccs"] == len(np.unique(want))
        if b % 2 == 0:
            _same_graph(deng.download_graph(), cur)


@pytest.mark.parametrize("k,seed", [(2, 3), (24, 4)])
def test_batch_in_degree_layout(deng, k, seed):
    """A batch uploaded in the degree layout is re-ordered per graph (each graph
    keeps its id range): labels and per-graph counts equal the oracle's."""
    gs = graphs.batched_config(k=k, seed=seed, n_lo=1_000, n_hi=60_000)
    u, off = graphs.disjoint_union(gs)
    print(f"  degree-layout batch of {k}: n={u.num_nodes} m={u.num_edges}", flush=True)
    deng.upload(u)
    deng.batch_set(off)
    deng.batch_run()
    want = oracle.scc_canonical(u, 8)
    assert np.array_equal(deng.labels(), want)
    counts = deng.batch_counts()
    for i in range(k):
        seg = want[off[i]:off[i + 1]]
        assert counts[i] == int(np.count_nonzero(seg == np.arange(off[i], off[i + 1])))
    raw = deng.labels(canonical=False)
    assert np.array_equal(want[raw], want)
    _same_graph(deng.download_graph(), u)
    # a second run and clearing the batch keep exact labels
    deng.batch_set(None)
    deng.run_method2()
    assert np.array_equal(deng.labels(), want)


def test_batch_in_degree_layout_rejects_cross_edges(deng):
    a, b = graphs.rmat(9, 4, seed=1), graphs.rmat(9, 4, seed=2)
    u, off = graphs.disjoint_union([a, b])
    src, dst = u.edges()
    bad = graphs.from_edges(u.num_nodes, np.append(src, 0), np.append(dst, u.num_nodes - 1))
    deng.upload(bad)
    with pytest.raises(SccError):
        deng.batch_set(off)
    deng.run_method2()  # still a usable single graph
    assert np.array_equal(deng.labels(), scipy_canonical(bad))


def test_rmat22_full_size(deng):
    """R-MAT scale 22, EF 16, seed 22 in the degree layout: bit-exact vs oracle."""
    g = graphs.rmat(22, 16, seed=22)
    want = oracle.scc_canonical(g, 16)
    lab, st = run(deng, g)
    assert np.array_equal(lab, want)
    assert st["num_sccs"] == len(np.unique(want))
    assert st["scc_first"] == np.bincount(want).max()


def test_auto_layout_follows_skew():
    """SCC_LAYOUT_AUTO relabels a skewed R-MAT graph and keeps a planted graph's
    ids; labels are exact either way."""
    from dynamicgraphs_scc_amd.engine import LAYOUT_AUTO, LAYOUT_INPUT
    with SccEngine() as e:
        e.set_layout(LAYOUT_AUTO)
        g = graphs.rmat(14, 8, seed=21)
        lab, _ = run(e, g)
        assert e.layout() == LAYOUT_DEGREE
        assert np.array_equal(lab, oracle.scc_canonical(g, 8))
        g, comp = graphs.planted(30000, seed=5, n_medium=600, with_components=True)
        lab, _ = run(e, g)
        assert e.layout() == LAYOUT_INPUT
        assert np.array_equal(lab, graphs.canonicalize(comp))

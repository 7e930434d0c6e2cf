"""GPU parity: the HIP engine's canonical labels against the oracle, the
reference's golden fixtures, the planted partition and phase-level fixpoints.
Integer work, so every comparison is bit-exact."""
import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import (PARTS_BOTH, PARTS_COLOR, PIVOT_DEGREE_PRODUCT, PIVOT_MIN_ID,
                                          SccEngine, SccError, default_opts)
from helpers import golden_cases, load_golden, random_graph, same_partition, scipy_canonical

pytestmark = pytest.mark.gpu


def run(engine, g, **kw):
    print(f"  run n={g.num_nodes} m={g.num_edges} {kw}", flush=True)
    engine.upload(g)
    st = engine.run_method2(**kw)
    return engine.labels(canonical=True), st


@pytest.mark.parametrize("name", golden_cases())
def test_golden_fixture(engine, name):
    g, canon, count = load_golden(name)
    lab, st = run(engine, g)
    assert np.array_equal(lab, canon)
    assert st["num_sccs"] == count == len(np.unique(canon))


@pytest.mark.parametrize("policy", [PIVOT_DEGREE_PRODUCT, PIVOT_MIN_ID])
@pytest.mark.parametrize("flags", [dict(), dict(use_trim2=0), dict(use_wcc=0),
                                   dict(trim_between=1), dict(trim_stop=100),
                                   dict(trim_stop=0), dict(trim_stop=-2), dict(trim_stop=-50),
                                   dict(trim_stop=-100)])
def test_pipeline_options_same_labels(engine, policy, flags):
    g = graphs.rmat(13, 6, seed=31)
    want = oracle.scc_canonical(g, 8)
    lab, _ = run(engine, g, pivot_policy=policy, **flags)
    assert np.array_equal(lab, want)


@pytest.mark.parametrize("seed,n,ef", [(1, 1, 0), (2, 2, 1), (3, 50, 0.5), (4, 300, 1.0),
                                       (5, 3000, 1.3), (6, 20000, 2.0), (7, 20000, 5.0)])
def test_random_graphs_with_loops_and_duplicates(engine, seed, n, ef):
    g = random_graph(n, int(ef * n), seed)
    lab, _ = run(engine, g)
    assert np.array_equal(lab, scipy_canonical(g))


def test_no_edges_and_all_self_loops(engine):
    g = graphs.from_edges(100, np.zeros(0, np.int32), np.zeros(0, np.int32))
    lab, st = run(engine, g)
    assert np.array_equal(lab, np.arange(100)) and st["num_sccs"] == 100
    v = np.arange(100, dtype=np.int32)
    g = graphs.from_edges(100, v, v)
    lab, _ = run(engine, g)
    assert np.array_equal(lab, np.arange(100))


def test_long_cycle_and_path(engine):
    n = 5000
    v = np.arange(n, dtype=np.int32)
    g = graphs.from_edges(n, v, (v + 1) % n)  # one long cycle: deep BFS, no trimming
    lab, _ = run(engine, g)
    assert (lab == 0).all()
    g = graphs.from_edges(n, v[:-1], v[1:])  # a path: all singletons
    lab, _ = run(engine, g)
    assert np.array_equal(lab, v)


def test_config1_rmat16(engine):
    """R-MAT scale 16, EF 16, seed 1: GPU == oracle == Tarjan (== reference fixture)."""
    g = graphs.rmat(16, 16, seed=1)
    want = oracle.scc_canonical(g, 8)
    assert np.array_equal(oracle.canonicalize(oracle.tarjan(g)), want)
    lab, st = run(engine, g)
    assert np.array_equal(lab, want)
    _, canon, _ = load_golden("rmat16_s1")
    assert np.array_equal(lab, canon)


def test_config2_rmat22_full_size(engine):
    """R-MAT scale 22, EF 16, seed 22 (4,194,304 vertices): bit-exact vs oracle."""
    g = graphs.rmat(22, 16, seed=22)
    want = oracle.scc_canonical(g, 16)
    lab, st = run(engine, g)
    assert np.array_equal(lab, want)
    assert st["num_sccs"] == len(np.unique(want))
    # the giant SCC is found by the global step
    assert st["scc_first"] == np.bincount(want).max()


def test_config3_planted_8m(engine):
    """Planted graph, 2^23 vertices: labels equal the planted partition."""
    g, comp = graphs.planted(1 << 23, seed=3, with_components=True)
    lab, st = run(engine, g)
    assert np.array_equal(lab, graphs.canonicalize(comp))


def test_planted_small_many_partitions(engine):
    g, comp = graphs.planted(300000, seed=5, n_medium=8000, with_components=True)
    for policy in (PIVOT_DEGREE_PRODUCT, PIVOT_MIN_ID):
        lab, _ = run(engine, g, pivot_policy=policy)
        assert np.array_equal(lab, graphs.canonicalize(comp))


@pytest.mark.parametrize("layout", [0, 1])
def test_cut_rows_same_labels_as_graph_rows(layout):
    """Once a label propagation has run, the partition steps follow rows cut to
    each class (8 (m + n) more bytes); a context whose budget has no room for
    them follows the graph's rows.  Both give the planted partition in both
    partition modes, on repeated runs (the cut rows are remade per run), and
    through the phase calls (WCC, which cuts the out-rows, then
    forward-backward rounds that follow them)."""
    g, comp = graphs.planted(120000, seed=11, n_medium=3000, with_components=True)
    want = graphs.canonicalize(comp)
    with SccEngine() as free, SccEngine() as tight:
        for e in (free, tight):
            e.set_layout(layout)
            e.upload(g)
        tight.set_budget(tight.bytes + (1 << 20))
        for mode in (PARTS_COLOR, PARTS_BOTH, PARTS_COLOR):
            for e in (free, tight):
                print(f"  layout {layout} mode {mode} budget {e is tight}", flush=True)
                e.reset()
                e.run_method2(partition_mode=mode)
                assert np.array_equal(e.labels(), want), (mode, e is tight)
        assert free.bytes >= tight.bytes + 8 * g.num_edges
        for e in (free, tight):
            e.reset()
            e.repeat_global_trim1(0)
            e.do_fw_bw_global_main()
            e.repeat_global_trim1(0)
            e.do_global_trim2_new()
            e.repeat_global_trim1(0)
            e.do_global_wcc()
            e.start_workers_fw_bw_dfs()
            assert np.array_equal(e.labels(), want), e is tight


def test_repeated_runs_and_geometry_stable(engine):
    """Run-to-run stability (atomics make work order nondeterministic)."""
    g = graphs.rmat(15, 8, seed=77)
    want = oracle.scc_canonical(g, 8)
    engine.upload(g)
    for _ in range(3):
        engine.reset()
        engine.run_method2()
        assert np.array_equal(engine.labels(), want)


# ---- phase-level parity --------------------------------------------------------

def test_phase_trim1_fixpoint(engine):
    g = graphs.rmat(14, 4, seed=8)
    engine.upload(g)
    t = engine.repeat_global_trim1(0)
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    want = oracle.trim1_fixpoint(g, col, lab)
    assert t == want
    assert np.array_equal(engine.colors(), col)
    assert np.array_equal(engine.labels(canonical=False), lab)


def test_phase_trim1_with_colours(engine):
    g = random_graph(20000, 30000, 9)
    rng = np.random.default_rng(9)
    colors = rng.integers(-1, 3, g.num_nodes).astype(np.int32)
    engine.upload(g)
    engine.set_colors(colors)
    engine.repeat_global_trim1(0)
    col = colors.copy()
    lab = np.full(g.num_nodes, -1, np.int32)
    oracle.trim1_fixpoint(g, col, lab)
    assert np.array_equal(engine.colors() == -2, col == -2)
    assert np.array_equal(engine.labels(canonical=False), lab)


def test_phase_fb_global_given_pivot(engine):
    g = graphs.rmat(14, 8, seed=12)
    canon = scipy_canonical(g)
    big = int(np.bincount(canon).argmax())
    pivot = int(np.flatnonzero(canon == big)[-1])
    engine.upload(g)
    size, p = engine.do_fw_bw_global_main(pivot=pivot)
    assert p == pivot
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    want = oracle.fb_global(g, col, lab, -1, pivot)
    assert size == want
    gcol = engine.colors()
    assert np.array_equal(gcol == -2, col == -2)
    assert np.array_equal(engine.labels(canonical=False)[col == -2], lab[col == -2])
    # same fw-only / bw-only / rest partition (colour values differ)
    live = col != -2
    assert same_partition(gcol[live], col[live])


def test_phase_fb_global_degree_pivot_and_trim2(engine):
    g = graphs.rmat(14, 8, seed=13)
    engine.upload(g)
    engine.repeat_global_trim1(0)
    size, pivot = engine.do_fw_bw_global_main(policy=PIVOT_DEGREE_PRODUCT)
    deg = np.diff(g.begin).astype(np.int64) * np.diff(g.r_begin)
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    oracle.trim1_fixpoint(g, col, lab)
    live = np.flatnonzero(col == -1)
    best = live[deg[live] == deg[live].max()].min()
    assert pivot == best
    engine.repeat_global_trim1(0)
    pre = engine.colors()
    removed = engine.do_global_trim2_new()
    gl = engine.labels(canonical=False)
    gc = engine.colors()
    canon = scipy_canonical(g)
    assert removed % 2 == 0
    # every vertex trim2 removed is in a 2-vertex SCC labelled with its minimum
    pairs = np.flatnonzero((gc == -2) & (gl != np.arange(g.num_nodes)) & (gl != pivot))
    for v in pairs[:2000]:
        members = np.flatnonzero(canon == canon[v])
        assert len(members) == 2 and gl[v] == members.min()
    # and it removed at least every pair that qualifies against the colours
    # before the pass (a pass in any order removes those)
    snap = oracle.trim2_snapshot(g, pre)
    assert (gc[snap] == -2).all()


@pytest.mark.parametrize("layout", [0, 1])
def test_phase_trim2_removes_every_pair_of_the_snapshot(layout):
    # sparse random edges plus 3,000 planted 2-cycles; after the trim1 fixpoint
    # (and after a global step) trim2 removes at least every pair that
    # qualifies against the colours before the pass, each labelled with one of
    # its members (the smaller id; in the degree layout the smaller internal
    # id), and nothing but 2-vertex SCCs
    rng = np.random.default_rng(21)
    n = 40_000
    src = rng.integers(0, n, 40_000)
    dst = rng.integers(0, n, 40_000)
    p = rng.permutation(n)[:6000].reshape(3000, 2)
    src = np.concatenate([src, p[:, 0], p[:, 1]]).astype(np.int32)
    dst = np.concatenate([dst, p[:, 1], p[:, 0]]).astype(np.int32)
    g = graphs.from_edges(n, src, dst, dedup=True)
    canon = scipy_canonical(g)
    with SccEngine() as engine:
        engine.set_layout(layout)
        engine.upload(g)
        _trim2_snapshot_rounds(engine, g, canon, layout)


def _trim2_snapshot_rounds(engine, g, canon, layout):
    for with_fb in (False, True):
        engine.reset()
        engine.repeat_global_trim1(0)
        if with_fb:
            engine.do_fw_bw_global_main()
            engine.repeat_global_trim1(0)
        pre = engine.colors()
        snap = oracle.trim2_snapshot(g, pre)
        assert snap.sum() > 1000
        removed = engine.do_global_trim2_new()
        gc = engine.colors()
        gl = engine.labels(canonical=False)
        now = (gc == -2) & (pre != -2)
        assert removed == now.sum() >= snap.sum()
        assert (gc[snap] == -2).all()
        for v in np.flatnonzero(now):
            members = np.flatnonzero(canon == canon[v])
            assert len(members) == 2 and now[members].all()
            assert gl[members[0]] == gl[members[1]] and gl[v] in members
            if layout == 0:
                assert gl[v] == members.min()


def test_phase_wcc_roots_equal_oracle(engine):
    g = graphs.rmat(13, 3, seed=14)
    engine.upload(g)
    engine.repeat_global_trim1(0)
    parts = engine.do_global_wcc()
    gc = engine.colors()
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    oracle.trim1_fixpoint(g, col, lab)
    W = oracle.wcc(g, col)
    live = col != -2
    assert np.array_equal(gc[live], W[live] & 0x1FFFFFFF)
    assert parts == int((W[live] & 0x1FFFFFFF == np.flatnonzero(live)).sum())


def test_phase_wcc_long_chains_and_hub_rows(engine):
    """Label propagation whose labels travel far past one push launch's local
    rounds (three chains of 2-cycles, 20,000 vertices each, ids permuted, each
    chain reaching the next one way) and whose pushes cross long in-rows (three
    stars of 3,000 spokes around a hub, each star reaching a chain one way):
    the WCC labels equal the oracle's."""
    rng = np.random.default_rng(21)
    L, S = 20_000, 3_000
    n = 3 * L + 3 * (S + 1)
    perm = rng.permutation(n).astype(np.int32)
    chains = [perm[c * L:(c + 1) * L] for c in range(3)]
    src, dst = [], []
    for ids in chains:
        src += [ids[:-1], ids[1:]]
        dst += [ids[1:], ids[:-1]]
    for c in range(2):
        src.append(chains[c][rng.integers(0, L, 20)])
        dst.append(chains[c + 1][rng.integers(0, L, 20)])
    base = 3 * L
    for s in range(3):
        hub = perm[base + s * (S + 1)]
        spokes = perm[base + s * (S + 1) + 1:base + (s + 1) * (S + 1)]
        src += [spokes, np.full(S, hub, np.int32), spokes[:5]]
        dst += [np.full(S, hub, np.int32), spokes, chains[s][rng.integers(0, L, 5)]]
    g = graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), dedup=True)
    print(f"  chains and stars n={g.num_nodes} m={g.num_edges}", flush=True)
    engine.upload(g)
    engine.repeat_global_trim1(0)
    parts = engine.do_global_wcc()
    gc = engine.colors()
    col = np.full(g.num_nodes, -1, np.int32)
    lab = np.full(g.num_nodes, -1, np.int32)
    oracle.trim1_fixpoint(g, col, lab)
    W = oracle.wcc(g, col)
    live = col != -2
    assert live.sum() == n
    assert np.array_equal(gc[live], W[live] & 0x1FFFFFFF)
    assert parts == int((W[live] & 0x1FFFFFFF == np.flatnonzero(live)).sum())
    engine.reset()
    engine.run_method2()
    assert np.array_equal(engine.labels(canonical=True), oracle.scc_canonical(g, 8))


def test_phase_sequence_matches_method2(engine):
    """Calling the scc.h-named phases in method-2 order gives the same labels."""
    g = graphs.rmat(13, 4, seed=15)
    engine.upload(g)
    engine.repeat_global_trim1(0)
    engine.do_fw_bw_global_main()
    engine.repeat_global_trim1(0)
    engine.do_global_trim2_new()
    engine.repeat_global_trim1(0)
    engine.do_global_wcc()
    engine.start_workers_fw_bw_dfs()
    assert np.array_equal(engine.labels(), oracle.scc_canonical(g))


# ---- error handling (forced from the host; every kernel exits normally) ---------

def test_invalid_graph_rejected(engine):
    g = random_graph(10, 20, 1)
    b = g.begin.copy()
    b[3] = b[4] + 1
    bad = graphs.Graph(b, g.node_idx, g.r_begin, g.r_node_idx)
    with pytest.raises(SccError) as ei:
        engine.upload(bad)
    assert ei.value.code == -1
    a = g.node_idx.copy()
    a[0] = 99
    with pytest.raises(SccError):
        engine.upload(graphs.Graph(g.begin, a, g.r_begin, g.r_node_idx))


def test_budget_rejection_and_fault_injection():
    g = graphs.rmat(12, 8, seed=1)
    with SccEngine(hbm_budget=1 << 16) as e:
        with pytest.raises(SccError) as ei:
            e.upload(g)
        assert ei.value.code == -4
    with SccEngine() as e:
        e.upload(g)
        e.inject_fault(3)
        with pytest.raises(SccError) as ei:
            e.run_method2()
        assert ei.value.code == -6
        e.inject_fault(-1)
        e.upload(g)
        e.run_method2()
        assert np.array_equal(e.labels(), oracle.scc_canonical(g))


def test_two_contexts_independent():
    g1 = graphs.rmat(12, 8, seed=2)
    g2 = graphs.planted(30000, seed=2, n_medium=700)
    with SccEngine() as a, SccEngine() as b:
        a.upload(g1)
        b.upload(g2)
        a.run_method2()
        b.run_method2()
        assert np.array_equal(a.labels(), oracle.scc_canonical(g1))
        assert np.array_equal(b.labels(), oracle.scc_canonical(g2))


@pytest.mark.parametrize("layout", [0, 1])
def test_huge_row_with_duplicate_entries(layout):
    """A pivot row above the huge-row threshold (8192 entries) that holds
    repeated ids: the chunked expansion claims each id once, so the searches'
    visited counts stay exact.  Input and degree layouts."""
    rng = np.random.default_rng(5)
    n = 20000
    hub_out = np.sort(rng.integers(1, n, 12000)).astype(np.int32)   # duplicates stay adjacent
    hub_in = rng.integers(1, n, 12000).astype(np.int32)
    rs = rng.integers(1, n, 2 * n).astype(np.int32)
    rd = rng.integers(0, n, 2 * n).astype(np.int32)
    src = np.concatenate([np.zeros(12000, np.int32), hub_in, rs])
    dst = np.concatenate([hub_out, np.zeros(12000, np.int32), rd])
    g = graphs.from_edges(n, src, dst, dedup=False)
    assert np.diff(g.begin)[0] == 12000 and len(np.unique(hub_out)) < 12000
    with SccEngine() as e:
        e.set_layout(layout)
        lab, st = run(e, g)
        assert np.array_equal(lab, scipy_canonical(g))
        assert st["num_sccs"] == len(np.unique(lab))

"""GPU behaviour of the request server: every result in the mix equals the
same request run alone (exact, integer labels), requests on different graphs
overlap, admission follows the HBM budget, a faulted graph does not disturb
the others, and latency statistics are reported."""
import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import SccEngine, SccError
from dynamicgraphs_scc_amd.serve import SccServer
from helpers import apply_update_host

pytestmark = pytest.mark.gpu


def _workload(k, batches):
    """k graphs of different kinds and sizes, each with its update batches."""
    out = []
    for i in range(k):
        if i % 2 == 0:
            g = graphs.rmat(11 + (i % 3), 8, seed=100 + i)
        else:
            g = graphs.planted(8000 + 2000 * i, seed=200 + i, n_medium=300)
        us = graphs.UpdateStream(g, seed=300 + i, lo=50, hi=4000)
        out.append((g, [us.next() for _ in range(batches)], us.graph()))
        us.close()
    return out


def _alone(g, batches):
    with SccEngine() as e:
        e.upload(g)
        e.run_method2()
        res = [e.labels()]
        for a, b, c, d in batches:
            e.update(a, b, c, d)
            res.append(e.labels())
    return res


def test_mix_equals_alone_and_overlaps():
    """Five rounds of interleaved update + labels requests over six graphs,
    then three updates per graph queued back to back (merged into one device
    update) and a final labels read: every read equals the same graph run
    alone, one update at a time."""
    work = _workload(6, 8)
    alone = [_alone(g, bs) for g, bs, _ in work]
    print("  alone done", flush=True)
    with SccServer(slots=4) as srv:
        gids, tickets = [], []
        for g, _, _ in work:
            gid, t = srv.load(g)
            gids.append(gid)
            tickets.append(srv.labels(gid, g.num_nodes))
        per = [[t] for t in tickets]
        # round-robin interleaving: every graph has work queued at once
        for j in range(5):
            for i, (g, bs, _) in enumerate(work):
                a, b, c, d = bs[j]
                srv.update(gids[i], a, b, c, d)
                per[i].append(srv.labels(gids[i], g.num_nodes))
        # back to back: no labels request between them, so they may merge
        merged = []
        for i, (g, bs, _) in enumerate(work):
            merged += [srv.update(gids[i], *bs[j]) for j in range(5, 8)]
            per[i].append(srv.labels(gids[i], g.num_nodes))
        assert srv.drain(120_000)
        for i in range(len(work)):
            steps = list(range(6)) + [8]
            for (t, buf), j in zip(per[i], steps):
                assert srv.result(t)["status"] == 0
                assert np.array_equal(buf, alone[i][j]), f"graph {i} step {j}"
        groups = [srv.result(t)["group"] for t in merged]
        print("  merged groups", groups, flush=True)
        assert max(groups) >= 2
        st = srv.stats()
        print("  stats", st, flush=True)
        assert st["failed"] == 0 and st["updates_done"] == 48
        assert st["max_concurrency"] >= 2
        assert 0 < st["update_p50_ms"] <= st["update_p99_ms"] <= st["update_max_ms"]
    # and the final labels are the recompute's
    for i, (g, bs, final) in enumerate(work):
        assert np.array_equal(alone[i][-1], oracle.scc_canonical(final, 8))


def test_queued_updates_merge_like_one_by_one():
    """UPDATEs queued back to back on one graph run as merged device updates.
    The labels after each run of queued updates, and the edge multiset after
    all of them (read through the number of entries a final delete removes),
    equal those of applying the requests one by one on the host.  A request
    that inserts an edge an earlier queued request deletes starts a new group;
    one with an out-of-range id fails alone and changes nothing."""
    rng = np.random.default_rng(77)
    n = 20_000
    s0 = rng.integers(0, n, 60_000).astype(np.int32)
    d0 = rng.integers(0, n, 60_000).astype(np.int32)
    rep = rng.integers(1, 4, s0.size)  # every edge one to three times
    loops = rng.choice(n, 20, replace=False).astype(np.int32)
    base = graphs.from_edges(n, np.concatenate([np.repeat(s0, rep), loops]),
                             np.concatenate([np.repeat(d0, rep), loops]), dedup=False)
    src, dst = base.edges()
    keys, cnt = np.unique(src.astype(np.int64) * n + dst, return_counts=True)
    dup = keys[(cnt > 1) & (keys // n != keys % n)]
    k1, k3 = int(dup[0]), int(dup[1])  # duplicated edges
    present = set(keys.tolist())
    k2 = next(int(k) for k in rng.integers(0, n * n, 100) if k not in present and k // n != k % n)

    def one(k):
        return np.array([k // n], np.int32), np.array([k % n], np.int32)

    def batch(ni, nd, ins_extra=(), del_extra=()):
        pick = rng.choice(src.size, nd, replace=False)
        ins = [(rng.integers(0, n, ni).astype(np.int32), rng.integers(0, n, ni).astype(np.int32))]
        dels = [(src[pick], dst[pick])]
        ins += list(ins_extra)
        dels += list(del_extra)
        return (np.concatenate([a for a, _ in ins]), np.concatenate([b for _, b in ins]),
                np.concatenate([a for a, _ in dels]), np.concatenate([b for _, b in dels]))

    loop0 = (loops[:1], loops[:1])
    ups = [
        batch(1500, 800, ins_extra=[one(k2)]),                       # 0
        batch(900, 1200, del_extra=[one(k1)]),                       # 1 deletes k1 (3 copies or 2)
        batch(700, 600, del_extra=[one(k2)]),                        # 2 deletes what 0 inserted
        batch(1200, 900, ins_extra=[one(k1)], del_extra=[loop0]),    # 3 inserts k1 again: new group
        batch(400, 1500),                                            # 4
        batch(800, 300, ins_extra=[loop0]),                          # 5 self-loop insert: no effect
        batch(1000, 1000),                                           # 6
        batch(600, 700, ins_extra=[one(k3)], del_extra=[one(k3)]),   # 7 insert beats delete
        batch(10, 10, del_extra=[(np.array([n], np.int32), np.array([0], np.int32))]),  # 8 bad id
        batch(900, 800, ins_extra=[(loops[1:2], loops[1:2])], del_extra=[one(k1)]),  # 9
        batch(1300, 400),                                            # 10
        batch(500, 1100),                                            # 11
    ]
    probe_keys = np.unique(np.concatenate([dup, [k1, k2, k3], loops.astype(np.int64) * (n + 1)]))
    probe = ((probe_keys // n).astype(np.int32), (probe_keys % n).astype(np.int32))

    # one by one on the host
    want, cs, cd = [], src, dst
    for i, (a, b, c, d) in enumerate(ups):
        if i != 8:
            cs, cd, _ = apply_update_host(n, cs, cd, (a, b), (c, d))
        if i in (5, 11):
            want.append(oracle.scc_canonical(graphs.from_edges(n, cs, cd, dedup=False), 8))
    cs, cd, probe_removed = apply_update_host(n, cs, cd, (np.empty(0, np.int32),) * 2, probe)
    want.append(oracle.scc_canonical(graphs.from_edges(n, cs, cd, dedup=False), 8))
    assert probe_removed > dup.size  # duplicated entries survive to the probe

    with SccServer(slots=2) as srv:
        gid, tl = srv.load(base)
        tu, reads = [], []
        for i, (a, b, c, d) in enumerate(ups):  # queued while the load runs
            tu.append(srv.update(gid, a, b, c, d))
            if i in (5, 11):
                reads.append(srv.labels(gid, n))
        tp = srv.update(gid, np.empty(0, np.int32), np.empty(0, np.int32), *probe)
        reads.append(srv.labels(gid, n))
        assert srv.drain(120_000)
        assert srv.result(tl)["status"] == 0
        info = [srv.wait(t) for t in tu]
        print("  groups", [x["group"] for x in info], flush=True)
        for i, x in enumerate(info):
            if i == 8:
                assert x["status"] < 0 and x["group"] == 1
            else:
                assert x["status"] == 0, (i, x)
        assert max(x["group"] for x in info) >= 2
        assert info[1]["start_ms"] != info[3]["start_ms"]  # never one group
        # members of a group share their start and end
        for x in info:
            same = [y for y in info if y["start_ms"] == x["start_ms"]]
            assert len(same) == x["group"] and all(y["end_ms"] == x["end_ms"] for y in same)
        p = srv.wait(tp)
        assert p["status"] == 0 and p["group"] == 1
        assert p["deleted"] == probe_removed
        for j, (t, buf) in enumerate(reads):
            assert srv.result(t)["status"] == 0
            assert np.array_equal(buf, want[j]), f"read {j}"
        st = srv.stats()
        assert st["updates_done"] == 12 and st["failed"] == 1


def test_admission_waits_for_memory():
    a = graphs.rmat(13, 8, seed=1)
    b = graphs.rmat(13, 8, seed=2)
    with SccServer(slots=2) as probe:
        ga, ta = probe.load(a)
        probe.result(ta)
        one = probe.stats()["reserved_bytes"]
    assert one > 0
    # room for one graph only
    with SccServer(slots=2, hbm_budget_bytes=int(one * 1.5)) as srv:
        ga, ta = srv.load(a)
        gb, tb = srv.load(b)
        ra = srv.result(ta)
        info = srv.wait(tb, 300)
        assert info["status"] == 1, "the second load must wait for memory"
        tu = srv.unload(ga)
        ru = srv.result(tu)
        rb = srv.result(tb)
        assert rb["start_ms"] >= ru["end_ms"]
        assert rb["num_sccs"] == len(np.unique(oracle.scc_canonical(b, 8)))
        assert ra["num_sccs"] == len(np.unique(oracle.scc_canonical(a, 8)))
        # a graph that can never fit is refused at submit
        with pytest.raises(SccError) as ei:
            srv.load(graphs.rmat(16, 16, seed=3))
        assert ei.value.code == -4


def test_fault_in_one_graph_leaves_others_running():
    work = _workload(2, 3)
    with SccServer(slots=2) as srv:
        gids = [srv.load(g)[0] for g, _, _ in work]
        assert srv.drain(60_000)
        srv.inject_fault(gids[0], 2)
        t_bad = srv.update(gids[0], *work[0][1][0])
        t_ok = [srv.update(gids[1], *bs) for bs in work[1][1]]
        assert srv.wait(t_bad)["status"] == -6
        for t in t_ok:
            assert srv.result(t)["status"] == 0
        t2 = srv.update(gids[0], *work[0][1][1])
        assert srv.wait(t2)["status"] == -5
        tl, buf = srv.labels(gids[1], work[1][0].num_nodes)
        srv.result(tl)
        assert np.array_equal(buf, oracle.scc_canonical(work[1][2], 8))
        srv.result(srv.unload(gids[0]))
        with pytest.raises(SccError):
            srv.update(gids[0], [0], [1])


def test_cancel_queued_and_growth_beyond_headroom():
    g = graphs.rmat(12, 4, seed=7)
    with SccServer(slots=1, growth=0.0) as srv:
        gid, t = srv.load(g)
        rng = np.random.default_rng(0)
        ins = rng.integers(0, g.num_nodes, (3000, 2)).astype(np.int32)
        ins = ins[ins[:, 0] != ins[:, 1]]
        t1 = srv.update(gid, ins[:, 0], ins[:, 1])
        t2 = srv.update(gid, [0], [1])
        srv.cancel(t2)
        assert srv.wait(t2)["status"] == -7
        r1 = srv.result(t1)
        assert r1["inserted"] > 0
        tl, buf = srv.labels(gid, g.num_nodes)
        srv.result(tl)
        src, dst = g.edges()
        want = graphs.from_edges(g.num_nodes, np.concatenate([src, ins[:, 0]]),
                                 np.concatenate([dst, ins[:, 1]]), dedup=True)
        assert np.array_equal(buf, oracle.scc_canonical(want, 8))


def test_poisson_stream_statistics():
    g = graphs.rmat(14, 8, seed=11)
    us = graphs.UpdateStream(g, seed=12, lo=100, hi=5000)
    batches = [us.next() for _ in range(40)]
    final = us.graph()
    us.close()
    rng = np.random.default_rng(13)
    arrivals = np.cumsum(rng.exponential(2.0, len(batches)))  # mean 2 ms apart
    with SccServer(slots=2) as srv:
        gid, t = srv.load(g)
        srv.result(t)
        srv.reset_stats()
        t0 = srv.now_ms()
        for (a, b, c, d), at in zip(batches, arrivals):
            while srv.now_ms() < t0 + at:
                pass
            srv.update(gid, a, b, c, d, arrival_ms=t0 + at)
        assert srv.drain(120_000)
        st = srv.stats()
        print("  stats", st, flush=True)
        assert st["updates_done"] == 40 and st["failed"] == 0
        assert st["update_p50_ms"] <= st["update_p99_ms"] and st["updates_per_s"] > 0
        tl, buf = srv.labels(gid, g.num_nodes)
        srv.result(tl)
    assert np.array_equal(buf, oracle.scc_canonical(final, 8))


def test_degree_layout_server_equals_input_layout():
    """Graphs loaded with the degree layout give the same labels after every
    update as the input layout alone (caller ids throughout)."""
    from dynamicgraphs_scc_amd.engine import LAYOUT_DEGREE
    work = _workload(3, 4)
    alone = [_alone(g, bs) for g, bs, _ in work]
    with SccServer(slots=2, layout=LAYOUT_DEGREE) as srv:
        for i, (g, bs, _) in enumerate(work):
            gid, t = srv.load(g)
            tickets = [srv.labels(gid, g.num_nodes)]
            for a, b, c, d in bs:
                srv.update(gid, a, b, c, d)
                tickets.append(srv.labels(gid, g.num_nodes))
            for j, (t, buf) in enumerate(tickets):
                assert srv.result(t)["status"] == 0
                assert np.array_equal(buf, alone[i][j]), f"graph {i} step {j}"

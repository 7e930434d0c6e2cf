"""Benchmark: SCC decomposition of directed R-MAT scale 22 on one MI355X.

One step = one full pass of the hot path over the graph already resident in
HBM: state reset, trim1 -> global forward-backward -> trim1/trim2/trim1 -> WCC
-> partition forward-backward rounds -> canonical labels in the caller's vertex
ids, on the device (the reference's run() window, src/common_main.h:801-807,
which ends with G_SCC indexed by the caller's vertices, src/scc.h:12).

Prints one JSON line.  value = SCC GTEPS = M / (seconds per step) / 1e9.
The graph is uploaded once, outside the timed region, into a context with the
degree vertex layout (relabelled by decreasing degree inside HBM; every result
is returned in the caller's ids); --layout input keeps the caller's ids.
roofline: the global forward-backward traversal (the reference's 76 % phase),
achieved = B_FB / (summed device time of its traversal kernels, HIP events on
the engine's stream), with B_FB = 20|F| + 8E_F + 20|B| + 8E_B + 4|F n B|.
cpu_baseline: the reference's own method 2 (its phase sources, timed by
oracle/_ref/ref_bench) when it was built, else the oracle's OpenMP restatement,
on this host's physical cores: median of 5 runs after a warm-up, a 1-thread
figure, the CPU model, one run on the planted graph of 8,388,608 vertices, and
R-MAT scale 16 (the reference's CPU-runnable case) timed on the CPU and the GPU.

dynamic: the serving workload (R-MAT scale 23 base resident in HBM, 2,000
insert/delete batches with sizes log-uniform in [1k, 100k], half uniform
inserts and half uniform deletes of present edges, Poisson arrivals at a fixed
rate) through the request server; latency = arrival to labels final; labels
read back every 100 batches and checked against a from-scratch decomposition
of the same edge set.  Then 500 batches of exactly 10k edges for the
10k-batch latency figure.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info() -> dict:
    """CPU model, the CPUs this process may run on, and the physical cores among
    them (distinct (package, core) pairs in sysfs)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                cores.add((pkg, f.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": max(1, len(cores))}


def graphs_mod():
    from dynamicgraphs_scc_amd import graphs
    return graphs


def gpu_decomposition_ms(g, reps: int = 6):
    """Median wall time of one synchronised method-2 run on the GPU (degree
    layout, graph resident), after a warm-up; and the caller-id labels."""
    from dynamicgraphs_scc_amd.engine import LAYOUT_DEGREE, SccEngine
    ts = []
    with SccEngine(device=0) as e:
        e.set_layout(LAYOUT_DEGREE)
        e.upload(g)
        for _ in range(reps):
            e.reset()
            e.synchronize()
            t0 = time.perf_counter()
            e.run_method2()
            e.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        lab = e.labels()
    return statistics.median(ts[1:]), lab


def cpu_baseline(g, gpu_labels, gp=None, gp_gpu_ms=None, want_ref: bool = True):
    """The reference's own method 2 on this host's cores, same graph: one load,
    one warm-up and five timed runs (median) on every physical core available
    to the process (at most 16, the host share of one GPU), two more at one
    thread; the labels of the last run are compared with the GPU's.  Then one
    run on the planted graph (n = 8,388,608), and the reference's own CPU case
    (R-MAT scale 16, seed 1) timed on both sides.  Falls back to the oracle's
    restatement when the reference was not built."""
    import numpy as np
    import oracle
    hi = host_info()
    threads = min(hi["physical_cores"], 16)
    res = {}
    if want_ref and oracle.have_ref_bench():
        td = tempfile.mkdtemp(prefix="sccref_")
        try:
            path = os.path.join(td, "g.gm")
            t = time.time()
            g.write_gm_binary(path)
            r = oracle.run_ref_bench(path, threads, 6, labels_out=os.path.join(td, "lab.bin"))
            runs = r["ms"][1:]
            med = statistics.median(runs)
            same = bool(np.array_equal(oracle.canonicalize(r["labels"]), gpu_labels))
            log(f"cpu_baseline: reference {threads} threads {', '.join(f'{x:.1f}' for x in r['ms'])} ms "
                f"({time.time() - t:.1f}s), labels equal to the GPU's: {same}")
            t = time.time()
            r1 = oracle.run_ref_bench(path, 1, 3)
            ms1 = statistics.median(r1["ms"][1:])
            log(f"cpu_baseline: reference 1 thread {', '.join(f'{x:.1f}' for x in r1['ms'])} ms "
                f"({time.time() - t:.1f}s)")
            res = {"value": g.num_edges / (med / 1e3) / 1e9, "unit": "GTEPS", "cores": threads,
                   "kind": "reference",
                   "sample": (f"the reference's method 2 (its phase sources, oracle/_ref/ref_bench) on the "
                              f"same graph (n={g.num_nodes}, m={g.num_edges}), loaded once; run() window; median of {len(runs)} runs "
                              f"after one warm-up at {threads} threads: "
                              f"{', '.join(f'{x:.1f}' for x in runs)} ms; 1 thread: median of "
                              f"{len(r1['ms']) - 1} after a warm-up"),
                   "median_ms": med, "runs_ms": [round(x, 3) for x in runs],
                   "ms_1thread": ms1, "gteps_1thread": g.num_edges / (ms1 / 1e3) / 1e9,
                   "labels_equal_gpu": same}
            if gp is not None:
                t = time.time()
                pp = os.path.join(td, "p.gm")
                gp.write_gm_binary(pp)
                rp = oracle.run_ref_bench(pp, threads, 1, timeout=900)
                res["planted8m"] = {"n": gp.num_nodes, "m": gp.num_edges, "threads": threads,
                                    "ms": rp["ms"][0], "gteps": gp.num_edges / (rp["ms"][0] / 1e3) / 1e9,
                                    "sccs": rp["sccs"][0], "runs": 1,
                                    "gpu_ms_plain": gp_gpu_ms,
                                    "gpu_speedup": (rp["ms"][0] / gp_gpu_ms) if gp_gpu_ms else None}
                log(f"cpu_baseline: planted 8M reference {rp['ms'][0]:.1f} ms ({time.time() - t:.1f}s)")
            # the reference's own CPU-runnable case: R-MAT scale 16, seed 1
            t = time.time()
            g16 = graphs_mod().rmat(16, 16, seed=1)
            gpu16, lab16 = gpu_decomposition_ms(g16)
            p16 = os.path.join(td, "r16.gm")
            g16.write_gm_binary(p16)
            r16 = oracle.run_ref_bench(p16, threads, 6, labels_out=os.path.join(td, "lab16.bin"))
            r16_1 = oracle.run_ref_bench(p16, 1, 6)
            m16 = statistics.median(r16["ms"][1:])
            m16_1 = statistics.median(r16_1["ms"][1:])
            res["rmat16"] = {"n": g16.num_nodes, "m": g16.num_edges, "threads": threads,
                             "ms": round(m16, 3), "ms_1thread": round(m16_1, 3),
                             "gteps": g16.num_edges / (m16 / 1e3) / 1e9,
                             "gpu_ms": round(gpu16, 4), "gpu_speedup": m16 / gpu16,
                             "labels_equal_gpu": bool(np.array_equal(oracle.canonicalize(r16["labels"]), lab16)),
                             "sample": "median of 5 runs after a warm-up, each side (GPU: one synchronised run each)"}
            log(f"cpu_baseline: R-MAT-16 reference {m16:.2f} ms ({threads} threads), {m16_1:.2f} ms (1 thread), "
                f"GPU {gpu16:.3f} ms ({time.time() - t:.1f}s)")
        except Exception as ex:  # fall back to the restatement
            log(f"cpu_baseline: reference harness failed ({ex}); using the restatement")
            res = {}
        finally:
            shutil.rmtree(td, ignore_errors=True)
    port_times = []
    for rep in range(6):
        _, ms, _ = oracle.method2(g, threads=threads)
        port_times.append(ms)
    port_ms = statistics.median(port_times[1:])
    if not res:
        res = {"value": g.num_edges / (port_ms / 1e3) / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
               "sample": (f"oracle OpenMP restatement of method 2 on the same graph, median of 5 runs after "
                          f"a warm-up at {threads} threads: {port_ms:.1f} ms"),
               "median_ms": port_ms}
    res.update({"model": hi["model"], "physical_cores": hi["physical_cores"],
                "logical_cpus": hi["logical_cpus"],
                "port_gteps": g.num_edges / (port_ms / 1e3) / 1e9, "port_ms": port_ms})
    return res


def learned_leg(args, eng, g, ms_plain_rmat, gp):
    """Plain vs learned scheduler on R-MAT-22 (the engine's graph, same step as
    `value`) and on the planted-SCC graph gp (8,388,608 vertices, seed 3): the
    phase head (SCC_PARTS_LEARNED, plain pivots) and the phase head with the
    pivot head (SCC_PIVOT_LEARNED ranks the pivots of the forward-backward steps
    it chooses).  Plain is the default schedule (the WCC phase, then colouring
    steps); the fixed alternation of forward-backward and colouring steps
    (SCC_PARTS_BOTH) is timed beside it.  Labels must equal the plain run's."""
    import numpy as np
    from dynamicgraphs_scc_amd.engine import (PARTS_BOTH, PARTS_LEARNED, PIVOT_DEGREE_PRODUCT, PIVOT_LEARNED,
                                              SccEngine, default_opts)
    from dynamicgraphs_scc_amd.sched import load_weights

    w = load_weights()
    if w is None:
        return None

    def timed(e, opts, reps):
        ts = []
        for _ in range(reps):
            e.reset()
            e.synchronize()
            t0 = time.perf_counter()
            e.run_method2(opts)
            e.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    variants = {"phase": default_opts(partition_mode=PARTS_LEARNED, pivot_policy=PIVOT_DEGREE_PRODUCT),
                "both": default_opts(partition_mode=PARTS_LEARNED, pivot_policy=PIVOT_LEARNED)}

    def heads(e, m, reps):
        # plain and learned runs timed alike: one run per timed call, each
        # between synchronisations (the headline's back-to-back steps overlap
        # one step's launches with the previous one's tail)
        po = default_opts(pivot_policy=PIVOT_DEGREE_PRODUCT)
        timed(e, po, 1)
        ms_plain = timed(e, po, reps)
        lab_plain = e.labels()
        ms_alt = timed(e, default_opts(partition_mode=PARTS_BOTH), reps)
        e.set_sched_weights(w)
        out = {"ms_plain": round(ms_plain, 4), "ms_alternating": round(ms_alt, 4)}
        for k, o in variants.items():
            timed(e, o, 1)
            ms = timed(e, o, reps)
            out[f"ms_{k}"] = round(ms, 4)
            out[f"speedup_{k}"] = round(ms_plain / ms, 3)
            out[f"labels_equal_{k}"] = bool(np.array_equal(lab_plain, e.labels()))
        e.set_sched_weights(None)
        best = min(variants, key=lambda k: out[f"ms_{k}"])
        out["ms_learned"], out["speedup"] = out[f"ms_{best}"], out[f"speedup_{best}"]
        out["learned_heads"] = best
        out["labels_equal"] = all(out[f"labels_equal_{k}"] for k in variants)
        out["gteps_plain"] = round(m / (ms_plain / 1e3) / 1e9, 4)
        out["learned_gteps"] = round(m / (out["ms_learned"] / 1e3) / 1e9, 4)
        out["speedup_vs_alternating"] = round(ms_alt / out["ms_learned"], 3)
        return out

    res = {"rmat22": heads(eng, g.num_edges, max(5, args.steps))}
    res["rmat22"]["ms_headline_step"] = round(ms_plain_rmat, 4)
    with SccEngine(device=0) as e:
        e.upload(gp)
        res["planted8m"] = heads(e, gp.num_edges, 5)
        res["planted8m"].update({"n": gp.num_nodes, "m": gp.num_edges})
    return res


def batched_leg(args):
    """The batched workload (512 graphs, vertex counts log-uniform in [1e3, 1e6],
    half R-MAT, half planted) as one disjoint union decomposed by shared
    launches: default launch shape, then after the autotuner."""
    import numpy as np
    from dynamicgraphs_scc_amd import graphs
    from dynamicgraphs_scc_amd.engine import SccEngine

    t = time.time()
    gs = graphs.batched_config()
    u, off = graphs.disjoint_union(gs)
    del gs
    log(f"batched: {len(off) - 1} graphs, n={u.num_nodes} m={u.num_edges} in {time.time() - t:.1f}s")

    from dynamicgraphs_scc_amd.engine import PARTS_BOTH, PARTS_LEARNED, default_opts
    from dynamicgraphs_scc_amd.sched import load_weights

    def timed(e, reps, **kw):
        ts = []
        for _ in range(reps):
            e.synchronize()
            t0 = time.perf_counter()
            e.batch_run(default_opts(**kw))
            e.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    weights = load_weights()
    with SccEngine(device=0) as e:
        e.upload(u)
        e.batch_set(off)
        timed(e, 1)
        ms_default = timed(e, 5)
        counts = e.batch_counts()
        lab_plain = e.labels()
        ms_alt = timed(e, 3, partition_mode=PARTS_BOTH)
        ms_learned = same = None
        if weights is not None:
            e.set_sched_weights(weights)
            timed(e, 1, partition_mode=PARTS_LEARNED)
            ms_learned = timed(e, 5, partition_mode=PARTS_LEARNED)
            same = bool(np.array_equal(lab_plain, e.labels()))
            e.set_sched_weights(None)
        del lab_plain
        tune = e.autotune(reps=2)
        timed(e, 1)
        ms_tuned = timed(e, 5)
    out = {"workload": "512 graphs, N log-uniform [1e3,1e6], even R-MAT EF U[8,32], odd planted, seeds 4000+g",
           "graphs": int(len(off) - 1), "n": int(u.num_nodes), "m": int(u.num_edges),
           "num_sccs": int(counts.sum()),
           "ms_default": round(ms_default, 3), "gteps_default": round(u.num_edges / ms_default / 1e6, 4),
           "ms_alternating": round(ms_alt, 3),
           "ms_tuned": round(ms_tuned, 3), "gteps_tuned": round(u.num_edges / ms_tuned / 1e6, 4),
           "tuning": tune["tuning"]}
    if ms_learned is not None:
        out.update({"ms_learned": round(ms_learned, 3), "gteps_learned": round(u.num_edges / ms_learned / 1e6, 4),
                    "learned_speedup": round(ms_default / ms_learned, 3), "learned_labels_equal": same})
    return out


def _sleep_until(srv, t_ms):
    while True:
        d = t_ms - srv.now_ms()
        if d <= 0:
            return
        if d > 2.0:
            time.sleep((d - 1.0) / 1e3)


def _stream(srv, gid, n, batches, rate_per_s, seed, check_at=()):
    """Submit batches at Poisson arrival times; LABELS right after check points."""
    import numpy as np
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1e3 / rate_per_s, len(batches)))
    srv.reset_stats()
    t0 = srv.now_ms() + 20.0
    reads = {}
    for i, (b, at) in enumerate(zip(batches, arr)):
        _sleep_until(srv, t0 + at)
        srv.update(gid, *b, arrival_ms=t0 + float(at))
        if i in check_at:
            reads[i] = srv.labels(gid, n)
        if (i + 1) % 250 == 0:
            log(f"  dynamic: submitted {i + 1}/{len(batches)}")
    if not srv.drain(600_000):
        raise RuntimeError("update stream did not drain")
    return srv.stats(), reads


def dynamic_leg(args):
    import hashlib

    from dynamicgraphs_scc_amd import graphs
    from dynamicgraphs_scc_amd.engine import SccEngine
    from dynamicgraphs_scc_amd.serve import SccServer

    t = time.time()
    base = graphs.rmat(args.dyn_scale, 16, seed=23)
    n = base.num_nodes
    log(f"dynamic: base R-MAT-{args.dyn_scale} n={n} m={base.num_edges} in {time.time() - t:.1f}s")
    us = graphs.UpdateStream(base, seed=5, lo=1000, hi=100_000)
    batches, want = [], {}
    ver = SccEngine(device=0)
    t = time.time()
    for i in range(args.dyn_batches):
        batches.append(us.next())
        if (i + 1) % args.dyn_check_every == 0:
            gi = us.graph()
            ver.upload(gi)
            ver.run_method2()
            want[i] = hashlib.sha256(ver.labels().tobytes()).hexdigest()
            log(f"  dynamic: reference labels after batch {i + 1} ({time.time() - t:.1f}s)")
    ver.close()
    us.set_sizes(10_000, 10_000)
    fixed = [us.next() for _ in range(args.dyn_fixed_batches)]
    us.close()
    n_upd = sum(len(b[0]) + len(b[2]) for b in batches)
    log(f"dynamic: {len(batches)} batches ({n_upd} edges) generated in {time.time() - t:.1f}s")

    from dynamicgraphs_scc_amd.engine import LAYOUT_DEGREE, LAYOUT_INPUT
    layout = LAYOUT_DEGREE if args.dyn_layout == "degree" else LAYOUT_INPUT
    with SccServer(device=0, slots=2, max_batch=100_000, layout=layout) as srv:
        gid, tk = srv.load(base)
        srv.result(tk)
        del base
        st, reads = _stream(srv, gid, n, batches, args.dyn_rate, seed=6, check_at=set(want))
        ok = 0
        for i, (tk, buf) in reads.items():
            srv.result(tk)
            ok += hashlib.sha256(buf.tobytes()).hexdigest() == want[i]
        log(f"dynamic: p50 {st['update_p50_ms']:.2f} p99 {st['update_p99_ms']:.2f} ms, "
            f"{st['updates_per_s'] / 1e6:.2f} M updates/s, checks {ok}/{len(want)}")
        st10, _ = _stream(srv, gid, n, fixed, args.dyn_fixed_rate, seed=7)
        log(f"dynamic 10k: p50 {st10['update_p50_ms']:.2f} p99 {st10['update_p99_ms']:.2f} ms")
    return {
        "workload": f"rmat{args.dyn_scale}_ef16_seed23 base, {len(batches)} batches "
                    f"log-uniform [1k,100k] edges (stream seed 5), Poisson arrivals seed 6",
        "layout": args.dyn_layout,
        "rate_batches_per_s": args.dyn_rate,
        "offered_updates_per_s": round(n_upd / len(batches) * args.dyn_rate, 1),
        "p50_ms": round(st["update_p50_ms"], 3), "p99_ms": round(st["update_p99_ms"], 3),
        "max_ms": round(st["update_max_ms"], 3), "mean_ms": round(st["update_mean_ms"], 3),
        "updates_per_s": round(st["updates_per_s"], 1),
        "edges_applied": int(st["update_edges"]),
        "checks_passed": ok, "checks": len(want),
        "batch10k": {"batches": len(fixed), "rate_batches_per_s": args.dyn_fixed_rate,
                     "p50_ms": round(st10["update_p50_ms"], 3),
                     "p99_ms": round(st10["update_p99_ms"], 3),
                     "max_ms": round(st10["update_max_ms"], 3)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=float, default=16)
    ap.add_argument("--seed", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ref", action="store_true", help="CPU baseline from the restatement only")
    ap.add_argument("--policy", default="degree", choices=["degree", "minid"])
    ap.add_argument("--layout", default="degree", choices=["input", "degree"],
                    help="vertex layout inside the context (relabelling happens at upload, untimed)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (reference point for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE runs)")
    ap.add_argument("--no-dyn", action="store_true", help="skip the dynamic serving leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-scheduler leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-graphs leg")
    ap.add_argument("--dyn-scale", type=int, default=23)
    ap.add_argument("--dyn-batches", type=int, default=2000)
    ap.add_argument("--dyn-check-every", type=int, default=100)
    ap.add_argument("--dyn-rate", type=float, default=60.0, help="Poisson arrival rate, batches/s")
    ap.add_argument("--dyn-fixed-batches", type=int, default=500)
    ap.add_argument("--dyn-fixed-rate", type=float, default=100.0)
    ap.add_argument("--dyn-layout", default="degree", choices=["input", "degree"],
                    help="vertex layout of the served graph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is designed but not built)"}))
        sys.exit(2)

    import numpy as np
    from dynamicgraphs_scc_amd import graphs
    from dynamicgraphs_scc_amd.engine import (LAYOUT_DEGREE, LAYOUT_INPUT, PIVOT_DEGREE_PRODUCT,
                                              PIVOT_MIN_ID, SccEngine, default_opts)

    if args.copy_calibration_mib:
        import torch
        nb = args.copy_calibration_mib << 20
        a = torch.ones(nb // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {nb} bytes read + {nb} bytes written")

    t = time.time()
    g = graphs.rmat(args.scale, args.edge_factor, seed=args.seed)
    log(f"generated R-MAT-{args.scale}: n={g.num_nodes} m={g.num_edges} in {time.time() - t:.1f}s")
    policy = PIVOT_DEGREE_PRODUCT if args.policy == "degree" else PIVOT_MIN_ID

    eng = SccEngine(device=0)
    eng.set_layout(LAYOUT_DEGREE if args.layout == "degree" else LAYOUT_INPUT)
    t = time.perf_counter()
    eng.upload(g)
    upload_ms = (time.perf_counter() - t) * 1e3
    opts = default_opts(pivot_policy=policy)
    for _ in range(args.warmup):
        eng.reset()
        eng.run_method2(opts)
    eng.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.reset()
        eng.run_method2(opts)
    eng.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    gteps = g.num_edges / (ms_step / 1e3) / 1e9

    # device time of the global FB traversal kernels (HIP events on the engine's stream)
    eng.set_kernel_timing(True)
    kt, launches, st = [], [], None
    for _ in range(max(3, min(args.steps, 10))):
        eng.reset()
        st = eng.run_method2(default_opts(pivot_policy=policy, time_phases=1))
        kt.append(st["ms_fb_kernels"])
        launches.append(st["fb_launches"])
    eng.set_kernel_timing(False)
    labels = eng.labels()
    nsccs = int(st["num_sccs"])
    F, B, EF, EB, FB = st["fb_f"], st["fb_b"], st["fb_ef"], st["fb_eb"], st["fb_fb"]
    b_fb = 20 * F + 8 * EF + 20 * B + 8 * EB + 4 * FB
    t_fb_ms = statistics.median(kt)
    achieved = b_fb / (t_fb_ms / 1e3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "rmat22_fb_traffic.json")
    if os.path.exists(tfile) and args.scale == 22:
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_level2 (one level of the forward and one of the backward search of the global step per launch)",
                "algorithmic_bytes": int(b_fb),
                # per launch: all timed runs' kernel time over all their launches (a
                # step runs 6 or 7 level launches, the last sometimes finding nothing)
                "algorithmic_bytes_per_launch": int(b_fb * len(launches) / max(1, sum(launches))),
                "avg_launch_us": round(1e3 * sum(kt) / max(1, sum(launches)), 3),
                "launches_per_step": round(sum(launches) / max(1, len(launches)), 2),
                "kernel_ms_per_step": round(t_fb_ms, 4),
                "sets": {"F": F, "B": B, "E_F": EF, "E_B": EB, "FnB": FB}}
    phases = {k: round(st[k], 4) for k in ("ms_trim1", "ms_fb_global", "ms_trim2", "ms_wcc",
                                           "ms_fb_parts", "ms_canon")}
    log(f"step {ms_step:.3f} ms  GTEPS {gteps:.2f}  phases {phases}  fb_kernels {t_fb_ms:.3f} ms")

    gp = None
    if not (args.no_learned and args.no_cpu):
        t = time.time()
        gp = graphs.planted(8_388_608, seed=3)
        log(f"generated planted 8M: n={gp.num_nodes} m={gp.num_edges} in {time.time() - t:.1f}s")
    learned = None
    if not args.no_learned:
        t = time.time()
        learned = learned_leg(args, eng, g, ms_step, gp)
        log(f"learned leg in {time.time() - t:.1f}s: {learned}")
    eng.close()
    cpu = None
    if not args.no_cpu:
        t = time.time()
        gp_ms = learned["planted8m"]["ms_plain"] if learned else None
        cpu = cpu_baseline(g, labels, gp, gp_ms, want_ref=not args.no_ref)
        log(f"cpu baseline {cpu['kind']}: {cpu['value']:.3f} GTEPS in {time.time() - t:.1f}s")
    del gp

    batched = None
    if not args.no_batched:
        t = time.time()
        batched = batched_leg(args)
        log(f"batched leg in {time.time() - t:.1f}s: {batched}")

    dyn = None
    if not args.no_dyn:
        t = time.time()
        dyn = dynamic_leg(args)
        log(f"dynamic leg in {time.time() - t:.1f}s")

    out = {
        "metric": "SCC GTEPS at R-MAT-22, learned vs plain scheduler; p50/p99 dyn-update latency",
        "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded R-MAT, (a,b,c,d)=(0.57,0.19,0.19,0.05), dedup, permuted)",
        "config": {"workload": f"rmat{args.scale}_ef{int(args.edge_factor)}_seed{args.seed}",
                   "n": g.num_nodes, "m": g.num_edges, "scheduler": "plain (max in-deg x out-deg pivot)",
                   "layout": args.layout, "upload_ms": round(upload_ms, 1),
                   "num_sccs": nsccs, "giant_scc": int(np.bincount(labels).max())},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phases_ms": phases,
        "learned_gteps": learned["rmat22"]["learned_gteps"] if learned else None,
        "learned": learned,
        "batched": batched,
        "dynamic": dyn,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

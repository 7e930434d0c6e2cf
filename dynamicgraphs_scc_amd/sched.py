"""Training and evaluation of the learned pivot / phase scheduler.

The network and its kernels live in csrc/device/sched.hip (16 features ->
64 -> 64 -> 2 heads, BF16 matrix-core arithmetic, FP32 accumulation).  This
module drives them:

* traces: counterfactual decompositions of synthetic graphs
  (SccEngine.sched_rollout) -- at each partition round, the forward-backward
  step and the colouring step are each taken from the same state and followed
  by the continuation policy to the end; the trace records which finished
  sooner, and every sampled vertex's SCC size relative to its partition;
* training: Adam on the device (SccEngine.sched_train), from seeded He
  initial weights; the first round of traces continues with colouring
  steps (the plain rule), and further rounds (--iterations) continue with the network trained
  so far (policy iteration: its one-step improvements are then measured
  against its own later choices);
* evaluation: wall time of the plain method (degree-product pivots, the WCC
  phase, then colouring steps) against the learned policy on held-out graphs,
  labels compared for equality.

The trained weights ship as sched_weights.npy next to this file (FP32, 5378
values, loaded with allow_pickle=False).  `python -m dynamicgraphs_scc_amd.sched
--rollouts 4 --epochs 80 --seed 4 --graph-seed 4100 --out DIR` is how they were
made on a GPU (the labels come from wall times, so runs differ).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import graphs
from .engine import (PARTS_BOTH, PARTS_LEARNED, SCHED_FEATURES, SCHED_HEADS, SCHED_HIDDEN, SCHED_NPARAMS,
                     SccEngine, default_opts)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "sched_weights.npy")


def load_weights(path: str = WEIGHTS):
    """The shipped weights, or None when the file is absent."""
    if not os.path.exists(path):
        return None
    w = np.load(path, allow_pickle=False)
    if w.shape != (SCHED_NPARAMS,) or w.dtype != np.float32 or not np.all(np.isfinite(w)):
        raise ValueError(f"{path}: not a scheduler weight vector")
    return w


def initial_weights(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, in the kernel's parameter order."""
    rng = np.random.default_rng(seed)
    nf, nh, no = SCHED_FEATURES, SCHED_HIDDEN, SCHED_HEADS
    parts = [rng.normal(0, np.sqrt(2 / nf), nf * nh), np.zeros(nh),
             rng.normal(0, np.sqrt(2 / nh), nh * nh), np.zeros(nh),
             rng.normal(0, np.sqrt(1 / nh), nh * no), np.zeros(no)]
    return np.concatenate(parts).astype(np.float32)


def batch_of(k: int, seed: int, n_hi: int):
    """A disjoint union of the batched workload's kind (half R-MAT, half
    planted) with its vertex offsets, for batch_set."""
    u, off = graphs.disjoint_union(graphs.batched_config(k=k, seed=seed, n_hi=n_hi))
    return u, off


def training_graphs(seed: int = 100):
    """Synthetic graphs whose decompositions need partition rounds: planted SCC
    graphs over a range of sizes and medium-SCC counts, and two small batches of
    the batched workload's kind.  (R-MAT graphs alone leave no partition rounds
    after the global step, trims and WCC, so they give no samples.)"""
    specs = []
    rng = np.random.default_rng(seed)
    for i in range(10):
        n = int(rng.integers(100_000, 1_500_000))
        med = int(n * rng.uniform(0.005, 0.04))
        specs.append((f"planted_n{n}_m{med}_s{seed + i}",
                      lambda n=n, med=med, s=seed + i: graphs.planted(n, seed=s, n_medium=med)))
    for i in range(2):
        specs.append((f"batch48_s{5000 + 100 * seed + 50 * i}",
                      lambda s=5000 + 100 * seed + 50 * i: batch_of(48, s, 200_000)))
    return specs


def held_out_graphs():
    return [("batch128_s4000", lambda: batch_of(128, 4000, 1_000_000)),
            ("planted8m_s3", lambda: graphs.planted(8_388_608, seed=3)),
            ("planted1m_s4", lambda: graphs.planted(1_000_000, seed=4, n_medium=25_000)),
            ("rmat22_ef16_s22", lambda: graphs.rmat(22, 16, seed=22)),
            ("rmat20_ef3_s7", lambda: graphs.rmat(20, 3, seed=7))]


def generate(specs, rollouts: int = 2, explore: float = 0.25, samples_per_round: int = 512, log=print,
             weights=None):
    Xs, Ys, Ms = [], [], []
    with SccEngine() as e:
        for name, make in specs:
            g, off = _graph(make())
            e.upload(g)
            if off is not None:
                e.batch_set(off)
            e.set_sched_weights(weights)
            for r in range(rollouts):
                X, Y, M, st = e.sched_rollout(seed=1000 * r + 7, explore=explore,
                                              samples_per_round=samples_per_round)
                Xs.append(X)
                Ys.append(Y)
                Ms.append(M)
                log(f"  trace {name} #{r}: {X.shape[0]} samples, {st['rounds']} rounds, "
                    f"colour better in {Y[:, 1].mean() if len(Y) else 0:.2f}")
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)


def train(X, Y, M, epochs: int = 40, batch: int = 1024, lr: float = 2e-3, seed: int = 0, start=None):
    with SccEngine() as e:
        e.set_sched_weights(initial_weights(seed) if start is None else start)
        loss = e.sched_train(X, Y, M, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return e.sched_weights(), loss


def _graph(x):
    """(graph, batch offsets or None) from a spec's result."""
    return x if isinstance(x, tuple) else (x, None)


def time_runs(e, opts, reps: int, batch: bool = False):
    ts = []
    for _ in range(reps):
        e.reset()
        e.synchronize()
        t = time.perf_counter()
        if batch:
            e.batch_run(opts)
        else:
            e.run_method2(opts)
        e.synchronize()
        ts.append((time.perf_counter() - t) * 1e3)
    return float(np.median(ts))


def evaluate(weights, specs, reps: int = 5, log=print):
    """Median wall ms of the plain schedule (colouring steps), the
    forward-backward/colouring alternation and the learned scheduler; labels
    must be equal."""
    out = []
    with SccEngine() as e:
        for name, make in specs:
            g, off = _graph(make())
            e.upload(g)
            if off is not None:
                e.batch_set(off)
            b = off is not None
            plain_opts = default_opts()
            e.set_sched_weights(None)
            time_runs(e, plain_opts, 1, b)
            t_plain = time_runs(e, plain_opts, reps, b)
            lab_plain = e.labels()
            alt_opts = default_opts(partition_mode=PARTS_BOTH)
            time_runs(e, alt_opts, 1, b)
            t_alt = time_runs(e, alt_opts, reps, b)
            same_alt = bool(np.array_equal(lab_plain, e.labels()))
            e.set_sched_weights(weights)
            learned_opts = default_opts(partition_mode=PARTS_LEARNED)
            time_runs(e, learned_opts, 1, b)
            t_learned = time_runs(e, learned_opts, reps, b)
            same = same_alt and bool(np.array_equal(lab_plain, e.labels()))
            e.set_sched_weights(None)
            rec = {"graph": name, "n": g.num_nodes, "m": g.num_edges, "ms_plain": round(t_plain, 3),
                   "ms_alternating": round(t_alt, 3), "ms_learned": round(t_learned, 3),
                   "speedup": round(t_plain / t_learned, 3),
                   "speedup_vs_alternating": round(t_alt / t_learned, 3), "labels_equal": same}
            log(f"  eval {rec}")
            out.append(rec)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=".", help="directory for sched_weights.npy and sched_report.json")
    ap.add_argument("--rollouts", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--iterations", type=int, default=1,
                    help="rounds of traces + training (policy iteration beyond 1)")
    ap.add_argument("--graph-seed", type=int, default=100, help="seed of the training graphs")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, flush=True)

    w = None
    for it in range(args.iterations):
        # each round learns from traces of its own continuation policy only
        t = time.time()
        X, Y, M = generate(training_graphs(args.graph_seed + 10 * it), rollouts=args.rollouts, log=log, weights=w)
        log(f"iteration {it}: {X.shape[0]} samples in {time.time() - t:.1f}s")
        w, loss = train(X, Y, M, epochs=args.epochs, seed=args.seed, start=w)
        log(f"iteration {it}: last batch loss {loss:.4f}")
    os.makedirs(args.out, exist_ok=True)
    np.save(os.path.join(args.out, "sched_weights.npy"), w, allow_pickle=False)
    ev = evaluate(w, held_out_graphs(), log=log)
    with open(os.path.join(args.out, "sched_report.json"), "w") as f:
        json.dump({"samples": int(X.shape[0]), "loss": loss, "eval": ev}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

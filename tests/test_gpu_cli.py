"""The `scc` command-line driver keeps the reference's CLI and stdout contract
(src/common_main.h:451-807, src/scc_main.cc:147-178) and its results match the
reference's own scc_list.txt (golden fixtures)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from helpers import load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
SCC = os.path.join(ROOT, "dynamicgraphs_scc_amd", "lib", "scc")


def run_scc(args, cwd):
    print("  scc", args, flush=True)
    r = subprocess.run([SCC] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, f"scc {args} exited {r.returncode}: {r.stdout[-2000:]} {r.stderr[-2000:]}"
    return r.stdout


def parse(out):
    """The fields exec_auto.py:10-33 scrapes."""
    info = {}
    m = re.search(r"data=(\S+) (\d+) (-?\d+)", out)
    if m:
        info.update(file=m.group(1), method=int(m.group(2)), threads=int(m.group(3)))
    m = re.search(r"running_time\(ms\)=([0-9.]+)", out)
    if m:
        info["runtime"] = float(m.group(1))
    m = re.search(r"Total # SCCs = (\d+)", out)
    if m:
        info["scc_count"] = int(m.group(1))
    return info


@pytest.mark.parametrize("name", ["rmat12_s13", "planted20k_s3", "pairs_loops_dups", "rmat12_ef3_s14"])
def test_cli_method2_matches_reference_scc_list(tmp_path, name):
    g, canon, count = load_golden(name)
    path = tmp_path / "edges.txt"
    g.write_edgelist(str(path))
    out = run_scc([path, 4, 2, "-p"], tmp_path)
    info = parse(out)
    assert info["method"] == 2 and info["threads"] == 4 and info["scc_count"] == count
    assert "runtime" in info and "graph loading time=" in out and "reverse edge creation time=" in out
    dense = np.loadtxt(tmp_path / "scc_list.txt", dtype=np.int64, ndmin=2)
    assert np.array_equal(dense[:, 0], np.arange(g.num_nodes))
    assert np.array_equal(oracle.dense_to_canonical(dense[:, 1].astype(np.int32)), canon)


def test_cli_detail_histogram_and_methods(tmp_path):
    g = graphs.rmat(12, 4, seed=3)
    path = tmp_path / "edges.txt"
    g.write_edgelist(str(path))
    n_ref = int(max(g.edges()[0].max(), g.edges()[1].max())) + 1
    want = oracle.scc_canonical(graphs.from_edges(n_ref, *g.edges()))
    out = run_scc([path, 1, 2, "-d"], tmp_path)
    assert "[TRIM1 phase:" in out and "[GLOBAL_BFS phase:" in out
    out = run_scc([path, 1, 2, "-a"], tmp_path)
    sizes = np.bincount(np.bincount(want)[np.unique(want)])
    hist = dict((int(a), int(b)) for a, b in re.findall(r"^(\d+) => (\d+)$", out, re.M))
    assert hist == {s: int(c) for s, c in enumerate(sizes) if c}
    for method in (0, 1, 3):
        assert parse(run_scc([path, 1, method], tmp_path))["scc_count"] == len(np.unique(want))


def test_cli_update_methods(tmp_path):
    base = graphs.rmat(11, 6, seed=9)
    base.write_edgelist(str(tmp_path / "refined_edges.txt"))
    n_ref = int(max(base.edges()[0].max(), base.edges()[1].max())) + 1
    src, dst = base.edges()
    rng = np.random.default_rng(1)
    ins = rng.integers(0, n_ref, (300, 2))
    ins = ins[ins[:, 0] != ins[:, 1]]
    with open(tmp_path / "insert.txt", "w") as f:
        for a, b in ins:
            f.write(f"{a + 1} {b + 1}\n")
    out = run_scc([tmp_path / "insert.txt", 2, 6], tmp_path)
    g2 = graphs.from_edges(n_ref, np.concatenate([src, ins[:, 0]]), np.concatenate([dst, ins[:, 1]]))
    assert parse(out)["scc_count"] == len(np.unique(oracle.scc_canonical(g2)))
    sel = rng.choice(len(src), 400, replace=False)
    with open(tmp_path / "delete.txt", "w") as f:
        for i in sel:
            f.write(f"{src[i] + 1} {dst[i] + 1}\n")
    out = run_scc([tmp_path / "delete.txt", 2, 9], tmp_path)
    keep = np.ones(len(src), bool)
    keep[sel] = False
    g3 = graphs.from_edges(n_ref, src[keep], dst[keep])
    assert parse(out)["scc_count"] == len(np.unique(oracle.scc_canonical(g3)))


@pytest.fixture(scope="module")
def shim_driver(tmp_path_factory):
    """tests/shim_driver.cpp (the reference driver's scc.h call sequence) built
    against include/scc.h and libsccshim.so (host g++, no device code)."""
    lib = os.path.join(ROOT, "dynamicgraphs_scc_amd", "lib")
    exe = str(tmp_path_factory.mktemp("drv") / "shim_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "shim_driver.cpp"), "-o", exe, "-L", lib, "-lsccshim",
                    "-lsccmi", "-lsccgen", f"-Wl,-rpath,{lib}"], check=True)
    return exe


def _run_driver(exe, args, cwd):
    print("  shim_driver", args, flush=True)
    r = subprocess.run([exe] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"{args}: {r.stdout[-2000:]} {r.stderr[-2000:]}"
    return r.stdout


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_reference_style_driver_methods(tmp_path, shim_driver, method):
    """Methods 0-4 as the reference's driver sequences its scc.h calls: the SCC
    count (G_SCC[i] == i) and scc_list.txt equal the oracle's partition."""
    g = graphs.planted(20000, seed=11, n_medium=300) if method == 2 else graphs.rmat(12, 5, seed=method + 20)
    path = tmp_path / "edges.txt"
    g.write_edgelist(str(path))
    src, dst = g.edges()
    n_ref = int(max(src.max(), dst.max())) + 1
    want = oracle.scc_canonical(graphs.from_edges(n_ref, src, dst))
    out = _run_driver(shim_driver, [path, method], tmp_path)
    assert parse(out)["scc_count"] == len(np.unique(want))
    if method == 2:
        assert re.search(r"#queue_size = \d+", out)
    dense = np.loadtxt(tmp_path / "scc_list.txt", dtype=np.int64, ndmin=2)
    assert np.array_equal(oracle.dense_to_canonical(dense[:, 1].astype(np.int32)), want)


@pytest.mark.parametrize("mode", ["ins", "del"])
def test_reference_style_driver_update_graphs(tmp_path, shim_driver, mode):
    """insert_idea1 / delete_idea1 build the updated graph as the reference's
    update methods do (every SCC group's kept edges, not group 5 alone); the
    decomposition of it equals the oracle's."""
    base = graphs.rmat(11, 6, seed=17)
    path = tmp_path / "edges.txt"
    base.write_edgelist(str(path))
    src, dst = base.edges()
    n_ref = int(max(src.max(), dst.max())) + 1
    rng = np.random.default_rng(5)
    if mode == "ins":
        e = rng.integers(0, n_ref, (400, 2))
        s2, d2 = np.concatenate([src, e[:, 0]]), np.concatenate([dst, e[:, 1]])
    else:
        sel = rng.choice(len(src), 500, replace=False)
        e = np.stack([src[sel], dst[sel]], 1)
        keys = src.astype(np.int64) * n_ref + dst
        keep = ~np.isin(keys, e[:, 0].astype(np.int64) * n_ref + e[:, 1])
        s2, d2 = src[keep], dst[keep]
    with open(tmp_path / "batch.txt", "w") as f:
        for a, b in e:
            f.write(f"{a + 1} {b + 1}\n")
    out = _run_driver(shim_driver, [path, mode, tmp_path / "batch.txt"], tmp_path)
    want = oracle.scc_canonical(graphs.from_edges(n_ref, s2, d2))
    assert parse(out)["scc_count"] == len(np.unique(want))
    dense = np.loadtxt(tmp_path / "scc_list.txt", dtype=np.int64, ndmin=2)
    assert np.array_equal(oracle.dense_to_canonical(dense[:, 1].astype(np.int32)), want)

"""scc_run_device_noalloc, the decomposition that allocates nothing (its scratch
is one static device array), called through the C ABI on graphs the test puts
in HBM itself.  Two calls in a row, the second on a smaller graph (the array
keeps the first call's contents), each equal to the context path
(scc_graph_upload + scc_run_method2 + canonical labels) and to the oracle.
Sizes above the limit are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import oracle
from dynamicgraphs_scc_amd import graphs
from dynamicgraphs_scc_amd.engine import SccEngine, load_library

pytestmark = pytest.mark.gpu

MAX_N = 262144  # SCC_NOALLOC_MAX_N in include/sccmi.h


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libsccmi.so links (the
    test allocates the inputs and outputs; the entry point allocates nothing)."""

    def __init__(self):
        load_library()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.int32)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(4, a.nbytes)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def empty(self, n):
        return self.put(np.full(n, -7, np.int32))

    def get(self, p, n):
        out = np.empty(n, np.int32)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(lib, hip, g):
    n = g.num_nodes
    b, a, rb, ra = (hip.put(x) for x in (g.begin, g.node_idx, g.r_begin, g.r_node_idx))
    out = hip.empty(n)
    cnt = C.c_int64(-1)
    err = C.create_string_buffer(256)
    print(f"  scc_run_device_noalloc n={n} m={g.num_edges}", flush=True)
    rc = lib.scc_run_device_noalloc(0, None, n, g.num_edges, b, a, rb, ra, out, C.byref(cnt), err, 256)
    assert rc == 0, err.value.decode()
    return hip.get(out, n), cnt.value


def _context(g):
    with SccEngine() as e:
        e.upload(g)
        st = e.run_method2()
        return e.labels(), st["num_sccs"]


def test_noalloc_equals_context_path_twice_second_smaller():
    lib = load_library()
    hip = _Hip()
    try:
        first = graphs.planted(200_000, seed=41, n_medium=3000)
        second = graphs.rmat(14, 8, seed=42)
        assert first.num_nodes <= MAX_N and second.num_nodes < first.num_nodes
        for g in (first, second):
            got, cnt = _noalloc(lib, hip, g)
            want, want_cnt = _context(g)
            assert np.array_equal(got, want)
            assert cnt == want_cnt == len(np.unique(want))
            assert np.array_equal(got, oracle.scc_canonical(g, 8))
    finally:
        hip.free()


def test_noalloc_refuses_beyond_the_limit():
    lib = load_library()
    err = C.create_string_buffer(256)
    cnt = C.c_int64(0)
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib.scc_run_device_noalloc(0, None, MAX_N + 1, 10, dummy, dummy, dummy, dummy, dummy, C.byref(cnt),
                                    err, 256)
    assert rc == -8 and b"no-alloc limits" in err.value

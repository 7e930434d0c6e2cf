"""Request loop over device-resident graphs (ctypes over libsccserve.so,
include/sccserve.h): continuous batching across worker slots, per-graph order,
admission by HBM budget, p50/p99 update latency.

No reference counterpart: the reference applies one update file per process
run (src/common_main.h:519-725); here many graphs stay resident and update
batches stream in.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .engine import PIVOT_DEGREE_PRODUCT, SccError, load_library

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsccserve.so")

REQ_LOAD, REQ_UPDATE, REQ_LABELS, REQ_UNLOAD = 1, 2, 3, 4
DONE, QUEUED, RUNNING = 0, 1, 2

EXPORTS = [
    "scc_server_config_default", "scc_server_create", "scc_server_destroy", "scc_server_last_error",
    "scc_server_now_ms", "scc_server_load", "scc_server_update", "scc_server_labels",
    "scc_server_unload", "scc_server_wait", "scc_server_drain", "scc_server_cancel",
    "scc_server_stats_get", "scc_server_stats_reset", "scc_server_inject_fault",
]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("queue_limit", C.c_int32),
                ("pivot_policy", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("max_batch", C.c_int64), ("growth", C.c_double), ("layout", C.c_int32),
                ("coalesce", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("graph", C.c_int32),
                ("group", C.c_int32), ("num_sccs", C.c_int64), ("inserted", C.c_int64),
                ("deleted", C.c_int64), ("affected", C.c_int64), ("arrival_ms", C.c_double),
                ("start_ms", C.c_double), ("end_ms", C.c_double), ("ms_latency", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("queued", C.c_int64), ("running", C.c_int64), ("resident_graphs", C.c_int64),
                ("reserved_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("updates_done", C.c_int64), ("update_edges", C.c_int64),
                ("update_p50_ms", C.c_double), ("update_p99_ms", C.c_double),
                ("update_max_ms", C.c_double), ("update_mean_ms", C.c_double),
                ("busy_ms", C.c_double), ("updates_per_s", C.c_double),
                ("max_concurrency", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_serve_library():
    global _lib
    if _lib is not None:
        return _lib
    load_library()  # libsccmi.so first (the server links it)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "scc_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "scc_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(vp)]),
        "scc_server_destroy": (C.c_int, [vp]),
        "scc_server_last_error": (C.c_char_p, [vp]),
        "scc_server_now_ms": (d, [vp]),
        "scc_server_load": (C.c_int, [vp, i32, i64, vp, vp, vp, vp, C.POINTER(i32), C.POINTER(i64)]),
        "scc_server_update": (C.c_int, [vp, i32, vp, vp, i64, vp, vp, i64, d, C.POINTER(i64)]),
        "scc_server_labels": (C.c_int, [vp, i32, vp, C.POINTER(i64)]),
        "scc_server_unload": (C.c_int, [vp, i32, C.POINTER(i64)]),
        "scc_server_wait": (C.c_int, [vp, i64, d, C.POINTER(RequestInfo)]),
        "scc_server_drain": (C.c_int, [vp, d]),
        "scc_server_cancel": (C.c_int, [vp, i64]),
        "scc_server_stats_get": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "scc_server_stats_reset": (C.c_int, [vp]),
        "scc_server_inject_fault": (C.c_int, [vp, i32, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _i32(a):
    return np.ascontiguousarray(a if a is not None else np.empty(0, np.int32), dtype=np.int32)


class SccServer:
    """Python face of scc_server.  Tickets are ints; results are dicts."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget_bytes: int = 0,
                 max_batch: int = 100_000, growth: float = 0.25, queue_limit: int = 1 << 16,
                 pivot_policy: int = PIVOT_DEGREE_PRODUCT, layout: int = 0, coalesce: int = 32):
        self.lib = load_serve_library()
        cfg = ServerConfig()
        self.lib.scc_server_config_default(C.byref(cfg))
        cfg.device, cfg.slots, cfg.queue_limit = device, slots, queue_limit
        cfg.pivot_policy, cfg.hbm_budget_bytes = pivot_policy, hbm_budget_bytes
        cfg.max_batch, cfg.growth, cfg.layout = max_batch, growth, layout
        cfg.coalesce = coalesce
        h = C.c_void_p()
        rc = self.lib.scc_server_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise SccError(rc, self.lib.scc_server_last_error(None).decode())
        self.h = h
        self._keep = {}  # label buffers alive until their request ends

    def close(self):
        if getattr(self, "h", None):
            self.lib.scc_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise SccError(rc, self.lib.scc_server_last_error(self.h).decode())
        return rc

    def now_ms(self) -> float:
        return self.lib.scc_server_now_ms(self.h)

    def load(self, g) -> tuple[int, int]:
        """Submit a LOAD; returns (graph id, ticket)."""
        b, a, rb, ra = (np.ascontiguousarray(x, np.int32) for x in g.arrays())
        gid, t = C.c_int32(), C.c_int64()
        self._check(self.lib.scc_server_load(self.h, g.num_nodes, g.num_edges, b.ctypes.data,
                                             a.ctypes.data, rb.ctypes.data, ra.ctypes.data,
                                             C.byref(gid), C.byref(t)))
        return gid.value, t.value

    def update(self, graph, ins_src=None, ins_dst=None, del_src=None, del_dst=None,
               arrival_ms: float = -1.0) -> int:
        isrc, idst, dsrc, ddst = (_i32(x) for x in (ins_src, ins_dst, del_src, del_dst))
        if isrc.shape != idst.shape or dsrc.shape != ddst.shape:
            raise ValueError("src/dst length mismatch")
        t = C.c_int64()
        self._check(self.lib.scc_server_update(self.h, graph, isrc.ctypes.data, idst.ctypes.data,
                                               isrc.shape[0], dsrc.ctypes.data, ddst.ctypes.data,
                                               dsrc.shape[0], float(arrival_ms), C.byref(t)))
        return t.value

    def labels(self, graph, n) -> tuple[int, np.ndarray]:
        """Submit a LABELS request; the array is filled when the ticket ends."""
        out = np.empty(n, np.int32)
        t = C.c_int64()
        self._check(self.lib.scc_server_labels(self.h, graph, out.ctypes.data, C.byref(t)))
        self._keep[t.value] = out
        return t.value, out

    def unload(self, graph) -> int:
        t = C.c_int64()
        self._check(self.lib.scc_server_unload(self.h, graph, C.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms: float = -1.0) -> dict:
        info = RequestInfo()
        rc = self.lib.scc_server_wait(self.h, ticket, float(timeout_ms), C.byref(info))
        if rc not in (QUEUED, RUNNING):
            self._keep.pop(ticket, None)
        d = info.as_dict()
        d["status"] = rc
        return d

    def result(self, ticket) -> dict:
        """Wait for a ticket and raise if it failed."""
        d = self.wait(ticket)
        if d["status"] < 0:
            raise SccError(d["status"], self.lib.scc_server_last_error(self.h).decode())
        return d

    def drain(self, timeout_ms: float = -1.0) -> bool:
        return self.lib.scc_server_drain(self.h, float(timeout_ms)) == 0

    def cancel(self, ticket) -> None:
        self._check(self.lib.scc_server_cancel(self.h, ticket))

    def stats(self) -> dict:
        st = ServerStats()
        self._check(self.lib.scc_server_stats_get(self.h, C.byref(st)))
        return st.as_dict()

    def reset_stats(self) -> None:
        self._check(self.lib.scc_server_stats_reset(self.h))

    def inject_fault(self, graph, after: int) -> None:
        self._check(self.lib.scc_server_inject_fault(self.h, graph, after))

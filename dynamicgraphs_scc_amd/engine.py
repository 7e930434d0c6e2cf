"""ctypes binding of libsccmi.so (include/sccmi.h) and a thin Python mirror of
the reference's scc.h entry points over one device context.

The product path is the HIP library: if libsccmi.so is missing or no GPU is
present, constructing an ``SccEngine`` raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsccmi.so")

# scc_status
SCC_OK = 0
SCC_ERR_ARG = -1
SCC_ERR_HIP = -2
SCC_ERR_OOM = -3
SCC_ERR_BUDGET = -4
SCC_ERR_STATE = -5
SCC_ERR_DEVICE_LOST = -6
SCC_ERR_CANCELLED = -7
SCC_ERR_LIMIT = -8
STATUS_NAMES = {0: "OK", -1: "ARG", -2: "HIP", -3: "OOM", -4: "BUDGET", -5: "STATE",
                -6: "DEVICE_LOST", -7: "CANCELLED", -8: "LIMIT"}

# scc_partition_mode
PARTS_FB = 0
PARTS_COLOR = 1
PARTS_BOTH = 2
PARTS_LEARNED = 3

SCHED_FEATURES = 16
SCHED_HIDDEN = 64
SCHED_HEADS = 2
SCHED_NPARAMS = 5378

PHASE_NAMES = {1: "trim1", 2: "fb_global", 3: "trim2", 4: "wcc", 5: "fb_round", 6: "color_round",
               7: "canon"}

# scc_pivot_policy
PIVOT_DEGREE_PRODUCT = 0
PIVOT_MIN_ID = 1
PIVOT_LEARNED = 2
# vertex layout inside a context (scc_ctx_set_layout)
LAYOUT_INPUT = 0
LAYOUT_DEGREE = 1
LAYOUT_AUTO = 2

# every symbol include/sccmi.h declares
EXPORTS = [
    "scc_ctx_create", "scc_ctx_destroy", "scc_last_error", "scc_ctx_bytes", "scc_bytes_needed",
    "scc_ctx_cancel", "scc_ctx_inject_fault", "scc_graph_upload", "scc_graph_upload_device",
    "scc_graph_info", "scc_graph_download", "scc_state_reset", "scc_state_set_colors",
    "scc_trim1", "scc_fb_global", "scc_trim2", "scc_wcc", "scc_color_round", "scc_fb_partitions",
    "scc_partitions_resolve",
    "scc_run_opts_default", "scc_run_method2", "scc_labels_download", "scc_colors_download",
    "scc_count", "scc_labels_device", "scc_update", "scc_ctx_set_kernel_timing", "scc_ctx_set_layout", "scc_graph_layout",
    "scc_ctx_synchronize", "scc_ctx_set_budget", "scc_update_bytes_needed",
    "scc_device_memory", "scc_phase_trace", "scc_sched_set_weights", "scc_sched_get_weights",
    "scc_sched_forward", "scc_sched_gradients", "scc_sched_train", "scc_sched_rollout",
    "scc_batch_set", "scc_batch_run", "scc_batch_counts", "scc_ctx_get_tuning", "scc_ctx_set_tuning",
    "scc_autotune", "scc_level_log", "scc_update_room_default", "scc_ctx_set_update_room",
    "scc_run_device_noalloc", "scc_ctx_set_chain",
]


class SccError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{STATUS_NAMES.get(code, code)}] {msg}")
        self.code = code


class RunOpts(C.Structure):
    _fields_ = [("pivot_policy", C.c_int32), ("trim_stop", C.c_int32), ("use_trim2", C.c_int32),
                ("trim_between", C.c_int32), ("use_wcc", C.c_int32), ("time_phases", C.c_int32),
                ("partition_mode", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("trimmed_first", C.c_int64), ("scc_first", C.c_int64),
                ("pivot_first", C.c_int32), ("chain_stop", C.c_int32),
                ("trimmed_second", C.c_int64), ("wcc_parts", C.c_int64),
                ("fb_rounds", C.c_int64), ("num_sccs", C.c_int64), ("fb_levels", C.c_int64),
                ("ms_total", C.c_double), ("ms_trim1", C.c_double),
                ("ms_fb_global", C.c_double), ("ms_trim2", C.c_double), ("ms_wcc", C.c_double),
                ("ms_fb_parts", C.c_double), ("ms_canon", C.c_double),
                ("ms_fb_kernels", C.c_double), ("fb_launches", C.c_int64),
                ("fb_f", C.c_int64), ("fb_b", C.c_int64), ("fb_ef", C.c_int64),
                ("fb_eb", C.c_int64), ("fb_fb", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PhaseRec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("round", C.c_int32), ("live_before", C.c_int64),
                ("live_after", C.c_int64), ("aux", C.c_int64), ("ms", C.c_double)]


class RolloutStats(C.Structure):
    _fields_ = [("rounds", C.c_int64), ("agree", C.c_int64), ("num_sccs", C.c_int64),
                ("ms_total", C.c_double), ("ms_best", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("light_row", C.c_int32), ("huge_row", C.c_int32), ("grid_cap", C.c_int32),
                ("wcc_dense", C.c_int32), ("bu_alpha", C.c_float), ("td_beta", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UpdateStats(C.Structure):
    _fields_ = [("inserted", C.c_int64), ("deleted", C.c_int64), ("affected", C.c_int64),
                ("num_sccs", C.c_int64), ("ms_total", C.c_double), ("ms_apply", C.c_double),
                ("ms_region", C.c_double), ("ms_recompute", C.c_double), ("giant", C.c_int64),
                ("rebuilds", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class UpdateRoom(C.Structure):
    _fields_ = [("row_spare", C.c_int32), ("row_spare_eighths", C.c_int32),
                ("overflow_eighths", C.c_int32), ("pad", C.c_int32), ("overflow_min", C.c_int64)]


_lib = None
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def load_library():
    """Load libsccmi.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
    sig = {
        "scc_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
        "scc_ctx_destroy": (C.c_int, [vp]),
        "scc_last_error": (C.c_char_p, [vp]),
        "scc_ctx_bytes": (C.c_uint64, [vp]),
        "scc_bytes_needed": (C.c_uint64, [i32, i64]),
        "scc_ctx_cancel": (None, [vp]),
        "scc_ctx_inject_fault": (None, [vp, C.c_int]),
        "scc_graph_upload": (C.c_int, [vp, i32, i64, vp, vp, vp, vp]),
        "scc_graph_upload_device": (C.c_int, [vp, i32, i64, vp, vp, vp, vp]),
        "scc_graph_info": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i64)]),
        "scc_graph_download": (C.c_int, [vp, vp, vp, vp, vp]),
        "scc_state_reset": (C.c_int, [vp]),
        "scc_state_set_colors": (C.c_int, [vp, vp]),
        "scc_trim1": (C.c_int, [vp, i64, C.POINTER(i64)]),
        "scc_fb_global": (C.c_int, [vp, C.c_int, i32, C.POINTER(i64), C.POINTER(i32)]),
        "scc_trim2": (C.c_int, [vp, C.POINTER(i64)]),
        "scc_wcc": (C.c_int, [vp, C.POINTER(i64)]),
        "scc_color_round": (C.c_int, [vp, C.POINTER(i64)]),
        "scc_fb_partitions": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(i64)]),
        "scc_partitions_resolve": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(i64)]),
        "scc_run_opts_default": (None, [C.POINTER(RunOpts)]),
        "scc_run_method2": (C.c_int, [vp, C.POINTER(RunOpts), C.POINTER(RunStats)]),
        "scc_labels_download": (C.c_int, [vp, vp, C.c_int]),
        "scc_colors_download": (C.c_int, [vp, vp]),
        "scc_count": (C.c_int, [vp, C.POINTER(i64)]),
        "scc_labels_device": (C.c_int, [vp, C.POINTER(vp)]),
        "scc_ctx_set_layout": (C.c_int, [vp, C.c_int]),
        "scc_graph_layout": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "scc_update": (C.c_int, [vp, vp, vp, i64, vp, vp, i64, C.POINTER(UpdateStats)]),
        "scc_ctx_set_kernel_timing": (C.c_int, [vp, C.c_int]),
        "scc_ctx_set_chain": (C.c_int, [vp, C.c_int, C.c_int32, C.c_int32, C.c_int32]),
        "scc_ctx_synchronize": (C.c_int, [vp]),
        "scc_ctx_set_budget": (C.c_int, [vp, C.c_uint64]),
        "scc_update_bytes_needed": (C.c_uint64, [i32, i64, i64, i64]),
        "scc_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "scc_phase_trace": (C.c_int, [vp, vp, i64, C.POINTER(i64)]),
        "scc_batch_set": (C.c_int, [vp, i32, vp]),
        "scc_level_log": (C.c_int, [vp, vp, C.POINTER(i32)]),
        "scc_batch_run": (C.c_int, [vp, C.POINTER(RunOpts), C.POINTER(RunStats)]),
        "scc_batch_counts": (C.c_int, [vp, vp]),
        "scc_ctx_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "scc_ctx_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "scc_autotune": (C.c_int, [vp, C.POINTER(RunOpts), i32, C.POINTER(Tuning),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "scc_update_room_default": (None, [C.POINTER(UpdateRoom)]),
        "scc_run_device_noalloc": (C.c_int, [C.c_int, vp, i32, i64, vp, vp, vp, vp, vp, C.POINTER(i64),
                                             C.c_char_p, C.c_size_t]),
        "scc_ctx_set_update_room": (C.c_int, [vp, C.POINTER(UpdateRoom)]),
        "scc_sched_set_weights": (C.c_int, [vp, vp]),
        "scc_sched_get_weights": (C.c_int, [vp, vp]),
        "scc_sched_forward": (C.c_int, [vp, vp, i64, vp]),
        "scc_sched_gradients": (C.c_int, [vp, vp, vp, vp, i64, C.POINTER(C.c_float), vp]),
        "scc_sched_train": (C.c_int, [vp, vp, vp, vp, i64, C.c_int32, C.c_int32, C.c_float,
                                      C.c_uint64, C.POINTER(C.c_float)]),
        "scc_sched_rollout": (C.c_int, [vp, C.POINTER(RunOpts), C.c_uint64, C.c_float, C.c_int32,
                                        vp, vp, vp, i64, C.POINTER(i64), C.POINTER(RolloutStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_opts(**kw) -> RunOpts:
    o = RunOpts()
    load_library().scc_run_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o


class SccEngine:
    """One device context: upload a graph, run phases or the whole method-2
    pipeline, download labels.  Method names mirror src/scc.h."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.scc_ctx_create(int(device), int(hbm_budget), C.byref(h))
        if rc != 0:
            raise SccError(rc, "scc_ctx_create failed (no HIP device?)")
        self.h = h
        self._graph = None

    # -- plumbing
    def _check(self, rc: int):
        if rc != 0:
            raise SccError(rc, self.lib.scc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.scc_ctx_destroy(self.h)
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

    @property
    def bytes(self) -> int:
        return int(self.lib.scc_ctx_bytes(self.h))

    # -- graph
    def set_layout(self, layout: int) -> None:
        """Vertex layout of the next upload (LAYOUT_INPUT or LAYOUT_DEGREE); ids,
        labels and colours seen through this object stay the caller's."""
        self._check(self.lib.scc_ctx_set_layout(self.h, int(layout)))

    def layout(self) -> int:
        """Layout the uploaded graph is held in (LAYOUT_INPUT or LAYOUT_DEGREE)."""
        x = C.c_int()
        self._check(self.lib.scc_graph_layout(self.h, C.byref(x)))
        return x.value

    def upload(self, g) -> None:
        b, a, rb, ra = (_i32(x) for x in (g.begin, g.node_idx, g.r_begin, g.r_node_idx))
        self._graph = (b, a, rb, ra)
        self._check(self.lib.scc_graph_upload(self.h, b.shape[0] - 1, a.shape[0], _ptr(b),
                                              _ptr(a), _ptr(rb), _ptr(ra)))
        self._graph = None

    def info(self):
        n = C.c_int32()
        m = C.c_int64()
        self._check(self.lib.scc_graph_info(self.h, C.byref(n), C.byref(m)))
        return n.value, m.value

    def download_graph(self):
        n, m = self.info()
        b = np.empty(n + 1, np.int32)
        a = np.empty(m, np.int32)
        rb = np.empty(n + 1, np.int32)
        ra = np.empty(m, np.int32)
        self._check(self.lib.scc_graph_download(self.h, _ptr(b), _ptr(a), _ptr(rb), _ptr(ra)))
        from .graphs import Graph
        return Graph(b, a, rb, ra)

    # -- state (src/scc_main.cc:96-118)
    def reset(self):
        self._check(self.lib.scc_state_reset(self.h))

    def set_colors(self, colors):
        c = _i32(colors)
        self._check(self.lib.scc_state_set_colors(self.h, _ptr(c)))

    # -- phases, named as in src/scc.h
    def repeat_global_trim1(self, trim_stop: int = 0) -> int:
        t = C.c_int64()
        self._check(self.lib.scc_trim1(self.h, int(trim_stop), C.byref(t)))
        return t.value

    def do_fw_bw_global_main(self, policy: int = PIVOT_DEGREE_PRODUCT, pivot: int = -1):
        s = C.c_int64()
        p = C.c_int32()
        self._check(self.lib.scc_fb_global(self.h, int(policy), int(pivot), C.byref(s),
                                           C.byref(p)))
        return s.value, p.value

    def do_global_trim2_new(self) -> int:
        t = C.c_int64()
        self._check(self.lib.scc_trim2(self.h, C.byref(t)))
        return t.value

    def do_global_wcc(self) -> int:
        t = C.c_int64()
        self._check(self.lib.scc_wcc(self.h, C.byref(t)))
        return t.value

    def color_round(self) -> int:
        """One colouring step over every live partition; returns the vertices resolved."""
        t = C.c_int64()
        self._check(self.lib.scc_color_round(self.h, C.byref(t)))
        return t.value

    def start_workers_fw_bw_dfs(self, policy: int = PIVOT_DEGREE_PRODUCT,
                                trim_between: bool = True) -> int:
        r = C.c_int64()
        self._check(self.lib.scc_fb_partitions(self.h, int(policy), int(trim_between),
                                               C.byref(r)))
        return r.value

    def resolve_partitions(self, mode: int = PARTS_BOTH, policy: int = PIVOT_DEGREE_PRODUCT,
                           trim_between: bool = True) -> int:
        r = C.c_int64()
        self._check(self.lib.scc_partitions_resolve(self.h, int(policy), int(mode),
                                                    int(trim_between), C.byref(r)))
        return r.value

    def run_method2(self, opts: RunOpts | None = None, **kw) -> dict:
        o = opts if opts is not None else default_opts(**kw)
        st = RunStats()
        self._check(self.lib.scc_run_method2(self.h, C.byref(o), C.byref(st)))
        return st.as_dict()

    def set_kernel_timing(self, on: bool):
        self._check(self.lib.scc_ctx_set_kernel_timing(self.h, int(on)))

    def set_chain(self, on: bool = True, shape: int = -1, level_launches: int = -1, sweeps: int = -1):
        """Queue a fresh decomposition as one chain of launches (default on);
        shape, level launches and sweeps per trim loop < 0 keep the values the
        last decomposition left (scc_ctx_set_chain)."""
        self._check(self.lib.scc_ctx_set_chain(self.h, int(on), shape, level_launches, sweeps))

    def synchronize(self):
        self._check(self.lib.scc_ctx_synchronize(self.h))

    # -- results
    def labels(self, canonical: bool = True) -> np.ndarray:
        n, _ = self.info()
        out = np.empty(n, np.int32)
        self._check(self.lib.scc_labels_download(self.h, _ptr(out), int(canonical)))
        return out

    def colors(self) -> np.ndarray:
        n, _ = self.info()
        out = np.empty(n, np.int32)
        self._check(self.lib.scc_colors_download(self.h, _ptr(out)))
        return out

    def count(self) -> int:
        c = C.c_int64()
        self._check(self.lib.scc_count(self.h, C.byref(c)))
        return c.value

    # -- dynamic updates
    def update(self, ins_src=None, ins_dst=None, del_src=None, del_dst=None) -> dict:
        e = np.empty(0, np.int32)
        isrc = _i32(ins_src if ins_src is not None else e)
        idst = _i32(ins_dst if ins_dst is not None else e)
        dsrc = _i32(del_src if del_src is not None else e)
        ddst = _i32(del_dst if del_dst is not None else e)
        if isrc.shape != idst.shape or dsrc.shape != ddst.shape:
            raise ValueError("src/dst length mismatch")
        st = UpdateStats()
        self._check(self.lib.scc_update(self.h, _ptr(isrc), _ptr(idst), isrc.shape[0],
                                        _ptr(dsrc), _ptr(ddst), dsrc.shape[0], C.byref(st)))
        return st.as_dict()

    def set_update_room(self, **kw) -> None:
        """Room of the gapped row layout updates keep (row_spare,
        row_spare_eighths, overflow_eighths, overflow_min); no arguments restore
        the defaults.  Takes effect at the next layout build."""
        r = UpdateRoom()
        self.lib.scc_update_room_default(C.byref(r))
        for k, v in kw.items():
            setattr(r, k, int(v))
        self._check(self.lib.scc_ctx_set_update_room(self.h, C.byref(r)))

    def set_budget(self, hbm_budget: int) -> None:
        self._check(self.lib.scc_ctx_set_budget(self.h, int(hbm_budget)))

    def level_log(self) -> list:
        """Per level of the last reachability search (both searches of a forward-
        backward step, forward first): direction, mode, input, admitted, edges."""
        out = np.zeros(256, np.int64)
        k = C.c_int32()
        self._check(self.lib.scc_level_log(self.h, out.ctypes.data, C.byref(k)))
        return [{"mode": "BU" if out[4 * i] & 1 else "TD", "dir": "B" if out[4 * i] & 256 else "F",
                 "input": int(out[4 * i + 1] & 0xFFFFFFFF),
                 "split_rows": int(out[4 * i + 1] >> 32) if not out[4 * i] & 1 else 0,
                 "admitted": int(out[4 * i + 2]), "edges": int(out[4 * i + 3])} for i in range(k.value)]

    # -- batched graphs and tuning
    def batch_set(self, offsets) -> None:
        if offsets is None:
            self._check(self.lib.scc_batch_set(self.h, 0, None))
            return
        o = np.ascontiguousarray(offsets, dtype=np.int32)
        self._check(self.lib.scc_batch_set(self.h, int(o.shape[0] - 1), o.ctypes.data))
        self._batch_k = int(o.shape[0] - 1)

    def batch_run(self, opts: RunOpts | None = None, **kw) -> dict:
        o = opts if opts is not None else default_opts(**kw)
        st = RunStats()
        self._check(self.lib.scc_batch_run(self.h, C.byref(o), C.byref(st)))
        return st.as_dict()

    def batch_counts(self) -> np.ndarray:
        out = np.empty(self._batch_k, np.int64)
        self._check(self.lib.scc_batch_counts(self.h, out.ctypes.data))
        return out

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.lib.scc_ctx_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        """Set knobs (light_row, huge_row, grid_cap, wcc_dense, bu_alpha, td_beta); none = defaults."""
        if not kw:
            self._check(self.lib.scc_ctx_set_tuning(self.h, None))
            return
        t = Tuning()
        self._check(self.lib.scc_ctx_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(self.lib.scc_ctx_set_tuning(self.h, C.byref(t)))

    def autotune(self, reps: int = 3, opts: RunOpts | None = None) -> dict:
        t = Tuning()
        b, d = C.c_double(), C.c_double()
        o = opts if opts is not None else default_opts()
        self._check(self.lib.scc_autotune(self.h, C.byref(o), int(reps), C.byref(t), C.byref(b),
                                          C.byref(d)))
        return {"tuning": t.as_dict(), "best_ms": b.value, "default_ms": d.value}

    # -- phase trace and learned scheduler
    def phase_trace(self) -> list:
        n = C.c_int64()
        self._check(self.lib.scc_phase_trace(self.h, None, 0, C.byref(n)))
        buf = (PhaseRec * max(1, n.value))()
        self._check(self.lib.scc_phase_trace(self.h, buf, n.value, C.byref(n)))
        return [{"phase": PHASE_NAMES.get(r.kind, r.kind), "round": r.round,
                 "live_before": r.live_before, "live_after": r.live_after, "aux": r.aux, "ms": r.ms}
                for r in buf[:n.value]]

    def set_sched_weights(self, params) -> None:
        if params is None:
            self._check(self.lib.scc_sched_set_weights(self.h, None))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (SCHED_NPARAMS,):
            raise ValueError(f"expected {SCHED_NPARAMS} parameters")
        self._check(self.lib.scc_sched_set_weights(self.h, p.ctypes.data))

    def sched_weights(self) -> np.ndarray:
        p = np.empty(SCHED_NPARAMS, np.float32)
        self._check(self.lib.scc_sched_get_weights(self.h, p.ctypes.data))
        return p

    def sched_forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty((X.shape[0], SCHED_HEADS), np.float32)
        self._check(self.lib.scc_sched_forward(self.h, X.ctypes.data, X.shape[0], out.ctypes.data))
        return out

    def sched_gradients(self, X, Y, M):
        X, Y, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, M))
        g = np.empty(SCHED_NPARAMS, np.float32)
        loss = C.c_float()
        self._check(self.lib.scc_sched_gradients(self.h, X.ctypes.data, Y.ctypes.data, M.ctypes.data,
                                                 X.shape[0], C.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def sched_train(self, X, Y, M, epochs=10, batch=1024, lr=1e-3, seed=0) -> float:
        X, Y, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, Y, M))
        loss = C.c_float()
        self._check(self.lib.scc_sched_train(self.h, X.ctypes.data, Y.ctypes.data, M.ctypes.data,
                                             X.shape[0], int(epochs), int(batch), float(lr),
                                             int(seed) & (2**64 - 1), C.byref(loss)))
        return float(loss.value)

    def sched_rollout(self, seed=0, explore=0.1, samples_per_round=256, cap=1 << 20,
                      opts: RunOpts | None = None):
        """(X, Y, M, stats) from one counterfactual decomposition of the uploaded graph."""
        X = np.zeros((cap, SCHED_FEATURES), np.float32)
        Y = np.zeros((cap, SCHED_HEADS), np.float32)
        M = np.zeros((cap, SCHED_HEADS), np.float32)
        n = C.c_int64()
        st = RolloutStats()
        o = opts if opts is not None else default_opts()
        self._check(self.lib.scc_sched_rollout(self.h, C.byref(o), int(seed) & (2**64 - 1),
                                               float(explore), int(samples_per_round), X.ctypes.data,
                                               Y.ctypes.data, M.ctypes.data, cap, C.byref(n),
                                               C.byref(st)))
        k = n.value
        return X[:k].copy(), Y[:k].copy(), M[:k].copy(), st.as_dict()

    # -- fault hooks (host-side failure tests)
    def cancel(self):
        self.lib.scc_ctx_cancel(self.h)

    def inject_fault(self, after: int):
        self.lib.scc_ctx_inject_fault(self.h, int(after))


def scc_labels(g, **kw) -> np.ndarray:
    """Canonical SCC labels of a graph through the GPU engine (one-shot)."""
    with SccEngine() as e:
        e.upload(g)
        e.run_method2(**kw)
        return e.labels(canonical=True)

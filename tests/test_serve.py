"""Host-side behaviour of the request server that needs no GPU: configuration
checks and a clean failure when no device is present."""
import ctypes

import pytest

from dynamicgraphs_scc_amd import serve
from dynamicgraphs_scc_amd.engine import SccError, load_library


def _have_device():
    lib = load_library()
    f, t = ctypes.c_uint64(), ctypes.c_uint64()
    return lib.scc_device_memory(0, ctypes.byref(f), ctypes.byref(t)) == 0


def test_bad_configuration_rejected():
    for kw in ({"slots": 0}, {"slots": 65}, {"max_batch": 0}, {"growth": -1.0}, {"queue_limit": 0},
               {"layout": 7}, {"coalesce": (1 << 20) + 1}):
        with pytest.raises(SccError) as ei:
            serve.SccServer(**kw)
        assert ei.value.code == -1, kw


def test_no_device_fails_cleanly():
    if _have_device():
        pytest.skip("a HIP device is present")
    with pytest.raises(SccError) as ei:
        serve.SccServer()
    assert ei.value.code == -2 and "no HIP device" in str(ei.value)


def test_default_config():
    lib = serve.load_serve_library()
    cfg = serve.ServerConfig()
    lib.scc_server_config_default(ctypes.byref(cfg))
    assert cfg.slots >= 1 and cfg.max_batch == 100000 and cfg.hbm_budget_bytes == 0
    assert cfg.layout == 0 and cfg.coalesce == 32


def test_update_host_model_matches_brute_force():
    """The host model of one update that the merged-update server test and the
    dynamic tests check against, against a per-edge loop over a small
    multigraph: deletes drop every copy unless the batch inserts the edge,
    inserts add an absent edge once, inserted self-loops do nothing."""
    import numpy as np

    from helpers import apply_update_host
    rng = np.random.default_rng(5)
    n = 12
    for _ in range(50):
        src = rng.integers(0, n, 40).astype(np.int32)
        dst = rng.integers(0, n, 40).astype(np.int32)
        ins = (rng.integers(0, n, 8).astype(np.int32), rng.integers(0, n, 8).astype(np.int32))
        dels = (rng.integers(0, n, 8).astype(np.int32), rng.integers(0, n, 8).astype(np.int32))
        edges = list(zip(src.tolist(), dst.tolist()))
        inserting = {(u, v) for u, v in zip(*[x.tolist() for x in ins]) if u != v}
        gone = {(u, v) for u, v in zip(*[x.tolist() for x in dels])} - inserting
        kept = [e for e in edges if e not in gone]
        removed = len(edges) - len(kept)
        for e in sorted(inserting):
            if e not in kept:
                kept.append(e)
        s2, d2, r2 = apply_update_host(n, src, dst, ins, dels)
        assert r2 == removed
        assert sorted(zip(s2.tolist(), d2.tolist())) == sorted(kept)

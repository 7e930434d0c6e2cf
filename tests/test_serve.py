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
               {"layout": 7}):
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
    assert cfg.layout == 0

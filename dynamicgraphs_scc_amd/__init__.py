"""MI355X-native SCC engine.

* ``graphs``: host CSR/reverse-CSR container and seeded workloads (libsccgen.so).
* ``engine``: ctypes binding of the gfx950 HIP engine behind include/sccmi.h
  (libsccmi.so) and ``SccEngine``, a mirror of the reference's scc.h calls.
"""
from . import graphs  # noqa: F401

__all__ = ["graphs", "engine"]


def __getattr__(name):
    if name in ("engine",):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)

"""The C ABI libraries load and export every symbol their headers declare.
No compute calls: these run without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LIB = os.path.join(ROOT, "dynamicgraphs_scc_amd", "lib")


def declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n.startswith(("scc_", "sccgen_"))))


@pytest.mark.parametrize("header,lib", [("sccmi.h", "libsccmi.so"), ("sccgen.h", "libsccgen.so"),
                                        ("sccserve.h", "libsccserve.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared(header)
    assert len(names) > 10
    so = ctypes.CDLL(os.path.join(LIB, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_covers_the_header():
    from dynamicgraphs_scc_amd import engine, serve
    assert sorted(engine.EXPORTS) == declared("sccmi.h")
    assert sorted(serve.EXPORTS) == declared("sccserve.h")


def test_struct_layouts_match_header():
    """ctypes mirrors of scc_run_opts / scc_run_stats / scc_update_stats have the
    sizes the C compiler gives the header's structs."""
    from dynamicgraphs_scc_amd import engine
    import subprocess
    import tempfile
    from dynamicgraphs_scc_amd import serve
    src = r"""
#include <stdio.h>
#include "sccmi.h"
#include "sccserve.h"
int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(scc_run_opts), sizeof(scc_run_stats),
  sizeof(scc_update_stats), sizeof(scc_server_config), sizeof(scc_request_info), sizeof(scc_server_stats),
  sizeof(scc_phase_rec), sizeof(scc_rollout_stats), sizeof(scc_tuning));}
"""
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "s.c")
        with open(c, "w") as f:
            f.write(src)
        exe = os.path.join(td, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [ctypes.sizeof(engine.RunOpts), ctypes.sizeof(engine.RunStats),
                                     ctypes.sizeof(engine.UpdateStats), ctypes.sizeof(serve.ServerConfig),
                                     ctypes.sizeof(serve.RequestInfo), ctypes.sizeof(serve.ServerStats),
                                     ctypes.sizeof(engine.PhaseRec), ctypes.sizeof(engine.RolloutStats),
                                     ctypes.sizeof(engine.Tuning)]


def test_no_cpu_fallback_in_product_path():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "dynamicgraphs_scc_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith(".py"):
                with open(os.path.join(dirpath, fn)) as f:
                    txt = f.read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", txt, re.M), fn


def test_shim_defines_every_scc_h_entry_point():
    """A C++ program that takes the address of every function scc.h declares
    links against libsccshim.so (a missing definition fails the link) and runs
    without touching the GPU."""
    import subprocess
    import tempfile
    with open(os.path.join(ROOT, "include", "scc.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = sorted(set(re.findall(r"^\s*(?:void|int|long long)\s+([a-z_][a-zA-Z0-9_]*)\s*\(", text, re.M)))
    assert {"repeat_global_trim1", "do_fw_bw_global_main", "start_workers_fw_bw_dfs",
            "do_global_wcc", "output_scc_list", "insert_edges"} <= set(names)
    body = "\n".join(f'  printf("%d\\n", (int)(reinterpret_cast<void*>(&{n}) != nullptr));' for n in names)
    src = f"""#include <stdio.h>
#include "scc.h"
int main() {{
{body}
  printf("%d\\n", G_num_nodes);
  return 0;
}}
"""
    with tempfile.TemporaryDirectory() as td:
        cc = os.path.join(td, "t.cpp")
        with open(cc, "w") as f:
            f.write(src)
        exe = os.path.join(td, "t")
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), cc, "-o", exe,
                        "-L", LIB, "-lsccshim", "-lsccmi", "-lsccgen", f"-Wl,-rpath,{LIB}"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["1"] * len(names) + ["0"]


def test_enum_values_match_header():
    """The Python constants for layouts, partition modes and pivot policies are
    the values the C compiler gives the header's enums."""
    import subprocess
    import tempfile
    from dynamicgraphs_scc_amd import engine
    src = r"""
#include <stdio.h>
#include "sccmi.h"
int main(){printf("%d %d %d %d %d %d %d %d %d %d\n", SCC_LAYOUT_INPUT, SCC_LAYOUT_DEGREE, SCC_LAYOUT_AUTO,
  SCC_PARTS_FB, SCC_PARTS_COLOR, SCC_PARTS_BOTH, SCC_PARTS_LEARNED, SCC_PIVOT_DEGREE_PRODUCT,
  SCC_PIVOT_MIN_ID, SCC_PIVOT_LEARNED);}
"""
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "e.c")
        with open(c, "w") as f:
            f.write(src)
        exe = os.path.join(td, "e")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out == [engine.LAYOUT_INPUT, engine.LAYOUT_DEGREE, engine.LAYOUT_AUTO, engine.PARTS_FB,
                   engine.PARTS_COLOR, engine.PARTS_BOTH, engine.PARTS_LEARNED,
                   engine.PIVOT_DEGREE_PRODUCT, engine.PIVOT_MIN_ID, engine.PIVOT_LEARNED]

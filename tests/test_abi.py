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


# Every function the reference's src/scc.h:19-92 declares, and the queue calls of
# src/my_work_queue.h:18-25 it includes, as (name, function-pointer type).
REFERENCE_SCC_H = [
    ("initialize_color", "void(*)()"), ("finalize_color", "void(*)()"),
    ("get_new_color", "int(*)()"), ("get_curr_color", "int(*)()"),
    ("choose_pivot_from_color", "node_t(*)(gm_graph&, int)"),
    ("generate_compact_set", "NODE_SET*(*)(gm_graph&, int)"),
    ("initialize_tarjan", "void(*)()"), ("finalize_tarjan", "void(*)()"),
    ("do_tarjan_all", "void(*)(gm_graph&)"), ("do_tarjan", "void(*)(gm_graph&, int, node_t)"),
    ("do_tarjan_parallel_color", "void(*)(gm_graph&, node_t*)"),
    ("initialize_trim1", "void(*)()"), ("do_global_trim1", "int(*)(gm_graph&)"),
    ("do_global_trim1_compact", "int(*)(gm_graph&)"),
    ("repeat_global_trim1", "int(*)(gm_graph&, int)"),
    ("repeat_global_trim1_compact", "int(*)(gm_graph&, int)"),
    ("get_compact_trim_targets", "std::vector<node_t>&(*)()"),
    ("repeat_local_trim1", "int(*)(gm_graph&, my_work*)"),
    ("initialize_trim2", "void(*)()"), ("finalize_trim2", "void(*)()"),
    ("do_global_trim2", "int(*)(gm_graph&)"), ("repeat_global_trim2", "int(*)(gm_graph&, int)"),
    ("do_global_trim2_new", "int(*)(gm_graph&)"), ("repeat_global_trim2_new", "int(*)(gm_graph&, int)"),
    ("initialize_WCC", "void(*)()"), ("finalize_WCC", "void(*)()"),
    ("do_global_wcc", "void(*)(gm_graph&)"), ("create_work_items_from_wcc", "void(*)(gm_graph&)"),
    ("get_WCC", "node_t*(*)()"),
    ("do_fw_bw_single_thread", "int(*)(gm_graph&, my_work*, std::vector<my_work*>&)"),
    ("start_workers_fw_bw", "void(*)(gm_graph&, int)"), ("start_workers_fw_bw_dfs", "void(*)(gm_graph&, int)"),
    ("initialize_global_fb", "void(*)()"),
    ("do_fw_bw_global_main", "int(*)(gm_graph&, int, int, bool)"),
    ("create_works_after_bfs_trim", "void(*)(gm_graph&)"),
    ("initialize_analyze", "void(*)()"), ("finalize_analyze", "void(*)()"),
    ("print_scc_of_size", "void(*)(int)"), ("print_scc_of_nontrivial_size", "void(*)(int)"),
    ("create_histogram_and_print", "void(*)()"), ("output_scc_list", "void(*)()"),
    ("insert_idea1", "void(*)(gm_graph&, std::vector<std::pair<int,int>>, std::vector<std::pair<int,int>>)"),
    ("insert_idea2", "void(*)(gm_graph&, std::vector<std::pair<int,int>>)"),
    ("delete_idea1", "void(*)(gm_graph&, std::vector<std::vector<std::pair<int,int>>>)"),
    ("is_work_q_empty_from_seq_context", "bool(*)()"), ("work_q_fetch", "my_work*(*)(int)"),
    ("work_q_fetch_N", "void(*)(int, int, std::vector<my_work*>&)"),
    ("work_q_put", "void(*)(int, my_work*)"), ("work_q_put_all", "void(*)(int, std::vector<my_work*>&)"),
    ("work_q_init", "void(*)(int)"), ("work_q_print_max_depth", "void(*)()"), ("work_q_size", "int(*)()"),
]


def _build_and_run(src, td, name):
    import subprocess
    cc = os.path.join(td, name + ".cpp")
    with open(cc, "w") as f:
        f.write(src)
    exe = os.path.join(td, name)
    subprocess.run(["g++", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "include"), cc, "-o", exe,
                    "-L", LIB, "-lsccshim", "-lsccmi", "-lsccgen", f"-Wl,-rpath,{LIB}"], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()


def test_shim_defines_every_scc_h_entry_point():
    """Every function of the reference's scc.h (and of the work queue it
    includes), bound to a pointer of the reference's exact type: a missing
    definition fails the link and a different signature fails the build.  The
    program runs without touching the GPU.  The other names include/scc.h
    declares link too."""
    import tempfile
    body = "\n".join(f'  {{ using T = {t}; T p = &{n}; printf("%d\\n", (int)(p != nullptr)); }}'
                     for n, t in REFERENCE_SCC_H)
    with open(os.path.join(ROOT, "include", "scc.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    ours = sorted(set(re.findall(r"^\s*(?:void|int|long long|bool)\s+([a-z_][a-zA-Z0-9_]*)\s*\(", text, re.M)))
    extra = [n for n in ours if n not in dict(REFERENCE_SCC_H)]
    assert {"insert_edges", "delete_edges", "run_method2", "scc_shim_sync"} <= set(extra)
    body += "\n" + "\n".join(f'  printf("%d\\n", (int)(reinterpret_cast<void*>(&{n}) != nullptr));'
                             for n in extra)
    src = f"""#include <stdio.h>
#include "scc.h"
int main() {{
{body}
  printf("%d %d\\n", G_num_nodes, get_curr_color());
  return 0;
}}
"""
    with tempfile.TemporaryDirectory() as td:
        out = _build_and_run(src, td, "t")
    assert out == ["1"] * (len(REFERENCE_SCC_H) + len(extra)) + ["0", "-1"]


def test_reference_style_driver_builds():
    """tests/shim_driver.cpp (the reference driver's call sequence) compiles and
    links against include/scc.h and libsccshim.so; without arguments it prints
    its usage and exits 2 (no GPU call)."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "drv")
        subprocess.run(["g++", "-std=c++17", "-O1", "-fopenmp", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "shim_driver.cpp"), "-o", exe, "-L", LIB, "-lsccshim",
                        "-lsccmi", "-lsccgen", f"-Wl,-rpath,{LIB}"], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


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

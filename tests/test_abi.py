"""The C-ABI library: it loads without a GPU and exports every entry point
include/fastrrt.h declares; the Python binding declares the same set."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastrrt.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
        if m.group(1) not in ("if", "while", "for", "sizeof"):
            names.add(m.group(1))
    return names


def test_header_declares_the_reference_abi():
    names = declared_symbols()
    # src/python_bind.cpp:7-199
    fastrrt = {"fastrrt_initialize", "fastrrt_destroy", "set_plan_data", "goal_reached", "search_init", "loop",
               "path_optimize", "export_graph_nodes", "release_export_graph_nodes", "get_planned_path",
               "interpolate_planned_path", "release_planned_path_data", "interpolate_planned_path_p",
               "ideal_curve", "compute_region_debug_performance", "save_current_graph_state", "load_graph_state"}
    # src/python_cuda_graph_bind.cpp:7-220, all 25
    graph = {"cudagraph_initialize", "cudagraph_destroy", "get_intrinsic_costs", "destroy_intrinsic_costs_ptr",
             "compute_apf_repulsion", "compute_apf_attraction", "can_connect_to_goal_using_hermite",
             "add", "add_temporary", "derivate_node", "accept_derived_nodes", "check_in_graph", "get_heading",
             "clear", "list_all", "free_list_all", "solve_collisions", "process_direct_goal_connection",
             "is_directly_connected_to_goal", "direct_connection_to_goal_cost", "direct_connection_to_goal_heading",
             "dump_graph_to_file", "read_from_dump_file", "get_parent", "free_parent_data", "get_cost", "count_all"}
    assert fastrrt <= names
    assert graph <= names


def test_library_exports_every_declared_symbol():
    from fastrrt_amd import _lib
    lib = _lib.load()  # loading needs no GPU
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared_symbols() <= exported


def test_binding_matches_header():
    from fastrrt_amd import _lib
    assert set(_lib.SIGNATURES) == declared_symbols()


def test_library_reports_missing_gpu_loudly():
    from fastrrt_amd import _lib
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    lib = _lib.load()
    h = lib.fastrrt_initialize(64, 64, 10.0, 10.0, 40.0, 5.4, -1, -1, -1, -1, None, -1.0)
    assert not h
    assert _lib.last_error() != ""
    # the graph and batch constructors fail part-way and free what they made
    for _ in range(3):
        g = lib.cudagraph_initialize(64, 64, 10.0, 10.0, 40.0, 5.4, 22, 40, -1, -1, -1, -1, None, -1.0)
        assert not g
        assert _lib.last_error() != ""
        b = lib.fastrrt_batch_create(4, 4 * 64 * 64)
        assert not b
        assert _lib.last_error() != ""

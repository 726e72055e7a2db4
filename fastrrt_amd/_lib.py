"""ctypes binding of libfastrrt.so (the C ABI declared in include/fastrrt.h).

The library is the in-tree build fastrrt_amd/lib/libfastrrt.so; importing
this module fails loudly when it is missing (there is no CPU fallback).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FASTRRT_LIB names another build of the library (an instrumented or
# experimental one); the default is the in-tree build
LIB_PATH = os.environ.get("FASTRRT_LIB") or os.path.join(HERE, "lib", "libfastrrt.so")

c_int, c_float, c_bool, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_bool, ctypes.c_void_p
c_char_p, c_u64, c_i64, c_double = ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double
c_size_t = ctypes.c_size_t
fp = ctypes.POINTER(ctypes.c_float)
ip = ctypes.POINTER(ctypes.c_int)
i32p = ctypes.POINTER(ctypes.c_int32)
u8p = ctypes.POINTER(ctypes.c_uint8)

# name -> (restype, argtypes); kept in the order of include/fastrrt.h
SIGNATURES = {
    # frames
    "fastrrt_frame_create": (c_void_p, [c_int, c_int, fp, fp, c_int]),
    "fastrrt_frame_destroy": (None, [c_void_p]),
    "fastrrt_frame_update": (c_bool, [c_void_p, fp]),
    "fastrrt_frame_read": (c_bool, [c_void_p, fp]),
    "fastrrt_frame_process_distance_to_goal": (c_bool, [c_void_p, c_int, c_int]),
    "fastrrt_frame_process_safe_zone": (c_bool, [c_void_p, c_int, c_int]),
    "fastrrt_last_error": (c_char_p, []),
    # FastRRT handle (src/python_bind.cpp)
    "fastrrt_initialize": (c_void_p, [c_int, c_int, c_float, c_float, c_float, c_float, c_int, c_int, c_int, c_int,
                                      fp, c_float]),
    "fastrrt_destroy": (None, [c_void_p]),
    "set_plan_data": (None, [c_void_p, c_void_p, c_int, c_int, c_float, c_int, c_int, c_float, c_float, c_int, c_int,
                             c_int, c_float, c_float, c_float]),
    "goal_reached": (c_bool, [c_void_p]),
    "search_init": (None, [c_void_p, c_bool]),
    "loop": (c_bool, [c_void_p, c_bool]),
    "path_optimize": (c_bool, [c_void_p]),
    "export_graph_nodes": (ip, [c_void_p]),
    "release_export_graph_nodes": (None, [ip]),
    "get_planned_path": (fp, [c_void_p]),
    "interpolate_planned_path": (fp, [c_void_p]),
    "release_planned_path_data": (None, [fp]),
    "interpolate_planned_path_p": (fp, [c_void_p, fp, c_int]),
    "ideal_curve": (fp, [c_void_p, c_int, c_int, c_float]),
    "compute_region_debug_performance": (None, [c_void_p]),
    "save_current_graph_state": (None, [c_void_p, c_char_p]),
    "load_graph_state": (None, [c_void_p, c_char_p]),
    "fastrrt_set_seed": (None, [c_void_p, c_u64, c_u64]),
    "fastrrt_set_density_region": (None, [c_void_p, c_int]),
    "fastrrt_iterations": (c_int, [c_void_p]),
    "fastrrt_iteration_stats": (None, [c_void_p, ip, ip, ip]),
    "fastrrt_export_state": (None, [c_void_p, i32p, i32p, i32p, i32p, fp, fp, u8p, fp, fp]),
    "fastrrt_profile_enable": (None, [c_bool]),
    "fastrrt_profile_reset": (None, []),
    "fastrrt_profile_split": (None, [c_bool]),
    "fastrrt_profile_read": (c_bool, [c_int, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(c_double)]),
    # graph handle (src/python_cuda_graph_bind.cpp)
    "cudagraph_initialize": (c_void_p, [c_int, c_int, c_float, c_float, c_float, c_float, c_int, c_int, c_int, c_int,
                                        c_int, c_int, fp, c_float]),
    "cudagraph_destroy": (None, [c_void_p]),
    "can_connect_to_goal_using_hermite": (c_bool, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float]),
    "cudagraph_can_connect_hermite": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_float, fp]),
    "compute_apf_repulsion": (None, [c_void_p, c_void_p, c_float, c_int]),
    "compute_apf_repulsion_noalloc": (c_bool, [c_void_p, c_void_p, c_float, c_int]),
    "fastrrt_apf_noalloc_scratch_bytes": (c_size_t, [c_int, c_int, c_int]),
    "fastrrt_apf_noalloc_scratch_capacity": (c_size_t, []),
    "compute_apf_attraction": (None, [c_void_p, c_void_p, c_float, c_int, c_int]),
    "get_intrinsic_costs": (fp, [c_void_p]),
    "destroy_intrinsic_costs_ptr": (None, [fp]),
    "add": (None, [c_void_p, c_int, c_int, c_float, c_int, c_int, c_float]),
    "add_temporary": (None, [c_void_p, c_int, c_int, c_float, c_int, c_int, c_float]),
    "derivate_node": (None, [c_void_p, c_void_p, c_float, c_int, c_float, c_int, c_int]),
    "accept_derived_nodes": (None, [c_void_p, c_int, c_int, c_float]),
    "check_in_graph": (c_bool, [c_void_p, c_int, c_int]),
    "get_heading": (c_float, [c_void_p, c_int, c_int]),
    "clear": (None, [c_void_p]),
    "list_all": (fp, [c_void_p]),
    "free_list_all": (None, [fp]),
    "cudagraph_check_feasible_connection": (c_bool, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int]),
    "cudagraph_check_feasible_connections": (c_int, [c_void_p, c_void_p, i32p, c_int, c_int, u8p,
                                                     ctypes.POINTER(c_double), fp]),
    "solve_collisions": (None, [c_void_p]),
    "process_direct_goal_connection": (None, [c_void_p, c_void_p, c_int, c_int, c_float, c_float]),
    "is_directly_connected_to_goal": (c_bool, [c_void_p, c_int, c_int]),
    "direct_connection_to_goal_cost": (c_float, [c_void_p, c_int, c_int]),
    "direct_connection_to_goal_heading": (c_float, [c_void_p, c_int, c_int]),
    "dump_graph_to_file": (None, [c_void_p, c_char_p]),
    "read_from_dump_file": (None, [c_void_p, c_char_p]),
    "get_parent": (ip, [c_void_p, c_int, c_int]),
    "free_parent_data": (None, [ip]),
    "get_cost": (c_float, [c_void_p, c_int, c_int]),
    "count_all": (c_int, [c_void_p]),
    "cudagraph_set_collision": (None, [c_void_p, c_int, c_int, c_int, c_int, c_float, c_float]),
    "cudagraph_get_type": (c_int, [c_void_p, c_int, c_int]),
    "cudagraph_get_child_count": (c_int, [c_void_p, c_int, c_int]),
    "cudagraph_count": (c_int, [c_void_p, c_int]),
    "cudagraph_check_consistent": (c_bool, [c_void_p]),
    "cudagraph_find_best_node": (None, [c_void_p, c_void_p, c_float, c_float, c_int, c_int, c_float, ip]),
    "cudagraph_check_goal_reached": (c_bool, [c_void_p, c_int, c_int, c_float, c_float, c_float]),
    "cudagraph_derivate_node": (None, [c_void_p, c_void_p, c_float, c_float, c_float, c_int, c_int, ip]),
    "cudagraph_export_state": (None, [c_void_p, i32p, i32p, i32p, i32p, fp, fp, u8p, fp, fp]),
}



class PlanRequest(ctypes.Structure):
    _fields_ = [
        ("frame", c_void_p),
        ("start_x", c_int), ("start_z", c_int), ("start_heading_rad", c_float),
        ("goal_x", c_int), ("goal_z", c_int), ("goal_heading_rad", c_float),
        ("velocity_m_s", c_float), ("min_dist_x", c_int), ("min_dist_z", c_int),
        ("lower_bound_x", c_int), ("lower_bound_z", c_int), ("upper_bound_x", c_int), ("upper_bound_z", c_int),
        ("max_steering_angle_deg", c_float), ("vehicle_length", c_float), ("max_curvature", c_float),
        ("max_path_size", c_float), ("dist_to_goal_tolerance", c_float), ("heading_error_tolerance_rad", c_float),
        ("max_iterations", c_int), ("optimize_rounds", c_int), ("smart", c_int), ("region_side", c_int),
        ("deadline_ms", c_double), ("seed", c_u64), ("stream", c_u64), ("user_tag", c_i64),
        ("expansion_budget", c_int),
    ]


class BatchHbm(ctypes.Structure):
    _fields_ = [("total_bytes", ctypes.c_uint64), ("free_bytes_at_create", ctypes.c_uint64),
                ("free_bytes_last", ctypes.c_uint64), ("reserve_bytes", ctypes.c_uint64),
                ("engine_bytes", ctypes.c_uint64), ("slab_bytes", ctypes.c_uint64),
                ("requested_cells", c_i64), ("cell_budget", c_i64), ("cells_in_use", c_i64),
                ("memory_stalls", c_i64), ("running", ctypes.c_int32), ("queued", ctypes.c_int32)]


class PlanResult(ctypes.Structure):
    _fields_ = [("user_tag", c_i64), ("status", c_int), ("iterations", c_int), ("path_len", c_int),
                ("expansions", c_int), ("submit_ms", c_double), ("done_ms", c_double), ("path_cost", c_float),
                ("work_rollout_cells", c_i64), ("work_goal_points", c_i64), ("work_goal_foot_points", c_i64)]


SIGNATURES.update({
    "fastrrt_batch_create": (c_void_p, [c_int, c_i64]),
    "fastrrt_batch_destroy": (None, [c_void_p]),
    "fastrrt_batch_submit": (c_int, [c_void_p, ctypes.POINTER(PlanRequest)]),
    "fastrrt_batch_step": (c_int, [c_void_p, c_int]),
    "fastrrt_batch_poll": (c_bool, [c_void_p, ctypes.POINTER(PlanResult), fp, c_int]),
    "fastrrt_batch_drain": (None, [c_void_p]),
    "fastrrt_batch_clock_ms": (c_double, [c_void_p]),
    "fastrrt_batch_set_launch": (c_bool, [c_void_p, c_int, c_int, c_int]),
    "fastrrt_batch_get_launch": (c_bool, [c_void_p, ctypes.POINTER(c_int)]),
    "fastrrt_batch_set_block_threads": (c_bool, [c_void_p, c_int]),
    "fastrrt_batch_get_block_threads": (c_int, [c_void_p]),
    "fastrrt_batch_memory": (c_bool, [c_void_p, ctypes.c_void_p]),
    "fastrrt_batch_set_hbm_reserve": (c_bool, [c_void_p, ctypes.c_uint64]),
})

_lib = None


def load():
    """Load libfastrrt.so and declare every entry point; raises when absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or "
                          "`python fastrrt_amd/build.py` (the planner has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error():
    msg = load().fastrrt_last_error()
    return msg.decode() if msg else ""


def check(what=""):
    msg = last_error()
    if msg:
        raise RuntimeError(f"fastrrt{(' ' + what) if what else ''}: {msg}")

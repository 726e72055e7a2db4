"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# FR_ORACLE_LIB points the suite at another build of the oracle (e.g. a
# deliberately broken copy, to check that the pins catch the mistake)
LIB_PATH = os.environ.get("FR_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "liboracle.so")

NULL, NODE, TEMP, PROCESSING, COLLISION, CONNECT_TO_GOAL = range(6)


class OgGraph(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int),
        ("px", ctypes.POINTER(ctypes.c_int32)), ("pz", ctypes.POINTER(ctypes.c_int32)),
        ("type", ctypes.POINTER(ctypes.c_int32)), ("derive", ctypes.POINTER(ctypes.c_int32)),
        ("heading", ctypes.POINTER(ctypes.c_float)), ("cost", ctypes.POINTER(ctypes.c_float)),
        ("intrinsic", ctypes.POINTER(ctypes.c_float)), ("direct", ctypes.POINTER(ctypes.c_float)),
        ("gc_flag", ctypes.POINTER(ctypes.c_float)), ("gc_heading", ctypes.POINTER(ctypes.c_float)),
        ("gc_cost", ctypes.POINTER(ctypes.c_float)),
        ("region_side", ctypes.c_int), ("dens_w", ctypes.c_int), ("dens_h", ctypes.c_int),
        ("region", ctypes.POINTER(ctypes.c_uint32)),
        ("rate_w", ctypes.c_double), ("inv_rate_w", ctypes.c_double), ("rate_h", ctypes.c_double),
        ("inv_rate_h", ctypes.c_double), ("max_steer_rad", ctypes.c_double), ("max_steer_deg", ctypes.c_double),
        ("lr", ctypes.c_double), ("max_curv", ctypes.c_double),
        ("has_physical", ctypes.c_int),
        ("min_dist_x", ctypes.c_int), ("min_dist_z", ctypes.c_int), ("lb_x", ctypes.c_int),
        ("lb_z", ctypes.c_int), ("ub_x", ctypes.c_int), ("ub_z", ctypes.c_int),
        ("class_costs", ctypes.POINTER(ctypes.c_float)), ("n_classes", ctypes.c_int),
        ("og_x", ctypes.c_float), ("og_y", ctypes.c_float), ("og_h", ctypes.c_float),
        ("goal_reached", ctypes.c_int), ("node_collision", ctypes.c_int),
        ("best_parent", ctypes.c_float * 4), ("best_child", ctypes.c_float * 4),
        ("n_footprint_checks", ctypes.c_uint64), ("n_footprint_samples", ctypes.c_uint64),
        ("n_rollouts", ctypes.c_uint64), ("n_hermite_walks", ctypes.c_uint64),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(OgGraph)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    c_int, c_float, c_double, u64, u32 = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32
    sig = {
        "og_graph_new": (P, [c_int, c_int]),
        "og_graph_free": (None, [P]),
        "og_set_region_side": (None, [P, c_int]),
        "og_set_physical_params": (None, [P, c_float, c_float, c_double, c_float, c_float]),
        "og_set_search_params": (None, [P, c_int, c_int, c_int, c_int, c_int, c_int]),
        "og_set_class_costs": (None, [P, fp, c_int]),
        "og_set_coordinate_start": (None, [P, c_int, c_int, c_float]),
        "og_add": (None, [P, c_int, c_int, c_float, c_int, c_int, c_float]),
        "og_add_temporary": (None, [P, c_int, c_int, c_float, c_int, c_int, c_float]),
        "og_remove": (None, [P, c_int, c_int]),
        "og_clear": (None, [P]),
        "og_set_collision": (None, [P, c_int, c_int, c_int, c_int, c_float, c_float]),
        "og_get_parent": (None, [P, c_int, c_int, ip, ip]),
        "og_get_cost": (c_float, [P, c_int, c_int]),
        "og_get_type": (c_int, [P, c_int, c_int]),
        "og_count": (ctypes.c_uint, [P, c_int]),
        "og_count_all": (ctypes.c_uint, [P]),
        "og_feasible": (c_int, [P, fp, c_int, c_int, c_float]),
        "og_kinematic_new_path": (None, [P, fp, c_int, c_int, c_float, c_float, c_float, fp]),
        "og_kinematic_new_path_pose": (None, [P, fp, c_int, c_int, c_float, c_float, c_float, fp,
                                               ctypes.POINTER(c_double)]),
        "og_derivate_node": (None, [P, fp, c_float, c_float, c_float, c_int, c_int, ip]),
        "og_accept_derived_nodes": (None, [P]),
        "og_solve_collisions": (None, [P]),
        "og_count_regions": (None, [P]),
        "og_expand": (None, [P, fp, c_float, c_float, c_int, c_int, c_int, u64, u64, u32]),
        "og_check_direct_connection": (c_float, [P, fp, c_float, c_int, c_int, c_float, c_int, c_int, c_float,
                                                 c_int, c_int]),
        "og_check_direct_connection_host": (c_float, [P, fp, c_float, c_int, c_int, c_float, c_int, c_int,
                                                      c_float, c_int]),
        "og_process_direct_goal_connection": (None, [P, fp, c_int, c_int, c_float, c_float]),
        "og_find_best_goal_direct_connection": (c_int, [P, fp, c_float, c_int]),
        "og_check_goal_reached": (c_int, [P, c_int, c_int, c_float, c_float, c_float]),
        "og_find_best_node": (None, [P, fp, c_float, c_float, c_int, c_int, c_float, ip]),
        "og_check_graph_consistent": (c_int, [P]),
        "og_can_connect_hermite": (c_float, [P, fp, c_float, c_int, c_int, c_int, c_int, c_int, ip]),
        "og_apf_repulsion": (None, [P, fp, c_float, c_int]),
        "og_apf_attraction": (None, [P, fp, c_float, c_int, c_int]),
        "og_check_feasible_connection": (c_int, [P, fp, c_int, c_int, c_int, c_int, c_float]),
        "og_check_feasible_connection_ex": (c_int, [P, fp, c_int, c_int, c_int, c_int, c_float,
                                                    ctypes.POINTER(c_double), fp]),
        "og_interpolate_hermite": (c_int, [c_int, c_int, c_int, c_int, c_float, c_int, c_int, c_float, fp, c_int]),
        "og_dump_graph": (None, [P, ctypes.c_char_p]),
        "og_read_dump": (None, [P, ctypes.c_char_p]),
        "og_rrt_new": (ctypes.c_void_p, [c_int, c_int, c_float, c_float, c_float, c_float, c_int, c_int, c_int,
                                         c_int, c_float]),
        "og_rrt_free": (None, [ctypes.c_void_p]),
        "og_rrt_graph": (P, [ctypes.c_void_p]),
        "og_rrt_set_seed": (None, [ctypes.c_void_p, u64, u64]),
        "og_rrt_set_plan_data": (None, [ctypes.c_void_p, fp, fp, c_int, c_int, c_int, c_float, c_int, c_int,
                                        c_float, c_float, c_int, c_int, c_int, c_float, c_float, c_float]),
        "og_rrt_search_init": (None, [ctypes.c_void_p]),
        "og_rrt_loop": (c_int, [ctypes.c_void_p, c_int]),
        "og_rrt_goal_reached": (c_int, [ctypes.c_void_p]),
        "og_rrt_load_graph_state": (None, [ctypes.c_void_p, ctypes.c_char_p]),
        "og_rrt_get_planned_path": (c_int, [ctypes.c_void_p, fp, c_int]),
        "og_rrt_path_optimize": (c_int, [ctypes.c_void_p]),
        "og_rrt_interpolate_planned_path": (c_int, [ctypes.c_void_p, fp, c_int]),
        "og_rrt_iterations": (c_int, [ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Graph:
    """Thin owner of an og_graph with numpy views of its planes."""

    def __init__(self, W, H):
        self.L = lib()
        self.g = self.L.og_graph_new(W, H)
        self.W, self.H = W, H

    def __del__(self):
        try:
            self.L.og_graph_free(self.g)
        except Exception:
            pass

    def plane(self, name):
        s = self.g.contents
        p = getattr(s, name)
        dt = np.int32 if name in ("px", "pz", "type", "derive") else np.float32
        return np.ctypeslib.as_array(p, shape=(self.H, self.W)).view(dt)

    def physical(self, perc_w, perc_h, steer_deg, L, max_curv=-1.0):
        self.L.og_set_physical_params(self.g, perc_w, perc_h, steer_deg * np.pi / 180.0, L, max_curv)

    def search(self, mdx, mdz, lbx=-1, lbz=-1, ubx=-1, ubz=-1):
        self.L.og_set_search_params(self.g, mdx, mdz, lbx, lbz, ubx, ubz)

    def costs(self, c):
        c = np.ascontiguousarray(c, dtype=np.float32)
        self.L.og_set_class_costs(self.g, fptr(c), len(c))

    def add(self, x, z, h, px, pz, cost):
        self.L.og_add(self.g, x, z, h, px, pz, cost)

    def apf_repulsion(self, frame, kr, radius):
        frame = np.ascontiguousarray(frame, np.float32)
        self.L.og_apf_repulsion(self.g, fptr(frame), kr, radius)

    def apf_attraction(self, frame, ka, gx, gz):
        frame = np.ascontiguousarray(frame, np.float32)
        self.L.og_apf_attraction(self.g, fptr(frame), ka, gx, gz)

    def add_temporary(self, x, z, h, px, pz, cost):
        self.L.og_add_temporary(self.g, x, z, h, px, pz, cost)

    def parent(self, x, z):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.L.og_get_parent(self.g, x, z, ctypes.byref(a), ctypes.byref(b))
        return (a.value, b.value)

    def type_at(self, x, z):
        return self.L.og_get_type(self.g, x, z)

    def child_count(self, x, z):
        return int(self.plane("derive")[z, x])

    def heading_at(self, x, z):
        return float(self.plane("heading")[z, x])


def empty_frame(W, H, cls=0.0):
    f = np.zeros((H, W, 3), dtype=np.float32)
    f[:, :, 0] = cls
    return f

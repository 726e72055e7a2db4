"""pyfastrrt-compatible Python interface over libfastrrt.so.

Mirrors the reference's ctypes package (python_bind/app/pyfastrrt/src/
fastrrt.py and cuda_graph.py): FastRRT and CudaGraph with the same method
names, arguments and return shapes.  The reference takes its parameter and
frame objects from pydriveless; this module provides the subset those two
classes read (angle, Waypoint, EgoParams, SearchParams, SearchFrame), with
the same attribute names, so caller code ports by changing the import.
"""
import ctypes
import math

import numpy as np

from . import _lib


# --------------------------------------------------------------------------
# parameter objects (the pydriveless attributes pyfastrrt reads)
# --------------------------------------------------------------------------

class angle:
    """An angle; stored in radians (pydriveless angle.new_rad / new_deg)."""
    __slots__ = ("_rad",)

    def __init__(self, rad=0.0):
        self._rad = float(rad)

    @staticmethod
    def new_rad(v):
        return angle(v)

    @staticmethod
    def new_deg(v):
        return angle(float(v) * math.pi / 180.0)

    def rad(self):
        return self._rad

    def deg(self):
        return self._rad * 180.0 / math.pi

    def __repr__(self):
        return f"angle({self._rad:.6f} rad)"


class Waypoint:
    __slots__ = ("x", "z", "heading")

    def __init__(self, x, z, heading=None):
        self.x = int(x)
        self.z = int(z)
        self.heading = heading if isinstance(heading, angle) else angle(heading or 0.0)

    def __repr__(self):
        return f"Waypoint({self.x}, {self.z}, {self.heading.rad():.4f})"


class EgoParams:
    """Vehicle/search-frame constants (fastrrt.py:10-34 reads these)."""

    def __init__(self, width, height, perception_width_m, perception_height_m, max_steering_angle_deg,
                 vehicle_length_m, ego_lower_bound=(-1, -1), ego_upper_bound=(-1, -1), max_curvature=-1.0,
                 segmentation_class_costs=None):
        self.search_frame_dimensions = (int(width), int(height))
        self.search_frame_physical_dimensions = (float(perception_width_m), float(perception_height_m))
        self.max_steering_angle = angle.new_deg(max_steering_angle_deg)
        self.vehicle_length_m = float(vehicle_length_m)
        self.ego_lower_bound = tuple(ego_lower_bound)
        self.ego_upper_bound = tuple(ego_upper_bound)
        self.max_curvature = float(max_curvature)
        self.segmentation_class_costs = (np.zeros(0, np.float32) if segmentation_class_costs is None
                                         else np.asarray(segmentation_class_costs, np.float32))


class SearchFrame:
    """A device-resident frame: float3 pixels + class-cost table.

    Replaces pydriveless.SearchFrame; get_cuda_ptr() returns the handle the
    C ABI's frame arguments take.
    """

    def __init__(self, pixels, class_costs):
        px = np.ascontiguousarray(pixels, dtype=np.float32)
        if px.ndim != 3 or px.shape[2] != 3:
            raise ValueError("pixels must be (height, width, 3) float32")
        cc = np.ascontiguousarray(class_costs, dtype=np.float32)
        L = _lib.load()
        self._h, self._w = px.shape[:2]
        self._ptr = L.fastrrt_frame_create(self._w, self._h, px.ctypes.data_as(_lib.fp), cc.ctypes.data_as(_lib.fp),
                                           len(cc))
        if not self._ptr:
            _lib.check("frame_create")
        self.class_costs = cc

    def width(self):
        return self._w

    def height(self):
        return self._h

    def get_cuda_ptr(self):
        return self._ptr

    def read(self) -> np.ndarray:
        """The device pixels, (height, width, 3) float32."""
        out = np.empty((self._h, self._w, 3), np.float32)
        if not _lib.load().fastrrt_frame_read(self._ptr, out.ctypes.data_as(_lib.fp)):
            _lib.check("frame_read")
        return out

    def process_distance_to_goal(self, goal_x: int, goal_z: int) -> None:
        """Fill channel 1 with the pixel distance to the goal (libdriveless
        SearchFrame::processDistanceToGoal; the curve length of
        CudaGraph.can_connect_to_goal_using_hermite)."""
        if not _lib.load().fastrrt_frame_process_distance_to_goal(self._ptr, goal_x, goal_z):
            _lib.check("frame_process_distance_to_goal")

    def process_safe_zone(self, min_dist_x: int, min_dist_z: int) -> None:
        """Set the safe-zone bit 0x100 of channel 2 where the footprint of
        min_dist is free at every heading (no obstacle pixel within Chebyshev
        distance ceil(hypot(min_dist_x // 2, min_dist_z // 2)) + 1); clear it
        elsewhere.  libdriveless's producer of the bit is not in the
        reference tree."""
        if not _lib.load().fastrrt_frame_process_safe_zone(self._ptr, int(min_dist_x), int(min_dist_z)):
            _lib.check("frame_process_safe_zone")

    def update(self, pixels):
        px = np.ascontiguousarray(pixels, dtype=np.float32)
        if px.shape != (self._h, self._w, 3):
            raise ValueError("pixel shape changed")
        if not _lib.load().fastrrt_frame_update(self._ptr, px.ctypes.data_as(_lib.fp)):
            _lib.check("frame_update")

    def __del__(self):
        try:
            if getattr(self, "_ptr", None):
                _lib.load().fastrrt_frame_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass


class SearchParams:
    """Per-plan parameters (fastrrt.py:172-198 reads these)."""

    def __init__(self, frame, start, goal, velocity_m_s=1.0, min_distance=(0, 0), timeout_ms=-1,
                 max_path_size_px=40.0, distance_to_goal_tolerance_px=20.0, heading_error_tolerance=None):
        self.frame = frame
        self.start = start
        self.goal = goal
        self.velocity_m_s = float(velocity_m_s)
        self.min_distance = tuple(min_distance)
        self.timeout_ms = int(timeout_ms)
        self.max_path_size_px = float(max_path_size_px)
        self.distance_to_goal_tolerance_px = float(distance_to_goal_tolerance_px)
        self.heading_error_tolerance = heading_error_tolerance or angle.new_deg(10)


def _path_from_ptr(L, ptr):
    if not ptr:
        _lib.check("path")
        return None
    size = int(ptr[0])
    res = None
    if size > 0:
        res = [Waypoint(int(ptr[3 * i + 1]), int(ptr[3 * i + 2]), angle.new_rad(float(ptr[3 * i + 3])))
               for i in range(size)]
    L.release_planned_path_data(ptr)
    return res


# --------------------------------------------------------------------------
# FastRRT (reference python_bind/app/pyfastrrt/src/fastrrt.py)
# --------------------------------------------------------------------------

class FastRRT:
    def __init__(self, ego_params: EgoParams):
        L = _lib.load()
        costs = np.ascontiguousarray(np.concatenate(([ego_params.segmentation_class_costs.shape[0]],
                                                     ego_params.segmentation_class_costs)), dtype=np.float32)
        w, h = ego_params.search_frame_dimensions
        pw, ph = ego_params.search_frame_physical_dimensions
        lb, ub = ego_params.ego_lower_bound, ego_params.ego_upper_bound
        self._ptr = L.fastrrt_initialize(w, h, pw, ph, ego_params.max_steering_angle.deg(),
                                         ego_params.vehicle_length_m, lb[0], lb[1], ub[0], ub[1],
                                         costs.ctypes.data_as(_lib.fp), ego_params.max_curvature)
        if not self._ptr:
            _lib.check("initialize")
        self._w, self._h = w, h
        self._frame = None

    def __del__(self):
        try:
            if getattr(self, "_ptr", None):
                _lib.load().fastrrt_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass

    def set_plan_data(self, search_params: SearchParams) -> bool:
        L = _lib.load()
        s, g = search_params.start, search_params.goal
        self._frame = search_params.frame  # borrowed by the planner: keep it alive
        L.set_plan_data(self._ptr, search_params.frame.get_cuda_ptr(), s.x, s.z, s.heading.rad(), g.x, g.z,
                        g.heading.rad(), search_params.velocity_m_s, search_params.min_distance[0],
                        search_params.min_distance[1], search_params.timeout_ms, search_params.max_path_size_px,
                        search_params.distance_to_goal_tolerance_px, search_params.heading_error_tolerance.rad())
        _lib.check("set_plan_data")

    def set_seed(self, seed: int, stream: int = 0) -> None:
        _lib.load().fastrrt_set_seed(self._ptr, seed, stream)

    def set_density_region(self, side_px: int) -> None:
        """Side of the density regions of smart expansion (128 in the
        reference; 32 for 1024x1024 searches)."""
        _lib.load().fastrrt_set_density_region(self._ptr, int(side_px))
        _lib.check("set_density_region")

    def search_init(self, copy_intrinsic_costs_from_frame: bool = False) -> None:
        _lib.load().search_init(self._ptr, copy_intrinsic_costs_from_frame)
        _lib.check("search_init")

    def loop(self, smart: bool) -> bool:
        r = _lib.load().loop(self._ptr, smart)
        if not r:
            _lib.check("loop")
        return r

    def path_optimize(self) -> bool:
        r = _lib.load().path_optimize(self._ptr)
        if not r:
            _lib.check("path_optimize")
        return r

    def goal_reached(self) -> bool:
        r = _lib.load().goal_reached(self._ptr)
        if not r:
            _lib.check("goal_reached")
        return r

    def iterations(self) -> int:
        return _lib.load().fastrrt_iterations(self._ptr)

    def iteration_stats(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.load().fastrrt_iteration_stats(self._ptr, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def get_planned_path(self, interpolate: bool = False):
        L = _lib.load()
        ptr = L.interpolate_planned_path(self._ptr) if interpolate else L.get_planned_path(self._ptr)
        return _path_from_ptr(L, ptr)

    def interpolate_planned_path_p(self, path) -> list:
        """Hermite resampling of a caller's path (python_bind.cpp:157-171):
        `path` is an (n, 3) array of (x, z, heading)."""
        L = _lib.load()
        flat = np.ascontiguousarray(np.asarray(path, np.float32).reshape(-1))
        return _path_from_ptr(L, L.interpolate_planned_path_p(self._ptr, flat.ctypes.data_as(_lib.fp), flat.size))

    def build_ideal_curve(self, goal_x: int, goal_z: int, goal_heading: float):
        L = _lib.load()
        return _path_from_ptr(L, L.ideal_curve(self._ptr, goal_x, goal_z, goal_heading))

    def export_graph_nodes(self):
        L = _lib.load()
        ptr = L.export_graph_nodes(self._ptr)
        if not ptr:
            _lib.check("export_graph_nodes")
        size = ptr[0]
        if size == 0:
            L.release_export_graph_nodes(ptr)
            return None
        arr = np.ctypeslib.as_array(ptr, shape=(3 * size + 1,))[1:].reshape(size, 3).astype(np.int32).copy()
        L.release_export_graph_nodes(ptr)
        return arr

    def compute_region_debug_performance(self) -> None:
        _lib.load().compute_region_debug_performance(self._ptr)

    def save_current_graph_state(self, filename: str):
        _lib.load().save_current_graph_state(self._ptr, filename.encode("utf-8"))
        _lib.check("save_current_graph_state")

    def load_graph_state(self, filename: str):
        _lib.load().load_graph_state(self._ptr, filename.encode("utf-8"))
        _lib.check("load_graph_state")

    def export_state(self, goal=True):
        """Grid planes as (H, W) arrays; goal=False leaves out the goal-connection
        planes (exporting them finishes a lazily evaluated goal pass)."""
        return _export_state(_lib.load().fastrrt_export_state, self._ptr, self._w, self._h, goal)


def _export_state(fn, ptr, w, h, goal=True):
    n = w * h
    out = {k: np.zeros(n, np.int32) for k in ("parent_x", "parent_z", "type", "child_count")}
    out.update({k: np.zeros(n, np.float32) for k in ("heading", "cost", "goal_heading", "goal_cost")})
    out["goal_flag"] = np.zeros(n, np.uint8)
    fn(ptr, *(out[k].ctypes.data_as(_lib.i32p) for k in ("parent_x", "parent_z", "type", "child_count")),
       out["heading"].ctypes.data_as(_lib.fp), out["cost"].ctypes.data_as(_lib.fp),
       out["goal_flag"].ctypes.data_as(_lib.u8p) if goal else None,
       out["goal_heading"].ctypes.data_as(_lib.fp) if goal else None,
       out["goal_cost"].ctypes.data_as(_lib.fp) if goal else None)
    _lib.check("export_state")
    return {k: v.reshape(h, w) for k, v in out.items()}


# --------------------------------------------------------------------------
# CudaGraph (reference python_bind/app/pyfastrrt/src/cuda_graph.py)
# --------------------------------------------------------------------------

class CudaGraph:
    def __init__(self, width, height, perception_width_m, perception_height_m, max_steering_angle_deg,
                 vehicle_length_m, min_dist_x, min_dist_z, lower_bound_x, lower_bound_z, upper_bound_x,
                 upper_bound_z, path_costs, max_curvature=-1):
        L = _lib.load()
        path_costs = np.asarray(path_costs, np.float32)
        costs = np.ascontiguousarray(np.concatenate(([path_costs.shape[0]], path_costs)), dtype=np.float32)
        self._ptr = L.cudagraph_initialize(width, height, perception_width_m, perception_height_m,
                                           max_steering_angle_deg, vehicle_length_m, min_dist_x, min_dist_z,
                                           lower_bound_x, lower_bound_z, upper_bound_x, upper_bound_z,
                                           costs.ctypes.data_as(_lib.fp), max_curvature)
        if not self._ptr:
            _lib.check("cudagraph_initialize")
        self._w, self._h = width, height

    def __del__(self):
        try:
            if getattr(self, "_ptr", None):
                _lib.load().cudagraph_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass

    def can_connect_to_goal_using_hermite(self, frame: SearchFrame, x: int, z: int, goal_x: int, goal_z: int,
                                          goal_heading: float) -> bool:
        """cuda_graph.py's binding of python_cuda_graph_bind.cpp:147-153: the
        curve length comes from the frame's distance-to-goal channel."""
        r = _lib.load().can_connect_to_goal_using_hermite(self._ptr, frame.get_cuda_ptr(), x, z, goal_x, goal_z,
                                                          goal_heading)
        _lib.check("can_connect_to_goal_using_hermite")
        return bool(r)

    def can_connect_hermite_value(self, frame: SearchFrame, x: int, z: int, goal_x: int, goal_z: int,
                                  goal_heading: float):
        """(goal pixel blocked, curve result: cost, -1 or -2)."""
        out = ctypes.c_float(0.0)
        r = _lib.load().cudagraph_can_connect_hermite(self._ptr, frame.get_cuda_ptr(), x, z, goal_x, goal_z,
                                                      goal_heading, ctypes.byref(out))
        if r < 0:
            _lib.check("cudagraph_can_connect_hermite")
        return (r == 1, out.value)

    def compute_apf_repulsion(self, frame: SearchFrame, kr: float, radius: int) -> None:
        """Potential-field repulsion into the intrinsic costs (cuda_graph.py:276-278)."""
        _lib.load().compute_apf_repulsion(self._ptr, frame.get_cuda_ptr(), kr, radius)
        _lib.check("compute_apf_repulsion")

    def compute_apf_repulsion_noalloc(self, frame: SearchFrame, kr: float, radius: int) -> None:
        """compute_apf_repulsion with its scratch in the library's static
        device array: allocates nothing; raises when the grid and radius need
        more scratch than the array holds."""
        if not _lib.load().compute_apf_repulsion_noalloc(self._ptr, frame.get_cuda_ptr(), kr, radius):
            raise RuntimeError(f"compute_apf_repulsion_noalloc: {_lib.last_error()}")

    def compute_apf_attraction(self, frame: SearchFrame, ka: float, goal_x: int, goal_z: int) -> None:
        """Potential-field attraction into the intrinsic costs (cuda_graph.py:280-282)."""
        _lib.load().compute_apf_attraction(self._ptr, frame.get_cuda_ptr(), ka, goal_x, goal_z)
        _lib.check("compute_apf_attraction")

    def get_intrinsic_costs(self) -> np.ndarray:
        L = _lib.load()
        p = L.get_intrinsic_costs(self._ptr)
        res = np.ctypeslib.as_array(p, shape=(self._h * self._w,)).reshape(self._h, self._w).copy()
        L.destroy_intrinsic_costs_ptr(p)
        return res

    def add(self, x, z, a: angle, parent_x, parent_z, cost) -> None:
        _lib.load().add(self._ptr, x, z, a.rad(), parent_x, parent_z, cost)

    def add_temporary(self, x, z, a: angle, parent_x, parent_z, cost) -> None:
        _lib.load().add_temporary(self._ptr, x, z, a.rad(), parent_x, parent_z, cost)

    def derivate_node(self, frame: SearchFrame, steering_angle: angle, path_size: int, velocity: float,
                      parent_x: int, parent_z: int) -> None:
        _lib.load().derivate_node(self._ptr, frame.get_cuda_ptr(), steering_angle.rad(), path_size, velocity,
                                  parent_x, parent_z)
        _lib.check("derivate_node")

    def derivate_node_end(self, frame: SearchFrame, steering_rad: float, path_size: float, velocity: float,
                          x: int, z: int):
        out = (ctypes.c_int * 2)()
        _lib.load().cudagraph_derivate_node(self._ptr, frame.get_cuda_ptr(), steering_rad, path_size, velocity, x,
                                            z, out)
        _lib.check("derivate_node")
        return (out[0], out[1])

    def check_feasible_connection(self, frame: SearchFrame, start, end, velocity: int) -> bool:
        """CudaGraph::checkFeasibleConnection (kinematics.cu:308-326)."""
        r = _lib.load().cudagraph_check_feasible_connection(self._ptr, frame.get_cuda_ptr(), int(start[0]),
                                                            int(start[1]), int(end[0]), int(end[1]), int(velocity))
        _lib.check("cudagraph_check_feasible_connection")
        return bool(r)

    def check_feasible_connections(self, frame: SearchFrame, queries, velocity: int):
        """Many (start x, start z, end x, end z) queries in one launch; returns
        (feasible bool[n], final heading float64[n], path cost float32[n])."""
        q = np.ascontiguousarray(np.asarray(queries, np.int32).reshape(-1, 4))
        n = q.shape[0]
        ok = np.zeros(n, np.uint8)
        fh = np.zeros(n, np.float64)
        pc = np.zeros(n, np.float32)
        r = _lib.load().cudagraph_check_feasible_connections(
            self._ptr, frame.get_cuda_ptr(), q.ctypes.data_as(_lib.i32p), n, int(velocity),
            ok.ctypes.data_as(_lib.u8p), fh.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            pc.ctypes.data_as(_lib.fp))
        if r != n:
            raise RuntimeError(f"check_feasible_connections: {_lib.last_error()}")
        return ok.astype(bool), fh, pc

    def accept_derived_nodes(self) -> None:
        _lib.load().accept_derived_nodes(self._ptr, 0, 0, 0.0)

    def check_in_graph(self, x, z) -> bool:
        return _lib.load().check_in_graph(self._ptr, x, z)

    def get_heading(self, x, z) -> angle:
        return angle.new_rad(_lib.load().get_heading(self._ptr, x, z))

    def clear(self) -> None:
        _lib.load().clear(self._ptr)

    def list_all(self) -> np.ndarray:
        L = _lib.load()
        ptr = L.list_all(self._ptr)
        count = int(ptr[0])
        res = np.ctypeslib.as_array(ptr, shape=(4 * count + 1,))[1:].reshape(count, 4).copy()
        L.free_list_all(ptr)
        return res

    def solve_collisions(self) -> None:
        _lib.load().solve_collisions(self._ptr)

    def process_direct_goal_connection(self, search_frame: SearchFrame, goal_x, goal_z, goal_heading: angle,
                                       max_curvature=-1) -> None:
        _lib.load().process_direct_goal_connection(self._ptr, search_frame.get_cuda_ptr(), goal_x, goal_z,
                                                   goal_heading.rad(), max_curvature)
        _lib.check("process_direct_goal_connection")

    def is_directly_connected_to_goal(self, x, z) -> bool:
        return _lib.load().is_directly_connected_to_goal(self._ptr, x, z)

    def direct_connection_to_goal_cost(self, x, z) -> float:
        return _lib.load().direct_connection_to_goal_cost(self._ptr, x, z)

    def direct_connection_to_goal_heading(self, x, z) -> angle:
        return angle.new_rad(_lib.load().direct_connection_to_goal_heading(self._ptr, x, z))

    def dump_graph_to_file(self, file: str) -> None:
        _lib.load().dump_graph_to_file(self._ptr, file.encode("utf-8"))
        _lib.check("dump_graph_to_file")

    def read_from_dump_file(self, file: str) -> None:
        _lib.load().read_from_dump_file(self._ptr, file.encode("utf-8"))
        _lib.check("read_from_dump_file")

    def get_parent(self, x, z):
        L = _lib.load()
        p = L.get_parent(self._ptr, x, z)
        res = (int(p[0]), int(p[1]))
        L.free_parent_data(p)
        return res

    def get_cost(self, x, z):
        return _lib.load().get_cost(self._ptr, x, z)

    def count_all(self) -> int:
        return _lib.load().count_all(self._ptr)

    # test hooks of the reference's C++ unit tests (graph.h)
    def set_collision(self, x, z, new_parent_x, new_parent_z, heading_rad, cost):
        _lib.load().cudagraph_set_collision(self._ptr, x, z, new_parent_x, new_parent_z, heading_rad, cost)

    def get_type(self, x, z) -> int:
        return _lib.load().cudagraph_get_type(self._ptr, x, z)

    def get_child_count(self, x, z) -> int:
        return _lib.load().cudagraph_get_child_count(self._ptr, x, z)

    def count(self, node_type=1) -> int:
        return _lib.load().cudagraph_count(self._ptr, node_type)

    def check_graph_is_consistent(self) -> bool:
        return _lib.load().cudagraph_check_consistent(self._ptr)

    def find_best_node(self, frame, heading_rad, radius, x, z, max_heading_error):
        out = (ctypes.c_int * 2)()
        _lib.load().cudagraph_find_best_node(self._ptr, frame.get_cuda_ptr() if frame else None, heading_rad,
                                             radius, x, z, max_heading_error, out)
        _lib.check("find_best_node")
        return (out[0], out[1])

    def check_goal_reached(self, goal_x, goal_z, heading_rad, tolerance, max_heading_error) -> bool:
        return _lib.load().cudagraph_check_goal_reached(self._ptr, goal_x, goal_z, heading_rad, tolerance,
                                                        max_heading_error)

    def export_state(self):
        return _export_state(_lib.load().cudagraph_export_state, self._ptr, self._w, self._h)

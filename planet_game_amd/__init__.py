"""planet_game_amd — MI355X (gfx950) implementation of PlanetSim's implicit
diffusive heat-transport path.

The product is the C ABI of ``include/psim.h`` implemented by the in-tree
``libpsim.so`` (host C++20 + hand-written HIP kernels). This module is a thin
ctypes binding of that ABI for Python callers (tests, bench.py). It does no
numerical work itself, and it raises at import when the library is missing
instead of falling back to anything.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import types
import weakref

import numpy as np

__all__ = [
    "LIB_PATH", "lib", "PsimError", "Mesh", "Context", "FV", "Graph", "System", "SolveInfo",
    "HOST", "DEVICE", "F64", "F32", "PLAIN_FP64", "MIXED", "instance_seed", "heat_instance",
    "heat_forcing", "Hierarchy", "Batch", "DeviceArray", "Server", "ServeRequest", "ServeRecord",
    "implicit_transport", "TransportSettings", "TransportResult", "Agglomeration", "Groups",
    "ResponseTable", "implicit_transport_table", "implicit_transport_table_device", "Ilu", "TracerResult",
    "tracer_step", "pcg_operator", "operator_solve", "Preconditioner", "OperatorInfo", "FORM_TRANSPORT",
    "FORM_BALANCE", "advection_source", "implicit_advection_diffusion", "AdvectionResult",
]

# PSIM_LIBRARY names another build of the same ABI (e.g. a tuning variant).
LIB_PATH = os.environ.get("PSIM_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpsim.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C planet_game_amd)")

lib = C.CDLL(LIB_PATH)

# Live native objects by class. At interpreter exit they are released here,
# dependants first, while ctypes and the HIP runtime are still whole (object
# finalisation during interpreter teardown is too late to call into them).
_LIVE = {}


def _track(obj):
    _LIVE.setdefault(type(obj).__name__, weakref.WeakSet()).add(obj)
    return obj


@atexit.register
def _release_all():
    for name in ("Server", "Batch", "System", "Groups", "Ilu", "FV", "Graph", "DeviceArray", "Context", "Agglomeration",
                 "Mesh"):
        for obj in list(_LIVE.get(name, ())):
            try:
                obj.close()
            except Exception:  # noqa: BLE001 - best effort at exit
                pass

HOST, DEVICE = 0, 1
F64, F32 = 0, 1
PLAIN_FP64, MIXED = 0, 1
STATUS = {0: "PSIM_OK", 1: "PSIM_INVALID_ARGUMENT", 2: "PSIM_LOGIC_ERROR", 3: "PSIM_DEVICE_ERROR",
          4: "PSIM_OUT_OF_MEMORY", 5: "PSIM_NOT_CONVERGED", 6: "PSIM_RUNTIME_ERROR"}

_vp = C.c_void_p
_u64 = C.c_uint64
_i32 = C.c_int32
_dbl = C.c_double
_pd = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_pf = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_pu32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_pu64 = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_pu8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class DeviceInfo(C.Structure):
    _fields_ = [("compute_units", C.c_int32), ("program_blocks_per_cu", C.c_int32), ("cooperative", C.c_int32),
                ("reserved", C.c_int32), ("lds_per_block", C.c_uint64), ("lds_per_cu", C.c_uint64),
                ("hbm_bytes", C.c_uint64), ("wall_clock_hz", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("restarts", C.c_int32), ("status", C.c_int32),
                ("v_cycles", C.c_int32), ("relative_residual", C.c_double), ("device_ms", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("psim_last_error", C.c_char_p)
_sig("psim_version", C.c_int)
_sig("psim_mesh_create", C.c_int, C.c_uint32, _dbl, C.POINTER(_vp))
_sig("psim_mesh_destroy", None, _vp)
_sig("psim_mesh_counts", C.c_int, _vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64))
_sig("psim_mesh_cells", C.c_int, _vp, _pd, _pd, _pd, _pd, _pu32, _pu8)
_sig("psim_mesh_cell_edges", C.c_int, _vp, _pu32, _pu32, _pf, _pf)
_sig("psim_mesh_edges", C.c_int, _vp, _pu32, _pu32, _pd, _pd)
_sig("psim_mesh_graph", C.c_int, _vp, _pd, _pu64, _pu64, _pd, _pd)
_sig("psim_agglomeration_create", C.c_int, _vp, C.POINTER(_vp))
_sig("psim_agglomeration_destroy", None, _vp)
_sig("psim_agglomeration_counts", C.c_int, _vp, C.POINTER(_u64), C.POINTER(_u64))
_sig("psim_agglomeration_groups", C.c_int, _vp, _pu32, _pd)
_sig("psim_agglomeration_mesh", _vp, _vp)
_sig("psim_hierarchy_create", C.c_int, _u64, _pu64, _pu64, _pd, _pd, C.POINTER(_vp))
_sig("psim_hierarchy_destroy", None, _vp)
_sig("psim_hierarchy_depth", _u64, _vp)
_sig("psim_hierarchy_level_counts", C.c_int, _vp, _u64, C.POINTER(_u64), C.POINTER(_u64))
_sig("psim_hierarchy_level", C.c_int, _vp, _u64, _pu64, _pu64, _pu64, _pu64, _pd, _pd)
_sig("psim_instance_seed", _u64, _u64, _u64)
_sig("psim_heat_instance", C.c_int, _vp, _u64, _dbl, _dbl, _pd, _pd, _pu8, _pd)
_sig("psim_heat_forcing", C.c_int, _vp, _u64, C.c_uint32, _pd)
_sig("psim_ctx_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("psim_ctx_destroy", None, _vp)
_sig("psim_ctx_synchronize", C.c_int, _vp)
_sig("psim_device_alloc", C.c_int, _vp, _u64, C.POINTER(_vp))
_sig("psim_device_free", C.c_int, _vp, _vp)
_sig("psim_copy_to_device", C.c_int, _vp, _vp, _vp, _u64)
_sig("psim_copy_to_host", C.c_int, _vp, _vp, _vp, _u64)
_sig("psim_device_copy", C.c_int, _vp, _vp, _vp, _u64)
_sig("psim_ctx_info", C.c_int, _vp, C.POINTER(DeviceInfo))
_sig("psim_ctx_events", C.c_int, _vp, C.c_uint32)
_sig("psim_ctx_record", C.c_int, _vp, C.c_uint32)
_sig("psim_ctx_elapsed", C.c_int, _vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_float))
_sig("psim_system_last_info", C.c_int, _vp, C.POINTER(SolveInfo))
_sig("psim_fv_create", C.c_int, _vp, _vp, C.POINTER(_vp))
_sig("psim_fv_create_from_arrays", C.c_int, _vp, _u64, _u64, _pd, _pu32, _pu8, _pu32, _pu32, _pf, _pf,
     _pu32, _pu32, _pd, _pd, C.POINTER(_vp))
_sig("psim_fv_destroy", None, _vp)
_sig("psim_edge_orientation", C.c_int, _vp, C.c_uint32, C.c_uint32, C.POINTER(_dbl))
_sig("psim_gradient", C.c_int, _vp, C.c_int, _u64, _vp, _vp, _vp, C.c_int)
_sig("psim_divergence", C.c_int, _vp, C.c_int, _u64, _vp, _vp, C.c_int)
_sig("psim_laplacian", C.c_int, _vp, C.c_int, _u64, _vp, _vp, C.c_int)
_sig("psim_graph_create", C.c_int, _vp, _u64, _pd, _pu64, _pu64, _pd, _pd, C.POINTER(_vp))
_sig("psim_graph_destroy", None, _vp)
_sig("psim_graph_size", _u64, _vp)
_sig("psim_graph_entries", _u64, _vp)
_sig("psim_graph_depth", _u64, _vp)
_sig("psim_diffusion_source", C.c_int, _vp, _dbl, _vp, _vp, C.c_int)
_sig("psim_ilu_create", C.c_int, _vp, _u64, _pu64, _pu64, _pu64, _u64, C.POINTER(_vp))
_sig("psim_ilu_destroy", None, _vp)
_sig("psim_ilu_factor", C.c_int, _vp, _vp, _vp, _dbl, C.c_int)
_sig("psim_ilu_solve", C.c_int, _vp, _vp, C.c_int)
_sig("psim_groups_create", C.c_int, _vp, _u64, _u64, _pu32, _pd, _pd, C.POINTER(_vp))
_sig("psim_groups_destroy", None, _vp)
for _name in ("psim_groups_mean", "psim_groups_max", "psim_groups_broadcast"):
    _sig(_name, C.c_int, _vp, C.c_uint32, _vp, _vp, C.c_int)
_sig("psim_system_create", C.c_int, _vp, C.c_int, C.POINTER(_vp))
_sig("psim_system_destroy", None, _vp)
_sig("psim_system_set_workgroups", C.c_int, _vp, _i32)
_sig("psim_system_workgroups", _i32, _vp)
_sig("psim_system_set_pre_sweeps", C.c_int, _vp, _i32)
_sig("psim_system_pre_sweeps", _i32, _vp)
_sig("psim_system_level_values", C.c_int, _vp, _u64, _vp, _vp)
_sig("psim_system_trace", C.c_int, _vp, _i32)
_sig("psim_system_trace_read", C.c_int, _vp, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"), _i32,
     C.POINTER(_i32))
_sig("psim_system_set_matrix", C.c_int, _vp, _vp, _u64, _vp, _u64, C.c_int)
_sig("psim_system_precondition", C.c_int, _vp, _vp, _vp, C.c_int)
_sig("psim_system_apply", C.c_int, _vp, _vp, _vp, C.c_int)
_sig("psim_system_solve", C.c_int, _vp, _vp, _vp, _dbl, _i32, C.POINTER(SolveInfo), C.c_int)
_sig("psim_heat_set_matrix", C.c_int, _vp, _vp, _vp, _dbl, C.c_int)
_sig("psim_heat_step", C.c_int, _vp, _vp, _vp, _dbl, _dbl, _i32, _vp, C.POINTER(SolveInfo), C.c_int)
_sig("psim_heat_chain", C.c_int, _vp, _vp, _vp, _vp, _vp, _i32, _dbl, _dbl, _i32, _vp, _vp, C.c_int)
_sig("psim_heat_step_noalloc", C.c_int, _vp, _vp, _vp, _dbl, _dbl, _i32, _vp, _vp)
_pp = C.POINTER(_vp)


class ServeRequest(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("graph", C.c_int32), ("session", C.c_int32), ("first", C.c_int32),
                ("last", C.c_int32), ("conductance_per_entry", _vp), ("capacity_area", _vp),
                ("temperature_old", _vp), ("forcing_W_m2", _vp), ("temperature_new", _vp),
                ("matrix_key", C.c_uint64)]


class ServeRecord(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("submitted_s", C.c_double), ("completed_s", C.c_double),
                ("device_ms", C.c_double), ("status", C.c_int32), ("iterations", C.c_int32),
                ("restarts", C.c_int32), ("v_cycles", C.c_int32), ("relative_residual", C.c_double)]


RESPONSE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                       C.c_uint64)


class TransportSettings(C.Structure):
    _fields_ = [("max_newton_iterations", C.c_int32), ("max_line_search_halvings", C.c_int32),
                ("rounding_floor_W_m2", C.c_double), ("max_non_monotone_steps", C.c_int32), ("pad", C.c_int32),
                ("max_non_monotone_growth", C.c_double), ("newton_tolerance_W_m2", C.c_double),
                ("cg_relative_tolerance", C.c_double), ("max_cg_iterations", C.c_int32), ("pad2", C.c_int32)]


class TransportResult(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("cg_iterations", C.c_int32),
                ("consistency_residual_W_m2", C.c_double), ("sum_W", C.c_double), ("absolute_sum_W", C.c_double),
                ("dissipation_W_K", C.c_double)]


_sig("psim_transport_settings_default", None, C.POINTER(TransportSettings))
_sig("psim_implicit_transport", C.c_int, _vp, _dbl, _pd, RESPONSE, _vp, C.POINTER(TransportSettings), _vp, _pd, _pd,
     C.POINTER(TransportResult))
_sig("psim_system_pcg", C.c_int, _vp, _vp, _vp, _vp, _dbl, _i32, C.POINTER(SolveInfo), C.c_int)


class ResponseTable(C.Structure):
    _fields_ = [("points", C.c_uint32), ("reserved", C.c_uint32), ("source_W_m2", _vp), ("mean_K", _vp)]


_sig("psim_implicit_transport_table", C.c_int, _vp, _dbl, _vp, C.POINTER(ResponseTable), C.POINTER(TransportSettings),
     _vp, _vp, _vp, C.POINTER(TransportResult), C.c_int)
_sig("psim_server_create", C.c_int, _vp, _pp, C.POINTER(_dbl), C.POINTER(_i32), _i32, C.c_int, _u64, C.POINTER(_vp))
_sig("psim_server_create_occupancy", C.c_int, _vp, _pp, C.POINTER(_dbl), C.POINTER(_i32), _i32, C.c_int, _u64, _i32,
     C.POINTER(_vp))
_sig("psim_server_destroy", None, _vp)
_sig("psim_server_plan", C.c_int, _vp, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_u64))
_sig("psim_server_run", C.c_int, _vp, C.POINTER(ServeRequest), _i32, _dbl, _dbl, _i32, C.POINTER(ServeRecord))
_sig("psim_batch_create", C.c_int, _vp, _pp, _i32, C.POINTER(_vp))
_sig("psim_batch_destroy", None, _vp)
_sig("psim_batch_size", C.c_int, _vp, C.POINTER(_i32), C.POINTER(_i32))
_sig("psim_batch_groups", C.c_int, _vp, C.POINTER(_i32))
_sig("psim_batch_set_occupancy", C.c_int, _vp, _i32)
_sig("psim_batch_occupancy", _i32, _vp)
_sig("psim_batch_set_groups", C.c_int, _vp, C.POINTER(_i32))
_sig("psim_batch_heat_step", C.c_int, _vp, _pp, _pp, _pp, _pp, _dbl, _dbl, _i32, _pp, C.POINTER(SolveInfo))
_sig("psim_batch_synchronize", C.c_int, _vp)
_sig("psim_batch_autotune", C.c_int, _vp, _pp, _pp, _pp, _pp, _dbl, _dbl, _i32, _pp, _i32, C.POINTER(_dbl))


class PsimError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def check(status):
    if status != 0:
        raise PsimError(status, lib.psim_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_vp) if isinstance(a, np.ndarray) else a


class Mesh:
    """make_icosphere(subdivision, radius) on the host, with its arrays."""

    def __init__(self, subdivision: int, radius: float = 6371000.0, _view=None):
        """_view: (handle, owner) of a mesh owned by another native object."""
        if _view is None:
            h = _vp()
            check(lib.psim_mesh_create(subdivision, radius, C.byref(h)))
            self._owner = None
        else:
            h, self._owner = _vp(_view[0]), _view[1]
        self.handle = h
        if self._owner is None:
            _track(self)
        self.subdivision = subdivision
        self.radius = radius
        n, e, d = _u64(), _u64(), _u64()
        check(lib.psim_mesh_counts(h, C.byref(n), C.byref(e), C.byref(d)))
        self.cells, self.edges, self.directed = n.value, e.value, d.value
        c = self.cells
        self.center, self.east, self.north = (np.zeros(3 * c) for _ in range(3))
        self.area = np.zeros(c)
        self.edge_offset = np.zeros(c, np.uint32)
        self.edge_count = np.zeros(c, np.uint8)
        check(lib.psim_mesh_cells(h, self.center, self.east, self.north, self.area,
                                  self.edge_offset, self.edge_count))
        self.neighbor = np.zeros(self.directed, np.uint32)
        self.edge = np.zeros(self.directed, np.uint32)
        self.normal_east = np.zeros(self.directed, np.float32)
        self.normal_north = np.zeros(self.directed, np.float32)
        check(lib.psim_mesh_cell_edges(h, self.neighbor, self.edge, self.normal_east,
                                       self.normal_north))
        self.first = np.zeros(self.edges, np.uint32)
        self.second = np.zeros(self.edges, np.uint32)
        self.length = np.zeros(self.edges)
        self.distance = np.zeros(self.edges)
        check(lib.psim_mesh_edges(h, self.first, self.second, self.length, self.distance))
        self.offset = np.zeros(c + 1, np.uint64)
        self.graph_neighbour = np.zeros(self.directed, np.uint64)
        self.weight = np.zeros(self.directed)
        self.weight_sum = np.zeros(c)
        g_area = np.zeros(c)
        check(lib.psim_mesh_graph(h, g_area, self.offset, self.graph_neighbour, self.weight,
                                  self.weight_sum))

    def entry_values(self, per_edge):
        """Expand a per-edge array to the CSR entries (cell-edge order)."""
        return np.ascontiguousarray(per_edge[self.edge])

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None) and getattr(self, "_owner", None) is None:
            lib.psim_mesh_destroy(self.handle)
        self.handle = None

    def __del__(self):
        self.close()


class Agglomeration:
    """build_agglomeration of a mesh (agglomerated_transport_graph and
    agglomerated_mesh): each cell's group, a cell of the mesh one level
    coarser (`coarse`), the groups' areas, and `graph`, the agglomerated
    transport graph (the coarse mesh's graph with the groups' areas), which
    Graph() accepts as it accepts a Mesh."""

    def __init__(self, mesh: Mesh):
        h = _vp()
        check(lib.psim_agglomeration_create(mesh.handle, C.byref(h)))
        self.handle = h
        _track(self)
        n, g = _u64(), _u64()
        check(lib.psim_agglomeration_counts(h, C.byref(n), C.byref(g)))
        self.cells, self.groups = n.value, g.value
        self.group_of_cell = np.zeros(self.cells, np.uint32)
        self.group_area = np.zeros(self.groups)
        check(lib.psim_agglomeration_groups(h, self.group_of_cell, self.group_area))
        self.coarse = Mesh(max(mesh.subdivision - 1, 0), mesh.radius,
                           _view=(lib.psim_agglomeration_mesh(h), self))
        c = self.coarse
        self.graph = types.SimpleNamespace(cells=self.groups, area=self.group_area, offset=c.offset,
                                           graph_neighbour=c.graph_neighbour, weight=c.weight,
                                           weight_sum=c.weight_sum)

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            self.coarse.close()
            lib.psim_agglomeration_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Hierarchy:
    """Host view of build_hierarchy for parity checks."""

    def __init__(self, mesh: Mesh):
        h = _vp()
        check(lib.psim_hierarchy_create(mesh.cells, mesh.offset, mesh.graph_neighbour, mesh.weight,
                                        mesh.weight_sum, C.byref(h)))
        try:
            depth = lib.psim_hierarchy_depth(h)
            self.levels = []
            for j in range(depth):
                s, z = _u64(), _u64()
                check(lib.psim_hierarchy_level_counts(h, j, C.byref(s), C.byref(z)))
                s, z = s.value, z.value
                last = j + 1 == depth
                lv = {"size": s, "nnz": z, "offset": np.zeros(s + 1, np.uint64),
                      "column": np.zeros(max(z, 1), np.uint64),
                      "aggregate": np.zeros(max(0 if last else s, 1), np.uint64),
                      "target": np.zeros(max(0 if last else z, 1), np.uint64),
                      "diagonal": np.zeros(s), "value": np.zeros(max(z, 1))}
                check(lib.psim_hierarchy_level(h, j, lv["offset"], lv["column"], lv["aggregate"],
                                               lv["target"], lv["diagonal"], lv["value"]))
                lv["column"] = lv["column"][:z]
                lv["value"] = lv["value"][:z]
                lv["aggregate"] = lv["aggregate"][: (0 if last else s)]
                lv["target"] = lv["target"][: (0 if last else z)]
                self.levels.append(lv)
        finally:
            lib.psim_hierarchy_destroy(h)


def instance_seed(index: int, base: int = 20261203) -> int:
    return lib.psim_instance_seed(base, index)


def heat_instance(mesh: Mesh, seed: int, k_base: float = 1.0e6, sigma: float = 0.3):
    """(conductance per edge, C*A, land mask, T_old) of one seeded instance."""
    cond = np.zeros(mesh.edges)
    ca = np.zeros(mesh.cells)
    land = np.zeros(mesh.cells, np.uint8)
    temp = np.zeros(mesh.cells)
    check(lib.psim_heat_instance(mesh.handle, seed, k_base, sigma, cond, ca, land, temp))
    return cond, ca, land, temp


def heat_forcing(mesh: Mesh, seed: int, step: int):
    f = np.zeros(mesh.cells)
    check(lib.psim_heat_forcing(mesh.handle, seed, step, f))
    return f


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        check(lib.psim_ctx_create(device, C.byref(h)))
        self.handle = h
        _track(self)

    def synchronize(self):
        check(lib.psim_ctx_synchronize(self.handle))

    def info(self):
        d = DeviceInfo()
        check(lib.psim_ctx_info(self.handle, C.byref(d)))
        return {k: getattr(d, k) for k, _ in DeviceInfo._fields_ if k != "reserved"}

    def events(self, count):
        check(lib.psim_ctx_events(self.handle, count))

    def record(self, index):
        check(lib.psim_ctx_record(self.handle, index))

    def copy(self, dst, src, nbytes):
        """Queue a device-to-device copy through the library's streaming kernel."""
        check(lib.psim_device_copy(self.handle, dst, src, nbytes))

    def elapsed_ms(self, start, end):
        ms = C.c_float()
        check(lib.psim_ctx_elapsed(self.handle, start, end, C.byref(ms)))
        return ms.value

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class DeviceArray:
    """A device buffer owned by the caller (resident inputs/outputs)."""

    def __init__(self, ctx: Context, count: int, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = count
        self.nbytes = count * self.dtype.itemsize
        p = _vp()
        check(lib.psim_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p
        _track(self)

    @classmethod
    def of(cls, ctx, array):
        a = np.ascontiguousarray(array)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, array):
        a = np.ascontiguousarray(array, dtype=self.dtype)
        assert a.size == self.count
        check(lib.psim_copy_to_device(self.ctx.handle, self.ptr, a.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        a = np.empty(self.count, self.dtype)
        check(lib.psim_copy_to_host(self.ctx.handle, a.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return a

    def offset(self, elements: int):
        return _vp(self.ptr.value + elements * self.dtype.itemsize)

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "ptr", None):
            lib.psim_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class FV:
    """Device FV operators on one mesh (finite_volume.hpp:21-44)."""

    def __init__(self, ctx: Context, mesh: Mesh):
        h = _vp()
        check(lib.psim_fv_create(ctx.handle, mesh.handle, C.byref(h)))
        self.handle, self.ctx, self.mesh = h, ctx, mesh
        _track(self)

    @staticmethod
    def _dtype(a):
        return F64 if a.dtype == np.float64 else F32

    def gradient(self, scalar):
        s = np.ascontiguousarray(scalar)
        fields = s.size // self.mesh.cells
        e, n = np.empty_like(s), np.empty_like(s)
        check(lib.psim_gradient(self.handle, self._dtype(s), fields, _ptr(s), _ptr(e), _ptr(n), HOST))
        return e, n

    def divergence(self, flux):
        f = np.ascontiguousarray(flux)
        fields = f.size // self.mesh.edges
        out = np.empty(fields * self.mesh.cells, f.dtype)
        check(lib.psim_divergence(self.handle, self._dtype(f), fields, _ptr(f), _ptr(out), HOST))
        return out

    def laplacian(self, scalar):
        s = np.ascontiguousarray(scalar)
        fields = s.size // self.mesh.cells
        out = np.empty_like(s)
        check(lib.psim_laplacian(self.handle, self._dtype(s), fields, _ptr(s), _ptr(out), HOST))
        return out

    def gradient_device(self, dtype, fields, scalar, east, north):
        """Queue the gradient of `fields` fields on device pointers (no synchronisation)."""
        check(lib.psim_gradient(self.handle, dtype, fields, scalar, east, north, DEVICE))

    def divergence_device(self, dtype, fields, flux, out):
        check(lib.psim_divergence(self.handle, dtype, fields, flux, out, DEVICE))

    def laplacian_device(self, dtype, fields, scalar, out):
        check(lib.psim_laplacian(self.handle, dtype, fields, scalar, out, DEVICE))

    def orientation(self, cell, edge):
        out = _dbl()
        check(lib.psim_edge_orientation(self.handle, cell, edge, C.byref(out)))
        return out.value

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_fv_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Graph:
    """Device transport graph + multigrid hierarchy structure."""

    def __init__(self, ctx: Context, mesh: Mesh):
        h = _vp()
        check(lib.psim_graph_create(ctx.handle, mesh.cells, mesh.area, mesh.offset,
                                    mesh.graph_neighbour, mesh.weight, mesh.weight_sum, C.byref(h)))
        self.handle, self.ctx, self.mesh = h, ctx, mesh
        _track(self)
        self.size = lib.psim_graph_size(h)
        self.entries = lib.psim_graph_entries(h)
        self.depth = lib.psim_graph_depth(h)

    def diffusion_source(self, conductance, temperature):
        t = np.ascontiguousarray(temperature, np.float64)
        out = np.empty_like(t)
        check(lib.psim_diffusion_source(self.handle, conductance, _ptr(t), _ptr(out), HOST))
        return out

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Groups:
    """psim_groups: a group map's restriction (area-weighted mean, highest
    value) and broadcast on the device. Host arrays are (fields, n) or (n,);
    the *_device forms take device pointers."""

    def __init__(self, ctx: Context, group_of_cell, cell_area, group_area):
        g = np.ascontiguousarray(group_of_cell, np.uint32)
        ca = np.ascontiguousarray(cell_area, np.float64)
        ga = np.ascontiguousarray(group_area, np.float64)
        h = _vp()
        check(lib.psim_groups_create(ctx.handle, g.size, ga.size, g, ca, ga, C.byref(h)))
        self.handle, self.ctx = h, ctx
        _track(self)
        self.cells, self.groups = g.size, ga.size

    @classmethod
    def of(cls, ctx: Context, agglomeration: "Agglomeration", mesh: Mesh):
        return cls(ctx, agglomeration.group_of_cell, mesh.area, agglomeration.group_area)

    def _host(self, fn, values, n_in, n_out):
        v = np.ascontiguousarray(values, np.float64)
        if v.shape[-1] != n_in:
            raise ValueError("field length does not match the map")
        fields = v.size // n_in
        out = np.empty(v.shape[:-1] + (n_out,))
        check(fn(self.handle, fields, _ptr(v), _ptr(out), HOST))
        return out

    def mean(self, values):
        return self._host(lib.psim_groups_mean, values, self.cells, self.groups)

    def max(self, values):
        return self._host(lib.psim_groups_max, values, self.cells, self.groups)

    def broadcast(self, values):
        return self._host(lib.psim_groups_broadcast, values, self.groups, self.cells)

    def mean_device(self, cell_values, group_values, fields=1):
        check(lib.psim_groups_mean(self.handle, fields, cell_values, group_values, DEVICE))

    def max_device(self, cell_values, group_values, fields=1):
        check(lib.psim_groups_max(self.handle, fields, cell_values, group_values, DEVICE))

    def broadcast_device(self, group_values, cell_values, fields=1):
        check(lib.psim_groups_broadcast(self.handle, fields, group_values, cell_values, DEVICE))

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_groups_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class System:
    """A linear system on a graph: GraphMultigrid + BiCGSTAB on the device."""

    def __init__(self, graph: Graph, precision: int = PLAIN_FP64):
        h = _vp()
        check(lib.psim_system_create(graph.handle, precision, C.byref(h)))
        self.handle, self.graph, self.precision = h, graph, precision
        _track(self)

    @property
    def workgroups(self):
        return lib.psim_system_workgroups(self.handle)

    @workgroups.setter
    def workgroups(self, value):
        check(lib.psim_system_set_workgroups(self.handle, value))

    @property
    def pre_sweeps(self):
        """Bit j set: level j pre-smooths with two Jacobi sweeps in solves."""
        return lib.psim_system_pre_sweeps(self.handle)

    @pre_sweeps.setter
    def pre_sweeps(self, mask):
        check(lib.psim_system_set_pre_sweeps(self.handle, mask))

    def trace(self, capacity):
        check(lib.psim_system_trace(self.handle, capacity))

    def trace_read(self, capacity=4096):
        ticks = np.zeros(capacity, np.uint64)
        count = _i32()
        check(lib.psim_system_trace_read(self.handle, ticks, capacity, C.byref(count)))
        return ticks[: min(count.value, capacity)]

    def set_matrix(self, diag, off):
        d = np.ascontiguousarray(diag, np.float64)
        o = np.ascontiguousarray(off, np.float64)
        check(lib.psim_system_set_matrix(self.handle, _ptr(d), d.size, _ptr(o), o.size, HOST))

    def level_values(self, level):
        n = self.graph.mesh.cells if level == 0 else None
        # sizes come from the host hierarchy view
        return self._level_values(level)

    def _level_values(self, level):
        if not hasattr(self, "_hier"):
            self._hier = Hierarchy(self.graph.mesh)
        lv = self._hier.levels[level]
        d = np.zeros(lv["size"])
        o = np.zeros(max(lv["nnz"], 1))
        check(lib.psim_system_level_values(self.handle, level, _ptr(d), _ptr(o)))
        return d, o[: lv["nnz"]]

    def precondition(self, rhs):
        r = np.ascontiguousarray(rhs, np.float64)
        out = np.empty_like(r)
        check(lib.psim_system_precondition(self.handle, _ptr(r), _ptr(out), HOST))
        return out

    def apply(self, x):
        v = np.ascontiguousarray(x, np.float64)
        out = np.empty_like(v)
        check(lib.psim_system_apply(self.handle, _ptr(v), _ptr(out), HOST))
        return out

    def solve(self, b, tolerance=1e-10, max_iterations=1000):
        bb = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(bb)
        info = SolveInfo()
        check(lib.psim_system_solve(self.handle, _ptr(bb), _ptr(x), tolerance, max_iterations,
                                    C.byref(info), HOST))
        return x, info

    def heat_set_matrix(self, conductance_per_entry, capacity_area, dt=2.63e6):
        k = np.ascontiguousarray(conductance_per_entry, np.float64)
        ca = np.ascontiguousarray(capacity_area, np.float64)
        check(lib.psim_heat_set_matrix(self.handle, _ptr(k), _ptr(ca), dt, HOST))

    def heat_set_matrix_device(self, conductance_per_entry: "DeviceArray", capacity_area: "DeviceArray",
                               dt=2.63e6):
        check(lib.psim_heat_set_matrix(self.handle, conductance_per_entry.ptr, capacity_area.ptr, dt,
                                       DEVICE))

    def heat_step_device(self, t_old, forcing, t_new, dt=2.63e6, tolerance=1e-10, max_iterations=1000):
        """Queue one solve on device-resident vectors (no synchronisation)."""
        check(lib.psim_heat_step(self.handle, t_old, forcing, dt, tolerance, max_iterations, t_new,
                                 None, DEVICE))

    def heat_step_noalloc(self, t_old, forcing, t_new, dt=2.63e6, tolerance=1e-10, max_iterations=1000):
        """One solve on device-resident vectors with the solver's workspace in
        the library's static device array (no allocation); returns SolveInfo."""
        info = SolveInfo()
        check(lib.psim_heat_step_noalloc(self.handle, t_old, forcing, dt, tolerance, max_iterations, t_new,
                                         C.byref(info)))
        return info

    def heat_chain(self, t_old, forcings, dt=2.63e6, tolerance=1e-10, max_iterations=1000,
                   conductance_per_entry=None, capacity_area=None):
        """Chained steps on the device in one launch; forcings is steps x n.
        With conductances and C*A given, the launch also sets the matrix."""
        t = np.ascontiguousarray(t_old, np.float64)
        f = np.ascontiguousarray(forcings, np.float64)
        steps = f.size // t.size
        out = np.empty_like(t)
        infos = (SolveInfo * steps)()
        k = None if conductance_per_entry is None else np.ascontiguousarray(conductance_per_entry, np.float64)
        ca = None if capacity_area is None else np.ascontiguousarray(capacity_area, np.float64)
        check(lib.psim_heat_chain(self.handle, _ptr(k) if k is not None else None,
                                  _ptr(ca) if ca is not None else None, _ptr(t), _ptr(f), steps, dt, tolerance,
                                  max_iterations, _ptr(out), C.cast(infos, _vp), HOST))
        return out, list(infos)

    def heat_chain_device(self, t_old, forcings, steps, t_new, dt=2.63e6, tolerance=1e-10,
                          max_iterations=1000, step_infos=None, conductance_per_entry=None, capacity_area=None):
        """Queue a chain on device-resident vectors (DeviceArray pointers);
        synchronises only when step_infos (a ctypes SolveInfo array) is given."""
        check(lib.psim_heat_chain(self.handle, conductance_per_entry, capacity_area, t_old, forcings, steps, dt,
                                  tolerance, max_iterations, t_new,
                                  C.cast(step_infos, _vp) if step_infos is not None else None, DEVICE))

    def last_info(self):
        info = SolveInfo()
        check(lib.psim_system_last_info(self.handle, C.byref(info)))
        return info

    def pcg(self, b, hold=None, tolerance=1e-6, max_iterations=2000):
        """Newton-system PCG on the matrix set by set_matrix (host arrays);
        hold: rows kept at 0 (identity rows of the matrix, b = 0 there)."""
        bb = np.ascontiguousarray(b, np.float64)
        hh = None if hold is None else np.ascontiguousarray(hold, np.uint8)
        x = np.empty_like(bb)
        info = SolveInfo()
        status = lib.psim_system_pcg(self.handle, bb.ctypes.data_as(_vp),
                                     None if hh is None else hh.ctypes.data_as(_vp), x.ctypes.data_as(_vp),
                                     tolerance, max_iterations, C.byref(info), HOST)
        if status not in (0, 5):
            check(status)
        return x, info

    def heat_step(self, t_old, forcing, dt=2.63e6, tolerance=1e-10, max_iterations=1000):
        t = np.ascontiguousarray(t_old, np.float64)
        f = np.ascontiguousarray(forcing, np.float64)
        out = np.empty_like(t)
        info = SolveInfo()
        check(lib.psim_heat_step(self.handle, _ptr(t), _ptr(f), dt, tolerance, max_iterations,
                                 _ptr(out), C.byref(info), HOST))
        return out, info

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_system_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

def _pointers(values):
    """A ctypes array of device pointers (ints / DeviceArray / None)."""
    out = (_vp * len(values))()
    for i, v in enumerate(values):
        out[i] = None if v is None else (v.ptr if isinstance(v, DeviceArray) else v)
    return out


class Batch:
    """Many systems in one launch of the persistent solver (psim_batch)."""

    def __init__(self, ctx: Context, systems):
        self.systems = list(systems)
        arr = (_vp * len(self.systems))(*[s.handle for s in self.systems])
        h = _vp()
        check(lib.psim_batch_create(ctx.handle, arr, len(self.systems), C.byref(h)))
        self.handle, self.ctx = h, ctx
        _track(self)

    @property
    def capacity(self):
        cap = _i32()
        check(lib.psim_batch_size(self.handle, None, C.byref(cap)))
        return cap.value

    @property
    def occupancy(self):
        """Workgroups per CU of the batch's program instance (2 or 3)."""
        return lib.psim_batch_occupancy(self.handle)

    @occupancy.setter
    def occupancy(self, workgroups_per_cu):
        check(lib.psim_batch_set_occupancy(self.handle, workgroups_per_cu))

    @property
    def groups(self):
        g = (_i32 * len(self.systems))()
        check(lib.psim_batch_groups(self.handle, g))
        return [int(x) for x in g]

    @groups.setter
    def groups(self, values):
        if values is None:
            check(lib.psim_batch_set_groups(self.handle, None))
        else:
            check(lib.psim_batch_set_groups(self.handle, (_i32 * len(values))(*values)))

    def _args(self, t_old, forcing, t_new, conductance, capacity):
        k = None if conductance is None else _pointers(conductance)
        c = None if capacity is None else _pointers(capacity)
        return k, c, _pointers(t_old), _pointers(forcing), _pointers(t_new)

    def heat_step(self, t_old, forcing, t_new, dt=2.63e6, tolerance=1e-10, max_iterations=1000,
                  conductance=None, capacity=None, collect=True):
        """One step of every system; per-system lists of device arrays or
        pointers. Returns the SolveInfo records when `collect`, else queues."""
        k, c, t0, f, t1 = self._args(t_old, forcing, t_new, conductance, capacity)
        infos = (SolveInfo * len(self.systems))() if collect else None
        check(lib.psim_batch_heat_step(self.handle, k, c, t0, f, dt, tolerance, max_iterations, t1, infos))
        return list(infos) if collect else None

    def synchronize(self):
        check(lib.psim_batch_synchronize(self.handle))

    def autotune(self, t_old, forcing, t_new, dt=2.63e6, tolerance=1e-10, max_iterations=1000,
                 conductance=None, capacity=None, repeats=2):
        k, c, t0, f, t1 = self._args(t_old, forcing, t_new, conductance, capacity)
        best = _dbl()
        check(lib.psim_batch_autotune(self.handle, k, c, t0, f, dt, tolerance, max_iterations, t1, repeats,
                                      C.byref(best)))
        return best.value

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Server:
    """Continuous batching of heat-step requests on one persistent launch
    (psim_server): slots per graph, pooled resident systems, real-time
    arrivals. `graphs` are Graph objects; `share` the expected fraction of
    requests per graph."""

    def __init__(self, ctx: Context, graphs, share=None, precision=MIXED, memory_budget=4 << 30, slot_sizes=None,
                 occupancy=0):
        self.ctx, self.graphs = ctx, list(graphs)
        arr = (_vp * len(self.graphs))(*[g.handle for g in self.graphs])
        sh = None if share is None else (_dbl * len(self.graphs))(*share)
        sz = None if slot_sizes is None else (_i32 * len(self.graphs))(*slot_sizes)
        h = _vp()
        check(lib.psim_server_create_occupancy(ctx.handle, arr, sh, sz, len(self.graphs), precision, memory_budget,
                                               occupancy, C.byref(h)))
        self.handle = h
        _track(self)

    def plan(self):
        k = len(self.graphs)
        slots, size, pool, nbytes = (_i32 * k)(), (_i32 * k)(), (_i32 * k)(), (_u64 * k)()
        check(lib.psim_server_plan(self.handle, slots, size, pool, nbytes))
        return [dict(slots=slots[q], slot_size=size[q], pool=pool[q], system_bytes=nbytes[q]) for q in range(k)]

    def run(self, requests, dt=2.63e6, tolerance=1e-10, max_iterations=1000):
        """requests: sequence of ServeRequest (sorted by arrival). Returns
        the ServeRecord array."""
        reqs = (ServeRequest * len(requests))(*requests)
        recs = (ServeRecord * len(requests))()
        check(lib.psim_server_run(self.handle, reqs, len(requests), dt, tolerance, max_iterations, recs))
        return recs

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def implicit_transport(graph: Graph, conductance, floor, response, settings=None, latent=None):
    """psim_implicit_transport: response(source ndarray) -> (mean, slope).
    Returns (source_W_m2, mean_K, TransportResult)."""
    n = graph.mesh.cells

    def call(_user, source, mean, slope, count):
        m, s = response(np.ctypeslib.as_array(source, shape=(count,)).copy())
        np.ctypeslib.as_array(mean, shape=(count,))[:] = m
        np.ctypeslib.as_array(slope, shape=(count,))[:] = s

    cb = RESPONSE(call)
    if settings is None:
        settings = TransportSettings()
        lib.psim_transport_settings_default(C.byref(settings))
    fl = np.ascontiguousarray(floor, np.float64)
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    src, mean, res = np.zeros(n), np.zeros(n), TransportResult()
    check(lib.psim_implicit_transport(graph.handle, conductance, fl, cb, None, C.byref(settings),
                                      None if lat is None else lat.ctypes.data_as(_vp), src, mean, C.byref(res)))
    return src, mean, res


def _settings(settings):
    if settings is None:
        settings = TransportSettings()
        lib.psim_transport_settings_default(C.byref(settings))
    return settings


def implicit_transport_table(graph: Graph, conductance, floor, table_source, table_mean, settings=None, latent=None):
    """psim_implicit_transport_table with host arrays: a device-resident
    piecewise-linear response, table_* of shape (points, n), sources
    increasing per node. Returns (source_W_m2, mean_K, TransportResult)."""
    n = graph.mesh.cells
    h = np.ascontiguousarray(table_source, np.float64)
    t = np.ascontiguousarray(table_mean, np.float64)
    if h.shape != t.shape or h.ndim != 2 or h.shape[1] != n:
        raise ValueError("response table must be (points, n) for both arrays")
    table = ResponseTable(h.shape[0], 0, h.ctypes.data_as(_vp), t.ctypes.data_as(_vp))
    fl = np.ascontiguousarray(floor, np.float64)
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    src, mean, res = np.zeros(n), np.zeros(n), TransportResult()
    check(lib.psim_implicit_transport_table(graph.handle, conductance, _ptr(fl), C.byref(table),
                                            C.byref(_settings(settings)), None if lat is None else _ptr(lat),
                                            _ptr(src), _ptr(mean), C.byref(res), HOST))
    return src, mean, res


def implicit_transport_table_device(graph: Graph, conductance, floor: "DeviceArray", table_source: "DeviceArray",
                                    table_mean: "DeviceArray", points, source_out: "DeviceArray",
                                    mean_out: "DeviceArray", settings=None, latent: "DeviceArray" = None):
    """The same with device arrays (table arrays hold points x n values)."""
    table = ResponseTable(points, 0, table_source.ptr, table_mean.ptr)
    res = TransportResult()
    check(lib.psim_implicit_transport_table(graph.handle, conductance, floor.ptr, C.byref(table),
                                            C.byref(_settings(settings)), None if latent is None else latent.ptr,
                                            source_out.ptr, mean_out.ptr, C.byref(res), DEVICE))
    return res


class Ilu:
    """psim_ilu: block ILU(0) of a CSR pattern on the device."""

    def __init__(self, ctx: Context, offset, neighbour, block_start):
        o = np.ascontiguousarray(offset, np.uint64)
        nb = np.ascontiguousarray(neighbour, np.uint64)
        bs = np.ascontiguousarray(block_start, np.uint64)
        h = _vp()
        check(lib.psim_ilu_create(ctx.handle, o.size - 1, o, nb, bs, bs.size - 1, C.byref(h)))
        self.handle, self.ctx, self.n = h, ctx, o.size - 1
        _track(self)

    def factor(self, diagonal, off, scale=1.0):
        d = np.ascontiguousarray(diagonal, np.float64)
        o = np.ascontiguousarray(off, np.float64)
        check(lib.psim_ilu_factor(self.handle, _ptr(d), _ptr(o), scale, HOST))

    def solve(self, x):
        y = np.ascontiguousarray(x, np.float64).copy()
        check(lib.psim_ilu_solve(self.handle, _ptr(y), HOST))
        return y

    def close(self):
        """Release the native object now (also done at exit and on collection)."""
        if getattr(self, "handle", None):
            lib.psim_ilu_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class TracerResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("reserved", C.c_int32), ("relative_residual", C.c_double),
                ("clipped_kg", C.c_double), ("column_divergence_kg_s", C.c_double)]


_sig("psim_tracer_step", C.c_int, _vp, C.c_uint32, _pd, _pd, _dbl, _pd, _pd, _dbl, _dbl, _i32, _pd,
     C.POINTER(TracerResult))


def tracer_step(graph: Graph, layers, conductance, outflow, layer_mass, tracer, dt, heat_capacity=1004.64,
                tolerance=1e-12, max_iterations=2000):
    """psim_tracer_step (solve_implicit_tracer): returns (tracer_out, TracerResult)."""
    t = np.ascontiguousarray(tracer, np.float64)
    out = np.zeros_like(t)
    res = TracerResult()
    check(lib.psim_tracer_step(graph.handle, layers, np.ascontiguousarray(conductance, np.float64),
                               np.ascontiguousarray(outflow, np.float64), heat_capacity,
                               np.ascontiguousarray(layer_mass, np.float64), t, dt, tolerance, max_iterations, out,
                               C.byref(res)))
    return out, res


_sig("psim_system_pcg_operator", C.c_int, _vp, _vp, _vp, _vp, _vp, _dbl, _i32, C.POINTER(SolveInfo), C.c_int)

FORM_TRANSPORT, FORM_BALANCE = 0, 1


class Preconditioner(C.Structure):
    _fields_ = [("multigrid", _vp), ("pre_scale", _vp), ("post_divide", _vp), ("ilu", _vp)]


class OperatorInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32), ("relative_residual", C.c_double)]


_sig("psim_operator_solve", C.c_int, _vp, _vp, _vp, C.POINTER(Preconditioner), C.c_int, _vp, _vp, _dbl, _i32,
     C.POINTER(OperatorInfo), C.c_int)


def pcg_operator(system: System, operator_diagonal, b, tolerance=1e-6, max_iterations=500, hold=None):
    """psim_system_pcg_operator with host arrays: (x, SolveInfo)."""
    d = np.ascontiguousarray(operator_diagonal, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(bb)
    h = None if hold is None else np.ascontiguousarray(hold, np.uint8)
    info = SolveInfo()
    check(lib.psim_system_pcg_operator(system.handle, _ptr(d), _ptr(bb), None if h is None else _ptr(h), _ptr(x),
                                       tolerance, max_iterations, C.byref(info), HOST))
    return x, info


def operator_solve(graph: Graph, diagonal, off, b, multigrid: System = None, ilu: "Ilu" = None, pre_scale=None,
                   post_divide=None, form=FORM_BALANCE, tolerance=1e-10, max_iterations=500):
    """psim_operator_solve with host arrays: (x, OperatorInfo)."""
    keep = [np.ascontiguousarray(a, np.float64) if a is not None else None
            for a in (diagonal, off, b, pre_scale, post_divide)]
    d, o, bb, ps, pd = keep
    pre = Preconditioner(multigrid.handle if multigrid is not None else None, _ptr(ps) if ps is not None else None,
                         _ptr(pd) if pd is not None else None, ilu.handle if ilu is not None else None)
    x = np.zeros_like(bb)
    info = OperatorInfo()
    check(lib.psim_operator_solve(graph.handle, _ptr(d), _ptr(o), C.byref(pre), form, _ptr(bb), _ptr(x), tolerance,
                                  max_iterations, C.byref(info), HOST))
    return x, info


ADVECTION_RESPONSE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64,
                                 C.c_uint32)


class AdvectionResult(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("linear_iterations", C.c_int32),
                ("consistency_residual_W_m2", C.c_double), ("sum_W", C.c_double), ("absolute_sum_W", C.c_double),
                ("eddy_absolute_sum_W", C.c_double), ("advective_absolute_sum_W", C.c_double)]


_sig("psim_advection_source", C.c_int, _vp, C.c_uint32, _pd, _pd, _vp, _pd, _pd, _pd, _pd, C.c_int)
_sig("psim_implicit_advection_diffusion", C.c_int, _vp, C.c_uint32, _pd, _pd, _vp, _pd, ADVECTION_RESPONSE, _vp,
     C.POINTER(TransportSettings), _vp, _pd, _pd, _pd, C.POINTER(AdvectionResult))


def advection_source(graph: Graph, layers, conductance, outflow, mean, energy, latent=None):
    """psim_advection_source with host arrays: (eddy, advective)."""
    n = graph.mesh.cells
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    eddy, adv = np.zeros(n), np.zeros(n)
    check(lib.psim_advection_source(graph.handle, layers, np.ascontiguousarray(conductance, np.float64),
                                    np.ascontiguousarray(outflow, np.float64), None if lat is None else _ptr(lat),
                                    np.ascontiguousarray(mean, np.float64), np.ascontiguousarray(energy, np.float64),
                                    eddy, adv, HOST))
    return eddy, adv


def implicit_advection_diffusion(graph: Graph, layers, conductance, outflow, floor, response, settings=None,
                                 latent=None, initial=None):
    """psim_implicit_advection_diffusion: response(source) -> (mean, slope,
    energy, energy_slope); returns (source, mean, energy, AdvectionResult)."""
    n = graph.mesh.cells

    def call(_user, source, mean, slope, energy, energy_slope, count, nl):
        th, sl, e, es = response(np.ctypeslib.as_array(source, shape=(count,)).copy())
        np.ctypeslib.as_array(mean, shape=(count,))[:] = th
        np.ctypeslib.as_array(slope, shape=(count,))[:] = sl
        np.ctypeslib.as_array(energy, shape=(nl * count,))[:] = e
        np.ctypeslib.as_array(energy_slope, shape=(nl * count,))[:] = es

    cb = ADVECTION_RESPONSE(call)
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    ini = None if initial is None else np.ascontiguousarray(initial, np.float64)
    src, mean, energy, res = np.zeros(n), np.zeros(n), np.zeros(layers * n), AdvectionResult()
    check(lib.psim_implicit_advection_diffusion(graph.handle, layers, np.ascontiguousarray(conductance, np.float64),
                                                np.ascontiguousarray(outflow, np.float64),
                                                None if lat is None else _ptr(lat),
                                                np.ascontiguousarray(floor, np.float64), cb, None,
                                                C.byref(_settings(settings)), None if ini is None else _ptr(ini),
                                                src, mean, energy, C.byref(res)))
    return src, mean, energy, res

"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the compiled
reference shim (oracle/_ref/libref.so) for the tests.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
from functools import lru_cache

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
RADIUS = 6371000.0

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u64 = C.c_uint64
_vp = C.c_void_p


def _declare(lib, prefix):
    """Give both libraries' common entry points their signatures."""
    p = prefix

    def fn(name, res, *args):
        f = getattr(lib, p + name)
        f.restype = res
        f.argtypes = list(args)

    fn("last_error", C.c_char_p)
    fn("mix_random_key", _u64, _u64)
    fn("keyed_random_u64", _u64, _u64, C.c_uint32, C.c_int64, C.c_uint32, C.c_uint32)
    fn("keyed_random_unit_double", C.c_double, _u64, C.c_uint32, C.c_int64, C.c_uint32, C.c_uint32)
    fn("mesh_build", _vp, C.c_uint32, C.c_double)
    fn("mesh_free", None, _vp)
    fn("mesh_counts", None, _vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64))
    fn("mesh_cells", None, _vp, _dp, _dp, _dp, _dp, _u32p, _u8p)
    fn("mesh_cell_edges", None, _vp, _u32p, _u32p, _fp, _fp)
    fn("mesh_edges", None, _vp, _u32p, _u32p, _dp, _dp)
    fn("graph", None, _vp, _dp, _u64p, _u64p, _dp, _dp)
    fn("gradient_f64", C.c_int, _vp, _dp, _dp, _dp, _u64)
    fn("gradient_f32", C.c_int, _vp, _fp, _fp, _fp, _u64)
    fn("divergence_f64", C.c_int, _vp, _dp, _dp, _u64)
    fn("divergence_f32", C.c_int, _vp, _fp, _fp, _u64)
    fn("laplacian_f64", C.c_int, _vp, _dp, _dp, _u64)
    fn("laplacian_f32", C.c_int, _vp, _fp, _fp, _u64)
    fn("edge_orientation", C.c_int, _vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_double))
    fn("diffusion_source", C.c_int, _vp, C.c_double, _dp, _dp, _u64)
    fn("hierarchy_build", _vp, _vp)
    fn("hierarchy_free", None, _vp)
    fn("hierarchy_depth", _u64, _vp)
    fn("hierarchy_level_counts", None, _vp, _u64, C.POINTER(_u64), C.POINTER(_u64))
    fn("hierarchy_level", None, _vp, _u64, _u64p, _u64p, _u64p, _u64p, _dp, _dp)
    fn("mg_create", _vp, _vp)
    fn("mg_free", None, _vp)
    fn("mg_set_matrix", C.c_int, _vp, _dp, _u64, _dp, _u64)
    fn("mg_precondition", C.c_int, _vp, _dp, _dp, _u64, _u64)
    fn("bicgstab", C.c_int, _vp, _vp, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _u64,
       C.POINTER(C.c_int))
    fn("newton_system", C.c_int, _vp, C.c_double, _dp, _dp, C.c_double, C.c_int, _u64, _dp,
       C.POINTER(C.c_int))
    fn("solve_implicit_transport", C.c_int, _vp, C.c_double, _dp, RESPONSE, _vp,
       C.POINTER(TransportSettings), _u64, _vp, _dp, _dp, C.POINTER(TransportResult))
    fn("advection_source", C.c_int, _vp, C.c_uint32, _dp, _dp, _vp, _dp, _dp, _dp, _dp)
    fn("solve_advection_diffusion", C.c_int, _vp, C.c_uint32, _dp, _dp, _vp, _dp, ADVECTION_RESPONSE, _vp,
       C.POINTER(TransportSettings), _vp, _u64, _dp, _dp, _dp, C.POINTER(AdvectionResult))


# solve_implicit_transport's C form (both libraries): a response callback
# response(user, source, mean_out, slope_out, n), settings and result.
RESPONSE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                       C.POINTER(C.c_double), C.c_uint64)


class TransportSettings(C.Structure):
    """ImplicitTransportSettings (heat_transport.hpp:82-106) defaults."""
    _fields_ = [("max_newton_iterations", C.c_int), ("max_line_search_halvings", C.c_int),
                ("rounding_floor", C.c_double), ("max_non_monotone_steps", C.c_int), ("pad", C.c_int),
                ("max_non_monotone_growth", C.c_double), ("newton_tolerance", C.c_double),
                ("cg_relative_tolerance", C.c_double), ("max_cg_iterations", C.c_int), ("pad2", C.c_int)]

    def __init__(self, **kw):
        defaults = dict(max_newton_iterations=30, max_line_search_halvings=12, rounding_floor=1e-3,
                        max_non_monotone_steps=4, max_non_monotone_growth=100.0, newton_tolerance=1e-6,
                        cg_relative_tolerance=1e-6, max_cg_iterations=2000)
        defaults.update(kw)
        super().__init__(**defaults)


class TransportResult(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("cg_iterations", C.c_int), ("consistency_residual", C.c_double),
                ("sum_W", C.c_double), ("absolute_sum_W", C.c_double), ("dissipation_W_K", C.c_double)]


def response_callback(fn):
    """Wrap fn(source ndarray) -> (mean, slope) as a C response callback."""

    def call(_user, source, mean, slope, n):
        h = np.ctypeslib.as_array(source, shape=(n,))
        m, s = fn(h.copy())
        np.ctypeslib.as_array(mean, shape=(n,))[:] = m
        np.ctypeslib.as_array(slope, shape=(n,))[:] = s

    return RESPONSE(call)


def implicit_transport(lib, mesh_handle, n, conductance, floor, fn, settings=None, latent=None):
    """One library's solve_implicit_transport: (source, mean, result)."""
    cb = response_callback(fn)
    src, mean = np.zeros(n), np.zeros(n)
    res = TransportResult()
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    lib.check(lib.solve_implicit_transport(mesh_handle, conductance, np.ascontiguousarray(floor, np.float64), cb,
                                           None, C.byref(settings or TransportSettings()), 1,
                                           None if lat is None else lat.ctypes.data, src, mean, C.byref(res)))
    return src, mean, res


def quartic_response(mesh_area_n, seed_values, a, radiative=0.75 * 0.97 * 5.670374419e-8):
    """A tile-like response (test_heat_transport.cpp:30-59 in spirit): each
    cell solves a x + r x^4 = b + h by Newton from 250 K (fixed iteration
    count, so every caller gets the same bits); cells with index % 7 == 0
    are held at 273.15 K (slope 0) whenever they would warm past it."""
    target = 200.0 + 120.0 * seed_values
    b = a * target + radiative * target ** 4

    def fn(h):
        x = np.full_like(h, 250.0)
        rhs = b + h
        for _ in range(40):
            f = a * x + radiative * x ** 4 - rhs
            x = x - f / (a + 4.0 * radiative * x ** 3)
        slope = 1.0 / (a + 4.0 * radiative * x ** 3)
        held = (np.arange(h.size) % 7 == 0) & (x > 273.15)
        mean = np.where(held, 273.15, x)
        slope = np.where(held, 0.0, slope)
        return mean, slope

    floor = a * 100.0 + radiative * 100.0 ** 4 - b
    return fn, floor


def table_response(table_h, table_t):
    """A piecewise-linear response per node, evaluated on the host exactly as
    psim_implicit_transport_table evaluates it on the device: the segment is
    the last one whose left end is <= h (first below the table, last above),
    s = (t1 - t0) / (h1 - h0), T = t0 + s (h - h0)."""
    points, n = table_h.shape
    cols = np.arange(n)

    def fn(h):
        k = np.zeros(n, np.int64)
        for j in range(1, points - 1):
            k += (h >= table_h[j]).astype(np.int64)
        h0, h1 = table_h[k, cols], table_h[k + 1, cols]
        t0, t1 = table_t[k, cols], table_t[k + 1, cols]
        s = (t1 - t0) / (h1 - h0)
        return t0 + s * (h - h0), s

    return fn


def response_table(n, points, seed):
    """A concave increasing table per node (slopes falling along h), with a
    flat (held) segment on every sixth node, and the floor at the table's
    first source."""
    rng = np.random.default_rng(seed)
    dh = rng.uniform(20.0, 80.0, (points - 1, n))
    table_h = np.vstack([np.full((1, n), -150.0) + rng.uniform(-20.0, 20.0, (1, n)),
                         np.zeros((points - 1, n))])
    table_h[1:] = table_h[0] + np.cumsum(dh, axis=0)
    slopes = -np.sort(-rng.uniform(0.01, 0.3, (points - 1, n)), axis=0)
    slopes[points // 2, ::6] = 0.0
    table_t = np.vstack([250.0 + 30.0 * rng.random((1, n)), np.zeros((points - 1, n))])
    table_t[1:] = table_t[0] + np.cumsum(slopes * dh, axis=0)
    return np.ascontiguousarray(table_h), np.ascontiguousarray(table_t), table_h[0].copy()


class TracerResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("pad", C.c_int), ("relative_residual", C.c_double),
                ("clipped_kg", C.c_double), ("column_divergence_kg_s", C.c_double)]


def tracer_case(m, layers, seed, heat_capacity=1004.64):
    """Synthetic tracer-step inputs on a mesh's graph: per-edge layer mass
    fluxes (antisymmetric per entry), symmetric eddy conductances, layer
    masses from the cell areas, positive mixing ratios."""
    rng = np.random.default_rng(seed)
    edges = m.n_edges
    sign = np.where(m.first[m.edge] == np.repeat(np.arange(m.n), np.diff(m.offset.astype(np.int64))), 1.0, -1.0)
    flux_e = rng.standard_normal((layers, edges)) * 2.0e7
    outflow = np.ascontiguousarray((flux_e[:, m.edge] * sign[None, :]).ravel())
    conductance = np.ascontiguousarray(rng.uniform(0.5, 1.5, edges)[m.edge] * 3.0e9)
    layer_mass = np.ascontiguousarray(m.area * 1.0e4)
    tracer = np.ascontiguousarray(rng.uniform(0.0, 2.0e-3, layers * m.n))
    return dict(layers=layers, conductance=conductance, outflow=outflow, heat_capacity=heat_capacity,
                layer_mass=layer_mass, tracer=tracer, dt=2.63e6)


ADVECTION_RESPONSE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64,
                                 C.c_uint32)


class AdvectionResult(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("linear_iterations", C.c_int), ("consistency_residual", C.c_double),
                ("sum_W", C.c_double), ("absolute_sum_W", C.c_double), ("eddy_absolute_sum_W", C.c_double),
                ("advective_absolute_sum_W", C.c_double)]


def advection_case(m, layers, seed, quartic=False):
    """Synthetic coupled-transport inputs: symmetric eddy conductances (W/K),
    antisymmetric layer mass fluxes (kg/s), and a column response per node:
    theta(h) (linear, or the quartic tile-like response), each layer's s_l =
    c_p (T_l + w_l (theta - 250)) with its slope."""
    rng = np.random.default_rng(seed)
    sign = np.where(m.first[m.edge] == np.repeat(np.arange(m.n), np.diff(m.offset.astype(np.int64))), 1.0, -1.0)
    conductance = np.ascontiguousarray(rng.uniform(0.5, 1.5, m.n_edges)[m.edge] * m.weight * 0.5 * RADIUS ** 2)
    outflow = np.ascontiguousarray((rng.standard_normal((layers, m.n_edges))[:, m.edge] * sign[None, :]
                                    * 3.0e8).ravel())
    cp = 1004.64
    w = rng.uniform(0.2, 1.0, layers)
    t0 = 250.0 + 10.0 * np.arange(layers)
    if quartic:
        theta_fn, floor = quartic_response(m.n, rng.random(m.n), 0.41)
    else:
        base = 250.0 + 40.0 * rng.random(m.n)
        slopes = 0.02 + 0.1 * rng.random(m.n)
        theta_fn = lambda h: (base + slopes * h, slopes.copy())   # noqa: E731
        floor = np.full(m.n, -np.inf)

    def fn(h):
        theta, slope = theta_fn(h)
        energy = cp * (t0[:, None] + w[:, None] * (theta[None, :] - 250.0))
        energy_slope = cp * w[:, None] * slope[None, :]
        return theta, slope, np.ascontiguousarray(energy.ravel()), np.ascontiguousarray(energy_slope.ravel())

    return dict(layers=layers, conductance=conductance, outflow=outflow, floor=floor, response=fn)


def advection_callback(fn):
    def call(_user, source, mean, slope, energy, energy_slope, n, layers):
        th, sl, e, es = fn(np.ctypeslib.as_array(source, shape=(n,)).copy())
        np.ctypeslib.as_array(mean, shape=(n,))[:] = th
        np.ctypeslib.as_array(slope, shape=(n,))[:] = sl
        np.ctypeslib.as_array(energy, shape=(layers * n,))[:] = e
        np.ctypeslib.as_array(energy_slope, shape=(layers * n,))[:] = es

    return ADVECTION_RESPONSE(call)


def solve_advection(lib, handle, n, case, settings=None, latent=None, initial=None):
    layers = case["layers"]
    src, mean, energy = np.zeros(n), np.zeros(n), np.zeros(layers * n)
    res = AdvectionResult()
    cb = advection_callback(case["response"])
    st = settings or TransportSettings()
    lat = None if latent is None else np.ascontiguousarray(latent, np.float64)
    ini = None if initial is None else np.ascontiguousarray(initial, np.float64)
    lib.check(lib.solve_advection_diffusion(handle, layers, case["conductance"], case["outflow"],
                                            None if lat is None else lat.ctypes.data_as(_vp), case["floor"], cb,
                                            None, C.byref(st), None if ini is None else ini.ctypes.data_as(_vp), 1,
                                            src, mean, energy, C.byref(res)))
    return src, mean, energy, res


def oracle_tracer(m, case, tol=1e-12, max_it=2000):
    out = np.zeros_like(case["tracer"])
    res = TracerResult()
    oracle().check(oracle().tracer_step(m.handle, case["layers"], case["conductance"], case["outflow"],
                                        case["heat_capacity"], case["layer_mass"], case["tracer"], case["dt"], tol,
                                        max_it, 1, out, C.byref(res)))
    return out, res


class Lib:
    """One of the two CPU libraries (oracle or reference shim)."""

    def __init__(self, path, prefix):
        self.lib = C.CDLL(path)
        self.p = prefix
        _declare(self.lib, prefix)
        if prefix == "orc_":
            f = self.lib
            f.orc_heat_instance.restype = C.c_int
            f.orc_heat_instance.argtypes = [_vp, _u64, C.c_double, C.c_double, _dp, _dp, _u8p, _dp]
            f.orc_heat_forcing.restype = C.c_int
            f.orc_heat_forcing.argtypes = [_vp, _u64, C.c_uint32, _dp]
            f.orc_heat_matrix.restype = C.c_int
            f.orc_heat_matrix.argtypes = [_vp, _dp, _dp, C.c_double, _dp, _dp]
            f.orc_heat_rhs.restype = C.c_int
            f.orc_heat_rhs.argtypes = [_vp, _dp, _dp, _dp, C.c_double, _dp]
            f.orc_relative_residual.restype = C.c_double
            f.orc_relative_residual.argtypes = [_vp, _dp, _dp, _dp, _dp]
            f.orc_mg_coarsest_size.restype = _u64
            f.orc_mg_coarsest_size.argtypes = [_vp]
            f.orc_mg_coarsest_factor.restype = None
            f.orc_mg_coarsest_factor.argtypes = [_vp, _dp]
            f.orc_mg_level_values.restype = None
            f.orc_mg_level_values.argtypes = [_vp, _u64, _dp, _dp]
            f.orc_agglomerate.restype = _vp
            f.orc_agglomerate.argtypes = [_vp, _u32p]
            f.orc_group_mean.restype = C.c_int
            f.orc_group_mean.argtypes = [_u64, _u64, _u32p, _dp, _dp, _dp, _dp]
            f.orc_group_max.restype = C.c_int
            f.orc_group_max.argtypes = [_u64, _u64, _u32p, _dp, _dp]
            f.orc_ilu_create.restype = _vp
            f.orc_ilu_create.argtypes = [_u64, _u64p, _u64p, _dp, _dp, _u64p, _u64, C.c_double]
            f.orc_ilu_free.restype = None
            f.orc_ilu_free.argtypes = [_vp]
            f.orc_ilu_solve.restype = C.c_int
            f.orc_ilu_solve.argtypes = [_vp, _dp]
            f.orc_tracer_step.restype = C.c_int
            f.orc_tracer_step.argtypes = [_vp, C.c_uint32, _dp, _dp, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                          C.c_int, _u64, _dp, C.POINTER(TracerResult)]
        else:
            f = self.lib
            f.ref_coupled_linear_solve.restype = C.c_int
            f.ref_coupled_linear_solve.argtypes = [_vp, _dp, _dp, _dp, C.c_double, C.c_double, _dp,
                                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]
            f.ref_ilu_create.restype = _vp
            f.ref_ilu_create.argtypes = [_u64, _u64p, _u64p, _dp, _dp, _u64p, _u64, C.c_double]
            f.ref_ilu_free.restype = None
            f.ref_ilu_free.argtypes = [_vp]
            f.ref_ilu_solve.restype = C.c_int
            f.ref_ilu_solve.argtypes = [_vp, _dp, _u64]
            _i, _d = C.POINTER(C.c_int), C.POINTER(C.c_double)
            f.ref_tracer_step.restype = C.c_int
            f.ref_tracer_step.argtypes = [_vp, C.c_uint32, _dp, _dp, C.c_double, _dp, _dp, C.c_double, C.c_double,
                                          C.c_int, _u64, _dp, _i, _d, _d, _d]
            f.ref_agglomeration_groups.restype = _u64
            f.ref_agglomeration_groups.argtypes = [_vp]
            f.ref_agglomeration.restype = C.c_int
            f.ref_agglomeration.argtypes = [_vp, _u64p, _dp, _u64p, _u64p, _dp, _dp, C.POINTER(C.c_uint32)]

    def __getattr__(self, name):
        return getattr(self.lib, self.p + name)

    def check(self, status):
        if status != 0:
            raise RuntimeError(self.last_error().decode())


@lru_cache(maxsize=None)
def oracle() -> Lib:
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return Lib(ORACLE_SO, "orc_")


def have_ref() -> bool:
    return os.path.exists(REF_SO)


@lru_cache(maxsize=None)
def ref() -> Lib:
    return Lib(REF_SO, "ref_")


class MeshArrays:
    """A mesh built by one library, with its arrays copied out."""

    def __init__(self, lib: Lib, level: int, radius: float = RADIUS, handle=None):
        self.lib = lib
        self.level = level
        self.handle = lib.mesh_build(level, radius) if handle is None else handle
        if not self.handle:
            raise RuntimeError(lib.last_error().decode())
        n, e, d = _u64(), _u64(), _u64()
        lib.mesh_counts(self.handle, C.byref(n), C.byref(e), C.byref(d))
        self.n, self.n_edges, self.n_directed = n.value, e.value, d.value
        self.center = np.zeros(3 * self.n)
        self.east = np.zeros(3 * self.n)
        self.north = np.zeros(3 * self.n)
        self.area = np.zeros(self.n)
        self.edge_offset = np.zeros(self.n, np.uint32)
        self.edge_count = np.zeros(self.n, np.uint8)
        lib.mesh_cells(self.handle, self.center, self.east, self.north, self.area,
                       self.edge_offset, self.edge_count)
        self.neighbor = np.zeros(self.n_directed, np.uint32)
        self.edge = np.zeros(self.n_directed, np.uint32)
        self.normal_east = np.zeros(self.n_directed, np.float32)
        self.normal_north = np.zeros(self.n_directed, np.float32)
        lib.mesh_cell_edges(self.handle, self.neighbor, self.edge, self.normal_east,
                            self.normal_north)
        self.first = np.zeros(self.n_edges, np.uint32)
        self.second = np.zeros(self.n_edges, np.uint32)
        self.length = np.zeros(self.n_edges)
        self.distance = np.zeros(self.n_edges)
        lib.mesh_edges(self.handle, self.first, self.second, self.length, self.distance)
        self.g_area = np.zeros(self.n)
        self.offset = np.zeros(self.n + 1, np.uint64)
        self.g_neighbour = np.zeros(self.n_directed, np.uint64)
        self.weight = np.zeros(self.n_directed)
        self.weight_sum = np.zeros(self.n)
        lib.graph(self.handle, self.g_area, self.offset, self.g_neighbour, self.weight,
                  self.weight_sum)

    def __del__(self):
        try:
            self.lib.mesh_free(self.handle)
        except Exception:
            pass

    def arrays(self):
        return {
            "center": self.center, "east": self.east, "north": self.north, "area": self.area,
            "edge_offset": self.edge_offset, "edge_count": self.edge_count,
            "neighbor": self.neighbor, "edge": self.edge, "normal_east": self.normal_east,
            "normal_north": self.normal_north, "first": self.first, "second": self.second,
            "length": self.length, "distance": self.distance, "offset": self.offset,
            "weight": self.weight, "weight_sum": self.weight_sum,
        }


_mesh_cache: dict = {}


def agglomerate(level: int):
    """The oracle's agglomeration of an L`level` mesh: (group_of_cell, the
    agglomerated graph as MeshArrays of the coarse mesh with group areas)."""
    fine = mesh("oracle", level)
    group = np.zeros(fine.n, np.uint32)
    h = oracle().agglomerate(fine.handle, group)
    if not h:
        raise RuntimeError(oracle().last_error().decode())
    return group, MeshArrays(oracle(), max(level - 1, 0), handle=h)


def ref_agglomeration(level: int):
    """The compiled reference's agglomerated_transport_graph of L`level`."""
    fine = mesh("ref", level)
    groups = ref().agglomeration_groups(fine.handle)
    nnz_guess = 6 * groups
    out = {"group": np.zeros(fine.n, np.uint64), "area": np.zeros(groups),
           "offset": np.zeros(groups + 1, np.uint64), "neighbour": np.zeros(nnz_guess, np.uint64),
           "weight": np.zeros(nnz_guess), "weight_sum": np.zeros(groups)}
    sub = C.c_uint32()
    ref().check(ref().agglomeration(fine.handle, out["group"], out["area"], out["offset"], out["neighbour"],
                                    out["weight"], out["weight_sum"], C.byref(sub)))
    nnz = int(out["offset"][-1])
    assert nnz <= nnz_guess
    out["neighbour"], out["weight"] = out["neighbour"][:nnz], out["weight"][:nnz]
    out["subdivision"] = sub.value
    return out


def group_mean(group, cell_area, group_area, values):
    out = np.zeros(group_area.size)
    oracle().check(oracle().group_mean(group.size, group_area.size, group, cell_area, group_area,
                                       np.ascontiguousarray(values, np.float64), out))
    return out


def group_max(group, groups, values):
    out = np.zeros(groups)
    oracle().check(oracle().group_max(group.size, groups, group, np.ascontiguousarray(values, np.float64), out))
    return out


def mesh(lib_name: str, level: int) -> MeshArrays:
    key = (lib_name, level)
    if key not in _mesh_cache:
        lib = oracle() if lib_name == "oracle" else ref()
        _mesh_cache[key] = MeshArrays(lib, level)
    return _mesh_cache[key]


def sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def crc32c(data: bytes) -> int:
    """CRC-32C (Castagnoli, reflected 0x82F63B78), table-driven."""
    table = _crc_table()
    crc = 0xFFFFFFFF
    arr = np.frombuffer(data, np.uint8)
    for b in arr.tolist():
        crc = table[(crc ^ b) & 0xFF] ^ (crc >> 8)
    return crc ^ 0xFFFFFFFF


@lru_cache(maxsize=1)
def _crc_table():
    table = []
    for i in range(256):
        r = i
        for _ in range(8):
            r = (r >> 1) ^ (0x82F63B78 if r & 1 else 0)
        table.append(r)
    return table


def hierarchy(lib: Lib, mesh_arrays: MeshArrays):
    """Levels of the unit-conductance hierarchy as dicts of arrays."""
    h = lib.hierarchy_build(mesh_arrays.handle)
    try:
        depth = lib.hierarchy_depth(h)
        levels = []
        for j in range(depth):
            size, nnz = _u64(), _u64()
            lib.hierarchy_level_counts(h, j, C.byref(size), C.byref(nnz))
            s, z = size.value, nnz.value
            last = j + 1 == depth
            lv = {
                "size": s, "nnz": z,
                "offset": np.zeros(s + 1, np.uint64), "column": np.zeros(max(z, 1), np.uint64),
                "aggregate": np.zeros(max(0 if last else s, 1), np.uint64),
                "target": np.zeros(max(0 if last else z, 1), np.uint64),
                "diagonal": np.zeros(s), "value": np.zeros(max(z, 1)),
            }
            lib.hierarchy_level(h, j, lv["offset"], lv["column"], lv["aggregate"], lv["target"],
                                lv["diagonal"], lv["value"])
            lv["column"] = lv["column"][:z]
            lv["value"] = lv["value"][:z]
            lv["aggregate"] = lv["aggregate"][: (0 if last else s)]
            lv["target"] = lv["target"][: (0 if last else z)]
            levels.append(lv)
        return levels
    finally:
        lib.hierarchy_free(h)


class Multigrid:
    def __init__(self, lib: Lib, mesh_arrays: MeshArrays):
        self.lib = lib
        self.mesh = mesh_arrays
        self.handle = lib.mg_create(mesh_arrays.handle)

    def set_matrix(self, diag, off):
        self.lib.check(self.lib.mg_set_matrix(self.handle, diag, len(diag), off, len(off)))

    def precondition(self, rhs, workers=1):
        out = np.zeros(self.mesh.n)
        self.lib.check(self.lib.mg_precondition(self.handle, rhs, out, self.mesh.n, workers))
        return out

    def bicgstab(self, diag, off, b, tol=1e-10, max_it=1000, workers=1):
        x = np.zeros(self.mesh.n)
        it = C.c_int()
        self.lib.check(self.lib.bicgstab(self.mesh.handle, self.handle, diag, off, b, x, tol,
                                         max_it, workers, C.byref(it)))
        return x, it.value

    def __del__(self):
        try:
            self.lib.mg_free(self.handle)
        except Exception:
            pass


BASE_SEED = 20261203
DT = 2.63e6


def instance_seed(index: int, base: int = BASE_SEED) -> int:
    return oracle().mix_random_key(base ^ index)


def heat_instance(m: MeshArrays, seed: int, k_base=1e6, sigma=0.3):
    """Oracle workload for one instance: conductance, C·A, land, T_old."""
    o = oracle()
    cond = np.zeros(m.n_edges)
    ca = np.zeros(m.n)
    land = np.zeros(m.n, np.uint8)
    temp = np.zeros(m.n)
    o.check(o.lib.orc_heat_instance(m.handle, seed, k_base, sigma, cond, ca, land, temp))
    return cond, ca, land, temp


def heat_forcing(m: MeshArrays, seed: int, step: int):
    o = oracle()
    f = np.zeros(m.n)
    o.check(o.lib.orc_heat_forcing(m.handle, seed, step, f))
    return f


def heat_matrix(m: MeshArrays, cond, ca, dt=DT):
    o = oracle()
    diag = np.zeros(m.n)
    off = np.zeros(m.n_directed)
    o.check(o.lib.orc_heat_matrix(m.handle, cond, ca, dt, diag, off))
    return diag, off


def heat_rhs(m: MeshArrays, ca, temp, forcing, dt=DT):
    o = oracle()
    b = np.zeros(m.n)
    o.check(o.lib.orc_heat_rhs(m.handle, ca, temp, forcing, dt, b))
    return b


def dense_matrix(m: MeshArrays, diag, off):
    A = np.diag(diag.copy())
    for c in range(m.n):
        for k in range(int(m.offset[c]), int(m.offset[c + 1])):
            A[c, int(m.neighbor[k])] += off[k]
    return A

"""psim_heat_step_noalloc (through the C ABI) against psim_heat_step.

The allocation-free entry point runs the same kernels with the solver's
workspace in the library's static device array, so its temperatures and
solve records must be bit-identical to psim_heat_step's on the same system
and inputs: two calls in a row, the second on a smaller system (the array
keeps the first call's contents), plus the FP32-inner form, plus the
rejection of a system whose workspace exceeds the array.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import planet_game_amd as P
import support as S

pytestmark = pytest.mark.gpu


def announce(*what):
    print("[gpu case]", *what, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _case(ctx, level, precision, seed_index, tol=1e-10):
    mesh = P.Mesh(level)
    graph = P.Graph(ctx, mesh)
    system = P.System(graph, precision)
    seed = P.instance_seed(seed_index)
    cond, ca, _, t0 = P.heat_instance(mesh, seed, 1.0e6, 0.3)
    forcing = P.heat_forcing(mesh, seed, 3)
    d_cond = P.DeviceArray.of(ctx, mesh.entry_values(cond))
    d_ca = P.DeviceArray.of(ctx, ca)
    d_t0 = P.DeviceArray.of(ctx, t0)
    d_f = P.DeviceArray.of(ctx, forcing)
    system.heat_set_matrix_device(d_cond, d_ca, 2.63e6)
    out_ref = P.DeviceArray(ctx, mesh.cells)
    out_na = P.DeviceArray(ctx, mesh.cells)
    # the existing entry point, device pointers
    system.heat_step_device(d_t0.ptr, d_f.ptr, out_ref.ptr, 2.63e6, tol, 1000)
    ref_info = system.last_info()
    # the allocation-free one, straight through the C ABI
    info = P.SolveInfo()
    status = P.lib.psim_heat_step_noalloc(system.handle, d_t0.ptr, d_f.ptr, C.c_double(2.63e6), C.c_double(tol),
                                          1000, out_na.ptr, C.byref(info))
    assert status == 0, P.lib.psim_last_error()
    return mesh, cond, ca, t0, forcing, out_ref.download(), ref_info, out_na.download(), info


@pytest.mark.parametrize("precision", [P.PLAIN_FP64, P.MIXED], ids=["fp64", "mixed"])
def test_noalloc_matches_heat_step_twice(ctx, precision):
    # first call at L6, then a smaller system (L4): the array holds the L6
    # call's workspace when the L4 call starts
    for level, seed_index in ((6, 31), (4, 32)):
        announce("noalloc vs heat_step", f"L{level}", "precision", precision)
        mesh, cond, ca, t0, forcing, ref, ref_info, got, info = _case(ctx, level, precision, seed_index)
        assert bits_equal(got, ref), f"L{level}: max |diff| {np.abs(got - ref).max():.3e}"
        for field in ("iterations", "restarts", "status", "v_cycles", "relative_residual"):
            assert getattr(info, field) == getattr(ref_info, field), field
        assert info.status == 0 and info.relative_residual <= 1e-10
        if precision == P.PLAIN_FP64:
            # and the step itself against the FP64 oracle (BiCGSTAB to 1e-13)
            om = S.mesh("oracle", level)
            diag, off = S.heat_matrix(om, cond, ca)
            mg = S.Multigrid(S.oracle(), om)
            mg.set_matrix(diag, off)
            x_ref, _ = mg.bicgstab(diag, off, S.heat_rhs(om, ca, t0, forcing), 1e-13, 200)
            err = np.linalg.norm(got - x_ref) / np.linalg.norm(x_ref)
            assert err <= 1e-11, err


@pytest.mark.parametrize("precision", [P.PLAIN_FP64, P.MIXED], ids=["fp64", "mixed"])
def test_noalloc_on_a_wide_graph(ctx, precision):
    """The widened L5 graph of test_gpu_wide (rows of 18, 11 and 8 entries on
    its three levels, so the solver's general row loops), stiff conductances
    (several iterations): the same bits and solve record as psim_heat_step."""
    from test_gpu_wide import STIFF, wide_case
    g, cond, ca, t_old, forcings = wide_case(STIFF)
    announce("noalloc vs heat_step", "wide L5 graph", "precision", precision)
    system = P.System(P.Graph(ctx, g), precision)
    d_cond, d_ca, d_t0, d_f = (P.DeviceArray.of(ctx, a) for a in (cond, ca, t_old, forcings[0]))
    system.heat_set_matrix_device(d_cond, d_ca, 2.63e6)
    out_ref, out_na = P.DeviceArray(ctx, g.cells), P.DeviceArray(ctx, g.cells)
    system.heat_step_device(d_t0.ptr, d_f.ptr, out_ref.ptr, 2.63e6, 1e-10, 1000)
    ref_info = system.last_info()
    info = P.SolveInfo()
    status = P.lib.psim_heat_step_noalloc(system.handle, d_t0.ptr, d_f.ptr, C.c_double(2.63e6), C.c_double(1e-10),
                                          1000, out_na.ptr, C.byref(info))
    assert status == 0, P.lib.psim_last_error()
    assert bits_equal(out_na.download(), out_ref.download())
    for field in ("iterations", "restarts", "status", "v_cycles", "relative_residual"):
        assert getattr(info, field) == getattr(ref_info, field), field
    assert info.status == 0 and info.iterations > 1


def test_noalloc_rejects_a_system_above_the_limit(ctx):
    announce("noalloc limit", "L7 fp64")
    mesh = P.Mesh(7)   # 163,842 cells: about 21 MB of FP64 workspace
    graph = P.Graph(ctx, mesh)
    system = P.System(graph, P.PLAIN_FP64)
    seed = P.instance_seed(33)
    cond, ca, _, t0 = P.heat_instance(mesh, seed, 1.0e6, 0.3)
    system.heat_set_matrix(mesh.entry_values(cond), ca)
    d_t0 = P.DeviceArray.of(ctx, t0)
    d_f = P.DeviceArray.of(ctx, P.heat_forcing(mesh, seed, 0))
    out = P.DeviceArray(ctx, mesh.cells)
    status = P.lib.psim_heat_step_noalloc(system.handle, d_t0.ptr, d_f.ptr, C.c_double(2.63e6), C.c_double(1e-10),
                                          1000, out.ptr, None)
    assert status == 1   # PSIM_INVALID_ARGUMENT
    assert b"maximum" in P.lib.psim_last_error()
    # the array is untouched by the rejected call: a small system still matches
    announce("noalloc after rejection", "L3 fp64")
    _, _, _, _, _, ref, _, got, info = _case(ctx, 3, P.PLAIN_FP64, 34)
    assert bits_equal(got, ref) and info.status == 0

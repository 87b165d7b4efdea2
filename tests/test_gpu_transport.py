"""The implicit transport on the device: the Newton-system PCG
(solve_newton_system, heat_transport.cpp:328-429) and the Newton driver
(solve_implicit_transport, :595-772), against the CPU oracle (itself pinned
bit for bit to the compiled reference in test_oracle.py).

The device V-cycle smooths with Jacobi where the reference uses
Gauss-Seidel, so inner iterates differ; what must agree is the answer:
- PCG: held rows exactly 0, and the solution within 1e-9 relative L2 of
  the oracle's when both run to tolerance 1e-12;
- Newton: both satisfy the reference test's bounds (consistency residual
  <= 1e-6 W/m^2, conservation, dissipation), and the transport H and the
  temperatures agree within the Newton tolerance's reach: |dH| <= 1e-5
  W/m^2 and |dT| <= 1e-6 K (the residual F = h - K lap T is <= 1e-6 W/m^2
  at both answers).
"""
import ctypes as C

import numpy as np
import pytest

import planet_game_amd as P
import support as S
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu

K = 0.5 * S.RADIUS ** 2


def newton_matrix(m, slope):
    free = slope > 0.0
    diag = np.where(free, m.g_area / np.where(free, slope, 1.0) + K * m.weight_sum, 1.0)
    nb = m.g_neighbour.astype(np.int64)
    rows = np.repeat(np.arange(m.n), np.diff(m.offset.astype(np.int64)))
    off = np.where(free[rows] & free[nb], -K * m.weight, 0.0)
    return diag, off


@pytest.mark.parametrize("level", [3, 4, 5])
def test_pcg_matches_oracle(ctx, level):
    announce("newton pcg", level)
    om = S.mesh("oracle", level)
    rng = np.random.default_rng(30 + level)
    slope = np.where(np.arange(om.n) % 5 == 0, 0.0, 0.01 + 0.2 * rng.random(om.n))
    rhs = rng.standard_normal(om.n) * om.g_area
    b = np.where(slope > 0.0, rhs, 0.0)
    x_ref = np.zeros(om.n)
    it = C.c_int()
    S.oracle().check(S.oracle().newton_system(om.handle, K, slope, rhs, 1e-12, 2000, 1, x_ref, C.byref(it)))
    diag, off = newton_matrix(om, slope)
    s = P.System(P.Graph(ctx, product_mesh(level)))
    s.set_matrix(diag, off)
    held = slope <= 0.0
    x, info = s.pcg(b, hold=held, tolerance=1e-12, max_iterations=2000)
    print("level", level, "device iterations", info.iterations, "oracle", it.value, flush=True)
    assert info.status == 0 and info.relative_residual <= 1e-12
    assert np.all(x[slope <= 0.0] == 0.0)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    # the default Newton tolerance stops early and still descends
    x6, info6 = s.pcg(b, hold=held, tolerance=1e-6)
    assert info6.status == 0 and info6.iterations <= info.iterations
    assert np.linalg.norm(x6 - x_ref) <= 1e-4 * np.linalg.norm(x_ref)


def linear_case(level):
    om = S.mesh("oracle", level)
    rng = np.random.default_rng(40 + level)
    base = 250.0 + 50.0 * rng.random(om.n)
    slopes = np.where(np.arange(om.n) % 5 == 0, 0.0, 0.01 + 0.2 * rng.random(om.n))
    return om, (lambda h: (base + slopes * h, slopes.copy())), np.full(om.n, -np.inf)


def compare(ctx, level, om, fn, floor, latent=None, newton_max=None):
    dev = P.implicit_transport(P.Graph(ctx, product_mesh(level)), K, floor, fn, latent=latent)
    ref = S.implicit_transport(S.oracle(), om.handle, om.n, K, floor, fn, latent=latent)
    (sd, md, rd), (sr, mr, rr) = dev, ref
    print("level", level, "newton", rd.newton_iterations, rr.newton_iterations, "cg", rd.cg_iterations,
          rr.cg_iterations, "residual", rd.consistency_residual_W_m2, rr.consistency_residual,
          "dH", float(np.max(np.abs(sd - sr))), "dT", float(np.max(np.abs(md - mr))), flush=True)
    assert rd.consistency_residual_W_m2 <= 1e-6
    expected = 0.0 if latent is None else -float(np.sum(om.g_area * latent))
    assert abs(rd.sum_W - expected) <= 1e-12 * rd.absolute_sum_W
    if latent is None:
        assert rd.dissipation_W_K <= 0.0
    if newton_max is not None:
        assert rd.newton_iterations <= newton_max
    assert np.max(np.abs(sd - sr)) <= 1e-5
    assert np.max(np.abs(md - mr)) <= 1e-6
    return dev


@pytest.mark.parametrize("level", [3, 4, 5])
def test_linear_transport_matches_oracle(ctx, level):
    announce("implicit transport linear", level)
    om, fn, floor = linear_case(level)
    compare(ctx, level, om, fn, floor, newton_max=3)


@pytest.mark.parametrize("a", [0.04 + 0.37, 1.0e5 / 600.0])
def test_quartic_transport_matches_oracle(ctx, a):
    level = 4
    announce("implicit transport quartic", level, a)
    om = S.mesh("oracle", level)
    u = np.random.default_rng(21).random(om.n)
    fn, floor = S.quartic_response(om.n, u, a)
    src, mean, res = compare(ctx, level, om, fn, floor)
    assert np.all(mean[::7] <= 273.15)
    latent = 5.0 * np.sin(np.arange(om.n))
    compare(ctx, level, om, fn, floor, latent=latent)


def test_transport_rejects_bad_conductance(ctx):
    announce("implicit transport errors")
    om, fn, floor = linear_case(3)
    g = P.Graph(ctx, product_mesh(3))
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(P.PsimError, match="transport conductance must be finite and positive"):
            P.implicit_transport(g, bad, floor, fn)


def same_result(a, b):
    return all(getattr(a, f) == getattr(b, f) for f, _ in P.TransportResult._fields_)


@pytest.mark.parametrize("level", [3, 5])
def test_table_response_equals_host_callback(ctx, level):
    """The device-resident table and the same table evaluated by a host
    callback drive the same Newton iterates: bit-identical outputs."""
    announce("implicit transport table vs callback", level)
    m = product_mesh(level)
    h, t, floor = S.response_table(m.cells, 6, 70 + level)
    g = P.Graph(ctx, m)
    st, mt, rt = P.implicit_transport_table(g, K, floor, h, t)
    sc, mc, rc = P.implicit_transport(g, K, floor, S.table_response(h, t))
    print("level", level, "newton", rt.newton_iterations, "cg", rt.cg_iterations, "residual",
          rt.consistency_residual_W_m2, flush=True)
    assert bits_equal(st, sc) and bits_equal(mt, mc) and same_result(rt, rc)
    latent = 3.0 * np.cos(np.arange(m.cells))
    st, mt, rt = P.implicit_transport_table(g, K, floor, h, t, latent=latent)
    sc, mc, rc = P.implicit_transport(g, K, floor, S.table_response(h, t), latent=latent)
    assert bits_equal(st, sc) and bits_equal(mt, mc) and same_result(rt, rc)


@pytest.mark.parametrize("level", [4, 6])
def test_table_transport_matches_oracle(ctx, level):
    announce("implicit transport table vs oracle", level)
    om = S.mesh("oracle", level)
    h, t, floor = S.response_table(om.n, 6, 70 + level)
    dev = P.implicit_transport_table(P.Graph(ctx, product_mesh(level)), K, floor, h, t)
    ref = S.implicit_transport(S.oracle(), om.handle, om.n, K, floor, S.table_response(h, t))
    (sd, md, rd), (sr, mr, rr) = dev, ref
    print("level", level, "newton", rd.newton_iterations, rr.newton_iterations, "cg", rd.cg_iterations,
          rr.cg_iterations, "residual", rd.consistency_residual_W_m2, rr.consistency_residual,
          "dH", float(np.max(np.abs(sd - sr))), "dT", float(np.max(np.abs(md - mr))), flush=True)
    assert rd.consistency_residual_W_m2 <= 1e-6
    assert abs(rd.sum_W) <= 1e-12 * rd.absolute_sum_W
    assert rd.dissipation_W_K <= 0.0
    assert np.max(np.abs(sd - sr)) <= 1e-5
    assert np.max(np.abs(md - mr)) <= 1e-6


def test_table_transport_device_pointers(ctx):
    announce("implicit transport table device pointers")
    level = 4
    m = product_mesh(level)
    h, t, floor = S.response_table(m.cells, 4, 90)
    g = P.Graph(ctx, m)
    expect = P.implicit_transport_table(g, K, floor, h, t)
    d_h, d_t, d_f = P.DeviceArray.of(ctx, h.ravel()), P.DeviceArray.of(ctx, t.ravel()), P.DeviceArray.of(ctx, floor)
    d_s, d_m = P.DeviceArray(ctx, m.cells), P.DeviceArray(ctx, m.cells)
    res = P.implicit_transport_table_device(g, K, d_f, d_h, d_t, 4, d_s, d_m)
    ctx.synchronize()
    assert bits_equal(d_s.download(), expect[0]) and bits_equal(d_m.download(), expect[1])
    assert same_result(res, expect[2])
    with pytest.raises(P.PsimError, match="response table needs 2 to 4096 points"):
        P.implicit_transport_table_device(g, K, d_f, d_h, d_t, 1, d_s, d_m)

"""Block ILU(0) (IncompleteLU, incomplete_lu.cpp:40-146) and the implicit
tracer step (solve_implicit_tracer, heat_transport.cpp:901-1081): the
oracle's restatements against the compiled reference, bit for bit, and the
properties the reference documents (exact solves on a block-diagonal
pattern, mass conservation of the flux form, uniform c staying uniform
under divergence-free fluxes).
"""
import ctypes as C

import numpy as np
import pytest

import support as S

needs_ref = pytest.mark.skipif(not S.have_ref(), reason="oracle/_ref not built (reference sources absent)")


def nonsymmetric(m, seed):
    rng = np.random.default_rng(seed)
    off = -m.weight * rng.uniform(0.5, 1.5, m.weight.size)
    rows = np.repeat(np.arange(m.n), np.diff(m.offset.astype(np.int64)))
    diag = -np.bincount(rows, weights=off, minlength=m.n) + m.area * 1e-9
    return np.ascontiguousarray(diag), np.ascontiguousarray(off)


def blocks_of(n, size=256):
    return np.array(list(range(0, n, size)) + [n], np.uint64)


def ilu_solve(lib, m, diag, off, block_start, scale, x):
    h = lib.ilu_create(m.n, m.offset, m.g_neighbour, diag, off, block_start, block_start.size - 1, scale)
    assert h, lib.last_error().decode()
    y = np.ascontiguousarray(x, np.float64).copy()
    try:
        if lib is S.oracle():
            lib.check(lib.ilu_solve(h, y))
        else:
            lib.check(lib.ilu_solve(h, y, y.size))
    finally:
        lib.ilu_free(h)
    return y


@needs_ref
@pytest.mark.parametrize("level,scale,size", [(3, 1.0, 256), (4, 1.0 + 1e-8, 256), (5, 1.0, 256), (4, 1.0, 40962)])
def test_oracle_ilu_matches_reference(level, scale, size):
    m = S.mesh("oracle", level)
    diag, off = nonsymmetric(m, level)
    bs = blocks_of(m.n, size)
    rng = np.random.default_rng(5)
    for _ in range(3):
        x = rng.standard_normal(m.n)
        a = ilu_solve(S.oracle(), m, diag, off, bs, scale, x)
        b = ilu_solve(S.ref(), S.mesh("ref", level), diag, off, bs, scale, x)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_oracle_ilu_exact_on_isolated_rows():
    """With blocks of one row every coupling is dropped: the solve is the
    scaled diagonal's inverse, exactly."""
    m = S.mesh("oracle", 3)
    diag, off = nonsymmetric(m, 1)
    x = np.random.default_rng(2).standard_normal(m.n)
    y = ilu_solve(S.oracle(), m, diag, off, np.arange(m.n + 1, dtype=np.uint64), 1.5, x)
    assert np.array_equal(y, x / (diag * 1.5))


def test_oracle_ilu_zero_pivot_raises():
    m = S.mesh("oracle", 3)
    diag, off = nonsymmetric(m, 1)
    diag[0] = 0.0
    bs = blocks_of(m.n)
    h = S.oracle().ilu_create(m.n, m.offset, m.g_neighbour, diag, off, bs, bs.size - 1, 1.0)
    assert not h and b"zero pivot" in S.oracle().last_error()


@needs_ref
@pytest.mark.parametrize("level,layers", [(3, 1), (3, 4), (4, 3)])
def test_oracle_tracer_matches_reference(level, layers):
    m = S.mesh("oracle", level)
    case = S.tracer_case(m, layers, 10 + level)
    out, res = S.oracle_tracer(m, case)
    ref_out = np.zeros_like(out)
    it, rr, clip, div = C.c_int(), C.c_double(), C.c_double(), C.c_double()
    S.ref().check(S.ref().tracer_step(S.mesh("ref", level).handle, layers, case["conductance"], case["outflow"],
                                      case["heat_capacity"], case["layer_mass"], case["tracer"], case["dt"], 1e-12,
                                      2000, 1, ref_out, C.byref(it), C.byref(rr), C.byref(clip), C.byref(div)))
    assert np.array_equal(out.view(np.uint64), ref_out.view(np.uint64))
    assert (res.iterations, res.relative_residual, res.clipped_kg, res.column_divergence_kg_s) == (
        it.value, rr.value, clip.value, div.value)


@pytest.mark.parametrize("layers", [1, 3])
def test_oracle_tracer_conserves_mass(layers):
    m = S.mesh("oracle", 4)
    case = S.tracer_case(m, layers, 3)
    out, res = S.oracle_tracer(m, case)
    before = np.sum(case["tracer"].reshape(layers, m.n) * case["layer_mass"][None, :])
    after = np.sum(out.reshape(layers, m.n) * case["layer_mass"][None, :])
    assert abs(after - before - res.clipped_kg) <= 1e-12 * before
    assert res.relative_residual <= 1e-11 and np.all(out >= 0.0)


def test_oracle_tracer_uniform_stays_uniform():
    """One layer, fluxes with no divergence (an edge circulation around
    every cell sums to zero only for purely eddy exchange): a uniform c stays
    uniform under eddies alone."""
    m = S.mesh("oracle", 3)
    case = S.tracer_case(m, 1, 4)
    case["outflow"] = np.zeros_like(case["outflow"])
    case["tracer"] = np.full_like(case["tracer"], 1.25e-3)
    out, res = S.oracle_tracer(m, case)
    assert np.max(np.abs(out - 1.25e-3)) <= 1e-15

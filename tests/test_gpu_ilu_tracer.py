"""Block ILU(0) and the implicit tracer step on the device, through the C
ABI, against the oracle (itself pinned bit for bit to the compiled
reference's IncompleteLU and solve_implicit_tracer in test_ilu_tracer.py).

- ILU(0): factors and solves are bit-identical (each row's operations in the
  reference's order; the level schedule only decides which rows run
  together);
- tracer step: the device BiCGSTAB's dot products are summed in another
  (fixed) order than the reference's, so iterates differ by rounding; both
  solve to relative residual 1e-12, and the stepped tracer agrees within
  1e-10 of its scale, with the flux form conserving mass to rounding.
"""
import numpy as np
import pytest

import planet_game_amd as P
import support as S
from test_gpu_parity import announce, bits_equal, product_mesh
from test_ilu_tracer import blocks_of, nonsymmetric

pytestmark = pytest.mark.gpu


def oracle_ilu(m, diag, off, bs, scale, x):
    h = S.oracle().ilu_create(m.n, m.offset, m.g_neighbour, diag, off, bs, bs.size - 1, scale)
    assert h
    y = np.ascontiguousarray(x, np.float64).copy()
    S.oracle().check(S.oracle().ilu_solve(h, y))
    S.oracle().ilu_free(h)
    return y


@pytest.mark.parametrize("level,size,scale", [(3, 256, 1.0), (5, 256, 1.0 + 1e-8), (6, 256, 1.0), (4, 100, 1.0),
                                              (4, 2562, 1.0)])
def test_ilu_matches_oracle(ctx, level, size, scale):
    announce("block ilu", level, size, scale)
    om = S.mesh("oracle", level)
    diag, off = nonsymmetric(om, level)
    bs = blocks_of(om.n, size)
    f = P.Ilu(ctx, om.offset, om.g_neighbour, bs)
    f.factor(diag, off, scale)
    rng = np.random.default_rng(level)
    for _ in range(2):
        x = rng.standard_normal(om.n)
        assert bits_equal(f.solve(x), oracle_ilu(om, diag, off, bs, scale, x))


def test_ilu_errors(ctx):
    announce("block ilu errors")
    om = S.mesh("oracle", 3)
    diag, off = nonsymmetric(om, 1)
    f = P.Ilu(ctx, om.offset, om.g_neighbour, blocks_of(om.n))
    with pytest.raises(P.PsimError, match="not factored"):
        f.solve(np.ones(om.n))
    bad = diag.copy()
    bad[0] = 0.0
    with pytest.raises(P.PsimError, match="zero pivot"):
        f.factor(bad, off)
    with pytest.raises(P.PsimError, match="pattern does not match"):
        P.Ilu(ctx, om.offset, om.g_neighbour, np.array([0, 5], np.uint64))
    # offsets that would read rows past the neighbour array are refused first
    decreasing = np.array(om.offset, np.uint64)
    decreasing[5] = decreasing[3]
    with pytest.raises(P.PsimError, match="must not decrease"):
        P.Ilu(ctx, decreasing, om.g_neighbour, blocks_of(om.n))
    f.factor(diag, off)
    assert np.all(np.isfinite(f.solve(np.ones(om.n))))


@pytest.mark.parametrize("level,layers", [(3, 1), (4, 3), (5, 4), (6, 2)])
def test_tracer_matches_oracle(ctx, level, layers):
    announce("tracer step", level, layers)
    om = S.mesh("oracle", level)
    case = S.tracer_case(om, layers, 10 + level)
    expect, eres = S.oracle_tracer(om, case)
    got, res = P.tracer_step(P.Graph(ctx, product_mesh(level)), layers, case["conductance"], case["outflow"],
                             case["layer_mass"], case["tracer"], case["dt"], case["heat_capacity"])
    scale = float(np.max(np.abs(expect)))
    print("level", level, "layers", layers, "iterations", res.iterations, eres.iterations, "residual",
          res.relative_residual, eres.relative_residual, "max diff", float(np.max(np.abs(got - expect))) / scale,
          flush=True)
    assert res.relative_residual <= 1e-11
    assert np.max(np.abs(got - expect)) <= 1e-10 * scale
    assert abs(res.iterations - eres.iterations) <= 2
    assert res.column_divergence_kg_s == eres.column_divergence_kg_s
    before = np.sum(case["tracer"].reshape(layers, om.n) * case["layer_mass"][None, :])
    after = np.sum(got.reshape(layers, om.n) * case["layer_mass"][None, :])
    assert abs(after - before - res.clipped_kg) <= 1e-12 * before

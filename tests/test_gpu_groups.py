"""The agglomerated transport on the device: the cell-to-group maps
(psim_groups: area-weighted mean, highest value, broadcast;
surface_energy.cpp:1226-1247,1285-1300,1306-1324) bit for bit against the
oracle, and a whole implicit transport on the agglomerated graph
(agglomerated_transport_graph, heat_transport.cpp:520-537) whose response
broadcasts the groups' sources to cells, answers per cell and restricts back,
as the reference's surface step does, against the oracle doing the same.
"""
import numpy as np
import pytest

import planet_game_amd as P
import support as S
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu

K = 0.5 * S.RADIUS ** 2


def groups_for(ctx, level):
    a = P.Agglomeration(product_mesh(level))
    return a, P.Groups.of(ctx, a, product_mesh(level))


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_group_maps_match_oracle(ctx, level):
    announce("group maps", level)
    a, g = groups_for(ctx, level)
    m = product_mesh(level)
    rng = np.random.default_rng(50 + level)
    v = rng.standard_normal((3, m.cells)) * 250.0
    v[1, ::9] = np.nan
    v[2, ::5] = -np.inf
    mean = g.mean(v)
    high = g.max(v)
    for f in range(3):
        assert bits_equal(mean[f], S.group_mean(a.group_of_cell, m.area, a.group_area, v[f])), f
        assert bits_equal(high[f], S.group_max(a.group_of_cell, a.groups, v[f])), f
    assert np.array_equal(g.mean(np.ones(m.cells)), np.ones(a.groups))
    w = rng.standard_normal((2, a.groups))
    assert bits_equal(g.broadcast(w), np.ascontiguousarray(w[:, a.group_of_cell]))
    # device pointers, one field
    d_in = P.DeviceArray.of(ctx, np.ascontiguousarray(v[0]))
    d_out = P.DeviceArray(ctx, a.groups)
    g.mean_device(d_in.ptr, d_out.ptr)
    ctx.synchronize()
    assert bits_equal(d_out.download(), mean[0])
    d_cells = P.DeviceArray(ctx, m.cells)
    g.broadcast_device(d_out.ptr, d_cells.ptr)
    ctx.synchronize()
    assert bits_equal(d_cells.download(), mean[0][a.group_of_cell])


def test_group_map_errors(ctx):
    announce("group map errors")
    area = np.ones(6)
    with pytest.raises(P.PsimError, match="group index out of range"):
        P.Groups(ctx, [0, 1, 2, 3, 4, 5], area, np.ones(3))
    with pytest.raises(P.PsimError, match="a group has no cell"):
        P.Groups(ctx, [0, 0, 2, 2, 0, 2], area, np.ones(3))
    g = P.Groups(ctx, [0, 0, 1, 1, 2, 2], area, np.full(3, 2.0))
    with pytest.raises(ValueError):
        g.mean(np.ones(5))
    assert np.array_equal(g.mean(np.arange(6.0)), np.array([0.5, 2.5, 4.5]))


def cell_response(level):
    """The quartic tile-like response per cell (seeded), and the groups'
    response and floor built from it as the reference's surface step does."""
    m = S.mesh("oracle", level)
    u = np.random.default_rng(60 + level).random(m.n)
    fn, cell_floor = S.quartic_response(m.n, u, 0.41)
    return fn, cell_floor


@pytest.mark.parametrize("level", [4, 5])
def test_agglomerated_transport_matches_oracle(ctx, level):
    announce("agglomerated transport", level)
    a, g = groups_for(ctx, level)
    fn, cell_floor = cell_response(level)
    group, og = S.agglomerate(level)
    fine = S.mesh("oracle", level)

    def device_response(h):
        mean, slope = fn(g.broadcast(h))
        both = g.mean(np.stack([mean, slope]))
        return both[0], both[1]

    def oracle_response(h):
        mean, slope = fn(h[group])
        return (S.group_mean(group, fine.area, og.area, mean), S.group_mean(group, fine.area, og.area, slope))

    floor_dev = g.max(cell_floor)
    floor_ref = S.group_max(group, og.n, cell_floor)
    assert bits_equal(floor_dev, floor_ref)
    sd, md, rd = P.implicit_transport(P.Graph(ctx, a.graph), K, floor_dev, device_response)
    sr, mr, rr = S.implicit_transport(S.oracle(), og.handle, og.n, K, floor_ref, oracle_response)
    print("level", level, "groups", a.groups, "newton", rd.newton_iterations, rr.newton_iterations, "cg",
          rd.cg_iterations, rr.cg_iterations, "residual", rd.consistency_residual_W_m2, rr.consistency_residual,
          "dH", float(np.max(np.abs(sd - sr))), "dT", float(np.max(np.abs(md - mr))), flush=True)
    assert rd.consistency_residual_W_m2 <= 1e-6 and rr.consistency_residual <= 1e-6
    assert abs(rd.sum_W) <= 1e-12 * rd.absolute_sum_W
    assert rd.dissipation_W_K <= 0.0
    assert np.max(np.abs(sd - sr)) <= 1e-5
    assert np.max(np.abs(md - mr)) <= 1e-6
    # the cells' sources, broadcast from the groups, conserve the area integral
    cells = g.broadcast(sd)
    assert abs(np.sum(product_mesh(level).area * cells)) <= 1e-12 * np.sum(product_mesh(level).area * np.abs(cells))

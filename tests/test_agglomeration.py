"""The agglomerated transport graph (build_agglomeration,
heat_transport.cpp:467-512, behind agglomerated_transport_graph and
agglomerated_mesh, :520-543) and the cell-to-group maps around a solve on it
(surface_energy.cpp:1285-1300,1306-1324), on the CPU.

Pins:
- the oracle's agglomeration against the compiled reference, bit for bit;
- the product's host agglomeration (libpsim) against the oracle, bit for bit;
- both against what the geometry fixes: every cell's group is a coarse cell
  of largest dot product with the cell's centre (brute force over all coarse
  cells), and the reference test's own properties (test_heat_transport.cpp:
  173-212: full cover, area sum, symmetric positive weights, about four
  cells per group);
- the oracle's group mean against the reference's group areas: the mean of
  the constant 1 is exactly 1.0 only when A_c are summed in the order the
  reference summed the group's area;
- the oracle's group maximum against numpy's fmax reduction from -inf.
"""
import numpy as np
import pytest

import planet_game_amd as P
import support as S

LEVELS = [0, 1, 2, 3, 4, 5, 6]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)


@pytest.mark.skipif(not S.have_ref(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("level", LEVELS)
def test_oracle_agglomeration_matches_reference(level):
    group, g = S.agglomerate(level)
    r = S.ref_agglomeration(level)
    assert r["subdivision"] == max(level - 1, 0)
    assert np.array_equal(group.astype(np.uint64), r["group"])
    assert np.array_equal(bits(g.area), bits(r["area"]))
    assert np.array_equal(g.offset, r["offset"])
    assert np.array_equal(g.g_neighbour, r["neighbour"])
    assert np.array_equal(bits(g.weight), bits(r["weight"]))
    assert np.array_equal(bits(g.weight_sum), bits(r["weight_sum"]))


@pytest.mark.parametrize("level", LEVELS)
def test_product_agglomeration_matches_oracle(level):
    group, g = S.agglomerate(level)
    a = P.Agglomeration(P.Mesh(level, S.RADIUS))
    assert a.cells == group.size and a.groups == g.n
    assert a.coarse.subdivision == max(level - 1, 0) and a.coarse.cells == g.n
    assert np.array_equal(a.group_of_cell, group)
    assert np.array_equal(bits(a.group_area), bits(g.area))
    assert np.array_equal(a.graph.offset, g.offset)
    assert np.array_equal(a.graph.graph_neighbour, g.g_neighbour)
    assert np.array_equal(bits(a.graph.weight), bits(g.weight))
    assert np.array_equal(bits(a.graph.weight_sum), bits(g.weight_sum))
    # the coarse mesh keeps its own cell areas (agglomerated_mesh)
    assert np.array_equal(bits(a.coarse.area), bits(S.mesh("oracle", max(level - 1, 0)).area))


def nearest_value(fine_center, coarse_center, chunk=2048):
    """Largest x*x' + y*y' + z*z' over all coarse centres, per fine cell
    (element-wise products and sums in the reference's order, no FMA)."""
    f = fine_center.reshape(-1, 3)
    c = coarse_center.reshape(-1, 3)
    best = np.empty(f.shape[0])
    for s in range(0, f.shape[0], chunk):
        p = f[s:s + chunk]
        d = p[:, 0:1] * c[None, :, 0] + p[:, 1:2] * c[None, :, 1]
        d = d + p[:, 2:3] * c[None, :, 2]
        best[s:s + chunk] = d.max(axis=1)
    return best


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5, 6])
def test_agglomeration_geometry(level):
    group, g = S.agglomerate(level)
    fine = S.mesh("oracle", level)
    coarse = S.mesh("oracle", level - 1)
    f = fine.center.reshape(-1, 3)
    c = coarse.center.reshape(-1, 3)[group]
    own = f[:, 0] * c[:, 0] + f[:, 1] * c[:, 1]
    own = own + f[:, 2] * c[:, 2]
    assert np.array_equal(own, nearest_value(fine.center, coarse.center))
    # test_heat_transport.cpp:173-212
    assert group.size == fine.n and int(group.max()) < g.n
    assert np.all(np.bincount(group, minlength=g.n) > 0)
    assert abs(g.area.sum() - fine.area.sum()) <= 1e-12 * fine.area.sum()
    rows = np.repeat(np.arange(g.n), np.diff(g.offset.astype(np.int64)))
    pairs = {(int(a), int(b)): w for a, b, w in zip(rows, g.g_neighbour, g.weight)}
    assert all(a != b and w > 0.0 and pairs.get((b, a)) == w for (a, b), w in pairs.items())
    if level >= 3:   # 42/12 and 162/42 cells per group below
        assert 3.9 < fine.n / g.n < 4.1


@pytest.mark.parametrize("level", [3, 5, 6])
def test_group_mean_oracle(level):
    group, g = S.agglomerate(level)
    fine = S.mesh("oracle", level)
    group_area = S.ref_agglomeration(level)["area"] if S.have_ref() else g.area
    ones = S.group_mean(group, fine.area, group_area, np.ones(fine.n))
    assert np.array_equal(ones, np.ones(g.n))
    # a single non-zero cell per group: its area over the group's, exactly
    first = np.full(g.n, -1)
    for c in range(fine.n - 1, -1, -1):
        first[group[c]] = c
    one_hot = np.zeros(fine.n)
    one_hot[first] = 1.0
    assert np.array_equal(S.group_mean(group, fine.area, group_area, one_hot), fine.area[first] / group_area)
    # conservation of the area integral
    v = np.random.default_rng(level).standard_normal(fine.n) * 300.0
    mean = S.group_mean(group, fine.area, group_area, v)
    assert abs(np.sum(group_area * mean) - np.sum(fine.area * v)) <= 1e-12 * np.sum(fine.area * np.abs(v))


def test_group_max_oracle():
    group, g = S.agglomerate(4)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(group.size)
    v[::11] = np.nan
    v[::13] = np.inf
    v[::17] = -np.inf
    expect = np.full(g.n, -np.inf)
    np.fmax.at(expect, group, v)
    got = S.group_max(group, g.n, v)
    assert np.array_equal(got, expect)
    # a group of NaN cells only keeps -inf, as std::max from -inf does
    all_nan = np.full(group.size, np.nan)
    assert np.all(S.group_max(group, g.n, all_nan) == -np.inf)

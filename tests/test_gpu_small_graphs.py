"""Degenerate transport graphs: a single cell (no entries at all), two
cells, a path of five, and the L3 icosphere graph with two isolated cells
appended (rows with no entries beside full ones).

The first three have a one-level hierarchy, whose V-cycle is the dense
coarsest solve; the last has two levels, and its isolated cells form
singleton aggregates. Each graph's heat step (plain FP64 and mixed) is checked
against scipy's sparse direct solve of the same system. The test requires a
true relative residual of at most 1e-9 at tolerance 1e-10, and an error within
kappa times that residual, with kappa bounded by Gershgorin's discs. Chains
must equal the same steps launched one by one, bit for bit, and the assembly
and SpMV must match the slot-order loops bit for bit. A server with three
queues (the widened L5 graph of test_gpu_wide, L3 with isolated cells, and the
one-cell graph) must serve interleaved sessions bit for bit as a stand-alone
system does the same steps.
"""
import ctypes as C
import types

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import planet_game_amd as P
from test_gpu_parity import announce, bits_equal, product_mesh
from test_gpu_wide import DT, STIFF, heat_system, kappa_bound, slot_loops, wide_case

pytestmark = pytest.mark.gpu

_cases = {}


def small_graph(rows, weight, area):
    n = len(rows)
    offset = np.zeros(n + 1, np.uint64)
    offset[1:] = np.cumsum([len(r) for r in rows])
    nb = np.array([c for r in rows for c in r], np.uint64)
    w = np.asarray(weight, np.float64)
    return types.SimpleNamespace(cells=n, area=np.asarray(area, np.float64), offset=offset, graph_neighbour=nb,
                                 weight=w, weight_sum=np.array([w[offset[i]:offset[i + 1]].sum() for i in range(n)]))


def symmetric_values(g, rng, scale):
    """One lognormal value per unordered pair, the same in both directions."""
    values = np.empty(g.graph_neighbour.size)
    pair = {}
    for i in range(g.cells):
        for k in range(int(g.offset[i]), int(g.offset[i + 1])):
            key = (min(i, int(g.graph_neighbour[k])), max(i, int(g.graph_neighbour[k])))
            if key not in pair:
                pair[key] = scale * float(np.exp(0.3 * rng.standard_normal()))
            values[k] = pair[key]
    return values


def case(name):
    """(graph arrays, conductance per entry, C*A, T_old, forcings)."""
    if name in _cases:
        return _cases[name]
    rng = np.random.default_rng(sum(map(ord, name)))
    if name == "l3_isolated":
        m = product_mesh(3)
        start = m.offset.astype(np.int64)
        seed = P.instance_seed(90)
        cond_edge, ca0, _, t0 = P.heat_instance(m, seed, STIFF, 0.3)
        rows = [m.graph_neighbour[start[c]:start[c + 1]].astype(np.int64).tolist() for c in range(m.cells)]
        g = small_graph(rows + [[], []], m.weight, np.concatenate([m.area, [m.area.mean()] * 2]))
        cond = m.entry_values(cond_edge)
        ca = np.concatenate([ca0, [ca0.mean()] * 2])
        t_old = np.concatenate([t0, [250.0, 300.0]])
        forcings = np.stack([np.concatenate([P.heat_forcing(m, seed, s), [40.0, -25.0]]) for s in range(4)])
    else:
        rows = {"one_cell": [[]], "two_cells": [[1], [0]],
                "path5": [[1], [0, 2], [1, 3], [2, 4], [3]]}[name]
        n = len(rows)
        w = np.array([0.5 + rng.random() for _ in range(sum(map(len, rows)))])
        g = small_graph(rows, w, 1.0e10 * (1.0 + rng.random(n)))
        g.weight = symmetric_values(g, rng, 1.0)
        g.weight_sum = np.array([g.weight[g.offset[i]:g.offset[i + 1]].sum() for i in range(n)])
        cond = symmetric_values(g, rng, 3.0e10)
        ca = 1.0e7 * g.area * (1.0 + rng.random(n))
        t_old = 250.0 + 50.0 * rng.random(n)
        forcings = 100.0 * (rng.random((4, n)) - 0.5)
    _cases[name] = (g, cond, ca, t_old, forcings)
    return _cases[name]


NAMES = ["one_cell", "two_cells", "path5", "l3_isolated"]


@pytest.mark.parametrize("name", NAMES)
def test_small_graph_matrix_and_apply(ctx, name):
    g, cond, ca, t_old, _ = case(name)
    announce("small graph matrix", name, g.cells, "cells")
    s = P.System(P.Graph(ctx, g))
    s.heat_set_matrix(cond, ca)
    diag, off, y = slot_loops(g, cond, ca, t_old)
    d0, o0 = s.level_values(0)
    assert bits_equal(d0, diag) and bits_equal(o0, off)
    assert bits_equal(s.apply(t_old), y)


@pytest.mark.parametrize("precision", [P.PLAIN_FP64, P.MIXED])
@pytest.mark.parametrize("name", NAMES)
def test_small_graph_steps_and_chain(ctx, name, precision):
    g, cond, ca, t_old, forcings = case(name)
    announce("small graph steps", name, g.cells, "cells, precision", precision)
    s = P.System(P.Graph(ctx, g), precision)
    s.heat_set_matrix(cond, ca)
    chained, infos = s.heat_chain(t_old, forcings, DT, 1e-10, 1000)
    t = t_old
    for k, f in enumerate(forcings):
        t_next, info = s.heat_step(t, f, DT, 1e-10, 1000)
        A, b = heat_system(g, cond, ca, t, f)
        ref = spla.spsolve(A.tocsc(), b) if g.cells > 1 else b / A.toarray()[0]
        err = np.linalg.norm(t_next - ref) / np.linalg.norm(ref)
        true = np.linalg.norm(b - A @ t_next) / np.linalg.norm(b)
        print("step", k, "iterations", info.iterations, "v_cycles", info.v_cycles, "true", true, "error", err,
              flush=True)
        assert info.status == 0 and infos[k].status == 0, k
        assert (infos[k].iterations, infos[k].v_cycles) == (info.iterations, info.v_cycles), k
        assert true <= 1e-9 and err <= kappa_bound(A) * true * (1 + 1e-6) + 1e-14, k
        t = t_next
    assert bits_equal(chained, t)


@pytest.mark.parametrize("precision", [P.MIXED, P.PLAIN_FP64])
def test_wide_and_degenerate_graphs_served_equal_alone(ctx, precision):
    cases = [wide_case(STIFF), case("l3_isolated"), case("one_cell")]
    announce("served wide and degenerate graphs, precision", precision, [c[0].cells for c in cases])
    steps = 3
    graphs = [P.Graph(ctx, c[0]) for c in cases]
    server = P.Server(ctx, graphs, precision=precision, memory_budget=8 << 30)
    dev = []
    for g, cond, ca, t_old, forcings in cases:
        dev.append({"cond": P.DeviceArray.of(ctx, cond), "ca": P.DeviceArray.of(ctx, ca),
                    "t0": P.DeviceArray.of(ctx, t_old),
                    "f": [P.DeviceArray.of(ctx, forcings[k]) for k in range(steps)],
                    "out": [P.DeviceArray(ctx, g.cells) for _ in range(steps)]})
    requests = []
    for step in range(steps):
        for q, d in enumerate(dev):
            r = P.ServeRequest()
            r.arrival_s = 0.002 * len(requests)
            r.graph, r.session = q, q
            r.first, r.last = int(step == 0), int(step == steps - 1)
            r.conductance_per_entry = d["cond"].ptr if step == 0 else None
            r.capacity_area = d["ca"].ptr if step == 0 else None
            r.temperature_old = (d["t0"] if step == 0 else d["out"][step - 1]).ptr
            r.forcing_W_m2 = d["f"][step].ptr
            r.temperature_new = d["out"][step].ptr
            requests.append(r)
    records = server.run(requests, DT, 1e-10, 1000)
    for k, rec in enumerate(records):
        assert rec.status == 0 and rec.relative_residual <= 1e-10, (k, rec.status, rec.relative_residual)
    for q, ((g, cond, ca, t_old, forcings), d) in enumerate(zip(cases, dev)):
        alone = P.System(graphs[q], precision)
        alone.heat_set_matrix(cond, ca)
        t = t_old
        for step in range(steps):
            t, info = alone.heat_step(t, forcings[step], DT, 1e-10, 1000)
            assert info.status == 0
            assert bits_equal(d["out"][step].download(), t), (q, step)


def test_graph_of_isolated_nodes_from_null_arrays(ctx):
    """psim_graph_create with no entries and NULL neighbour and weight
    arrays: each node's heat step is its own scalar equation,
    T' = (C A T + dt A S) / (C A), which one coarsest solve gives exactly
    (relative 1e-15)."""
    n = 7
    announce("graph of isolated nodes from NULL arrays", n)
    rng = np.random.default_rng(11)
    area = 1.0e10 * (1.0 + rng.random(n))
    ca = 1.0e7 * area
    t_old = 250.0 + 50.0 * rng.random(n)
    forcing = 100.0 * (rng.random(n) - 0.5)
    f = C.CDLL(P.LIB_PATH).psim_graph_create
    f.restype = C.c_int
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    handle = C.c_void_p()
    offset = np.zeros(n + 1, np.uint64)
    status = f(ctx.handle, C.c_uint64(n), ptr(area), ptr(offset), None, None, ptr(np.zeros(n)), C.byref(handle))
    assert status == 0, P.lib.psim_last_error()
    assert P.lib.psim_graph_depth(handle) == 1 and P.lib.psim_graph_entries(handle) == 0
    s = C.c_void_p()
    assert P.lib.psim_system_create(handle, P.PLAIN_FP64, C.byref(s)) == 0, P.lib.psim_last_error()
    info = P.SolveInfo()
    out = np.empty(n)
    empty = np.zeros(1)
    status = P.lib.psim_heat_set_matrix(s, ptr(empty), ptr(ca), C.c_double(DT), P.HOST)
    assert status == 0, P.lib.psim_last_error()
    status = P.lib.psim_heat_step(s, ptr(t_old), ptr(forcing), C.c_double(DT), C.c_double(1e-10), 1000, ptr(out),
                                  C.byref(info), P.HOST)
    assert status == 0 and info.status == 0, P.lib.psim_last_error()
    expect = (ca * t_old + DT * area * forcing) / ca
    assert np.max(np.abs(out - expect) / np.abs(expect)) <= 1e-15
    P.lib.psim_system_destroy(s)
    P.lib.psim_graph_destroy(handle)

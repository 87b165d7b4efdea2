"""Serving (psim_server): continuous batching on one persistent launch.

The property under test is that each served request's result is the same
alone and in the mix: bit for bit against a stand-alone system of the same
graph and precision doing the same steps with psim_heat_step (sessions that
continue a resident state included). Admission is exercised with a pool no
larger than the slots, so later sessions wait for a system to free up.
"""
import numpy as np
import pytest

import planet_game_amd as P
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu

DT = 2.63e6
_graphs = {}


def graph(ctx, level):
    if level not in _graphs:
        _graphs[level] = P.Graph(ctx, product_mesh(level))
    return _graphs[level]


class Session:
    """A session: its inputs, one output buffer per step, resident on the device."""

    def __init__(self, ctx, level, steps, index):
        m = product_mesh(level)
        self.level, self.steps = level, steps
        seed = P.instance_seed(7000 + index)
        cond, ca, _, t0 = P.heat_instance(m, seed, 1.0e6, 0.3)
        self.d_cond = P.DeviceArray.of(ctx, m.entry_values(cond))
        self.d_ca = P.DeviceArray.of(ctx, ca)
        self.d_t0 = P.DeviceArray.of(ctx, t0)
        self.d_f = [P.DeviceArray.of(ctx, P.heat_forcing(m, seed, s)) for s in range(steps)]
        self.d_out = [P.DeviceArray(ctx, m.cells) for _ in range(steps)]

    def request(self, graph_index, session_id, step, arrival):
        r = P.ServeRequest()
        r.arrival_s = arrival
        r.graph = graph_index
        r.session = session_id
        r.first = int(step == 0)
        r.last = int(step == self.steps - 1)
        r.conductance_per_entry = self.d_cond.ptr if step == 0 else None
        r.capacity_area = self.d_ca.ptr if step == 0 else None
        r.temperature_old = (self.d_t0 if step == 0 else self.d_out[step - 1]).ptr
        r.forcing_W_m2 = self.d_f[step].ptr
        r.temperature_new = self.d_out[step].ptr
        return r

    def alone(self, ctx, precision, tolerance):
        s = P.System(graph(ctx, self.level), precision)
        s.heat_set_matrix_device(self.d_cond, self.d_ca, DT)
        out, src = [], self.d_t0
        for step in range(self.steps):
            dst = P.DeviceArray(ctx, product_mesh(self.level).cells)
            s.heat_step_device(src.ptr, self.d_f[step].ptr, dst.ptr, DT, tolerance, 1000)
            out.append(dst.download())
            src = dst
        return out


def serve(ctx, levels, plan, precision, budget, tolerance=1e-10, spacing=0.002):
    """plan: list of (level, steps); requests interleave the sessions' steps."""
    server = P.Server(ctx, [graph(ctx, L) for L in levels], precision=precision, memory_budget=budget)
    sessions = [Session(ctx, L, steps, i) for i, (L, steps) in enumerate(plan)]
    order = []
    for step in range(max(s.steps for s in sessions)):
        for i, s in enumerate(sessions):
            if step < s.steps:
                order.append((i, step))
    requests = [sessions[i].request(levels.index(sessions[i].level), i, step, k * spacing)
                for k, (i, step) in enumerate(order)]
    records = server.run(requests, DT, tolerance, 1000)
    return server, sessions, order, records


@pytest.mark.parametrize("precision", [P.MIXED, P.PLAIN_FP64])
def test_served_results_equal_alone(ctx, precision):
    levels = [3, 4, 5, 6]
    plan = [(3, 1), (4, 3), (5, 2), (6, 3), (3, 4), (6, 1), (5, 1), (4, 1)]
    announce("serve alone vs mix", precision, plan)
    server, sessions, order, records = serve(ctx, levels, plan, precision, 2 << 30)
    print("plan", server.plan(), flush=True)
    for (i, step), rec in zip(order, records):
        assert rec.status == 0 and rec.relative_residual <= 1e-10, (i, step, rec.status, rec.relative_residual)
        assert rec.completed_s >= rec.submitted_s >= rec.arrival_s
    for i, s in enumerate(sessions):
        expect = s.alone(ctx, precision, 1e-10)
        for step in range(s.steps):
            assert bits_equal(s.d_out[step].download(), expect[step]), (i, s.level, step)


def test_admission_waits_for_a_free_system(ctx):
    """A budget of one system per slot: sessions beyond the pool wait for a
    released system, and every result still matches the stand-alone run."""
    levels = [3, 4]
    announce("serve admission")
    probe = P.Server(ctx, [graph(ctx, L) for L in levels], precision=P.MIXED, memory_budget=0)
    plan0 = probe.plan()
    assert all(p["pool"] == p["slots"] for p in plan0), plan0
    minimum = sum(p["slots"] * p["system_bytes"] for p in plan0)
    del probe
    plan = [(3, 2)] * (plan0[0]["slots"] + 6) + [(4, 2)] * (plan0[1]["slots"] + 3)
    server, sessions, order, records = serve(ctx, levels, plan, P.MIXED, minimum, spacing=0.0)
    pools = [p["pool"] for p in server.plan()]
    assert pools == [p["slots"] for p in plan0], pools
    with pytest.raises(P.PsimError, match="below one system per slot"):
        P.Server(ctx, [graph(ctx, L) for L in levels], precision=P.MIXED, memory_budget=minimum - 1)
    for rec in records:
        assert rec.status == 0 and rec.relative_residual <= 1e-10
    for s in sessions[:: max(1, len(sessions) // 6)]:
        expect = s.alone(ctx, P.MIXED, 1e-10)
        for step in range(s.steps):
            assert bits_equal(s.d_out[step].download(), expect[step])


def test_serve_rejects_bad_streams(ctx):
    announce("serve error paths")
    server = P.Server(ctx, [graph(ctx, 3)], precision=P.MIXED, memory_budget=1 << 30)
    s = Session(ctx, 3, 2, 99)
    first, second = s.request(0, 1, 0, 0.0), s.request(0, 1, 1, 0.001)
    late_first = s.request(0, 1, 0, 0.002)
    with pytest.raises(P.PsimError, match="sorted by arrival"):
        server.run([late_first, second])
    with pytest.raises(P.PsimError, match="continues before it starts"):
        server.run([second])
    with pytest.raises(P.PsimError, match="no last request"):
        server.run([first])
    bad = s.request(0, 1, 0, 0.0)
    bad.graph = 3
    with pytest.raises(P.PsimError, match="graph index out of range"):
        server.run([bad])
    records = server.run([first, second])
    assert all(r.status == 0 for r in records)

"""Batched solves (psim_batch): many systems of mixed mesh levels and
precisions in one launch of the persistent solver.

The property under test is that a system's result is the same alone and in
the mix: bit for bit against its own psim_heat_step, whatever the workgroup
allocation (default, all-ones, the autotuner's). The plain FP64 members are
also checked against the CPU oracle (1e-12 relative L2 at tolerance 1e-13),
the mixed ones against the FP64 oracle (1e-8 at tolerance 1e-10).
"""
import numpy as np
import pytest

import planet_game_amd as P
import support as S
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu

DT = 2.63e6
_graphs = {}


def graph(ctx, level):
    if level not in _graphs:
        _graphs[level] = P.Graph(ctx, product_mesh(level))
    return _graphs[level]


class Member:
    """One instance: its system, resident inputs and output buffer."""

    def __init__(self, ctx, level, precision, index):
        m = product_mesh(level)
        self.level, self.precision, self.mesh = level, precision, m
        self.seed = P.instance_seed(1000 + index)
        cond, ca, _, t0 = P.heat_instance(m, self.seed, 1.0e6, 1.0 if precision == P.MIXED else 0.3)
        self.cond, self.ca, self.t0 = cond, ca, t0
        self.forcing = [P.heat_forcing(m, self.seed, s) for s in range(2)]
        self.system = P.System(graph(ctx, level), precision)
        self.d_cond = P.DeviceArray.of(ctx, m.entry_values(cond))
        self.d_ca = P.DeviceArray.of(ctx, ca)
        self.d_t0 = P.DeviceArray.of(ctx, t0)
        self.d_f = [P.DeviceArray.of(ctx, f) for f in self.forcing]
        self.d_out = P.DeviceArray(ctx, m.cells)

    def alone(self, step, tolerance, t_in=None):
        """The system's own psim_heat_step on the same inputs."""
        self.system.heat_set_matrix_device(self.d_cond, self.d_ca, DT)
        src = self.d_t0 if t_in is None else t_in
        out = P.DeviceArray(self.d_out.ctx, self.mesh.cells)
        self.system.heat_step_device(src.ptr, self.d_f[step].ptr, out.ptr, DT, tolerance, 1000)
        return out.download()


def members(ctx, spec):
    out = []
    for index, (level, precision) in enumerate(spec):
        announce("batch member", index, "level", level, "precision", precision)
        out.append(Member(ctx, level, precision, index))
    return out


MIX = [(3, P.PLAIN_FP64), (4, P.MIXED), (5, P.PLAIN_FP64), (6, P.MIXED), (6, P.PLAIN_FP64), (3, P.MIXED),
       (4, P.PLAIN_FP64), (5, P.MIXED)]


def run_batch(batch, ms, step, tolerance, assemble, t_in=None):
    infos = batch.heat_step([m.d_t0 if t_in is None else t_in[i] for i, m in enumerate(ms)],
                            [m.d_f[step] for m in ms], [m.d_out for m in ms], DT, tolerance, 1000,
                            conductance=[m.d_cond for m in ms] if assemble else None,
                            capacity=[m.d_ca for m in ms] if assemble else None)
    return [m.d_out.download() for m in ms], infos


def test_batch_results_equal_alone(ctx):
    announce("batch alone vs mix", len(MIX), "members")
    ms = members(ctx, MIX)
    batch = P.Batch(ctx, [m.system for m in ms])
    groups = batch.groups
    assert len(groups) == len(ms) and min(groups) >= 1 and sum(groups) <= batch.capacity
    # step 0 with assembly in the launch
    mixed, infos = run_batch(batch, ms, 0, 1e-10, assemble=True)
    for m, x, info in zip(ms, mixed, infos):
        assert info.status == 0 and info.relative_residual <= 1e-10, (m.level, m.precision, info.status)
        assert bits_equal(x, m.alone(0, 1e-10)), (m.level, m.precision)
    # step 1 on the kept matrices, from step 0's T (device-resident states)
    states = [P.DeviceArray.of(ctx, x) for x in mixed]
    second, infos = run_batch(batch, ms, 1, 1e-10, assemble=False, t_in=states)
    for m, x, s in zip(ms, second, states):
        assert bits_equal(x, m.alone(1, 1e-10, t_in=s)), (m.level, m.precision)


def test_batch_allocation_does_not_change_bits(ctx):
    announce("batch allocation invariance")
    ms = members(ctx, MIX[:6])
    batch = P.Batch(ctx, [m.system for m in ms])
    ref, _ = run_batch(batch, ms, 0, 1e-11, assemble=True)
    for groups in ([1] * len(ms), [3, 1, 7, 2, 40, 1]):
        batch.groups = groups
        assert batch.groups == groups
        got, _ = run_batch(batch, ms, 0, 1e-11, assemble=True)
        assert all(bits_equal(a, b) for a, b in zip(ref, got)), groups
    batch.groups = None
    best = batch.autotune([m.d_t0 for m in ms], [m.d_f[0] for m in ms], [m.d_out for m in ms], DT, 1e-11,
                          conductance=[m.d_cond for m in ms], capacity=[m.d_ca for m in ms])
    assert best > 0.0 and sum(batch.groups) <= batch.capacity
    assert all(bits_equal(a, m.d_out.download()) for a, m in zip(ref, ms))
    got, _ = run_batch(batch, ms, 0, 1e-11, assemble=True)
    assert all(bits_equal(a, b) for a, b in zip(ref, got))


def test_batch_against_oracle(ctx):
    """Plain members at 1e-13 match the oracle to 1e-12; mixed members at
    1e-10 are within 1e-8 of the FP64 oracle."""
    spec = [(4, P.PLAIN_FP64), (5, P.MIXED), (6, P.PLAIN_FP64), (6, P.MIXED)]
    announce("batch vs oracle", spec)
    ms = members(ctx, spec)
    batch = P.Batch(ctx, [m.system for m in ms])
    for m in ms:
        tol = 1e-13 if m.precision == P.PLAIN_FP64 else 1e-10
        om = S.mesh("oracle", m.level)
        diag, off = S.heat_matrix(om, m.cond, m.ca)
        mg = S.Multigrid(S.oracle(), om)
        mg.set_matrix(diag, off)
        m.ref, _ = mg.bicgstab(diag, off, S.heat_rhs(om, m.ca, m.t0, m.forcing[0]), 1e-13, 2000, 8)
        m.tol = tol
    plain, _ = run_batch(batch, ms, 0, 1e-13, assemble=True)
    mixed, _ = run_batch(batch, ms, 0, 1e-10, assemble=True)
    for m, xp, xm in zip(ms, plain, mixed):
        x = xp if m.precision == P.PLAIN_FP64 else xm
        err = np.linalg.norm(x - m.ref) / np.linalg.norm(m.ref)
        print("level", m.level, "precision", m.precision, "error", err, flush=True)
        assert err <= (1e-12 if m.precision == P.PLAIN_FP64 else 1e-8), (m.level, m.precision, err)


def test_batch_errors(ctx):
    announce("batch error paths")
    ms = members(ctx, [(3, P.PLAIN_FP64), (4, P.PLAIN_FP64)])
    batch = P.Batch(ctx, [m.system for m in ms])
    with pytest.raises(P.PsimError, match="null argument"):
        batch.heat_step([ms[0].d_t0, None], [m.d_f[0] for m in ms], [m.d_out for m in ms])
    with pytest.raises(P.PsimError, match="has not been called"):
        batch.heat_step([m.d_t0 for m in ms], [m.d_f[0] for m in ms], [m.d_out for m in ms])
    with pytest.raises(P.PsimError, match="both conductance and capacity"):
        batch.heat_step([m.d_t0 for m in ms], [m.d_f[0] for m in ms], [m.d_out for m in ms],
                        conductance=[m.d_cond for m in ms], capacity=[ms[0].d_ca, None])
    with pytest.raises(P.PsimError, match="at least one workgroup"):
        batch.groups = [0, 1]
    with pytest.raises(P.PsimError, match="co-resident capacity"):
        batch.groups = [batch.capacity, 1]
    with pytest.raises(P.PsimError, match="twice"):
        P.Batch(ctx, [ms[0].system, ms[0].system])
    other = P.Context(0)
    with pytest.raises(P.PsimError, match="share its context"):
        P.Batch(other, [m.system for m in ms])

#!/usr/bin/env python3
"""Headline benchmark: implicit heat-transport solves per second at L6 to a
true relative residual of 1e-10, with p50/p99 per-solve latency.

Workload (one bench "step" = one pass): the L6 icosahedral mesh (40,962
cells, 122,880 edges, radius 6,371 km), one seeded instance of the literal
heat-step distributions (K_e = 1e6 (1 + 0.5 sin^2 lat) exp(0.3 z) m^2/s,
30 % land, C = 1e7 / 4e8 J/m^2/K, dt = 2.63e6 s), the matrix set once
(device assembly, Galerkin values, coarse inverse), then 120 chained monthly
steps, each step's T feeding the next, solved with the plain FP64 path.
Inputs (C A, K per entry, T_0 and all 120 forcings) are resident in HBM
before the timed region.

One pass is ONE device launch (psim_heat_chain): the persistent solver
assembles the matrix and then runs the 120 solves back to back, each step's
T read on the device by the next. Prints one JSON line. `value` is solves/s
over the timed passes (wall clock between two stream synchronisations);
p50/p99 are per-solve durations on the device clock inside the chained
launch (the same passes repeated with per-step records);
`one_launch_per_solve` is the same work issued as 120 separate launches of
psim_heat_step (HIP events per launch). `roofline` relates the algorithmic
bytes of a pass to the mean duration of its launch (HIP events on the
library's stream); `cpu_baseline` times the CPU oracle (a bit-identical
restatement of the reference algorithm) on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "implicit heat-transport solves/sec at L6, rel resid 1e-10; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
DT = 2.63e6
TOL = 1e-10
CHAIN = 120


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "program_kernel_l6_chain_traffic.json")
BATCH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "program_kernel_batch256_mixed_traffic.json")
COPY_BYTES = 256 << 20
_T0 = time.perf_counter()


def progress(what):
    """One stderr line per finished section, so a long run shows its steps."""
    print(f"bench: {what} ({time.perf_counter() - _T0:.1f} s)", file=sys.stderr, flush=True)




def measured_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch from a committed counter profile (FETCH_SIZE and
    WRITE_SIZE, each scaled by the copy-kernel calibration), or None when
    that profile is absent."""
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def algorithmic_bytes(levels, v_cycles, restarts, mixed=False):
    """Bytes one solve must move at least (int32 indices, each array once per
    operation, neighbour gathers from cache). The BiCGSTAB vectors are FP64
    in both forms; the V-cycle's values are FP64 (plain) or, in the mixed
    form's split V-cycle, FP64 for the first level-0 sweep and FP32 after it:
      SpMV at level j, v-byte values   B_A(j, v) = (3 v + 4) n_j + (v + 4) nnz_j + 4
      V-cycle, plain     sum_{j<coarsest} (3 B_A(j, 8) + 8 n_j + 16 n_{j+1}) + 8 n_c^2
      V-cycle, split     B_A(0, 8) + 2 B_A(0, 4) + 28 n_0 + 8 n_1     (r1 out and in, w and
                                                                       D^-1 again, x1 + y out)
                         + sum_{1<=j<coarsest} (3 B_A(j, 4) + 4 n_j + 8 n_{j+1}) + 8 n_c^2
      half-iteration     B_A(0, 8) + B_V + 88 n (the halves' vector updates)
      per solve          48 n (RHS assembly and ||b||) + halves
                         + (1 + restarts)(B_A(0, 8) + 16 n) (true residuals)
    """
    def b_a(n, nnz, w):
        return (3 * w + 4) * n + (w + 4) * nnz + 4

    n0, nnz0 = levels[0]
    b_v = 8 * levels[-1][0] ** 2
    for j in range(len(levels) - 1):
        n, nnz = levels[j]
        nn = levels[j + 1][0]
        if not mixed:
            b_v += 3 * b_a(n, nnz, 8) + 8 * n + 16 * nn
        elif j == 0:
            b_v += b_a(n, nnz, 8) + 2 * b_a(n, nnz, 4) + 28 * n + 8 * nn
        else:
            b_v += 3 * b_a(n, nnz, 4) + 4 * n + 8 * nn
    half = b_a(n0, nnz0, 8) + b_v + 88 * n0
    return 48 * n0 + v_cycles * half + (1 + restarts) * (b_a(n0, nnz0, 8) + 16 * n0)


def set_matrix_bytes(levels, edges, mixed=False):
    """Assembly (read C A, K_e, w; write diag and off), Galerkin sweeps and
    the dense coarsest level; the mixed form also writes FP32 copies of every
    level's diagonal, inverse diagonal and off-diagonals."""
    n0, nnz0 = levels[0]
    total = 16 * n0 + 8 * edges + 16 * nnz0 + 8 * levels[-1][0] ** 2
    for j in range(len(levels) - 1):
        n, nnz = levels[j]
        nc, nnzc = levels[j + 1]
        total += 12 * n + 12 * nnz + 8 * nc + 8 * nnzc
    if mixed:
        total += sum(8 * n + 4 * nnz for n, nnz in levels)
    return total


def step_errors_vs_oracle(mesh, cond, ca, temps, forcings):
    """Each GPU step's error against the oracle solving that step (BiCGSTAB to
    1e-14) from the GPU's own T_k: relative L2 over the chain."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    om = S.mesh("oracle", mesh.subdivision)
    diag, off = S.heat_matrix(om, cond, ca)
    mg = S.Multigrid(S.oracle(), om)
    mg.set_matrix(diag, off)
    errors = []
    for k in range(len(temps) - 1):
        ref, _ = mg.bicgstab(diag, off, S.heat_rhs(om, ca, temps[k], forcings[k]), 1e-14, 1000, 1)
        errors.append(float(np.linalg.norm(temps[k + 1] - ref) / np.linalg.norm(ref)))
    return {"max": max(errors), "median": float(np.median(errors)), "first": errors[0], "last": errors[-1],
            "oracle": "FP64 oracle converged to 1e-14 from the same T_k"}


def oracle_chain(mesh, cond, ca, t0, forcings, tol):
    """The oracle's 120-step chain at tolerance `tol` (final T)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    om = S.mesh("oracle", mesh.subdivision)
    diag, off = S.heat_matrix(om, cond, ca)
    mg = S.Multigrid(S.oracle(), om)
    mg.set_matrix(diag, off)
    t = t0.copy()
    for step in range(CHAIN):
        t, _ = mg.bicgstab(diag, off, S.heat_rhs(om, ca, t, forcings[step]), tol, 1000, 1)
    return t


def host_cpu():
    """The cores this process may use and what they are: the affinity mask,
    nproc, the cgroup's CPU quota where one is set, and lscpu's model name."""
    import subprocess

    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        info["nproc"] = None
    for path in ("/sys/fs/cgroup/cpu.max", "/sys/fs/cgroup/cpu/cpu.cfs_quota_us"):
        try:
            with open(path) as f:
                info["cgroup_" + os.path.basename(path)] = f.read().strip()
        except OSError:
            pass
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        info["model"] = next((line.split(":", 1)[1].strip() for line in text.splitlines()
                              if line.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        info["model"] = platform.processor() or platform.machine()
    return info


def cpu_baseline(mesh, cond, ca, t0, forcings, seconds, workers=1):
    """The reference CPU path on `workers` host threads: the compiled
    reference's own BiCGSTAB template and GraphMultigrid (oracle/_ref, built
    from the reference sources where they were available) when present, else
    the oracle (a bit-identical restatement of the same algorithm). Each
    solve: RHS assembly, solve to 1e-10, true residual; the matrix is set once
    per chain (as on the GPU). Returns (rate, solves, chains, elapsed, first
    chain's final T, kind)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import support as S

    o = S.oracle()
    kind = "reference" if S.have_ref() else "port"
    lib = S.ref() if kind == "reference" else o
    om = S.mesh("oracle", mesh.subdivision)
    lm = S.mesh("ref", mesh.subdivision) if kind == "reference" else om
    diag, off = S.heat_matrix(om, cond, ca)
    mg = S.Multigrid(lib, lm)
    solves, chains = 0, 0
    start = time.perf_counter()
    final = None
    while True:
        mg.set_matrix(diag, off)
        t = t0.copy()
        for step in range(CHAIN):
            b = S.heat_rhs(om, ca, t, forcings[step])
            t, _ = mg.bicgstab(diag, off, b, TOL, 1000, workers)
            o.lib.orc_relative_residual(om.handle, diag, off, b, t)
            solves += 1
        chains += 1
        if final is None:
            final = t
        if time.perf_counter() - start >= seconds:
            break
    elapsed = time.perf_counter() - start
    return solves / elapsed, solves, chains, elapsed, final, kind


def l4_chain_section(P, ctx, seconds=3.0):
    """The L4 chain (2,562 cells), the literal distributions: the
    120-step chain on the GPU (plain FP64, one launch per pass) beside the
    reference CPU path on a bounded sample of the same chain."""
    mesh = P.Mesh(4)
    seed = P.instance_seed(0)
    cond, ca, _, t0 = P.heat_instance(mesh, seed, 1.0e6, 0.3)
    forcings = [P.heat_forcing(mesh, seed, s) for s in range(CHAIN)]
    graph = P.Graph(ctx, mesh)
    system = P.System(graph, P.PLAIN_FP64)
    d_cond = P.DeviceArray.of(ctx, mesh.entry_values(cond))
    d_ca = P.DeviceArray.of(ctx, ca)
    d_forcing = P.DeviceArray.of(ctx, np.concatenate(forcings))
    d_t0 = P.DeviceArray.of(ctx, t0)
    d_out = P.DeviceArray(ctx, mesh.cells)
    infos = (P.SolveInfo * CHAIN)()

    def run_pass(collect):
        system.heat_chain_device(d_t0.ptr, d_forcing.ptr, CHAIN, d_out.ptr, DT, TOL, 1000,
                                 step_infos=infos if collect else None,
                                 conductance_per_entry=d_cond.ptr, capacity_area=d_ca.ptr)

    run_pass(True)
    converged = all(i.status == 0 and i.relative_residual <= TOL for i in infos)
    step_ms = np.array([i.device_ms for i in infos])
    ctx.synchronize()
    passes = 20
    t = time.perf_counter()
    for _ in range(passes):
        run_pass(False)
    ctx.synchronize()
    gpu_rate = passes * CHAIN / (time.perf_counter() - t)
    rate, solves, _, elapsed, cpu_final, kind = cpu_baseline(mesh, cond, ca, t0, forcings, seconds, 1)
    return {"config": {"workload": "L4 plain FP64: set_matrix + 120 chained monthly implicit heat steps per pass",
                       "cells": mesh.cells, "tolerance": TOL, "workgroups": system.workgroups},
            "gpu_solves_per_s": round(gpu_rate, 1),
            "gpu_p50_ms": round(float(np.percentile(step_ms, 50)), 5),
            "gpu_p99_ms": round(float(np.percentile(step_ms, 99)), 5),
            "converged": converged,
            "cpu_solves_per_s": round(rate, 1), "cpu_kind": kind, "cpu_threads": 1,
            "cpu_sample": f"{solves} solves in {elapsed:.1f} s"}


def fv_bytes(fields, n, edges, directed):
    """Algorithmic bytes of the operator sweeps for F fields (int32 indices,
    geometry read once per launch, neighbour gathers from cache):
      gradient    24 F n + 4 n + 20 nnz   (phi in, east and north out; row,
                                           neighbour, two float normals, distance)
      divergence  F (8 E + 8 n) + 12 n + 12 nnz  (flux in, result out; area,
                                           row; edge id and signed length)"""
    return 24 * fields * n + 4 * n + 20 * directed, fields * (8 * edges + 8 * n) + 12 * n + 12 * directed


def fv_section(P, ctx, mesh, with_cpu, cpu_workers, repeats=50):
    """The operator sweeps beside the L6 chain: 64 seeded FP64 fields through the LSQ
    gradient and the edge-flux divergence at L6 (cell scalars and edge fluxes
    u - 0.5, as test_operator_determinism.cpp:15-26), one launch per operator
    for all 64 fields; HIP events on the library's stream around `repeats`
    back-to-back launches (the mean launch-to-launch time)."""
    fields = 64
    n, e = mesh.cells, mesh.edges
    rng = np.random.default_rng(20261205)
    phi = rng.random(fields * n) - 0.5
    flux = rng.random(fields * e) - 0.5
    fv = P.FV(ctx, mesh)
    d_phi, d_flux = P.DeviceArray.of(ctx, phi), P.DeviceArray.of(ctx, flux)
    d_east, d_north, d_div = (P.DeviceArray(ctx, fields * n) for _ in range(3))
    out = {"config": {"workload": "64 FP64 fields at L6: LSQ gradient and edge-flux divergence, one launch per "
                                  "operator for all 64 fields", "cells": n, "edges": e, "fields": fields,
                      "repeats": repeats}}
    timings = {}
    for name, call in (("gradient", lambda: fv.gradient_device(P.F64, fields, d_phi.ptr, d_east.ptr, d_north.ptr)),
                       ("divergence", lambda: fv.divergence_device(P.F64, fields, d_flux.ptr, d_div.ptr))):
        call()
        ctx.events(2)
        ctx.synchronize()
        ctx.record(0)
        for k in range(repeats):
            call()
        ctx.record(1)
        ctx.synchronize()
        timings[name] = ctx.elapsed_ms(0, 1) / repeats
    g_bytes, d_bytes = fv_bytes(fields, n, e, mesh.directed)
    for name, nbytes in (("gradient", g_bytes), ("divergence", d_bytes)):
        gbs = nbytes / (timings[name] / 1e3) / 1e9
        out[name] = {"mean_launch_ms": round(timings[name], 5), "algorithmic_bytes": nbytes,
                     "achieved_GBps": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["fields_per_s"] = round(fields / ((timings["gradient"] + timings["divergence"]) / 1e3), 1)
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import support as S

        o = S.oracle()
        om = S.mesh("oracle", mesh.subdivision)
        east_ref, north_ref, div_ref = np.zeros(fields * n), np.zeros(fields * n), np.zeros(fields * n)
        t = time.perf_counter()
        for f in range(fields):
            o.gradient_f64(om.handle, phi[f * n:(f + 1) * n], east_ref[f * n:(f + 1) * n],
                           north_ref[f * n:(f + 1) * n], cpu_workers)
            o.divergence_f64(om.handle, flux[f * e:(f + 1) * e], div_ref[f * n:(f + 1) * n], cpu_workers)
        cpu_s = time.perf_counter() - t
        same = (np.array_equal(d_east.download().view(np.uint64), east_ref.view(np.uint64))
                and np.array_equal(d_north.download().view(np.uint64), north_ref.view(np.uint64))
                and np.array_equal(d_div.download().view(np.uint64), div_ref.view(np.uint64)))
        out["cpu_oracle"] = {"fields_per_s": round(fields / cpu_s, 1), "workers": cpu_workers,
                             "bit_identical_to_oracle": bool(same)}
    return out


def mixed_section(P, ctx, mesh, graph, levels, with_oracle):
    """Mixed precision at L6 on the widened log-normal spread (sigma = 1.0):
    the 120-step chain in one launch, against the plain FP64 path on the same
    inputs, with each step's error against the FP64 oracle solving that step
    from the same T (the step's own error) and the trajectory's drift."""
    seed = P.instance_seed(0)
    cond, ca, _, t0 = P.heat_instance(mesh, seed, 1.0e6, 1.0)
    forcings = [P.heat_forcing(mesh, seed, s) for s in range(CHAIN)]
    d_cond = P.DeviceArray.of(ctx, mesh.entry_values(cond))
    d_ca = P.DeviceArray.of(ctx, ca)
    d_f = P.DeviceArray.of(ctx, np.concatenate(forcings))
    d_t0 = P.DeviceArray.of(ctx, t0)
    d_out = P.DeviceArray(ctx, mesh.cells)
    out = {}
    for name, precision in (("mixed", P.MIXED), ("plain_fp64", P.PLAIN_FP64)):
        system = P.System(graph, precision)
        infos = (P.SolveInfo * CHAIN)()
        system.heat_chain_device(d_t0.ptr, d_f.ptr, CHAIN, d_out.ptr, DT, TOL, 1000, step_infos=infos,
                                 conductance_per_entry=d_cond.ptr, capacity_area=d_ca.ptr)
        census = list(infos)
        ctx.events(2)
        ctx.synchronize()
        ctx.record(0)
        for _ in range(3):
            system.heat_chain_device(d_t0.ptr, d_f.ptr, CHAIN, d_out.ptr, DT, TOL, 1000,
                                     conductance_per_entry=d_cond.ptr, capacity_area=d_ca.ptr)
        ctx.record(1)
        ctx.synchronize()
        ms = ctx.elapsed_ms(0, 1) / 3
        mixed = precision == P.MIXED
        nbytes = set_matrix_bytes(levels, mesh.edges, mixed) + sum(
            algorithmic_bytes(levels, i.v_cycles, i.restarts, mixed) for i in census)
        out[name] = {
            "solves_per_s": round(CHAIN / ms * 1e3, 2),
            "iterations_per_solve": float(np.mean([i.iterations for i in census])),
            "restarts_per_solve": float(np.mean([i.restarts for i in census])),
            "v_cycles_per_solve": float(np.mean([i.v_cycles for i in census])),
            "max_true_residual": float(max(i.relative_residual for i in census)),
            "converged": all(i.status == 0 for i in census),
            "launch_GBps": round(nbytes / (ms / 1e3) / 1e9, 2),
        }
        if precision == P.MIXED:
            # each step's T (one launch per step, same kernels) for the error
            # report; the launches are timed too (one launch per solve)
            steps = [P.DeviceArray(ctx, mesh.cells) for _ in range(CHAIN)]
            system.heat_set_matrix_device(d_cond, d_ca, DT)
            ctx.events(2)
            ctx.synchronize()
            ctx.record(0)
            src = d_t0
            for k in range(CHAIN):
                system.heat_step_device(src.ptr, d_f.offset(k * mesh.cells), steps[k].ptr, DT, TOL, 1000)
                src = steps[k]
            ctx.record(1)
            ctx.synchronize()
            out[name]["one_launch_per_solve_solves_per_s"] = round(CHAIN / ctx.elapsed_ms(0, 1) * 1e3, 2)
            mixed_steps = [s.download() for s in steps]
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import support as S

        om = S.mesh("oracle", mesh.subdivision)
        diag, off = S.heat_matrix(om, cond, ca)
        mg = S.Multigrid(S.oracle(), om)
        mg.set_matrix(diag, off)
        errors = []
        prev = t0
        for k in range(CHAIN):
            ref, _ = mg.bicgstab(diag, off, S.heat_rhs(om, ca, prev, forcings[k]), 1e-13, 1000, 1)
            errors.append(float(np.linalg.norm(mixed_steps[k] - ref) / np.linalg.norm(ref)))
            prev = mixed_steps[k]
        out["mixed"]["error_vs_fp64_oracle_per_step"] = {
            "max": max(errors), "median": float(np.median(errors)), "first": errors[0], "last": errors[-1],
            "oracle": "FP64 oracle converged to 1e-13 from the same T_k"}
    out["config"] = {"workload": "L6, sigma = 1.0, set_matrix + 120 chained steps per launch", "tolerance": TOL}
    return out


def batched_section(P, ctx, meshes, graphs, hierarchy_levels, precision=None, autotune=True):
    """256 independent instances, L3:L4:L5:L6 = 64 each, own seeds, matrix
    assembled in the same launch; default allocation and the autotuner's.
    Mixed precision unless another precision is given."""
    precision = P.MIXED if precision is None else precision
    members = []
    for i, level in enumerate([3, 4, 5, 6] * 64):
        m = meshes[level]
        seed = P.instance_seed(1 + i)
        cond, ca, _, t0 = P.heat_instance(m, seed, 1.0e6, 0.3)
        members.append(dict(level=level, system=P.System(graphs[level], precision),
                            cond=P.DeviceArray.of(ctx, m.entry_values(cond)), ca=P.DeviceArray.of(ctx, ca),
                            t0=P.DeviceArray.of(ctx, t0), f=P.DeviceArray.of(ctx, P.heat_forcing(m, seed, 0)),
                            out=P.DeviceArray(ctx, m.cells)))
    batch = P.Batch(ctx, [x["system"] for x in members])
    args = ([x["t0"] for x in members], [x["f"] for x in members], [x["out"] for x in members])
    kw = dict(conductance=[x["cond"] for x in members], capacity=[x["ca"] for x in members])

    def timed(reps=5):
        infos = batch.heat_step(*args, DT, TOL, 1000, **kw)
        ctx.events(2)
        ctx.synchronize()
        ctx.record(0)
        for _ in range(reps):
            batch.heat_step(*args, DT, TOL, 1000, collect=False, **kw)
        ctx.record(1)
        batch.synchronize()
        return ctx.elapsed_ms(0, 1) / reps, infos

    ms_default, infos = timed()
    if autotune:
        batch.autotune(*args, DT, TOL, 1000, repeats=2, **kw)
        ms_tuned, infos = timed()
    else:
        ms_tuned = ms_default
    nbytes = 0
    for x, info in zip(members, infos):
        lv = hierarchy_levels[x["level"]]
        nbytes += set_matrix_bytes(lv, meshes[x["level"]].edges, precision == P.MIXED) + algorithmic_bytes(
            lv, info.v_cycles, info.restarts, precision == P.MIXED)
    best = min(ms_default, ms_tuned)
    groups = batch.groups
    by_level = {}
    for x, gsize in zip(members, groups):
        by_level.setdefault(str(x["level"]), []).append(gsize)
    out = {
        "config": {"workload": "256 instances, L3/L4/L5/L6 x 64, " +
                               ("mixed precision" if precision == P.MIXED else "plain FP64") +
                               ", set_matrix + 1 solve each, one launch", "tolerance": TOL},
        "v_cycles_per_solve": round(float(np.mean([i.v_cycles for i in infos])), 3),
        "solves_per_s": round(len(members) / best * 1e3, 1),
        "ms_per_launch": round(best, 4),
        "default_allocation_ms": round(ms_default, 4),
        "autotuned_ms": round(ms_tuned, 4),
        "workgroups_per_instance": {k: [min(v), max(v)] for k, v in by_level.items()},
        "converged": all(i.status == 0 and i.relative_residual <= TOL for i in infos),
        "launch_GBps": round(nbytes / (best / 1e3) / 1e9, 2),
    }
    traffic = measured_traffic(BATCH_TRAFFIC_FILE) if precision == P.MIXED else None
    achieved = nbytes / (ms_default / 1e3) / 1e9
    out["roofline"] = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "algorithmic_bytes_per_launch": int(nbytes),
        "mean_launch_ms": round(ms_default, 5), "allocation": "default",
        "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
        "traffic_GBps": round(traffic["hbm_bytes_per_launch"] / (ms_default / 1e3) / 1e9, 1) if traffic else None,
        "traffic_source": os.path.relpath(BATCH_TRAFFIC_FILE, ROOT) if traffic else None,
    }
    return out


def autotune_section(P, ctx, meshes, graphs):
    """The autotuner on lopsided batches (mixed precision, one heat step per
    member): the default allocation against the tuned plan."""
    out = []
    for name, spec in (("8 L6", [6] * 8), ("16 L6 + 240 L3", [6] * 16 + [3] * 240)):
        members = []
        for i, level in enumerate(spec):
            m = meshes[level]
            seed = P.instance_seed(1000 + i)
            cond, ca, _, t0 = P.heat_instance(m, seed, 1.0e6, 0.3)
            members.append(dict(system=P.System(graphs[level], P.MIXED),
                                cond=P.DeviceArray.of(ctx, m.entry_values(cond)), ca=P.DeviceArray.of(ctx, ca),
                                t0=P.DeviceArray.of(ctx, t0), f=P.DeviceArray.of(ctx, P.heat_forcing(m, seed, 0)),
                                out=P.DeviceArray(ctx, m.cells)))
        batch = P.Batch(ctx, [x["system"] for x in members])
        args = ([x["t0"] for x in members], [x["f"] for x in members], [x["out"] for x in members])
        kw = dict(conductance=[x["cond"] for x in members], capacity=[x["ca"] for x in members])

        def timed(reps=5):
            batch.heat_step(*args, DT, TOL, 1000, **kw)
            ctx.events(2)
            ctx.synchronize()
            ctx.record(0)
            for _ in range(reps):
                batch.heat_step(*args, DT, TOL, 1000, collect=False, **kw)
            ctx.record(1)
            batch.synchronize()
            return ctx.elapsed_ms(0, 1) / reps

        default_ms = timed()
        default_groups = list(batch.groups)
        batch.autotune(*args, DT, TOL, 1000, repeats=2, **kw)
        tuned_ms = timed()
        out.append({"batch": name, "default_ms": round(default_ms, 4), "tuned_ms": round(tuned_ms, 4),
                    "gain": round(default_ms / tuned_ms, 3),
                    "default_workgroups_per_instance": [min(default_groups), max(default_groups)],
                    "tuned_workgroups_per_instance": [min(batch.groups), max(batch.groups)],
                    "tuned_workgroups_per_cu": batch.occupancy})
        batch.close()
        for x in members:
            x["system"].close()
    return out


class OutputRing:
    """Device output buffers for served requests: one block per level carved
    into `slots` temperature vectors, handed out in order from the start of
    each run, so no buffer is reused within a run."""

    def __init__(self, P, ctx, meshes, levels, slots=16384):
        self.n = {L: meshes[L].cells for L in levels}
        self.blocks = {L: P.DeviceArray(ctx, slots * self.n[L]) for L in levels}
        self.slots = slots
        self.next = {L: 0 for L in levels}

    def reset(self):
        self.next = {L: 0 for L in self.next}

    def take(self, level):
        k = self.next[level]
        if k >= self.slots:
            raise RuntimeError(f"more than {self.slots} L{level} requests in one served run")
        self.next[level] += 1
        return self.blocks[level].offset(k * self.n[level])


def serve_requests(P, bank, levels, shares, rate, count, rng, ring):
    """Poisson arrivals at `rate` (inf: all at once); level by `shares`; half
    the requests continue a live 12-step session (its resident T), the rest
    start sessions (1 in 11 of them 12 steps long). Each session's matrix is
    one of its level's bank entries, named by its key (the server keeps
    assembled matrices by key)."""
    times = np.cumsum(rng.exponential(1.0 / rate, count)) if np.isfinite(rate) else np.zeros(count)
    times -= times[0]
    ring.reset()
    live, reqs, session = [], [], 0
    for k in range(count):
        if rng.random() < 0.5 and live:
            s = live[int(rng.integers(len(live)))]
        else:
            level = levels[int(rng.choice(len(levels), p=shares))]
            s = [session, level, int(rng.integers(len(bank[level]))), 0, None, 12 if rng.random() < 1 / 11 else 1]
            session += 1
            if s[5] > 1:
                live.append(s)
        sid, level, b, step, prev, total = s
        inst = bank[level][b]
        out = ring.take(level)
        r = P.ServeRequest()
        r.arrival_s = float(times[k])
        r.graph = levels.index(level)
        r.session = sid
        r.first = int(step == 0)
        r.last = int(step == total - 1)
        r.conductance_per_entry = inst["cond"].ptr if step == 0 else None
        r.capacity_area = inst["ca"].ptr if step == 0 else None
        r.temperature_old = inst["t0"].ptr if step == 0 else prev
        r.forcing_W_m2 = inst["f"][step].ptr
        r.temperature_new = out
        r.matrix_key = inst["key"] if step == 0 else 0
        reqs.append(r)
        s[3], s[4] = step + 1, out
        if s[3] >= total and s in live:
            live.remove(s)
    unfinished = {s[0] for s in live}
    return [r for r in reqs if r.session not in unfinished]


SWEEP = ((100, 1000), (200, 2000), (500, 5000), (1000, 10000), (2000, 20000), (5000, 20000))


def served_section(P, ctx, meshes, graphs, plain_value, sweep=SWEEP, profile=False):
    """Served stream: Poisson arrivals of single-step requests, levels
    L3:L4:L5:L6 = 40:30:20:10, half continuing a resident 12-step session,
    mixed precision, one persistent server launch per run; each offered rate
    runs >= 10 s or 20,000 requests. Matrices come from a bank of 8 per level
    (the per-level cached matrix of the workload), kept on the device by key."""
    levels, shares = [3, 4, 5, 6], [0.4, 0.3, 0.2, 0.1]
    bank = {}
    for level in levels:
        m = meshes[level]
        bank[level] = []
        for b in range(8):
            seed = P.instance_seed(50000 + 100 * level + b)
            cond, ca, _, t0 = P.heat_instance(m, seed, 1.0e6, 0.3)
            bank[level].append(dict(key=100 * level + b + 1, cond=P.DeviceArray.of(ctx, m.entry_values(cond)),
                                    ca=P.DeviceArray.of(ctx, ca), t0=P.DeviceArray.of(ctx, t0),
                                    f=[P.DeviceArray.of(ctx, P.heat_forcing(m, seed, s)) for s in range(12)]))
    ring = OutputRing(P, ctx, meshes, levels)

    def run(server, lv, sh, rate, count, seed):
        reqs = serve_requests(P, bank, lv, sh, rate, count, np.random.default_rng(seed), ring)
        rec = server.run(reqs, DT, TOL, 1000)
        lat = np.array([(r.completed_s - r.arrival_s) * 1e3 for r in rec])
        queue = np.array([(r.submitted_s - r.arrival_s) * 1e3 for r in rec])
        device = np.array([r.device_ms for r in rec])
        rest = lat - queue - device
        span = max(r.completed_s for r in rec) - min(r.arrival_s for r in rec)
        return {
            "offered_per_s": rate if np.isfinite(rate) else "all at t=0", "requests": len(rec),
            "seconds": round(span, 2), "solves_per_s": round(len(rec) / span, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "breakdown_p50_ms": {"admission_queue": round(float(np.percentile(queue, 50)), 3),
                                 "device_solve": round(float(np.percentile(device, 50)), 3),
                                 "claim_and_completion": round(float(np.percentile(rest, 50)), 3)},
            "breakdown_p99_ms": {"admission_queue": round(float(np.percentile(queue, 99)), 3),
                                 "device_solve": round(float(np.percentile(device, 99)), 3),
                                 "claim_and_completion": round(float(np.percentile(rest, 99)), 3)},
            "v_cycles_per_solve": round(float(np.mean([r.v_cycles for r in rec])), 3),
            "converged": all(r.status == 0 and r.relative_residual <= TOL for r in rec),
        }

    out = {"config": {"workload": "Poisson single-step requests, L3:L4:L5:L6 = 40:30:20:10, half continuing a "
                                  "resident 12-step session, mixed precision, persistent server, matrices kept "
                                  "by key (8 per level)", "tolerance": TOL}}
    if profile:
        # two server launches of the same all-at-once stream: cold, then with
        # the bank's matrices kept
        server = P.Server(ctx, [graphs[L] for L in levels], share=shares, precision=P.MIXED, memory_budget=8 << 30)
        out["cold"] = run(server, levels, shares, float("inf"), 4000, 17)
        out["warm"] = run(server, levels, shares, float("inf"), 4000, 17)
        server.close()
        return out
    # The slot plan is the deployment's latency/throughput choice: slots of
    # ceil(tiles / 2) workgroups (two 256-row tiles per workgroup per phase)
    # for the offered-rate sweep, whose load needs few slots; the default plan
    # (about 12 tiles per workgroup, many small slots) for saturation.
    latency_sizes = [max(1, -(-((meshes[L].cells + 255) // 256) // 2)) for L in levels]
    server = P.Server(ctx, [graphs[L] for L in levels], share=shares, precision=P.MIXED, memory_budget=8 << 30,
                      slot_sizes=latency_sizes)
    # first an all-at-once stream, which also assembles every bank matrix
    # once (a running server holds them by key), then the offered rates
    out["all_at_once_on_sweep_plan"] = run(server, levels, shares, float("inf"), 10000, 17)
    out["sweep"] = []
    for i, (rate, count) in enumerate(sweep):
        out["sweep"].append(run(server, levels, shares, rate, count, 11 + i))
        progress(f"served sweep at {rate} req/s")
    out["sweep_plan"] = {str(L): [p["slots"], p["slot_size"]] for L, p in zip(levels, server.plan())}
    server.close()
    server = P.Server(ctx, [graphs[L] for L in levels], share=shares, precision=P.MIXED, memory_budget=8 << 30)
    out["all_at_once"] = run(server, levels, shares, float("inf"), 10000, 17)
    plan = server.plan()
    out["config"]["slots"] = {str(L): [p["slots"], p["slot_size"], p["pool"]] for L, p in zip(levels, plan)}
    out["config"]["system_bytes"] = {str(L): p["system_bytes"] for L, p in zip(levels, plan)}
    server.close()
    # L6 only, all requests at t = 0: a cold run (every session assembles its
    # matrix) and the same stream again on the same server (the matrices of
    # the bank stay by key, as a running server keeps them); pool sized for
    # 32 GiB of HBM; slots of 4 workgroups. The server program's occupancy
    # (two or three workgroups per CU) is a deployment choice: both are run
    # and the faster warm one is reported, with the other beside it.
    for name, precision in (("l6_only_saturated", P.MIXED), ("l6_only_saturated_plain_fp64", P.PLAIN_FP64)):
        runs = {}
        for occupancy in (3, 2):
            l6 = P.Server(ctx, [graphs[6]], precision=precision, memory_budget=32 << 30, slot_sizes=[4],
                          occupancy=occupancy)
            cold = run(l6, [6], [1.0], float("inf"), 6000, 99)
            runs[occupancy] = run(l6, [6], [1.0], float("inf"), 6000, 99)
            runs[occupancy]["cold_start_solves_per_s"] = cold["solves_per_s"]
            runs[occupancy]["workgroups_per_cu"] = occupancy
            l6.close()
        best = max(runs, key=lambda o: runs[o]["solves_per_s"])
        out[name] = runs[best]
        out[name]["other_occupancy_solves_per_s"] = runs[5 - best]["solves_per_s"]
    out["l6_served_mixed_vs_plain_fp64_chain"] = round(out["l6_only_saturated"]["solves_per_s"] / plain_value, 2)
    out["l6_served_plain_vs_plain_fp64_chain"] = round(
        out["l6_only_saturated_plain_fp64"]["solves_per_s"] / plain_value, 2)
    return out


def response_table(n, points, seed):
    """A concave increasing piecewise-linear response per node (the
    device-resident response form), a flat (held) segment on every sixth
    node, the floor at the table's first source."""
    rng = np.random.default_rng(seed)
    dh = rng.uniform(20.0, 80.0, (points - 1, n))
    h = np.empty((points, n))
    h[0] = -150.0 + rng.uniform(-20.0, 20.0, n)
    h[1:] = h[0] + np.cumsum(dh, axis=0)
    slopes = -np.sort(-rng.uniform(0.01, 0.3, (points - 1, n)), axis=0)
    slopes[points // 2, ::6] = 0.0
    t = np.empty((points, n))
    t[0] = 250.0 + 30.0 * rng.random(n)
    t[1:] = t[0] + np.cumsum(slopes * dh, axis=0)
    return h, t, h[0].copy()


def host_table_response(h_tab, t_tab):
    """The same table evaluated by a host callback (the caller's response in
    the reference's form), with the device kernel's operations."""
    points, n = h_tab.shape
    cols = np.arange(n)

    def fn(h):
        k = np.zeros(n, np.int64)
        for j in range(1, points - 1):
            k += (h >= h_tab[j]).astype(np.int64)
        h0, h1, t0, t1 = h_tab[k, cols], h_tab[k + 1, cols], t_tab[k, cols], t_tab[k + 1, cols]
        s = (t1 - t0) / (h1 - h0)
        return t0 + s * (h - h0), s

    return fn


def transport_section(P, ctx, meshes, with_oracle, repeats=5):
    """solve_implicit_transport (Newton + line search, PCG with the V-cycle)
    on the L6 mesh's agglomerated graph (10,242 groups, the production
    configuration) and on the L6 mesh itself, with a 6-point table response:
    device-resident (psim_implicit_transport_table, inputs in HBM) and as a
    host callback (psim_implicit_transport); the oracle, on one host thread,
    with the same callback."""
    K = 0.5 * 6371000.0 ** 2
    agg = P.Agglomeration(meshes[6])
    cases = {"l6_groups": (P.Graph(ctx, agg.graph), agg.groups), "l6_cells": (P.Graph(ctx, meshes[6]), meshes[6].cells)}
    out = {"config": {"workload": "implicit transport, 6-point concave table response per node, every sixth node "
                                  "held on one segment, conductance 0.5 R^2, default Newton settings",
                      "repeats": repeats}}
    for name, (graph, n) in cases.items():
        h, t, floor = response_table(n, 6, 4242 + n)
        d_h, d_t, d_f = (P.DeviceArray.of(ctx, np.ascontiguousarray(a.ravel())) for a in (h, t, floor))
        d_s, d_m = P.DeviceArray(ctx, n), P.DeviceArray(ctx, n)
        res = P.implicit_transport_table_device(graph, K, d_f, d_h, d_t, 6, d_s, d_m)
        ctx.synchronize()
        start = time.perf_counter()
        for _ in range(repeats):
            res = P.implicit_transport_table_device(graph, K, d_f, d_h, d_t, 6, d_s, d_m)
        ctx.synchronize()
        device_ms = (time.perf_counter() - start) / repeats * 1e3
        fn = host_table_response(h, t)
        start = time.perf_counter()
        for _ in range(repeats):
            src, _, cres = P.implicit_transport(graph, K, floor, fn)
        callback_ms = (time.perf_counter() - start) / repeats * 1e3
        entry = {"nodes": n, "newton_iterations": res.newton_iterations, "cg_iterations": res.cg_iterations,
                 "consistency_residual_W_m2": res.consistency_residual_W_m2,
                 "device_table_ms": round(device_ms, 3), "host_callback_ms": round(callback_ms, 3),
                 "table_equals_callback": bool(np.array_equal(d_s.download(), src))}
        if with_oracle:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import support as S

            if name == "l6_groups":
                _, om = S.agglomerate(6)
            else:
                om = S.mesh("oracle", 6)
            start = time.perf_counter()
            osrc, _, ores = S.implicit_transport(S.oracle(), om.handle, om.n, K, floor, fn)
            oracle_ms = (time.perf_counter() - start) * 1e3
            entry.update({"oracle_ms_1_thread": round(oracle_ms, 1), "oracle_newton_iterations": ores.newton_iterations,
                          "max_abs_dH_vs_oracle_W_m2": float(np.max(np.abs(src - osrc))),
                          "device_table_speedup_vs_oracle": round(oracle_ms / device_ms, 1)})
        out[name] = entry
    return out


def _entry_sign(mesh):
    """+1 where a CSR entry's cell is its edge's first cell, else -1."""
    rows = np.repeat(np.arange(mesh.cells), np.diff(mesh.offset.astype(np.int64)))
    return np.where(mesh.first[mesh.edge] == rows, 1.0, -1.0), rows


def _tracer_inputs(mesh, layers, seed):
    """Tracer-step inputs: antisymmetric layer mass fluxes per entry (kg/s),
    symmetric eddy conductances, layer masses from the areas, mixing ratios."""
    rng = np.random.default_rng(seed)
    sign, _ = _entry_sign(mesh)
    flux = rng.standard_normal((layers, mesh.edges)) * 2.0e7
    return dict(layers=layers, outflow=np.ascontiguousarray((flux[:, mesh.edge] * sign[None, :]).ravel()),
                conductance=np.ascontiguousarray(rng.uniform(0.5, 1.5, mesh.edges)[mesh.edge] * 3.0e9),
                layer_mass=np.ascontiguousarray(mesh.area * 1.0e4), heat_capacity=1004.64, dt=2.63e6,
                tracer=np.ascontiguousarray(rng.uniform(0.0, 2.0e-3, layers * mesh.cells)))


def _advection_inputs(mesh, layers, seed):
    """Coupled-transport inputs with a linear column response per node."""
    rng = np.random.default_rng(seed)
    sign, _ = _entry_sign(mesh)
    conductance = np.ascontiguousarray(rng.uniform(0.5, 1.5, mesh.edges)[mesh.edge] * mesh.weight * 0.5
                                       * mesh.radius ** 2)
    outflow = np.ascontiguousarray((rng.standard_normal((layers, mesh.edges))[:, mesh.edge] * sign[None, :]
                                    * 3.0e8).ravel())
    cp, w, t0 = 1004.64, rng.uniform(0.2, 1.0, layers), 250.0 + 10.0 * np.arange(layers)
    base, slopes = 250.0 + 40.0 * rng.random(mesh.cells), 0.02 + 0.1 * rng.random(mesh.cells)

    def response(h):
        theta = base + slopes * h
        energy = cp * (t0[:, None] + w[:, None] * (theta[None, :] - 250.0))
        return (theta, slopes.copy(), np.ascontiguousarray(energy.ravel()),
                np.ascontiguousarray((cp * w[:, None] * slopes[None, :]).ravel()))

    return dict(layers=layers, conductance=conductance, outflow=outflow, floor=np.full(mesh.cells, -np.inf),
                response=response)


def _flux_operator(mesh, seed, epsilon=1e-9):
    """Eddy diffusion plus upwind advection in flux form (the balance
    operator's shape) and its symmetric part, on a graph's CSR pattern."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(mesh.cells), np.diff(mesh.offset.astype(np.int64)))
    nb = mesh.graph_neighbour.astype(np.int64)
    key, back = rows * mesh.cells + nb, nb * mesh.cells + rows
    order = np.argsort(key)
    reverse = order[np.searchsorted(key[order], back)]
    g = mesh.weight * rng.uniform(0.5, 1.5, mesh.weight.size)
    g = 0.5 * (g + g[reverse])
    face = rng.standard_normal(mesh.cells)
    flow = 0.8 * (face[nb] - face[rows]) * mesh.weight
    off = -g - np.maximum(-flow, 0.0)
    div = np.bincount(rows, weights=flow, minlength=mesh.cells)
    diag = (np.bincount(rows, weights=g + np.maximum(flow, 0.0), minlength=mesh.cells) + np.maximum(-div, 0.0)
            + epsilon * mesh.area)
    sym_diag = np.bincount(rows, weights=g, minlength=mesh.cells) + epsilon * mesh.area
    return np.ascontiguousarray(diag), np.ascontiguousarray(off), sym_diag, -g


def _timed(fn, repeats):
    out = fn()
    start = time.perf_counter()
    for _ in range(repeats):
        out = fn()
    return out, (time.perf_counter() - start) / repeats * 1e3


def transport_rows_section(P, ctx, meshes, with_oracle, repeats=3):
    """The other consumers of the path, each on the device against the
    oracle on one host thread where the oracle restates it: the implicit
    tracer step (BiCGSTAB + block ILU(0) on the stacked layers), block
    ILU(0) factor and solve, the coupled advection-diffusion Newton solve,
    the pressure-Poisson PCG on the agglomerated groups and the balance
    BiCGSTAB (V-cycle of the symmetric part, then block ILU(0)). Host
    arrays in and out (PCIe included): these entry points take the
    reference's host fields."""
    S = None
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import support as S
    out = {"config": {"workload": "consumers of the path (tracer step, block ILU, advection-diffusion, operator solves) on the L6 mesh (40,962 cells) and its agglomerated groups "
                                  "(10,242); host arrays, timings per call", "repeats": repeats}}
    m6 = meshes[6]
    g6 = P.Graph(ctx, m6)
    blocks = np.array(list(range(0, m6.cells, 256)) + [m6.cells], np.uint64)

    # implicit tracer step, 2 layers (81,924 unknowns)
    case = _tracer_inputs(m6, 2, 16)
    (tr, tres), dev_ms = _timed(lambda: P.tracer_step(g6, case["layers"], case["conductance"], case["outflow"],
                                                      case["layer_mass"], case["tracer"], case["dt"],
                                                      case["heat_capacity"]), repeats)
    entry = {"unknowns": 2 * m6.cells, "iterations": tres.iterations, "relative_residual": tres.relative_residual,
             "device_ms": round(dev_ms, 3)}
    if S is not None:
        om = S.mesh("oracle", 6)
        start = time.perf_counter()
        expect, eres = S.oracle_tracer(om, case)
        oracle_ms = (time.perf_counter() - start) * 1e3
        entry.update({"oracle_ms_1_thread": round(oracle_ms, 1), "oracle_iterations": eres.iterations,
                      "max_rel_diff_vs_oracle": float(np.max(np.abs(tr - expect)) / np.max(np.abs(expect))),
                      "speedup_vs_oracle": round(oracle_ms / dev_ms, 1)})
    out["tracer_step_l6_2_layers"] = entry

    # block ILU(0): factor and one solve of a non-symmetric L6 operator
    rng = np.random.default_rng(6)
    off = -m6.weight * rng.uniform(0.5, 1.5, m6.weight.size)
    _, rows = _entry_sign(m6)
    diag = -np.bincount(rows, weights=off, minlength=m6.cells) + m6.area * 1e-9
    ilu = P.Ilu(ctx, m6.offset, m6.graph_neighbour, blocks)
    _, factor_ms = _timed(lambda: ilu.factor(diag, off), repeats)
    x = rng.standard_normal(m6.cells)
    y, solve_ms = _timed(lambda: ilu.solve(x), repeats)
    entry = {"blocks": int(blocks.size - 1), "factor_ms": round(factor_ms, 3), "solve_ms": round(solve_ms, 3)}
    if S is not None:
        lib = S.oracle()
        start = time.perf_counter()
        h = lib.ilu_create(m6.cells, m6.offset, m6.graph_neighbour, np.ascontiguousarray(diag),
                           np.ascontiguousarray(off), blocks, blocks.size - 1, 1.0)
        mid = time.perf_counter()
        z = x.copy()
        lib.check(lib.ilu_solve(h, z))
        end = time.perf_counter()
        lib.ilu_free(h)
        entry.update({"oracle_factor_ms_1_thread": round((mid - start) * 1e3, 3),
                      "oracle_solve_ms_1_thread": round((end - mid) * 1e3, 3),
                      "solve_bit_identical_to_oracle": bool(np.array_equal(y.view(np.uint64), z.view(np.uint64)))})
    out["block_ilu_l6"] = entry

    # coupled advection-diffusion (Newton with a host response), L5 and L6
    for level in (5, 6):
        m = meshes[level]
        g = g6 if level == 6 else P.Graph(ctx, m)
        case = _advection_inputs(m, 2, 9 + level)
        (sd, md, _, rd), dev_ms = _timed(lambda: P.implicit_advection_diffusion(
            g, case["layers"], case["conductance"], case["outflow"], case["floor"], case["response"]), repeats)
        entry = {"nodes": m.cells, "layers": 2, "newton_iterations": rd.newton_iterations,
                 "linear_iterations": rd.linear_iterations,
                 "consistency_residual_W_m2": rd.consistency_residual_W_m2, "device_ms": round(dev_ms, 3)}
        if S is not None:
            om = S.mesh("oracle", level)
            start = time.perf_counter()
            sr, mr, _, rr = S.solve_advection(S.oracle(), om.handle, om.n, case)
            oracle_ms = (time.perf_counter() - start) * 1e3
            entry.update({"oracle_ms_1_thread": round(oracle_ms, 1), "oracle_newton_iterations": rr.newton_iterations,
                          "max_abs_dH_vs_oracle_W_m2": float(np.max(np.abs(sd - sr))),
                          "max_abs_dT_vs_oracle_K": float(np.max(np.abs(md - mr))),
                          "speedup_vs_oracle": round(oracle_ms / dev_ms, 1)})
        out[f"advection_diffusion_l{level}"] = entry

    # pressure-Poisson PCG on the agglomerated L6 groups (10,242 nodes)
    agg = P.Agglomeration(m6)
    gg, am = P.Graph(ctx, agg.graph), agg.graph
    s = P.System(gg)
    s.set_matrix(am.weight_sum * (1.0 + 1.0e-10), -am.weight)
    rhs = np.random.default_rng(6).standard_normal(am.cells) * am.area
    rhs -= rhs.mean()
    (_, pinfo), pcg_ms = _timed(lambda: P.pcg_operator(s, am.weight_sum, rhs, tolerance=1e-6, max_iterations=500),
                                repeats)
    out["pressure_poisson_pcg_l6_groups"] = {"nodes": am.cells, "iterations": pinfo.iterations,
                                             "relative_residual": pinfo.relative_residual,
                                             "device_ms": round(pcg_ms, 3)}

    # balance BiCGSTAB on L6 cells: V-cycle of the symmetric part, then ILU(0)
    diag, off, sym_diag, sym_off = _flux_operator(m6, 13)
    mg = P.System(g6)
    mg.set_matrix(sym_diag * (1.0 + 1.0e-8), sym_off)
    bilu = P.Ilu(ctx, m6.offset, m6.graph_neighbour, blocks)
    bilu.factor(diag, off, 1.0 + 1.0e-8)
    b = np.random.default_rng(13).standard_normal(m6.cells) * m6.area
    (_, binfo), bal_ms = _timed(lambda: P.operator_solve(g6, diag, off, b, multigrid=mg, ilu=bilu,
                                                         form=P.FORM_BALANCE, tolerance=1e-10), repeats)
    out["balance_bicgstab_l6"] = {"nodes": m6.cells, "iterations": binfo.iterations,
                                  "true_relative_residual": binfo.relative_residual, "status": binfo.status,
                                  "device_ms": round(bal_ms, 3)}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed passes of the 120-step chain")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--workgroups", type=int, default=0, help="override the solver's workgroup count")
    ap.add_argument("--only-consumers", action="store_true",
                    help="run only the consumer measurements (tracer, ILU, advection, operator solves)")
    ap.add_argument("--only", choices=["chain", "fv", "mixed", "batched", "served", "kernels"],
                    help="run only one section (profiling runs): the headline chain's census, warm-up and timed "
                         "passes, the 64-field FV sweeps, the mixed-precision chain beside the plain one (no "
                         "oracle), the 256-instance batch, a short served run, or 'kernels': the chain passes, "
                         "then the batch, then the served run, in one process")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-precision, batched and served measurements")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (instances shard "
              "across GPUs by instance, without communication)", file=sys.stderr)
        sys.exit(2)

    import planet_game_amd as P

    level = 6
    if args.only_consumers:
        ctx = P.Context(0)
        meshes = {L: P.Mesh(L) for L in (5, 6)}
        print(json.dumps({"transport_consumers": transport_rows_section(P, ctx, meshes, not args.no_cpu)}))
        return
    if args.only and args.only not in ("chain", "kernels"):
        ctx = P.Context(0)
        mesh = P.Mesh(level)
        # one 256 MiB streaming copy first: the calibration kernel of the
        # counter runs (FETCH_SIZE / WRITE_SIZE per known byte)
        calibrate = P.DeviceArray(ctx, COPY_BYTES // 8), P.DeviceArray(ctx, COPY_BYTES // 8)
        ctx.copy(calibrate[1].ptr, calibrate[0].ptr, COPY_BYTES)
        ctx.synchronize()
        if args.only == "fv":
            print(json.dumps({"fv_sweeps": fv_section(P, ctx, mesh, False, 1)}))
            return
        if args.only == "mixed":
            levels = [(lv["size"], lv["nnz"]) for lv in P.Hierarchy(mesh).levels]
            print(json.dumps({"mixed_precision": mixed_section(P, ctx, mesh, P.Graph(ctx, mesh), levels, False)}))
            return
        meshes = {L: (mesh if L == level else P.Mesh(L)) for L in (3, 4, 5, 6)}
        graphs = {L: P.Graph(ctx, meshes[L]) for L in meshes}
        hierarchy_levels = {L: [(lv["size"], lv["nnz"]) for lv in P.Hierarchy(meshes[L]).levels] for L in meshes}
        if args.only == "batched":
            print(json.dumps({"batched": batched_section(P, ctx, meshes, graphs, hierarchy_levels, autotune=False)}))
        else:
            print(json.dumps({"served": served_section(P, ctx, meshes, graphs, 1.0, profile=True)}))
        return
    mesh = P.Mesh(level)
    seed = P.instance_seed(0)
    cond, ca, land, t0 = P.heat_instance(mesh, seed, 1.0e6, 0.3)
    forcings = [P.heat_forcing(mesh, seed, s) for s in range(CHAIN)]
    n = mesh.cells

    ctx = P.Context(0)
    graph = P.Graph(ctx, mesh)
    system = P.System(graph, P.PLAIN_FP64)
    if args.workgroups:
        system.workgroups = args.workgroups
    hier = P.Hierarchy(mesh)
    levels = [(lv["size"], lv["nnz"]) for lv in hier.levels]

    d_cond = P.DeviceArray.of(ctx, mesh.entry_values(cond))
    d_ca = P.DeviceArray.of(ctx, ca)
    d_forcing = P.DeviceArray.of(ctx, np.concatenate(forcings))
    d_t0 = P.DeviceArray.of(ctx, t0)
    d_out = P.DeviceArray(ctx, n)
    infos = (P.SolveInfo * CHAIN)()

    def run_pass(collect=True, tol=TOL):
        """One launch: set the matrix, then 120 chained steps T_k -> T_{k+1}."""
        system.heat_chain_device(d_t0.ptr, d_forcing.ptr, CHAIN, d_out.ptr, DT, tol, 1000,
                                 step_infos=infos if collect else None,
                                 conductance_per_entry=d_cond.ptr, capacity_area=d_ca.ptr)
        return [(i.iterations, i.v_cycles, i.restarts, i.relative_residual, i.status, i.device_ms)
                for i in infos] if collect else None

    census = run_pass()
    progress("census pass")
    for step, (_, _, _, res, status, _) in enumerate(census):
        if status != 0 or not res <= TOL:
            raise SystemExit(f"step {step} did not converge: status {status}, residual {res:.3e}")
    gpu_final = d_out.download()

    for _ in range(args.warmup):
        run_pass(collect=False)
    ctx.synchronize()

    ctx.events(2 * args.steps)
    ctx.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ctx.record(2 * k)
        run_pass(collect=False)
        ctx.record(2 * k + 1)
    ctx.synchronize()
    wall = time.perf_counter() - t_start
    progress(f"{args.steps} timed passes")
    launch_ms = np.array([ctx.elapsed_ms(2 * k, 2 * k + 1) for k in range(args.steps)])
    solves = CHAIN * args.steps
    value = solves / wall
    if args.only in ("chain", "kernels"):
        # the counter runs' calibration kernel: one 256 MiB streaming copy
        calibrate = P.DeviceArray(ctx, COPY_BYTES // 8), P.DeviceArray(ctx, COPY_BYTES // 8)
        ctx.copy(calibrate[1].ptr, calibrate[0].ptr, COPY_BYTES)
        ctx.synchronize()
        print(json.dumps({"chain_passes": {"solves_per_s": round(value, 2),
                                           "mean_launch_ms": round(float(launch_ms.mean()), 5)}}), flush=True)
        if args.only == "kernels":
            # the batched and served kernels in the same process (their
            # program instances have names of their own in a trace)
            meshes = {L: (mesh if L == level else P.Mesh(L)) for L in (3, 4, 5, 6)}
            graphs = {L: (graph if L == level else P.Graph(ctx, meshes[L])) for L in meshes}
            hierarchy_levels = {L: [(lv["size"], lv["nnz"]) for lv in P.Hierarchy(meshes[L]).levels]
                                for L in meshes}
            print(json.dumps({"batched": batched_section(P, ctx, meshes, graphs, hierarchy_levels,
                                                         autotune=False)}), flush=True)
            progress("batched")
            print(json.dumps({"served": served_section(P, ctx, meshes, graphs, 1.0, profile=True)}), flush=True)
            progress("served")
        return
    # per-solve latencies: the same passes again, with per-step records
    step_ms = []
    for _ in range(args.steps):
        records = run_pass(collect=True)
        if any(r[4] != 0 or not r[3] <= TOL for r in records):
            raise SystemExit("a timed step did not converge")
        step_ms.extend(r[5] for r in records)
    step_ms = np.array(step_ms)
    # the same chain converged to 1e-13, for the parity check against the oracle
    run_pass(collect=True, tol=1e-13)
    tight_final = d_out.download()

    bytes_per_launch = set_matrix_bytes(levels, mesh.edges) + sum(
        algorithmic_bytes(levels, v, r) for _, v, r, _, _, _ in census)
    achieved = bytes_per_launch / (launch_ms.mean() / 1e3) / 1e9

    # Secondary: one launch per solve (the psim_heat_step entry point), same work.
    system.heat_set_matrix_device(d_cond, d_ca, DT)
    d_t = [P.DeviceArray(ctx, n), P.DeviceArray(ctx, n)]
    ctx.events(CHAIN + 1)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.record(0)
    src = d_t0.ptr
    for step in range(CHAIN):
        dst = d_t[step % 2]
        system.heat_step_device(src, d_forcing.offset(step * n), dst.ptr, DT, TOL, 1000)
        ctx.record(step + 1)
        src = dst.ptr
    ctx.synchronize()
    single_wall = time.perf_counter() - t1
    single_ms = np.array([ctx.elapsed_ms(s, s + 1) for s in range(CHAIN)])
    single_err = float(np.abs(d_t[(CHAIN - 1) % 2].download() - gpu_final).max())
    # every step's T (the same launches again, untimed) for the per-step error
    step_temps = [t0]
    src = d_t0.ptr
    for step in range(CHAIN):
        dst = d_t[step % 2]
        system.heat_step_device(src, d_forcing.offset(step * n), dst.ptr, DT, TOL, 1000)
        step_temps.append(dst.download())
        src = dst.ptr

    # HBM reference: one 256 MiB streaming copy (the counter calibration kernel)
    copy_src = P.DeviceArray(ctx, COPY_BYTES // 8)
    copy_dst = P.DeviceArray(ctx, COPY_BYTES // 8)
    ctx.copy(copy_dst.ptr, copy_src.ptr, COPY_BYTES)   # first touch of both buffers
    ctx.events(2)
    ctx.synchronize()
    ctx.record(0)
    for _ in range(4):
        ctx.copy(copy_dst.ptr, copy_src.ptr, COPY_BYTES)
    ctx.record(1)
    ctx.synchronize()
    copy_gbs = 4 * 2 * COPY_BYTES / (ctx.elapsed_ms(0, 1) / 1e3) / 1e9
    traffic = measured_traffic()
    progress("latency, single-launch and copy runs")

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded heat-step inputs, BASE_SEED 20261203, instance 0)",
        "config": {
            "workload": "L6 plain FP64: set_matrix + 120 chained monthly implicit heat steps per pass, "
                        "one device launch per pass",
            "cells": n, "edges": mesh.edges, "levels": levels, "tolerance": TOL,
            "solves_per_step": CHAIN, "workgroups": system.workgroups,
        },
        "p50_ms": round(float(np.percentile(step_ms, 50)), 5),
        "p99_ms": round(float(np.percentile(step_ms, 99)), 5),
        "latency_clock": "device wall clock per solve inside the chained launch",
        "iterations_per_solve": float(np.mean([c[0] for c in census])),
        "v_cycles_per_solve": float(np.mean([c[1] for c in census])),
        "max_true_residual": float(max(c[3] for c in census)),
        "one_launch_per_solve": {
            "solves_per_s": round(CHAIN / single_wall, 2),
            "p50_ms": round(float(np.percentile(single_ms, 50)), 5),
            "p99_ms": round(float(np.percentile(single_ms, 99)), 5),
            "max_abs_diff_vs_chain_K": single_err,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
            "kernel": "program_kernel (one launch per pass: set_matrix + 120 solves)",
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
            "mean_launch_ms": round(float(launch_ms.mean()), 5),
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic else None,
            "copy_kernel_GBps": round(copy_gbs, 1),
        },
    }
    if not args.no_cpu:
        # the reference's worker pool gives bit-identical results for any
        # count: time one worker and the cores this process may use (capped
        # at the 16-thread host share of one GPU), report the faster
        host = host_cpu()
        usable = max(1, min(16, host["affinity_cpus"], host["nproc"] or host["affinity_cpus"]))
        runs = {}
        for workers in sorted({1, usable}):
            runs[workers] = cpu_baseline(mesh, cond, ca, t0, forcings, args.cpu_seconds / 2, workers)
        best = max(runs, key=lambda w: runs[w][0])
        rate, cpu_solves, chains, elapsed, cpu_final, kind = runs[best]
        what = ("the compiled reference's BiCGSTAB + GraphMultigrid (Gauss-Seidel V-cycle, built from the "
                "reference sources)" if kind == "reference" else
                "the oracle's BiCGSTAB + Gauss-Seidel V-cycle (restatement of the reference)")
        line["cpu_baseline"] = {
            "value": round(rate, 2), "unit": "solves/s", "cores": best, "kind": kind,
            "sample": f"{cpu_solves} solves of the 120-step L6 chain ({elapsed:.1f} s, matrix set per chain) with "
                      f"the reference worker pool at {best} worker thread(s), the faster of "
                      f"{sorted(runs)}: {what}; host {host['model']}, {host['affinity_cpus']} CPUs in this "
                      f"process's affinity mask",
            "by_workers": {str(w): {"solves_per_s": round(r[0], 2), "solves": r[1], "seconds": round(r[3], 1)}
                           for w, r in runs.items()},
            "host": host,
        }
        err = float(np.linalg.norm(gpu_final - cpu_final) / np.linalg.norm(cpu_final))
        # Both sides stop near the 1e-10 residual with different iterates,
        # and the near-identity chain carries each step's difference forward.
        line["drift_vs_oracle_after_120_steps_tol_1e-10"] = err
        tight_cpu = oracle_chain(mesh, cond, ca, t0, forcings, 1e-13)
        line["error_vs_oracle_after_120_steps_tol_1e-13"] = float(
            np.linalg.norm(tight_final - tight_cpu) / np.linalg.norm(tight_cpu))
        line["error_vs_fp64_oracle_per_step"] = step_errors_vs_oracle(mesh, cond, ca, step_temps, forcings)
        progress("CPU baseline and oracle checks")
        line["fv_sweeps"] = fv_section(P, ctx, mesh, True, best)
        progress("FV sweeps")
        line["l4_chain"] = l4_chain_section(P, ctx)
        progress("L4 chain")
    else:
        line["fv_sweeps"] = fv_section(P, ctx, mesh, False, 1)
    if not args.no_extras:
        meshes = {L: (mesh if L == level else P.Mesh(L)) for L in (3, 4, 5, 6)}
        graphs = {L: (graph if L == level else P.Graph(ctx, meshes[L])) for L in meshes}
        hierarchy_levels = {L: [(lv["size"], lv["nnz"]) for lv in P.Hierarchy(meshes[L]).levels] for L in meshes}
        line["mixed_precision"] = mixed_section(P, ctx, mesh, graph, levels, not args.no_cpu)
        progress("mixed precision")
        line["batched"] = batched_section(P, ctx, meshes, graphs, hierarchy_levels)
        line["batched_plain_fp64"] = batched_section(P, ctx, meshes, graphs, hierarchy_levels, P.PLAIN_FP64)
        line["batched"]["autotuner_lopsided"] = autotune_section(P, ctx, meshes, graphs)
        progress("batched")
        line["served"] = served_section(P, ctx, meshes, graphs, value)
        progress("served")
        line["implicit_transport"] = transport_section(P, ctx, meshes, not args.no_cpu)
        progress("implicit transport")
        line["transport_consumers"] = transport_rows_section(P, ctx, meshes, not args.no_cpu)
        progress("transport consumers")
    print(json.dumps(line))


if __name__ == "__main__":
    main()

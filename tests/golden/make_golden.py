#!/usr/bin/env python3
"""Writes tests/golden/reference_fixtures.json from the compiled reference.

Run in a container where /root/reference exists and oracle/_ref/libref.so
is built (``make -C oracle ref``). Every value comes from the reference's own
code (mesh, transport graph, hierarchy, FV operators, diffusion source,
GraphMultigrid::precondition, the BiCGSTAB template and the public coupled
solve), driven through oracle/ref_shim.cpp. Arrays are stored as SHA-256
digests of their little-endian bytes, plus a few scalars for diagnostics.

The heat-step inputs come from this project's generator (the reference has
none); their digests are recorded from the oracle so drift across hosts is
caught.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import support as S  # noqa: E402

LEVELS = range(0, 7)
OP_LEVELS = [3, 4, 5, 6]
SOLVE_LEVELS = [3, 4, 5]
STIFF_K = 0.5 * S.RADIUS ** 2


def digest(a):
    return S.sha(np.ascontiguousarray(a))


def seeded_cells(lib, n, seed=7, sample=0):
    return np.array([lib.keyed_random_unit_double(seed, 1, 0, c, sample) - 0.5 for c in range(n)])


def seeded_edges(lib, e, seed=11, tick=1):
    return np.array([lib.keyed_random_unit_double(seed, 1, tick, k, 0) - 0.5 for k in range(e)])


def main():
    if not S.have_ref():
        raise SystemExit("oracle/_ref/libref.so is not built")
    ref = S.ref()
    out = {"about": __doc__.strip().splitlines()[0], "meshes": {}, "hierarchies": {}, "operators": {},
           "workload": {}, "multigrid": {}, "solves": {}}
    # Eight golden vectors of the reference's RNG test, as data
    # (tests/unit/test_counter_rng.cpp:24-51): seed, stream, tick, cell,
    # sample, mixed seed, u64, unit-double bits.
    out["rng_golden_vectors"] = [
        ["0x0000000000000000", 1, 0, 0x00000000, 0x00000000, "0xE220A8397B1DCDAF", "0xF9AC6E337C2F0F1E", "0x3FEF358DC66F85E1"],
        ["0xFFFFFFFFFFFFFFFF", 0x10001, 1, 0x00000001, 0x00000001, "0xE4D971771B652C20", "0x6855357F5CA0ACE7", "0x3FDA154D5FD7282A"],
        ["0x0000000000000001", 0x20001, -1, 0xFFFFFFFF, 0xFFFFFFFF, "0x910A2DEC89025CC1", "0x997F76C36C15318E", "0x3FE32FEED86D82A6"],
        ["0x0123456789ABCDEF", 1, 9223372036854775807, 0x80000000, 0x7FFFFFFF, "0x157A3807A48FAA9D", "0x9308C34A9CA34D9A", "0x3FE2611869539469"],
        ["0xFEDCBA9876543210", 0x10001, -9223372036854775808, 0x0000002A, 0x00000000, "0x7AE893B5E32FEE86", "0xBD4FDBA488769EA3", "0x3FE7A9FB74910ED3"],
        ["0x9B97F4A7C1500011", 0x20001, 123456, 0x0000007B, 0x00000008, "0x52E385A41D40AD4A", "0x04EC07CDD7AA3A1A", "0x3F93B01F375EA8E0"],
        ["0xAAAAAAAAAAAAAAAA", 1, -123456, 0xFFFFFFFF, 0x00000011, "0x2B02C6E039F1B1BF", "0xF7C531A2A41F0BAA", "0x3FEEF8A6345483E1"],
        ["0x5555555555555555", 0x10001, 987654321, 0x00000007, 0xFFFFFFFF, "0x3D23DCE41C588F8C", "0x8C3E575A07DC5172", "0x3FE187CAEB40FB8A"],
    ]
    for level in LEVELS:
        m = S.mesh("ref", level)
        entry = {"cells": m.n, "edges": m.n_edges, "directed": m.n_directed,
                 "center_crc32c": S.crc32c(m.center.astype("<f8").tobytes())}
        entry.update({k: digest(v) for k, v in m.arrays().items()})
        out["meshes"][str(level)] = entry
        print("mesh", level, flush=True)
    for level in OP_LEVELS:
        m = S.mesh("ref", level)
        levels = S.hierarchy(ref, m)
        out["hierarchies"][str(level)] = [
            {"size": lv["size"], "nnz": lv["nnz"],
             **{k: digest(lv[k]) for k in ("offset", "column", "aggregate", "target", "diagonal", "value")}}
            for lv in levels]
        phi = seeded_cells(ref, m.n)
        flux = seeded_edges(ref, m.n_edges)
        ops = {}
        for dtype, suffix in ((np.float64, "f64"), (np.float32, "f32")):
            p = phi.astype(dtype)
            f = flux.astype(dtype)
            e, no, d, lap = (np.zeros(m.n, dtype) for _ in range(4))
            getattr(ref, "gradient_" + suffix)(m.handle, p, e, no, 1)
            getattr(ref, "divergence_" + suffix)(m.handle, f, d, 1)
            getattr(ref, "laplacian_" + suffix)(m.handle, p, lap, 1)
            ops[suffix] = {"east": digest(e), "north": digest(no), "divergence": digest(d),
                           "laplacian": digest(lap), "east_0": float(e[0]), "laplacian_0": float(lap[0])}
        t = 288.0 + 30.0 * phi
        ds = np.zeros(m.n)
        ref.diffusion_source(m.handle, 3.5e12, t, ds, 1)
        ops["diffusion_source"] = digest(ds)
        out["operators"][str(level)] = ops
        print("operators", level, flush=True)
    # Workload digests (project generator, oracle) and reference multigrid
    # and solver outputs on them.
    for level in OP_LEVELS:
        om = S.mesh("oracle", level)
        rm = S.mesh("ref", level)
        seed = S.instance_seed(0)
        cond, ca, land, temp = S.heat_instance(om, seed)
        forcing = S.heat_forcing(om, seed, 0)
        out["workload"][str(level)] = {
            "seed": seed, "conductance": digest(cond), "capacity_area": digest(ca), "land": digest(land),
            "temperature": digest(temp), "forcing_0": digest(forcing),
            "land_area_fraction": float((land * om.area).sum() / om.area.sum())}
        if level not in SOLVE_LEVELS:
            continue
        cases = {}
        for name, k_base in (("literal", 1.0e6), ("stiff", STIFF_K)):
            cond, ca, land, temp = S.heat_instance(om, seed, k_base, 0.3)
            diag, off = S.heat_matrix(om, cond, ca)
            b = S.heat_rhs(om, ca, temp, forcing)
            mg = S.Multigrid(ref, rm)
            mg.set_matrix(diag, off)
            pre = mg.precondition(b, 1)
            x, it = mg.bicgstab(diag, off, b, 1e-10, 1000, 1)
            x_tight, it_tight = mg.bicgstab(diag, off, b, 1e-13, 1000, 1)
            res = S.oracle().lib.orc_relative_residual(om.handle, diag, off, b, x)
            cases[name] = {"diagonal": digest(diag), "off": digest(off), "rhs": digest(b),
                           "precondition": digest(pre), "x": digest(x), "iterations": it,
                           "relative_residual": res, "x_tight": digest(x_tight),
                           "iterations_tight": it_tight, "x_norm": float(np.linalg.norm(x))}
            # The reference's public coupled solve with the linear response
            # theta(h) = T_old + dt/C (S + h): an independent route to the
            # same temperatures.
            if level in (3, 4):
                graph_area = om.area
                conductance = np.zeros(om.n_directed)
                for c in range(om.n):
                    for k in range(int(om.offset[c]), int(om.offset[c + 1])):
                        conductance[k] = cond[om.edge[k]] * om.weight[k]
                slope = S.DT * graph_area / ca
                t0 = temp + slope * forcing
                theta = np.zeros(om.n)
                newton, linear = C.c_int(), C.c_int()
                ref.check(ref.lib.ref_coupled_linear_solve(rm.handle, conductance, t0, slope, 1e-12, 1e-13,
                                                            theta, C.byref(newton), C.byref(linear)))
                rel = float(np.linalg.norm(theta - x_tight) / np.linalg.norm(x_tight))
                cases[name]["coupled_route_relative_l2"] = rel
                cases[name]["coupled_route_newton"] = newton.value
                cases[name]["coupled_route_theta"] = digest(theta)
        out["solves"][str(level)] = cases
        print("solves", level, flush=True)
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

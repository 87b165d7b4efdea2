"""The CPU oracle pinned against the reference: its golden vectors and
published values, fixtures made from the compiled reference sources
(tests/golden/make_golden.py), closed forms and a dense direct solve."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import support as S

FIXTURES = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))


def test_rng_golden_vectors():
    """tests/unit/test_counter_rng.cpp:24-51."""
    o = S.oracle()
    for seed, stream, tick, cell, sample, mixed, u64, bits in FIXTURES["rng_golden_vectors"]:
        seed = int(seed, 16)
        assert o.mix_random_key(seed) == int(mixed, 16)
        assert o.keyed_random_u64(seed, stream, tick, cell, sample) == int(u64, 16)
        unit = o.keyed_random_unit_double(seed, stream, tick, cell, sample)
        assert np.float64(unit).view(np.uint64) == int(bits, 16)


def test_rng_planted_mistake_is_caught():
    """A dropped mixing round would not reproduce the golden vectors."""
    seed = 0x0123456789ABCDEF
    o = S.oracle()
    wrong = o.mix_random_key(o.mix_random_key(seed) ^ 1)
    assert wrong != int(FIXTURES["rng_golden_vectors"][3][6], 16)


@pytest.mark.parametrize("level", range(0, 7))
def test_mesh_counts_pentagons_and_area_closure(level):
    """icosphere.cpp:168-198 counts, twelve pentagons, and the published
    area closure at Earth radius (README.md:93-101): Kahan sum of the cell
    areas against 4 pi R^2."""
    m = S.mesh("oracle", level)
    assert m.n == 10 * 4 ** level + 2
    assert m.n_edges == 30 * 4 ** level
    assert m.n_directed == 60 * 4 ** level
    assert int((m.edge_count == 5).sum()) == 12 and int((m.edge_count == 6).sum()) == m.n - 12
    total, carry = 0.0, 0.0
    for a in m.area.tolist():
        piece = a - carry
        nxt = total + piece
        carry = (nxt - total) - piece
        total = nxt
    expected = 4.0 * math.pi * S.RADIUS * S.RADIUS
    closure = abs(total - expected) / expected
    assert closure == (0.0 if level == 1 else 1.2253352946929809e-16)


def test_l0_geometry_checksum():
    """CRC32C of the L0 cell centres, recorded in every golden snapshot
    manifest (tests/regression/test_golden_snapshot.cpp:182-188)."""
    assert S.crc32c(b"123456789") == 0xE3069283   # CRC-32C check value
    m = S.mesh("oracle", 0)
    assert S.crc32c(m.center.astype("<f8").tobytes()) == 1362977666


@pytest.mark.parametrize("level", range(0, 7))
def test_mesh_matches_reference_fixtures(level):
    m = S.mesh("oracle", level)
    f = FIXTURES["meshes"][str(level)]
    assert (m.n, m.n_edges, m.n_directed) == (f["cells"], f["edges"], f["directed"])
    for name, array in m.arrays().items():
        assert S.sha(array) == f[name], name


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_hierarchy_matches_reference_fixtures(level):
    levels = S.hierarchy(S.oracle(), S.mesh("oracle", level))
    ref = FIXTURES["hierarchies"][str(level)]
    assert [(lv["size"], lv["nnz"]) for lv in levels] == [(r["size"], r["nnz"]) for r in ref]
    for lv, r in zip(levels, ref):
        for key in ("offset", "column", "aggregate", "target", "diagonal", "value"):
            assert S.sha(lv[key]) == r[key], key


def test_hierarchy_invariants():
    """Every row aggregated, coarse levels shrink to <= 400 rows, and
    unit-conductance Galerkin rows sum to zero (a graph Laplacian)."""
    levels = S.hierarchy(S.oracle(), S.mesh("oracle", 6))
    for j, lv in enumerate(levels[:-1]):
        nxt = levels[j + 1]["size"]
        assert lv["aggregate"].max() == nxt - 1 and nxt < lv["size"]
    assert levels[-1]["size"] <= 400
    for lv in levels:
        rows = np.repeat(np.arange(lv["size"]), np.diff(lv["offset"].astype(np.int64)))
        sums = lv["diagonal"].copy()
        np.add.at(sums, rows, lv["value"])
        assert np.abs(sums).max() <= 1e-12 * np.abs(lv["diagonal"]).max()


def seeded_cells(n):
    o = S.oracle()
    return np.array([o.keyed_random_unit_double(7, 1, 0, c, 0) - 0.5 for c in range(n)])


def seeded_edges(e):
    o = S.oracle()
    return np.array([o.keyed_random_unit_double(11, 1, 1, k, 0) - 0.5 for k in range(e)])


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_operators_match_reference_fixtures(level):
    m = S.mesh("oracle", level)
    o = S.oracle()
    f = FIXTURES["operators"][str(level)]
    phi, flux = seeded_cells(m.n), seeded_edges(m.n_edges)
    for dtype, suffix in ((np.float64, "f64"), (np.float32, "f32")):
        p, fl = phi.astype(dtype), flux.astype(dtype)
        e, no, d, lap = (np.zeros(m.n, dtype) for _ in range(4))
        getattr(o, "gradient_" + suffix)(m.handle, p, e, no, 3)
        getattr(o, "divergence_" + suffix)(m.handle, fl, d, 2)
        getattr(o, "laplacian_" + suffix)(m.handle, p, lap, 1)
        assert S.sha(e) == f[suffix]["east"] and S.sha(no) == f[suffix]["north"]
        assert S.sha(d) == f[suffix]["divergence"] and S.sha(lap) == f[suffix]["laplacian"]
    ds = np.zeros(m.n)
    o.diffusion_source(m.handle, 3.5e12, 288.0 + 30.0 * phi, ds, 4)
    assert S.sha(ds) == f["diffusion_source"]


def test_operator_closed_forms():
    """Constant fields: gradient and Laplacian exactly 0; zero flux: 0
    divergence (test_operator_conservation.cpp:36-57); the Laplacian is
    conservative, sum A lap = 0 to 1e-13 of sum A |lap|, and dissipative."""
    m = S.mesh("oracle", 5)
    o = S.oracle()
    e, no, lap, d = (np.ones(m.n) for _ in range(4))
    o.gradient_f64(m.handle, np.full(m.n, 287.5), e, no, 1)
    o.laplacian_f64(m.handle, np.full(m.n, 287.5), lap, 1)
    o.divergence_f64(m.handle, np.zeros(m.n_edges), d, 1)
    assert not e.any() and not no.any() and not lap.any() and not d.any()
    phi = seeded_cells(m.n)
    o.laplacian_f64(m.handle, phi, lap, 1)
    assert abs((m.area * lap).sum()) <= 1e-13 * (m.area * np.abs(lap)).sum()
    assert (m.area * phi * lap).sum() <= 0.0


def test_operator_accuracy_on_harmonics():
    """Analytic fields of operator_validation.cpp:52-101 (unrotated frame):
    g = xyz + (x^2 - y^2)/2 has surface gradient from its ambient gradient
    and Laplacian -(12 xyz + 3 (x^2 - y^2)) / R^2. Bounds of
    test_operator_accuracy.cpp:53,71-75 at L5: gradient relative L2 <= 2e-3,
    Laplacian relative max <= 2.5e-2."""
    m = S.mesh("oracle", 5)
    o = S.oracle()
    p = m.center.reshape(-1, 3)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    g = x * y * z + 0.5 * (x * x - y * y)
    amb = np.stack([y * z + x, x * z - y, x * y], axis=1)
    surf = (amb - p * (amb * p).sum(1, keepdims=True)) / S.RADIUS
    east, north = m.east.reshape(-1, 3), m.north.reshape(-1, 3)
    ge, gn = (surf * east).sum(1), (surf * north).sum(1)
    e, no = np.zeros(m.n), np.zeros(m.n)
    o.gradient_f64(m.handle, g, e, no, 1)
    err = np.sqrt((m.area * ((e - ge) ** 2 + (no - gn) ** 2)).sum() / (m.area * (ge ** 2 + gn ** 2)).sum())
    assert err <= 2e-3, err
    lap_exact = -(12 * x * y * z + 3 * (x * x - y * y)) / S.RADIUS ** 2
    lap = np.zeros(m.n)
    o.laplacian_f64(m.handle, g, lap, 1)
    assert np.abs(lap - lap_exact).max() / np.abs(lap_exact).max() <= 2.5e-2


@pytest.mark.parametrize("level", [3, 4, 5])
@pytest.mark.parametrize("case", ["literal", "stiff"])
def test_multigrid_and_bicgstab_match_reference_fixtures(level, case):
    """GraphMultigrid::precondition bits and the BiCGSTAB template's iterate
    bits and iteration counts, on the heat-step matrix (worker count 3 on the
    oracle vs 1 in the fixtures: the block structure makes them identical)."""
    m = S.mesh("oracle", level)
    f = FIXTURES["solves"][str(level)][case]
    seed = S.instance_seed(0)
    k_base = 1.0e6 if case == "literal" else 0.5 * S.RADIUS ** 2
    cond, ca, land, temp = S.heat_instance(m, seed, k_base, 0.3)
    forcing = S.heat_forcing(m, seed, 0)
    diag, off = S.heat_matrix(m, cond, ca)
    b = S.heat_rhs(m, ca, temp, forcing)
    assert S.sha(diag) == f["diagonal"] and S.sha(off) == f["off"] and S.sha(b) == f["rhs"]
    mg = S.Multigrid(S.oracle(), m)
    mg.set_matrix(diag, off)
    assert S.sha(mg.precondition(b, 3)) == f["precondition"]
    x, it = mg.bicgstab(diag, off, b, 1e-10, 1000, 3)
    assert S.sha(x) == f["x"] and it == f["iterations"]
    x, it = mg.bicgstab(diag, off, b, 1e-13, 1000, 2)
    assert S.sha(x) == f["x_tight"] and it == f["iterations_tight"]


def test_reference_coupled_route_agrees():
    """The reference's public coupled solve (heat_transport.hpp:246-249) with
    a linear column response reaches the same temperatures as the BiCGSTAB
    solve: recorded relative L2 differences at L3 and L4."""
    for level in ("3", "4"):
        for case in ("literal", "stiff"):
            assert FIXTURES["solves"][level][case]["coupled_route_relative_l2"] <= 1e-12


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("k_base", [1.0e6, 0.5 * S.RADIUS ** 2])
def test_bicgstab_against_dense_direct_solve(level, k_base):
    """Brute force: numpy's dense LU solve of the assembled matrix."""
    m = S.mesh("oracle", level)
    seed = S.instance_seed(1)
    cond, ca, land, temp = S.heat_instance(m, seed, k_base, 0.3)
    forcing = S.heat_forcing(m, seed, 0)
    diag, off = S.heat_matrix(m, cond, ca)
    b = S.heat_rhs(m, ca, temp, forcing)
    A = S.dense_matrix(m, diag, off)
    x_direct = np.linalg.solve(A, b)
    mg = S.Multigrid(S.oracle(), m)
    mg.set_matrix(diag, off)
    x, it = mg.bicgstab(diag, off, b, 1e-14, 1000, 1)
    assert np.linalg.norm(x - x_direct) / np.linalg.norm(x_direct) <= 1e-11


def test_workload_digests():
    """The seeded heat-step inputs are reproducible across hosts; land is
    30 % of the area."""
    for level, f in FIXTURES["workload"].items():
        m = S.mesh("oracle", int(level))
        cond, ca, land, temp = S.heat_instance(m, f["seed"])
        assert S.sha(cond) == f["conductance"] and S.sha(ca) == f["capacity_area"]
        assert S.sha(land) == f["land"] and S.sha(temp) == f["temperature"]
        assert S.sha(S.heat_forcing(m, f["seed"], 0)) == f["forcing_0"]
        assert abs(f["land_area_fraction"] - 0.3) < 0.01


@pytest.mark.skipif(not S.have_ref(), reason="compiled reference not built here")
def test_oracle_equals_compiled_reference_directly():
    """Where oracle/_ref is built: same bits on fresh random inputs."""
    rng = np.random.default_rng(5)
    for level in (2, 4):
        a, b = S.mesh("oracle", level), S.mesh("ref", level)
        phi = rng.random(a.n) - 0.5
        out = []
        for lib, m in ((S.oracle(), a), (S.ref(), b)):
            e, no, lap = np.zeros(a.n), np.zeros(a.n), np.zeros(a.n)
            lib.gradient_f64(m.handle, phi, e, no, 1)
            lib.laplacian_f64(m.handle, phi, lap, 1)
            out.append((e, no, lap))
        for u, v in zip(*out):
            assert np.array_equal(u.view(np.uint64), v.view(np.uint64))


# ---------------------------------------------------------------- Newton transport
# solve_newton_system (heat_transport.cpp:328-429) and solve_implicit_transport
# (:595-772): the oracle's restatement against the compiled reference, bit for
# bit, plus the properties the reference's own tests assert
# (test_heat_transport.cpp:73-107).

NEWTON_K = 0.5 * S.RADIUS ** 2


def newton_case(level, seed=3):
    om = S.mesh("oracle", level)
    rng = np.random.default_rng(seed + level)
    slope = np.where(np.arange(om.n) % 5 == 0, 0.0, 0.01 + 0.2 * rng.random(om.n))
    rhs = rng.standard_normal(om.n) * om.g_area
    return om, slope, rhs


def newton_solve(lib, mesh, slope, rhs, tol, max_it=2000):
    x = np.zeros(mesh.n)
    it = C.c_int()
    lib.check(lib.newton_system(mesh.handle, NEWTON_K, slope, rhs, tol, max_it, 1, x, C.byref(it)))
    return x, it.value


@pytest.mark.skipif(not S.have_ref(), reason="reference shim not built")
@pytest.mark.parametrize("level", [3, 4])
def test_newton_system_matches_reference(level):
    om, slope, rhs = newton_case(level)
    rm = S.mesh("ref", level)
    for tol in (1e-6, 1e-12):
        xo, io = newton_solve(S.oracle(), om, slope, rhs, tol)
        xr, ir = newton_solve(S.ref(), rm, slope, rhs, tol)
        assert io == ir and np.array_equal(xo.view(np.uint64), xr.view(np.uint64)), (tol, io, ir)


@pytest.mark.parametrize("level", [3, 4])
def test_newton_system_solves_the_held_system(level):
    """The PCG answer solves the Newton matrix of heat_transport.cpp:350-363:
    held rows stay 0 and the true residual is at the CG tolerance."""
    om, slope, rhs = newton_case(level)
    x, it = newton_solve(S.oracle(), om, slope, rhs, 1e-12)
    free = slope > 0.0
    assert np.all(x[~free] == 0.0)
    diag = np.where(free, om.g_area / np.where(free, slope, 1.0) + NEWTON_K * om.weight_sum, 1.0)
    nb = om.g_neighbour.astype(np.int64)
    rows = np.repeat(np.arange(om.n), np.diff(om.offset.astype(np.int64)))
    off = np.where(free[rows] & free[nb], -NEWTON_K * om.weight, 0.0)
    ax = diag * x + np.bincount(rows, weights=off * x[nb], minlength=om.n)
    b = np.where(free, rhs, 0.0)
    assert np.linalg.norm(b - ax) <= 1e-10 * np.linalg.norm(b)
    assert 1 <= it < 100


def linear_transport_case(level, seed=9):
    om = S.mesh("oracle", level)
    rng = np.random.default_rng(seed + level)
    base = 250.0 + 50.0 * rng.random(om.n)
    slopes = np.where(np.arange(om.n) % 5 == 0, 0.0, 0.01 + 0.2 * rng.random(om.n))

    def fn(h):
        return base + slopes * h, slopes.copy()

    return om, fn, np.full(om.n, -np.inf)


def same_transport(a, b):
    (sa, ma, ra), (sb, mb, rb) = a, b
    return (np.array_equal(sa.view(np.uint64), sb.view(np.uint64)) and
            np.array_equal(ma.view(np.uint64), mb.view(np.uint64)) and
            all(getattr(ra, f) == getattr(rb, f) for f, _ in S.TransportResult._fields_))


@pytest.mark.skipif(not S.have_ref(), reason="reference shim not built")
@pytest.mark.parametrize("level", [3, 4])
def test_implicit_transport_linear_matches_reference(level):
    om, fn, floor = linear_transport_case(level)
    rm = S.mesh("ref", level)
    got = S.implicit_transport(S.oracle(), om.handle, om.n, NEWTON_K, floor, fn)
    ref = S.implicit_transport(S.ref(), rm.handle, rm.n, NEWTON_K, floor, fn)
    assert same_transport(got, ref)
    src, mean, res = got
    # test_heat_transport.cpp:99-106
    assert res.newton_iterations <= 3
    assert res.consistency_residual <= 1e-6
    assert abs(res.sum_W) <= 1e-13 * res.absolute_sum_W
    assert res.dissipation_W_K <= 0.0


@pytest.mark.skipif(not S.have_ref(), reason="reference shim not built")
@pytest.mark.parametrize("a", [0.04 + 0.37, 1.0e5 / 600.0])
def test_implicit_transport_quartic_matches_reference(a):
    level = 3
    om = S.mesh("oracle", level)
    rm = S.mesh("ref", level)
    u = np.random.default_rng(21).random(om.n)
    fn, floor = S.quartic_response(om.n, u, a)
    latent = 5.0 * np.sin(np.arange(om.n))
    for lat in (None, latent):
        got = S.implicit_transport(S.oracle(), om.handle, om.n, NEWTON_K, floor, fn, latent=lat)
        ref = S.implicit_transport(S.ref(), rm.handle, rm.n, NEWTON_K, floor, fn, latent=lat)
        assert same_transport(got, ref)
        src, mean, res = got
        assert np.all(mean[::7] <= 273.15)
        assert res.consistency_residual <= 1e-6
        # the diffusion conserves energy; the returned source is diffusion - latent
        expected = 0.0 if lat is None else -float(np.sum(om.g_area * lat))
        assert abs(res.sum_W - expected) <= 1e-12 * res.absolute_sum_W


def test_implicit_transport_rejects_bad_conductance():
    om, fn, floor = linear_transport_case(3)
    with pytest.raises(RuntimeError, match="transport conductance must be finite and positive"):
        S.implicit_transport(S.oracle(), om.handle, om.n, 0.0, floor, fn)

"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer/index work and the FV operators, diffusion source, SpMV and Galerkin
values are compared bit for bit; solves are compared on the converged
solution (relative L2) with the tolerance written in each test. Every case
prints its name and sizes before it touches the GPU.
"""
import sys

import numpy as np
import pytest

import planet_game_amd as P
import support as S

pytestmark = pytest.mark.gpu

LEVELS = [3, 4, 5, 6]
_meshes = {}


def announce(*what):
    print("[gpu case]", *what, flush=True)
    sys.stdout.flush()


def product_mesh(level):
    if level not in _meshes:
        _meshes[level] = P.Mesh(level)
    return _meshes[level]


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def seeded_cells(level, fields, dtype, seed=7):
    """u - 0.5 from the reference RNG, as test_operator_conservation.cpp:14-24."""
    o = S.oracle()
    n = product_mesh(level).cells
    out = np.empty(fields * n)
    for f in range(fields):
        out[f * n:(f + 1) * n] = [o.keyed_random_unit_double(seed, 1, 0, c, f) - 0.5 for c in range(n)]
    return out.astype(dtype)


def seeded_edges(level, fields, dtype, seed=11):
    o = S.oracle()
    e = product_mesh(level).edges
    out = np.empty(fields * e)
    for f in range(fields):
        out[f * e:(f + 1) * e] = [o.keyed_random_unit_double(seed, 1, 1 + f, k, 0) - 0.5 for k in range(e)]
    return out.astype(dtype)


@pytest.mark.parametrize("level", LEVELS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fv_operators_bit_exact(ctx, level, dtype):
    fields = 3
    announce("fv operators", level, np.dtype(dtype).name, fields)
    m = product_mesh(level)
    om = S.mesh("oracle", level)
    fv = P.FV(ctx, m)
    phi = seeded_cells(level, fields, dtype)
    flux = seeded_edges(level, fields, dtype)
    east, north = fv.gradient(phi)
    div = fv.divergence(flux)
    lap = fv.laplacian(phi)
    o = S.oracle()
    n, e = m.cells, m.edges
    suffix = "f64" if dtype == np.float64 else "f32"
    for f in range(fields):
        re, rn, rd, rl = (np.zeros(n, dtype) for _ in range(4))
        getattr(o, "gradient_" + suffix)(om.handle, phi[f * n:(f + 1) * n], re, rn, 1)
        getattr(o, "divergence_" + suffix)(om.handle, flux[f * e:(f + 1) * e], rd, 1)
        getattr(o, "laplacian_" + suffix)(om.handle, phi[f * n:(f + 1) * n], rl, 1)
        assert bits_equal(east[f * n:(f + 1) * n], re)
        assert bits_equal(north[f * n:(f + 1) * n], rn)
        assert bits_equal(div[f * n:(f + 1) * n], rd)
        assert bits_equal(lap[f * n:(f + 1) * n], rl)


def test_fv_sixty_four_fields_l6(ctx):
    """The L6 operator sweep: 64 seeded FP64 fields at L6 in one call,
    every field bit-identical to the oracle."""
    level, fields = 6, 64
    announce("fv 64 fields", level)
    m = product_mesh(level)
    om = S.mesh("oracle", level)
    fv = P.FV(ctx, m)
    rng = np.random.default_rng(64)
    phi = rng.random(fields * m.cells) - 0.5
    flux = rng.random(fields * m.edges) - 0.5
    east, north = fv.gradient(phi)
    div = fv.divergence(flux)
    o = S.oracle()
    n, e = m.cells, m.edges
    for f in range(fields):
        re, rn, rd = np.zeros(n), np.zeros(n), np.zeros(n)
        o.gradient_f64(om.handle, phi[f * n:(f + 1) * n], re, rn, 8)
        o.divergence_f64(om.handle, flux[f * e:(f + 1) * e], rd, 8)
        assert bits_equal(east[f * n:(f + 1) * n], re)
        assert bits_equal(north[f * n:(f + 1) * n], rn)
        assert bits_equal(div[f * n:(f + 1) * n], rd)


def test_fv_constant_and_zero_fields(ctx):
    """test_operator_conservation.cpp:36-57: constants give exactly 0."""
    m = product_mesh(4)
    announce("fv constants", 4)
    fv = P.FV(ctx, m)
    east, north = fv.gradient(np.full(m.cells, 287.5))
    lap = fv.laplacian(np.full(m.cells, 287.5))
    div = fv.divergence(np.zeros(m.edges))
    assert not east.any() and not north.any() and not lap.any() and not div.any()


@pytest.mark.parametrize("level", LEVELS)
def test_diffusion_source_bit_exact(ctx, level):
    announce("diffusion_source", level)
    m = product_mesh(level)
    om = S.mesh("oracle", level)
    g = P.Graph(ctx, m)
    t = 288.0 + seeded_cells(level, 1, np.float64) * 30.0
    out = g.diffusion_source(3.5e12, t)
    ref = np.zeros(m.cells)
    S.oracle().diffusion_source(om.handle, 3.5e12, t, ref, 1)
    assert bits_equal(out, ref)


def heat_case(level, index=0, k_base=1e6, sigma=0.3):
    m = product_mesh(level)
    seed = P.instance_seed(index)
    cond, ca, land, t_old = P.heat_instance(m, seed, k_base, sigma)
    forcing = P.heat_forcing(m, seed, 0)
    return m, seed, cond, ca, land, t_old, forcing


def csr_apply(m, diag, off, x):
    """diag*x, then entries in CSR order (heat_transport.cpp:1204-1214)."""
    y = diag * x
    start = m.offset[:-1].astype(np.int64)
    deg = (m.offset[1:] - m.offset[:-1]).astype(np.int64)
    for slot in range(int(deg.max())):
        rows = np.nonzero(deg > slot)[0]
        k = start[rows] + slot
        y[rows] = y[rows] + off[k] * x[m.neighbor[k].astype(np.int64)]
    return y


@pytest.mark.parametrize("level", [3, 4, 6])
def test_heat_matrix_apply_and_galerkin_bit_exact(ctx, level):
    announce("heat set_matrix / apply / Galerkin", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level)
    om = S.mesh("oracle", level)
    diag, off = S.heat_matrix(om, cond, ca)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    d0, o0 = s.level_values(0)
    assert bits_equal(d0, diag) and bits_equal(o0, off)
    x = t_old.copy()
    assert bits_equal(s.apply(x), csr_apply(m, diag, off, x))
    mg = S.Multigrid(S.oracle(), om)
    mg.set_matrix(diag, off)
    for j in range(1, g.depth):
        dj, oj = s.level_values(j)
        rd, ro = np.zeros(dj.size), np.zeros(max(oj.size, 1))
        S.oracle().lib.orc_mg_level_values(mg.handle, j, rd, ro)
        assert bits_equal(dj, rd)
        assert bits_equal(oj, ro[: oj.size])


@pytest.mark.parametrize("level", [0, 1, 2])
def test_coarse_inverse_is_the_dense_inverse(ctx, level):
    """Up to 400 cells the hierarchy is one dense level, so the V-cycle is the
    coarsest solve itself: the matrix-core block sweep's inverse (12, 42 and
    162 rows: one, three and eleven 16 x 16 blocks, with identity padding)
    against a dense LU solve, on the stiff variant, and its symmetry."""
    announce("coarse inverse", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level, k_base=0.5 * 6371000.0 ** 2)
    om = S.mesh("oracle", level)
    diag, off = S.heat_matrix(om, cond, ca)
    g = P.Graph(ctx, m)
    assert g.depth == 1
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    A = S.dense_matrix(om, diag, off)
    rng = np.random.default_rng(level)
    for _ in range(2):
        rhs = rng.standard_normal(m.cells)
        got = s.precondition(rhs)
        want = np.linalg.solve(A, rhs)
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.cond(A) * np.linalg.norm(want)
    u, v = rng.standard_normal(m.cells), rng.standard_normal(m.cells)
    assert abs(u @ s.precondition(v) - s.precondition(u) @ v) <= 1e-14 * np.linalg.norm(u) * np.linalg.norm(v) * \
        np.linalg.norm(np.linalg.inv(A), 2)


@pytest.mark.parametrize("level", [3, 4, 6])
def test_precondition_is_symmetric_positive(ctx, level):
    """The V(1,1) cycle with the same Jacobi smoother before and after the
    coarse correction is a symmetric positive definite operator."""
    announce("precondition symmetry", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level, k_base=0.5 * 6371000.0 ** 2)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    rng = np.random.default_rng(level)
    u, v = rng.standard_normal(m.cells), rng.standard_normal(m.cells)
    mu, mv = s.precondition(u), s.precondition(v)
    lhs, rhs = u @ mv, mu @ v
    assert abs(lhs - rhs) <= 1e-10 * np.sqrt(abs(u @ mu) * abs(v @ mv))
    assert u @ mu > 0 and v @ mv > 0


def oracle_solution(level, cond, ca, t_old, forcing, tol=1e-14):
    om = S.mesh("oracle", level)
    diag, off = S.heat_matrix(om, cond, ca)
    b = S.heat_rhs(om, ca, t_old, forcing)
    mg = S.Multigrid(S.oracle(), om)
    mg.set_matrix(diag, off)
    x, it = mg.bicgstab(diag, off, b, tol, 2000, 8)
    return x, it, diag, off, b


@pytest.mark.parametrize("level", LEVELS)
def test_heat_step_matches_oracle(ctx, level):
    """FP64 plain path: the converged solution agrees with the reference
    algorithm's to 1e-12 relative L2 (both converged to <= 1e-13); at the
    metric's 1e-10 the true residual is <= 1e-10."""
    announce("heat step vs oracle", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level)
    x_ref, _, diag, off, b = oracle_solution(level, cond, ca, t_old, forcing)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    x, info = s.heat_step(t_old, forcing, tolerance=1e-13)
    assert info.status == 0
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-12, err
    x10, info10 = s.heat_step(t_old, forcing, tolerance=1e-10)
    assert info10.status == 0 and info10.relative_residual <= 1e-10
    true_res = np.linalg.norm(b - csr_apply(m, diag, off, x10)) / np.linalg.norm(b)
    assert true_res <= 1e-10, true_res


@pytest.mark.parametrize("level", [4, 6])
def test_stiff_heat_step_matches_oracle(ctx, level):
    """Laplacian-dominated variant (K = 0.5 R^2): the V-cycle and BiCGSTAB
    do real work; solutions agree to 1e-9 relative L2 at tolerance 1e-12."""
    announce("stiff heat step vs oracle", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level, k_base=0.5 * 6371000.0 ** 2)
    x_ref, it_ref, diag, off, b = oracle_solution(level, cond, ca, t_old, forcing, tol=1e-13)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    x, info = s.heat_step(t_old, forcing, tolerance=1e-12)
    print("gpu iterations", info.iterations, "oracle iterations", it_ref, flush=True)
    assert info.status == 0 and info.relative_residual <= 1e-12
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-9, err


def test_workgroup_count_does_not_change_bits(ctx):
    announce("workgroup invariance", 5)
    m, seed, cond, ca, land, t_old, forcing = heat_case(5, k_base=0.5 * 6371000.0 ** 2)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    default = s.workgroups
    x0, i0 = s.heat_step(t_old, forcing, tolerance=1e-11)
    results = []
    for wg in (1, 7, default):
        s.workgroups = wg
        s.heat_set_matrix(m.entry_values(cond), ca)
        x, info = s.heat_step(t_old, forcing, tolerance=1e-11)
        results.append((wg, x, info.iterations))
    for wg, x, it in results:
        assert it == i0.iterations and bits_equal(x, x0), wg


@pytest.mark.parametrize("level", [4, 6])
def test_mixed_precision_error_vs_fp64_oracle(ctx, level):
    """Mixed form (FP64 BiCGSTAB, split V-cycle: FP64 first level-0 sweep,
    FP32 rest) reaches the FP64 true residual 1e-10 and <= 1e-8 relative
    error vs the FP64 oracle, in one V-cycle and no restart, as the plain
    form does on these inputs (an FP32 V-cycle without the FP64 first sweep
    leaves about 6e-8 and needs a second one)."""
    announce("mixed precision", level)
    m, seed, cond, ca, land, t_old, forcing = heat_case(level, sigma=1.0)
    x_ref, _, diag, off, b = oracle_solution(level, cond, ca, t_old, forcing)
    g = P.Graph(ctx, m)
    s = P.System(g, P.MIXED)
    s.heat_set_matrix(m.entry_values(cond), ca)
    x, info = s.heat_step(t_old, forcing, tolerance=1e-10)
    assert info.status == 0 and info.relative_residual <= 1e-10
    true_res = np.linalg.norm(b - csr_apply(m, diag, off, x)) / np.linalg.norm(b)
    assert true_res <= 1e-10
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print("mixed error", err, "restarts", info.restarts, "iterations", info.iterations, "v-cycles", info.v_cycles,
          flush=True)
    assert err <= 1e-8, err
    plain = P.System(g, P.PLAIN_FP64)
    plain.heat_set_matrix(m.entry_values(cond), ca)
    _, plain_info = plain.heat_step(t_old, forcing, tolerance=1e-10)
    assert info.restarts == 0 and info.v_cycles == plain_info.v_cycles, (info.v_cycles, plain_info.v_cycles)


def test_chained_steps_track_oracle(ctx):
    """The chained form at L4: twelve monthly steps, each T feeding the next."""
    level = 4
    announce("chained steps", level)
    m, seed, cond, ca, land, t_old, _ = heat_case(level)
    om = S.mesh("oracle", level)
    diag, off = S.heat_matrix(om, cond, ca)
    mg = S.Multigrid(S.oracle(), om)
    mg.set_matrix(diag, off)
    g = P.Graph(ctx, m)
    s = P.System(g)
    s.heat_set_matrix(m.entry_values(cond), ca)
    t_gpu, t_cpu = t_old.copy(), t_old.copy()
    for step in range(12):
        forcing = P.heat_forcing(m, seed, step)
        t_gpu, info = s.heat_step(t_gpu, forcing, tolerance=1e-13)
        b = S.heat_rhs(om, ca, t_cpu, forcing)
        t_cpu, _ = mg.bicgstab(diag, off, b, 1e-14, 100)
        assert info.status == 0
    err = np.linalg.norm(t_gpu - t_cpu) / np.linalg.norm(t_cpu)
    assert err <= 1e-12, err


def test_invalid_arguments_raise(ctx):
    announce("error paths")
    m = product_mesh(3)
    g = P.Graph(ctx, m)
    s = P.System(g)
    with pytest.raises(P.PsimError, match="multigrid matrix does not match the graph"):
        s.set_matrix(np.ones(m.cells - 1), np.zeros(m.directed))
    with pytest.raises(P.PsimError, match="set_matrix has not been called"):
        s.precondition(np.ones(m.cells))
    fv = P.FV(ctx, m)
    far = next(e for e in range(m.edges) if m.first[e] != 0 and m.second[e] != 0)
    with pytest.raises(P.PsimError, match="cell is not incident to edge"):
        fv.orientation(0, far)
    assert fv.orientation(int(m.first[0]), 0) == 1.0
    assert fv.orientation(int(m.second[0]), 0) == -1.0


@pytest.mark.parametrize("level,precision,case", [(3, P.PLAIN_FP64, "literal"), (6, P.PLAIN_FP64, "literal"),
                                                   (4, P.PLAIN_FP64, "stiff"), (6, P.MIXED, "sigma1"),
                                                   (4, P.MIXED, "stiff"), (5, P.PLAIN_FP64, "switch"),
                                                   (5, P.MIXED, "switch")])
def test_chain_launch_equals_separate_steps(ctx, level, precision, case):
    """psim_heat_chain (set_matrix + N steps in one launch, each step's last
    phase preparing the next step's right-hand side and first level-0 sweep)
    gives the same bits as psim_heat_set_matrix followed by N psim_heat_step
    launches, in both forms; the stiff inputs take several iterations per
    step, so the prepared values are written while a step still iterates; at
    L5 with conductance scale 1e8 the first steps end after one V-cycle and
    the later ones need a full iteration, so a prepared step goes on past the
    phase that prepares the next one."""
    announce("chained launch", level, "mixed" if precision == P.MIXED else "plain", case)
    if case == "stiff":
        m, seed, cond, ca, land, t_old, _ = heat_case(level, k_base=0.5 * 6371000.0 ** 2)
    elif case == "switch":
        m, seed, cond, ca, land, t_old, _ = heat_case(level, k_base=1e8)
    else:
        m, seed, cond, ca, land, t_old, _ = heat_case(level, sigma=1.0 if case == "sigma1" else 0.3)
    tol = 1e-10 if precision == P.MIXED or case == "switch" else 1e-11
    steps = 12 if case == "switch" else 6
    forcings = np.concatenate([P.heat_forcing(m, seed, k) for k in range(steps)])
    g = P.Graph(ctx, m)
    s = P.System(g, precision)
    chained, infos = s.heat_chain(t_old, forcings, tolerance=tol,
                                  conductance_per_entry=m.entry_values(cond), capacity_area=ca)
    assert len(infos) == steps and all(i.status == 0 and i.relative_residual <= tol for i in infos)
    assert all(i.device_ms > 0 for i in infos)
    if case == "stiff":
        assert all(i.iterations > 1 for i in infos), [i.iterations for i in infos]
    if case == "switch":
        v = [i.v_cycles for i in infos]
        assert any(a == 1 and b > 1 for a, b in zip(v, v[1:])), v
    s2 = P.System(g, precision)
    s2.heat_set_matrix(m.entry_values(cond), ca)
    t = t_old.copy()
    for k in range(steps):
        t, info = s2.heat_step(t, forcings[k * m.cells:(k + 1) * m.cells], tolerance=tol)
        assert info.iterations == infos[k].iterations
        assert info.v_cycles == infos[k].v_cycles
    assert bits_equal(chained, t)

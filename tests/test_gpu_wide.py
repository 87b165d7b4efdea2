"""Graphs whose rows are wider than the solver's unrolled row slots.

The solver program takes an unrolled, early-issued form on levels whose rows
have at most seven entries (every icosphere level and its aggregates) and a
general row loop on wider ones; the drop-in accepts any transport graph
(heat_transport.hpp's TransportGraph), so both are parity-tested. The graph
here is the L5 icosphere graph with every fourth cell also joined to its
second ring of neighbours (both directions, one weight and one conductance per
pair), which makes every level of its hierarchy wider than seven (18, 11 and 8
entries a row). On it:

- the assembled heat matrix and the SpMV bit for bit against the slot-order
  loops of heat_transport.cpp (the restatement the icosphere tests pin against
  the oracle), and each coarse level against the Galerkin product P^T A P of
  the device's level above (within 1e-14 of the sum of the absolute terms: the
  device sums in its own order);
- plain FP64 and mixed solves against scipy's sparse direct solve of the same
  system: true relative residual 1e-9 at tolerance 1e-10, and the error within
  the bound the residual fixes, kappa * residual, with kappa bounded by
  Gershgorin's discs (below 60 at conductance 1e6, where one V-cycle
  converges; about 3,400 for the stiff variant K = 0.5 R^2, where BiCGSTAB
  takes several iterations);
- a chain against the same steps launched one by one, bit for bit, and each
  step against its direct solve as above;
- the V-cycle's symmetry and positivity, the workgroup count's and a batch's
  invariance (bit for bit).
"""
import types

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import planet_game_amd as P
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu

DT = 2.63e6
LEVEL, EVERY = 5, 4
STIFF = 0.5 * 6371000.0 ** 2   # the stiff variant K = 0.5 R^2
_case = {}


def wide_case(k_base=1.0e6):
    """(graph arrays, conductance per entry, C*A, T_old, forcings) of the
    widened L5 graph; new pairs get half the mesh's mean weight and a
    lognormal conductance around k_base."""
    if k_base in _case:
        return _case[k_base]
    m = product_mesh(LEVEL)
    n = m.cells
    start = m.offset.astype(np.int64)
    seed = P.instance_seed(77)
    cond_edge, ca, _, t_old = P.heat_instance(m, seed, k_base, 0.3)
    ring1 = [m.graph_neighbour[start[c]:start[c + 1]].astype(np.int64).tolist() for c in range(n)]
    rows = [list(r) for r in ring1]
    wts = [m.weight[start[c]:start[c + 1]].tolist() for c in range(n)]
    conds = [cond_edge[m.edge[start[c]:start[c + 1]]].tolist() for c in range(n)]
    adj = [set(r) for r in ring1]
    w_new = 0.5 * float(m.weight.mean())
    rng = np.random.default_rng(5)
    for c in range(0, n, EVERY):
        for d in sorted({e for b in ring1[c] for e in ring1[b]} - adj[c] - {c}):
            k = k_base * float(np.exp(0.3 * rng.standard_normal()))
            adj[c].add(d)
            adj[d].add(c)
            for a, b in ((c, d), (d, c)):
                rows[a].append(b)
                wts[a].append(w_new)
                conds[a].append(k)
    offset = np.zeros(n + 1, np.uint64)
    offset[1:] = np.cumsum([len(r) for r in rows])
    g = types.SimpleNamespace(
        cells=n, area=m.area.copy(), offset=offset,
        graph_neighbour=np.concatenate(rows).astype(np.uint64),
        weight=np.concatenate([np.asarray(w, np.float64) for w in wts]),
        weight_sum=np.array([sum(w) for w in wts]))
    cond = np.concatenate([np.asarray(k, np.float64) for k in conds])
    forcings = np.stack([P.heat_forcing(m, seed, s) for s in range(6)])
    _case[k_base] = (g, cond, ca, t_old, forcings)
    return _case[k_base]


def slot_loops(g, cond, ca, x):
    """heat_transport.cpp's assembly (diag = C A + dt sum g_k, off = -dt g_k,
    g_k = K_k w_k, summed in row order) and SpMV (diag x, then the row's
    entries in order), vectorised over rows one slot at a time."""
    start = g.offset[:-1].astype(np.int64)
    deg = (g.offset[1:] - g.offset[:-1]).astype(np.int64)
    nb = g.graph_neighbour.astype(np.int64)
    gk = cond * g.weight
    s = np.zeros(g.cells)
    for slot in range(int(deg.max())):
        rows = np.nonzero(deg > slot)[0]
        s[rows] = s[rows] + gk[start[rows] + slot]
    diag = ca + DT * s
    off = -DT * gk
    y = diag * x
    for slot in range(int(deg.max())):
        rows = np.nonzero(deg > slot)[0]
        k = start[rows] + slot
        y[rows] = y[rows] + off[k] * x[nb[k]]
    return diag, off, y


def csr(n, offset, column, diag, off):
    rows = np.repeat(np.arange(n), np.diff(offset.astype(np.int64)))
    return (sp.csr_matrix((off, (rows, column.astype(np.int64))), shape=(n, n)) + sp.diags(diag)).tocsr()


def heat_system(g, cond, ca, t_old, forcing):
    diag, off, _ = slot_loops(g, cond, ca, t_old)
    A = csr(g.cells, g.offset, g.graph_neighbour, diag, off)
    return A, ca * t_old + DT * g.area * forcing


def kappa_bound(A):
    """Gershgorin bound on the 2-norm condition number of the symmetric,
    strictly diagonally dominant A: max(d + r) / min(d - r)."""
    d = A.diagonal()
    r = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    assert (d - r).min() > 0
    return float((d + r).max() / (d - r).min())


def wide_system(ctx, precision=P.PLAIN_FP64, k_base=1.0e6):
    g, cond, ca, t_old, forcings = wide_case(k_base)
    s = P.System(P.Graph(ctx, g), precision)
    s.heat_set_matrix(cond, ca)
    return s


def test_wide_levels_matrix_and_galerkin(ctx):
    g, cond, ca, t_old, _ = wide_case()
    h = P.Hierarchy(g)
    widths = [int(np.diff(lv["offset"].astype(np.int64)).max()) for lv in h.levels]
    announce("wide graph levels", [lv["size"] for lv in h.levels], "widths", widths)
    assert len(widths) >= 3 and min(widths) > 7
    s = wide_system(ctx)
    diag, off, y = slot_loops(g, cond, ca, t_old)
    d0, o0 = s.level_values(0)
    assert bits_equal(d0, diag) and bits_equal(o0, off)
    assert bits_equal(s.apply(t_old), y)
    fine = (d0, o0)
    for j in range(1, len(h.levels)):
        up, lv = h.levels[j - 1], h.levels[j]
        A = csr(up["size"], up["offset"], up["column"], fine[0], fine[1])
        agg = up["aggregate"].astype(np.int64)
        R = sp.csr_matrix((np.ones(up["size"]), (np.arange(up["size"]), agg)), shape=(up["size"], lv["size"]))
        Ac = (R.T @ A @ R).tocsr()
        scale = (R.T @ abs(A) @ R).tocsr()
        dj, oj = s.level_values(j)
        B = csr(lv["size"], lv["offset"], lv["column"], dj, oj)
        assert (Ac != 0).sum() == (B != 0).sum() == lv["size"] + lv["nnz"]
        err = abs(Ac - B)
        bound = 1e-14 * scale
        assert (err > bound).sum() == 0, j
        fine = (dj, oj)


@pytest.mark.parametrize("k_base", [1.0e6, STIFF])
@pytest.mark.parametrize("precision", [P.PLAIN_FP64, P.MIXED])
def test_wide_solve_matches_direct_solve(ctx, precision, k_base):
    announce("wide graph solve", "precision", precision, "k_base", k_base)
    g, cond, ca, t_old, forcings = wide_case(k_base)
    s = wide_system(ctx, precision, k_base)
    x, info = s.heat_step(t_old, forcings[0], DT, 1e-10, 1000)
    A, b = heat_system(g, cond, ca, t_old, forcings[0])
    ref = spla.spsolve(A.tocsc(), b)
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    kappa = kappa_bound(A)
    print("iterations", info.iterations, "v_cycles", info.v_cycles, "recursive", info.relative_residual,
          "true", true, "error", err, "kappa bound", kappa, flush=True)
    assert info.status == 0 and info.relative_residual <= 1e-10
    assert true <= 1e-9 and err <= kappa * true * (1 + 1e-6) + 1e-14


@pytest.mark.parametrize("k_base", [1.0e6, STIFF])
@pytest.mark.parametrize("precision", [P.PLAIN_FP64, P.MIXED])
def test_wide_chain_equals_separate_steps(ctx, precision, k_base):
    g, cond, ca, t_old, forcings = wide_case(k_base)
    announce("wide graph chain", len(forcings), "steps, precision", precision, "k_base", k_base)
    s = wide_system(ctx, precision, k_base)
    chained, infos = s.heat_chain(t_old, forcings, DT, 1e-10, 1000)
    t = t_old
    for k, f in enumerate(forcings):
        t_next, info = s.heat_step(t, f, DT, 1e-10, 1000)
        assert info.status == 0 and infos[k].status == 0, k
        assert (infos[k].iterations, infos[k].v_cycles) == (info.iterations, info.v_cycles), k
        # each step against the direct solve of the same step
        A, b = heat_system(g, cond, ca, t, f)
        ref = spla.spsolve(A.tocsc(), b)
        err = np.linalg.norm(t_next - ref) / np.linalg.norm(ref)
        true = np.linalg.norm(b - A @ t_next) / np.linalg.norm(b)
        print("step", k, "iterations", info.iterations, "v_cycles", info.v_cycles, "true", true, "error", err,
              flush=True)
        assert true <= 1e-9 and err <= kappa_bound(A) * true * (1 + 1e-6) + 1e-14, k
        t = t_next
    assert bits_equal(chained, t)


def test_wide_precondition_is_symmetric_positive(ctx):
    announce("wide graph precondition symmetry")
    g = wide_case(STIFF)[0]
    s = wide_system(ctx, k_base=STIFF)
    rng = np.random.default_rng(3)
    u, v = rng.standard_normal(g.cells), rng.standard_normal(g.cells)
    mu, mv = s.precondition(u), s.precondition(v)
    assert abs(u @ mv - mu @ v) <= 1e-10 * np.sqrt(abs(u @ mu) * abs(v @ mv))
    assert u @ mu > 0 and v @ mv > 0


def test_wide_workgroups_and_batch_do_not_change_bits(ctx):
    g, cond, ca, t_old, forcings = wide_case(STIFF)
    announce("wide graph workgroup and batch invariance")
    s = wide_system(ctx, k_base=STIFF)
    x0, i0 = s.heat_step(t_old, forcings[0], DT, 1e-11, 1000)
    for wg in (1, 5, s.workgroups):
        s.workgroups = wg
        s.heat_set_matrix(cond, ca)
        x, info = s.heat_step(t_old, forcings[0], DT, 1e-11, 1000)
        assert info.iterations == i0.iterations and bits_equal(x, x0), wg
    # a batch of the wide system (plain and mixed) beside an icosphere system
    m4 = product_mesh(4)
    c4, ca4, _, t4 = P.heat_instance(m4, P.instance_seed(78), 1.0e6, 0.3)
    f4 = P.heat_forcing(m4, P.instance_seed(78), 0)
    members = [(P.System(P.Graph(ctx, g), P.PLAIN_FP64), cond, ca, t_old, forcings[0]),
               (P.System(P.Graph(ctx, g), P.MIXED), cond, ca, t_old, forcings[0]),
               (P.System(P.Graph(ctx, m4), P.PLAIN_FP64), m4.entry_values(c4), ca4, t4, f4)]
    arrays = [[P.DeviceArray.of(ctx, a) for a in (k, c, t, f)] + [P.DeviceArray(ctx, t.size)]
              for _, k, c, t, f in members]
    batch = P.Batch(ctx, [m[0] for m in members])
    infos = batch.heat_step([a[2] for a in arrays], [a[3] for a in arrays], [a[4] for a in arrays], DT, 1e-10,
                            1000, conductance=[a[0] for a in arrays], capacity=[a[1] for a in arrays])
    for (system, k, c, t, f), a, info in zip(members, arrays, infos):
        system.heat_set_matrix(k, c)
        alone, ainfo = system.heat_step(t, f, DT, 1e-10, 1000)
        assert info.status == 0 and info.iterations == ainfo.iterations, system.precision
        assert bits_equal(a[4].download(), alone), system.precision

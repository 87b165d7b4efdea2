"""Host half of libpsim (no GPU): the library loads and exports every
symbol its header declares; the product's mesh, hierarchy and workload are
bit-identical to the oracle's (and so to the reference's); argument errors
map to the reference's exception messages."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import planet_game_amd as P
import support as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_fixtures.json")))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "psim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psim_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 40
    lib = C.CDLL(P.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_compiles_as_c():
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "t.c")
        with open(src, "w") as fh:
            fh.write('#include "psim.h"\nint main(void) { return psim_version() == 1 ? 0 : 1; }\n')
        exe = os.path.join(tmp, "t")
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src,
                        P.LIB_PATH, "-o", exe, f"-Wl,-rpath,{os.path.dirname(P.LIB_PATH)}"], check=True)
        assert subprocess.run([exe]).returncode == 0


@pytest.mark.parametrize("level", range(0, 7))
def test_product_mesh_bit_identical(level):
    m = P.Mesh(level)
    o = S.mesh("oracle", level)
    f = FIXTURES["meshes"][str(level)]
    ours = dict(center=m.center, east=m.east, north=m.north, area=m.area, edge_offset=m.edge_offset,
                edge_count=m.edge_count, neighbor=m.neighbor, edge=m.edge, normal_east=m.normal_east,
                normal_north=m.normal_north, first=m.first, second=m.second, length=m.length,
                distance=m.distance, offset=m.offset, weight=m.weight, weight_sum=m.weight_sum)
    for name, array in ours.items():
        assert S.sha(array) == f[name], name
        assert np.array_equal(array.view(np.uint8), o.arrays()[name].view(np.uint8)), name


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_product_hierarchy_bit_identical(level):
    levels = P.Hierarchy(P.Mesh(level)).levels
    ref = FIXTURES["hierarchies"][str(level)]
    assert [(lv["size"], lv["nnz"]) for lv in levels] == [(r["size"], r["nnz"]) for r in ref]
    for lv, r in zip(levels, ref):
        for key in ("offset", "column", "aggregate", "target", "diagonal", "value"):
            assert S.sha(lv[key]) == r[key], key


@pytest.mark.parametrize("level", [3, 4, 5, 6])
def test_product_workload_bit_identical(level):
    m = P.Mesh(level)
    f = FIXTURES["workload"][str(level)]
    assert P.instance_seed(0) == f["seed"]
    cond, ca, land, temp = P.heat_instance(m, f["seed"])
    assert S.sha(cond) == f["conductance"] and S.sha(ca) == f["capacity_area"]
    assert S.sha(land) == f["land"] and S.sha(temp) == f["temperature"]
    assert S.sha(P.heat_forcing(m, f["seed"], 0)) == f["forcing_0"]
    o = S.mesh("oracle", level)
    for step in (1, 57, 119):
        assert np.array_equal(P.heat_forcing(m, f["seed"], step), S.heat_forcing(o, f["seed"], step))


def test_mesh_argument_errors():
    with pytest.raises(P.PsimError, match="icosphere radius must be finite and positive"):
        P.Mesh(2, -1.0)
    with pytest.raises(P.PsimError, match="icosphere radius must be finite and positive"):
        P.Mesh(2, float("nan"))


def test_hierarchy_rejects_bad_graph():
    m = P.Mesh(2)
    bad = m.graph_neighbour.copy()
    bad[3] = m.cells + 5
    h = C.c_void_p()
    status = P.lib.psim_hierarchy_create(m.cells, m.offset, bad, m.weight, m.weight_sum, C.byref(h))
    assert status == 1 and b"out of range" in P.lib.psim_last_error()
    # offsets: not starting at 0, decreasing, or past the entry arrays
    for offset, message in ((m.offset + 1, b"start at 0"),
                            (np.concatenate([m.offset[:5], m.offset[3:4], m.offset[6:]]), b"must not decrease")):
        offset = np.ascontiguousarray(offset, np.uint64)
        status = P.lib.psim_hierarchy_create(m.cells, offset, m.graph_neighbour, m.weight, m.weight_sum,
                                             C.byref(h))
        assert status == 1 and message in P.lib.psim_last_error(), message
    status = _raw_hierarchy_create(m.cells, m.offset, None, None, m.weight_sum, h)
    assert status == 1 and b"null argument" in P.lib.psim_last_error()


def _raw_hierarchy_create(n, offset, neighbour, weight, weight_sum, h):
    """psim_hierarchy_create without the binding's array types, so that NULL
    arrays can be passed."""
    f = C.CDLL(P.LIB_PATH).psim_hierarchy_create
    f.restype = C.c_int
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    return f(C.c_uint64(n), ptr(offset), ptr(neighbour), ptr(weight), ptr(weight_sum), C.byref(h))


def test_hierarchy_of_a_graph_without_entries():
    """Isolated nodes only: no neighbour or weight arrays are needed, and
    every node is its own coarsest row."""
    n = 5
    h = C.c_void_p()
    status = _raw_hierarchy_create(n, np.zeros(n + 1, np.uint64), None, None, np.zeros(n), h)
    assert status == 0, P.lib.psim_last_error()
    assert P.lib.psim_hierarchy_depth(h) == 1
    P.lib.psim_hierarchy_destroy(h)


def test_no_gpu_is_a_device_error():
    """Without a visible GPU the device entry points fail loudly (no CPU
    fallback exists)."""
    h = C.c_void_p()
    status = P.lib.psim_ctx_create(0, C.byref(h))
    if status == 0:
        P.lib.psim_ctx_destroy(h)
        pytest.skip("a GPU is visible here")
    assert status == 3 and b"no GPU" in P.lib.psim_last_error()

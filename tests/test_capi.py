"""Handle API (include/cedar_capi.h) on the CPU: config files, handles, topologies, vectors,
matrix entry and the text dumps.  None of these calls touch the GPU (vectors and matrices are
host arrays until a matvec or a solve)."""
import ctypes
import json
import re

import numpy as np
import pytest

from cedar_amd import capi as C
from tests.capi_entries import entries2, random_symmetric2, random_symmetric3, entries3

NUM = re.compile(r"-?\d\.\d{7}e[-+]\d\d")


def _write(tmp_path, obj, name="conf.json"):
    p = tmp_path / name
    p.write_text(json.dumps(obj) if not isinstance(obj, str) else obj)
    return p


def test_every_declared_function_is_exported_and_bound():
    text = open(C._abi.HEADER.replace("cedar_hip.h", "cedar_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(cedar_\w+)\s*\(", text))
    assert len(names) == 26
    lib = ctypes.CDLL(C._abi.LIB_PATH)
    assert all(hasattr(lib, n) for n in names)
    assert names == set(C.PROTOTYPES)


def test_config_reads_like_the_reference(tmp_path):
    lib = C.load()
    # the reference's own test configuration (test/2d/mpi/test-cgsolve-1.json), retyped as data
    conf = {"log": ["status", "error", "info"],
            "cg-config": {"solver": {"cycle": {"nrelax-pre": 2, "nrelax-post": 1},
                                     "relaxation": "point", "num-levels": 5}},
            "solver": {"cycle": {"nrelax-pre": 2, "nrelax-post": 1}, "relaxation": "point",
                       "tol": 1e-10, "max-iter": 1, "min-coarse": 3, "num-levels": 3,
                       "cg-solver": "cedar"},
            "grid": {"local": False, "n": [400, 400], "np": [2, 2]},
            "misc": {"f": 2.75, "neg": -3.5, "t": True, "s": "x", "esc": "a\\u00e9\"b"}}
    h = C.config(_write(tmp_path, conf))
    assert (h.value >> 28) == 5
    v = ctypes.c_int()

    def get(path):
        st = lib.cedar_config_get_int(h, path.encode(), ctypes.byref(v))
        return v.value if st == C.SUCCESS else st
    assert get("solver.max-iter") == 1
    assert get("solver.cycle.nrelax-pre") == 2
    assert get("cg-config.solver.num-levels") == 5
    assert get("misc.f") == 2 and get("misc.neg") == -3      # truncation, as get<int>
    assert get("misc.t") == 1                                # a boolean reads as 0/1
    assert get("misc.s") == C.ERR_CONFIG_PATH                # a string does not convert
    assert get("solver.nope") == C.ERR_CONFIG_PATH
    assert get("grid.n") == C.ERR_CONFIG_PATH                # an array is not an int
    assert lib.cedar_log_init(h) == C.SUCCESS
    assert lib.cedar_config_free(ctypes.byref(h)) == C.SUCCESS
    assert h.value == C.CONFIG_NULL
    assert lib.cedar_config_get_int(h, b"solver.max-iter", ctypes.byref(v)) == C.ERR_CONFIG


def test_config_errors(tmp_path):
    lib = C.load()
    h = ctypes.c_int(7)
    assert lib.cedar_config_create(str(tmp_path / "absent.json").encode(),
                                   ctypes.byref(h)) == C.ERR_FNAME
    assert h.value == C.CONFIG_NULL
    for bad in ['{"a": 1,}', '{"a": }', '[1, 2', '{"a": 01}', '{"a": 1} x', '']:
        p = _write(tmp_path, bad, "bad.json")
        assert lib.cedar_config_create(str(p).encode(), ctypes.byref(h)) == C.ERR_CONFIG, bad


def test_handles_kinds_and_nulls(tmp_path):
    lib = C.load()
    t = C.topo2d(5, 4)
    assert (t.value >> 28) == 1
    v = C.vec(t)
    assert (v.value >> 28) == 4
    # a handle of another kind is rejected by kind
    nd = ctypes.c_int()
    assert lib.cedar_vec_getdim(t, ctypes.byref(nd)) == C.ERR_VEC
    assert lib.cedar_topo_free(ctypes.byref(v)) == C.ERR_TOPO
    assert lib.cedar_vec_getdim(v, ctypes.byref(nd)) == C.SUCCESS and nd.value == 2
    # freeing the topology while a vector uses it keeps the vector valid
    assert lib.cedar_topo_free(ctypes.byref(t)) == C.SUCCESS and t.value == C.TOPO_NULL
    a = C.vec_array(v)
    assert a.shape == (6, 7) and not a.any()
    assert lib.cedar_vec_free(ctypes.byref(v)) == C.SUCCESS and v.value == C.VEC_NULL
    assert lib.cedar_vec_free(ctypes.byref(v)) == C.ERR_VEC
    assert lib.cedar_vec_getdim(0x4000FFFF, ctypes.byref(nd)) == C.ERR_VEC


def test_topology_is_one_process(tmp_path):
    lib = C.load()
    h = ctypes.c_int()
    one = (ctypes.c_uint * 1)(8)
    two = (ctypes.c_uint * 2)(4, 4)
    assert lib.cedar_topo_create2d(None, 8, 8, two, two, 2, 2, ctypes.byref(h)) == C.ERR_TOPO
    assert h.value == C.TOPO_NULL
    assert lib.cedar_topo_create2d(None, 8, 9, one, one, 1, 1, ctypes.byref(h)) == C.ERR_TOPO
    assert lib.cedar_topo_create2d(None, 0, 0, (ctypes.c_uint * 1)(0), (ctypes.c_uint * 1)(0),
                                   1, 1, ctypes.byref(h)) == C.ERR_TOPO
    t3 = C.topo3d(3, 4, 5)
    v3 = C.vec(t3, nd=3)
    assert C.vec_array(v3).shape == (7, 6, 5)
    a, b = ctypes.c_uint(), ctypes.c_uint()
    assert lib.cedar_vec_len2d(v3, ctypes.byref(a), ctypes.byref(b)) == C.ERR_DIM
    w = ctypes.c_int()
    assert lib.cedar_vec_create2d(t3, ctypes.byref(w)) == C.ERR_DIM


def _mat2(tmp_path, nx, ny, sten):
    lib = C.load()
    conf = C.config(_write(tmp_path, {"solver": {"relaxation": "point"}}))
    topo = C.topo2d(nx, ny)
    m = ctypes.c_int()
    C.check(lib.cedar_mat_create2d(conf, topo, sten, ctypes.byref(m)))
    return m, topo


def test_five_point_rejects_diagonal_directions(tmp_path):
    m, _ = _mat2(tmp_path, 4, 4, C.FIVE_PT)
    for d in (C.SW, C.NW, C.NE, C.SE):
        assert C.set2d(m, [(1, 1, d, -1.0)]) == C.ERR_STENCIL
    assert C.set2d(m, [(1, 1, 9, -1.0)]) == C.ERR_STENCIL
    assert C.set2d(m, [(4, 1, C.C, 4.0)]) == C.ERR_DIM
    assert C.set2d(m, [(1, 4, C.C, 4.0)]) == C.ERR_DIM
    for d in (C.C, C.W, C.S, C.N, C.E):
        assert C.set2d(m, [(1, 1, d, 1.0)]) == C.SUCCESS


def test_set2d_leaves_values_unchanged(tmp_path):
    m, _ = _mat2(tmp_path, 3, 3, C.NINE_PT)
    coords = (C.Coord2 * 2)(C.Coord2(0, 0, C.E), C.Coord2(0, 0, C.C))
    vals = (ctypes.c_double * 2)(-1.5, 4.0)
    assert C.load().cedar_mat_set2d(m, 2, coords, vals) == C.SUCCESS
    assert list(vals) == [-1.5, 4.0]


@pytest.mark.parametrize("ns,sten", [(3, C.FIVE_PT), (5, C.NINE_PT)])
def test_mat_dump_rows_are_matrix_rows(tmp_path, monkeypatch, ns, sten):
    """Each line of op-0-0.txt holds a row of A in the vertex order NW N NE W C E SW S SE (5-point:
    N W C E S), so entering a matrix by rows and dumping it gives those rows back."""
    nx, ny = 6, 5
    so = random_symmetric2(nx, ny, ns, seed=ns, integer=True)   # exact in 8 significant digits
    ent = entries2(so)
    m, _ = _mat2(tmp_path, nx, ny, sten)
    assert C.set2d(m, ent) == C.SUCCESS
    monkeypatch.chdir(tmp_path)
    assert C.load().cedar_mat_dump(m) == C.SUCCESS
    lines = (tmp_path / "op-0-0.txt").read_text().splitlines()
    assert len(lines) == nx * ny
    order = [C.NW, C.N, C.NE, C.W, C.C, C.E, C.SW, C.S, C.SE] if ns == 5 else \
        [C.N, C.W, C.C, C.E, C.S]
    row = {}
    for (i, j, d, v) in ent:
        row[(i, j, d)] = v
    for ln in lines:
        idx, i1, j1 = [int(t) for t in re.findall(r"-?\d+(?=[ ,])", ln[:16])[:3]]
        i, j = i1 - 2, j1 - 2
        assert idx == (j1 - 2) * (nx + 2) + i1 - 2
        got = [float(t) for t in NUM.findall(ln)]
        want = [row.get((i, j, d), 0.0) for d in order]
        assert got == want, (i, j)


def test_mat_dump_3d_and_set3d(tmp_path, monkeypatch):
    lib = C.load()
    conf = C.config(_write(tmp_path, {}))
    t = C.topo3d(3, 2, 2)
    m = ctypes.c_int()
    C.check(lib.cedar_mat_create3d(conf, t, C.XXVII_PT, ctypes.byref(m)))
    so = random_symmetric3(3, 2, 2, 14, seed=1)
    assert C.set3d(m, entries3(so)) == C.SUCCESS
    assert C.set3d(m, [(0, 0, 0, 14, 1.0)]) == C.ERR_STENCIL
    assert C.set3d(m, [(0, 2, 0, 0, 1.0)]) == C.ERR_DIM
    monkeypatch.chdir(tmp_path)
    assert lib.cedar_mat_dump(m) == C.SUCCESS
    lines = (tmp_path / "so-0.0.0").read_text().split()
    assert int(lines[0]) == 14 and int(lines[1]) == so.size
    vals = np.array([float(x) for x in lines[2:]])
    assert np.allclose(vals, so.ravel(), rtol=1e-5, atol=0)    # default stream precision
    top = ctypes.c_int()
    assert lib.cedar_mat_gettopo(m, ctypes.byref(top)) == C.SUCCESS and top.value == t.value
    m7 = ctypes.c_int()
    C.check(lib.cedar_mat_create3d(conf, t, C.SEVEN_PT, ctypes.byref(m7)))
    assert C.set3d(m7, [(0, 0, 0, 4, 1.0)]) == C.ERR_STENCIL
    assert lib.cedar_mat_free(ctypes.byref(m)) == C.SUCCESS and m.value == C.MAT_NULL
    assert lib.cedar_mat_gettopo(m, ctypes.byref(top)) == C.ERR_MAT


def test_solver_create_rejects_bad_settings_before_any_gpu_work(tmp_path):
    lib = C.load()
    topo = C.topo2d(7, 7)
    for conf in ({"solver": {"relaxation": "zebra"}}, {"solver": {"cycle": {"type": "w"}}},
                 {"solver": {"cg-solver": "qr"}},
                 {"grid": {"periodic": [False, False, True]}},
                 {"gpu": {"precision": "fp16"}}):
        c = C.config(_write(tmp_path, conf))
        m = ctypes.c_int()
        C.check(lib.cedar_mat_create2d(c, topo, C.FIVE_PT, ctypes.byref(m)))
        s = ctypes.c_int(3)
        assert lib.cedar_solver_create(m, ctypes.byref(s)) == C.ERR_CONFIG, conf
        assert s.value == C.SOLVER_NULL
    s = ctypes.c_int()
    assert lib.cedar_solver_create(C.MAT_NULL, ctypes.byref(s)) == C.ERR_MAT
    assert lib.cedar_solver_run(C.SOLVER_NULL, C.VEC_NULL, C.VEC_NULL) == C.ERR_SOLVER
    assert lib.cedar_solver_destroy(ctypes.byref(s)) == C.ERR_SOLVER


def test_headers_are_plain_c(tmp_path):
    """Both headers compile as C99 (a C program or an FFI generator can include them)."""
    import os
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    inc = os.path.dirname(C._abi.HEADER)
    src = tmp_path / "h.c"
    src.write_text('#include "cedar_hip.h"\n#include "cedar_capi.h"\n'
                   "int main(void) { cedar_coord_2d c = {0, 0, BMG2_E}; cdr_solver_settings s;\n"
                   "  (void)c; (void)s; return CEDAR_SUCCESS; }\n")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{inc}",
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

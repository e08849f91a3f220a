"""ctypes binding of the handle-based C API in include/cedar_capi.h (libcedar_hip.so).

The functions keep the reference's C names and semantics (include/cedar/capi.h); this module
only declares their prototypes and the constants, plus `Mat2`/`Vec` helpers used by tests.
"""
import ctypes

import numpy as np

from . import _abi

SUCCESS, ERR_TOPO, ERR_KIND, ERR_HANDLE, ERR_MAT, ERR_VEC, ERR_DIM, ERR_STENCIL, ERR_SOLVER, \
    ERR_CONFIG, ERR_FNAME, ERR_CONFIG_PATH, ERR_OTHER = range(13)

TOPO_NULL, SOLVER_NULL, MAT_NULL, VEC_NULL, CONFIG_NULL = (0x10000000, 0x20000000, 0x30000000,
                                                           0x40000000, 0x50000000)
# bmg2_dir (include/cedar/2d/base_types.h)
C, W, S, SW, NW, SE, N, NE, E = range(9)
FIVE_PT, NINE_PT = 0, 1
SEVEN_PT, XXVII_PT = 0, 1


class Coord2(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint), ("j", ctypes.c_uint), ("dir", ctypes.c_int)]


class Coord3(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint), ("j", ctypes.c_uint), ("k", ctypes.c_uint),
                ("dir", ctypes.c_int)]


_i = ctypes.c_int
_u = ctypes.c_uint
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)
_dp = ctypes.POINTER(ctypes.c_double)

PROTOTYPES = {
    "cedar_log_init": (_i, [_i]),
    "cedar_config_create": (_i, [ctypes.c_char_p, _ip]),
    "cedar_config_get_int": (_i, [_i, ctypes.c_char_p, _ip]),
    "cedar_config_free": (_i, [_ip]),
    "cedar_topo_create2d": (_i, [ctypes.c_void_p, _u, _u, _up, _up, _i, _i, _ip]),
    "cedar_topo_create3d": (_i, [ctypes.c_void_p, _u, _u, _u, _up, _up, _up, _i, _i, _i, _ip]),
    "cedar_topo_free": (_i, [_ip]),
    "cedar_vec_create2d": (_i, [_i, _ip]),
    "cedar_vec_create3d": (_i, [_i, _ip]),
    "cedar_vec_baseptr": (_i, [_i, ctypes.POINTER(_dp)]),
    "cedar_vec_len2d": (_i, [_i, _up, _up]),
    "cedar_vec_len3d": (_i, [_i, _up, _up, _up]),
    "cedar_vec_getdim": (_i, [_i, _ip]),
    "cedar_vec_free": (_i, [_ip]),
    "cedar_mat_create2d": (_i, [_i, _i, _i, _ip]),
    "cedar_mat_create3d": (_i, [_i, _i, _i, _ip]),
    "cedar_mat_set2d": (_i, [_i, _u, ctypes.POINTER(Coord2), _dp]),
    "cedar_mat_set3d": (_i, [_i, _u, ctypes.POINTER(Coord3), _dp]),
    "cedar_matvec": (_i, [_i, _i, _i]),
    "cedar_mat_gettopo": (_i, [_i, _ip]),
    "cedar_mat_dump": (_i, [_i]),
    "cedar_mat_free": (_i, [_ip]),
    "cedar_solver_create": (_i, [_i, _ip]),
    "cedar_solver_run": (_i, [_i, _i, _i]),
    "cedar_solver_destroy": (_i, [_ip]),
    "cedar_solver_free": (_i, [_ip]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        lib = _abi.load()
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class CedarError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what} returned {code}")


def check(code, what=""):
    if code != SUCCESS:
        raise CedarError(code, what)


def handle():
    return ctypes.c_int(0)


def config(path):
    h = handle()
    check(load().cedar_config_create(str(path).encode(), ctypes.byref(h)), "cedar_config_create")
    return h


def topo2d(nx, ny):
    h = handle()
    lnx, lny = (ctypes.c_uint * 1)(nx), (ctypes.c_uint * 1)(ny)
    check(load().cedar_topo_create2d(None, nx, ny, lnx, lny, 1, 1, ctypes.byref(h)), "topo2d")
    return h


def topo3d(nx, ny, nz):
    h = handle()
    lnx, lny, lnz = (ctypes.c_uint * 1)(nx), (ctypes.c_uint * 1)(ny), (ctypes.c_uint * 1)(nz)
    check(load().cedar_topo_create3d(None, nx, ny, nz, lnx, lny, lnz, 1, 1, 1, ctypes.byref(h)),
          "topo3d")
    return h


def vec(topo, nd=2):
    h = handle()
    fn = load().cedar_vec_create2d if nd == 2 else load().cedar_vec_create3d
    check(fn(topo, ctypes.byref(h)), "vec_create")
    return h


def vec_array(v):
    """numpy view (Fortran-order grid as (.., jj, ii)) of a vector's host storage."""
    lib = load()
    nd = ctypes.c_int()
    check(lib.cedar_vec_getdim(v, ctypes.byref(nd)))
    base = _dp()
    check(lib.cedar_vec_baseptr(v, ctypes.byref(base)))
    if nd.value == 2:
        a, b = ctypes.c_uint(), ctypes.c_uint()
        check(lib.cedar_vec_len2d(v, ctypes.byref(a), ctypes.byref(b)))
        shape = (b.value, a.value)
    else:
        a, b, c = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        check(lib.cedar_vec_len3d(v, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        shape = (c.value, b.value, a.value)
    return np.ctypeslib.as_array(base, shape=shape)


def set2d(mat, entries):
    """entries: iterable of (i, j, dir, value) in vertex form."""
    entries = list(entries)
    coords = (Coord2 * max(len(entries), 1))()
    vals = (ctypes.c_double * max(len(entries), 1))()
    for n, (i, j, d, v) in enumerate(entries):
        coords[n] = Coord2(i, j, d)
        vals[n] = v
    return load().cedar_mat_set2d(mat, len(entries), coords, vals)


def set3d(mat, entries):
    entries = list(entries)
    coords = (Coord3 * max(len(entries), 1))()
    vals = (ctypes.c_double * max(len(entries), 1))()
    for n, (i, j, k, d, v) in enumerate(entries):
        coords[n] = Coord3(i, j, k, d)
        vals[n] = v
    return load().cedar_mat_set3d(mat, len(entries), coords, vals)

"""ctypes bindings for the 3D FP64 CPU oracle (bmg3_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product (cedar_amd) never imports this module.

Array convention: a 3D grid function with ghost-inclusive extents (ii, jj, kk) is a C-contiguous
numpy array of shape (kk, jj, ii) -- the bytes of the reference's Fortran array Q(ii, jj, kk).
A stencil is (nstncl, kk, jj, ii) with nstncl 4 (7-point) or 14 (27-point), an interpolation
operator (26, kkc, jjc, iic).
"""
import ctypes

import numpy as np

from . import oracle as _o

DOWN, UP = 0, 1
RELAX3 = {"point": 0, "plane-xy": 4, "plane-xz": 5, "plane-yz": 6, "plane-xyz": 7}
PLANE = {"xy": 0, "xz": 1, "yz": 2}

_i = ctypes.c_int
_p = _o._ptr


class _Settings3(ctypes.Structure):
    _fields_ = [("relax", _i), ("nrelax_pre", _i), ("nrelax_post", _i), ("num_levels", _i),
                ("maxiter", _i), ("tol", ctypes.c_double), ("min_coarse", _i),
                ("plane", _o._Settings), ("cycle", _i), ("periodic", _i)]


_init = False


def lib():
    global _init
    L = _o.lib()
    if not _init:
        L.orc3_norm.restype = ctypes.c_double
        L.orc3_solver_create.restype = ctypes.c_void_p
        L.orc3_solver_create.argtypes = [_o._dp, _i, _i, _i, _i, ctypes.POINTER(_Settings3)]
        for f in ("orc3_solver_nlevels", "orc3_solver_vcycle", "orc3_solver_solve",
                  "orc3_solver_destroy", "orc3_solver_level_dims", "orc3_solver_get_level",
                  "orc3_solver_relax_planes"):
            getattr(L, f).argtypes = None
        _init = True
    return L


def coarse_dims3(ii, jj, kk):
    """Ghost-inclusive coarse extents: n_c = (n - 1) / 2 + 1 (include/cedar/3d/solver.h:94-96)."""
    return tuple(int((n - 3) / 2.0 + 1) + 2 for n in (ii, jj, kk))


def residual(so, qf, q):
    kk, jj, ii = q.shape
    r = np.zeros_like(q)
    lib().orc3_residual(_p(so), _i(so.shape[0]), _p(qf), _p(q), _p(r), _i(ii), _i(jj), _i(kk))
    return r


def norm(x):
    kk, jj, ii = x.shape
    return lib().orc3_norm(_p(x), _i(ii), _i(jj), _i(kk))


def matvec(so, q, y=None):
    """y = A q over the interior; ghosts of y kept (zero when y is not given)."""
    kk, jj, ii = q.shape
    y = np.zeros_like(q) if y is None else y.copy()
    lib().orc3_matvec(_p(so), _i(so.shape[0]), _p(q), _p(y), _i(ii), _i(jj), _i(kk))
    return y


def setup_recip(so):
    _, kk, jj, ii = so.shape
    sor = np.zeros((2, kk, jj, ii))
    lib().orc3_setup_recip(_p(so), _p(sor), _i(ii), _i(jj), _i(kk))
    return sor


def relax_gs(so, qf, q, sor, updown):
    kk, jj, ii = q.shape
    q = q.copy()
    lib().orc3_relax_gs(_p(so), _i(so.shape[0]), _p(qf), _p(q), _p(sor), _i(ii), _i(jj), _i(kk),
                        _i(updown))
    return q


def restrict(ci, q):
    kkf, jjf, iif = q.shape
    _, kkc, jjc, iic = ci.shape
    qc = np.zeros((kkc, jjc, iic))
    lib().orc3_restrict(_p(ci), _p(q), _p(qc), _i(iif), _i(jjf), _i(kkf), _i(iic), _i(jjc),
                        _i(kkc))
    return qc


def interp_add(q, qc, res, so, ci):
    """Returns (q + P qc + res/SO_p, res divided in place as the reference does)."""
    kkf, jjf, iif = q.shape
    kkc, jjc, iic = qc.shape
    q = q.copy()
    res = res.copy()
    lib().orc3_interp_add(_p(q), _p(qc), _p(res), _p(so), _i(so.shape[0]), _p(ci), _i(iic),
                          _i(jjc), _i(kkc), _i(iif), _i(jjf), _i(kkf))
    return q, res


def setup_interp(so, cdims=None):
    ns, kkf, jjf, iif = so.shape
    iic, jjc, kkc = cdims or coarse_dims3(iif, jjf, kkf)
    ci = np.zeros((26, kkc, jjc, iic))
    lib().orc3_setup_interp(_p(so), _i(ns), _p(ci), _i(iif), _i(jjf), _i(kkf), _i(iic), _i(jjc),
                            _i(kkc))
    return ci


def setup_galerkin(so, ci):
    ns, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    soc = np.zeros((14, kkc, jjc, iic))
    lib().orc3_setup_galerkin(_p(so), _i(ns), _p(ci), _p(soc), _i(iif), _i(jjf), _i(kkf),
                              _i(iic), _i(jjc), _i(kkc))
    return soc


def cg_bandwidth(ii, jj):
    return lib().orc3_cg_bandwidth(_i(ii), _i(jj))


def setup_cg(so):
    """Returns (band factor (n, ldab) C-order = ABD(ldab, n), info)."""
    ns, kk, jj, ii = so.shape
    kd = cg_bandwidth(ii, jj)
    n = (ii - 2) * (jj - 2) * (kk - 2)
    abd = np.zeros((n, kd + 1))
    info = lib().orc3_setup_cg(_p(so), _i(ns), _p(abd), _i(kd + 1), _i(ii), _i(jj), _i(kk))
    return abd, info


def solve_cg(abd, qf):
    kk, jj, ii = qf.shape
    q = np.zeros_like(qf)
    bbd = np.zeros(abd.shape[0])
    lib().orc3_solve_cg(_p(q), _p(qf), _p(abd), _i(abd.shape[1]), _p(bbd), _i(ii), _i(jj), _i(kk))
    return q


# ---- periodic branches (bmg3_periodic.c); jpn is the BoxMG boundary id of the mask ----
def bc_from_mask(mask):
    """BMG_get_bc for the 3D periodic mask (bits x = 1, y = 2, z = 4)."""
    return lib().orc3_bc_from_mask(_i(mask))


def images(q, jpn):
    kk, jj, ii = q.shape
    q = q.copy()
    lib().orc3_images(_p(q), _i(ii), _i(jj), _i(kk), _i(jpn))
    return q


def relax_gs_bc(so, qf, q, sor, updown, jpn):
    kk, jj, ii = q.shape
    q = q.copy()
    lib().orc3_relax_gs_bc(_p(so), _i(so.shape[0]), _p(qf), _p(q), _p(sor), _i(ii), _i(jj),
                           _i(kk), _i(updown), _i(jpn))
    return q


def restrict_bc(ci, q, jpn):
    """Returns (qc, q with its images refreshed, as the reference leaves it)."""
    kkf, jjf, iif = q.shape
    _, kkc, jjc, iic = ci.shape
    q = q.copy()
    qc = np.zeros((kkc, jjc, iic))
    lib().orc3_restrict_bc(_p(ci), _p(q), _p(qc), _i(iif), _i(jjf), _i(kkf), _i(iic), _i(jjc),
                           _i(kkc), _i(jpn))
    return qc, q


def interp_add_bc(q, qc, res, so, ci, jpn):
    kkf, jjf, iif = q.shape
    kkc, jjc, iic = qc.shape
    q = q.copy()
    res = res.copy()
    lib().orc3_interp_add_bc(_p(q), _p(qc), _p(res), _p(so), _i(so.shape[0]), _p(ci), _i(iic),
                             _i(jjc), _i(kkc), _i(iif), _i(jjf), _i(kkf), _i(jpn))
    return q, res


def setup_interp_bc(so, jpn, cdims=None):
    ns, kkf, jjf, iif = so.shape
    iic, jjc, kkc = cdims or coarse_dims3(iif, jjf, kkf)
    ci = np.zeros((26, kkc, jjc, iic))
    lib().orc3_setup_interp_bc(_p(so), _i(ns), _p(ci), _i(iif), _i(jjf), _i(kkf), _i(iic),
                               _i(jjc), _i(kkc), _i(jpn))
    return ci


def setup_galerkin_bc(so, ci, jpn):
    ns, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    soc = np.zeros((14, kkc, jjc, iic))
    lib().orc3_setup_galerkin_bc(_p(so), _i(ns), _p(ci), _p(soc), _i(iif), _i(jjf), _i(kkf),
                                 _i(iic), _i(jjc), _i(kkc), _i(jpn))
    return soc


def assemble_cg_periodic(so, jpn):
    """Dense n x n periodic matrix, upper triangle (A[r, c] for r <= c), C order [c, r]."""
    ns, kk, jj, ii = so.shape
    n = (ii - 2) * (jj - 2) * (kk - 2)
    a = np.zeros((n, n))
    lib().orc3_assemble_cg_periodic(_p(so), _i(ns), _p(a), _i(ii), _i(jj), _i(kk), _i(jpn))
    return a


def setup_cg_bc(so, jpn):
    """Periodic: (dense upper Cholesky factor, C order [c, r] = ABD(r, c), info)."""
    ns, kk, jj, ii = so.shape
    n = (ii - 2) * (jj - 2) * (kk - 2)
    abd = np.zeros((n, n))
    info = lib().orc3_setup_cg_bc(_p(so), _i(ns), _p(abd), _i(n), _i(ii), _i(jj), _i(kk),
                                  _i(jpn))
    return abd, info


def solve_cg_bc(abd, qf, jpn):
    kk, jj, ii = qf.shape
    q = np.zeros_like(qf)
    bbd = np.zeros(abd.shape[0])
    lib().orc3_solve_cg_bc(_p(q), _p(qf), _p(abd), _i(abd.shape[1]), _p(bbd), _i(ii), _i(jj),
                           _i(kk), _i(jpn))
    return q


def plane_dims(direction, ii, jj, kk):
    """(ii2, jj2, number of planes) of the planes of a direction."""
    d = PLANE[direction]
    return [(ii, jj, kk - 2), (ii, kk, jj - 2), (jj, kk, ii - 2)][d]


def copy_coeff(so, direction, ipl):
    ns, kk, jj, ii = so.shape
    ii2, jj2, _ = plane_dims(direction, ii, jj, kk)
    so2 = np.zeros((5 if ns == 14 else 3, jj2, ii2))
    lib().orc3_copy_coeff(_i(PLANE[direction]), _p(so), _i(ns), _p(so2), _i(ipl), _i(ii), _i(jj),
                          _i(kk))
    return so2


def copy_rhs(so, x, b, direction, ipl):
    ns, kk, jj, ii = so.shape
    ii2, jj2, _ = plane_dims(direction, ii, jj, kk)
    b2 = np.zeros((jj2, ii2))
    lib().orc3_copy_rhs(_i(PLANE[direction]), _p(so), _i(ns), _p(x), _p(b), _p(b2), _i(ipl),
                        _i(ii), _i(jj), _i(kk))
    return b2


def copy32(x, direction, ipl):
    kk, jj, ii = x.shape
    ii2, jj2, _ = plane_dims(direction, ii, jj, kk)
    x2 = np.zeros((jj2, ii2))
    lib().orc3_copy32(_i(PLANE[direction]), _p(x), _p(x2), _i(ipl), _i(ii), _i(jj), _i(kk))
    return x2


def copy23(x2, x, direction, ipl):
    kk, jj, ii = x.shape
    x = x.copy()
    lib().orc3_copy23(_i(PLANE[direction]), _p(x2), _p(x), _i(ipl), _i(ii), _i(jj), _i(kk))
    return x


def num_levels(nx, ny, nz, min_coarse=3):
    return lib().orc3_num_levels(_i(nx), _i(ny), _i(nz), _i(min_coarse))


def plane_settings(relaxation="line-xy", nrelax_pre=2, nrelax_post=1, maxiter=1, tol=1e-8):
    """Cedar's default plane configuration (src/kernel_params.cc:83-89)."""
    return _o._Settings(_o.RELAX[relaxation], nrelax_pre, nrelax_post, -1, maxiter, tol, 3)


class Solver3:
    """3D FP64 oracle solver: cdr3::solver<fsten>(so, conf) + solve(b, x)."""

    def __init__(self, so, relaxation="point", nrelax_pre=2, nrelax_post=1, num_levels=-1,
                 maxiter=10, tol=1e-8, min_coarse=3, plane=None, cycle="v", periodic=0):
        """periodic: per_mask (bits x = 1, y = 2, z = 4), point relaxation only."""
        self.so = np.ascontiguousarray(so, dtype=np.float64)
        ns, kk, jj, ii = self.so.shape
        self.shape = (kk, jj, ii)
        if periodic and relaxation != "point":
            raise ValueError("periodic 3D problems take point relaxation")
        s = _Settings3(RELAX3[relaxation], nrelax_pre, nrelax_post, num_levels, maxiter, tol,
                       min_coarse, plane or plane_settings(), {"v": 0, "f": 1}[cycle],
                       int(periodic))
        self.maxiter = maxiter
        self._h = ctypes.c_void_p(lib().orc3_solver_create(_p(self.so), _i(ns), _i(ii - 2),
                                                           _i(jj - 2), _i(kk - 2),
                                                           ctypes.byref(s)))

    @property
    def nlevels(self):
        return lib().orc3_solver_nlevels(self._h)

    def level_dims(self, lvl):
        n = [ctypes.c_int() for _ in range(3)]
        lib().orc3_solver_level_dims(self._h, _i(lvl), *[ctypes.byref(v) for v in n])
        return tuple(v.value for v in n)

    def level(self, lvl):
        nx, ny, nz = self.level_dims(lvl)
        shp = (nz + 2, ny + 2, nx + 2)
        ns = self.so.shape[0] if lvl == 0 else 14
        so = np.zeros((ns,) + shp)
        ci = np.zeros((26,) + shp) if lvl > 0 else None
        lib().orc3_solver_get_level(self._h, _i(lvl), _p(so), _p(ci))
        return so, ci

    def vcycle(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().orc3_solver_vcycle(self._h, _p(x), _p(b))
        return x

    def relax_planes(self, lvl, direction, x, b, updown):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().orc3_solver_relax_planes(self._h, _i(lvl), _i(PLANE[direction]), _p(x), _p(b),
                                       _i(updown))
        return x

    def solve(self, b, x=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        hist = np.zeros(max(self.maxiter, 1))
        n = lib().orc3_solver_solve(self._h, _p(b), _p(x), _p(hist), _i(len(hist)))
        return x, hist[:n].copy()

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc3_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

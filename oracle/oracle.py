"""ctypes bindings for the FP64 CPU oracle and for the reference kernels built from source.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product (cedar_amd) never imports this module.

Array convention: a grid function with ghost-inclusive extents (ii, jj) is a C-contiguous numpy
array of shape (jj, ii) -- the same bytes as the reference's Fortran array Q(ii, jj).  A stencil
is (nstncl, jj, ii), an interpolation operator (8, jjc, iic), a relaxation stencil (2, jj, ii).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcedar_ref.so")

DOWN, UP = 0, 1
RELAX = {"point": 0, "line-x": 1, "line-y": 2, "line-xy": 3}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class _Settings(ctypes.Structure):
    _fields_ = [("relax", _i), ("nrelax_pre", _i), ("nrelax_post", _i), ("num_levels", _i),
                ("maxiter", _i), ("tol", ctypes.c_double), ("min_coarse", _i), ("cycle", _i),
                ("periodic", _i)]


def bc_from_mask(mask):
    """BMG_get_bc (src/2d/ftn/BMG_get_bc.f90): per_mask bit 0 = x, bit 1 = y -> BoxMG BC id."""
    return (0, 2, 1, 3)[mask]


_lib = None


def lib():
    """Load the oracle shared library (built by `make -C oracle`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        L = ctypes.CDLL(ORACLE_SO)
        L.orc2_norm.restype = ctypes.c_double
        L.orc2_solver_create.restype = ctypes.c_void_p
        L.orc2_solver_create.argtypes = [_dp, _i, _i, _i, ctypes.POINTER(_Settings)]
        for f in ("orc2_solver_nlevels", "orc2_solver_vcycle", "orc2_solver_solve",
                  "orc2_solver_destroy", "orc2_solver_level_dims", "orc2_solver_get_level",
                  "orc2_solver_abd"):
            getattr(L, f).argtypes = None
        L.orc2_solver_abd.restype = _dp
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


_ref = None


def ref():
    """The reference's own serial kernels compiled from their sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference kernels not built: `make -C oracle ref`")
        _ref = ctypes.CDLL(REF_SO)
    return _ref


# ---------------------------------------------------------------------------------------------
# oracle kernels (numpy in / numpy out)
# ---------------------------------------------------------------------------------------------
def residual(so, qf, q):
    jj, ii = q.shape
    r = np.zeros_like(q)
    lib().orc2_residual(_ptr(so), _i(so.shape[0]), _ptr(qf), _ptr(q), _ptr(r), _i(ii), _i(jj))
    return r


def norm(x):
    jj, ii = x.shape
    return lib().orc2_norm(_ptr(x), _i(ii), _i(jj))


def matvec(so, q, y=None):
    """y = A q over the interior; ghosts of y kept (zero when y is not given)."""
    jj, ii = q.shape
    y = np.zeros_like(q) if y is None else y.copy()
    lib().orc2_matvec(_ptr(so), _i(so.shape[0]), _ptr(q), _ptr(y), _i(ii), _i(jj))
    return y


def setup_recip(so):
    _, jj, ii = so.shape
    sor = np.zeros((2, jj, ii))
    lib().orc2_setup_recip(_ptr(so), _ptr(sor), _i(ii), _i(jj))
    return sor


def relax_gs(so, qf, q, sor, updown, jpn=0):
    jj, ii = q.shape
    q = q.copy()
    lib().orc2_relax_gs_bc(_ptr(so), _i(so.shape[0]), _ptr(qf), _ptr(q), _ptr(sor), _i(ii),
                           _i(jj), _i(updown), _i(jpn))
    return q


def setup_lines(so, direction):
    _, jj, ii = so.shape
    sor = np.zeros((2, jj, ii))
    fn = lib().orc2_setup_lines_x if direction == "x" else lib().orc2_setup_lines_y
    info = fn(_ptr(so), _ptr(sor), _i(ii), _i(jj))
    return sor, info


def setup_lines_periodic(so, direction, jpn):
    _, jj, ii = so.shape
    sor = np.zeros((2, jj, ii))
    fn = lib().orc2_setup_lines_x_bc if direction == "x" else lib().orc2_setup_lines_y_bc
    info = fn(_ptr(so), _ptr(sor), _i(ii), _i(jj), _i(jpn))
    return sor, info


def relax_lines_periodic(so, qf, q, sor, updown, direction, jpn):
    jj, ii = q.shape
    q = q.copy()
    b = np.zeros(2 * max(ii, jj))
    fn = lib().orc2_relax_lines_x_bc if direction == "x" else lib().orc2_relax_lines_y_bc
    fn(_ptr(so), _i(so.shape[0]), _ptr(qf), _ptr(q), _ptr(sor), _ptr(b), _i(ii), _i(jj),
       _i(updown), _i(jpn))
    return q


def relax_lines(so, qf, q, sor, updown, direction):
    jj, ii = q.shape
    q = q.copy()
    if direction == "x":
        lib().orc2_relax_lines_x(_ptr(so), _i(so.shape[0]), _ptr(qf), _ptr(q), _ptr(sor),
                                 _i(ii), _i(jj), _i(updown))
    else:
        b = np.zeros(jj)
        lib().orc2_relax_lines_y(_ptr(so), _i(so.shape[0]), _ptr(qf), _ptr(q), _ptr(sor), _ptr(b),
                                 _i(ii), _i(jj), _i(updown))
    return q


def restrict(ci, q, jpn=0):
    """P^T q; with periodic jpn the ghosts of q are refreshed first (on a copy)."""
    jjf, iif = q.shape
    _, jjc, iic = ci.shape
    qc = np.zeros((jjc, iic))
    q = np.ascontiguousarray(q, dtype=np.float64).copy()
    lib().orc2_restrict_bc(_ptr(ci), _ptr(q), _ptr(qc), _i(iif), _i(jjf), _i(iic), _i(jjc),
                           _i(jpn))
    return qc


def interp_add(q, qc, res, so, ci, jpn=0):
    """Returns (q + P qc + affine term, res divided by SO_c in place as the reference does)."""
    jjf, iif = q.shape
    jjc, iic = qc.shape
    q = q.copy()
    res = res.copy()
    lib().orc2_interp_add_bc(_ptr(q), _ptr(qc), _ptr(res), _ptr(so), _i(so.shape[0]), _ptr(ci),
                             _i(iic), _i(jjc), _i(iif), _i(jjf), _i(jpn))
    return q, res


def coarse_dims(ii, jj):
    nx, ny = ii - 2, jj - 2
    return int((nx - 1) / 2.0 + 1) + 2, int((ny - 1) / 2.0 + 1) + 2


def setup_interp(so, iic=None, jjc=None, jpn=0):
    ns, jjf, iif = so.shape
    if iic is None:
        iic, jjc = coarse_dims(iif, jjf)
    ci = np.zeros((8, jjc, iic))
    lib().orc2_setup_interp_bc(_ptr(so), _i(ns), _ptr(ci), _i(iif), _i(jjf), _i(iic), _i(jjc),
                               _i(jpn))
    return ci


def setup_galerkin(so, ci, jpn=0):
    ns, jjf, iif = so.shape
    _, jjc, iic = ci.shape
    soc = np.zeros((5, jjc, iic))
    lib().orc2_setup_galerkin_bc(_ptr(so), _i(ns), _ptr(ci), _ptr(soc), _i(iif), _i(jjf),
                                 _i(iic), _i(jjc), _i(jpn))
    return soc


def setup_cg(so):
    ns, jj, ii = so.shape
    n = (ii - 2) * (jj - 2)
    abd = np.zeros((n, ii))  # ABD(ii, n) in Fortran order
    info = lib().orc2_setup_cg(_ptr(so), _i(ns), _ptr(abd), _i(ii), _i(jj))
    return abd, info


def solve_cg(abd, qf):
    jj, ii = qf.shape
    q = np.zeros_like(qf)
    bbd = np.zeros(abd.shape[0])
    lib().orc2_solve_cg(_ptr(q), _ptr(qf), _ptr(abd), _ptr(bbd), _i(ii), _i(jj))
    return q


def assemble_cg_periodic(so, jpn):
    """The dense upper triangle the periodic coarse setup factors: (n, n) C array == A(n, n)."""
    ns, jj, ii = so.shape
    n = (ii - 2) * (jj - 2)
    abd = np.zeros((n, n))
    lib().orc2_assemble_cg_periodic(_ptr(so), _i(ns), _ptr(abd), _i(ii), _i(jj), _i(jpn))
    return abd


def setup_cg_periodic(so, jpn):
    ns, jj, ii = so.shape
    n = (ii - 2) * (jj - 2)
    abd = np.zeros((n, n))
    info = lib().orc2_setup_cg_bc(_ptr(so), _i(ns), _ptr(abd), _i(ii), _i(jj), _i(jpn))
    return abd, info


def solve_cg_periodic(abd, qf, jpn):
    jj, ii = qf.shape
    q = np.zeros_like(qf)
    bbd = np.zeros(abd.shape[0])
    lib().orc2_solve_cg_bc(_ptr(q), _ptr(np.ascontiguousarray(qf)), _ptr(abd), _ptr(bbd), _i(ii),
                           _i(jj), _i(jpn))
    return q


def dpotrf_u(a):
    """Upper Cholesky of a symmetric (n, n) array (upper triangle read, Fortran order = a.T)."""
    a = np.array(a, dtype=np.float64)
    n = a.shape[0]
    info = lib().orc_dpotrf_u(_i(n), _ptr(a), _i(n))
    return a, info


def dpttrf(d, e):
    d = np.array(d, dtype=np.float64)
    e = np.array(e, dtype=np.float64)
    info = lib().orc_dpttrf(_i(len(d)), _ptr(d), _ptr(e))
    return d, e, info


def dpttrs(d, e, b):
    b = np.array(b, dtype=np.float64)
    lib().orc_dpttrs(_i(len(d)), _ptr(np.ascontiguousarray(d, dtype=np.float64)),
                     _ptr(np.ascontiguousarray(e, dtype=np.float64)), _ptr(b))
    return b


def dpbtrf_u(ab):
    """ab: (n, ldab) C array == Fortran AB(ldab, n); kd = ldab - 1."""
    ab = np.array(ab, dtype=np.float64)
    n, ldab = ab.shape
    info = lib().orc_dpbtrf_u(_i(n), _i(ldab - 1), _ptr(ab), _i(ldab))
    return ab, info


def dpbtrs_u(ab, b):
    n, ldab = ab.shape
    b = np.array(b, dtype=np.float64)
    lib().orc_dpbtrs_u(_i(n), _i(ldab - 1), _ptr(np.ascontiguousarray(ab)), _i(ldab), _ptr(b))
    return b


# ---------------------------------------------------------------------------------------------
# multilevel solver
# ---------------------------------------------------------------------------------------------
def num_levels(nx, ny, min_coarse=3):
    return lib().orc2_num_levels(_i(nx), _i(ny), _i(min_coarse))


class Solver:
    """cdr2::solver<fsten>(so, conf): setup in the constructor, solve(b) from x0 = 0."""

    def __init__(self, so, relaxation="point", nrelax_pre=2, nrelax_post=1, num_levels=-1,
                 maxiter=10, tol=1e-8, min_coarse=3, cycle="v", periodic=(False, False)):
        self.so = np.ascontiguousarray(so, dtype=np.float64)  # held for the solver's lifetime
        ns, jj, ii = self.so.shape
        self.ii, self.jj = ii, jj
        mask = int(bool(periodic[0])) | (int(bool(periodic[1])) << 1)
        s = _Settings(RELAX[relaxation], nrelax_pre, nrelax_post, num_levels, maxiter, tol,
                      min_coarse, {"v": 0, "f": 1}[cycle], mask)
        L = lib()
        self._h = ctypes.c_void_p(L.orc2_solver_create(_ptr(self.so), ns, ii - 2, jj - 2,
                                                       ctypes.byref(s)))
        self.maxiter = maxiter

    @property
    def nlevels(self):
        return lib().orc2_solver_nlevels(self._h)

    def level_dims(self, lvl):
        nx, ny = _i(), _i()
        lib().orc2_solver_level_dims(self._h, _i(lvl), ctypes.byref(nx), ctypes.byref(ny))
        return nx.value, ny.value

    def level(self, lvl):
        """(SO, CI, SOR0, SOR1) of a level as built by setup."""
        nx, ny = self.level_dims(lvl)
        ii, jj = nx + 2, ny + 2
        ns = self.so.shape[0] if lvl == 0 else 5
        so = np.zeros((ns, jj, ii))
        ci = np.zeros((8, jj, ii))
        s0 = np.zeros((2, jj, ii))
        s1 = np.zeros((2, jj, ii))
        lib().orc2_solver_get_level(self._h, _i(lvl), _ptr(so), _ptr(ci), _ptr(s0), _ptr(s1))
        return so, (ci if lvl > 0 else None), s0, s1

    def abd(self):
        n1, n2 = _i(), _i()
        p = lib().orc2_solver_abd(self._h, ctypes.byref(n1), ctypes.byref(n2))
        return np.ctypeslib.as_array(p, shape=(n2.value, n1.value)).copy()

    def vcycle(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc2_solver_vcycle(self._h, _ptr(x), _ptr(np.ascontiguousarray(b)))
        return x

    def solve(self, b, x=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        hist = np.zeros(max(self.maxiter, 1))
        it = lib().orc2_solver_solve(self._h, _ptr(b), _ptr(x), _ptr(hist), _i(len(hist)))
        return x, hist[:it].copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.orc2_solver_destroy(h)
            self._h = None

"""Python face of the 3D path: Cedar's cdr3 kernels and solver on device tensors.

A grid function with nx*ny*nz interior points is a (nz+2, ny+2, nx+2) tensor (the bytes of the
Fortran array Q(nx+2, ny+2, nz+2)); a stencil is (4 | 14, nz+2, ny+2, nx+2), an interpolation
operator (26, nzc+2, nyc+2, nxc+2).  Every function is one C-ABI call on torch's current stream.
`Solver3` mirrors cdr3::solver<fsten> (include/cedar/3d/solver.h:46-57): setup in the
constructor, `solve(b)` from x0 = 0 (include/cedar/multilevel.h:268-298).
"""
import ctypes

import numpy as np
import torch

from . import _abi
from .cdr2 import _memcpy_d2h, _p, _stream

DOWN, UP = _abi.CDR_DOWN, _abi.CDR_UP


def _dims(q):
    kk, jj, ii = q.shape
    return ii, jj, kk


def _sfx(t):
    return {torch.float64: "f64", torch.float32: "f32"}[t.dtype]


def coarse_dims3(ii, jj, kk):
    """Ghost-inclusive coarse extents (include/cedar/3d/solver.h:94-96)."""
    return tuple(int((n - 3) / 2.0 + 1) + 2 for n in (ii, jj, kk))


def residual(so, qf, q, res):
    _abi.call(f"cdr_hip_residual3d_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q), _p(res),
              *_dims(q), _stream())
    return res


def residual_norm(so, qf, q, res):
    """Residual into res; returns ||r||_2 over the interior (synchronises)."""
    ii, jj, kk = _dims(q)
    work = torch.empty(max(_abi.load().cdr_hip_norm_work3d(ii, jj, kk), 1), dtype=torch.float64,
                       device=q.device)
    out = torch.empty(1, dtype=torch.float64, device=q.device)
    _abi.call("cdr_hip_residual_norm3d_f64", _p(so), so.shape[0], _p(qf), _p(q), _p(res), ii, jj,
              kk, _p(work), _p(out), _stream())
    return float(out.sqrt().item())


def setup_recip(so, sor):
    _abi.call("cdr_hip_setup_recip3d_f64", _p(so), _p(sor), so.shape[3], so.shape[2],
              so.shape[1], so.shape[0], _stream())
    return sor


def relax_gs(so, qf, q, sor, updown):
    _abi.call(f"cdr_hip_relax_gs3d_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q), _p(sor),
              *_dims(q), int(updown), _stream())
    return q


def restrict(ci, q, qc):
    _abi.call(f"cdr_hip_restrict3d_{_sfx(q)}", _p(ci), _p(q), _p(qc), *_dims(q), *_dims(qc),
              _stream())
    return qc


def interp_add(q, qc, res, so, ci):
    _abi.call(f"cdr_hip_interp_add3d_{_sfx(q)}", _p(q), _p(qc), _p(res), _p(so), so.shape[0],
              _p(ci), *_dims(qc), *_dims(q), _stream())
    return q


def setup_interp(so, ci):
    _, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    _abi.call("cdr_hip_setup_interp3d_f64", _p(so), so.shape[0], _p(ci), iif, jjf, kkf, iic, jjc,
              kkc, _stream())
    return ci


def setup_galerkin(so, ci, soc):
    _, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    _abi.call("cdr_hip_setup_galerkin3d_f64", _p(so), so.shape[0], _p(ci), _p(soc), iif, jjf, kkf,
              iic, jjc, kkc, _stream())
    return soc


def cg_bandwidth(ii, jj):
    return _abi.load().cdr_hip_cg_bandwidth3d(ii, jj)


def setup_cg(so):
    """Band factor (n, kd+1) tensor of the coarsest operator (ABD(kd+1, n) in Fortran order)."""
    _, kk, jj, ii = so.shape
    kd = cg_bandwidth(ii, jj)
    n = (ii - 2) * (jj - 2) * (kk - 2)
    abd = torch.zeros((n, kd + 1), dtype=torch.float64, device=so.device)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call("cdr_hip_setup_cg3d_f64", _p(so), so.shape[0], _p(abd), kd + 1, ii, jj, kk,
              _p(info), _stream())
    return abd


def solve_cg(q, qf, abd):
    """Coarsest solve with the FP64 band abd; q and qf FP64 or FP32 (then solved in FP64)."""
    bbd = torch.zeros(abd.shape[0], dtype=torch.float64, device=q.device)
    _abi.call(f"cdr_hip_solve_cg3d_{_sfx(q)}", _p(q), _p(qf), _p(abd), abd.shape[1], _p(bbd),
              *_dims(q), _stream())
    return q


# ---- periodic boundaries (the reference's JPN != 0 branches) --------------------------------
def bc_from_mask(mask):
    """BMG_get_bc for the 3D periodic mask (bits x = 1, y = 2, z = 4): 2, 1, 3, 5, 6, 7, 8."""
    return (0, 2, 1, 3, 5, 6, 7, 8)[int(mask)]


def relax_gs_per(so, qf, q, sor, updown, jpn):
    _abi.call(f"cdr_hip_relax_gs3d_per_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q), _p(sor),
              *_dims(q), int(updown), int(jpn), _stream())
    return q


def restrict_per(ci, q, qc, jpn):
    """Restriction; refreshes the images of q in place first."""
    _abi.call(f"cdr_hip_restrict3d_per_{_sfx(q)}", _p(ci), _p(q), _p(qc), *_dims(q), *_dims(qc),
              int(jpn), _stream())
    return qc


def interp_add_per(q, qc, res, so, ci, jpn):
    _abi.call(f"cdr_hip_interp_add3d_per_{_sfx(q)}", _p(q), _p(qc), _p(res), _p(so), so.shape[0],
              _p(ci), *_dims(qc), *_dims(q), int(jpn), _stream())
    return q


def setup_interp_per(so, ci, jpn):
    """ci must be zero on entry."""
    _, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    _abi.call("cdr_hip_setup_interp3d_per_f64", _p(so), so.shape[0], _p(ci), iif, jjf, kkf, iic,
              jjc, kkc, int(jpn), _stream())
    return ci


def setup_galerkin_per(so, ci, soc, jpn):
    _, kkf, jjf, iif = so.shape
    _, kkc, jjc, iic = ci.shape
    _abi.call("cdr_hip_setup_galerkin3d_per_f64", _p(so), so.shape[0], _p(ci), _p(soc), iif, jjf,
              kkf, iic, jjc, kkc, int(jpn), _stream())
    return soc


def setup_cg_per(so, jpn):
    """Dense upper Cholesky factor (n, n) tensor, [c, r] = ABD(r, c) (synchronises)."""
    _, kk, jj, ii = so.shape
    n = (ii - 2) * (jj - 2) * (kk - 2)
    abd = torch.zeros((n, n), dtype=torch.float64, device=so.device)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call("cdr_hip_setup_cg3d_per_f64", _p(so), so.shape[0], _p(abd), ii, jj, kk, int(jpn),
              _p(info), _stream())
    return abd


def solve_cg_per(q, qf, abd, jpn):
    bbd = torch.zeros(abd.shape[0], dtype=torch.float64, device=q.device)
    _abi.call(f"cdr_hip_solve_cg3d_per_{_sfx(q)}", _p(q), _p(qf), _p(abd), _p(bbd), *_dims(q),
              int(jpn), _stream())
    return q


def _registry(fn, h):
    import json
    n = fn(h, None, 0)
    if n < 0:
        raise _abi.CedarHipError(-n, "kernel registry")
    buf = ctypes.create_string_buffer(n + 1)
    fn(h, buf, n + 1)
    return json.loads(buf.value.decode())


def plane_settings(relaxation="line-xy", nrelax_pre=2, nrelax_post=1, maxiter=1, tol=1e-8,
                   min_coarse=3, num_levels=-1, fused_coarse=True):
    """plane-config of the nested 2D plane solves (default: src/kernel_params.cc:83-89)."""
    return _abi.Settings(_abi.RELAX[relaxation], nrelax_pre, nrelax_post, num_levels, maxiter, tol,
                         min_coarse, 0, 0, 0, int(bool(fused_coarse)))


class Solver3:
    """cdr3::solver<fsten>(so, conf) on the GPU; precision 'f64' (Cedar's algorithm) or 'mixed'
    (FP32 3D cycle with FP32 plane solves inside FP64 iterative refinement).

    Settings use Cedar's keys and defaults: relaxation ('point' | 'plane-xy' | 'plane-xz' |
    'plane-yz' | 'plane-xyz'), nrelax_pre=2, nrelax_post=1, num_levels=-1, maxiter=10, tol=1e-8,
    min_coarse=3, plane = plane_settings() (line-xy, one V(2,1) cycle per plane solve);
    stall_cycles as for the 2D Solver.  periodic: Cedar's grid.periodic as a per_mask (bits
    x = 1, y = 2, z = 4) with point relaxation; the operator, b and x carry the periodic images
    in their ghosts, and periodic extents must be even on every coarsened level.
    """

    def __init__(self, so, relaxation="point", nrelax_pre=2, nrelax_post=1, num_levels=-1,
                 maxiter=10, tol=1e-8, min_coarse=3, plane=None, cycle="v", precision="f64",
                 stall_cycles=0, periodic=0):
        so = so if isinstance(so, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(so))
        self.so = so.to(device="cuda", dtype=torch.float64).contiguous()  # held by reference
        ns, kk, jj, ii = self.so.shape
        self.maxiter = maxiter
        s = _abi.Settings3(_abi.RELAX[relaxation], nrelax_pre, nrelax_post, num_levels, maxiter, tol,
                           min_coarse, plane or plane_settings(), _abi.CYCLE[cycle],
                           _abi.PRECISION[precision], int(stall_cycles), int(periodic))
        h = ctypes.c_void_p()
        _abi.call("cdr_hip_solver3d_create", _p(self.so), ns, ii - 2, jj - 2, kk - 2,
                  ctypes.byref(s), _stream(), ctypes.byref(h))
        self._h = h

    @property
    def nlevels(self):
        return _abi.load().cdr_hip_solver3d_nlevels(self._h)

    def level(self, lvl):
        n = [ctypes.c_int() for _ in range(3)]
        ptrs = [ctypes.c_void_p() for _ in range(2)]
        _abi.call("cdr_hip_solver3d_level", self._h, lvl, *[ctypes.byref(v) for v in n],
                  *[ctypes.byref(p) for p in ptrs])
        return tuple(v.value for v in n) + tuple(p.value for p in ptrs)

    def level_arrays(self, lvl):
        """Host copies (so, ci) of a level's operator and interpolation."""
        nx, ny, nz, so, ci = self.level(lvl)
        shp = (nz + 2, ny + 2, nx + 2)
        ns = self.so.shape[0] if lvl == 0 else 14
        torch.cuda.synchronize()

        def grab(ptr, planes):
            if not ptr:
                return None
            out = np.empty((planes,) + shp)
            _memcpy_d2h(out, ptr)
            return out
        return grab(so, ns), grab(ci, 26)

    def times(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _abi.call("cdr_hip_solver3d_times", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def vcycle(self, x, b):
        _abi.call("cdr_hip_solver3d_vcycle", self._h, _p(b), _p(x))
        return x

    def kernels(self):
        """The kernel registry: {interface: {"impls": [...], "default": name}}."""
        return _registry(_abi.load().cdr_hip_solver3d_kernels, self._h)

    def set_kernel(self, interface, impl):
        """kman->set<I>(impl) for the interface named `interface`."""
        _abi.call("cdr_hip_solver3d_set_kernel", self._h, interface.encode(), impl.encode())

    def relax_planes(self, lvl, family, x, b, updown):
        _abi.call("cdr_hip_solver3d_relax_planes", self._h, lvl, _abi.PLANE_FAMILY[family], _p(x),
                  _p(b), int(updown))
        return x

    def solve(self, b, x=None):
        """Returns (x, rho history as a numpy array)."""
        if x is None:
            x = torch.zeros_like(b)
        hist = (ctypes.c_double * max(self.maxiter, 1))()
        nit = ctypes.c_int(0)
        st = _abi.load().cdr_hip_solver3d_solve(self._h, _p(b), _p(x), hist, len(hist),
                                                ctypes.byref(nit))
        h = np.array(hist[:nit.value])
        if st != _abi.CDR_OK:
            err = _abi.CedarHipError(st, "cdr_hip_solver3d_solve")
            err.x, err.hist = x, h
            raise err
        return x, h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _abi.load().cdr_hip_solver3d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

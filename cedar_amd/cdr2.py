"""Python face of the 2D path: Cedar's cdr2 objects on device memory and the kernel calls.

`StencilOp` / `GridFunc` hold torch tensors on the GPU in the solver's layout (a grid function
with nx*ny interior points is a (ny+2, nx+2) tensor = Fortran Q(nx+2, ny+2); a stencil is
(nstncl, ny+2, nx+2)).  Every function below is one C-ABI call on torch's current stream.
`Solver` mirrors cdr2::solver<fsten>: setup in the constructor, `solve(b)` from x0 = 0
(include/cedar/2d/solver.h:55-61, include/cedar/multilevel.h:268-298).
"""
import ctypes

import numpy as np
import torch

from . import _abi

DOWN, UP = _abi.CDR_DOWN, _abi.CDR_UP


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("cedar_amd works on device tensors only")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _sfx(t):
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {t.dtype}")


class GridFunc:
    """cdr2::grid_func on the device: (ny+2, nx+2) tensor with one ghost layer."""

    def __init__(self, data):
        self.data = data

    @classmethod
    def zeros(cls, nx, ny, dtype=torch.float64, device="cuda"):
        return cls(torch.zeros((ny + 2, nx + 2), dtype=dtype, device=device))

    @classmethod
    def from_numpy(cls, a, dtype=torch.float64, device="cuda"):
        return cls(torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device))

    @property
    def ii(self):
        return self.data.shape[1]

    @property
    def jj(self):
        return self.data.shape[0]

    def numpy(self):
        return self.data.detach().cpu().numpy()


class StencilOp:
    """cdr2::stencil_op<five_pt|nine_pt> on the device: (nstncl, ny+2, nx+2)."""

    def __init__(self, data):
        if data.dim() != 3 or data.shape[0] not in (3, 5):
            raise ValueError("stencil must be (3|5, ny+2, nx+2)")
        self.data = data

    @classmethod
    def from_numpy(cls, a, dtype=torch.float64, device="cuda"):
        return cls(torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device))

    @property
    def nstncl(self):
        return self.data.shape[0]

    @property
    def ii(self):
        return self.data.shape[2]

    @property
    def jj(self):
        return self.data.shape[1]


def _t(x):
    return x.data if isinstance(x, (GridFunc, StencilOp)) else x


# ---------------------------------------------------------------------------------------------
# kernels (arrays are torch tensors or GridFunc/StencilOp)
# ---------------------------------------------------------------------------------------------
def residual(so, qf, q, res):
    so, qf, q, res = map(_t, (so, qf, q, res))
    _abi.call(f"cdr_hip_residual2d_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q), _p(res),
              q.shape[1], q.shape[0], _stream())
    return res


def residual_norm(so, qf, q, res):
    """Residual into res; returns ||r||_2 over the interior (synchronises)."""
    so, qf, q, res = map(_t, (so, qf, q, res))
    ii, jj = q.shape[1], q.shape[0]
    work = torch.empty(max(_abi.load().cdr_hip_norm_work2d(ii, jj), 1), dtype=torch.float64,
                       device=q.device)
    out = torch.empty(1, dtype=torch.float64, device=q.device)
    _abi.call("cdr_hip_residual_norm2d_f64", _p(so), so.shape[0], _p(qf), _p(q), _p(res), ii, jj,
              _p(work), _p(out), _stream())
    return float(out.sqrt().item())


def setup_recip(so, sor):
    so, sor = _t(so), _t(sor)
    _abi.call("cdr_hip_setup_recip2d_f64", _p(so), _p(sor), so.shape[2], so.shape[1],
              so.shape[0], _stream())
    return sor


def bc_from_mask(periodic):
    """BMG_get_bc (src/2d/ftn/BMG_get_bc.f90) of periodic flags (x, y): 0, 2 (x), 1 (y), 3."""
    return (0, 2, 1, 3)[int(bool(periodic[0])) | (int(bool(periodic[1])) << 1)]


def relax_gs(so, qf, q, sor, updown, jpn=0):
    so, qf, q, sor = map(_t, (so, qf, q, sor))
    if jpn:
        _abi.call(f"cdr_hip_relax_gs2d_per_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q),
                  _p(sor), q.shape[1], q.shape[0], int(updown), int(jpn), _stream())
        return q
    _abi.call(f"cdr_hip_relax_gs2d_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q), _p(sor),
              q.shape[1], q.shape[0], int(updown), _stream())
    return q


def setup_lines(so, sor, direction):
    so, sor = _t(so), _t(sor)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call(f"cdr_hip_setup_lines_{direction}2d_f64", _p(so), _p(sor), so.shape[2],
              so.shape[1], so.shape[0], _p(info), _stream())
    return sor


def setup_lines_periodic(so, sor, direction, jpn):
    so, sor = _t(so), _t(sor)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call(f"cdr_hip_setup_lines_{direction}2d_per_f64", _p(so), _p(sor), so.shape[2],
              so.shape[1], so.shape[0], int(jpn), _p(info), _stream())
    return sor


def relax_lines_periodic(so, qf, q, sor, updown, direction, jpn):
    """One periodic zebra line sweep (work: an internal scratch grid)."""
    so, qf, q, sor = map(_t, (so, qf, q, sor))
    work = torch.zeros_like(q)
    _abi.call(f"cdr_hip_relax_lines_{direction}2d_per_{_sfx(q)}", _p(so), so.shape[0], _p(qf),
              _p(q), _p(sor), _p(work), q.shape[1], q.shape[0], int(updown), int(jpn), _stream())
    return q


def relax_lines(so, qf, q, sor, updown, direction):
    so, qf, q, sor = map(_t, (so, qf, q, sor))
    _abi.call(f"cdr_hip_relax_lines_{direction}2d_{_sfx(q)}", _p(so), so.shape[0], _p(qf), _p(q),
              _p(sor), None, q.shape[1], q.shape[0], int(updown), _stream())
    return q


def relax_lines_batched(so, qf, q, sor, updown, direction):
    """One zebra line sweep over nb problems of the same extents, stacked along a leading batch
    dimension: so (nb, nstncl, jj, ii), qf and q (nb, jj, ii), sor (nb, 2, ...)."""
    so, qf, q, sor = map(_t, (so, qf, q, sor))
    _abi.call(f"cdr_hip_relax_lines2d_batched_{_sfx(q)}", 0 if direction == "x" else 1, _p(so),
              so.shape[1], _p(qf), _p(q), _p(sor), q.shape[-1], q.shape[-2], q.shape[0],
              int(updown), _stream())
    return q


def relax_lines_xy(so, qf, q, sorx, sory, updown):
    """One line-xy smoothing step: x then y going down, y then x going up."""
    so, qf, q, sorx, sory = map(_t, (so, qf, q, sorx, sory))
    _abi.call(f"cdr_hip_relax_lines_xy2d_{_sfx(q)}", _p(so), so.shape[-3], _p(qf), _p(q), _p(sorx),
              _p(sory), None, q.shape[-1], q.shape[-2], int(updown), _stream())
    return q


def autotune(on):
    """Turns line-kernel shape tuning on or off (process-wide); returns the previous setting."""
    return bool(_abi.load().cdr_hip_autotune(1 if on else 0))


def autotune_report():
    """The recorded line-kernel shape choices (list of dicts)."""
    import json
    n = _abi.load().cdr_hip_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    _abi.load().cdr_hip_autotune_report(buf, n + 1)
    return json.loads(buf.value.decode())


def restrict(ci, q, qc, jpn=0):
    """qc = P^T q; with a periodic jpn the images in the ghosts of q are refreshed in place."""
    ci, q, qc = map(_t, (ci, q, qc))
    if jpn:
        _abi.call(f"cdr_hip_restrict2d_per_{_sfx(q)}", _p(ci), _p(q), _p(qc), q.shape[1],
                  q.shape[0], qc.shape[1], qc.shape[0], int(jpn), _stream())
        return qc
    _abi.call(f"cdr_hip_restrict2d_{_sfx(q)}", _p(ci), _p(q), _p(qc), q.shape[1], q.shape[0],
              qc.shape[1], qc.shape[0], _stream())
    return qc


def interp_add(q, qc, res, so, ci, jpn=0):
    q, qc, res, so, ci = map(_t, (q, qc, res, so, ci))
    name = "interp_add2d_per" if jpn else "interp_add2d"
    extra = (int(jpn),) if jpn else ()
    _abi.call(f"cdr_hip_{name}_{_sfx(q)}", _p(q), _p(qc), _p(res), _p(so), so.shape[0],
              _p(ci), qc.shape[1], qc.shape[0], q.shape[1], q.shape[0], *extra, _stream())
    return q


def setup_interp(so, ci, jpn=0):
    so, ci = _t(so), _t(ci)
    if jpn:
        _abi.call("cdr_hip_setup_interp2d_per_f64", _p(so), so.shape[0], _p(ci), so.shape[2],
                  so.shape[1], ci.shape[2], ci.shape[1], int(jpn), _stream())
        return ci
    _abi.call("cdr_hip_setup_interp2d_f64", _p(so), so.shape[0], _p(ci), so.shape[2], so.shape[1],
              ci.shape[2], ci.shape[1], _stream())
    return ci


def setup_galerkin(so, ci, soc, jpn=0):
    so, ci, soc = map(_t, (so, ci, soc))
    if jpn:
        _abi.call("cdr_hip_setup_galerkin2d_per_f64", _p(so), so.shape[0], _p(ci), _p(soc),
                  so.shape[2], so.shape[1], ci.shape[2], ci.shape[1], int(jpn), _stream())
        return soc
    _abi.call("cdr_hip_setup_galerkin2d_f64", _p(so), so.shape[0], _p(ci), _p(soc), so.shape[2],
              so.shape[1], ci.shape[2], ci.shape[1], _stream())
    return soc


def setup_cg(so):
    """Returns the band factor ABD as an (nx*ny, ii) device tensor (Fortran ABD(ii, n))."""
    so = _t(so)
    ii, jj = so.shape[2], so.shape[1]
    abd = torch.zeros(((ii - 2) * (jj - 2), ii), dtype=torch.float64, device=so.device)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call("cdr_hip_setup_cg2d_f64", _p(so), so.shape[0], _p(abd), ii, jj, _p(info), _stream())
    return abd


def solve_cg(abd, qf, q):
    qf, q = _t(qf), _t(q)
    bbd = torch.zeros(abd.shape[0], dtype=torch.float64, device=q.device)
    _abi.call(f"cdr_hip_solve_cg2d_{_sfx(q)}", _p(q), _p(qf), _p(abd), _p(bbd), q.shape[1],
              q.shape[0], _stream())
    return q


def setup_cg_periodic(so, jpn):
    """The factored dense periodic coarse matrix as an (n, n) device tensor (Fortran ABD(n, n))."""
    so = _t(so)
    ii, jj = so.shape[2], so.shape[1]
    n = (ii - 2) * (jj - 2)
    abd = torch.zeros((n, n), dtype=torch.float64, device=so.device)
    info = torch.zeros(1, dtype=torch.int32, device=so.device)
    _abi.call("cdr_hip_setup_cg2d_per_f64", _p(so), so.shape[0], _p(abd), ii, jj, int(jpn),
              _p(info), _stream())
    return abd


def solve_cg_periodic(abd, qf, q, jpn):
    qf, q = _t(qf), _t(q)
    bbd = torch.zeros(abd.shape[0], dtype=torch.float64, device=q.device)
    _abi.call(f"cdr_hip_solve_cg2d_per_{_sfx(q)}", _p(q), _p(qf), _p(abd), _p(bbd), q.shape[1],
              q.shape[0], int(jpn), _stream())
    return q


# ---------------------------------------------------------------------------------------------
# solver
# ---------------------------------------------------------------------------------------------
class Solver:
    """cdr2::solver<fsten>(so, conf) on the GPU.

    Settings use Cedar's config keys and defaults (src/multilevel_settings.cc:15-61):
    relaxation ('point' | 'line-x' | 'line-y' | 'line-xy'), nrelax_pre=2, nrelax_post=1,
    num_levels=-1, maxiter=10, tol=1e-8, min_coarse=3; precision 'f64' (Cedar's algorithm) or
    'mixed' (FP32 V-cycle preconditioning FP64 iterative refinement); graph=True replays each
    V-cycle from a captured hipGraph; cycle 'v' or 'f' (solver.cycle.type); periodic = (x, y)
    flags of grid.periodic (the operator, b and x hold the periodic images in their ghosts, as in
    examples/basic-2d-ser/periodic.cc); stall_cycles=k > 0 stops a problem whose residual has
    not reached a new minimum for k cycles (solve() then raises CedarHipError with status
    CDR_ERR_STALLED, carrying .x and .hist), 0 (default) keeps the reference's loop.

    A 4-D operator (nb, nstncl, ny+2, nx+2) makes a batched solver: nb independent problems of
    the same extents solved together (every kernel launch covers the batch); solve() then takes
    and returns (nb, ny+2, nx+2) grids and a list of nb residual histories, each the history of
    solving that problem alone.  nrhs > 1 solves every operator for nrhs right-hand sides at once
    (cdr_hip_solver2d_create_shared): the grids then hold nops * nrhs problems, right-hand side r
    of operator k at index k * nrhs + r, and the operators are read once per group.
    """

    def __init__(self, so, relaxation="point", nrelax_pre=2, nrelax_post=1, num_levels=-1,
                 maxiter=10, tol=1e-8, min_coarse=3, precision="f64", graph=True, cycle="v",
                 fused_coarse=True, periodic=(False, False), stall_cycles=0, nrhs=1):
        so = _t(so) if not isinstance(so, np.ndarray) else torch.as_tensor(so).cuda()
        if so.dtype != torch.float64:
            so = so.to(torch.float64)
        self.so = so.contiguous()  # held by reference for the solver's lifetime
        self.maxiter = maxiter
        self.batched = self.so.dim() == 4 or nrhs > 1
        nops = self.so.shape[0] if self.so.dim() == 4 else 1
        self.nops, self.nrhs = nops, int(nrhs)
        nb = nops * self.nrhs
        self.nb = nb
        ns, jj, ii = self.so.shape[-3:]
        s = _abi.Settings(_abi.RELAX[relaxation], nrelax_pre, nrelax_post, num_levels, maxiter,
                          tol, min_coarse, _abi.PRECISION[precision], int(bool(graph)),
                          _abi.CYCLE[cycle], int(bool(fused_coarse)),
                          int(bool(periodic[0])) | (int(bool(periodic[1])) << 1),
                          int(stall_cycles))
        h = ctypes.c_void_p()
        _abi.call("cdr_hip_solver2d_create_shared", _p(self.so), ns, ii - 2, jj - 2, nops,
                  self.nrhs, ctypes.byref(s), _stream(), ctypes.byref(h))
        self._h = h

    @property
    def nlevels(self):
        return _abi.load().cdr_hip_solver2d_nlevels(self._h)

    def level(self, lvl):
        """(nx, ny, so_ptr, ci_ptr, sor0_ptr, sor1_ptr) of a level (device pointers)."""
        nx, ny = ctypes.c_int(), ctypes.c_int()
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        _abi.call("cdr_hip_solver2d_level", self._h, lvl, ctypes.byref(nx), ctypes.byref(ny),
                  *[ctypes.byref(p) for p in ptrs])
        return (nx.value, ny.value) + tuple(p.value for p in ptrs)

    def level_arrays(self, lvl):
        """Host copies (so, ci, sor0, sor1) of a level's FP64 data."""
        nx, ny, so, ci, s0, s1 = self.level(lvl)
        ii, jj = nx + 2, ny + 2
        ns = self.so.shape[-3] if lvl == 0 else 5
        lead = (self.nops,) if self.batched else ()

        def grab(ptr, planes):
            if not ptr:
                return None
            out = np.empty(lead + (planes, jj, ii))
            torch.cuda.synchronize()
            _memcpy_d2h(out, ptr)
            return out
        return grab(so, ns), grab(ci, 8), grab(s0, 2), grab(s1, 2)

    def times(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _abi.call("cdr_hip_solver2d_times", self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def _check_grids(self, *gs):
        jj, ii = self.so.shape[-2:]
        for g in gs:
            if g.numel() != self.nb * jj * ii or g.dtype != torch.float64:
                raise ValueError(f"expected {self.nb} float64 grids of {jj} x {ii}, got "
                                 f"{tuple(g.shape)} {g.dtype}")

    def vcycle(self, x, b):
        x, b = _t(x), _t(b)
        self._check_grids(x, b)
        _abi.call("cdr_hip_solver2d_vcycle", self._h, _p(b), _p(x))
        return x

    def solve(self, b, x=None):
        """Returns (x, rho history as a numpy array)."""
        b = _t(b)
        if x is None:
            x = torch.zeros_like(b)
        x = _t(x)
        self._check_grids(x, b)
        m = max(self.maxiter, 1)
        hist = (ctypes.c_double * (m * self.nb))()
        nit = (ctypes.c_int * self.nb)()
        st = _abi.load().cdr_hip_solver2d_solve(self._h, _p(b), _p(x), hist, m, nit)
        hs = [np.array(hist[k * m:k * m + nit[k]]) for k in range(self.nb)]
        out = hs if self.batched else hs[0]
        if st != _abi.CDR_OK:
            err = _abi.CedarHipError(st, "cdr_hip_solver2d_solve")
            err.x, err.hist = x, out
            raise err
        return x, out

    def kernels(self):
        """The kernel registry: {interface: {"impls": [...], "default": name}}."""
        from .cdr3 import _registry
        return _registry(_abi.load().cdr_hip_solver2d_kernels, self._h)

    def set_kernel(self, interface, impl):
        """kman->set<I>(impl) for the interface named `interface` (both precisions)."""
        _abi.call("cdr_hip_solver2d_set_kernel", self._h, interface.encode(), impl.encode())

    def timers(self, on=True):
        """Per-level device timers of the V-cycle phases (cycles then run without graph replay
        and fused coarse levels; the results are the same)."""
        _abi.call("cdr_hip_solver2d_timers", self._h, 1 if on else 0)

    def save_timers(self, path):
        """Writes the timers as JSON in the reference's time_log layout and resets them."""
        _abi.call("cdr_hip_solver2d_save_timers", self._h, str(path).encode())

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _abi.load().cdr_hip_solver2d_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _memcpy_d2h(out, ptr):
    """Copy out.nbytes from device pointer ptr into the numpy array out (synchronous)."""
    hip = _hip()
    st = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr),
                       ctypes.c_size_t(out.nbytes), 2)  # hipMemcpyDeviceToHost
    if st != 0:
        raise RuntimeError(f"hipMemcpy failed ({st})")


_hiprt = None


def _hip():
    global _hiprt
    if _hiprt is None:
        _hiprt = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded (by soname)
        _hiprt.hipMemcpy.restype = ctypes.c_int
        _hiprt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hiprt


def write_stencil(so, path):
    """Text dump of a 2D stencil in the reference's layout (src/2d/stencil_op.cc:5-51)."""
    so = _t(so)
    so = so if isinstance(so, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(so))
    so = so.to(device="cuda", dtype=torch.float64).contiguous()
    ns, jj, ii = so.shape
    _abi.call("cdr_hip_stencil2d_write", _p(so), ns, ii, jj, str(path).encode())

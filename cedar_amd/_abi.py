"""ctypes binding of the C ABI in include/cedar_hip.h (libcedar_hip.so, built in-tree).

load() (called on first use) loads the HIP library and raises if it is missing: there is no
CPU fallback anywhere in cedar_amd.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcedar_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cedar_hip.h")

CDR_OK = 0
CDR_DOWN, CDR_UP = 0, 1
RELAX = {"point": 0, "line-x": 1, "line-y": 2, "line-xy": 3,
         "plane-xy": 4, "plane-xz": 5, "plane-yz": 6, "plane-xyz": 7}
PLANE_FAMILY = {"xy": 0, "xz": 1, "yz": 2}
PRECISION = {"f64": 0, "mixed": 1}
CYCLE = {"v": 0, "f": 1}

_i = ctypes.c_int
_d = ctypes.c_double
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


class Settings(ctypes.Structure):
    """cdr_solver_settings"""
    _fields_ = [("relaxation", _i), ("nrelax_pre", _i), ("nrelax_post", _i), ("num_levels", _i),
                ("maxiter", _i), ("tol", _d), ("min_coarse", _i), ("precision", _i), ("graph", _i),
                ("cycle", _i), ("fused_coarse", _i), ("periodic", _i), ("stall_cycles", _i)]


class Settings3(ctypes.Structure):
    """cdr_solver3_settings"""
    _fields_ = [("relaxation", _i), ("nrelax_pre", _i), ("nrelax_post", _i), ("num_levels", _i),
                ("maxiter", _i), ("tol", _d), ("min_coarse", _i), ("plane", Settings),
                ("cycle", _i), ("precision", _i), ("stall_cycles", _i), ("periodic", _i)]


# name -> (restype, argtypes); pointers to device memory are passed as c_void_p
_P = _vp
PROTOTYPES = {
    "cdr_hip_status_string": (ctypes.c_char_p, [_i]),
    "cdr_hip_version": (ctypes.c_char_p, []),
    "cdr_hip_autotune": (_i, [_i]),
    "cdr_hip_autotune_report": (_i, [ctypes.c_char_p, _i]),
    "cdr_hip_line_shapes2d": (_i, [_i, _i, _i, _i, _i]),
    "cdr_hip_set_line_shape2d": (_i, [_i, _i, _i, _i, _i, _i, _i]),
    "cdr_hip_norm_work2d": (_sz, [_i, _i]),
    "cdr_hip_residual2d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _vp]),
    "cdr_hip_residual2d_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _vp]),
    "cdr_hip_residual_norm2d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _P, _P, _vp]),
    "cdr_hip_residual_norm2d_f64_noalloc": (_i, [_P, _i, _P, _P, _P, _i, _i, _P, _vp]),
    "cdr_hip_sumsq2d_f64": (_i, [_P, _i, _i, _P, _P, _vp]),
    "cdr_hip_setup_recip2d_f64": (_i, [_P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs2d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs2d_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_setup_lines_x2d_f64": (_i, [_P, _P, _i, _i, _i, _P, _vp]),
    "cdr_hip_setup_lines_y2d_f64": (_i, [_P, _P, _i, _i, _i, _P, _vp]),
    "cdr_hip_relax_lines2d_batched_f64": (_i, [_i, _P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines2d_batched_f32": (_i, [_i, _P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_x2d_f64": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_y2d_f64": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_x2d_f32": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_y2d_f32": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_xy2d_f64": (_i, [_P, _i, _P, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_xy2d_f32": (_i, [_P, _i, _P, _P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_restrict2d_f64": (_i, [_P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict2d_f32": (_i, [_P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add2d_f64": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add2d_f32": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_interp2d_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_galerkin2d_f64": (_i, [_P, _i, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_cg2d_f64": (_i, [_P, _i, _P, _i, _i, _P, _vp]),
    "cdr_hip_solve_cg2d_f64": (_i, [_P, _P, _P, _P, _i, _i, _vp]),
    "cdr_hip_solve_cg2d_f32": (_i, [_P, _P, _P, _P, _i, _i, _vp]),
    "cdr_hip_default_settings": (None, [ctypes.POINTER(Settings)]),
    "cdr_hip_stop_rule": (_i, [ctypes.POINTER(_d), _i, _d, _i, ctypes.POINTER(_i)]),
    "cdr_hip_solver2d_create": (_i, [_P, _i, _i, _i, ctypes.POINTER(Settings), _vp,
                                     ctypes.POINTER(_vp)]),
    "cdr_hip_solver2d_create_batched": (_i, [_P, _i, _i, _i, _i, ctypes.POINTER(Settings), _vp,
                                             ctypes.POINTER(_vp)]),
    "cdr_hip_solver2d_create_shared": (_i, [_P, _i, _i, _i, _i, _i, ctypes.POINTER(Settings),
                                            _vp, ctypes.POINTER(_vp)]),
    "cdr_hip_solver2d_nbatch": (_i, [_vp]),
    "cdr_hip_solver2d_solve": (_i, [_vp, _P, _P, ctypes.POINTER(_d), _i, ctypes.POINTER(_i)]),
    "cdr_hip_solver2d_vcycle": (_i, [_vp, _P, _P]),
    "cdr_hip_solver2d_nlevels": (_i, [_vp]),
    "cdr_hip_solver2d_level": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp)]),
    "cdr_hip_solver2d_times": (_i, [_vp, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "cdr_hip_solver2d_timers": (_i, [_vp, _i]),
    "cdr_hip_solver2d_save_timers": (_i, [_vp, ctypes.c_char_p]),
    "cdr_hip_stencil2d_write": (_i, [_P, _i, _i, _i, ctypes.c_char_p]),
    "cdr_hip_solver2d_kernels": (_i, [_vp, ctypes.c_char_p, _i]),
    "cdr_hip_solver2d_set_kernel": (_i, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "cdr_hip_solver2d_destroy": (None, [_vp]),
    # 3D kernels
    "cdr_hip_norm_work3d": (_sz, [_i, _i, _i]),
    "cdr_hip_residual3d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_residual3d_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_residual_norm3d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _P, _P, _vp]),
    "cdr_hip_setup_recip3d_f64": (_i, [_P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs3d_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs3d_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict3d_f64": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict3d_f32": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add3d_f64": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add3d_f32": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_interp3d_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_galerkin3d_f64": (_i, [_P, _i, _P, _P, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_cg_bandwidth3d": (_i, [_i, _i]),
    "cdr_hip_setup_cg3d_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _P, _vp]),
    "cdr_hip_solve_cg3d_f64": (_i, [_P, _P, _P, _i, _P, _i, _i, _i, _vp]),
    "cdr_hip_solve_cg3d_f32": (_i, [_P, _P, _P, _i, _P, _i, _i, _i, _vp]),
    # 3D solver
    "cdr_hip_default_settings3": (None, [ctypes.POINTER(Settings3)]),
    "cdr_hip_relax_gs3d_per_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs3d_per_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict3d_per_f64": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict3d_per_f32": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add3d_per_f64": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _i, _i,
                                          _vp]),
    "cdr_hip_interp_add3d_per_f32": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _i, _i,
                                          _vp]),
    "cdr_hip_setup_interp3d_per_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_galerkin3d_per_f64": (_i, [_P, _i, _P, _P, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_cg3d_per_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _P, _vp]),
    "cdr_hip_solve_cg3d_per_f64": (_i, [_P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_solve_cg3d_per_f32": (_i, [_P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_solver3d_create": (_i, [_P, _i, _i, _i, _i, ctypes.POINTER(Settings3), _vp,
                                     ctypes.POINTER(_vp)]),
    "cdr_hip_solver3d_solve": (_i, [_vp, _P, _P, ctypes.POINTER(_d), _i, ctypes.POINTER(_i)]),
    "cdr_hip_solver3d_vcycle": (_i, [_vp, _P, _P]),
    "cdr_hip_solver3d_relax_planes": (_i, [_vp, _i, _i, _P, _P, _i]),
    "cdr_hip_solver3d_nlevels": (_i, [_vp]),
    "cdr_hip_solver3d_level": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                    ctypes.POINTER(_i), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "cdr_hip_solver3d_times": (_i, [_vp, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "cdr_hip_solver3d_kernels": (_i, [_vp, ctypes.c_char_p, _i]),
    "cdr_hip_solver3d_set_kernel": (_i, [_vp, ctypes.c_char_p, ctypes.c_char_p]),
    "cdr_hip_solver3d_destroy": (None, [_vp]),
    # periodic boundaries
    "cdr_hip_relax_gs2d_per_f64": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_gs2d_per_f32": (_i, [_P, _i, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_lines_x2d_per_f64": (_i, [_P, _P, _i, _i, _i, _i, _P, _vp]),
    "cdr_hip_setup_lines_y2d_per_f64": (_i, [_P, _P, _i, _i, _i, _i, _P, _vp]),
    "cdr_hip_relax_lines_x2d_per_f64": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_y2d_per_f64": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_x2d_per_f32": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_relax_lines_y2d_per_f32": (_i, [_P, _i, _P, _P, _P, _P, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict2d_per_f64": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_restrict2d_per_f32": (_i, [_P, _P, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add2d_per_f64": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_interp_add2d_per_f32": (_i, [_P, _P, _P, _P, _i, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_interp2d_per_f64": (_i, [_P, _i, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_galerkin2d_per_f64": (_i, [_P, _i, _P, _P, _i, _i, _i, _i, _i, _vp]),
    "cdr_hip_setup_cg2d_per_f64": (_i, [_P, _i, _P, _i, _i, _i, _P, _vp]),
    "cdr_hip_solve_cg2d_per_f64": (_i, [_P, _P, _P, _P, _i, _i, _i, _vp]),
    "cdr_hip_solve_cg2d_per_f32": (_i, [_P, _P, _P, _P, _i, _i, _i, _vp]),
    # matvec
    "cdr_hip_matvec2d_f64": (_i, [_P, _i, _P, _P, _i, _i, _vp]),
    "cdr_hip_matvec3d_f64": (_i, [_P, _i, _P, _P, _i, _i, _i, _vp]),
    # differentiated 2D solves
    "cdr_hip_q1_assemble2d_f64": (_i, [_P, _P, _i, _i, _i, _vp]),
    "cdr_hip_q1_gradient2d_f64": (_i, [_P, _P, _P, _i, _i, _i, _vp]),
}


class CedarHipError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


_lib = None


def _hip_runtimes():
    """Distinct libamdhip64 files mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for ln in f:
                if "libamdhip64" in ln:
                    paths.add(os.path.realpath(ln.split()[-1]))
    except OSError:
        pass
    return paths


def load():
    """Load libcedar_hip.so (raises if it has not been built).

    torch is imported first so that the library binds to the HIP runtime torch already uses
    (same soname): device pointers, streams and events are then shared by one runtime.
    """
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (loads the process's HIP runtime before ours)
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"cedar_amd: {LIB_PATH} is missing; build it with "
                              "`python -m cedar_amd.build` (no CPU fallback exists)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        rts = _hip_runtimes()
        if len(rts) > 1:
            raise ImportError(f"cedar_amd: two HIP runtimes are loaded ({sorted(rts)}); import "
                              "torch before any other HIP library")
        _lib = lib
    return _lib


def status_string(status):
    return load().cdr_hip_status_string(status).decode()


def call(name, *args):
    """Call an int-returning entry point and raise CedarHipError on a non-zero status."""
    st = getattr(load(), name)(*args)
    if st != CDR_OK:
        raise CedarHipError(st, name)
    return st

"""cedar_amd: Cedar's BoxMG solver path on one MI355X (gfx950).

The numerical work runs in hand-written HIP kernels behind the C ABI in include/cedar_hip.h
(libcedar_hip.so); Python only allocates device memory through torch and passes pointers.
The library is loaded on first use (or by `cedar_amd.load()`), which raises if it is missing --
there is no CPU path.  Importing the package alone does not load it, so that
`python -m cedar_amd.build` can rebuild a stale library.
"""
from . import _abi
from ._abi import CedarHipError, RELAX, PRECISION, load  # noqa: F401

__version__ = "0.1.0"

from .cdr2 import (  # noqa: E402,F401
    GridFunc, StencilOp, Solver, residual, residual_norm, setup_recip, relax_gs, setup_lines,
    relax_lines, relax_lines_xy, restrict, interp_add, setup_interp, setup_galerkin, setup_cg, solve_cg,
    write_stencil, autotune, autotune_report, bc_from_mask, setup_cg_periodic, solve_cg_periodic,
    setup_lines_periodic, relax_lines_periodic,
)
from . import cdr3  # noqa: E402,F401
from .cdr3 import Solver3  # noqa: E402,F401

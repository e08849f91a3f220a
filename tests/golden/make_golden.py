"""Writes tests/golden/bmg2_kernels.npz: inputs and outputs of the reference's own serial 2D kernels.

The kernels are Cedar's Fortran sources compiled as they are (`make -C oracle ref`, which reads
/root/reference); this script only runs in the container that has them.  The fixture is data
(seeded random inputs and the reference's outputs) and is what the CPU tests compare the oracle
against when the compiled reference is absent, e.g. on a machine without /root/reference.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import refkernels as R  # noqa: E402

# odd and even interior sizes, both stencil widths
CASES = [(7, 9), (8, 6), (16, 11), (15, 15)]


def rand_so(rng, nx, ny, ns):
    """Diagonally dominant M-matrix stencil with Dirichlet couplings zeroed, ghosts zero."""
    so = np.zeros((ns, ny + 2, nx + 2))
    so[1:, 1:ny + 1, 1:nx + 1] = rng.uniform(0.1, 1.0, (ns - 1, ny, nx))
    so[1, :, 1] = 0.0
    so[2, 1, :] = 0.0
    if ns == 5:
        so[3, 1, :] = 0.0
        so[3, :, 1] = 0.0
        so[4, :, 1] = 0.0
        so[4, 1, :] = 0.0
    so[0, 1:ny + 1, 1:nx + 1] = 2.0 * (ns - 1) + rng.uniform(0.0, 1.0, (ny, nx))
    return so


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for nx, ny in CASES:
        for ns in (3, 5):
            tag = f"{nx}x{ny}_{ns}"
            so = rand_so(rng, nx, ny, ns)
            q = rng.standard_normal((ny + 2, nx + 2))
            qf = rng.standard_normal((ny + 2, nx + 2))
            res = rng.standard_normal((ny + 2, nx + 2))
            sor = R.setup_recip(so)
            ci = R.setup_interp(so)
            jjc, iic = ci.shape[1:]
            qc = rng.standard_normal((jjc, iic))
            qi, resi = R.interp_add(q, qc, res, so, ci)
            out.update({
                f"{tag}/so": so, f"{tag}/q": q, f"{tag}/qf": qf, f"{tag}/res": res,
                f"{tag}/qc": qc,
                f"{tag}/residual": R.residual(so, qf, q),
                f"{tag}/recip": sor,
                f"{tag}/gs_down": R.relax_gs(so, qf, q, sor, O.DOWN),
                f"{tag}/gs_up": R.relax_gs(so, qf, q, sor, O.UP),
                f"{tag}/ci": ci,
                f"{tag}/soc": R.setup_galerkin(so, ci),
                f"{tag}/restrict": R.restrict(ci, q),
                f"{tag}/interp_q": qi,
                f"{tag}/interp_res": resi,
            })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bmg2_kernels.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

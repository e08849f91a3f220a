"""Writes tests/golden/bmg2_periodic.npz: inputs and outputs of the periodic branches (JPN != 0)
of the reference's own serial 2D kernels, compiled from their Fortran sources (`make -C oracle
ref`, which reads /root/reference; this script only runs where that build exists).

    python tests/golden/make_golden_periodic.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import refkernels as R  # noqa: E402

# odd, even and ragged interior sizes; masks x (1), y (2), both (3)
CASES = [(8, 8), (7, 9), (12, 6), (16, 15)]


def rand_periodic_so(rng, nx, ny, ns, mask):
    """Random M-matrix stencil; couplings across a periodic boundary kept, Dirichlet ones zeroed,
    and the ghost rows/columns of every plane holding the periodic images."""
    px, py = bool(mask & 1), bool(mask & 2)
    so = np.zeros((ns, ny + 2, nx + 2))
    so[1:, 1:ny + 1, 1:nx + 1] = rng.uniform(0.1, 1.0, (ns - 1, ny, nx))
    if not px:
        so[1, :, 1] = 0.0
        if ns == 5:
            so[3, :, 1] = 0.0
            so[4, :, 1] = 0.0
    if not py:
        so[2, 1, :] = 0.0
        if ns == 5:
            so[3, 1, :] = 0.0
            so[4, 1, :] = 0.0
    so[0, 1:ny + 1, 1:nx + 1] = 2.0 * (ns - 1) + rng.uniform(0.0, 1.0, (ny, nx))
    if px:
        so[:, :, 0] = so[:, :, nx]
        so[:, :, nx + 1] = so[:, :, 1]
    if py:
        so[:, 0, :] = so[:, ny, :]
        so[:, ny + 1, :] = so[:, 1, :]
    return so


def main():
    rng = np.random.default_rng(20251202)
    out = {}
    for nx, ny in CASES:
        for ns in (3, 5):
            for mask in (1, 2, 3):
                jpn = O.bc_from_mask(mask)
                tag = f"{nx}x{ny}_{ns}_{mask}"
                so = rand_periodic_so(rng, nx, ny, ns, mask)
                q = rng.standard_normal((ny + 2, nx + 2))
                qf = rng.standard_normal((ny + 2, nx + 2))
                res = rng.standard_normal((ny + 2, nx + 2))
                sor = R.setup_recip(so)
                ci = R.setup_interp(so, jpn=jpn)
                jjc, iic = ci.shape[1:]
                qc = rng.standard_normal((jjc, iic))
                qi, resi = R.interp_add(q, qc, res, so, ci, jpn=jpn)
                out.update({
                    f"{tag}/so": so, f"{tag}/q": q, f"{tag}/qf": qf, f"{tag}/res": res,
                    f"{tag}/qc": qc,
                    f"{tag}/gs_down": R.relax_gs(so, qf, q, sor, O.DOWN, jpn=jpn),
                    f"{tag}/gs_up": R.relax_gs(so, qf, q, sor, O.UP, jpn=jpn),
                    f"{tag}/ci": ci,
                    f"{tag}/soc": R.setup_galerkin(so, ci, jpn=jpn),
                    f"{tag}/restrict": R.restrict(ci, q, jpn=jpn),
                    f"{tag}/interp_q": qi,
                    f"{tag}/interp_res": resi,
                })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bmg2_periodic.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

"""Writes tests/golden/bmg3_kernels.npz: inputs and outputs of the reference's own serial 3D
kernels, compiled from their sources (`make -C oracle ref`); only runs where /root/reference is.

    python tests/golden/make_golden3.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle3 as O3  # noqa: E402
from oracle import refkernels3 as R3  # noqa: E402
from tests.golden.gen3 import rand_so3  # noqa: E402

# odd and even interior sizes in each direction, both stencils
CASES = [(5, 5, 5), (6, 7, 8), (8, 6, 7)]


def main():
    rng = np.random.default_rng(20240602)
    out = {}
    for nx, ny, nz in CASES:
        for ns in (4, 14):
            tag = f"{nx}x{ny}x{nz}_{ns}"
            so = rand_so3(rng, nx, ny, nz, ns)
            shp = (nz + 2, ny + 2, nx + 2)
            q = rng.standard_normal(shp)
            qf = rng.standard_normal(shp)
            res = rng.standard_normal(shp)
            sor = R3.setup_recip(so)
            ci = R3.setup_interp(so)
            qc = rng.standard_normal(ci.shape[1:])
            qi, resi = R3.interp_add(q, qc, res, so, ci)
            out.update({
                f"{tag}/so": so, f"{tag}/q": q, f"{tag}/qf": qf, f"{tag}/res": res,
                f"{tag}/qc": qc,
                f"{tag}/residual": R3.residual(so, qf, q),
                f"{tag}/recip": sor,
                f"{tag}/gs_down": R3.relax_gs(so, qf, q, sor, O3.DOWN),
                f"{tag}/gs_up": R3.relax_gs(so, qf, q, sor, O3.UP),
                f"{tag}/ci": ci,
                f"{tag}/soc": R3.setup_galerkin(so, ci),
                f"{tag}/restrict": R3.restrict(ci, q),
                f"{tag}/interp_q": qi, f"{tag}/interp_res": resi,
            })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bmg3_kernels.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.nbytes for v in out.values()) // 1024, "KiB raw")


if __name__ == "__main__":
    main()

"""Writes tests/golden/bmg3_periodic.npz: outputs of the reference's own serial 3D kernels with
periodic boundary ids, compiled from their sources (`make -C oracle ref`); only runs where
/root/reference is.  The inputs are regenerated from their seeds by tests.golden.gen3 (numpy's
PCG64 streams), so only the outputs are stored.

    python tests/golden/make_golden3_periodic.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle3 as O3  # noqa: E402
from oracle import refkernels3 as R3  # noqa: E402
from tests.golden.gen3 import periodic_case  # noqa: E402

# (nx, ny, nz), stencil planes, periodic mask; the last two exercise the 7-point branch's
# image row past column jjc-1 and the 27-point branch's copies on the plane kkc
CASES = ([((6, 6, 6), ns, m) for m in (1, 2, 3, 4, 7) for ns in (4, 14)] +
         [((7, 9, 5), 4, 2), ((8, 6, 10), 14, 6)])


def tag(n, ns, mask):
    return f"{n[0]}x{n[1]}x{n[2]}_{ns}_m{mask}"


def main():
    out = {}
    for n, ns, mask in CASES:
        t = tag(n, ns, mask)
        so, q, qf = periodic_case(n, ns, mask)
        jpn = O3.bc_from_mask(mask)
        sor = O3.setup_recip(so)
        ci = R3.setup_interp(so, jpn=jpn)
        qc, qi = R3.restrict(ci, q, jpn=jpn, with_q=True)
        out.update({
            f"{t}/gs_down": R3.relax_gs(so, qf, q, sor, O3.DOWN, jpn),
            f"{t}/gs_up": R3.relax_gs(so, qf, q, sor, O3.UP, jpn),
            f"{t}/ci": ci,
            f"{t}/restrict": qc, f"{t}/restrict_q": qi,
            f"{t}/soc": R3.setup_galerkin(so, ci, jpn=jpn),
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bmg3_periodic.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.nbytes for v in out.values()) // 1024, "KiB raw")


if __name__ == "__main__":
    main()

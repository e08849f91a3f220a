"""The FP64 oracle against the reference's own kernels (golden fixtures + live compiled reference).

The fixtures in tests/golden/bmg2_kernels.npz were written by tests/golden/make_golden.py from
Cedar's Fortran kernels compiled from source; every comparison here is bitwise.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import refkernels as R

CASES = ["7x9_3", "7x9_5", "8x6_3", "8x6_5", "16x11_3", "16x11_5", "15x15_3", "15x15_5"]


@pytest.mark.parametrize("tag", CASES)
def test_kernels_match_reference_fixtures(golden2d, tag):
    g = {k.split("/", 1)[1]: golden2d[k] for k in golden2d.files if k.startswith(tag + "/")}
    so, q, qf, res, qc = g["so"], g["q"], g["qf"], g["res"], g["qc"]
    assert np.array_equal(O.residual(so, qf, q), g["residual"])
    sor = O.setup_recip(so)
    assert np.array_equal(sor, g["recip"])
    assert np.array_equal(O.relax_gs(so, qf, q, sor, O.DOWN), g["gs_down"])
    assert np.array_equal(O.relax_gs(so, qf, q, sor, O.UP), g["gs_up"])
    ci = O.setup_interp(so)
    assert np.array_equal(ci, g["ci"])
    assert np.array_equal(O.setup_galerkin(so, ci), g["soc"])
    assert np.array_equal(O.restrict(ci, q), g["restrict"])
    qi, resi = O.interp_add(q, qc, res, so, ci)
    assert np.array_equal(qi, g["interp_q"])
    assert np.array_equal(resi, g["interp_res"])


@pytest.mark.skipif(not O.ref_available(), reason="reference kernels not compiled (make -C oracle ref)")
@pytest.mark.parametrize("nx,ny", [(5, 5), (12, 13), (33, 31), (64, 40)])
@pytest.mark.parametrize("ns", [3, 5])
def test_kernels_match_live_reference(nx, ny, ns):
    from tests.golden.make_golden import rand_so
    rng = np.random.default_rng(nx * 1000 + ny * 10 + ns)
    so = rand_so(rng, nx, ny, ns)
    q = rng.standard_normal((ny + 2, nx + 2))
    qf = rng.standard_normal((ny + 2, nx + 2))
    assert np.array_equal(O.residual(so, qf, q), R.residual(so, qf, q))
    sor = O.setup_recip(so)
    for ud in (O.DOWN, O.UP):
        assert np.array_equal(O.relax_gs(so, qf, q, sor, ud), R.relax_gs(so, qf, q, sor, ud))
    ci = O.setup_interp(so)
    assert np.array_equal(ci, R.setup_interp(so))
    assert np.array_equal(O.setup_galerkin(so, ci), R.setup_galerkin(so, ci))
    assert np.array_equal(O.restrict(ci, q), R.restrict(ci, q))
    qc = rng.standard_normal(ci.shape[1:])
    res = rng.standard_normal(q.shape)
    a, b = O.interp_add(q, qc, res, so, ci), R.interp_add(q, qc, res, so, ci)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])

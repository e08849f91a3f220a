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


@pytest.mark.parametrize("ns", [3, 5])
def test_matvec_pins(ns):
    """The matvec restatement (the reference's BMG2_SymStd_UTILS_matvec includes mpif.h and does
    not build here) against the reference-built residual and the symmetry of A."""
    from oracle import oracle3 as O3
    from oracle import refkernels3 as R3
    rng = np.random.default_rng(ns)
    nx, ny = 13, 10
    so = np.zeros((ns, ny + 2, nx + 2))
    so[1:, 1:ny + 1, 1:nx + 1] = rng.uniform(0.1, 1.0, (ns - 1, ny, nx))
    so[0, 1:ny + 1, 1:nx + 1] = 8.0
    q = rng.standard_normal((ny + 2, nx + 2))
    y = O.matvec(so, q)
    assert not y[0].any() and not y[:, 0].any() and not y[-1].any() and not y[:, -1].any()
    if O.ref_available():
        qf = rng.standard_normal(q.shape)
        r = R.residual(so, qf, q)
        np.testing.assert_allclose(y[1:-1, 1:-1], (qf - r)[1:-1, 1:-1], rtol=1e-12, atol=1e-12)
    # interior-supported vectors: x^T A z == z^T A x exposes a wrongly placed coupling
    x, z = np.zeros_like(q), np.zeros_like(q)
    x[1:-1, 1:-1] = rng.standard_normal((ny, nx))
    z[1:-1, 1:-1] = rng.standard_normal((ny, nx))
    assert abs((x * O.matvec(so, z)).sum() - (z * O.matvec(so, x)).sum()) < 1e-10
    # ghosts of y are kept
    y0 = np.full_like(q, 7.0)
    assert O.matvec(so, q, y0)[0, 0] == 7.0
    # 3D: same two pins against the reference-built 3D residual
    for ns3 in (4, 14):
        s3 = np.zeros((ns3, 7, 6, 8))
        s3[1:, 1:-1, 1:-1, 1:-1] = rng.uniform(0.1, 1.0, (ns3 - 1, 5, 4, 6))
        s3[0, 1:-1, 1:-1, 1:-1] = 30.0
        q3 = rng.standard_normal((7, 6, 8))
        y3 = O3.matvec(s3, q3)
        if O.ref_available():
            qf3 = rng.standard_normal(q3.shape)
            r3 = R3.residual(s3, qf3, q3)
            np.testing.assert_allclose(y3[1:-1, 1:-1, 1:-1], (qf3 - r3)[1:-1, 1:-1, 1:-1],
                                       rtol=1e-12, atol=1e-12)
        a, b = np.zeros_like(q3), np.zeros_like(q3)
        a[1:-1, 1:-1, 1:-1] = rng.standard_normal((5, 4, 6))
        b[1:-1, 1:-1, 1:-1] = rng.standard_normal((5, 4, 6))
        assert abs((a * O3.matvec(s3, b)).sum() - (b * O3.matvec(s3, a)).sum()) < 1e-10

"""The 3D FP64 oracle against the reference's own kernels and against the mathematics.

* Kernels: bitwise against tests/golden/bmg3_kernels.npz (written by make_golden3.py from the
  Fortran kernels compiled from source) and, where oracle/_ref is built, against the live
  compiled kernels on more sizes.  The Galerkin operator is a triple product P^T A P here and
  explicit formulas in the reference (ITLI07/ITLI27): they agree to rounding (<= 1e-14 rel).
* Coarse solve: the band Cholesky solve satisfies A q = f to rounding (A applied by the residual
  kernel).
* Plane relaxation: restates test/3d/test_planes.cc: with exact plane solves (point smoother, 50
  cycles, tol 1e-30 as in test/3d/test-planes-ser.json) the planes of the second colour solve
  their plane system, checked with scipy's sparse direct solver on the textbook plane stencil
  (test/3d/pyplanes.pyx:8-141).
* Whole solves: 3D Poisson to the reference test's tolerances (test/3d/test_poisson.cc:73-104)
  at a size the CPU runs in seconds; plane relaxation on an anisotropic problem.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from cedar_amd import gallery as G
from oracle import oracle as O
from oracle import oracle3 as O3
from tests.golden.gen3 import rand_so3

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [f"{c}_{ns}" for c in ("5x5x5", "6x7x8", "8x6x7") for ns in (4, 14)]


@pytest.fixture(scope="module")
def golden3():
    return np.load(os.path.join(HERE, "golden", "bmg3_kernels.npz"))


def check_kernels(so, q, qf, res, qc, ref):
    assert np.array_equal(O3.residual(so, qf, q), ref["residual"])
    sor = O3.setup_recip(so)
    assert np.array_equal(sor, ref["recip"])
    assert np.array_equal(O3.relax_gs(so, qf, q, sor, O3.DOWN), ref["gs_down"])
    assert np.array_equal(O3.relax_gs(so, qf, q, sor, O3.UP), ref["gs_up"])
    ci = O3.setup_interp(so)
    assert np.array_equal(ci, ref["ci"])
    soc = O3.setup_galerkin(so, ci)
    assert np.abs(soc - ref["soc"]).max() <= 1e-14 * np.abs(ref["soc"]).max()
    assert np.array_equal(O3.restrict(ci, q), ref["restrict"])
    qi, resi = O3.interp_add(q, qc, res, so, ci)
    assert np.array_equal(qi, ref["interp_q"])
    assert np.array_equal(resi, ref["interp_res"])


@pytest.mark.parametrize("tag", CASES)
def test_kernels_match_reference_fixtures(golden3, tag):
    g = {k.split("/", 1)[1]: golden3[k] for k in golden3.files if k.startswith(tag + "/")}
    check_kernels(g["so"], g["q"], g["qf"], g["res"], g["qc"], g)


@pytest.mark.skipif(not O.ref_available(), reason="reference kernels not compiled (make -C oracle ref)")
@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 5, 6), (9, 9, 9), (10, 12, 11), (17, 8, 13)])
@pytest.mark.parametrize("ns", [4, 14])
def test_kernels_match_live_reference(dims, ns):
    from oracle import refkernels3 as R3
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 10000 + ny * 100 + nz + ns)
    so = rand_so3(rng, nx, ny, nz, ns)
    shp = (nz + 2, ny + 2, nx + 2)
    q, qf, res = (rng.standard_normal(shp) for _ in range(3))
    sor = R3.setup_recip(so)
    ci = R3.setup_interp(so)
    qc = rng.standard_normal(ci.shape[1:])
    qi, resi = R3.interp_add(q, qc, res, so, ci)
    ref = {"residual": R3.residual(so, qf, q), "recip": sor,
           "gs_down": R3.relax_gs(so, qf, q, sor, O3.DOWN),
           "gs_up": R3.relax_gs(so, qf, q, sor, O3.UP), "ci": ci,
           "soc": R3.setup_galerkin(so, ci), "restrict": R3.restrict(ci, q),
           "interp_q": qi, "interp_res": resi}
    check_kernels(so, q, qf, res, qc, ref)


def apply_A(so, x):
    return -O3.residual(so, np.zeros_like(x), x)


@pytest.mark.parametrize("dims,ns", [((3, 3, 3), 4), ((4, 3, 5), 14), ((5, 5, 4), 14)])
def test_coarse_solve_solves(dims, ns):
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    so = rand_so3(rng, nx, ny, nz, ns)
    f = np.zeros((nz + 2, ny + 2, nx + 2))
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal((nz, ny, nx))
    abd, info = O3.setup_cg(so)
    assert info == 0
    q = O3.solve_cg(abd, f)
    r = apply_A(so, q) - f
    assert np.abs(r[1:-1, 1:-1, 1:-1]).max() <= 1e-12 * np.abs(f).max()
    # the band holds exactly the upper triangle of A: compare with a dense factorisation
    n = nx * ny * nz
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        x = np.zeros_like(f)
        x[1:-1, 1:-1, 1:-1] = e.reshape(nz, ny, nx)
        A[:, c] = apply_A(so, x)[1:-1, 1:-1, 1:-1].ravel()
    assert np.allclose(A, A.T, rtol=0, atol=1e-14)
    L = np.linalg.cholesky(A)
    kd = O3.cg_bandwidth(nx + 2, ny + 2)
    U = np.zeros((n, n))
    for c in range(n):
        for r in range(max(0, c - kd), c + 1):
            U[r, c] = abd[c, kd + r - c]
    assert np.abs(U - L.T).max() <= 1e-13 * np.abs(L).max()


def stencil_matrix(nx, ny, st):
    """Sparse matrix of a constant 3x3 stencil st[dy][dx] on nx x ny with Dirichlet boundary
    (what pyamg.gallery.stencil_grid builds for test/3d/pyplanes.pyx)."""
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            v = st[dy + 1][dx + 1]
            if v == 0:
                continue
            j0, j1 = max(0, -dy), min(ny, ny - dy)
            i0, i1 = max(0, -dx), min(nx, nx - dx)
            r = idx[j0:j1, i0:i1].ravel()
            c = idx[j0 + dy:j1 + dy, i0 + dx:i1 + dx].ravel()
            rows.append(r)
            cols.append(c)
            vals.append(np.full(r.size, v))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


PLANE_CASES = [("xy", (31, 35, 5), False), ("xy", (31, 35, 5), True),
               ("xz", (31, 5, 35), False), ("xz", (31, 5, 35), True),
               ("yz", (5, 31, 35), False), ("yz", (5, 31, 35), True)]


@pytest.mark.parametrize("updown", [O3.DOWN, O3.UP])
@pytest.mark.parametrize("direction,dims,fe", PLANE_CASES)
def test_plane_relaxation_solves_planes(direction, dims, fe, updown):
    """The colour solved last satisfies its plane systems exactly: the even planes going DOWN
    (colours 1 then 2) and the odd planes going UP (colours 2 then 1), as in
    include/cedar/3d/relax_planes.h:44-56."""
    nx, ny, nz = dims
    so = G.fe3(nx, ny, nz) if fe else G.poisson3(nx, ny, nz)
    kk, jj, ii = nz + 2, ny + 2, nx + 2
    K, J, I = np.meshgrid(np.arange(kk), np.arange(jj), np.arange(ii), indexing="ij")
    x = np.zeros((kk, jj, ii))
    x[1:-1, 1:-1, 1:-1] = (K * nx * ny + J * nx + I)[1:-1, 1:-1, 1:-1].astype(float)
    b = x.copy()
    exact_planes = O3.plane_settings(relaxation="point", maxiter=50, tol=1e-30)
    s = O3.Solver3(so, relaxation="plane-" + direction, num_levels=2, plane=exact_planes)
    xr = s.relax_planes(0, direction, x, b, updown)
    hx, hy, hz = 1.0 / (nx + 1), 1.0 / (ny + 1), 1.0 / (nz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    d = 2 * xh + 2 * yh + 2 * zh
    if fe:
        st = [[-1.0, -1.0, -1.0], [-1.0, 26.0, -1.0], [-1.0, -1.0, -1.0]]
    elif direction == "xy":
        st = [[0, -yh, 0], [-xh, d, -xh], [0, -yh, 0]]
    elif direction == "xz":
        st = [[0, -zh, 0], [-xh, d, -xh], [0, -zh, 0]]
    else:
        st = [[0, -zh, 0], [-yh, d, -yh], [0, -zh, 0]]
    _, _, npl = O3.plane_dims(direction, ii, jj, kk)
    for ipl in range(2 if updown == O3.DOWN else 1, npl + 1, 2):
        b2 = O3.copy_rhs(so, xr, b, direction, ipl)
        n1, n2 = b2.shape[1] - 2, b2.shape[0] - 2
        sol = sla.spsolve(stencil_matrix(n1, n2, st).tocsc(), b2[1:-1, 1:-1].ravel())
        got = O3.copy32(xr, direction, ipl)[1:-1, 1:-1].ravel()
        assert np.abs(got - sol).max() < 1e-8, ipl


def compose_vcycle2(s, so, b, x, families_down, families_up):
    """One two-level V(1,1) cycle assembled from separately pinned pieces: plane sweeps of the
    given families, residual, restriction, band Cholesky coarse solve, interp_add."""
    for d in families_down:
        x = s.relax_planes(0, d, x, b, O3.DOWN)
    r = O3.residual(so, b, x)
    so1, ci1 = s.level(1)
    abd, info = O3.setup_cg(so1)
    assert info == 0
    xc = O3.solve_cg(abd, O3.restrict(ci1, r))
    x, _ = O3.interp_add(x, xc, r, so, ci1)
    for d in families_up:
        x = s.relax_planes(0, d, x, b, O3.UP)
    return x


@pytest.mark.parametrize("relax,down,up", [
    # include/cedar/multilevel.h:179-183 (pre) and :208-212 (post)
    ("plane-xyz", ("xy", "yz", "xz"), ("xz", "yz", "xy")),
    ("plane-xy", ("xy",), ("xy",)), ("plane-xz", ("xz",), ("xz",)),
    ("plane-yz", ("yz",), ("yz",))])
def test_plane_cycle_is_the_family_sweeps_in_reference_order(relax, down, up):
    """The oracle's plane V-cycle equals, bit for bit, the family sweeps in the order
    include/cedar/multilevel.h gives them (pre: xy, yz, xz going DOWN; post: xz, yz, xy going UP)
    around the pinned residual / restriction / coarse solve / interpolation; any other order of
    the three families gives a different result."""
    so, b, _ = G.heterogeneous_7pt(13, seed=8, block=4, sigma=1.0, aniso=(1.0, 0.3, 0.1))
    s = O3.Solver3(so, relaxation=relax, nrelax_pre=1, nrelax_post=1, num_levels=2)
    x0 = np.zeros_like(b)
    x0[1:-1, 1:-1, 1:-1] = np.random.default_rng(9).standard_normal(tuple(n - 2 for n in b.shape))
    want = s.vcycle(x0, b)
    assert np.array_equal(compose_vcycle2(s, so, b, x0, down, up), want)
    if relax == "plane-xyz":
        for d2, u2 in ((("xy", "xz", "yz"), up), (down, ("xy", "yz", "xz")),
                       (("yz", "xy", "xz"), up)):
            assert not np.array_equal(compose_vcycle2(s, so, b, x0, d2, u2), want)


def test_poisson3_point_solve():
    """test/3d/test_poisson.cc: default solver (point V(2,1), 10 cycles, tol 1e-8); the residual
    falls below 1e-8 and the error against sin(2 pi x) sin(2 pi y) sin(2 pi z) is at the level of
    the O(h^2) discretisation error (1e-4 at n = 200 in the reference; (201/32)^2 times that at
    n = 31)."""
    n = 31
    so, b, exact = G.poisson3_sine(n)
    s = O3.Solver3(so)
    x, hist = s.solve(b)
    r = O3.residual(so, b, x)
    assert O3.norm(r) < 1e-8
    assert np.abs(x - exact).max() < 1e-4 * (201 / 32) ** 2


@pytest.mark.parametrize("relax", ["plane-xy", "plane-xyz"])
def test_anisotropic_plane_solves_converge(relax):
    """kz = 1e-3 k: point smoothing stalls, xy-plane smoothing converges fast."""
    so, b, _ = G.heterogeneous_7pt(23, seed=3, block=4)
    s = O3.Solver3(so, relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=12, tol=1e-10)
    x, hist = s.solve(b)
    assert hist[-1] < 1e-10 and len(hist) <= 8
    sp_ = O3.Solver3(so, relaxation="point", nrelax_pre=1, nrelax_post=1, maxiter=8, tol=1e-10)
    _, hp = sp_.solve(b)
    assert hp[-1] > 1e-3


def test_num_levels_3d():
    assert O3.num_levels(255, 255, 255) == 7
    assert O3.num_levels(31, 35, 5) == 2
    assert O3.coarse_dims3(257, 257, 257) == (130, 130, 130)

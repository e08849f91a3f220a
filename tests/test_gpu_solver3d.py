"""Whole 3D solves on the GPU against the FP64 oracle and the reference's own 3D tests.

Bar: residual history within 1e-12 of the oracle's per cycle (relative residuals), the same
cycle count, the solution within 1e-10 relative l2 (point relaxation uses only kernels that are
bitwise equal to the oracle, so there the solution agrees to 1e-12)."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from cedar_amd import gallery as G
from oracle import oracle3 as O3

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def rel_l2(a, b):
    s = (slice(1, -1),) * 3
    return np.linalg.norm((a - b)[s]) / np.linalg.norm(b[s])


def compare(C, so, b, **kw):
    o = O3.Solver3(so, **kw)
    xo, ho = o.solve(b)
    g = C.Solver3(dev(so), **kw)
    assert g.nlevels == o.nlevels
    xg, hg = g.solve(dev(b))
    torch.cuda.synchronize()
    return xo, ho, xg.cpu().numpy(), hg


@pytest.mark.parametrize("n,fe", [(31, False), (24, False), (17, True), (20, True)])
def test_point_solve_matches_oracle(C, n, fe):
    if fe:
        so = G.fe3(n, n, n)
        b = G.rhs_uniform3(n, n, n, 5)
    else:
        so, b, _ = G.poisson3_sine(n)
    xo, ho, xg, hg = compare(C, so, b, relaxation="point", maxiter=20, tol=1e-12)
    assert len(hg) == len(ho)
    assert np.abs(hg - ho).max() <= 1e-12
    assert rel_l2(xg, xo) <= 1e-12


@pytest.mark.parametrize("relax,n", [("plane-xy", 31), ("plane-xy", 24), ("plane-xz", 19),
                                     ("plane-yz", 19), ("plane-xyz", 15)])
def test_plane_solve_matches_oracle(C, relax, n):
    so, b, _ = G.heterogeneous_7pt(n, seed=3, block=4)
    xo, ho, xg, hg = compare(C, so, b, relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=12,
                             tol=1e-10)
    assert len(hg) == len(ho)
    assert np.abs(hg - ho).max() <= 1e-12
    assert rel_l2(xg, xo) <= 1e-10
    if relax in ("plane-xy", "plane-xyz"):  # kz is the weak direction: xy planes hold the strong
        assert hg[-1] < 1e-10                # couplings; xz / yz planes do not and converge slowly


def test_poisson3_reference_test(C):
    """test/3d/test_poisson.cc:73-104: 200^3 Poisson, default solver; ||r|| < 1e-8 and the
    error against the continuous solution < 1e-4."""
    n = 200
    so, b, exact = G.poisson3_sine(n)
    s = C.Solver3(dev(so))
    x, hist = s.solve(dev(b))
    r = torch.zeros_like(x)
    nrm = C.cdr3.residual_norm(s.so, dev(b), x, r)
    assert nrm < 1e-8
    assert np.abs(x.cpu().numpy() - exact).max() < 1e-4


@pytest.mark.parametrize("updown", [0, 1])
@pytest.mark.parametrize("family", ["xy", "xz", "yz"])
def test_exact_plane_solves(C, family, updown):
    """test/3d/test_planes.cc restated: with exact plane solves (point, 50 cycles, tol 1e-30)
    the planes of the colour solved last solve their plane systems: the even planes going DOWN,
    the odd planes going UP (include/cedar/3d/relax_planes.h:44-56)."""
    from tests.test_oracle3 import stencil_matrix
    dims = {"xy": (31, 35, 5), "xz": (31, 5, 35), "yz": (5, 31, 35)}[family]
    nx, ny, nz = dims
    so = G.poisson3(nx, ny, nz)
    kk, jj, ii = nz + 2, ny + 2, nx + 2
    K, J, I = np.meshgrid(np.arange(kk), np.arange(jj), np.arange(ii), indexing="ij")
    x = np.zeros((kk, jj, ii))
    x[1:-1, 1:-1, 1:-1] = (K * nx * ny + J * nx + I)[1:-1, 1:-1, 1:-1].astype(float)
    b = x.copy()
    plane = C.cdr3.plane_settings(relaxation="point", maxiter=50, tol=1e-30)
    g = C.Solver3(dev(so), relaxation="plane-" + family, num_levels=2, plane=plane)
    xg = dev(x)
    g.relax_planes(0, family, xg, dev(b), updown)
    xr = xg.cpu().numpy()
    hx, hy, hz = 1.0 / (nx + 1), 1.0 / (ny + 1), 1.0 / (nz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    d = 2 * xh + 2 * yh + 2 * zh
    st = {"xy": [[0, -yh, 0], [-xh, d, -xh], [0, -yh, 0]],
          "xz": [[0, -zh, 0], [-xh, d, -xh], [0, -zh, 0]],
          "yz": [[0, -zh, 0], [-yh, d, -yh], [0, -zh, 0]]}[family]
    _, _, npl = O3.plane_dims(family, ii, jj, kk)
    for ipl in range(2 if updown == 0 else 1, npl + 1, 2):
        b2 = O3.copy_rhs(so, xr, b, family, ipl)
        n1, n2 = b2.shape[1] - 2, b2.shape[0] - 2
        sol = sla.spsolve(stencil_matrix(n1, n2, st).tocsc(), b2[1:-1, 1:-1].ravel())
        got = O3.copy32(xr, family, ipl)[1:-1, 1:-1].ravel()
        assert np.abs(got - sol).max() < 1e-8, ipl


def test_setup_hierarchy_matches_oracle(C):
    so, b, _ = G.heterogeneous_7pt(23, seed=5, block=4)
    o = O3.Solver3(so)
    g = C.Solver3(dev(so))
    for lvl in range(1, o.nlevels):
        oso, oci = o.level(lvl)
        gso, gci = g.level_arrays(lvl)
        assert np.array_equal(gci, oci), lvl
        assert np.array_equal(gso, oso), lvl


def test_errors_become_status_codes(C):
    so = G.poisson3(15, 15, 15)
    with pytest.raises(C.CedarHipError) as e:
        C.Solver3(dev(so), num_levels=9)
    assert e.value.status == 6
    s = C.Solver3(dev(so))
    b = G.rhs_uniform3(15, 15, 15, 1)
    b[3, 3, 3] = np.nan
    with pytest.raises(C.CedarHipError) as e:
        s.solve(dev(b))
    assert e.value.status == 4


@pytest.mark.parametrize("relax", ["point", "plane-xy"])
def test_fcycle_matches_oracle(C, relax):
    so, b, _ = G.heterogeneous_7pt(23, seed=8, block=4)
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=3, cycle="f")
    xo, ho, xg, hg = compare(C, so, b, **kw)
    assert len(hg) == len(ho) == 3
    assert np.abs(hg - ho).max() <= 1e-12
    assert rel_l2(xg, xo) <= 1e-10


@pytest.mark.parametrize("relax,n", [("point", 31), ("plane-xy", 31), ("plane-xyz", 15)])
def test_mixed_precision_3d(C, relax, n):
    """FP32 3D cycle (FP32 plane solves) inside FP64 refinement: reaches the FP64 tolerance in
    about the FP64 cycle count, solution within 1e-9 of the FP64 one."""
    if relax == "point":
        so, b, _ = G.poisson3_sine(n)
    else:
        so, b, _ = G.heterogeneous_7pt(n, seed=3, block=4)
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10)
    g64 = C.Solver3(dev(so), **kw)
    x64, h64 = g64.solve(dev(b))
    g32 = C.Solver3(dev(so), precision="mixed", **kw)
    x32, h32 = g32.solve(dev(b))
    assert h32[-1] < 1e-10
    assert abs(len(h32) - len(h64)) <= 1
    assert rel_l2(x32.cpu().numpy(), x64.cpu().numpy()) < 1e-9


@pytest.mark.parametrize("relax,down,up", [
    # include/cedar/multilevel.h:179-183 (pre) and :208-212 (post)
    ("plane-xyz", ("xy", "yz", "xz"), ("xz", "yz", "xy")), ("plane-xy", ("xy",), ("xy",))])
def test_plane_cycle_is_the_family_sweeps_in_reference_order(C, relax, down, up):
    """One two-level V(1,1) plane cycle of the GPU solver equals, bit for bit, the GPU family
    sweeps in the reference's order around the GPU residual, restriction, coarse band solve and
    interp_add (each of which is pinned to the oracle on its own)."""
    K = C.cdr3
    so, b, _ = G.heterogeneous_7pt(13, seed=8, block=4, sigma=1.0, aniso=(1.0, 0.3, 0.1))
    g = C.Solver3(dev(so), relaxation=relax, nrelax_pre=1, nrelax_post=1, num_levels=2)
    x0 = np.zeros_like(b)
    x0[1:-1, 1:-1, 1:-1] = np.random.default_rng(9).standard_normal(tuple(n - 2 for n in b.shape))
    want = g.vcycle(dev(x0), dev(b))
    db, dso = dev(b), dev(so)
    x = dev(x0)
    for d in down:
        g.relax_planes(0, d, x, db, 0)
    r = torch.zeros_like(x)
    K.residual(dso, db, x, r)
    so1, ci1 = g.level_arrays(1)
    so1, ci1 = dev(so1), dev(ci1)
    bc = torch.zeros(so1.shape[1:], dtype=torch.float64, device="cuda")
    K.restrict(ci1, r, bc)
    xc = K.solve_cg(torch.zeros_like(bc), bc, K.setup_cg(so1))
    K.interp_add(x, xc, r, dso, ci1)
    for d in up:
        g.relax_planes(0, d, x, db, 1)
    torch.cuda.synchronize()
    assert torch.equal(x, want)
    # and the oracle's cycle agrees to rounding (its plane solves use the sequential line order)
    o = O3.Solver3(so, relaxation=relax, nrelax_pre=1, nrelax_post=1, num_levels=2)
    xo = o.vcycle(x0, b)
    assert rel_l2(want.cpu().numpy(), xo) <= 1e-12


def test_c3_shaped_plane_solve_history(C):
    """The C3 problem family at 63^3 (7-point FV, (kx, ky, kz) = (1, 1, 1e-3) k, lognormal k with
    sigma 1.5 on 8^3 blocks, plane-xy V(1,1) to 1e-10), FP64 GPU against the oracle: the same
    cycle count, every cycle's relative residual within 1e-12 (absolute, of rho = |r_i| / |r_0|),
    the solution within 1e-10; the largest per-cycle relative deviation |drho_i| / rho_i is
    printed (the plane solves' line scans round differently from the sequential substitution)."""
    so, b, _ = G.heterogeneous_7pt(63, seed=3, block=8, sigma=1.5, aniso=(1.0, 1.0, 1e-3))
    xo, ho, xg, hg = compare(C, so, b, relaxation="plane-xy", nrelax_pre=1, nrelax_post=1,
                             maxiter=60, tol=1e-10)
    rel = np.abs(hg - ho) / ho
    print(f"C3-shaped 63^3: {len(hg)} cycles (oracle {len(ho)}), max |drho| {np.abs(hg - ho).max():.2e}, "
          f"max |drho_i|/rho_i {rel.max():.2e}, rel l2 {rel_l2(xg, xo):.2e}", flush=True)
    assert len(hg) == len(ho) and hg[-1] < 1e-10
    assert np.abs(hg - ho).max() <= 1e-12
    assert rel_l2(xg, xo) <= 1e-10

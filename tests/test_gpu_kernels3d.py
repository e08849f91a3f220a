"""3D HIP kernels (C ABI) against the FP64 oracle, which is pinned bitwise to the reference's
Fortran kernels (tests/test_oracle3.py).  Every kernel is bitwise except the norm reduction
(blocked sum, compared to 1e-13) and the plane relaxation sweep, which nests 2D line relaxation
(a parallel scan; compared to 1e-11 relative)."""
import numpy as np
import pytest

from cedar_amd import gallery as G
from oracle import oracle3 as O3
from tests.golden.gen3 import rand_so3

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = [(5, 5, 5), (6, 7, 8), (8, 6, 7), (9, 9, 9), (16, 12, 10), (17, 8, 13), (33, 20, 18)]


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def case(dims, ns, seed):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    so = rand_so3(rng, nx, ny, nz, ns)
    shp = (nz + 2, ny + 2, nx + 2)
    return so, rng.standard_normal(shp), rng.standard_normal(shp), rng.standard_normal(shp), rng


@pytest.mark.parametrize("dims", SIZES)
@pytest.mark.parametrize("ns", [4, 14])
def test_pointwise_kernels_bitwise(C, dims, ns):
    K = C.cdr3
    so, q, qf, res, rng = case(dims, ns, 100 + sum(dims) + ns)
    dso = dev(so)
    r = torch.zeros_like(dev(q))
    K.residual(dso, dev(qf), dev(q), r)
    assert np.array_equal(host(r), O3.residual(so, qf, q))
    sor = torch.zeros((2,) + q.shape, dtype=torch.float64, device="cuda")
    K.setup_recip(dso, sor)
    osor = O3.setup_recip(so)
    assert np.array_equal(host(sor), osor)
    for ud in (K.DOWN, K.UP):
        x = dev(q)
        K.relax_gs(dso, dev(qf), x, sor, ud)
        assert np.array_equal(host(x), O3.relax_gs(so, qf, q, osor, ud)), ud


@pytest.mark.parametrize("dims", SIZES)
@pytest.mark.parametrize("ns", [4, 14])
def test_transfer_and_setup_kernels_bitwise(C, dims, ns):
    K = C.cdr3
    so, q, qf, res, rng = case(dims, ns, 200 + sum(dims) + ns)
    dso = dev(so)
    oci = O3.setup_interp(so)
    ci = torch.zeros(oci.shape, dtype=torch.float64, device="cuda")
    K.setup_interp(dso, ci)
    assert np.array_equal(host(ci), oci)
    soc = torch.zeros((14,) + oci.shape[1:], dtype=torch.float64, device="cuda")
    K.setup_galerkin(dso, ci, soc)
    assert np.array_equal(host(soc), O3.setup_galerkin(so, oci))
    qc = torch.zeros(oci.shape[1:], dtype=torch.float64, device="cuda")
    K.restrict(ci, dev(q), qc)
    assert np.array_equal(host(qc), O3.restrict(oci, q))
    qcv = rng.standard_normal(oci.shape[1:])
    x = dev(q)
    dres = dev(res)
    K.interp_add(x, dev(qcv), dres, dso, ci)
    oq, _ = O3.interp_add(q, qcv, res, so, oci)
    assert np.array_equal(host(x), oq)
    assert np.array_equal(host(dres), res)  # the GPU kernel leaves res unchanged


@pytest.mark.parametrize("dims,ns", [((3, 3, 3), 4), ((4, 3, 5), 14), ((5, 5, 4), 14),
                                     ((4, 4, 4), 14)])
def test_coarse_solve_bitwise(C, dims, ns):
    K = C.cdr3
    so, q, qf, res, rng = case(dims, ns, 300 + sum(dims))
    dso = dev(so)
    abd = K.setup_cg(dso)
    oabd, info = O3.setup_cg(so)
    assert info == 0
    assert np.array_equal(host(abd), oabd)
    x = torch.zeros_like(dev(qf))
    K.solve_cg(x, dev(qf), abd)
    assert np.array_equal(host(x), O3.solve_cg(oabd, qf))


def test_residual_norm(C):
    so, q, qf, res, rng = case((40, 33, 29), 14, 5)
    r = torch.zeros_like(dev(q))
    nrm = C.cdr3.residual_norm(dev(so), dev(qf), dev(q), r)
    assert abs(nrm - O3.norm(O3.residual(so, qf, q))) <= 1e-13 * nrm


@pytest.mark.parametrize("family,dims,ns", [("xy", (31, 35, 5), 4), ("xy", (20, 17, 9), 14),
                                            ("xz", (31, 5, 35), 4), ("xz", (18, 9, 16), 14),
                                            ("yz", (5, 31, 35), 4), ("yz", (9, 16, 18), 14)])
def test_plane_sweep_matches_oracle(C, family, dims, ns):
    """One plane-relaxation sweep with the default plane solver (line-xy, one V(2,1) cycle),
    both directions, against the oracle's nested 2D solves."""
    nx, ny, nz = dims
    so = G.fe3(nx, ny, nz) if ns == 14 else G.heterogeneous_7pt(max(dims), block=4)[0][
        :, :nz + 2, :ny + 2, :nx + 2].copy()
    if ns == 4:
        so = G.fv7_diffusion(G.block_lognormal3(nx, ny, nz, 4, 1.0, 11), (1.0, 1.0, 0.1))
    rng = np.random.default_rng(9)
    b = G.rhs_uniform3(nx, ny, nz, 12)
    x0 = np.zeros_like(b)
    x0[1:-1, 1:-1, 1:-1] = rng.standard_normal((nz, ny, nx))
    o = O3.Solver3(so, relaxation="plane-" + family, num_levels=2)
    g = C.Solver3(dev(so), relaxation="plane-" + family, num_levels=2)
    for ud in (0, 1):
        xo = o.relax_planes(0, family, x0, b, ud)
        xg = dev(x0)
        g.relax_planes(0, family, xg, dev(b), ud)
        xg = host(xg)
        scale = np.abs(xo).max()
        assert np.abs(xg - xo).max() <= 1e-11 * scale, ud

"""The plugin registry of the 2D and 3D solvers (kernel_manager: add / set / setup / run,
include/cedar/kernel_manager.h:38-72; interface list of ser_kernels,
include/cedar/kernel_registry.h:23-35), and the FP32 3D transfer and coarse-solve entry points
the mixed-precision 3D cycle uses.

Every interface a solve uses is registered with the "hip" implementation; selecting it by name
and running a cycle gives bitwise the cycle of a solver left at its defaults; an unknown
interface or implementation is rejected."""
import numpy as np
import pytest

from cedar_amd import gallery as G
from oracle import oracle3 as O3
from tests.golden.gen3 import rand_so3

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

IFACES2 = {"residual", "point-relax", "line-relax-x", "line-relax-y", "setup-interp",
           "coarsen-op", "restriction", "interp-add", "solve-cg"}
IFACES3 = {"residual (3D)", "point-relax (3D)", "plane-relax-xy", "plane-relax-xz",
           "plane-relax-yz", "setup-interp (3D)", "coarsen-op (3D)", "restriction (3D)",
           "interp-add (3D)", "solve-cg (3D)"}


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def test_registry_2d_select_hip_by_name(C):
    so, b, _ = G.heterogeneous_9pt(95, seed=12, block=8, sigma=1.0)
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=8, tol=1e-12)
    s0 = C.Solver(dev(so), **kw)
    s1 = C.Solver(dev(so), **kw)
    reg = s1.kernels()
    assert set(reg) == IFACES2
    assert all(v["impls"] == ["hip"] and v["default"] == "hip" for v in reg.values())
    for name in reg:
        s1.set_kernel(name, "hip")
    with pytest.raises(C.CedarHipError):
        s1.set_kernel("residual", "system")
    with pytest.raises(C.CedarHipError):
        s1.set_kernel("no-such-interface", "hip")
    x0, h0 = s0.solve(dev(b))
    x1, h1 = s1.solve(dev(b))
    assert np.array_equal(h0, h1) and torch.equal(x0, x1)


@pytest.mark.parametrize("relax,precision", [("plane-xyz", "f64"), ("plane-xy", "mixed"),
                                             ("point", "f64")])
def test_registry_3d_select_hip_by_name(C, relax, precision):
    n = 23
    so, b, _ = G.heterogeneous_7pt(n, seed=13, block=4, sigma=1.0, aniso=(1.0, 1.0, 0.1))
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=6, tol=1e-12,
              precision=precision)
    s0 = C.Solver3(dev(so), **kw)
    s1 = C.Solver3(dev(so), **kw)
    reg = s1.kernels()
    assert set(reg) == IFACES3
    assert all(v["impls"] == ["hip"] and v["default"] == "hip" for v in reg.values())
    for name in reg:
        s1.set_kernel(name, "hip")
    with pytest.raises(C.CedarHipError):
        s1.set_kernel("plane-relax-xy", "system")
    # one cycle through the registry, then a full solve: bitwise those of the default handle
    x0 = s0.vcycle(torch.zeros_like(dev(b)), dev(b))
    x1 = s1.vcycle(torch.zeros_like(dev(b)), dev(b))
    assert torch.equal(x0, x1)
    y0, h0 = s0.solve(dev(b))
    y1, h1 = s1.solve(dev(b))
    assert np.array_equal(h0, h1) and torch.equal(y0, y1)


@pytest.mark.parametrize("dims", [(9, 9, 9), (16, 12, 10), (17, 8, 13)])
def test_fp32_transfers_3d(C, dims):
    """FP32 restriction and interp_add against the FP64 oracle on the same (float-rounded)
    inputs: the float kernels follow the same operation order, so they agree to FP32 rounding."""
    K = C.cdr3
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims))
    so = rand_so3(rng, nx, ny, nz, 14).astype(np.float32).astype(np.float64)
    shp = (nz + 2, ny + 2, nx + 2)
    q = rng.standard_normal(shp).astype(np.float32).astype(np.float64)
    res = rng.standard_normal(shp).astype(np.float32).astype(np.float64)
    ii, jj, kk = nx + 2, ny + 2, nz + 2
    iic, jjc, kkc = K.coarse_dims3(ii, jj, kk)
    ci = O3.setup_interp(so, (iic, jjc, kkc)).astype(np.float32).astype(np.float64)
    qc_ref = O3.restrict(ci, q)
    qc = torch.zeros(qc_ref.shape, dtype=torch.float32, device="cuda")
    K.restrict(dev(ci, torch.float32), dev(q, torch.float32), qc)
    torch.cuda.synchronize()
    err = np.abs(qc.cpu().numpy() - qc_ref).max() / np.abs(qc_ref).max()
    assert err <= 2e-6, err
    xc = rng.standard_normal(qc_ref.shape).astype(np.float32).astype(np.float64)
    out_ref, _ = O3.interp_add(q, xc, res, so, ci)
    out = dev(q, torch.float32)
    K.interp_add(out, dev(xc, torch.float32), dev(res, torch.float32), dev(so, torch.float32),
                 dev(ci, torch.float32))
    torch.cuda.synchronize()
    err = np.abs(out.cpu().numpy() - out_ref).max() / np.abs(out_ref).max()
    assert err <= 2e-6, err


def test_fp32_coarse_solve_3d(C):
    K = C.cdr3
    so = G.poisson3(3, 4, 3)
    abd = K.setup_cg(dev(so))
    qf = np.random.default_rng(3).standard_normal(so.shape[1:]).astype(np.float32)
    q64 = K.solve_cg(torch.zeros(qf.shape, dtype=torch.float64, device="cuda"),
                     dev(qf.astype(np.float64)), abd)
    q32 = K.solve_cg(torch.zeros(qf.shape, dtype=torch.float32, device="cuda"),
                     dev(qf, torch.float32), abd)
    torch.cuda.synchronize()
    a, b = q64.cpu().numpy(), q32.cpu().numpy().astype(np.float64)
    # the solve runs in FP64 on the widened right-hand side; only the result is rounded
    assert np.array_equal(a.astype(np.float32).astype(np.float64), b)

"""Differentiated solves (cedar_amd.adjoint) on the GPU: the Q1 assembly kernel equals the
gallery's assembly, and dL/dtheta from the adjoint solve plus the gradient kernel matches the FP64
CPU gradient (oracle forward and adjoint solves, tests/adjoint_ref.py) and central finite
differences (h = 1e-6) of the GPU loss on 32 random cells."""
import numpy as np
import pytest

from cedar_amd import gallery as G
from tests import adjoint_ref as R

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    import cedar_amd.adjoint as adj
    return adj


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def test_assembly_matches_gallery(A):
    rng = np.random.default_rng(3)
    for shape in ((5, 5), (7, 4)):
        ks = [np.exp(rng.standard_normal((shape[1] + 1, shape[0] + 1))) for _ in range(3)]
        so = A.assemble(dev(np.stack(ks))).cpu().numpy()
        for t, k in enumerate(ks):
            assert np.array_equal(so[t], G.q1_diffusion(k)), (shape, t)


def test_cell_gradient_kernel(A):
    rng = np.random.default_rng(4)
    n = 21
    u = np.stack([R.zero_boundary(rng.standard_normal((n + 2, n + 2))) for _ in range(2)])
    lam = np.stack([R.zero_boundary(rng.standard_normal((n + 2, n + 2))) for _ in range(2)])
    g = A.cell_gradient(dev(u), dev(lam)).cpu().numpy()
    for t in range(2):
        ref = R.cell_gradient(u[t], lam[t])
        assert np.allclose(g[t], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def loss_and_grad(A, theta, f, uobs, **kw):
    th = dev(theta).requires_grad_(True)
    u = A.solve(th, dev(f), **kw)
    loss = 0.5 * ((u - dev(uobs)) ** 2)[..., 1:-1, 1:-1].sum()
    loss.backward()
    return float(loss.item()), th.grad.cpu().numpy()


@pytest.mark.parametrize("precision,tol,bound", [("f64", 1e-13, 1e-9), ("mixed", 1e-12, 1e-6)])
def test_gradient_matches_cpu_fp64(A, precision, tol, bound):
    n, nb = 63, 3
    probs = [R.problem(n, 100 + 10 * t) for t in range(nb)]
    theta, f, uobs = (np.stack([p[i] for p in probs]) for i in range(3))
    _, g = loss_and_grad(A, theta, f, uobs, precision=precision, tol=tol, maxiter=80)
    for t in range(nb):
        gc, _, _ = R.gradient(theta[t], f[t], uobs[t])
        err = np.abs(g[t] - gc).max() / np.abs(gc).max()
        print(f"problem {t} {precision}: max |g - g_cpu| / max |g_cpu| = {err:.3e}")
        assert err <= bound, (t, err)


def test_gradient_matches_finite_differences(A):
    n = 31
    theta, f, uobs = R.problem(n, 21)
    kw = dict(precision="f64", tol=1e-13, maxiter=80)
    _, g = loss_and_grad(A, theta[None], f[None], uobs[None], **kw)
    g = g[0]
    rng = np.random.default_rng(22)
    cells = rng.integers(0, n + 1, size=(32, 2))
    h = 1e-6
    scale = np.abs(g).max()
    for cj, ci in cells:
        tp, tm = theta.copy(), theta.copy()
        tp[cj, ci] += h
        tm[cj, ci] -= h
        lp, _ = loss_and_grad(A, tp[None], f[None], uobs[None], **kw)
        lm, _ = loss_and_grad(A, tm[None], f[None], uobs[None], **kw)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[cj, ci]) <= 1e-6 * scale + 1e-5 * abs(fd), (cj, ci, fd, g[cj, ci])

"""Differentiated solves (cedar_amd.adjoint): the GPU operator assembly equals the gallery's, and
dL/dtheta from the adjoint solve plus the gradient kernel matches central finite differences
(h = 1e-6) and an FP64 CPU gradient built from oracle solves."""
import numpy as np
import pytest

from cedar_amd import gallery as G
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

K6 = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]], float)


@pytest.fixture(scope="module")
def A():
    import cedar_amd.adjoint as adj
    return adj


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def cpu_gradient(u, lam):
    """dL/dk per cell from state and adjoint (numpy), boundary nodes zero."""
    def pad(v):
        w = v.copy()
        w[0, :] = w[-1, :] = 0.0
        w[:, 0] = w[:, -1] = 0.0
        return w
    u, lam = pad(u), pad(lam)
    corners = lambda v: np.stack([v[:-1, :-1], v[:-1, 1:], v[1:, 1:], v[1:, :-1]])
    U, L = corners(u), corners(lam)
    return -np.einsum("a...,ab,b...->...", L, K6, U) / 6.0


def problem(n, seed):
    theta = G.smooth_theta(n, seed)
    f = G.rhs_uniform(n, n, seed + 17)
    return theta, f


def test_assembly_matches_gallery(A):
    rng = np.random.default_rng(3)
    ks = [np.exp(rng.standard_normal((n + 1, n + 1))) for n in (5, 5)]
    so = A.assemble(dev(np.stack(ks))).cpu().numpy()
    for t, k in enumerate(ks):
        assert np.array_equal(so[t], G.q1_diffusion(k))


def loss_and_grad(A, theta, f, uobs, **kw):
    th = dev(theta).requires_grad_(True)
    u = A.solve(th, dev(f), **kw)
    loss = 0.5 * ((u - dev(uobs)) ** 2)[..., 1:-1, 1:-1].sum()
    loss.backward()
    return float(loss.item()), th.grad.cpu().numpy(), u.detach().cpu().numpy()


@pytest.mark.parametrize("precision,tol", [("f64", 1e-14), ("mixed", 1e-10)])
def test_gradient_matches_cpu_fp64(A, precision, tol):
    n, nb = 63, 3
    ths, fs, obs = [], [], []
    for t in range(nb):
        theta, f = problem(n, 100 + t)
        th2, _ = problem(n, 200 + t)
        so2 = G.q1_diffusion(np.exp(th2))
        uo, _ = O.Solver(so2, relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=60,
                         tol=1e-13).solve(f)
        uo = uo + 1e-3 * np.random.default_rng(300 + t).standard_normal(uo.shape)
        ths.append(theta)
        fs.append(f)
        obs.append(uo)
    theta, f, uobs = np.stack(ths), np.stack(fs), np.stack(obs)
    loss, g, u = loss_and_grad(A, theta, f, uobs, precision=precision, tol=tol, maxiter=80)
    for t in range(nb):
        k = np.exp(theta[t])
        so = G.q1_diffusion(k)
        s = O.Solver(so, relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=80, tol=1e-14)
        uc, _ = s.solve(f[t])
        gu = np.zeros_like(uc)
        gu[1:-1, 1:-1] = (uc - uobs[t])[1:-1, 1:-1]
        lam, _ = s.solve(gu)
        gc = cpu_gradient(uc, lam) * k
        err = np.abs(g[t] - gc).max() / np.abs(gc).max()
        assert err <= (1e-10 if precision == "f64" else 1e-6), (t, err)


def test_gradient_matches_finite_differences(A):
    n = 47
    theta, f = problem(n, 7)
    th2, _ = problem(n, 8)
    uo, _ = O.Solver(G.q1_diffusion(np.exp(th2)), relaxation="line-xy", maxiter=60,
                     tol=1e-13).solve(f)
    uobs = uo + 1e-3 * np.random.default_rng(9).standard_normal(uo.shape)
    kw = dict(precision="f64", tol=1e-14, maxiter=80)
    _, g, _ = loss_and_grad(A, theta[None], f[None], uobs[None], **kw)
    rng = np.random.default_rng(11)
    cells = [tuple(c) for c in rng.integers(0, n + 1, size=(8, 2))]
    h = 1e-6
    for (cj, ci) in cells:
        tp, tm = theta.copy(), theta.copy()
        tp[cj, ci] += h
        tm[cj, ci] -= h
        lp = loss_and_grad(A, tp[None], f[None], uobs[None], **kw)[0]
        lm = loss_and_grad(A, tm[None], f[None], uobs[None], **kw)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[0, cj, ci]) <= 1e-5 * np.abs(g).max() + 1e-4 * abs(fd), (cj, ci, fd,
                                                                                   g[0, cj, ci])

"""CPU checks of the adjoint-gradient reference (tests/adjoint_ref.py): the element formula
reproduces the bilinear form of the assembled operator, and the FP64 adjoint gradient agrees with
central finite differences of the loss (h = 1e-6) on 32 random cells."""
import numpy as np

from cedar_amd import gallery as G
from oracle import oracle as O
from tests import adjoint_ref as R


def apply_A(k, u):
    """A(k) u through the oracle's residual with a zero right-hand side."""
    so = G.q1_diffusion(k)
    return -O.residual(so, np.zeros_like(u), u)


def test_element_formula_is_the_bilinear_form():
    # lambda^T A(k) u = sum_e k_e lambda_e^T K u_e, so -sum_e k_e g_e(u, lambda) must equal it
    rng = np.random.default_rng(5)
    n = 13
    k = np.exp(rng.standard_normal((n + 1, n + 1)))
    u = R.zero_boundary(rng.standard_normal((n + 2, n + 2)))
    lam = R.zero_boundary(rng.standard_normal((n + 2, n + 2)))
    lhs = float((lam * apply_A(k, u))[1:-1, 1:-1].sum())
    rhs = -float((k * R.cell_gradient(u, lam)).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # linearity in k: one cell changed by dk changes lambda^T A u by -dk g_e
    e = (4, 7)
    dk = np.zeros_like(k)
    dk[e] = 0.37
    d = float((lam * (apply_A(k + dk, u) - apply_A(k, u)))[1:-1, 1:-1].sum())
    assert abs(d + 0.37 * R.cell_gradient(u, lam)[e]) <= 1e-12 * max(abs(d), 1e-30) + 1e-14


def test_operator_is_symmetric():
    rng = np.random.default_rng(6)
    n = 9
    k = np.exp(rng.standard_normal((n + 1, n + 1)))
    u = R.zero_boundary(rng.standard_normal((n + 2, n + 2)))
    v = R.zero_boundary(rng.standard_normal((n + 2, n + 2)))
    a = float((v * apply_A(k, u))[1:-1, 1:-1].sum())
    b = float((u * apply_A(k, v))[1:-1, 1:-1].sum())
    assert abs(a - b) <= 1e-12 * abs(a)


def test_adjoint_gradient_matches_finite_differences():
    n = 31
    theta, f, uobs = R.problem(n, 21)
    g, _, _ = R.gradient(theta, f, uobs)
    rng = np.random.default_rng(22)
    cells = rng.integers(0, n + 1, size=(32, 2))
    h = 1e-6
    scale = np.abs(g).max()
    for cj, ci in cells:
        tp, tm = theta.copy(), theta.copy()
        tp[cj, ci] += h
        tm[cj, ci] -= h
        fd = (R.loss(tp, f, uobs) - R.loss(tm, f, uobs)) / (2 * h)
        assert abs(fd - g[cj, ci]) <= 1e-6 * scale + 1e-5 * abs(fd), (cj, ci, fd, g[cj, ci])

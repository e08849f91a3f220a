"""Whole-solve parity of the GPU solver (C ABI) with the FP64 oracle and Cedar's known answer.

Bar (north star): FP64 GPU residual history within 1e-12 of the oracle's per iteration
(|rho_gpu - rho_cpu|, rho normalised by ||r_0||), the same iteration count, and the solution
within 1e-10 relative l2.  The mixed-precision solve must stop at the same iteration count
with its FP64 residual below the tolerance.
"""
import json
import os

import numpy as np
import pytest

from cedar_amd import gallery as G
from oracle import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def rel_l2(a, b):
    return np.linalg.norm((a - b)[1:-1, 1:-1]) / np.linalg.norm(b[1:-1, 1:-1])


def compare(C, so, b, **kw):
    s_cpu = O.Solver(so, **{k: v for k, v in kw.items() if k != "precision"})
    x_cpu, h_cpu = s_cpu.solve(b)
    s_gpu = C.Solver(dev(so), **kw)
    assert s_gpu.nlevels == s_cpu.nlevels
    x_gpu, h_gpu = s_gpu.solve(dev(b))
    x_gpu = x_gpu.cpu().numpy()
    return x_cpu, h_cpu, x_gpu, h_gpu, s_gpu


def test_readme_known_answer(C):
    with open(os.path.join(HERE, "golden", "readme_history.json")) as f:
        g = json.load(f)
    so, b, exact = G.readme_problem(400)
    s = C.Solver(dev(so), relaxation="point", nrelax_pre=1, nrelax_post=1, maxiter=10, tol=1e-10,
                 num_levels=g["levels"])
    x, hist = s.solve(dev(b))
    assert [float("%.6g" % h) for h in hist] == g["rho"]
    d = (exact - x.cpu().numpy())[1:-1, 1:-1].ravel()
    assert float("%.6g" % abs(d[np.argmax(np.abs(d))])) == g["solution_norm"]


@pytest.mark.parametrize("nx,ny,relax", [(200, 200, "point"), (800, 200, "line-x"),
                                          (200, 800, "line-y")])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_reference_poisson_tests(C, nx, ny, relax, precision):
    """test/2d/test_poisson.cc:64-159 (Isotropic, StretchX, StretchY) with the reference's default
    settings: ||r||_2 < 1e-8 after the solve and the error against the continuous solution
    sin(2 pi x) sin(2 pi y) below 1e-4."""
    so, b, exact = G.sine_problem(nx, ny)
    x, _ = C.Solver(dev(so), relaxation=relax, precision=precision).solve(dev(b))
    x = x.cpu().numpy()
    assert O.norm(O.residual(so, b, x)) < 1e-8
    assert np.abs(x - exact)[1:-1, 1:-1].max() < 1e-4


@pytest.mark.parametrize("relax,n,nine", [("point", 511, False), ("point", 200, True),
                                          ("line-xy", 255, True), ("line-x", 130, True),
                                          ("line-y", 97, False), ("line-xy", 256, False)])
def test_fp64_solve_matches_oracle(C, relax, n, nine):
    if nine:
        so, b, _ = G.heterogeneous_9pt(n, seed=3, block=8, sigma=1.5)
    else:
        so, b = G.poisson_uniform_rhs(n, seed=1)
    x_cpu, h_cpu, x_gpu, h_gpu, _ = compare(C, so, b, relaxation=relax, nrelax_pre=1,
                                            nrelax_post=1, maxiter=40, tol=1e-10)
    assert len(h_gpu) == len(h_cpu)
    assert np.abs(h_gpu - h_cpu).max() <= 1e-12
    assert rel_l2(x_gpu, x_cpu) <= 1e-10
    assert h_gpu[-1] < 1e-10


def test_fp64_point_solve_is_bitwise_for_pointwise_kernels(C):
    """Point relaxation uses only bitwise-exact kernels except the Galerkin setup and the norm
    reduction; with a 2-level hierarchy on a 5-point operator the coarse operator comes from the
    5-to-9 product, so allow rounding only there."""
    so, b = G.poisson_uniform_rhs(63, seed=5)
    x_cpu, h_cpu, x_gpu, h_gpu, _ = compare(C, so, b, relaxation="point", maxiter=12, tol=1e-14)
    assert np.abs(h_gpu - h_cpu).max() <= 1e-13
    assert rel_l2(x_gpu, x_cpu) <= 1e-12


@pytest.mark.parametrize("relax,n", [("line-xy", 255), ("point", 255)])
def test_mixed_precision_same_iterations(C, relax, n):
    so, b, _ = G.heterogeneous_9pt(n, seed=2, block=16, sigma=2.0)
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=60, tol=1e-10)
    s64 = C.Solver(dev(so), **kw)
    x64, h64 = s64.solve(dev(b))
    s32 = C.Solver(dev(so), precision="mixed", **kw)
    x32, h32 = s32.solve(dev(b))
    assert h32[-1] < 1e-10
    assert len(h32) == len(h64)
    assert rel_l2(x32.cpu().numpy(), x64.cpu().numpy()) < 1e-9


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 4), (5, 17), (17, 5), (2, 9)])
def test_small_and_ragged_grids(C, nx, ny):
    so = G.poisson(nx, ny)
    b = G.rhs_uniform(nx, ny, seed=nx * 100 + ny)
    x_cpu, h_cpu, x_gpu, h_gpu, s = compare(C, so, b, relaxation="point", maxiter=20, tol=1e-12)
    assert len(h_gpu) == len(h_cpu)
    np.testing.assert_allclose(h_gpu, h_cpu, rtol=0, atol=1e-12)
    assert rel_l2(x_gpu, x_cpu) <= 1e-10


def test_setup_hierarchy_matches_oracle(C):
    so, b, _ = G.heterogeneous_9pt(127, seed=9, block=4)
    o = O.Solver(so, relaxation="line-xy")
    s = C.Solver(dev(so), relaxation="line-xy")
    for lvl in range(o.nlevels):
        oso, oci, os0, os1 = o.level(lvl)
        gso, gci, gs0, gs1 = s.level_arrays(lvl)
        scale = np.abs(oso).max()
        assert np.abs(gso - oso).max() <= 1e-12 * scale, lvl
        if lvl > 0:
            assert np.abs(gci - oci).max() <= 1e-12, lvl
        if lvl < o.nlevels - 1:
            assert np.abs(gs0 - os0).max() <= 1e-12 * max(1.0, np.abs(os0).max()), lvl
            assert np.abs(gs1 - os1).max() <= 1e-12 * max(1.0, np.abs(os1).max()), lvl


def test_errors_become_status_codes(C):
    so = G.poisson(31, 31)
    with pytest.raises(C.CedarHipError) as e:
        C.Solver(dev(so), num_levels=9)
    assert e.value.status == 6
    bad = so.copy()
    bad[0, 1:-1, 1:-1] = -1.0
    with pytest.raises(C.CedarHipError) as e:
        C.Solver(dev(bad), relaxation="line-x")
    assert e.value.status == 3
    s = C.Solver(dev(so))
    b = G.rhs_uniform(31, 31, 1)
    b[5, 5] = np.nan
    with pytest.raises(C.CedarHipError) as e:
        s.solve(dev(b))
    assert e.value.status == 4


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_graph_replay_matches_eager(C, precision):
    """Every V-cycle replayed from the captured graph gives the eager result bit for bit, over
    several solves in a row (with eager work and synchronisation in between) and after the
    right-hand side / solution buffers change (re-capture)."""
    so, b, _ = G.heterogeneous_9pt(127, seed=4, block=8, sigma=1.0)
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=30, tol=1e-10,
              precision=precision)
    eager = C.Solver(dev(so), graph=False, **kw)
    graph = C.Solver(dev(so), graph=True, **kw)
    db = dev(b)
    x_e, h_e = eager.solve(db)
    for rep in range(3):
        x_g, h_g = graph.solve(db)
        assert np.array_equal(h_g, h_e), rep
        assert torch.equal(x_g, x_e), rep
        torch.cuda.synchronize()
        _ = (db * 2.0).sum().item()  # unrelated eager work between replays
    db2 = dev(b) * 0.5
    x_e2, h_e2 = eager.solve(db2)
    x_g2, h_g2 = graph.solve(db2, torch.zeros_like(db2))
    assert np.array_equal(h_g2, h_e2)
    assert torch.equal(x_g2, x_e2)


@pytest.mark.parametrize("relax,n,nine,precision", [("point", 127, False, "f64"),
                                                    ("line-xy", 96, True, "f64"),
                                                    ("line-xy", 127, True, "mixed")])
def test_fcycle_matches_oracle(C, relax, n, nine, precision):
    """solver.cycle.type = f (include/cedar/cycle/fcycle.h): each cycle recomputes x from b, so
    the FP64 history repeats and matches the oracle to the usual bar."""
    if nine:
        so, b, _ = G.heterogeneous_9pt(n, seed=6, block=8, sigma=1.0)
    else:
        so, b = G.poisson_uniform_rhs(n, seed=6)
    o = O.Solver(so, relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=3, cycle="f")
    xo, ho = o.solve(b)
    g = C.Solver(dev(so), relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=3, cycle="f",
                 precision=precision)
    xg, hg = g.solve(dev(b))
    assert len(hg) == len(ho) == 3
    if precision == "f64":
        assert np.abs(hg - ho).max() <= 1e-12
        assert rel_l2(xg.cpu().numpy(), xo) <= 1e-10
    else:
        # inside iterative refinement the F-cycle solves for the correction, so the first cycle
        # matches the plain F-cycle up to the FP32 rounding of that correction (rho is relative to
        # |b| here: a few units of 2^-24) and the later ones keep reducing the residual
        assert abs(hg[0] - ho[0]) <= 2.0 ** -22
        assert hg[2] < hg[1] < hg[0]
        # run to the tolerance: the FP64 oracle residual of the returned x meets it
        g = C.Solver(dev(so), relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=60,
                     tol=1e-10, cycle="f", precision="mixed")
        xg, hg = g.solve(dev(b))
        assert hg[-1] < 1e-10 and len(hg) < 60
        rr = O.norm(O.residual(so, b, xg.cpu().numpy())) / O.norm(b)
        assert rr < 1e-10, rr


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_stall_guard_stops_a_solve_that_cannot_reach_tol(C, precision):
    """stall_cycles = 3 with tol = 0: once rounding floors the residual, three cycles without a
    new minimum stop the solve with CDR_ERR_STALLED; the cycles before the stop are those of the
    unguarded solve, and the stopping rule applied to the history gives the same stop."""
    import ctypes
    so, b = G.poisson_uniform_rhs(63, seed=3)
    kw = dict(relaxation="point", nrelax_pre=2, nrelax_post=1, tol=0.0, precision=precision)
    s = C.Solver(dev(so), maxiter=200, stall_cycles=3, **kw)
    with pytest.raises(C.CedarHipError) as ei:
        s.solve(dev(b))
    assert ei.value.status == 7
    h = ei.value.hist
    assert 3 < len(h) < 200
    assert min(h[-3:]) >= min(h[:-3])
    assert h[-4] < 1e-12  # it stalled at the rounding floor, not early
    arr = (ctypes.c_double * len(h))(*h)
    n = ctypes.c_int()
    assert C.load().cdr_hip_stop_rule(arr, len(h), 0.0, 3, ctypes.byref(n)) == 7
    assert n.value == len(h)
    x2, h2 = C.Solver(dev(so), maxiter=len(h) + 4, **kw).solve(dev(b))
    assert len(h2) == len(h) + 4
    assert np.array_equal(h2[:len(h)], h)


@pytest.mark.parametrize("relaxation", ["point", "line-x", "line-y", "line-xy"])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("nx,ny,cycle", [(255, 255, "v"), (200, 131, "v"), (257, 257, "f")])
def test_fused_coarse_cycle_is_bitwise_unfused(relaxation, precision, nx, ny, cycle):
    """The fused coarse-level cycle (one launch for all levels of at most 32 points per side)
    gives bitwise the iterates and residual histories of one launch per step."""
    import cedar_amd
    from cedar_amd import gallery as G
    k = G.block_lognormal(nx + 1, ny + 1, 8, 1.5, 31 + nx, 1e3)
    so = torch.as_tensor(G.q1_diffusion(k)).cuda()
    b = torch.as_tensor(G.rhs_uniform(nx, ny, 5 + ny)).cuda()
    kw = dict(relaxation=relaxation, nrelax_pre=2, nrelax_post=1, maxiter=12, tol=1e-12,
              precision=precision, cycle=cycle)
    x0, h0 = cedar_amd.Solver(so, fused_coarse=False, **kw).solve(b)
    x1, h1 = cedar_amd.Solver(so, fused_coarse=True, **kw).solve(b)
    assert np.array_equal(h0, h1)
    assert torch.equal(x0, x1)


def test_fused_coarse_cycle_batched():
    import cedar_amd
    from cedar_amd import gallery as G
    n, nb = 63, 3
    sos = torch.stack([torch.as_tensor(G.q1_diffusion(G.block_lognormal(n + 1, n + 1, 8, 2.0, 60 + t, 1e4)))
                       for t in range(nb)]).cuda()
    bs = torch.stack([torch.as_tensor(G.rhs_uniform(n, n, 70 + t)) for t in range(nb)]).cuda()
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=30, tol=1e-10,
              precision="mixed")
    x0, h0 = cedar_amd.Solver(sos, fused_coarse=False, **kw).solve(bs)
    x1, h1 = cedar_amd.Solver(sos, fused_coarse=True, **kw).solve(bs)
    for t in range(nb):
        assert np.array_equal(h0[t], h1[t])
    assert torch.equal(x0, x1)

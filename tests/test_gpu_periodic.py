"""Periodic boundaries on the GPU (grid.periodic): every periodic kernel against the oracle's
periodic branches (which equal the reference's Fortran bit for bit, tests/test_oracle_periodic.py),
and periodic solves against the oracle's driver."""
import ctypes
import json

import numpy as np
import pytest

from cedar_amd import gallery as Gal
from oracle import oracle as O
from tests.golden.make_golden_periodic import rand_periodic_so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SIZES = [(8, 8), (7, 9), (12, 6), (33, 20), (64, 31)]


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def host(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("mask", [1, 2, 3])
@pytest.mark.parametrize("ns", [3, 5])
def test_periodic_kernels_bitwise(nx, ny, mask, ns):
    import cedar_amd as C
    print(f"\nperiodic kernels {nx}x{ny} ns={ns} mask={mask}", flush=True)
    rng = np.random.default_rng(nx * 1000 + ny * 10 + mask + ns)
    jpn = O.bc_from_mask(mask)
    so = rand_periodic_so(rng, nx, ny, ns, mask)
    q = rng.standard_normal((ny + 2, nx + 2))
    qf = rng.standard_normal(q.shape)
    sor = O.setup_recip(so)
    for ud in (O.DOWN, O.UP):
        g = C.relax_gs(dev(so), dev(qf), dev(q), dev(sor), ud, jpn=jpn)
        assert np.array_equal(host(g), O.relax_gs(so, qf, q, sor, ud, jpn)), ud
    ci = O.setup_interp(so, jpn=jpn)
    gci = torch.zeros((8,) + ci.shape[1:], dtype=torch.float64, device="cuda")
    C.setup_interp(dev(so), gci, jpn=jpn)
    assert np.array_equal(host(gci), ci)
    soc = O.setup_galerkin(so, ci, jpn)
    gsoc = torch.zeros(soc.shape, dtype=torch.float64, device="cuda")
    C.setup_galerkin(dev(so), dev(ci), gsoc, jpn=jpn)
    assert np.array_equal(host(gsoc), soc)
    qc = torch.zeros(ci.shape[1:], dtype=torch.float64, device="cuda")
    gq = dev(q)
    C.restrict(dev(ci), gq, qc, jpn=jpn)
    assert np.array_equal(host(qc), O.restrict(ci, q, jpn))
    qcr = rng.standard_normal(ci.shape[1:])
    res = rng.standard_normal(q.shape)
    gi = C.interp_add(dev(q), dev(qcr), dev(res), dev(so), dev(ci), jpn=jpn)
    assert np.array_equal(host(gi), O.interp_add(q, qcr, res, so, ci, jpn)[0])


@pytest.mark.parametrize("nx,ny", [(4, 4), (5, 4), (6, 7), (11, 9)])
@pytest.mark.parametrize("mask", [1, 2, 3])
def test_periodic_coarse_solve_bitwise(nx, ny, mask):
    import cedar_amd as C
    jpn = O.bc_from_mask(mask)
    for ns in (3, 5):
        so = rand_periodic_so(np.random.default_rng(nx + ny + mask + ns), nx, ny, ns, mask)
        abd, info = O.setup_cg_periodic(so, jpn)
        assert info == 0
        gabd = C.setup_cg_periodic(dev(so), jpn)
        assert np.array_equal(host(gabd), abd)
        qf = np.random.default_rng(ns).standard_normal((ny + 2, nx + 2))
        q = C.solve_cg_periodic(gabd, dev(qf), torch.zeros((ny + 2, nx + 2), dtype=torch.float64,
                                                            device="cuda"), jpn)
        assert np.array_equal(host(q), O.solve_cg_periodic(abd, qf, jpn))


def _problem(case):
    if case == "xy":
        so, _, _ = Gal.periodic_problem(96, 80, (True, True))
        so[0] += 0.05  # definite (Cedar has no regularisation for the singular fully periodic case)
        b = np.zeros_like(so[0])
        b[1:-1, 1:-1] = np.random.default_rng(5).standard_normal((80, 96))
        b[:, 0], b[:, -1] = b[:, -2], b[:, 1]
        b[0, :], b[-1, :] = b[-2, :], b[1, :]
        return so, b, None, (True, True)
    per = (True, False) if case == "x" else (False, True)
    so, b, exact = Gal.periodic_problem(127, 100, per)
    return so, b, exact, per


@pytest.mark.parametrize("case", ["x", "y", "xy"])
@pytest.mark.parametrize("cycle", ["v", "f"])
def test_periodic_solve_matches_oracle(case, cycle):
    import cedar_amd as C
    print(f"\nperiodic solve {case} {cycle}", flush=True)
    so, b, exact, per = _problem(case)
    kw = dict(relaxation="point", nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10, cycle=cycle)
    xo, ho = O.Solver(so, periodic=per, **kw).solve(b)
    s = C.Solver(dev(so), periodic=per, **kw)
    x, h = s.solve(dev(b))
    assert len(h) == len(ho)
    np.testing.assert_allclose(h, ho, rtol=1e-12, atol=0)
    x = host(x)
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    if exact is not None and cycle == "v":
        assert np.abs(x - exact)[1:-1, 1:-1].max() < 5e-3


@pytest.mark.parametrize("case", ["x", "xy"])
def test_periodic_mixed_precision(case):
    import cedar_amd as C
    so, b, _, per = _problem(case)
    kw = dict(relaxation="point", nrelax_pre=2, nrelax_post=1, maxiter=40, tol=1e-10,
              periodic=per)
    x64, h64 = C.Solver(dev(so), **kw).solve(dev(b))
    x32, h32 = C.Solver(dev(so), precision="mixed", **kw).solve(dev(b))
    assert h32[-1] < 1e-10 and len(h32) <= len(h64) + 2
    assert torch.linalg.norm(x32 - x64) <= 1e-8 * torch.linalg.norm(x64)


def test_periodic_batch_equals_single():
    import cedar_amd as C
    ops, rhs = [], []
    for k in range(3):
        so = rand_periodic_so(np.random.default_rng(40 + k), 40, 36, 5, 1)
        so[0] += 1.0
        ops.append(so)
        r = np.random.default_rng(50 + k).standard_normal(so[0].shape)
        r[:, 0], r[:, -1] = r[:, -2], r[:, 1]
        rhs.append(r)
    kw = dict(relaxation="point", maxiter=20, tol=1e-11, periodic=(True, False))
    xb, hb = C.Solver(dev(np.stack(ops)), **kw).solve(dev(np.stack(rhs)))
    for k in range(3):
        x, h = C.Solver(dev(ops[k]), **kw).solve(dev(rhs[k]))
        assert torch.equal(xb[k], x) and np.array_equal(hb[k], h)


def test_capi_periodic_solve(tmp_path):
    """The handle API honours grid.periodic (the reference's examples/basic-2d-ser/
    periodic-config.json): a matrix entered by rows, including the couplings across the periodic
    boundary, solves exactly as the device solver does on the example's operator."""
    import cedar_amd as C
    from cedar_amd import capi as A
    from tests.capi_entries import entries2
    nx, ny = 63, 50
    so, b, _ = Gal.periodic_problem(nx, ny, (True, False))
    conf = {"solver": {"cycle": {"nrelax-pre": 1, "nrelax-post": 1}, "relaxation": "point",
                       "tol": 1e-10, "max-iter": 30, "min-coarse": 3, "cg-solver": "LU"},
            "grid": {"periodic": [True, False]}}
    p = tmp_path / "periodic-config.json"
    p.write_text(json.dumps(conf))
    lib = A.load()
    topo = A.topo2d(nx, ny)
    m = ctypes.c_int()
    A.check(lib.cedar_mat_create2d(A.config(p), topo, A.FIVE_PT, ctypes.byref(m)))
    ent = entries2(so)          # couplings between interior points
    for j in range(ny):         # and across the periodic boundary (matrix entries)
        ent.append((0, j, A.W, -so[1, j + 1, 1]))
        ent.append((nx - 1, j, A.E, -so[1, j + 1, nx + 1]))
    A.check(A.set2d(m, ent))
    slv = ctypes.c_int()
    A.check(lib.cedar_solver_create(m, ctypes.byref(slv)), "solver_create")
    x, bv = A.vec(topo), A.vec(topo)
    A.vec_array(bv)[...] = b
    A.check(lib.cedar_solver_run(slv, x, bv), "solver_run")
    ref, _ = C.Solver(dev(so), relaxation="point", nrelax_pre=1, nrelax_post=1, tol=1e-10,
                      maxiter=30, periodic=(True, False)).solve(dev(b))
    assert np.array_equal(A.vec_array(x), host(ref))
    # matvec with the periodic images of x equals the oracle's on the imaged vector
    xv, yv = A.vec(topo), A.vec(topo)
    xa = A.vec_array(xv)
    xa[1:-1, 1:-1] = np.random.default_rng(1).standard_normal((ny, nx))
    A.check(lib.cedar_matvec(m, xv, yv))
    xi = xa.copy()
    xi[:, 0], xi[:, -1] = xi[:, -2], xi[:, 1]
    assert np.array_equal(A.vec_array(yv), O.matvec(so, xi))


@pytest.mark.parametrize("nx,ny", [(12, 10), (33, 20), (64, 31)])
@pytest.mark.parametrize("mask", [1, 2, 3])
@pytest.mark.parametrize("ns", [3, 5])
def test_periodic_line_kernels_bitwise(nx, ny, mask, ns):
    import cedar_amd as C
    print(f"\nperiodic lines {nx}x{ny} ns={ns} mask={mask}", flush=True)
    rng = np.random.default_rng(nx + ny * 7 + mask * 3 + ns)
    jpn = O.bc_from_mask(mask)
    so = rand_periodic_so(rng, nx, ny, ns, mask)
    q = rng.standard_normal((ny + 2, nx + 2))
    qf = rng.standard_normal(q.shape)
    for d in ("x", "y"):
        sor, info = O.setup_lines_periodic(so, d, jpn)
        assert info == 0
        gsor = torch.zeros(sor.shape, dtype=torch.float64, device="cuda")
        C.setup_lines_periodic(dev(so), gsor, d, jpn)
        assert np.array_equal(host(gsor), sor), d
        for ud in (O.DOWN, O.UP):
            g = C.relax_lines_periodic(dev(so), dev(qf), dev(q), dev(sor), ud, d, jpn)
            assert np.array_equal(host(g), O.relax_lines_periodic(so, qf, q, sor, ud, d, jpn)), (d, ud)


@pytest.mark.parametrize("relax", ["line-x", "line-y", "line-xy"])
@pytest.mark.parametrize("mask", [1, 2, 3])
def test_periodic_line_solve_matches_oracle(relax, mask):
    import cedar_amd as C
    print(f"\nperiodic line solve {relax} mask={mask}", flush=True)
    so = rand_periodic_so(np.random.default_rng(mask + 11), 95, 70, 5, mask)
    so[0] += 0.2
    nx, ny = 95, 70
    if mask & 1:
        so[:, :, 0], so[:, :, nx + 1] = so[:, :, nx], so[:, :, 1]
    if mask & 2:
        so[:, 0, :], so[:, ny + 1, :] = so[:, ny, :], so[:, 1, :]
    b = np.random.default_rng(4).standard_normal((ny + 2, nx + 2))
    if mask & 1:
        b[:, 0], b[:, -1] = b[:, -2], b[:, 1]
    if mask & 2:
        b[0, :], b[-1, :] = b[-2, :], b[1, :]
    per = (bool(mask & 1), bool(mask & 2))
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=30, tol=1e-10)
    xo, ho = O.Solver(so, periodic=per, **kw).solve(b)
    x, h = C.Solver(dev(so), periodic=per, **kw).solve(dev(b))
    assert len(h) == len(ho)
    np.testing.assert_allclose(h, ho, rtol=1e-12, atol=0)
    assert np.linalg.norm(host(x) - xo) <= 1e-10 * np.linalg.norm(xo)
    xm, hm = C.Solver(dev(so), periodic=per, precision="mixed", **kw).solve(dev(b))
    assert hm[-1] < 1e-10
    assert torch.linalg.norm(xm - x) <= 1e-8 * torch.linalg.norm(x)

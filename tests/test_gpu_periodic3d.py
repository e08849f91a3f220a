"""3D periodic boundaries on the GPU (the reference's JPN != 0 branches).

Kernels: Gauss-Seidel, the periodic interpolation setup and restriction are bitwise equal to the
outputs of the reference's own Fortran kernels (tests/golden/bmg3_periodic.npz, written by
tests/golden/make_golden3_periodic.py from the sources compiled here) for x-, y-, z-, xy-, xz-,
yz- and xyz-periodic 7- and 27-point operators, including the reference's image-row quirks.  The
Galerkin product over the periodic geometry, interp_add with its image refresh and the dense
periodic coarse solve are bitwise equal to the oracle (oracle/bmg3_periodic.c, pinned on the CPU
against the reference and the mathematics); the Galerkin product also matches the reference's
hand-expanded products to 1e-13.

Solver: Cedar's periodic example (examples/basic-3d-mpi/periodic.cc) on one GPU for every mask:
the FP64 GPU history equals the oracle's to 1e-12 per cycle with the same cycle count, the
solution to 1e-10, and the discretisation error is O(h^2); F-cycles match the oracle; mixed
precision reaches the FP64 tolerance.
"""
import os

import numpy as np
import pytest

from cedar_amd import gallery as G
from oracle import oracle3 as O3
from tests.golden.gen3 import periodic_case, rand_so3_periodic
from tests.golden.make_golden3_periodic import CASES, tag

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MASKS = [1, 2, 3, 4, 5, 6, 7]


@pytest.fixture(scope="module")
def K():
    from cedar_amd import cdr3
    return cdr3


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "bmg3_periodic.npz"))


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("n,ns,mask", CASES)
def test_kernels_match_reference_vectors(K, golden, n, ns, mask):
    t = tag(n, ns, mask)
    print(f"case {t}", flush=True)
    so, q, qf = periodic_case(n, ns, mask)
    jpn = K.bc_from_mask(mask)
    dso = dev(so)
    sor = K.setup_recip(dso, torch.zeros((2,) + so.shape[1:], dtype=torch.float64, device="cuda"))
    for ud, key in ((K.DOWN, "gs_down"), (K.UP, "gs_up")):
        x = dev(q)
        K.relax_gs_per(dso, dev(qf), x, sor, ud, jpn)
        assert np.array_equal(host(x), golden[f"{t}/{key}"]), key
    ci = torch.zeros(golden[f"{t}/ci"].shape, dtype=torch.float64, device="cuda")
    K.setup_interp_per(dso, ci, jpn)
    cih = host(ci)
    assert np.array_equal(cih, golden[f"{t}/ci"], equal_nan=True)
    xq = dev(q)
    qc = torch.zeros(ci.shape[1:], dtype=torch.float64, device="cuda")
    K.restrict_per(ci, xq, qc, jpn)
    assert np.array_equal(host(qc), golden[f"{t}/restrict"])
    assert np.array_equal(host(xq), golden[f"{t}/restrict_q"])
    if all(n[a] % 2 == 0 for a in range(3) if mask & (1 << a)):
        soc = torch.zeros((14,) + tuple(ci.shape[1:]), dtype=torch.float64, device="cuda")
        K.setup_galerkin_per(dso, ci, soc, jpn)
        got = host(soc)
        assert np.array_equal(got, O3.setup_galerkin_bc(so, cih, jpn))
        ref = golden[f"{t}/soc"]
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("mask", MASKS)
@pytest.mark.parametrize("ns", [4, 14])
def test_interp_add_and_coarse_solve_match_oracle(K, mask, ns):
    rng = np.random.default_rng(40 + mask)
    so = rand_so3_periodic(rng, 8, 6, 10, ns, mask)
    jpn = K.bc_from_mask(mask)
    ci = O3.setup_interp_bc(so, jpn)
    kk, jj, ii = so.shape[1:]
    q = rng.standard_normal((kk, jj, ii))
    res = rng.standard_normal((kk, jj, ii))
    qc = rng.standard_normal(ci.shape[1:])
    want, _ = O3.interp_add_bc(q, qc, res, so, ci, jpn)
    x = dev(q)
    K.interp_add_per(x, dev(qc), dev(res), dev(so), dev(ci), jpn)
    assert np.array_equal(host(x), want)
    # coarsest grid: the dense factor and the solve with mean removal and images
    cs = rand_so3_periodic(rng, 4, 4, 2, ns, mask)
    fac, info = O3.setup_cg_bc(cs, jpn)
    assert info == 0
    gfac = K.setup_cg_per(dev(cs), jpn)
    assert np.array_equal(host(gfac), fac)
    b = rng.standard_normal(cs.shape[1:])
    xg = torch.zeros_like(dev(b))
    K.solve_cg_per(xg, dev(b), gfac, jpn)
    assert np.array_equal(host(xg), O3.solve_cg_bc(fac, b, jpn))


def rel_l2(a, b):
    s = (slice(1, -1),) * 3
    return np.linalg.norm((a - b)[s]) / np.linalg.norm(b[s])


@pytest.mark.parametrize("mask", MASKS)
def test_periodic_example_matches_oracle(mask):
    import cedar_amd
    per = tuple(bool(mask & (1 << a)) for a in range(3))
    so, b, exact = G.periodic_problem3(16, 16, 16, per)
    kw = dict(relaxation="point", maxiter=40, tol=1e-10, periodic=mask)
    xo, ho = O3.Solver3(so, **kw).solve(b)
    g = cedar_amd.Solver3(dev(so), **kw)
    xg, hg = g.solve(dev(b))
    xg = host(xg)
    rel = np.abs(hg - ho) / ho
    print(f"mask {mask}: {len(hg)} cycles (oracle {len(ho)}), max |drho| {np.abs(hg - ho).max():.2e}, "
          f"max |drho_i|/rho_i {rel.max():.2e}, rel l2 {rel_l2(xg, xo):.2e}", flush=True)
    assert len(hg) == len(ho) and hg[-1] < 1e-10
    assert np.abs(hg - ho).max() <= 1e-12
    assert rel_l2(xg, xo) <= 1e-10
    e = (xg - exact)[1:-1, 1:-1, 1:-1]
    if mask == 7:
        e = e - e.mean()
    assert np.abs(e).max() < 0.014


@pytest.mark.parametrize("mask,cycle,precision", [(7, "f", "f64"), (3, "v", "mixed"),
                                                  (5, "v", "mixed"), (6, "f", "mixed")])
def test_periodic_fcycle_and_mixed(mask, cycle, precision):
    """An F-cycle restarts from x = 0 each cycle (include/cedar/cycle/fcycle.h:49-83), so its FP64
    history is one value repeated, as the oracle's; as the correction cycle of the mixed
    refinement it converges.  Mixed V-cycles take the FP64 cycle count to within one."""
    import cedar_amd
    per = tuple(bool(mask & (1 << a)) for a in range(3))
    so, b, _ = G.periodic_problem3(16, 16, 16, per)
    kw = dict(relaxation="point", maxiter=40, tol=1e-10, periodic=mask)
    if precision == "f64":
        x64, h64 = cedar_amd.Solver3(dev(so), cycle=cycle, maxiter=3, **{
            k: v for k, v in kw.items() if k != "maxiter"}).solve(dev(b))
        xo, ho = O3.Solver3(so, cycle=cycle, maxiter=3, **{
            k: v for k, v in kw.items() if k != "maxiter"}).solve(b)
        assert len(h64) == len(ho) == 3 and np.abs(h64 - ho).max() <= 1e-12
        assert rel_l2(host(x64), xo) <= 1e-10
        return
    x64, h64 = cedar_amd.Solver3(dev(so), **kw).solve(dev(b))
    xm, hm = cedar_amd.Solver3(dev(so), precision="mixed", cycle=cycle, **kw).solve(dev(b))
    print(f"mask {mask} {cycle}: mixed {len(hm)} cycles, f64 V {len(h64)}", flush=True)
    assert hm[-1] < 1e-10
    if cycle == "v":
        assert abs(len(hm) - len(h64)) <= 1
    assert rel_l2(host(xm), host(x64)) < 1e-9


def test_periodic_settings_are_checked():
    import cedar_amd
    so, _, _ = G.periodic_problem3(15, 16, 16, (True, False, False))
    with pytest.raises(cedar_amd.CedarHipError) as e:  # odd periodic extent
        cedar_amd.Solver3(dev(so), periodic=1)
    assert e.value.status == 1
    so, _, _ = G.periodic_problem3(16, 16, 16, (True, False, False))
    with pytest.raises(cedar_amd.CedarHipError):  # plane relaxation
        cedar_amd.Solver3(dev(so), relaxation="plane-xy", periodic=1)


def test_capi_periodic_solve_3d(tmp_path):
    """The handle API honours grid.periodic in 3D: the example's operator entered by its stored
    planes at interior points (the couplings across the periodic boundaries included) and its
    right-hand side solve exactly as the device solver does; the matvec uses x's images."""
    import ctypes
    import json

    import cedar_amd
    from cedar_amd import capi as A
    from tests.capi_entries import entries3
    n = 16
    per = (True, False, True)
    so, b, _ = G.periodic_problem3(n, n, n, per)
    conf = {"solver": {"cycle": {"nrelax-pre": 2, "nrelax-post": 1}, "relaxation": "point",
                       "tol": 1e-10, "max-iter": 40, "min-coarse": 3, "cg-solver": "LU"},
            "grid": {"periodic": list(per)}}
    p = tmp_path / "periodic3-config.json"
    p.write_text(json.dumps(conf))
    lib = A.load()
    topo = A.topo3d(n, n, n)
    m = ctypes.c_int()
    A.check(lib.cedar_mat_create3d(A.config(p), topo, A.SEVEN_PT, ctypes.byref(m)))
    A.check(A.set3d(m, entries3(so)))
    slv = ctypes.c_int()
    A.check(lib.cedar_solver_create(m, ctypes.byref(slv)), "solver_create")
    x, bv = A.vec(topo, 3), A.vec(topo, 3)
    A.vec_array(bv)[...] = b
    A.check(lib.cedar_solver_run(slv, x, bv), "solver_run")
    ref, _ = cedar_amd.Solver3(dev(so), relaxation="point", tol=1e-10, maxiter=40,
                               periodic=5).solve(dev(b))
    assert np.array_equal(A.vec_array(x), host(ref))
    xv, yv = A.vec(topo, 3), A.vec(topo, 3)
    xa = A.vec_array(xv)
    xa[1:-1, 1:-1, 1:-1] = np.random.default_rng(2).standard_normal((n, n, n))
    A.check(lib.cedar_matvec(m, xv, yv))
    xi = xa.copy()
    xi[:, :, 0], xi[:, :, -1] = xi[:, :, -2], xi[:, :, 1]
    xi[0], xi[-1] = xi[-2], xi[1]
    assert np.array_equal(A.vec_array(yv), O3.matvec(so, xi))

"""Parity of every 2D HIP kernel (through the C ABI) with the FP64 oracle.

Bitwise: residual, 1/c, Gauss-Seidel, restriction, interpolation, operator-induced
interpolation, line factorisation, coarse Cholesky (same operation order, no FMA contraction).
Within 1e-13 (relative to the largest entry): Galerkin coarse operator (formed as P^T A P) and
line relaxation (wavefront affine scans instead of the sequential LDL^T substitution; every
launch shape of the line kernels gives bitwise identical results).
FP32 kernels: within 1e-5 relative of the FP64 oracle.
"""
import numpy as np
import pytest

from cedar_amd import _abi
from oracle import oracle as O
from tests.golden.make_golden import rand_so

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = ["7x9_3", "7x9_5", "8x6_3", "8x6_5", "16x11_3", "16x11_5", "15x15_3", "15x15_5"]
SIZES = [(3, 3), (4, 4), (5, 8), (33, 17), (64, 65), (130, 129), (257, 200)]


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def case(golden, tag):
    return {k.split("/", 1)[1]: golden[k] for k in golden.files if k.startswith(tag + "/")}


def random_case(nx, ny, ns, seed):
    rng = np.random.default_rng(seed)
    so = rand_so(rng, nx, ny, ns)
    q = rng.standard_normal((ny + 2, nx + 2))
    qf = rng.standard_normal((ny + 2, nx + 2))
    return so, q, qf, rng


def assert_close(got, want, rtol):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want).max() / scale
    assert err <= rtol, f"max rel err {err:.3e} > {rtol:.1e}"


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("tag", CASES)
def test_golden_fixtures_bitwise(C, golden2d, tag):
    g = case(golden2d, tag)
    so, q, qf, res, qc = (dev(g[k]) for k in ("so", "q", "qf", "res", "qc"))
    r = torch.zeros_like(q)
    C.residual(so, qf, q, r)
    assert np.array_equal(host(r), g["residual"])
    sor = torch.zeros((2,) + tuple(q.shape), dtype=torch.float64, device="cuda")
    C.setup_recip(so, sor)
    assert np.array_equal(host(sor)[1, 1:-1, 1:-1], g["recip"][1, 1:-1, 1:-1])
    for ud, key in ((0, "gs_down"), (1, "gs_up")):
        qq = q.clone()
        C.relax_gs(so, qf, qq, sor, ud)
        assert np.array_equal(host(qq), g[key]), key
    ci = torch.zeros(tuple(g["ci"].shape), dtype=torch.float64, device="cuda")
    C.setup_interp(so, ci)
    assert np.array_equal(host(ci), g["ci"])
    soc = torch.zeros(tuple(g["soc"].shape), dtype=torch.float64, device="cuda")
    C.setup_galerkin(so, ci, soc)
    assert_close(host(soc), g["soc"], 1e-13)
    qcr = torch.zeros_like(qc)
    C.restrict(ci, q, qcr)
    assert np.array_equal(host(qcr), g["restrict"])
    qi = q.clone()
    C.interp_add(qi, qc, res, so, ci)
    assert np.array_equal(host(qi), g["interp_q"])


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("ns", [3, 5])
def test_pointwise_kernels_bitwise(C, nx, ny, ns):
    so, q, qf, rng = random_case(nx, ny, ns, 10 * nx + ny + ns)
    dso, dq, dqf = dev(so), dev(q), dev(qf)
    r = torch.zeros_like(dq)
    C.residual(dso, dqf, dq, r)
    assert np.array_equal(host(r), O.residual(so, qf, q))
    nrm = C.residual_norm(dso, dqf, dq, r)
    np.testing.assert_allclose(nrm, O.norm(O.residual(so, qf, q)), rtol=1e-13)
    sor = O.setup_recip(so)
    dsor = torch.zeros_like(dev(sor))
    C.setup_recip(dso, dsor)
    assert np.array_equal(host(dsor), sor)
    for ud in (0, 1):
        qq = dq.clone()
        C.relax_gs(dso, dqf, qq, dsor, ud)
        assert np.array_equal(host(qq), O.relax_gs(so, qf, q, sor, ud))
    if min(nx, ny) >= 3:
        ci = O.setup_interp(so)
        dci = torch.zeros_like(dev(ci))
        C.setup_interp(dso, dci)
        assert np.array_equal(host(dci), ci)
        soc = O.setup_galerkin(so, ci)
        dsoc = torch.zeros_like(dev(soc))
        C.setup_galerkin(dso, dci, dsoc)
        assert_close(host(dsoc), soc, 1e-13)
        qc = np.zeros(ci.shape[1:])
        dqc = torch.zeros_like(dev(qc))
        C.restrict(dci, dq, dqc)
        assert np.array_equal(host(dqc), O.restrict(ci, q))
        qc = rng.standard_normal(ci.shape[1:])
        res = rng.standard_normal(q.shape)
        qq = dq.clone()
        C.interp_add(qq, dev(qc), dev(res), dso, dci)
        assert np.array_equal(host(qq), O.interp_add(q, qc, res, so, ci)[0])


LINE_SIZES = SIZES + [(1000, 7), (7, 1000), (700, 33), (33, 700), (300, 5000), (4095, 9),
              (1500, 9), (9, 1500), (9, 4095)]


@pytest.mark.parametrize("nx,ny", LINE_SIZES)
@pytest.mark.parametrize("ns", [3, 5])
@pytest.mark.parametrize("direction", ["x", "y"])
def test_line_relaxation(C, nx, ny, ns, direction):
    so, q, qf, _ = random_case(nx, ny, ns, 7 * nx + ny + ns + (direction == "y"))
    sor, info = O.setup_lines(so, direction)
    assert info == 0
    dso = dev(so)
    dsor = torch.zeros_like(dev(sor))
    C.setup_lines(dso, dsor, direction)
    assert np.array_equal(host(dsor), sor)
    for ud in (0, 1):
        qq = dev(q)
        C.relax_lines(dso, dev(qf), qq, dsor, ud, direction)
        assert_close(host(qq), O.relax_lines(so, qf, q, sor, ud, direction), 1e-13)


@pytest.mark.parametrize("nx,ny", [(700, 33), (33, 700), (4095, 5), (5, 4095), (130, 129),
                                   (2047, 7), (7, 2047), (1500, 9), (9, 1500)])
@pytest.mark.parametrize("direction", ["x", "y"])
def test_line_relaxation_fp32(C, nx, ny, direction):
    so, q, qf, _ = random_case(nx, ny, 5, 3 * nx + ny + (direction == "y"))
    sor, info = O.setup_lines(so, direction)
    f = lambda a: dev(a, torch.float32)
    for ud in (0, 1):
        qq = f(q)
        C.relax_lines(f(so), f(qf), qq, f(sor), ud, direction)
        assert_close(host(qq), O.relax_lines(so, qf, q, sor, ud, direction), 2e-5)


@pytest.mark.parametrize("direction", ["x", "y"])
def test_reference_line_smoothing_tests(C, direction):
    """test/2d/test_relax.cc:96-165 (LineX5, LineY5) on the GPU kernels: 15 DOWN line sweeps on
    the 132^2 operator strong along the lines, b = 0, random x, leave high frequencies below 1e-10
    (measured as test/2d/pyrelax.pyx get_high_freq does)."""
    from tests.test_oracle_math import high_freq
    from cedar_amd import gallery as G
    dx, dy = (1.0, 1e-4) if direction == "x" else (1e-4, 1.0)
    so = G.diag_diffusion(132, 132, dx, dy)
    x0 = np.zeros((134, 134))
    x0[1:-1, 1:-1] = np.random.default_rng(5).uniform(size=(132, 132))
    dso, b = dev(so), torch.zeros((134, 134), dtype=torch.float64, device="cuda")
    sor = torch.zeros((2, 134, 134), dtype=torch.float64, device="cuda")
    C.setup_lines(dso, sor, direction)
    x = dev(x0)
    for _ in range(15):
        C.relax_lines(dso, b, x, sor, 0, direction)
    assert high_freq(host(x)) < 1e-10


def test_line_factorisation_reports_non_spd(C):
    so = np.zeros((3, 7, 7))
    so[0, 1:-1, 1:-1] = 1.0
    so[1, 1:-1, 2:-1] = 2.0  # |off-diagonal| > diagonal: not positive definite
    dsor = torch.zeros((2, 7, 7), dtype=torch.float64, device="cuda")
    with pytest.raises(C.CedarHipError) as e:
        C.setup_lines(dev(so), dsor, "x")
    assert e.value.status == 3


@pytest.mark.parametrize("nx,ny,ns", [(3, 3, 5), (4, 4, 5), (4, 4, 3), (5, 3, 5), (9, 7, 5)])
def test_coarse_cholesky(C, nx, ny, ns):
    so, q, qf, _ = random_case(nx, ny, ns, nx * 31 + ny)
    abd, info = O.setup_cg(so)
    dabd = C.setup_cg(dev(so))
    assert np.array_equal(host(dabd), abd)
    qf[0, :] = qf[-1, :] = 0.0
    qf[:, 0] = qf[:, -1] = 0.0
    dq = torch.zeros_like(dev(qf))
    C.solve_cg(dabd, dev(qf), dq)
    assert np.array_equal(host(dq), O.solve_cg(abd, qf))


def test_coarse_cholesky_reports_non_spd(C):
    so = np.zeros((5, 6, 6))
    so[0, 1:-1, 1:-1] = -1.0
    with pytest.raises(C.CedarHipError) as e:
        C.setup_cg(dev(so))
    assert e.value.status == 3


@pytest.mark.parametrize("ns", [3, 5])
def test_fp32_kernels_close_to_fp64_oracle(C, ns):
    nx, ny = 129, 96
    so, q, qf, rng = random_case(nx, ny, ns, 99 + ns)
    f = lambda a: dev(a, torch.float32)
    r = torch.zeros_like(f(q))
    C.residual(f(so), f(qf), f(q), r)
    assert_close(host(r), O.residual(so, qf, q), 2e-6)
    sor = O.setup_recip(so)
    qq = f(q)
    C.relax_gs(f(so), f(qf), qq, f(sor), 0)
    assert_close(host(qq), O.relax_gs(so, qf, q, sor, 0), 2e-6)
    for d in ("x", "y"):
        sl, _ = O.setup_lines(so, d)
        qq = f(q)
        C.relax_lines(f(so), f(qf), qq, f(sl), 1, d)
        assert_close(host(qq), O.relax_lines(so, qf, q, sl, 1, d), 1e-5)
    ci = O.setup_interp(so)
    qc = torch.zeros(ci.shape[1:], dtype=torch.float32, device="cuda")
    C.restrict(f(ci), f(q), qc)
    assert_close(host(qc), O.restrict(ci, q), 2e-6)
    qcn = rng.standard_normal(ci.shape[1:])
    res = rng.standard_normal(q.shape)
    qq = f(q)
    C.interp_add(qq, f(qcn), f(res), f(so), f(ci))
    assert_close(host(qq), O.interp_add(q, qcn, res, so, ci)[0], 2e-6)


def test_bad_arguments_are_rejected(C):
    from cedar_amd import _abi
    z = torch.zeros((5, 5), dtype=torch.float64, device="cuda")
    so = torch.zeros((4, 5, 5), dtype=torch.float64, device="cuda")  # 4 planes: invalid
    st = _abi.load().cdr_hip_residual2d_f64(so.data_ptr(), 4, z.data_ptr(), z.data_ptr(),
                                            z.data_ptr(), 5, 5, None)
    assert st == 1
    st = _abi.load().cdr_hip_relax_gs2d_f64(so.data_ptr(), 5, z.data_ptr(), z.data_ptr(),
                                            z.data_ptr(), 5, 5, 7, None)
    assert st == 1
    st = _abi.load().cdr_hip_restrict2d_f64(z.data_ptr(), z.data_ptr(), z.data_ptr(), 5, 5, 9, 9,
                                            None)
    assert st == 1


XY_SIZES = [(3, 3), (4, 4), (5, 8), (9, 7), (33, 17), (64, 65), (130, 129), (257, 200),
            (1000, 7), (7, 1000), (700, 33), (33, 700), (300, 2000), (4095, 9)]


@pytest.mark.parametrize("nx,ny", XY_SIZES)
@pytest.mark.parametrize("ns", [3, 5])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_line_xy_equals_sequential_sweeps(C, nx, ny, ns, dtype):
    """One line-xy step (x then y down, y then x up) is bit for bit the two line kernels, ghost
    ring untouched."""
    so, q, qf, rng = random_case(nx, ny, ns, 11 * nx + ny + ns)
    q[0, :] = rng.standard_normal(nx + 2)  # nonzero ghosts must pass through unchanged
    q[:, -1] = rng.standard_normal(ny + 2)
    sx, _ = O.setup_lines(so, "x")
    sy, _ = O.setup_lines(so, "y")
    f = lambda a: dev(a, dtype)
    for ud in (0, 1):
        q1 = f(q)
        C.relax_lines_xy(f(so), f(qf), q1, f(sx), f(sy), ud)
        q2 = f(q)
        order = ("x", "y") if ud == 0 else ("y", "x")
        for d in order:
            C.relax_lines(f(so), f(qf), q2, f(sx if d == "x" else sy), ud, d)
        assert torch.equal(q1, q2), ud
        assert torch.equal(q1[0], q2[0]) and torch.equal(q1[:, -1], f(q)[:, -1])


# line lengths covering every chunk class (1..64 chunks of 64, and the x-line classes of 3, 5, 9,
# 17 and 33 chunks of the 2^k + 1 lines) and partial last segments
SHAPE_SIZES = [(5, 9), (100, 9), (129, 9), (200, 9), (257, 9), (500, 9), (513, 9), (1000, 9),
               (1025, 9), (1500, 9), (2047, 9), (2049, 9), (3000, 9), (4095, 9)]


@pytest.mark.parametrize("n,m", SHAPE_SIZES)
@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_every_line_shape_gives_identical_results(C, n, m, direction, dtype):
    """All launch shapes of the line kernels (the autotuner's candidates) give bitwise identical
    results, each within the oracle tolerance."""
    nx, ny = (n, m) if direction == "x" else (m, n)
    ii, jj = nx + 2, ny + 2
    bits = 64 if dtype == torch.float64 else 32
    lib = _abi.load()
    for ns in (3, 5):
        so, q, qf, _ = random_case(nx, ny, ns, 5 * n + m + ns + (direction == "y"))
        sor, _ = O.setup_lines(so, direction)
        ref = O.relax_lines(so, qf, q, sor, 1, direction)
        d = 0 if direction == "x" else 1
        nshape = lib.cdr_hip_line_shapes2d(d, bits, ns, ii, jj)
        assert nshape >= 1
        f = lambda a: dev(a, dtype)
        outs = []
        try:
            for shape in range(nshape):
                assert lib.cdr_hip_set_line_shape2d(d, bits, ns, ii, jj, 1, shape) == 0
                qq = f(q)
                C.relax_lines(f(so), f(qf), qq, f(sor), 1, direction)
                outs.append(qq)
        finally:
            lib.cdr_hip_set_line_shape2d(d, bits, ns, ii, jj, 1, -1)
        for shape in range(1, nshape):
            assert torch.equal(outs[0], outs[shape]), (ns, shape)
        assert_close(host(outs[0]), ref, 1e-13 if bits == 64 else 2e-5)
        assert lib.cdr_hip_set_line_shape2d(d, bits, ns, ii, jj, 1, nshape) == 1  # out of range


def test_autotune_records_shapes(C):
    """Solver setup with line relaxation records a shape for every smoothing level and direction
    that has more than one candidate (lines of 1023 points have several of each)."""
    import cedar_amd
    from cedar_amd import gallery as G
    so, b, _ = G.heterogeneous_9pt(1023, seed=4, block=16, sigma=1.0, contrast=1e2)
    prev = cedar_amd.cdr2.autotune(True)
    try:
        s = cedar_amd.Solver(so, relaxation="line-xy", precision="mixed", nrelax_pre=1,
                             nrelax_post=1, maxiter=30, tol=1e-10)
        rep = cedar_amd.cdr2.autotune_report()
        fine = [r for r in rep if r["ii"] == 1025 and r["jj"] == 1025 and r["fp"] == 32]
        # x, y, and the choice between direct and transposed y sweeps ("t")
        assert {r["dir"] for r in fine} == {"x", "y", "t"}
        x, h = s.solve(torch.as_tensor(b).cuda())
        assert h[-1] < 1e-10
    finally:
        cedar_amd.cdr2.autotune(prev)


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("ns", [3, 5])
@pytest.mark.parametrize("pre,post,nb", [(1, 1, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)])
def test_transposed_y_sweeps_are_bitwise_direct(C, precision, ns, pre, post, nb):
    """Line-xy solves with the y sweeps run as x sweeps of the transposed grid (pinned at every
    level) give bitwise the iterates and histories of the direct y-line kernels.  On that path
    the iterate stays transposed from the last pre-smoothing y sweep through the residual and
    interp_add to the first post-smoothing y sweep, so this also pins the residual and
    interp_add kernels that read and update it in that layout."""
    from cedar_amd import gallery as G
    n = 257
    if ns == 5:
        so, b, _ = G.heterogeneous_9pt(n, seed=8, block=16, sigma=1.5, contrast=1e3)
    else:
        so, b = G.poisson_uniform_rhs(n, seed=8)
    if nb > 1:
        so2, b2, _ = G.heterogeneous_9pt(n, seed=9, block=16, sigma=0.5, contrast=1e2)
        so = np.stack([so, so2 if ns == 5 else so2[:3]])
        b = np.stack([b, b2])
    lib = _abi.load()
    bits = 64 if precision == "f64" else 32
    sizes, m = [], n
    while m >= 3:
        sizes.append(m)
        m = int((m - 1) / 2.0 + 1)
    prev = lib.cdr_hip_autotune(0)
    kw = dict(relaxation="line-xy", nrelax_pre=pre, nrelax_post=post, maxiter=40, tol=1e-10,
              precision=precision)
    try:
        x0, h0 = C.Solver(so, **kw).solve(dev(b))
        for lvl, m in enumerate(sizes):
            stencil = ns if lvl == 0 else 5
            if lib.cdr_hip_line_shapes2d(2, bits, stencil, m + 2, m + 2) > 0:
                assert lib.cdr_hip_set_line_shape2d(2, bits, stencil, m + 2, m + 2, nb, 0) == 0
        x1, h1 = C.Solver(so, **kw).solve(dev(b))
    finally:
        for lvl, m in enumerate(sizes):
            lib.cdr_hip_set_line_shape2d(2, bits, ns if lvl == 0 else 5, m + 2, m + 2, nb, -1)
        lib.cdr_hip_autotune(prev)
    h0, h1 = [(h if nb > 1 else [h]) for h in (h0, h1)]
    assert len(h0) == len(h1) == nb
    for a, c in zip(h0, h1):
        assert np.array_equal(a, c)
        assert a[-1] < 1e-10
    assert torch.equal(x0, x1)

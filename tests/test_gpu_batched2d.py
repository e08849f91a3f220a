"""Batched 2D solves: nb problems of the same extents in one solver, every kernel launch covering
the batch.  Each problem of a batch must get exactly (bit for bit) the iterate and residual
history of solving it alone, including problems that converge in fewer cycles than others
(they stop and keep their iterate while the rest continue)."""
import numpy as np
import pytest

from cedar_amd import gallery as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    import cedar_amd
    return cedar_amd


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).cuda()


def problems(n, nb, nine):
    sos, bs = [], []
    for k in range(nb):
        if nine:
            # contrasts differ per problem, so the problems need different cycle counts
            so, b, _ = G.heterogeneous_9pt(n, seed=20 + k, block=8, sigma=0.5 + 0.6 * k)
        else:
            so, b = G.poisson_uniform_rhs(n, seed=30 + k)
        sos.append(so)
        bs.append(b)
    return np.stack(sos), np.stack(bs)


@pytest.mark.parametrize("relax,n,nine,precision", [
    ("line-xy", 127, True, "f64"), ("line-xy", 127, True, "mixed"),
    ("point", 96, False, "f64"), ("line-x", 65, True, "mixed")])
def test_batch_equals_individual_solves(C, relax, n, nine, precision):
    nb = 4
    so, b = problems(n, nb, nine)
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10,
              precision=precision)
    sb = C.Solver(dev(so), **kw)
    assert sb.nb == nb and sb.batched
    xb, hb = sb.solve(dev(b))
    counts = []
    for k in range(nb):
        s1 = C.Solver(dev(so[k]), **kw)
        x1, h1 = s1.solve(dev(b[k]))
        assert np.array_equal(hb[k], h1), k
        assert torch.equal(xb[k], x1), k
        assert h1[-1] < 1e-10
        counts.append(len(h1))
    if relax == "line-xy":
        assert len(set(counts)) > 1, counts  # the stopping mask was exercised


def test_batched_setup_matches_individual(C):
    so, _ = problems(63, 3, True)
    sb = C.Solver(dev(so), relaxation="line-xy")
    for k in range(3):
        s1 = C.Solver(dev(so[k]), relaxation="line-xy")
        assert s1.nlevels == sb.nlevels
        for lvl in range(s1.nlevels):
            for a, c in zip(sb.level_arrays(lvl), s1.level_arrays(lvl)):
                if c is None:
                    continue
                assert np.array_equal(a[k], c), (k, lvl)


def test_batched_not_spd_reported(C):
    so, _ = problems(31, 3, False)
    so[1, 0, 1:-1, 1:-1] = -1.0
    with pytest.raises(C.CedarHipError) as e:
        C.Solver(dev(so), relaxation="line-y")
    assert e.value.status == 3


@pytest.mark.parametrize("concurrent", [True, False])
def test_batchset_mixed_sizes_equal_individual(C, concurrent):
    """Mixed sizes grouped into batched solvers (the groups on streams of their own, driven from
    one host thread each, or one after another): each problem gets its individual result, on the
    first solve and again on the second (graph replay)."""
    from cedar_amd.batch import BatchSet
    sizes = [33, 65, 33, 47, 65, 33]
    ops, rhs = [], []
    for k, n in enumerate(sizes):
        so, b, _ = G.heterogeneous_9pt(n, seed=70 + k, block=8, sigma=1.0 + 0.2 * k)
        ops.append(dev(so))
        rhs.append(dev(b))
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10,
              precision="mixed")
    bs = BatchSet(ops, concurrent=concurrent, **kw)
    assert len(bs.groups) == 3
    runs = [bs.solve(rhs), bs.solve(rhs)]
    for k in range(len(sizes)):
        x1, h1 = C.Solver(ops[k], **kw).solve(rhs[k])
        for xs, hs in runs:
            assert np.array_equal(hs[k], h1), k
            assert torch.equal(xs[k], x1), k


@pytest.mark.parametrize("direction", ["x", "y"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("nine", [True, False])
def test_batched_line_sweep_entry_points(C, direction, dtype, nine):
    """cdr_hip_relax_lines2d_batched_*: one call over a stack of problems gives every problem
    bitwise the result of its own single-problem sweep, both colour orders."""
    K = C.cdr2
    so, b = problems(77, 3, nine)
    rng = np.random.default_rng(3)
    q = np.zeros_like(b)
    q[:, 1:-1, 1:-1] = rng.standard_normal((3, 77, 77))
    sor = torch.stack([K.setup_lines(dev(so[k]), torch.zeros((2,) + b.shape[1:],
                                                           dtype=torch.float64, device="cuda"),
                                     direction) for k in range(3)])
    for ud in (0, 1):
        want = [K.relax_lines(dev(so[k]).to(dtype), dev(b[k]).to(dtype), dev(q[k]).to(dtype),
                              sor[k].to(dtype), ud, direction) for k in range(3)]
        got = K.relax_lines_batched(dev(so).to(dtype), dev(b).to(dtype), dev(q).to(dtype),
                                    sor.to(dtype).contiguous(), ud, direction)
        for k in range(3):
            assert torch.equal(got[k], want[k]), (k, ud)


@pytest.mark.parametrize("relax,n,nine,precision", [
    ("line-xy", 127, True, "mixed"), ("line-xy", 96, True, "f64"), ("point", 65, False, "f64")])
def test_shared_operator_rhs_equal_individual_solves(C, relax, n, nine, precision):
    """nrhs right-hand sides per operator (cdr_hip_solver2d_create_shared): the hierarchy is built
    once per operator and every cycle kernel reads it once for the operator's right-hand sides;
    each (operator, right-hand side) problem gets bitwise the iterate and history of its own
    single solve, including problems that stop earlier than others."""
    nops, nrhs = 2, 3
    so, b0 = problems(n, nops, nine)
    rng = np.random.default_rng(41)
    rhs = np.zeros((nops, nrhs) + b0.shape[1:])
    for k in range(nops):
        for r in range(nrhs):
            rhs[k, r, 1:-1, 1:-1] = rng.uniform(-1, 1, (n, n)) * (1.0 + 3.0 * r)
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10,
              precision=precision)
    s = C.Solver(dev(so), nrhs=nrhs, **kw)
    assert s.nb == nops * nrhs
    x, h = s.solve(dev(rhs.reshape((nops * nrhs,) + b0.shape[1:])))
    lens = set()
    for k in range(nops):
        single = C.Solver(dev(so[k]), **kw)
        for r in range(nrhs):
            xs, hs = single.solve(dev(rhs[k, r]))
            assert np.array_equal(h[k * nrhs + r], hs), (k, r)
            assert torch.equal(x[k * nrhs + r], xs), (k, r)
            lens.add(len(hs))
    assert min(lens) < 40


def test_shared_operator_argument_errors(C):
    so, b = problems(33, 2, True)
    with pytest.raises(C.CedarHipError):
        C.Solver(dev(so), nrhs=2, relaxation="point", periodic=(True, False))
    s = C.Solver(dev(so), nrhs=2, relaxation="point")
    with pytest.raises(ValueError):
        s.solve(dev(b))  # 2 grids given, 4 problems

"""The fused forms of the V-cycle against the eager step-by-step cycle.

Without timers the cycles run as graph replays and the coarse levels as one fused launch
(coarse_cycle_k, data in LDS where it fits); with the per-level timers on, every step of every
level is its own launch, run eagerly.  Both must give bit for bit the same residual histories and
iterates: odd, even and ragged extents, 5- and 9-point operators, point and line relaxation
(line-xy keeps the iterate in the transposed layout between the smoothings of the levels whose y
sweeps run through the transposed grid), FP64 and mixed precision, batches and shared operators.
"""
import numpy as np
import pytest

from cedar_amd import gallery as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


CASES = [
    # (nx, ny, 9-point, relaxation, precision, problems, right-hand sides per operator)
    (255, 255, True, "line-xy", "mixed", 1, 1),
    (256, 256, True, "line-xy", "f64", 1, 1),
    (200, 131, True, "line-xy", "mixed", 1, 1),
    (257, 257, False, "point", "f64", 1, 1),
    (130, 97, False, "line-x", "mixed", 1, 1),
    (127, 127, True, "line-xy", "f64", 2, 1),
    (129, 129, True, "line-xy", "mixed", 2, 2),
]


def problem(nx, ny, nine, seed):
    so, b, _ = G.heterogeneous_9pt(max(nx, ny), seed=seed, block=8, sigma=1.0, contrast=1e2)
    so, b = so[:, :ny + 2, :nx + 2].copy(), b[:ny + 2, :nx + 2].copy()
    so[:, -1, :] = 0.0
    so[:, :, -1] = 0.0
    b[-1, :] = 0.0
    b[:, -1] = 0.0
    if not nine:
        so = so[:3].copy()
    return so, b


@pytest.mark.parametrize("nx,ny,nine,relax,precision,nops,nrhs", CASES)
def test_fused_cycle_steps_equal_unfused(nx, ny, nine, relax, precision, nops, nrhs):
    import cedar_amd
    print(f"case {nx}x{ny} nine={nine} {relax} {precision} nops={nops} nrhs={nrhs}", flush=True)
    ops, rhs = [], []
    for k in range(nops):
        so, b = problem(nx, ny, nine, 11 + k)
        ops.append(so)
        for r in range(nrhs):
            rhs.append(b * (1.0 + 0.25 * r) + 0.01 * r)
    so = np.stack(ops) if nops > 1 else ops[0]
    b = np.stack(rhs) if len(rhs) > 1 else rhs[0]
    kw = dict(relaxation=relax, nrelax_pre=1, nrelax_post=1, maxiter=40, tol=1e-10,
              precision=precision)
    if nrhs > 1:
        kw["nrhs"] = nrhs
    s0 = cedar_amd.Solver(dev(so), **kw)
    x0, h0 = s0.solve(dev(b))
    s1 = cedar_amd.Solver(dev(so), **kw)
    s1.timers(True)
    x1, h1 = s1.solve(dev(b))
    h0 = h0 if isinstance(h0, list) else [h0]
    h1 = h1 if isinstance(h1, list) else [h1]
    assert len(h0) == len(h1) == nops * nrhs
    for a, c in zip(h0, h1):
        assert np.array_equal(a, c)
        assert a[-1] < 1e-10
    assert torch.equal(x0, x1)


ZS_CASES = [
    # (nx, ny, 9-point, pre-smoothing sweeps, problems, right-hand sides per operator)
    (255, 255, True, 1, 1, 1),
    (127, 255, False, 1, 1, 1),
    (1023, 1023, True, 2, 1, 1),
    (129, 129, True, 1, 2, 2),
]


@pytest.mark.parametrize("nx,ny,nine,npre,nops,nrhs", ZS_CASES)
def test_zero_start_equals_zero_filled_start(monkeypatch, nx, ny, nine, npre, nops, nrhs):
    """The mixed solve's FP32 cycle starts from a zero correction: by default the first colour of
    the finest level's first x sweep takes the zeros as operands instead of a fill of the
    correction before each cycle (CEDAR_AMD_ZERO_START=0 at construction keeps the fill).  Both
    must give the same residual histories and iterates, bit for bit, in graph replays and with
    the per-level timers on (eager)."""
    import cedar_amd
    print(f"zero start {nx}x{ny} nine={nine} npre={npre} nops={nops} nrhs={nrhs}", flush=True)
    ops, rhs = [], []
    for k in range(nops):
        so, b = problem(nx, ny, nine, 21 + k)
        ops.append(so)
        for r in range(nrhs):
            rhs.append(b * (1.0 + 0.5 * r) - 0.02 * r)
    so = np.stack(ops) if nops > 1 else ops[0]
    b = np.stack(rhs) if len(rhs) > 1 else rhs[0]
    kw = dict(relaxation="line-xy", nrelax_pre=npre, nrelax_post=1, maxiter=40, tol=1e-10,
              precision="mixed")
    if nrhs > 1:
        kw["nrhs"] = nrhs
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("CEDAR_AMD_ZERO_START", flag)
        for timers in (False, True):
            s = cedar_amd.Solver(dev(so), **kw)
            s.timers(timers)
            x, h = s.solve(dev(b))
            out.append((x, h if isinstance(h, list) else [h]))
    x0, h0 = out[0]
    assert len(h0) == nops * nrhs
    for x, h in out[1:]:
        for a, c in zip(h0, h):
            assert np.array_equal(a, c)
        assert torch.equal(x0, x)
    assert all(a[-1] < 1e-9 for a in h0)

"""cdr_hip_residual_norm2d_f64_noalloc: the allocation-free residual + squared norm, through the C
ABI, against cdr_hip_residual_norm2d_f64 (the same kernels with a caller-supplied work buffer).

Bitwise equality of r and of sum(r^2) on two calls in a row, the second on a smaller grid (the
device array then still holds the first call's partials beyond the ones the second call
writes), plus the size limit, which the host rejects before any device work.
"""
import ctypes

import numpy as np
import pytest

from cedar_amd import _abi
from cedar_amd import gallery as G

CDR_ERR_TOO_LARGE = 8
MAX_PARTIALS = 524288


def test_noalloc_limit_matches_header():
    text = open(_abi.HEADER).read()
    assert f"#define CDR_NOALLOC_NORM_PARTIALS {MAX_PARTIALS}" in text
    assert "CDR_ERR_TOO_LARGE = 8" in text


def test_noalloc_rejects_grids_beyond_its_buffer():
    """Argument and size checks run on the host before any device call (no GPU needed): the
    dummy pointers below are never dereferenced."""
    lib = _abi.load()
    fake = ctypes.c_void_p(0x1000)
    # 16386 x 16386 needs 256 x 4096 partials, twice the array
    big = 16386
    assert lib.cdr_hip_norm_work2d(big, big) > MAX_PARTIALS
    st = lib.cdr_hip_residual_norm2d_f64_noalloc(fake, 5, fake, fake, fake, big, big, fake, None)
    assert st == CDR_ERR_TOO_LARGE
    assert "fixed-size" in lib.cdr_hip_status_string(st).decode()
    # the largest square grid that fits: 8190 interior points -> 128 x 2048 partials
    assert lib.cdr_hip_norm_work2d(8194, 8194) <= MAX_PARTIALS
    assert lib.cdr_hip_residual_norm2d_f64_noalloc(None, 5, fake, fake, fake, 10, 10, fake,
                                                   None) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("stencil", ["9pt", "5pt"])
def test_noalloc_matches_work_buffer_entry_point(stencil):
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    lib = _abi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    for n in (1023, 301):   # second call smaller: stale partials of the first stay in the array
        if stencil == "9pt":
            so_np, b_np, _ = G.heterogeneous_9pt(n, seed=n, block=8, sigma=1.0)
        else:
            so_np, b_np = G.poisson_uniform_rhs(n, seed=n)
        rng = np.random.default_rng(n + 1)
        q_np = np.zeros_like(b_np)
        q_np[1:-1, 1:-1] = rng.uniform(-1, 1, (n, n))
        so, qf, q = (torch.as_tensor(a).cuda() for a in (so_np, b_np, q_np))
        ii, jj = n + 2, n + 2
        print(f"noalloc {stencil} n={n}", flush=True)
        work = torch.empty(lib.cdr_hip_norm_work2d(ii, jj), dtype=torch.float64, device="cuda")
        r_ref = torch.full_like(qf, 7.0)
        r_new = torch.full_like(qf, 7.0)
        s_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
        s_new = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
        assert lib.cdr_hip_residual_norm2d_f64(p(so), so.shape[0], p(qf), p(q), p(r_ref), ii, jj,
                                               p(work), p(s_ref), stream) == 0
        assert lib.cdr_hip_residual_norm2d_f64_noalloc(p(so), so.shape[0], p(qf), p(q), p(r_new),
                                                       ii, jj, p(s_new), stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(r_new, r_ref), n
        assert s_new.item() == s_ref.item(), (n, s_new.item(), s_ref.item())
        # and against the oracle's norm (same fixed-point data, different summation order)
        from oracle import oracle as O
        r_o = O.residual(so_np, b_np, q_np)
        assert abs(np.sqrt(s_new.item()) - O.norm(r_o)) <= 1e-12 * O.norm(r_o)

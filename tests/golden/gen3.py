"""Seeded random 3D operators for the oracle and kernel tests (test infrastructure)."""
import numpy as np

# plane d of a 27-point stencil at g couples g+a with g+b (0-based plane order
# p, pw, ps, b, psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw)
PAIRS3 = {1: ((0, 0, 0), (-1, 0, 0)), 2: ((0, 0, 0), (0, -1, 0)), 3: ((0, 0, 0), (0, 0, -1)),
          4: ((0, 0, 0), (-1, -1, 0)), 5: ((0, -1, 0), (-1, 0, 0)),
          6: ((0, 0, 0), (-1, 0, -1)), 7: ((0, -1, 0), (-1, 0, -1)),
          8: ((0, -1, 0), (0, 0, -1)), 9: ((-1, -1, 0), (0, 0, -1)),
          10: ((-1, 0, 0), (0, 0, -1)), 11: ((-1, 0, 0), (0, -1, -1)),
          12: ((0, 0, 0), (0, -1, -1)), 13: ((0, 0, 0), (-1, -1, -1))}


def coupling_mask(nx, ny, nz, d):
    """True where plane d couples two interior points."""
    a, b = PAIRS3[d]
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    ok = np.ones_like(i, dtype=bool)
    for off in (a, b):
        ii, jj, kk = i + off[0], j + off[1], k + off[2]
        ok &= (ii >= 1) & (ii <= nx) & (jj >= 1) & (jj <= ny) & (kk >= 1) & (kk <= nz)
    return ok


def rand_so3(rng, nx, ny, nz, ns, lo=0.1, hi=1.0):
    """Diagonally dominant M-matrix 7-point (ns 4) or 27-point (ns 14) stencil with couplings to
    ghost points zero and a zero ghost layer."""
    so = np.zeros((ns, nz + 2, ny + 2, nx + 2))
    for d in range(1, ns):
        so[d] = np.where(coupling_mask(nx, ny, nz, d), rng.uniform(lo, hi, so[d].shape), 0.0)
    so[0, 1:-1, 1:-1, 1:-1] = 2.0 * (ns - 1) + rng.uniform(0.0, 1.0, (nz, ny, nx))
    return so


def images3(a, mask):
    """Periodic images of a (..., kk, jj, ii) in place for the mask (bits x = 1, y = 2, z = 4),
    y first, then x, then z, over the full extents."""
    if mask & 2:
        a[..., 0, :] = a[..., -2, :]
        a[..., -1, :] = a[..., 1, :]
    if mask & 1:
        a[..., :, 0] = a[..., :, -2]
        a[..., :, -1] = a[..., :, 1]
    if mask & 4:
        a[..., 0, :, :] = a[..., -2, :, :]
        a[..., -1, :, :] = a[..., 1, :, :]
    return a


def rand_so3_periodic(rng, nx, ny, nz, ns, mask, lo=0.1, hi=1.0):
    """Diagonally dominant M-matrix stencil that couples across the boundaries of the periodic
    directions (mask bits x = 1, y = 2, z = 4), with every plane's periodic images filled."""
    so = np.zeros((ns, nz + 2, ny + 2, nx + 2))
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    interior = (i >= 1) & (i <= nx) & (j >= 1) & (j <= ny) & (k >= 1) & (k <= nz)
    for d in range(1, ns):
        ok = interior.copy()
        for off in PAIRS3[d]:
            ii_, jj_, kk_ = i + off[0], j + off[1], k + off[2]
            ok &= ((ii_ >= 1) & (ii_ <= nx)) | bool(mask & 1)
            ok &= ((jj_ >= 1) & (jj_ <= ny)) | bool(mask & 2)
            ok &= ((kk_ >= 1) & (kk_ <= nz)) | bool(mask & 4)
        so[d] = np.where(ok, rng.uniform(lo, hi, so[d].shape), 0.0)
    so[0, 1:-1, 1:-1, 1:-1] = 2.0 * (ns - 1) + rng.uniform(0.0, 1.0, (nz, ny, nx))
    return images3(so, mask)


def periodic_case(n, ns, mask):
    """(so, q, qf) of a periodic kernel case: seeded by its extents, stencil and mask."""
    nx, ny, nz = n
    rng = np.random.default_rng(1000 * mask + 100 * ns + nx * 7 + ny * 5 + nz * 3)
    so = rand_so3_periodic(rng, nx, ny, nz, ns, mask)
    shp = (nz + 2, ny + 2, nx + 2)
    return so, rng.standard_normal(shp), rng.standard_normal(shp)

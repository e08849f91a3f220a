"""Vertex-form matrix entries for the handle API's cedar_mat_set2d/3d, derived from a stencil
operator through the oracle's matvec (A applied to probe vectors), not through the storage map
that cedar_mat_set2d implements: entry (i, j, dir) is the matrix element A[(i, j), (i, j) + off]
with off the offset the direction names.  Couplings to ghost points are not matrix elements and
are left out (the operators here have none)."""
import numpy as np

from oracle import oracle as O
from oracle import oracle3 as O3

# bmg2_dir -> (di, dj)
OFFSETS2 = {0: (0, 0), 1: (-1, 0), 2: (0, -1), 3: (-1, -1), 4: (-1, 1), 5: (1, -1), 6: (0, 1),
            7: (1, 1), 8: (1, 0)}


def random_symmetric2(nx, ny, ns, seed, integer=False):
    """Random diagonally dominant symmetric operator without boundary couplings."""
    rng = np.random.default_rng(seed)
    draw = (lambda *s: rng.integers(1, 9, size=s).astype(float)) if integer else \
        (lambda *s: rng.uniform(0.1, 1.0, size=s))
    so = np.zeros((ns, ny + 2, nx + 2))
    so[1, 1:ny + 1, 2:nx + 1] = draw(ny, nx - 1)
    so[2, 2:ny + 1, 1:nx + 1] = draw(ny - 1, nx)
    if ns == 5:
        so[3, 2:ny + 1, 2:nx + 1] = draw(ny - 1, nx - 1)
        so[4, 2:ny + 1, 2:nx + 1] = draw(ny - 1, nx - 1)
    so[0, 1:ny + 1, 1:nx + 1] = 2 * ns + draw(ny, nx)
    return so


def entries2(so):
    """[(i, j, dir, value)] for every interior row and every direction whose neighbour is an
    interior point."""
    _, jj, ii = so.shape
    nx, ny = ii - 2, jj - 2
    cols = {}
    for a in range(3):
        for b in range(3):
            q = np.zeros((jj, ii))
            q[1 + b:ny + 1:3, 1 + a:nx + 1:3] = 1.0
            cols[(a, b)] = O.matvec(so, q)
    dirs = range(9) if so.shape[0] == 5 else (0, 1, 2, 6, 8)
    out = []
    for j in range(ny):
        for i in range(nx):
            for d in dirs:
                di, dj = OFFSETS2[d]
                ni, nj = i + di, j + dj
                if not (0 <= ni < nx and 0 <= nj < ny):
                    continue
                y = cols[(ni % 3, nj % 3)]
                out.append((i, j, d, float(y[j + 1, i + 1])))
    return out


def random_symmetric3(nx, ny, nz, ns, seed):
    """Random 3D operator (4 or 14 planes) in the stored (symmetric) form; interior positions
    only, so couplings to ghost points are included (matvec reads the ghosts of q)."""
    rng = np.random.default_rng(seed)
    so = np.zeros((ns, nz + 2, ny + 2, nx + 2))
    for d in range(1, ns):
        so[d, 1:nz + 1, 1:ny + 1, 1:nx + 1] = rng.uniform(0.1, 1.0, size=(nz, ny, nx))
    so[0, 1:nz + 1, 1:ny + 1, 1:nx + 1] = 4.0 * ns
    return so


def entries3(so):
    """[(i, j, k, dir, -stored)] for the stored planes at interior positions: cedar_mat_set3d
    takes the symmetric storage directly (values are matrix elements, so off-diagonals are the
    negated stored couplings)."""
    ns, kk, jj, ii = so.shape
    out = []
    for k in range(kk - 2):
        for j in range(jj - 2):
            for i in range(ii - 2):
                for d in range(ns):
                    v = so[d, k + 1, j + 1, i + 1]
                    out.append((i, j, k, d, float(v if d == 0 else -v)))
    return out


def matvec3(so, q):
    return O3.matvec(so, q)

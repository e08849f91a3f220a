"""Test operators and seeded synthetic problems, as numpy arrays in the solver's HBM layout.

Layout: a grid function with nx*ny interior points is a C-contiguous (ny+2, nx+2) float64 array
(the bytes of the Fortran array Q(nx+2, ny+2)); a 2D stencil is (nstncl, ny+2, nx+2) with planes
{c, w, s} (5-point) or {c, w, s, sw, nw} (9-point).  Off-diagonal couplings are stored negated,
couplings to the Dirichlet boundary are zero and all ghost entries are zero.

`poisson`, `diag_diffusion` and `fe` restate the reference gallery (src/2d/gallery.cc:7-113).
The diffusion problems below are defined here (the reference gallery has no heterogeneous
operator); the random stream is SplitMix64 so that any language can regenerate it.
"""
import numpy as np

C, W, S, SW, NW = 0, 1, 2, 3, 4

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n successive SplitMix64 outputs from `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_pm1(seed, n):
    return 2.0 * uniform01(seed, n) - 1.0


def normal(seed, n):
    """Standard normals by Box-Muller on pairs of uniforms (first output of each pair)."""
    u = uniform01(seed, 2 * n).reshape(n, 2)
    u1 = 1.0 - u[:, 0]  # (0, 1]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[:, 1])


# ---------------------------------------------------------------------------------------------
# reference gallery (src/2d/gallery.cc)
# ---------------------------------------------------------------------------------------------
def diag_diffusion(nx, ny, dx=1.0, dy=1.0):
    """5-point -dx u_xx - dy u_yy scaled by h^2 (src/2d/gallery.cc:42-74)."""
    so = np.zeros((3, ny + 2, nx + 2))
    hx = 1.0 / (nx + 1)
    hy = 1.0 / (ny + 1)
    xh = hy / hx
    yh = hx / hy
    # C++ index k  <->  array index k (ghost at 0): s for j in [2, ny], i in [1, nx]
    so[S, 2:ny + 1, 1:nx + 1] = dy * yh
    so[W, 1:ny + 1, 2:nx + 1] = dx * xh
    so[C, 1:ny + 1, 1:nx + 1] = 2 * dx * xh + 2 * dy * yh
    return so


def poisson(nx, ny):
    """5-point Poisson (src/2d/gallery.cc:7-39)."""
    so = np.zeros((3, ny + 2, nx + 2))
    hx = 1.0 / (nx + 1)
    hy = 1.0 / (ny + 1)
    xh = hy / hx
    yh = hx / hy
    so[S, 2:ny + 1, 1:nx + 1] = 1.0 * yh
    so[W, 1:ny + 1, 2:nx + 1] = 1.0 * xh
    so[C, 1:ny + 1, 1:nx + 1] = 2 * xh + 2 * yh
    return so


def fe(nx, ny):
    """All-ones 9-point finite-element stencil (src/2d/gallery.cc:77-113)."""
    so = np.zeros((5, ny + 2, nx + 2))
    so[S, 2:ny + 1, 1:nx + 1] = 1.0
    so[W, 1:ny + 1, 2:nx + 1] = 1.0
    so[SW, 2:ny + 1, 2:nx + 1] = 1.0
    so[NW, 2:ny + 1, 2:nx + 1] = 1.0
    so[C, 1:ny + 1, 1:nx + 1] = 8.0
    return so


# ---------------------------------------------------------------------------------------------
# heterogeneous diffusion problems
# ---------------------------------------------------------------------------------------------
def block_lognormal(ncx, ncy, block, sigma, seed, contrast=1e4):
    """Cell coefficients k (shape (ncy, ncx)), constant on block x block cells, log k ~ N(0, sigma^2)
    clipped to +-ln(contrast)/2 so that max k / min k <= contrast."""
    nbx = -(-ncx // block)
    nby = -(-ncy // block)
    g = sigma * normal(seed, nbx * nby).reshape(nby, nbx)
    lim = 0.5 * np.log(contrast)
    g = np.clip(g, -lim, lim)
    kb = np.exp(g)
    return np.repeat(np.repeat(kb, block, axis=0), block, axis=1)[:ncy, :ncx].copy()


def q1_diffusion(kcell):
    """9-point bilinear (Q1) finite-element stiffness of -div(k grad u) on a uniform grid.

    kcell has shape (ny+1, nx+1): cell (cj, ci) has corners at array indices
    (cj, ci), (cj, ci+1), (cj+1, ci), (cj+1, ci+1) of the (ny+2, nx+2) node grid (row/col 0 and
    ny+1/nx+1 are the Dirichlet boundary).  Element matrix k/6 [4 -1 -2 -1; ...]: an edge gets
    k/6 from each of its two cells, a diagonal pair k/3, the centre 2k/3 per cell.
    """
    ncy, ncx = kcell.shape
    ny, nx = ncy - 1, ncx - 1
    so = np.zeros((5, ny + 2, nx + 2))
    k = kcell
    # node (j, i) interior: 1..n.  Cells touching node (j,i): (j-1,i-1),(j-1,i),(j,i-1),(j,i)
    so[C, 1:ny + 1, 1:nx + 1] = (2.0 / 3.0) * (k[:-1, :-1] + k[:-1, 1:] + k[1:, :-1] + k[1:, 1:])
    # w at node (j,i): edge to (j, i-1), cells (j-1, i-1) and (j, i-1); zero when i-1 is boundary
    w = (k[:-1, :-1] + k[1:, :-1]) / 6.0           # indexed by node (j,i) for j,i in 1..n
    so[W, 1:ny + 1, 2:nx + 1] = w[:, 1:]
    # s at node (j,i): edge to (j-1, i), cells (j-1, i-1) and (j-1, i); zero when j-1 is boundary
    s = (k[:-1, :-1] + k[:-1, 1:]) / 6.0
    so[S, 2:ny + 1, 1:nx + 1] = s[1:, :]
    # sw at node (j,i): pair (j,i)-(j-1,i-1) in cell (j-1,i-1); nw at node (j,i): pair
    # (j-1,i)-(j,i-1) in cell (j-1,i-1).  Both zero when either end is on the boundary.
    d = k[:-1, :-1] / 3.0
    so[SW, 2:ny + 1, 2:nx + 1] = d[1:, 1:]
    so[NW, 2:ny + 1, 2:nx + 1] = d[1:, 1:]
    # nw at (j,i) links (j,i-1) and (j-1,i): (j-1,i) is interior for j>=2, (j,i-1) for i>=2;
    # also both must be interior rows/cols, already ensured by the slices above.
    return so


def rhs_uniform(nx, ny, seed):
    b = np.zeros((ny + 2, nx + 2))
    b[1:ny + 1, 1:nx + 1] = uniform_pm1(seed, nx * ny).reshape(ny, nx)
    return b


def heterogeneous_9pt(n, seed, block=16, sigma=2.0, contrast=1e4):
    """-div(k grad u) = f on n x n interior nodes: lognormal k on block x block cells, f ~ U[-1,1]."""
    k = block_lognormal(n + 1, n + 1, block, sigma, seed, contrast)
    so = q1_diffusion(k)
    b = rhs_uniform(n, n, seed + 1000003)
    return so, b, k


def poisson_uniform_rhs(n, seed=1):
    """poisson(n, n) with f ~ U[-1, 1] (seeded)."""
    return poisson(n, n), rhs_uniform(n, n, seed)


def readme_problem(n=400):
    """Cedar's example problem: poisson(n, n), b = 8 pi^2 sin(2 pi x) sin(2 pi y) h^2
    (examples/basic-2d-ser/poisson.cc:15-37); returns (so, b, exact)."""
    so = poisson(n, n)
    h = 1.0 / (n + 1)
    idx = np.arange(n + 2) * h
    X, Y = np.meshgrid(idx, idx)  # X[j, i] = i*h, Y[j, i] = j*h
    b = np.zeros((n + 2, n + 2))
    rhs = 8 * (np.pi * np.pi) * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    b[1:n + 1, 1:n + 1] = (rhs * (h * h))[1:n + 1, 1:n + 1]
    exact = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    return so, b, exact


def sine_problem(nx, ny):
    """poisson(nx, ny) with b = 8 pi^2 sin(2 pi x) sin(2 pi y) hx hy, hx = 1/(nx+1), hy = 1/(ny+1),
    and the continuous solution sin(2 pi x) sin(2 pi y) (test/2d/test_poisson.cc:13-61); returns
    (so, b, exact)."""
    so = poisson(nx, ny)
    hx, hy = 1.0 / (nx + 1), 1.0 / (ny + 1)
    X, Y = np.meshgrid(np.arange(nx + 2) * hx, np.arange(ny + 2) * hy)  # X[j, i] = i hx
    s = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    b = np.zeros((ny + 2, nx + 2))
    b[1:ny + 1, 1:nx + 1] = (8 * (np.pi * np.pi) * s * (hx * hy))[1:ny + 1, 1:nx + 1]
    return so, b, s


def periodic_problem(nx, ny, periodic=(True, False)):
    """Cedar's periodic example (examples/basic-2d-ser/periodic.cc:16-155): 5-point Poisson on
    nx x ny points, periodic in the flagged directions (a periodic direction of n points has
    period n, h = 1/n; its ghost rows/columns of the operator and of b hold the periodic images),
    b = 8 pi^2 sin(2 pi x) sin(2 pi y) h_x h_y; returns (so, b, exact), exact = sin(2 pi x)
    sin(2 pi y)."""
    px, py = bool(periodic[0]), bool(periodic[1])
    so = np.zeros((3, ny + 2, nx + 2))
    mx, my = nx - 1 if px else nx, ny - 1 if py else ny
    hx, hy = 1.0 / (mx + 1), 1.0 / (my + 1)
    xh, yh = hy / hx, hx / hy
    ibeg, jbeg = (1 if px else 2), (1 if py else 2)
    so[S, jbeg:ny + 1, 1:nx + 1] = 1.0 * yh
    so[W, 1:ny + 1, ibeg:nx + 1] = 1.0 * xh
    so[C, 1:ny + 1, 1:nx + 1] = 2 * xh + 2 * yh
    if px:
        so[:, 1:ny + 1, ibeg - 1] = so[:, 1:ny + 1, nx]
        so[:, 1:ny + 1, nx + 1] = so[:, 1:ny + 1, ibeg]
    if py:
        so[:, jbeg - 1, 1:nx + 1] = so[:, ny, 1:nx + 1]
        so[:, ny + 1, 1:nx + 1] = so[:, jbeg, 1:nx + 1]
    X, Y = np.meshgrid(np.arange(nx + 2) * hx, np.arange(ny + 2) * hy)
    b = np.zeros((ny + 2, nx + 2))
    b[1:ny + 1, 1:nx + 1] = (8 * (np.pi * np.pi) * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
                             * (hx * hy))[1:ny + 1, 1:nx + 1]
    if px:
        b[:, 0] = b[:, nx]
        b[:, nx + 1] = b[:, 1]
    if py:
        b[0, :] = b[ny, :]
        b[ny + 1, :] = b[1, :]
    exact = np.zeros((ny + 2, nx + 2))
    exact[1:ny + 1, 1:nx + 1] = (np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y))[1:ny + 1, 1:nx + 1]
    return so, b, exact


# ---------------------------------------------------------------------------------------------
# 3D: a grid function is (nz+2, ny+2, nx+2), a stencil (4, ...) {p, pw, ps, b} (7-point) or
# (14, ...) {p, pw, ps, b, psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw} (27-point)
# ---------------------------------------------------------------------------------------------
P3, PW3, PS3, B3 = 0, 1, 2, 3


def diag_diffusion3(nx, ny, nz, dx=1.0, dy=1.0, dz=1.0):
    """7-point -dx u_xx - dy u_yy - dz u_zz scaled by h^3 / h^2 (src/3d/gallery.cc:55-101)."""
    so = np.zeros((4, nz + 2, ny + 2, nx + 2))
    hx, hy, hz = 1.0 / (nx + 1), 1.0 / (ny + 1), 1.0 / (nz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    so[PS3, 1:nz + 1, 2:ny + 1, 1:nx + 1] = dy * yh
    so[PW3, 1:nz + 1, 1:ny + 1, 2:nx + 1] = dx * xh
    so[B3, 2:nz + 1, 1:ny + 1, 1:nx + 1] = dz * zh
    so[P3, 1:nz + 1, 1:ny + 1, 1:nx + 1] = 2.0 * dx * xh + 2.0 * dy * yh + 2.0 * dz * zh
    return so


def poisson3(nx, ny, nz):
    """7-point Poisson (src/3d/gallery.cc:7-52)."""
    so = np.zeros((4, nz + 2, ny + 2, nx + 2))
    hx, hy, hz = 1.0 / (nx + 1), 1.0 / (ny + 1), 1.0 / (nz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    so[PS3, 1:nz + 1, 2:ny + 1, 1:nx + 1] = 1.0 * yh
    so[PW3, 1:nz + 1, 1:ny + 1, 2:nx + 1] = 1.0 * xh
    so[B3, 2:nz + 1, 1:ny + 1, 1:nx + 1] = 1.0 * zh
    so[P3, 1:nz + 1, 1:ny + 1, 1:nx + 1] = 2.0 * xh + 2.0 * yh + 2.0 * zh
    return so


def periodic_problem3(nx, ny, nz, periodic=(True, True, True)):
    """Cedar's 3D periodic example (examples/basic-3d-mpi/periodic.cc:17-214) on one process:
    7-point Poisson on nx x ny x nz points, periodic in the flagged directions (a periodic
    direction of n points has period n, h = 1/n); the couplings are set up to the upper ghost
    plane of a periodic direction, as the example does, and the lower ghost planes of the
    operator and the ghosts of b hold the periodic images too.  b = 12 pi^2 sin(2 pi x)
    sin(2 pi y) sin(2 pi z) hx hy hz; returns (so, b, exact) with exact = sin(2 pi x) sin(2 pi y)
    sin(2 pi z) at x = i hx (i the array index)."""
    px, py, pz = (bool(p) for p in periodic)
    mx, my, mz = nx - px, ny - py, nz - pz
    hx, hy, hz = 1.0 / (mx + 1), 1.0 / (my + 1), 1.0 / (mz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    so = np.zeros((4, nz + 2, ny + 2, nx + 2))
    ib, jb, kb = (1 if px else 2), (1 if py else 2), (1 if pz else 2)
    ie, je, ke = nx + 1 + px, ny + 1 + py, nz + 1 + pz
    so[PS3, 1:ke, jb:je, 1:ie] = 1.0 * yh
    so[PW3, 1:ke, 1:je, ib:ie] = 1.0 * xh
    so[B3, kb:ke, 1:je, 1:ie] = 1.0 * zh
    so[P3, 1:nz + 1, 1:ny + 1, 1:nx + 1] = 2 * xh + 2 * yh + 2 * zh
    def images(a):
        if py:
            a[..., 0, :] = a[..., ny, :]
            a[..., ny + 1, :] = a[..., 1, :]
        if px:
            a[..., :, 0] = a[..., :, nx]
            a[..., :, nx + 1] = a[..., :, 1]
        if pz:
            a[..., 0, :, :] = a[..., nz, :, :]
            a[..., nz + 1, :, :] = a[..., 1, :, :]
        return a

    so = images(so)
    Z, Y, X = np.meshgrid(np.arange(nz + 2) * hz, np.arange(ny + 2) * hy, np.arange(nx + 2) * hx,
                          indexing="ij")
    u = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    b = np.zeros((nz + 2, ny + 2, nx + 2))
    b[1:-1, 1:-1, 1:-1] = (12 * np.pi ** 2 * u * (hx * hy * hz))[1:-1, 1:-1, 1:-1]
    exact = np.zeros_like(b)
    exact[1:-1, 1:-1, 1:-1] = u[1:-1, 1:-1, 1:-1]
    return so, images(b), exact


def fe3(nx, ny, nz):
    """All-ones 27-point stencil, centre 26 (src/3d/gallery.cc:104-187)."""
    so = np.zeros((14, nz + 2, ny + 2, nx + 2))
    I = slice(1, nx + 1)
    J = slice(1, ny + 1)
    K = slice(1, nz + 1)
    I2, J2, K2 = slice(2, nx + 1), slice(2, ny + 1), slice(2, nz + 1)
    so[1][K, J, I2] = 1.0                    # pw
    so[2][K, J2, I] = 1.0                    # ps
    so[3][K2, J, I] = 1.0                    # b
    so[5][K, J2, I2] = 1.0                   # pnw
    so[4][K, J2, I2] = 1.0                   # psw
    so[6][K2, J, I2] = 1.0                   # bw
    so[10][K2, J, I2] = 1.0                  # be
    so[8][K2, J2, I] = 1.0                   # bn
    so[12][K2, J2, I] = 1.0                  # bs
    for d in (7, 9, 11, 13):                 # bnw, bne, bse, bsw
        so[d][K2, J2, I2] = 1.0
    so[0][K, J, I] = 26.0
    return so


def rhs_uniform3(nx, ny, nz, seed):
    b = np.zeros((nz + 2, ny + 2, nx + 2))
    b[1:nz + 1, 1:ny + 1, 1:nx + 1] = uniform_pm1(seed, nx * ny * nz).reshape(nz, ny, nx)
    return b


def block_lognormal3(ncx, ncy, ncz, block, sigma, seed, contrast=None):
    """Cell coefficients (ncz, ncy, ncx), constant on block^3 cells, log k ~ N(0, sigma^2)
    (clipped to +-ln(contrast)/2 when contrast is given)."""
    nb = [-(-n // block) for n in (ncx, ncy, ncz)]
    g = sigma * normal(seed, nb[0] * nb[1] * nb[2]).reshape(nb[2], nb[1], nb[0])
    if contrast is not None:
        lim = 0.5 * np.log(contrast)
        g = np.clip(g, -lim, lim)
    kb = np.exp(g)
    k = np.repeat(np.repeat(np.repeat(kb, block, axis=0), block, axis=1), block, axis=2)
    return k[:ncz, :ncy, :ncx].copy()


def fv7_diffusion(kcell, aniso=(1.0, 1.0, 1.0)):
    """7-point cell-centred finite-volume -div(K grad u) on n^3 unknowns with Dirichlet ghosts.

    kcell has shape (nz, ny, nx), one coefficient per unknown; K = diag(aniso) * k.  A face
    between two unknowns gets the harmonic mean of their coefficients times h^2 / h (unit cube,
    h = 1/(n+1) per direction as in the reference gallery); couplings to the boundary are dropped
    from the stencil but their Dirichlet faces (coefficient of the cell itself) stay in the
    diagonal."""
    nz, ny, nx = kcell.shape
    hx, hy, hz = 1.0 / (nx + 1), 1.0 / (ny + 1), 1.0 / (nz + 1)
    xh, yh, zh = hy * hz / hx, hx * hz / hy, hx * hy / hz
    kx, ky, kz = (a * kcell for a in aniso)

    def harm(a, b):
        return 2.0 * a * b / (a + b)
    so = np.zeros((4, nz + 2, ny + 2, nx + 2))
    fw = xh * harm(kx[:, :, 1:], kx[:, :, :-1])      # face between i-1 and i (i = 1..nx-1)
    fs = yh * harm(ky[:, 1:, :], ky[:, :-1, :])
    fb = zh * harm(kz[1:, :, :], kz[:-1, :, :])
    so[PW3, 1:nz + 1, 1:ny + 1, 2:nx + 1] = fw
    so[PS3, 1:nz + 1, 2:ny + 1, 1:nx + 1] = fs
    so[B3, 2:nz + 1, 1:ny + 1, 1:nx + 1] = fb
    diag = np.zeros((nz, ny, nx))
    diag[:, :, 1:] += fw
    diag[:, :, :-1] += fw
    diag[:, 1:, :] += fs
    diag[:, :-1, :] += fs
    diag[1:, :, :] += fb
    diag[:-1, :, :] += fb
    # Dirichlet faces on the domain boundary use the cell's own coefficient
    diag[:, :, 0] += xh * kx[:, :, 0]
    diag[:, :, -1] += xh * kx[:, :, -1]
    diag[:, 0, :] += yh * ky[:, 0, :]
    diag[:, -1, :] += yh * ky[:, -1, :]
    diag[0, :, :] += zh * kz[0, :, :]
    diag[-1, :, :] += zh * kz[-1, :, :]
    so[P3, 1:nz + 1, 1:ny + 1, 1:nx + 1] = diag
    return so


def heterogeneous_7pt(n, seed=3, block=8, sigma=1.5, aniso=(1.0, 1.0, 1e-3)):
    """n^3 7-point finite-volume problem with anisotropy (kx, ky, kz) = aniso * k, log k ~
    N(0, sigma^2) on block^3 blocks, harmonic face averaging, f ~ U[-1, 1]."""
    k = block_lognormal3(n, n, n, block, sigma, seed)
    so = fv7_diffusion(k, aniso)
    b = rhs_uniform3(n, n, n, seed + 1000003)
    return so, b, k


def poisson3_sine(n):
    """poisson3(n, n, n) with b = 12 pi^2 sin(2 pi x) sin(2 pi y) sin(2 pi z) h^3 and the exact
    continuous solution (test/3d/test_poisson.cc:14-67); returns (so, b, exact)."""
    so = poisson3(n, n, n)
    h = 1.0 / (n + 1)
    t = np.arange(n + 2) * h
    Z, Y, X = np.meshgrid(t, t, t, indexing="ij")
    s = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    b = np.zeros((n + 2,) * 3)
    b[1:-1, 1:-1, 1:-1] = (12 * np.pi ** 2 * s * h ** 3)[1:-1, 1:-1, 1:-1]
    exact = np.zeros_like(b)
    exact[1:-1, 1:-1, 1:-1] = s[1:-1, 1:-1, 1:-1]
    return so, b, exact


def box_filter(a, w=5):
    """w x w moving average with edge replication (same shape as a)."""
    r = w // 2
    p = np.pad(a, r, mode="edge")
    c = np.cumsum(np.cumsum(p, axis=0), axis=1)
    c = np.pad(c, ((1, 0), (1, 0)))
    n0, n1 = a.shape
    s = c[w:w + n0, w:w + n1] - c[:n0, w:w + n1] - c[w:w + n0, :n1] + c[:n0, :n1]
    return s / (w * w)


def smooth_theta(n, seed):
    """Per-cell log-coefficients on (n+1) x (n+1) cells: theta ~ N(0, 1) smoothed by a 5-cell
    box filter."""
    return box_filter(normal(seed, (n + 1) * (n + 1)).reshape(n + 1, n + 1), 5)


# ---------------------------------------------------------------------------------------------
# workloads of many problems (batched and differentiated solves)
# ---------------------------------------------------------------------------------------------
def batched_set(nprob=96, sizes=(257, 513, 1025, 2049), seed=4, nrhs=4):
    """Many independent 9-point diffusion problems of mixed sizes.

    Side lengths are drawn uniformly from `sizes` (numpy Generator(seed)); problem k has its own
    lognormal block coefficients (sigma 2, 16x16-cell blocks, contrast <= 1e4, seed
    1000 seed + k) and `nrhs` right-hand sides f ~ U[-1, 1] (seed 10000 seed + nrhs k + r).
    Returns (side lengths, operators, rhs) with rhs[r][k] the r-th right-hand side of problem k.
    """
    rng = np.random.default_rng(seed)
    ns = [int(n) for n in rng.choice(list(sizes), size=nprob)]
    ops, rhs = [], [[] for _ in range(nrhs)]
    for k, n in enumerate(ns):
        so, _, _ = heterogeneous_9pt(n, seed=1000 * seed + k, block=16, sigma=2.0, contrast=1e4)
        ops.append(so)
        for r in range(nrhs):
            rhs[r].append(rhs_uniform(n, n, 10000 * seed + nrhs * k + r))
    return ns, ops, rhs


def differentiated_set(nb=16, n=1025):
    """Inputs of the differentiated solves: per problem t, theta_t = smooth_theta(n, 700 + t)
    (k = exp(theta) per cell), a second field smooth_theta(n, 900 + t) whose solution gives the
    observations, f ~ U[-1, 1] (seed 800 + t) and N(0, 1) observation noise (seed 1000 + t).
    Returns (theta, theta_obs, f, noise) stacked over the nb problems."""
    theta = np.stack([smooth_theta(n, 700 + t) for t in range(nb)])
    theta_obs = np.stack([smooth_theta(n, 900 + t) for t in range(nb)])
    f = np.stack([rhs_uniform(n, n, 800 + t) for t in range(nb)])
    noise = np.stack([normal(1000 + t, (n + 2) ** 2).reshape(n + 2, n + 2) for t in range(nb)])
    return theta, theta_obs, f, noise

// Differentiated 2D diffusion solves: assembly of the 9-point Q1 operator from per-cell
// coefficients, and the adjoint gradient kernel.  Not part of the reference (Cedar has no
// differentiation); these serve the torch.autograd wrapper in cedar_amd/adjoint.py.
//
// Operator: bilinear finite elements for -div(k grad u) on a uniform grid of nx x ny interior
// nodes, cell (cj, ci) (0 <= ci <= nx, 0 <= cj <= ny) having corner nodes (ci, cj), (ci+1, cj),
// (ci+1, cj+1), (ci, cj+1) (array indices, ghosts = boundary nodes).  Element matrix
// k_e K with K = 1/6 [4 -1 -2 -1; -1 4 -1 -2; -2 -1 4 -1; -1 -2 -1 4] (corners in that order);
// Dirichlet couplings dropped.  Stored as Cedar's 9-point stencil {c, w, s, sw, nw}, negated
// off-diagonals.
//
// Gradient: for L(u) with u = A(k)^{-1} f and lambda = A^{-1} dL/du (A is symmetric),
// dL/dk_e = -lambda_e^T K u_e over the element's corners, boundary corners contributing zero.
#include "../../include/cedar_hip.h"

#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline std::size_t pidx(int i, int j, int ii) {
    return (std::size_t)j * (std::size_t)ii + (std::size_t)i;
}

// one thread per node of the ghost-inclusive grid; kcell is (ny+1) x (nx+1), row-major in cj
__global__ __launch_bounds__(256) void q1_assemble_k(const double *__restrict__ kc,
                                                     double *__restrict__ so, int nx, int ny) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64);
    const int j = blockIdx.y * 4 + (threadIdx.x / 64);
    const int ii = nx + 2, jj = ny + 2, ic = nx + 1;
    if (i >= ii || j >= jj) return;
    const std::size_t P = (std::size_t)ii * jj, bz = blockIdx.z;
    kc += bz * (std::size_t)(nx + 1) * (ny + 1);
    so += bz * 5 * P;
    auto K = [&](int cj, int ci) { return kc[(std::size_t)cj * ic + ci]; };
    const bool in = i >= 1 && i <= nx && j >= 1 && j <= ny;
    const std::size_t p = pidx(i, j, ii);
    // centre: 4/6 k per touching cell; an edge: 1/6 k from each of its two cells; a diagonal
    // pair: 2/6 k from its cell (summation order of the gallery's numpy assembly)
    so[p] = in ? (2.0 / 3.0) * (K(j - 1, i - 1) + K(j - 1, i) + K(j, i - 1) + K(j, i)) : 0.0;
    so[P + p] = (in && i >= 2) ? (K(j - 1, i - 1) + K(j, i - 1)) / 6.0 : 0.0;
    so[2 * P + p] = (in && j >= 2) ? (K(j - 1, i - 1) + K(j - 1, i)) / 6.0 : 0.0;
    const double d = (in && i >= 2 && j >= 2) ? K(j - 1, i - 1) / 3.0 : 0.0;
    so[3 * P + p] = d;
    so[4 * P + p] = d;
}

// one thread per cell: g[cj, ci] = -(1/6) lambda_e^T K6 u_e, K6 = 6 K
__global__ __launch_bounds__(256) void q1_gradient_k(const double *__restrict__ u,
                                                     const double *__restrict__ lam,
                                                     double *__restrict__ g, int nx, int ny) {
    const int ci = blockIdx.x * 64 + (threadIdx.x % 64);
    const int cj = blockIdx.y * 4 + (threadIdx.x / 64);
    if (ci > nx || cj > ny) return;
    const int ii = nx + 2, jj = ny + 2;
    const std::size_t P = (std::size_t)ii * jj, bz = blockIdx.z;
    u += bz * P;
    lam += bz * P;
    g += bz * (std::size_t)(nx + 1) * (ny + 1);
    const int ni[4] = {ci, ci + 1, ci + 1, ci};
    const int nj[4] = {cj, cj, cj + 1, cj + 1};
    double uv[4], lv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const bool in = ni[a] >= 1 && ni[a] <= nx && nj[a] >= 1 && nj[a] <= ny;
        uv[a] = in ? u[pidx(ni[a], nj[a], ii)] : 0.0;
        lv[a] = in ? lam[pidx(ni[a], nj[a], ii)] : 0.0;
    }
    constexpr double K6[4][4] = {{4, -1, -2, -1}, {-1, 4, -1, -2}, {-2, -1, 4, -1}, {-1, -2, -1, 4}};
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) t += K6[a][b] * uv[b];
        s += lv[a] * t;
    }
    g[(std::size_t)cj * (nx + 1) + ci] = -s / 6.0;
}

inline dim3 grid2(int nx, int ny, int nb) {
    return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nb);
}

}  // namespace

extern "C" {

int cdr_hip_q1_assemble2d_f64(const double *kcell, double *so, int nx, int ny, int nb, void *st) {
    if (!kcell || !so || nx < 1 || ny < 1 || nb < 1 || nb > 65535 ||
        (long long)(nx + 2) * (ny + 2) >= (1LL << 31))
        return CDR_ERR_ARG;
    q1_assemble_k<<<grid2(nx + 2, ny + 2, nb), 256, 0, S(st)>>>(kcell, so, nx, ny);
    return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP;
}

int cdr_hip_q1_gradient2d_f64(const double *u, const double *lam, double *grad_k, int nx, int ny,
                              int nb, void *st) {
    if (!u || !lam || !grad_k || nx < 1 || ny < 1 || nb < 1 || nb > 65535 ||
        (long long)(nx + 2) * (ny + 2) >= (1LL << 31))
        return CDR_ERR_ARG;
    q1_gradient_k<<<grid2(nx + 1, ny + 1, nb), 256, 0, S(st)>>>(u, lam, grad_k, nx, ny);
    return hipGetLastError() == hipSuccess ? CDR_OK : CDR_ERR_HIP;
}

}  // extern "C"

// HIP kernels for the 3D BoxMG path on gfx950 (MI355X), FP64 unless templated on T.
//
// Layout and indices: arrays are Fortran order with one ghost layer; inside the kernels the
// reference's Fortran indices are used (1-based, ghosts at 1 and ii), so that each kernel can
// be read against the Fortran it restates (cited per kernel).  Sums are evaluated in the
// reference's left-to-right order and the library is built with -ffp-contract=off, so every
// kernel here reproduces the reference's rounding exactly, except the Galerkin product (a
// triple product P^T A P where the reference has per-direction formulas; equal to rounding).
//
// Grids: point kernels use 64 x 4 thread tiles in (i, j) and one grid z-slice per k plane.
#include "band_chol.hpp"
#include "bmg2_launch.hpp"
#include "bmg3_launch.hpp"

#include <hip/hip_runtime.h>

namespace cdr {
namespace k3 {

namespace {

constexpr int WAVE = 64;

// stencil planes, 1-based (src/3d/ftn/BMG_stencils_f90.h:46-48)
enum { kp = 1, kpw, kps, kb, kpsw, kpnw, kbw, kbnw, kbn, kbne, kbe, kbse, kbs, kbsw };
// prolongation planes, 1-based (src/3d/ftn/BMG_stencils_f90.h:59-65)
enum {
    lxyl = 1, lxyr, lxya, lxyb, lxza, lxzb, lxyne, lxyse, lxysw, lxynw, lxzsw, lxznw, lxzne,
    lxzse, lyzsw, lyznw, lyzne, lyzse, lbsw, lbnw, lbne, lbse, ltsw, ltnw, ltne, ltse
};

constexpr double EMACH = 0x1.0p-53;  // D1MACH(3)

__host__ __device__ inline std::size_t ix3(int i, int j, int k, int ii, int jj) {
    return (std::size_t)(i - 1) + (std::size_t)ii * ((std::size_t)(j - 1) + (std::size_t)jj * (std::size_t)(k - 1));
}

// fine-grid (ii, jj, kk; P points) and coarse-grid (iic, jjc, kkc; Pc points) accessors
#define S3(i, j, k, d) so[(std::size_t)((d)-1) * P + ix3((i), (j), (k), ii, jj)]
#define Q3(a, i, j, k) (a)[ix3((i), (j), (k), ii, jj)]
#define CI3(i, j, k, d) ci[(std::size_t)((d)-1) * Pc + ix3((i), (j), (k), iic, jjc)]
#define QC3(a, i, j, k) (a)[ix3((i), (j), (k), iic, jjc)]

__device__ inline double dmin2(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax2(double a, double b) { return a > b ? a : b; }
__device__ inline double min4(double a, double b, double c, double d) {
    return dmin2(dmin2(dmin2(a, b), c), d);
}
// lumped, stabilised OI denominators (src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:154-157,
// :273-275)
__device__ inline double lump3(double a, double b, double c, double p, double ep) {
    return a + b + (p - c) * dmax2(p - (1.0 + ep) * c, 0.0) / (fabs(p - (1.0 + ep) * c) + EMACH);
}
__device__ inline double lumpd(double dp, double s, double ep) {
    return dp + (s - dp) * dmax2(s - (1.0 + ep) * dp, 0.0) / (fabs(s - (1.0 + ep) * dp) + EMACH);
}

// ------------------------------------------------------------------------------------------
// The 26 neighbour terms of the 27-point operator in the reference's order
// (src/3d/ftn/BMG3_SymStd_residual.f90:71-97): X(si, sj, sk, plane, oi, oj, ok) = coefficient
// at (i,j,k) + s times q at (i,j,k) + o.  The 7-point operator is the sub-list below
// (src/3d/ftn/BMG3_SymStd_residual.f90:106-112).
// ------------------------------------------------------------------------------------------
#define NB27_LIST(X)                                                                              \
    X(0, 0, 0, kpw, -1, 0, 0) X(0, 1, 0, kpnw, -1, 1, 0) X(0, 1, 0, kps, 0, 1, 0)                 \
    X(1, 1, 0, kpsw, 1, 1, 0) X(1, 0, 0, kpw, 1, 0, 0) X(1, 0, 0, kpnw, 1, -1, 0)                 \
    X(0, 0, 0, kps, 0, -1, 0) X(0, 0, 0, kpsw, -1, -1, 0) X(0, 0, 0, kb, 0, 0, -1)                \
    X(0, 0, 0, kbw, -1, 0, -1) X(0, 1, 0, kbnw, -1, 1, -1) X(0, 1, 0, kbn, 0, 1, -1)              \
    X(1, 1, 0, kbne, 1, 1, -1) X(1, 0, 0, kbe, 1, 0, -1) X(1, 0, 0, kbse, 1, -1, -1)              \
    X(0, 0, 0, kbs, 0, -1, -1) X(0, 0, 0, kbsw, -1, -1, -1) X(0, 0, 1, kb, 0, 0, 1)               \
    X(0, 0, 1, kbe, -1, 0, 1) X(0, 1, 1, kbse, -1, 1, 1) X(0, 1, 1, kbs, 0, 1, 1)                 \
    X(1, 1, 1, kbsw, 1, 1, 1) X(1, 0, 1, kbw, 1, 0, 1) X(1, 0, 1, kbnw, 1, -1, 1)                 \
    X(0, 0, 1, kbn, 0, -1, 1) X(0, 0, 1, kbne, -1, -1, 1)
#define NB7_LIST(X)                                                                               \
    X(0, 0, 0, kpw, -1, 0, 0) X(0, 1, 0, kps, 0, 1, 0) X(1, 0, 0, kpw, 1, 0, 0)                   \
    X(0, 0, 0, kps, 0, -1, 0) X(0, 0, 0, kb, 0, 0, -1) X(0, 0, 1, kb, 0, 0, 1)

#define CDR3_ADD(si, sj, sk, d, oi, oj, ok)                                                       \
    v = v + S3(i + (si), j + (sj), k + (sk), d) * Q3(q, i + (oi), j + (oj), k + (ok));

// qf + sum of the off-diagonal terms at Fortran point (i, j, k)
template <typename T, bool NINE>
__device__ inline T offdiag3(const T *__restrict__ so, const T *__restrict__ qf,
                             const T *__restrict__ q, int i, int j, int k, int ii, int jj,
                             std::size_t P) {
    T v = Q3(qf, i, j, k);
    if (NINE) {
        NB27_LIST(CDR3_ADD)
    } else {
        NB7_LIST(CDR3_ADD)
    }
    return v;
}

// ------------------------------------------------------------------------------------------
// deterministic block reduction (256 threads)
// ------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ inline void block_sum_store(double v, double *dst) {
    __shared__ double part[4];
    v = wave_sum(v);
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) *dst = part[0] + part[1] + part[2] + part[3];
}

// ------------------------------------------------------------------------------------------
// residual (src/3d/ftn/BMG3_SymStd_residual.f90:63-121); with NORM also per-block sums of r^2
// ------------------------------------------------------------------------------------------
template <typename T, bool NINE, bool NORM>
__global__ __launch_bounds__(256) void residual3_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                   const T *__restrict__ q, T *__restrict__ res,
                                                   int ii, int jj, int kk,
                                                   double *__restrict__ partials) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int k = blockIdx.z + 2;
    const std::size_t P = (std::size_t)ii * jj * kk;
    double sq = 0.0;
    if (i <= ii - 1 && j <= jj - 1) {
        const T r = offdiag3<T, NINE>(so, qf, q, i, j, k, ii, jj, P) - S3(i, j, k, kp) * Q3(q, i, j, k);
        Q3(res, i, j, k) = r;
        sq = (double)r * (double)r;
    }
    if (NORM)
        block_sum_store(sq, &partials[((std::size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x]);
}

// Mixed-precision outer step: xnew = x + (double) e (e may be null: xnew = x), r = b - A xnew in
// FP64 (the residual's term order), r32 = (float) r, per-block sums of r^2.
#define CDR3_ADDX(si, sj, sk, d, oi, oj, ok)                                                      \
    v = v + S3(i + (si), j + (sj), k + (sk), d) * X(ix3(i + (oi), j + (oj), k + (ok), ii, jj));

template <bool NINE, bool HAS_E>
__global__ __launch_bounds__(256) void residual_ir3_k(const double *__restrict__ so,
                                                      const double *__restrict__ qf,
                                                      const double *__restrict__ x,
                                                      const float *__restrict__ e,
                                                      double *__restrict__ xnew,
                                                      float *__restrict__ r32, int ii, int jj,
                                                      int kk, double *__restrict__ partials) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int k = blockIdx.z + 2;
    const std::size_t P = (std::size_t)ii * jj * kk;
    double sq = 0.0;
    if (i <= ii - 1 && j <= jj - 1) {
        auto X = [&](std::size_t p) -> double { return HAS_E ? x[p] + (double)e[p] : x[p]; };
        double v = Q3(qf, i, j, k);
        if (NINE) {
            NB27_LIST(CDR3_ADDX)
        } else {
            NB7_LIST(CDR3_ADDX)
        }
        const std::size_t p = ix3(i, j, k, ii, jj);
        const double xc = X(p);
        const double r = v - S3(i, j, k, kp) * xc;
        xnew[p] = xc;
        r32[p] = (float)r;
        sq = r * r;
    }
    block_sum_store(sq, &partials[((std::size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x]);
}

// SOR(:,:,:,2) = 1 / SO_p over the interior (src/3d/ftn/BMG3_SymStd_SETUP_recip.f90:63-69)
__global__ __launch_bounds__(256) void recip3_k(const double *__restrict__ so,
                                                double *__restrict__ sor, int ii, int jj, int kk) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int k = blockIdx.z + 2;
    if (i > ii - 1 || j > jj - 1) return;
    const std::size_t P = (std::size_t)ii * jj * kk;
    sor[P + ix3(i, j, k, ii, jj)] = 1.0 / S3(i, j, k, kp);
}

// ------------------------------------------------------------------------------------------
// Gauss-Seidel, one launch per colour (src/3d/ftn/BMG3_SymStd_relax_GS.f90:79-181).
// 27-point: colour (pi, pj, pk) = points with i = pi + 2t (Fortran), likewise j, k.
// 7-point: colour pts: i starts at mod(j + k + pts, 2) + 2 in each (j, k) line.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void gs27_colour_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                     T *__restrict__ q, const T *__restrict__ sor,
                                                     int ii, int jj, int kk, int pi, int pj, int pk,
                                                     int rev) {
    // rev: planes visited in reverse, so that a colour starts where the previous one ended (in
    // the memory-side cache); the points of a colour are independent
    const int bz = rev ? (int)(gridDim.z - 1 - blockIdx.z) : (int)blockIdx.z;
    const int i = pi + 2 * (int)(blockIdx.x * 64 + (threadIdx.x % 64));
    const int j = pj + 2 * (int)(blockIdx.y * 4 + (threadIdx.x / 64));
    const int k = pk + 2 * bz;
    if (i > ii - 1 || j > jj - 1 || k > kk - 1) return;
    const std::size_t P = (std::size_t)ii * jj * kk;
    Q3(q, i, j, k) = offdiag3<T, true>(so, qf, q, i, j, k, ii, jj, P) * sor[P + ix3(i, j, k, ii, jj)];
}

template <typename T>
__global__ __launch_bounds__(256) void gs7_colour_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                    T *__restrict__ q, const T *__restrict__ sor,
                                                    int ii, int jj, int kk, int pts, int rev) {
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int k = (rev ? (int)(gridDim.z - 1 - blockIdx.z) : (int)blockIdx.z) + 2;
    if (j > jj - 1) return;
    const int i = (j + k + pts) % 2 + 2 + 2 * (int)(blockIdx.x * 64 + (threadIdx.x % 64));
    if (i > ii - 1) return;
    const std::size_t P = (std::size_t)ii * jj * kk;
    Q3(q, i, j, k) = offdiag3<T, false>(so, qf, q, i, j, k, ii, jj, P) * sor[P + ix3(i, j, k, ii, jj)];
}

// ------------------------------------------------------------------------------------------
// restriction qc = R q = P^T q (src/3d/ftn/BMG3_SymStd_restrict.f90:110-141), per coarse point
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void restrict3_k(const T *__restrict__ ci, const T *__restrict__ q,
                                                   T *__restrict__ qc, int iif, int jjf, int kkf,
                                                   int iic, int jjc, int kkc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int kc = blockIdx.z + 2;
    if (ic > iic - 1 || jc > jjc - 1) return;
    const int ii = iif, jj = jjf;
    const std::size_t Pc = (std::size_t)iic * jjc * kkc;
    const int i = 2 * (ic - 1), j = 2 * (jc - 1), k = 2 * (kc - 1);
    QC3(qc, ic, jc, kc) =
        CI3(ic, jc, kc, lxyne) * Q3(q, i - 1, j - 1, k) + CI3(ic, jc, kc, lxya) * Q3(q, i, j - 1, k) +
        CI3(ic + 1, jc, kc, lxynw) * Q3(q, i + 1, j - 1, k) + CI3(ic, jc, kc, lxyr) * Q3(q, i - 1, j, k) +
        Q3(q, i, j, k) + CI3(ic + 1, jc, kc, lxyl) * Q3(q, i + 1, j, k) +
        CI3(ic, jc + 1, kc, lxyse) * Q3(q, i - 1, j + 1, k) + CI3(ic, jc + 1, kc, lxyb) * Q3(q, i, j + 1, k) +
        CI3(ic + 1, jc + 1, kc, lxysw) * Q3(q, i + 1, j + 1, k) +
        CI3(ic, jc, kc, ltne) * Q3(q, i - 1, j - 1, k - 1) + CI3(ic, jc, kc, lyznw) * Q3(q, i, j - 1, k - 1) +
        CI3(ic + 1, jc, kc, ltnw) * Q3(q, i + 1, j - 1, k - 1) +
        CI3(ic, jc, kc, lxzne) * Q3(q, i - 1, j, k - 1) + CI3(ic, jc, kc, lxza) * Q3(q, i, j, k - 1) +
        CI3(ic + 1, jc, kc, lxznw) * Q3(q, i + 1, j, k - 1) +
        CI3(ic, jc + 1, kc, ltse) * Q3(q, i - 1, j + 1, k - 1) +
        CI3(ic, jc + 1, kc, lyzne) * Q3(q, i, j + 1, k - 1) +
        CI3(ic + 1, jc + 1, kc, ltsw) * Q3(q, i + 1, j + 1, k - 1) +
        CI3(ic, jc, kc + 1, lbne) * Q3(q, i - 1, j - 1, k + 1) +
        CI3(ic, jc, kc + 1, lyzsw) * Q3(q, i, j - 1, k + 1) +
        CI3(ic + 1, jc, kc + 1, lbnw) * Q3(q, i + 1, j - 1, k + 1) +
        CI3(ic, jc, kc + 1, lxzse) * Q3(q, i - 1, j, k + 1) + CI3(ic, jc, kc + 1, lxzb) * Q3(q, i, j, k + 1) +
        CI3(ic + 1, jc, kc + 1, lxzsw) * Q3(q, i + 1, j, k + 1) +
        CI3(ic, jc + 1, kc + 1, lbse) * Q3(q, i - 1, j + 1, k + 1) +
        CI3(ic, jc + 1, kc + 1, lyzse) * Q3(q, i, j + 1, k + 1) +
        CI3(ic + 1, jc + 1, kc + 1, lbsw) * Q3(q, i + 1, j + 1, k + 1);
}

// ------------------------------------------------------------------------------------------
// interp_add q += P qc + res / SO_p (src/3d/ftn/BMG3_SymStd_interp_add.f90:86-212), one thread
// per fine point.  The reference's three loop nests cover: fine planes k = 2(kc-1) carrying
// coarse points (2D pattern), and fine-only planes k = 2kc-3 split by the parity of i.  Their
// index ranges reach the fine ghost layer for even extents (coarse index IIC = IICF1); those
// points take the raw res there, as the reference divides res only in the interior.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void interp3_k(T *__restrict__ q, const T *__restrict__ qc,
                                                 const T *__restrict__ res, const T *__restrict__ so,
                                                 const T *__restrict__ ci, int iic, int jjc, int kkc,
                                                 int iif, int jjf, int kkf) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int k = blockIdx.z + 2;
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    const int imax1 = 2 * (iicf1 - 1), jmax1 = 2 * (jjcf1 - 1);
    if (j > jmax1) return;
    const bool keven = (k % 2) == 0;
    if (keven) {
        if (k > 2 * (kkc - 2) || i > imax1) return;
    } else {
        if (k < 3 || k > 2 * kkcf1 - 3) return;
        if ((i % 2) == 0 ? i > 2 * (iic - 2) : (i < 3 || i > 2 * iicf1 - 3)) return;
    }
    const bool interior = i <= iif - 1 && j <= jjf - 1 && k <= kkf - 1;
    const T r = interior ? Q3(res, i, j, k) / S3(i, j, k, kp) : Q3(res, i, j, k);
    T &qv = Q3(q, i, j, k);
    if (keven) {
        const int kc = k / 2 + 1;
        const bool ie = (i % 2) == 0, je = (j % 2) == 0;
        if (ie && je) {
            qv = qv + QC3(qc, i / 2 + 1, j / 2 + 1, kc);
        } else if (!ie && je) {
            const int ic = (i + 1) / 2 + 1, jc = j / 2 + 1;
            const T a = CI3(ic, jc, kc, lxyr) * QC3(qc, ic, jc, kc) + CI3(ic, jc, kc, lxyl) * QC3(qc, ic - 1, jc, kc);
            qv = qv + a + r;
        } else if (ie && !je) {
            const int ic = i / 2 + 1, jc = (j + 1) / 2 + 1;
            const T a = CI3(ic, jc, kc, lxya) * QC3(qc, ic, jc, kc) + CI3(ic, jc, kc, lxyb) * QC3(qc, ic, jc - 1, kc);
            qv = qv + a + r;
        } else {
            const int ic = (i + 1) / 2 + 1, jc = (j + 1) / 2 + 1;
            const T a = CI3(ic, jc, kc, lxysw) * QC3(qc, ic - 1, jc - 1, kc) +
                        CI3(ic, jc, kc, lxynw) * QC3(qc, ic - 1, jc, kc) +
                        CI3(ic, jc, kc, lxyne) * QC3(qc, ic, jc, kc) +
                        CI3(ic, jc, kc, lxyse) * QC3(qc, ic, jc - 1, kc);
            qv = qv + a + r;
        }
        return;
    }
    const int kc = (k + 3) / 2;
    const bool je = (j % 2) == 0;
    if ((i % 2) == 0) {
        const int ic = i / 2 + 1;
        if (je) {
            const int jc = j / 2 + 1;
            qv = qv + CI3(ic, jc, kc, lxza) * QC3(qc, ic, jc, kc) + CI3(ic, jc, kc, lxzb) * QC3(qc, ic, jc, kc - 1) + r;
        } else {
            const int jc = (j + 1) / 2 + 1;
            qv = qv + CI3(ic, jc, kc, lyznw) * QC3(qc, ic, jc, kc) + CI3(ic, jc, kc, lyzne) * QC3(qc, ic, jc - 1, kc) +
                 CI3(ic, jc, kc, lyzsw) * QC3(qc, ic, jc, kc - 1) +
                 CI3(ic, jc, kc, lyzse) * QC3(qc, ic, jc - 1, kc - 1) + r;
        }
    } else {
        const int ic = (i + 3) / 2;
        if (je) {
            const int jc = j / 2 + 1;
            qv = qv + CI3(ic, jc, kc, lxznw) * QC3(qc, ic - 1, jc, kc) + CI3(ic, jc, kc, lxzne) * QC3(qc, ic, jc, kc) +
                 CI3(ic, jc, kc, lxzsw) * QC3(qc, ic - 1, jc, kc - 1) +
                 CI3(ic, jc, kc, lxzse) * QC3(qc, ic, jc, kc - 1) + r;
        } else {
            const int jc = (j + 1) / 2 + 1;
            qv = qv + CI3(ic, jc, kc, ltnw) * QC3(qc, ic - 1, jc, kc) + CI3(ic, jc, kc, ltne) * QC3(qc, ic, jc, kc) +
                 CI3(ic, jc, kc, ltsw) * QC3(qc, ic - 1, jc - 1, kc) +
                 CI3(ic, jc, kc, ltse) * QC3(qc, ic, jc - 1, kc) +
                 CI3(ic, jc, kc, lbnw) * QC3(qc, ic - 1, jc, kc - 1) +
                 CI3(ic, jc, kc, lbne) * QC3(qc, ic, jc, kc - 1) +
                 CI3(ic, jc, kc, lbsw) * QC3(qc, ic - 1, jc - 1, kc - 1) +
                 CI3(ic, jc, kc, lbse) * QC3(qc, ic, jc - 1, kc - 1) + r;
        }
    }
}

// ------------------------------------------------------------------------------------------
// operator-induced interpolation (src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:86-808), one
// thread per coarse point in three launches: edge weights, then face weights (they read edge
// weights of neighbouring coarse points), then cell-centre weights (they read both).  Each
// device function below is one loop body of the reference (27-point branch when NINE, 7-point
// branch otherwise) for the fine point tied to coarse (ic, jc, kc) through the fine anchor
// (i, j, k): 2 (ic-1), ... in the definite branch, wrapped indices in the periodic one.
// ------------------------------------------------------------------------------------------
#define OI_ARGS const double *__restrict__ so, double *__restrict__ ci, int ii, int jj, std::size_t P, \
                int iic, int jjc, std::size_t Pc, int ic, int jc, int kc, int i, int j, int k
#define OI_PASS so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, 2 * (ic - 1), 2 * (jc - 1), 2 * (kc - 1)

template <bool NINE>
__device__ inline void oi_xedge(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i - 1, j, k, kp);
    double a, b, c, ep;
    if (nine) {
        a = S3(i - 1, j + 1, k, kpnw) + S3(i - 1, j, k, kpw) + S3(i - 1, j, k, kpsw) +
            S3(i - 1, j + 1, k, kbnw) + S3(i - 1, j, k, kbw) + S3(i - 1, j, k, kbsw) +
            S3(i - 1, j + 1, k + 1, kbse) + S3(i - 1, j, k + 1, kbe) +
            S3(i - 1, j, k + 1, kbne);
        b = S3(i, j + 1, k, kpsw) + S3(i, j, k, kpw) + S3(i, j, k, kpnw) +
            S3(i, j + 1, k, kbne) + S3(i, j, k, kbe) + S3(i, j, k, kbse) +
            S3(i, j + 1, k + 1, kbsw) + S3(i, j, k + 1, kbw) + S3(i, j, k + 1, kbnw);
        c = a + b + S3(i - 1, j, k, kps) + S3(i - 1, j + 1, k, kps) +
            S3(i - 1, j + 1, k, kbn) + S3(i - 1, j, k, kb) + S3(i - 1, j, k, kbs) +
            S3(i - 1, j + 1, k + 1, kbs) + S3(i - 1, j, k + 1, kb) +
            S3(i - 1, j, k + 1, kbn);
        ep = dmin2(fabs(a / p0), fabs(b / p0));
    } else {
        a = S3(i - 1, j, k, kpw);
        b = S3(i, j, k, kpw);
        ep = dmin2(fabs(a / p0), fabs(b) / p0);
        c = a + b + S3(i - 1, j, k, kps) + S3(i - 1, j + 1, k, kps) + S3(i - 1, j, k, kb) +
            S3(i - 1, j, k + 1, kb);
    }
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxyl) = a / c;
    CI3(ic, jc, kc, lxyr) = b / c;
}

template <bool NINE>
__device__ inline void oi_yedge(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i, j - 1, k, kp);
    double a, b, c;
    if (nine) {
        a = S3(i, j, k, kpnw) + S3(i, j, k, kps) + S3(i + 1, j, k, kpsw) +
            S3(i, j, k, kbnw) + S3(i, j, k, kbn) + S3(i + 1, j, k, kbne) +
            S3(i, j, k + 1, kbse) + S3(i, j, k + 1, kbs) + S3(i + 1, j, k + 1, kbsw);
        b = S3(i, j - 1, k, kpsw) + S3(i, j - 1, k, kps) + S3(i + 1, j - 1, k, kpnw) +
            S3(i, j - 1, k, kbsw) + S3(i, j - 1, k, kbs) + S3(i + 1, j - 1, k, kbse) +
            S3(i, j - 1, k + 1, kbne) + S3(i, j - 1, k + 1, kbn) +
            S3(i + 1, j - 1, k + 1, kbnw);
        c = a + b + S3(i, j - 1, k, kpw) + S3(i + 1, j - 1, k, kpw) +
            S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kb) + S3(i + 1, j - 1, k, kbe) +
            S3(i, j - 1, k + 1, kbe) + S3(i, j - 1, k + 1, kb) +
            S3(i + 1, j - 1, k + 1, kbw);
    } else {
        a = S3(i, j, k, kps);
        b = S3(i, j - 1, k, kps);
        c = a + b + S3(i, j - 1, k, kpw) + S3(i + 1, j - 1, k, kpw) + S3(i, j - 1, k, kb) +
            S3(i, j - 1, k + 1, kb);
    }
    const double ep = dmin2(fabs(a / p0), fabs(b / p0));
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxya) = a / c;
    CI3(ic, jc, kc, lxyb) = b / c;
}

template <bool NINE>
__device__ inline void oi_zedge(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i, j, k - 1, kp);
    double a, b, c;
    if (nine) {
        a = S3(i, j + 1, k, kbse) + S3(i, j + 1, k, kbs) + S3(i + 1, j + 1, k, kbsw) +
            S3(i, j, k, kbe) + S3(i, j, k, kb) + S3(i + 1, j, k, kbw) +
            S3(i, j, k, kbne) + S3(i, j, k, kbn) + S3(i + 1, j, k, kbnw);
        b = S3(i, j + 1, k - 1, kbnw) + S3(i, j + 1, k - 1, kbn) +
            S3(i + 1, j + 1, k - 1, kbne) + S3(i, j, k - 1, kbw) + S3(i, j, k - 1, kb) +
            S3(i + 1, j, k - 1, kbe) + S3(i, j, k - 1, kbsw) + S3(i, j, k - 1, kbs) +
            S3(i + 1, j, k - 1, kbse);
        c = a + b + S3(i, j, k - 1, kpw) + S3(i + 1, j, k - 1, kpw) +
            S3(i, j + 1, k - 1, kpnw) + S3(i, j + 1, k - 1, kps) +
            S3(i + 1, j + 1, k - 1, kpsw) + S3(i, j, k - 1, kpsw) +
            S3(i, j, k - 1, kps) + S3(i + 1, j, k - 1, kpnw);
    } else {
        a = S3(i, j, k, kb);
        b = S3(i, j, k - 1, kb);
        c = a + b + S3(i, j, k - 1, kpw) + S3(i + 1, j, k - 1, kpw) +
            S3(i, j + 1, k - 1, kps) + S3(i, j, k - 1, kps);
    }
    const double ep = dmin2(fabs(a / p0), fabs(b / p0));
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxza) = a / c;
    CI3(ic, jc, kc, lxzb) = b / c;
}

template <bool NINE>
__device__ inline void oi_xyface(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i - 1, j - 1, k, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i - 1, j, k, kpnw) + S3(i - 1, j, k, kbnw) + S3(i - 1, j, k + 1, kbse);
        dn = S3(i - 1, j, k, kps) + S3(i - 1, j, k, kbn) + S3(i - 1, j, k + 1, kbs);
        dne = S3(i, j, k, kpsw) + S3(i, j, k, kbne) + S3(i, j, k + 1, kbsw);
        dw = S3(i - 1, j - 1, k, kpw) + S3(i - 1, j - 1, k, kbw) + S3(i - 1, j - 1, k + 1, kbe);
        de = S3(i, j - 1, k, kpw) + S3(i, j - 1, k, kbe) + S3(i, j - 1, k + 1, kbw);
        dsw = S3(i - 1, j - 1, k, kpsw) + S3(i - 1, j - 1, k, kbsw) +
              S3(i - 1, j - 1, k + 1, kbne);
        ds = S3(i - 1, j - 1, k, kps) + S3(i - 1, j - 1, k, kbs) + S3(i - 1, j - 1, k + 1, kbn);
        dse = S3(i, j - 1, k, kpnw) + S3(i, j - 1, k, kbse) + S3(i, j - 1, k + 1, kbnw);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i - 1, j, k, kps);
        dw = S3(i - 1, j - 1, k, kpw);
        de = S3(i, j - 1, k, kpw);
        ds = S3(i - 1, j - 1, k, kps);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i - 1, j - 1, k, kb) - S3(i - 1, j - 1, k + 1, kb);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lxynw) = dp * (dnw + CI3(ic - 1, jc, kc, lxya) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxyne) = dp * (dne + CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxya) * de);
        CI3(ic, jc, kc, lxyse) = dp * (dse + CI3(ic, jc, kc, lxyb) * de + CI3(ic, jc - 1, kc, lxyr) * ds);
        CI3(ic, jc, kc, lxysw) = dp * (dsw + CI3(ic, jc - 1, kc, lxyl) * ds + CI3(ic - 1, jc, kc, lxyb) * dw);
    } else {
        CI3(ic, jc, kc, lxynw) = dp * (CI3(ic - 1, jc, kc, lxya) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxyne) = dp * (CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxya) * de);
        CI3(ic, jc, kc, lxyse) = dp * (CI3(ic, jc, kc, lxyb) * de + CI3(ic, jc - 1, kc, lxyr) * ds);
        CI3(ic, jc, kc, lxysw) = dp * (CI3(ic, jc - 1, kc, lxyl) * ds + CI3(ic - 1, jc, kc, lxyb) * dw);
    }
}

template <bool NINE>
__device__ inline void oi_xzface(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i - 1, j, k - 1, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i - 1, j + 1, k, kbse) + S3(i - 1, j, k, kbe) + S3(i - 1, j, k, kbne);
        dn = S3(i - 1, j + 1, k, kbs) + S3(i - 1, j, k, kb) + S3(i - 1, j, k, kbn);
        dne = S3(i, j + 1, k, kbsw) + S3(i, j, k, kbw) + S3(i, j, k, kbnw);
        dw = S3(i - 1, j + 1, k - 1, kpnw) + S3(i - 1, j, k - 1, kpw) + S3(i - 1, j, k - 1, kpsw);
        de = S3(i, j + 1, k - 1, kpsw) + S3(i, j, k - 1, kpw) + S3(i, j, k - 1, kpnw);
        dsw = S3(i - 1, j + 1, k - 1, kbnw) + S3(i - 1, j, k - 1, kbw) +
              S3(i - 1, j, k - 1, kbsw);
        ds = S3(i - 1, j + 1, k - 1, kbn) + S3(i - 1, j, k - 1, kb) + S3(i - 1, j, k - 1, kbs);
        dse = S3(i, j + 1, k - 1, kbne) + S3(i, j, k - 1, kbe) + S3(i, j, k - 1, kbse);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i - 1, j, k, kb);
        dw = S3(i - 1, j, k - 1, kpw);
        de = S3(i, j, k - 1, kpw);
        ds = S3(i - 1, j, k - 1, kb);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i - 1, j + 1, k - 1, kps) - S3(i - 1, j, k - 1, kps);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lxznw) = dp * (dnw + CI3(ic - 1, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxzne) = dp * (dne + CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxza) * de);
        CI3(ic, jc, kc, lxzse) = dp * (dse + CI3(ic, jc, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyr) * ds);
        CI3(ic, jc, kc, lxzsw) = dp * (dsw + CI3(ic, jc, kc - 1, lxyl) * ds + CI3(ic - 1, jc, kc, lxzb) * dw);
    } else {
        CI3(ic, jc, kc, lxznw) = dp * (CI3(ic - 1, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxzne) = dp * (CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxza) * de);
        CI3(ic, jc, kc, lxzse) = dp * (CI3(ic, jc, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyr) * ds);
        CI3(ic, jc, kc, lxzsw) = dp * (CI3(ic, jc, kc - 1, lxyl) * ds + CI3(ic - 1, jc, kc, lxzb) * dw);
    }
}

template <bool NINE>
__device__ inline void oi_yzface(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i, j - 1, k - 1, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i, j, k, kbse) + S3(i, j, k, kbs) + S3(i + 1, j, k, kbsw);
        dn = S3(i, j - 1, k, kbe) + S3(i, j - 1, k, kb) + S3(i + 1, j - 1, k, kbw);
        dne = S3(i, j - 1, k, kbne) + S3(i, j - 1, k, kbn) + S3(i + 1, j - 1, k, kbnw);
        dw = S3(i, j, k - 1, kpnw) + S3(i, j, k - 1, kps) + S3(i + 1, j, k - 1, kpsw);
        de = S3(i, j - 1, k - 1, kpsw) + S3(i, j - 1, k - 1, kps) + S3(i + 1, j - 1, k - 1, kpnw);
        dsw = S3(i, j, k - 1, kbnw) + S3(i, j, k - 1, kbn) + S3(i + 1, j, k - 1, kbne);
        ds = S3(i, j - 1, k - 1, kbw) + S3(i, j - 1, k - 1, kb) + S3(i + 1, j - 1, k - 1, kbe);
        dse = S3(i, j - 1, k - 1, kbsw) + S3(i, j - 1, k - 1, kbs) +
              S3(i + 1, j - 1, k - 1, kbse);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i, j - 1, k, kb);
        dw = S3(i, j, k - 1, kps);
        de = S3(i, j - 1, k - 1, kps);
        ds = S3(i, j - 1, k - 1, kb);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i, j - 1, k - 1, kpw) - S3(i + 1, j - 1, k - 1, kpw);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lyznw) = dp * (dnw + CI3(ic, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxya) * dn);
        CI3(ic, jc, kc, lyzne) = dp * (dne + CI3(ic, jc, kc, lxyb) * dn + CI3(ic, jc - 1, kc, lxza) * de);
        CI3(ic, jc, kc, lyzse) = dp * (dse + CI3(ic, jc - 1, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyb) * ds);
        CI3(ic, jc, kc, lyzsw) = dp * (dsw + CI3(ic, jc, kc - 1, lxya) * ds + CI3(ic, jc, kc, lxzb) * dw);
    } else {
        CI3(ic, jc, kc, lyznw) = dp * (CI3(ic, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxya) * dn);
        CI3(ic, jc, kc, lyzne) = dp * (CI3(ic, jc, kc, lxyb) * dn + CI3(ic, jc - 1, kc, lxza) * de);
        CI3(ic, jc, kc, lyzse) = dp * (CI3(ic, jc - 1, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyb) * ds);
        CI3(ic, jc, kc, lyzsw) = dp * (CI3(ic, jc, kc - 1, lxya) * ds + CI3(ic, jc, kc, lxzb) * dw);
    }
}

template <bool NINE>
__device__ inline void oi_cube(OI_ARGS) {
    constexpr bool nine = NINE;
    const double p0 = S3(i - 1, j - 1, k - 1, kp);
    if (nine) {
        double yp = S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kpnw) +
                    S3(i - 1, j, k - 1, kps) + S3(i, j, k - 1, kpsw) +
                    S3(i, j - 1, k - 1, kpw) + S3(i, j - 1, k - 1, kpnw) +
                    S3(i - 1, j - 1, k - 1, kps) + S3(i - 1, j - 1, k - 1, kpsw) +
                    S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k - 1, kbw) +
                    S3(i - 1, j, k - 1, kbnw) + S3(i - 1, j, k - 1, kbn) +
                    S3(i, j, k - 1, kbne) + S3(i, j - 1, k - 1, kbe) +
                    S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                    S3(i - 1, j - 1, k - 1, kbsw) + S3(i - 1, j - 1, k, kb) +
                    S3(i - 1, j - 1, k, kbe) + S3(i - 1, j, k, kbse) +
                    S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) + S3(i, j - 1, k, kbw) +
                    S3(i, j - 1, k, kbnw) + S3(i - 1, j - 1, k, kbn) +
                    S3(i - 1, j - 1, k, kbne);
        /* as in the reference, the third and fifth sums are not divided by SO_p,
           and kbse enters the fourth sum twice */
        double ye = min4(
            fabs(S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kpnw) +
                 S3(i - 1, j, k, kbse) + S3(i - 1, j - 1, k, kbe) +
                 S3(i - 1, j - 1, k, kbne) + S3(i - 1, j - 1, k - 1, kpsw) +
                 S3(i - 1, j - 1, k - 1, kbsw) + S3(i - 1, j - 1, k - 1, kbw) +
                 S3(i - 1, j, k - 1, kbnw)) / p0,
            fabs(S3(i, j - 1, k - 1, kpw) + S3(i, j, k - 1, kpsw) + S3(i, j, k, kbsw) +
                 S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kbnw) + S3(i, j - 1, k - 1, kpnw) +
                 S3(i, j - 1, k - 1, kbse) + S3(i, j - 1, k - 1, kbe) +
                 S3(i, j, k - 1, kbne)) / p0,
            fabs(S3(i - 1, j, k - 1, kps) + S3(i - 1, j, k - 1, kpnw) +
                 S3(i - 1, j, k, kbse) + S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) +
                 S3(i, j, k - 1, kpsw) + S3(i, j, k - 1, kbne) + S3(i - 1, j, k - 1, kbn) +
                 S3(i - 1, j, k - 1, kbnw)),
            fabs(S3(i - 1, j - 1, k - 1, kps) + S3(i - 1, j - 1, k - 1, kpsw) +
                 S3(i - 1, j - 1, k, kbne) + S3(i - 1, j - 1, k, kbn) +
                 S3(i, j - 1, k, kbnw) + S3(i, j - 1, k - 1, kpnw) +
                 S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                 S3(i, j - 1, k - 1, kbse)) / p0);
        ye = dmin2(dmin2(ye,
                         fabs(S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k - 1, kbw) +
                              S3(i - 1, j, k - 1, kbnw) + S3(i - 1, j, k - 1, kbn) +
                              S3(i, j, k - 1, kbne) + S3(i, j - 1, k - 1, kbe) +
                              S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                              S3(i - 1, j - 1, k - 1, kbsw))),
                   fabs(S3(i - 1, j - 1, k, kb) + S3(i - 1, j - 1, k, kbe) +
                        S3(i - 1, j, k, kbse) + S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) +
                        S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kbnw) +
                        S3(i - 1, j - 1, k, kbn) + S3(i - 1, j - 1, k, kbne)) / p0);
        yp = lumpd(yp, p0, ye);
        yp = 1.0 / yp;
        CI3(ic, jc, kc, ltnw) = yp * (S3(i - 1, j, k, kbse) +
            CI3(ic - 1, jc, kc, lyznw) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic - 1, jc, kc, lxza) * S3(i - 1, j, k - 1, kpnw) +
            CI3(ic, jc, kc, lxznw) * S3(i - 1, j, k - 1, kps) +
            CI3(ic - 1, jc, kc, lxya) * S3(i - 1, j - 1, k, kbe) +
            CI3(ic, jc, kc, lxyl) * S3(i - 1, j, k, kbs) +
            CI3(ic, jc, kc, lxynw) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, ltne) = yp * (S3(i, j, k, kbsw) +
            CI3(ic, jc, kc, lxzne) * S3(i - 1, j, k - 1, kps) +
            CI3(ic, jc, kc, lxza) * S3(i, j, k - 1, kpsw) +
            CI3(ic, jc, kc, lyznw) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc, kc, lxyr) * S3(i - 1, j, k, kbs) +
            CI3(ic, jc, kc, lxya) * S3(i, j - 1, k, kbw) +
            CI3(ic, jc, kc, lxyne) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, lbnw) = yp * (S3(i - 1, j, k - 1, kbnw) +
            CI3(ic - 1, jc, kc - 1, lxya) * S3(i - 1, j - 1, k - 1, kbw) +
            CI3(ic, jc, kc - 1, lxyl) * S3(i - 1, j, k - 1, kbn) +
            CI3(ic, jc, kc - 1, lxynw) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic - 1, jc, kc, lyzsw) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic - 1, jc, kc, lxzb) * S3(i - 1, j, k - 1, kpnw) +
            CI3(ic, jc, kc, lxzsw) * S3(i - 1, j, k - 1, kps));
        CI3(ic, jc, kc, lbne) = yp * (S3(i, j, k - 1, kbne) +
            CI3(ic, jc, kc - 1, lxyne) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc, kc - 1, lxyr) * S3(i - 1, j, k - 1, kbn) +
            CI3(ic, jc, kc - 1, lxya) * S3(i, j - 1, k - 1, kbe) +
            CI3(ic, jc, kc, lxzse) * S3(i - 1, j, k - 1, kps) +
            CI3(ic, jc, kc, lxzb) * S3(i, j, k - 1, kpsw) +
            CI3(ic, jc, kc, lyzsw) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbsw) = yp * (S3(i - 1, j - 1, k - 1, kbsw) +
            CI3(ic - 1, jc, kc - 1, lxyb) * S3(i - 1, j - 1, k - 1, kbw) +
            CI3(ic, jc, kc - 1, lxysw) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc - 1, kc - 1, lxyl) * S3(i - 1, j - 1, k - 1, kbs) +
            CI3(ic - 1, jc, kc, lyzse) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxzsw) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic - 1, jc - 1, kc, lxzb) * S3(i - 1, j - 1, k - 1, kpsw));
        CI3(ic, jc, kc, ltsw) = yp * (S3(i - 1, j - 1, k, kbne) +
            CI3(ic - 1, jc, kc, lxyb) * S3(i - 1, j - 1, k, kbe) +
            CI3(ic, jc, kc, lxysw) * S3(i - 1, j - 1, k, kb) +
            CI3(ic, jc - 1, kc, lxyl) * S3(i - 1, j - 1, k, kbn) +
            CI3(ic - 1, jc, kc, lyzne) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxznw) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic - 1, jc - 1, kc, lxza) * S3(i - 1, j - 1, k - 1, kpsw));
        CI3(ic, jc, kc, ltse) = yp * (S3(i, j - 1, k, kbnw) +
            CI3(ic, jc - 1, kc, lxyr) * S3(i - 1, j - 1, k, kbn) +
            CI3(ic, jc, kc, lxyse) * S3(i - 1, j - 1, k, kb) +
            CI3(ic, jc, kc, lxyb) * S3(i, j - 1, k, kbw) +
            CI3(ic, jc - 1, kc, lxzne) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic, jc, kc, lyzne) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxza) * S3(i, j - 1, k - 1, kpnw));
        CI3(ic, jc, kc, lbse) = yp * (S3(i, j - 1, k - 1, kbse) +
            CI3(ic, jc - 1, kc - 1, lxyr) * S3(i - 1, j - 1, k - 1, kbs) +
            CI3(ic, jc, kc - 1, lxyse) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc, kc - 1, lxyb) * S3(i, j - 1, k - 1, kbe) +
            CI3(ic, jc - 1, kc, lxzse) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic, jc, kc, lyzse) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxzb) * S3(i, j - 1, k - 1, kpnw));
    } else {
        double dp = S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kps) +
                    S3(i, j - 1, k - 1, kpw) + S3(i - 1, j - 1, k - 1, kps) +
                    S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k, kb);
        const double ep = dmin2(
            dmin2(min4(fabs(S3(i - 1, j - 1, k - 1, kpw) / p0), fabs(S3(i - 1, j, k - 1, kps) / p0),
                       fabs(S3(i, j - 1, k - 1, kpw) / p0), fabs(S3(i - 1, j - 1, k - 1, kps) / p0)),
                  fabs(S3(i - 1, j - 1, k - 1, kb) / p0)),
            fabs(S3(i - 1, j - 1, k, kb) / p0));
        /* here the reference adds the correction before dp */
        dp = (p0 - dp) * dmax2(p0 - (1.0 + ep) * dp, 0.0) /
                 (fabs(p0 - (1.0 + ep) * dp) + EMACH) + dp;
        dp = 1.0 / dp;
        CI3(ic, jc, kc, ltnw) = dp * (CI3(ic - 1, jc, kc, lyznw) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxznw) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lxynw) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, ltne) = dp * (CI3(ic, jc, kc, lxzne) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lyznw) * S3(i, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxyne) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, lbnw) = dp * (CI3(ic, jc, kc - 1, lxynw) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic - 1, jc, kc, lyzsw) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxzsw) * S3(i - 1, j, k - 1, kps));
        CI3(ic, jc, kc, lbne) = dp * (CI3(ic, jc, kc - 1, lxyne) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic, jc, kc, lxzse) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzsw) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbsw) = dp * (CI3(ic, jc, kc - 1, lxysw) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic - 1, jc, kc, lyzse) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc - 1, kc, lxzsw) * S3(i - 1, j - 1, k - 1, kps));
        CI3(ic, jc, kc, ltsw) = dp * (CI3(ic, jc, kc, lxysw) * S3(i - 1, j - 1, k, kb) +
                                      CI3(ic - 1, jc, kc, lyzne) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc - 1, kc, lxznw) * S3(i - 1, j - 1, k - 1, kps));
        CI3(ic, jc, kc, ltse) = dp * (CI3(ic, jc, kc, lxyse) * S3(i - 1, j - 1, k, kb) +
                                      CI3(ic, jc - 1, kc, lxzne) * S3(i - 1, j - 1, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzne) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbse) = dp * (CI3(ic, jc, kc - 1, lxyse) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic, jc - 1, kc, lxzse) * S3(i - 1, j - 1, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzse) * S3(i, j - 1, k - 1, kpw));
    }
}

template <bool NINE>
__global__ __launch_bounds__(256) void oi_edges_k(const double *__restrict__ so, double *__restrict__ ci,
                                                  int iif, int jjf, int kkf, int iic, int jjc, int kkc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int kc = blockIdx.z + 2;
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    if (kc <= kkc1 && jc <= jjc1 && ic >= 3 && ic <= iicf1) oi_xedge<NINE>(OI_PASS);
    if (kc <= kkc1 && jc >= 3 && jc <= jjcf1 && ic <= iic1) oi_yedge<NINE>(OI_PASS);
    if (kc >= 3 && kc <= kkcf1 && jc <= jjc1 && ic <= iic1) oi_zedge<NINE>(OI_PASS);
}

template <bool NINE>
__global__ __launch_bounds__(256) void oi_faces_k(const double *__restrict__ so, double *__restrict__ ci,
                                                  int iif, int jjf, int kkf, int iic, int jjc, int kkc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int kc = blockIdx.z + 2;
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    if (kc <= kkc1 && jc >= 3 && jc <= jjcf1 && ic >= 3 && ic <= iicf1) oi_xyface<NINE>(OI_PASS);
    if (kc >= 3 && kc <= kkcf1 && jc <= jjc1 && ic >= 3 && ic <= iicf1) oi_xzface<NINE>(OI_PASS);
    if (kc >= 3 && kc <= kkcf1 && jc >= 3 && jc <= jjcf1 && ic <= iic1) oi_yzface<NINE>(OI_PASS);
}

template <bool NINE>
__global__ __launch_bounds__(256) void oi_cube_k(const double *__restrict__ so, double *__restrict__ ci,
                                                 int iif, int jjf, int kkf, int iic, int jjc, int kkc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 3;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 3;
    const int kc = blockIdx.z + 3;
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    if (kc <= kkcf1 && jc <= jjcf1 && ic <= iicf1) oi_cube<NINE>(OI_PASS);
}

// ------------------------------------------------------------------------------------------
// Galerkin coarse operator as the triple product P^T A P (the reference forms the same product
// with explicit formulas, src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90 and
// BMG3_SymStd_SETUP_ITLI27_ex.f90).  P is the prolongation applied by interp_add: fine offset
// d from coarse point C has weight CI(C + max(d, 0), plane(d)), weight 1 at d = 0.  A is the fine
// operator, A(f, f + o) = -so(f + s(o), plane(o)) with (s, plane) from the neighbour list.
// Fine points outside the interior carry no unknown and are skipped.  One thread per coarse
// point computes its 14 entries; loop order fixes the rounding.
// ------------------------------------------------------------------------------------------
// tables indexed by (dz+1)*9 + (dy+1)*3 + (dx+1)
__constant__ int PW3_PLANE[27] = {
    ltne, lyznw, ltnw, lxzne, lxza, lxznw, ltse, lyzne, ltsw,
    lxyne, lxya, lxynw, lxyr, 0, lxyl, lxyse, lxyb, lxysw,
    lbne, lyzsw, lbnw, lxzse, lxzb, lxzsw, lbse, lyzse, lbsw};
// stored position s and plane of A(f, f + o)
__constant__ signed char A_SI[27] = {0, 0, 1, 0, 0, 1, 0, 0, 1,  0, 0, 1, 0, 0, 1, 0, 0, 1,  0, 0, 1, 0, 0, 1, 0, 0, 1};
__constant__ signed char A_SJ[27] = {0, 0, 0, 0, 0, 0, 1, 1, 1,  0, 0, 0, 0, 0, 0, 1, 1, 1,  0, 0, 0, 0, 0, 0, 1, 1, 1};
__constant__ signed char A_SK[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0,  0, 0, 0, 0, 0, 0, 0, 0, 0,  1, 1, 1, 1, 1, 1, 1, 1, 1};
__constant__ signed char A_PL[27] = {
    kbsw, kbs, kbse, kbw, kb, kbe, kbnw, kbn, kbne,
    kpsw, kps, kpnw, kpw, kp, kpw, kpnw, kps, kpsw,
    kbne, kbn, kbnw, kbe, kb, kbw, kbse, kbs, kbsw};

__device__ inline double pweight3(const double *__restrict__ ci, std::size_t Pc, int iic, int jjc,
                                  int ic, int jc, int kc, int dx, int dy, int dz) {
    if (dx == 0 && dy == 0 && dz == 0) return 1.0;
    const int d = PW3_PLANE[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
    return CI3(ic + (dx > 0), jc + (dy > 0), kc + (dz > 0), d);
}

template <bool NINE>
__device__ inline double aentry3(const double *__restrict__ so, std::size_t P, int ii, int jj, int i,
                                 int j, int k, int ox, int oy, int oz) {
    if (ox == 0 && oy == 0 && oz == 0) return S3(i, j, k, kp);
    const int t = (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1);
    return -S3(i + A_SI[t], j + A_SJ[t], k + A_SK[t], A_PL[t]);
}

template <bool NINE>
__device__ double kpair3(const double *__restrict__ so, const double *__restrict__ ci, int iif, int jjf,
                         int kkf, int iic, int jjc, int kkc, int icx, int jcx, int kcx, int DX, int DY,
                         int DZ) {
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const int X0 = 2 * (icx - 1), Y0 = 2 * (jcx - 1), Z0 = 2 * (kcx - 1);
    double sum = 0.0;
#pragma unroll 1
    for (int fz = -1; fz <= 1; ++fz)
#pragma unroll 1
        for (int fy = -1; fy <= 1; ++fy)
#pragma unroll 1
            for (int fx = -1; fx <= 1; ++fx) {
                const int x = X0 + fx, y = Y0 + fy, z = Z0 + fz;
                if (x < 2 || x > iif - 1 || y < 2 || y > jjf - 1 || z < 2 || z > kkf - 1) continue;
                double inner = 0.0;
                bool any = false;
#pragma unroll 1
                for (int oz = -1; oz <= 1; ++oz)
#pragma unroll 1
                    for (int oy = -1; oy <= 1; ++oy)
#pragma unroll 1
                        for (int ox = -1; ox <= 1; ++ox) {
                            const int gx = fx + ox - 2 * DX, gy = fy + oy - 2 * DY, gz = fz + oz - 2 * DZ;
                            if (gx < -1 || gx > 1 || gy < -1 || gy > 1 || gz < -1 || gz > 1) continue;
                            const int xg = x + ox, yg = y + oy, zg = z + oz;
                            if (xg < 2 || xg > iif - 1 || yg < 2 || yg > jjf - 1 || zg < 2 || zg > kkf - 1)
                                continue;
                            if (!NINE && (ox != 0) + (oy != 0) + (oz != 0) > 1) continue;
                            inner += aentry3<NINE>(so, P, ii, jj, x, y, z, ox, oy, oz) *
                                     pweight3(ci, Pc, iic, jjc, icx + DX, jcx + DY, kcx + DZ, gx, gy, gz);
                            any = true;
                        }
                if (any) sum += pweight3(ci, Pc, iic, jjc, icx, jcx, kcx, fx, fy, fz) * inner;
            }
    return sum;
}

// coarse plane d at g couples g + a with g + b
__constant__ signed char CP_D[13] = {kpw, kps, kb, kpsw, kpnw, kbw, kbnw, kbn, kbne, kbe, kbse, kbs, kbsw};
__constant__ signed char CP_A[13][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, -1, 0},
                                        {0, -1, 0}, {-1, -1, 0}, {-1, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {0, 0, 0}};
__constant__ signed char CP_B[13][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 0, 0}, {-1, 0, -1},
                                        {-1, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, -1, -1}, {0, -1, -1},
                                        {-1, -1, -1}};

template <bool NINE>
__global__ __launch_bounds__(256) void galerkin3_k(const double *__restrict__ so,
                                                   const double *__restrict__ ci,
                                                   double *__restrict__ soc, int iif, int jjf, int kkf,
                                                   int iic, int jjc, int kkc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int kc = blockIdx.z + 2;
    if (ic > iic - 1 || jc > jjc - 1) return;
    const std::size_t Pc = (std::size_t)iic * jjc * kkc, pc = ix3(ic, jc, kc, iic, jjc);
    soc[pc] = kpair3<NINE>(so, ci, iif, jjf, kkf, iic, jjc, kkc, ic, jc, kc, 0, 0, 0);
#pragma unroll 1
    for (int t = 0; t < 13; ++t) {
        const int xa = ic + CP_A[t][0], ya = jc + CP_A[t][1], za = kc + CP_A[t][2];
        const int xb = ic + CP_B[t][0], yb = jc + CP_B[t][1], zb = kc + CP_B[t][2];
        double v = 0.0;
        if (xa >= 2 && ya >= 2 && za >= 2 && xb >= 2 && yb >= 2 && zb >= 2)
            v = -kpair3<NINE>(so, ci, iif, jjf, kkf, iic, jjc, kkc, xa, ya, za, xb - xa, yb - ya, zb - za);
        soc[(std::size_t)(CP_D[t] - 1) * Pc + pc] = v;
    }
}

// ------------------------------------------------------------------------------------------
// coarsest grid (src/3d/ftn/BMG3_SymStd_SETUP_cg_LU.f90:100-193,
// src/3d/ftn/BMG3_SymStd_SOLVE_cg.f90:78-160): band assembly in the reference's row order
// (later rows win where rows coincide on tiny grids), then the shared band Cholesky
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void setup_cg3_k(const double *__restrict__ so, int nstncl,
                                                   double *__restrict__ abd, int ld, int ii, int jj,
                                                   int kk, int *__restrict__ info) {
    const std::size_t P = (std::size_t)ii * jj * kk;
    const int i1 = ii - 1, j1 = jj - 1, i2 = i1 - 1;
    const int ibw = i2 * j1 + 1;
    const int nx = ii - 2, ny = jj - 2, n = nx * ny * (kk - 2);
    const bool nine = nstncl == 14;
    auto AB = [&](int r, int c) -> double & { return abd[(std::size_t)(c - 1) * ld + (r - 1)]; };
    for (int kl0 = threadIdx.x; kl0 < n; kl0 += blockDim.x) {
        const int i = kl0 % nx + 2, j = (kl0 / nx) % ny + 2, k = kl0 / (nx * ny) + 2;
        const int kl = kl0 + 1;
        for (int r = 1; r <= ld; ++r) AB(r, kl) = 0.0;
        AB(ibw + 1, kl) = S3(i, j, k, kp);
        AB(ibw, kl) = -S3(i, j, k, kpw);
        AB(ibw - i1 + 3, kl) = nine ? -S3(i + 1, j, k, kpnw) : 0.0;
        AB(ibw - i1 + 2, kl) = -S3(i, j, k, kps);
        AB(ibw - i1 + 1, kl) = nine ? -S3(i, j, k, kpsw) : 0.0;
        AB(ibw - (j1 - 2) * i2 + 2, kl) = nine ? -S3(i + 1, j + 1, k, kbne) : 0.0;
        AB(ibw - (j1 - 2) * i2 + 1, kl) = nine ? -S3(i, j + 1, k, kbn) : 0.0;
        AB(ibw - (j1 - 2) * i2, kl) = nine ? -S3(i, j + 1, k, kbnw) : 0.0;
        AB(ibw - (j1 - 1) * i2 + 2, kl) = nine ? -S3(i + 1, j, k, kbe) : 0.0;
        AB(ibw - (j1 - 1) * i2 + 1, kl) = -S3(i, j, k, kb);
        AB(ibw - (j1 - 1) * i2, kl) = nine ? -S3(i, j, k, kbw) : 0.0;
        AB(3, kl) = nine ? -S3(i + 1, j, k, kbse) : 0.0;
        AB(2, kl) = nine ? -S3(i, j, k, kbs) : 0.0;
        AB(1, kl) = nine ? -S3(i, j, k, kbsw) : 0.0;
    }
    __syncthreads();
    __shared__ int bad;
    band_factor_upper(abd, n, ibw, ld, bad);
    if (threadIdx.x == 0 && bad) atomicMax(info, bad);
}

// One thread runs the two triangular solves, each step of which reads what the previous one
// wrote.  Where the factor (ld * n) and the iterate (n) fit in CG3_LDS doubles, the workgroup
// first stages both into LDS, so that those steps wait on LDS and not on global-memory round
// trips (4^3 coarsest grid of the 255^3 hierarchy: about 300 us per solve through memory);
// otherwise the iterate lives in the global scratch x.  Same operations either way.
constexpr int CG3_LDS = 6144;
template <typename T>
__global__ __launch_bounds__(256) void solve_cg3_k(T *__restrict__ q, const T *__restrict__ qf,
                                                   const double *__restrict__ abd, int ld,
                                                   double *__restrict__ x, int ii, int jj, int kk) {
    __shared__ double sh[CG3_LDS];
    const int nx = ii - 2, ny = jj - 2, n = nx * ny * (kk - 2);
    const int kd = (ii - 2) * (jj - 1) + 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    const bool lds = (std::size_t)n * (ld + 1) <= (std::size_t)CG3_LDS;
    double *xx = lds ? sh : x;
    const double *ab = abd;
    if (lds) {
        for (int t = tid; t < ld * n; t += nt) sh[n + t] = abd[t];
        ab = sh + n;
    }
    for (int t = tid; t < n; t += nt)
        xx[t] = (double)Q3(qf, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2);
    __syncthreads();
    if (lds && kd < 64 && nt >= 64) {  // bitwise the same sweeps, one wavefront
        if (tid < 64) band_solve_upper_wave(ab, n, kd, ld, xx, tid);
    } else if (tid == 0) {
        band_solve_upper(ab, n, kd, ld, xx);
    }
    __syncthreads();
    for (int t = tid; t < n; t += nt)
        Q3(q, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2) = (T)xx[t];
}

// ------------------------------------------------------------------------------------------
// Periodic boundaries: the JPN != 0 branches of the 3D kernels.  jpn is the BoxMG boundary id
// BMG_get_bc returns for the periodic mask (bits x = 1, y = 2, z = 4): 2 x, 1 y, 3 xy, 5 z,
// 6 xz, 7 yz, 8 xyz (include/cedar/2d/ftn/BMG_parameters_c.h:193-204).  The images of a
// periodic direction live in its ghost layer: q(1) = q(n-1), q(n) = q(2).
// ------------------------------------------------------------------------------------------
// one axis' images, q(1) = q(n-1) then q(n) = q(2) along `axis`, on the lines a = a0 + astep t
// (t < na), b = b0 + bstep u (u < nb) of the other two axes (x: a = j, b = k; y: a = i, b = k;
// z: a = i, b = j); lines are independent
template <typename T>
__global__ __launch_bounds__(256) void images3_k(T *__restrict__ q, int ii, int jj, int kk,
                                                 int axis, int a0, int astep, int na, int b0,
                                                 int bstep, int nb) {
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t >= na * nb) return;
    const int a = a0 + astep * (t % na), b = b0 + bstep * (t / na);
    if (axis == 0) {
        Q3(q, 1, a, b) = Q3(q, ii - 1, a, b);
        Q3(q, ii, a, b) = Q3(q, 2, a, b);
    } else if (axis == 1) {
        Q3(q, a, 1, b) = Q3(q, a, jj - 1, b);
        Q3(q, a, jj, b) = Q3(q, a, 2, b);
    } else {
        Q3(q, a, b, 1) = Q3(q, a, b, kk - 1);
        Q3(q, a, b, kk) = Q3(q, a, b, 2);
    }
}

// One loop nest of the periodic interpolation setup (src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90
// :809-2784): per axis, the coarse index runs c0..c1 with the fine index 2 (c-1) (mode 0) or
// advanced by I = MAX(MOD(I+2, M), MIN(I+2, 3)) from f0 (mode 1), or is the single pair
// (c0, f0) (mode 2).  The loop bodies are the definite branch's (oi_* above); a nest writes
// only its own weight planes and reads others, so its points are independent.
struct oi_axis {
    int mode, c0, c1, f0, m;
};
__device__ inline int oi_wnext(int prev, int n) {
    const int a = (prev + 2) % n, b = prev + 2 < 3 ? prev + 2 : 3;
    return a > b ? a : b;
}
__device__ inline void oi_at(const oi_axis &a, int t, int &c, int &f) {
    if (a.mode == 2) {
        c = a.c0;
        f = a.f0;
        return;
    }
    c = a.c0 + t;
    if (a.mode == 0) {
        f = 2 * (c - 1);
        return;
    }
    f = a.f0;
    for (int u = 0; u <= t; ++u) f = oi_wnext(f, a.m);
}

template <bool NINE, int TYPE>
__global__ __launch_bounds__(256) void oi_nest_k(const double *__restrict__ so, double *__restrict__ ci,
                                                 int iif, int jjf, int kkf, int iic, int jjc, int kkc,
                                                 oi_axis ax, oi_axis ay, oi_axis az, int nx, int ny,
                                                 int nz) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nx * ny * nz) return;
    int ic, i, jc, j, kc, k;
    oi_at(ax, (int)(t % nx), ic, i);
    oi_at(ay, (int)((t / nx) % ny), jc, j);
    oi_at(az, (int)(t / ((long long)nx * ny)), kc, k);
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    if (TYPE == 0) oi_xedge<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 1) oi_yedge<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 2) oi_zedge<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 3) oi_xyface<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 4) oi_xzface<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 5) oi_yzface<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
    if (TYPE == 6) oi_cube<NINE>(so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc, i, j, k);
}

// A group of up to four weight copies CI(dst) = CI(src) of the periodic interpolation setup,
// executed in order for each value v = lo..hi of a loop variable; a coordinate of -1 stands
// for v.  The groups copy within one row, column or plane per v, so values of v are
// independent.
struct ci_copy {
    int n, lo, hi;
    int dst[4][4], src[4][4];  // (ic, jc, kc, plane)
};
__global__ __launch_bounds__(256) void ci_copy_k(double *__restrict__ ci, int iic, int jjc, int kkc,
                                                 ci_copy cp) {
    const int v = cp.lo + (int)(blockIdx.x * 256 + threadIdx.x);
    if (v > cp.hi) return;
    const std::size_t Pc = (std::size_t)iic * jjc * kkc;
    for (int s = 0; s < cp.n; ++s) {
        int d[4], r[4];
        for (int a = 0; a < 4; ++a) {
            d[a] = cp.dst[s][a] == -1 ? v : cp.dst[s][a];
            r[a] = cp.src[s][a] == -1 ? v : cp.src[s][a];
        }
        CI3(d[0], d[1], d[2], d[3]) = CI3(r[0], r[1], r[2], r[3]);
    }
}

// Galerkin product with periodic directions (bit 0 x, 1 y, 2 z of pm): the fine points of the
// product are taken in the unwrapped geometry between X and Y, so each stored coupling is the
// one across its own side, as the reference's hand-expanded products form it; only the array
// reads wrap (coarse indices into 2..n+1, fine points into the interior).  Summation order as
// in kpair3.
__device__ inline int wrap_in(int n, int m) { return ((n - 2) % m + m) % m + 2; }

template <bool NINE>
__device__ double kpair3w(const double *__restrict__ so, const double *__restrict__ ci, int iif,
                          int jjf, int kkf, int iic, int jjc, int kkc, int icx, int jcx, int kcx,
                          int DX, int DY, int DZ, int pm) {
    const int ii = iif, jj = jjf;
    const std::size_t P = (std::size_t)iif * jjf * kkf, Pc = (std::size_t)iic * jjc * kkc;
    const bool px = pm & 1, py = pm & 2, pz = pm & 4;
    if (px) icx = wrap_in(icx, iic - 2);
    if (py) jcx = wrap_in(jcx, jjc - 2);
    if (pz) kcx = wrap_in(kcx, kkc - 2);
    int icy = icx + DX, jcy = jcx + DY, kcy = kcx + DZ;
    if (px) icy = wrap_in(icy, iic - 2);
    if (py) jcy = wrap_in(jcy, jjc - 2);
    if (pz) kcy = wrap_in(kcy, kkc - 2);
    const int X0 = 2 * (icx - 1), Y0 = 2 * (jcx - 1), Z0 = 2 * (kcx - 1);
    double sum = 0.0;
#pragma unroll 1
    for (int fz = -1; fz <= 1; ++fz)
#pragma unroll 1
        for (int fy = -1; fy <= 1; ++fy)
#pragma unroll 1
            for (int fx = -1; fx <= 1; ++fx) {
                int x = X0 + fx, y = Y0 + fy, z = Z0 + fz;
                if (x < 2 || x > iif - 1) {
                    if (!px) continue;
                    x = wrap_in(x, iif - 2);
                }
                if (y < 2 || y > jjf - 1) {
                    if (!py) continue;
                    y = wrap_in(y, jjf - 2);
                }
                if (z < 2 || z > kkf - 1) {
                    if (!pz) continue;
                    z = wrap_in(z, kkf - 2);
                }
                double inner = 0.0;
                bool any = false;
#pragma unroll 1
                for (int oz = -1; oz <= 1; ++oz)
#pragma unroll 1
                    for (int oy = -1; oy <= 1; ++oy)
#pragma unroll 1
                        for (int ox = -1; ox <= 1; ++ox) {
                            const int gx = fx + ox - 2 * DX, gy = fy + oy - 2 * DY, gz = fz + oz - 2 * DZ;
                            if (gx < -1 || gx > 1 || gy < -1 || gy > 1 || gz < -1 || gz > 1) continue;
                            const int xg = x + ox, yg = y + oy, zg = z + oz;
                            if ((!px && (xg < 2 || xg > iif - 1)) || (!py && (yg < 2 || yg > jjf - 1)) ||
                                (!pz && (zg < 2 || zg > kkf - 1)))
                                continue;
                            if (!NINE && (ox != 0) + (oy != 0) + (oz != 0) > 1) continue;
                            inner += aentry3<NINE>(so, P, ii, jj, x, y, z, ox, oy, oz) *
                                     pweight3(ci, Pc, iic, jjc, icy, jcy, kcy, gx, gy, gz);
                            any = true;
                        }
                if (any) sum += pweight3(ci, Pc, iic, jjc, icx, jcx, kcx, fx, fy, fz) * inner;
            }
    return sum;
}

template <bool NINE>
__global__ __launch_bounds__(256) void galerkin3_per_k(const double *__restrict__ so,
                                                       const double *__restrict__ ci,
                                                       double *__restrict__ soc, int iif, int jjf,
                                                       int kkf, int iic, int jjc, int kkc, int pm) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int kc = blockIdx.z + 2;
    if (ic > iic - 1 || jc > jjc - 1) return;
    const std::size_t Pc = (std::size_t)iic * jjc * kkc, pc = ix3(ic, jc, kc, iic, jjc);
    soc[pc] = kpair3w<NINE>(so, ci, iif, jjf, kkf, iic, jjc, kkc, ic, jc, kc, 0, 0, 0, pm);
#pragma unroll 1
    for (int t = 0; t < 13; ++t) {
        const int xa = ic + CP_A[t][0], ya = jc + CP_A[t][1], za = kc + CP_A[t][2];
        const int xb = ic + CP_B[t][0], yb = jc + CP_B[t][1], zb = kc + CP_B[t][2];
        double v = 0.0;
        if (((pm & 1) || (xa >= 2 && xb >= 2)) && ((pm & 2) || (ya >= 2 && yb >= 2)) &&
            ((pm & 4) || (za >= 2 && zb >= 2)))
            v = -kpair3w<NINE>(so, ci, iif, jjf, kkf, iic, jjc, kkc, xa, ya, za, xb - xa, yb - ya,
                               zb - za, pm);
        soc[(std::size_t)(CP_D[t] - 1) * Pc + pc] = v;
    }
}

// images of the coarse operator, all 14 planes, as the hand-expanded products end
// (src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90:1002-1113): per line of `axis`, the upper image
// from 2 first, then the lower image from n-1
__global__ __launch_bounds__(256) void soc_images_k(double *__restrict__ soc, int ii, int jj, int kk,
                                                    int axis) {
    const int na = axis == 0 ? jj : ii, nb = axis == 2 ? jj : kk;
    const int t = (int)(blockIdx.x * 256 + threadIdx.x);
    if (t >= na * nb) return;
    const int a = t % na + 1, b = t / na + 1;
    const std::size_t P = (std::size_t)ii * jj * kk;
    auto at = [&](int u, int d) -> double & {
        const std::size_t p = axis == 0 ? ix3(u, a, b, ii, jj) : axis == 1 ? ix3(a, u, b, ii, jj)
                                                                           : ix3(a, b, u, ii, jj);
        return soc[(std::size_t)(d - 1) * P + p];
    };
    const int n = axis == 0 ? ii : axis == 1 ? jj : kk;
    for (int d = 1; d <= 14; ++d) at(n, d) = at(2, d);
    for (int d = 1; d <= 14; ++d) at(1, d) = at(n - 1, d);
}

// coarsest grid with periodic directions: the dense matrix of the operator with its periodic
// couplings (row r's terms in the residual's order, a term whose neighbour wraps onto c added
// at (r, c), r <= c kept), then a Cholesky in DPOTF2's order (src/3d/ftn/
// BMG3_SymStd_SETUP_cg_LU.f90:190-612 writes the same matrix term by term and calls DPOTRF)
__constant__ signed char NB_SI[26] = {0, 0, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0};
__constant__ signed char NB_SJ[26] = {0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 0};
__constant__ signed char NB_SK[26] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
__constant__ signed char NB_D[26] = {kpw, kpnw, kps, kpsw, kpw, kpnw, kps, kpsw, kb, kbw, kbnw, kbn, kbne,
                                     kbe, kbse, kbs, kbsw, kb, kbe, kbse, kbs, kbsw, kbw, kbnw, kbn, kbne};
__constant__ signed char NB_OI[26] = {-1, -1, 0, 1, 1, 1, 0, -1, 0, -1, -1, 0, 1, 1, 1, 0, -1, 0, -1, -1, 0, 1, 1, 1, 0, -1};
__constant__ signed char NB_OJ[26] = {0, 1, 1, 1, 0, -1, -1, -1, 0, 0, 1, 1, 1, 0, -1, -1, -1, 0, 0, 1, 1, 1, 0, -1, -1, -1};
__constant__ signed char NB_OK[26] = {0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
__constant__ signed char NB7_T[6] = {0, 2, 4, 6, 8, 17};

__global__ __launch_bounds__(256) void setup_cg3_per_k(const double *__restrict__ so, int nstncl,
                                                       double *__restrict__ abd, int ii, int jj,
                                                       int kk, int pm, int *__restrict__ info) {
    const int nx = ii - 2, ny = jj - 2, nz = kk - 2, n = nx * ny * nz;
    const std::size_t P = (std::size_t)ii * jj * kk;
    for (std::size_t t = threadIdx.x; t < (std::size_t)n * n; t += blockDim.x) abd[t] = 0.0;
    __syncthreads();
    auto A = [&](int r, int c) -> double & { return abd[(std::size_t)c * n + r]; };  // 0-based
    const int nt = nstncl == 14 ? 26 : 6;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {  // row r is written by its thread only
        const int i = r % nx + 2, j = (r / nx) % ny + 2, k = r / (nx * ny) + 2;
        A(r, r) += S3(i, j, k, kp);
        for (int u = 0; u < nt; ++u) {
            const int t = nstncl == 14 ? u : NB7_T[u];
            const double a = -S3(i + NB_SI[t], j + NB_SJ[t], k + NB_SK[t], NB_D[t]);
            int ci_ = i + NB_OI[t], cj = j + NB_OJ[t], ck = k + NB_OK[t];
            if (ci_ < 2 || ci_ > ii - 1) {
                if (!(pm & 1)) continue;
                ci_ = ci_ < 2 ? ci_ + nx : ci_ - nx;
            }
            if (cj < 2 || cj > jj - 1) {
                if (!(pm & 2)) continue;
                cj = cj < 2 ? cj + ny : cj - ny;
            }
            if (ck < 2 || ck > kk - 1) {
                if (!(pm & 4)) continue;
                ck = ck < 2 ? ck + nz : ck - nz;
            }
            const int c = (ci_ - 2) + nx * ((cj - 2) + ny * (ck - 2));
            if (r <= c) A(r, c) += a;
        }
    }
    __syncthreads();
    __shared__ double piv;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (threadIdx.x == 0) {
            double ajj = A(j, j);
            for (int k = 0; k < j; ++k) ajj -= A(k, j) * A(k, j);
            if (!(ajj > 0.0)) {
                A(j, j) = ajj;
                bad = j + 1;
            } else {
                ajj = sqrt(ajj);
                A(j, j) = ajj;
            }
            piv = ajj;
        }
        __syncthreads();
        if (bad) break;
        for (int c = j + 1 + (int)threadIdx.x; c < n; c += (int)blockDim.x) {
            double s = A(j, c);
            for (int k = 0; k < j; ++k) s -= A(k, j) * A(k, c);
            A(j, c) = s / piv;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && bad) atomicMax(info, bad);
}

// DPOTRS on the dense factor, then the mean removal and the images in the order x, y, z
// (src/3d/ftn/BMG3_SymStd_SOLVE_cg.f90:103-196); one lane (the sums are sequential)
template <typename T>
__global__ void solve_cg3_per_k(T *__restrict__ q, const T *__restrict__ qf,
                                const double *__restrict__ abd, double *__restrict__ x, int ii,
                                int jj, int kk, int pm) {
    if (threadIdx.x != 0) return;
    const int nx = ii - 2, ny = jj - 2, n = nx * ny * (kk - 2);
    auto A = [&](int r, int c) { return abd[(std::size_t)c * n + r]; };
    for (int t = 0; t < n; ++t) x[t] = (double)Q3(qf, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2);
    for (int j = 0; j < n; ++j) {
        double s = x[j];
        for (int k = 0; k < j; ++k) s -= A(k, j) * x[k];
        x[j] = s / A(j, j);
    }
    for (int j = n - 1; j >= 0; --j) {
        double s = x[j];
        for (int k = j + 1; k < n; ++k) s -= A(j, k) * x[k];
        x[j] = s / A(j, j);
    }
    // the mean of the FP64 solution, removed before q takes it (rounded once in FP32)
    double cint = 0.0, qint = 0.0;
    for (int t = 0; t < n; ++t) {
        qint = qint + x[t];
        cint = cint + 1.0;
    }
    const double c = -qint / cint;
    for (int t = 0; t < n; ++t) Q3(q, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2) = (T)(x[t] + c);
    if (pm & 1)
        for (int k = 1; k <= kk; ++k)
            for (int j = 1; j <= jj; ++j) {
                Q3(q, 1, j, k) = Q3(q, ii - 1, j, k);
                Q3(q, ii, j, k) = Q3(q, 2, j, k);
            }
    if (pm & 2)
        for (int k = 1; k <= kk; ++k)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, 1, k) = Q3(q, i, jj - 1, k);
                Q3(q, i, jj, k) = Q3(q, i, 2, k);
            }
    if (pm & 4)
        for (int j = 1; j <= jj; ++j)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, j, 1) = Q3(q, i, j, kk - 1);
                Q3(q, i, j, kk) = Q3(q, i, j, 2);
            }
}

// ------------------------------------------------------------------------------------------
// plane relaxation data movement (include/cedar/3d/relax_planes.h:75-163,
// src/3d/relax_planes.cc:25-238); C++ indices here (ghost = 0), grid z = batch entry b,
// plane ipl = first + 2b.  Plane extents: XY (ii, jj), XZ (ii, kk), YZ (jj, kk).
// ------------------------------------------------------------------------------------------
#define X3C(a, i, j, k) (a)[(std::size_t)(i) + (std::size_t)ii * ((std::size_t)(j) + (std::size_t)jj * (std::size_t)(k))]
#define S3C(i, j, k, d) so3[(std::size_t)((d)-1) * P + (std::size_t)(i) + (std::size_t)ii * ((std::size_t)(j) + (std::size_t)jj * (std::size_t)(k))]

__device__ inline void plane_ext(int dir, int ii, int jj, int kk, int &n1, int &n2) {
    n1 = dir == YZ ? jj : ii;
    n2 = dir == XY ? jj : kk;
}

__global__ __launch_bounds__(256) void copy_coeff_k(int dir, const double *__restrict__ so3, int ns3,
                                                    double *__restrict__ so2, int first, int ii, int jj,
                                                    int kk) {
    int n1, n2;
    plane_ext(dir, ii, jj, kk, n1, n2);
    const int a = blockIdx.x * 64 + (threadIdx.x % 64);
    const int c = blockIdx.y * 4 + (threadIdx.x / 64);
    if (a >= n1 || c >= n2) return;
    const int ipl = first + 2 * (int)blockIdx.z;
    const std::size_t P = (std::size_t)ii * jj * kk, P2 = (std::size_t)n1 * n2;
    const bool nine = ns3 == 14;
    so2 += (std::size_t)blockIdx.z * (nine ? 5 : 3) * P2;
    const std::size_t p = (std::size_t)a + (std::size_t)n1 * c;
    if (dir == XY) {
        so2[p] = S3C(a, c, ipl, kp);
        so2[P2 + p] = S3C(a, c, ipl, kpw);
        so2[2 * P2 + p] = S3C(a, c, ipl, kps);
        if (nine) {
            so2[3 * P2 + p] = S3C(a, c, ipl, kpsw);
            so2[4 * P2 + p] = S3C(a, c, ipl, kpnw);
        }
    } else if (dir == XZ) {
        if (c < 1 || c > kk - 2) return;  // the reference copies interior k only
        so2[p] = S3C(a, ipl, c, kp);
        so2[P2 + p] = S3C(a, ipl, c, kpw);
        so2[2 * P2 + p] = S3C(a, ipl, c, kb);
        if (nine) {
            so2[3 * P2 + p] = S3C(a, ipl, c, kbw);
            so2[4 * P2 + p] = S3C(a, ipl, c, kbe);
        }
    } else {
        so2[p] = S3C(ipl, a, c, kp);
        so2[P2 + p] = S3C(ipl, a, c, kps);
        so2[2 * P2 + p] = S3C(ipl, a, c, kb);
        if (nine) {
            so2[3 * P2 + p] = S3C(ipl, a, c, kbs);
            so2[4 * P2 + p] = S3C(ipl, a, c, kbn);
        }
    }
}

template <typename T, bool TO2D>
__global__ __launch_bounds__(256) void copy_plane_k(int dir, T *__restrict__ x, T *__restrict__ x2, int first,
                                                    int ii, int jj, int kk) {
    int n1, n2;
    plane_ext(dir, ii, jj, kk, n1, n2);
    const int a = blockIdx.x * 64 + (threadIdx.x % 64);
    const int c = blockIdx.y * 4 + (threadIdx.x / 64);
    if (a >= n1 || c >= n2) return;
    const int ipl = first + 2 * (int)blockIdx.z;
    const std::size_t P2 = (std::size_t)n1 * n2;
    T *p2 = x2 + (std::size_t)blockIdx.z * P2 + (std::size_t)a + (std::size_t)n1 * c;
    T *p3 = dir == XY ? &X3C(x, a, c, ipl) : dir == XZ ? &X3C(x, a, ipl, c) : &X3C(x, ipl, a, c);
    if (TO2D)
        *p2 = *p3;
    else
        *p3 = *p2;
}

template <typename T, bool NINE>
__global__ __launch_bounds__(256) void copy_rhs_k(int dir, const T *__restrict__ so3, const T *__restrict__ x,
                                                  const T *__restrict__ b, T *__restrict__ b2, int first,
                                                  int ii, int jj, int kk) {
    int n1, n2;
    plane_ext(dir, ii, jj, kk, n1, n2);
    const int a = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int c = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    if (a > n1 - 2 || c > n2 - 2) return;
    const int ipl = first + 2 * (int)blockIdx.z;
    const std::size_t P = (std::size_t)ii * jj * kk, P2 = (std::size_t)n1 * n2;
#define X(i, j, k) X3C(x, i, j, k)
#define S(i, j, k, d) S3C(i, j, k, d)
    T v;
    if (dir == XY) {
        const int i = a, j = c, k = ipl;
        if (!NINE) {
            v = X3C(b, i, j, k) + S(i, j, k, kb) * X(i, j, k - 1) + S(i, j, k + 1, kb) * X(i, j, k + 1);
        } else {
            v = X3C(b, i, j, k) + S(i, j, k, kb) * X(i, j, k - 1) + S(i, j, k, kbw) * X(i - 1, j, k - 1) +
                S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) + S(i, j + 1, k, kbn) * X(i, j + 1, k - 1) +
                S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) + S(i + 1, j, k, kbe) * X(i + 1, j, k - 1) +
                S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) + S(i, j, k, kbs) * X(i, j - 1, k - 1) +
                S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) + S(i, j, k + 1, kbe) * X(i - 1, j, k + 1) +
                S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) + S(i, j + 1, k + 1, kbs) * X(i, j + 1, k + 1) +
                S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) + S(i + 1, j, k + 1, kbw) * X(i + 1, j, k + 1) +
                S(i, j, k + 1, kb) * X(i, j, k + 1) + S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1) +
                S(i, j, k + 1, kbn) * X(i, j - 1, k + 1) + S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1);
        }
    } else if (dir == XZ) {
        const int i = a, j = ipl, k = c;
        if (!NINE) {
            v = X3C(b, i, j, k) + S(i, j, k, kps) * X(i, j - 1, k) + S(i, j + 1, k, kps) * X(i, j + 1, k);
        } else {
            v = X3C(b, i, j, k) + S(i, j + 1, k, kpnw) * X(i - 1, j + 1, k) + S(i, j + 1, k, kps) * X(i, j + 1, k) +
                S(i + 1, j + 1, k, kpsw) * X(i + 1, j + 1, k) + S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) +
                S(i, j + 1, k, kbn) * X(i, j + 1, k - 1) + S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) +
                S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) + S(i, j + 1, k + 1, kbs) * X(i, j + 1, k + 1) +
                S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) + S(i, j, k, kpsw) * X(i - 1, j - 1, k) +
                S(i, j, k, kps) * X(i, j - 1, k) + S(i + 1, j, k, kpnw) * X(i + 1, j - 1, k) +
                S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) + S(i, j, k, kbs) * X(i, j - 1, k - 1) +
                S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) + S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1) +
                S(i, j, k + 1, kbn) * X(i, j - 1, k + 1) + S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1);
        }
    } else {
        const int i = ipl, j = a, k = c;
        if (!NINE) {
            v = X3C(b, i, j, k) + S(i, j, k, kpw) * X(i - 1, j, k) + S(i + 1, j, k, kpw) * X(i + 1, j, k);
        } else {
            v = X3C(b, i, j, k) + S(i, j + 1, k, kpnw) * X(i - 1, j + 1, k) + S(i, j, k, kpw) * X(i - 1, j, k) +
                S(i, j, k, kpsw) * X(i - 1, j - 1, k) + S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) +
                S(i, j, k, kbw) * X(i - 1, j, k - 1) + S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) +
                S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) + S(i, j, k + 1, kbe) * X(i - 1, j, k + 1) +
                S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1) + S(i + 1, j + 1, k, kpsw) * X(i + 1, j + 1, k) +
                S(i + 1, j, k, kpw) * X(i + 1, j, k) + S(i + 1, j, k, kpnw) * X(i + 1, j - 1, k) +
                S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) + S(i + 1, j, k, kbe) * X(i + 1, j, k - 1) +
                S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) + S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) +
                S(i + 1, j, k + 1, kbw) * X(i + 1, j, k + 1) + S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1);
        }
    }
#undef X
#undef S
    b2[(std::size_t)blockIdx.z * P2 + (std::size_t)a + (std::size_t)n1 * c] = v;
}

inline dim3 grid3(int nx, int ny, int nz) {
    return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)(nz > 0 ? nz : 1));
}

}  // namespace

// ==========================================================================================
// launchers
// ==========================================================================================
std::size_t norm_partials(int ii, int jj, int kk) {
    const dim3 g = grid3(ii - 2, jj - 2, kk - 2);
    return (std::size_t)g.x * g.y * g.z;
}

template <typename T>
void residual(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj, int kk,
              hipStream_t s) {
    const dim3 g = grid3(ii - 2, jj - 2, kk - 2);
    if (nstncl == 14)
        residual3_k<T, true, false><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, kk, nullptr);
    else
        residual3_k<T, false, false><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, kk, nullptr);
}

void residual_norm(const double *so, int nstncl, const double *qf, const double *q, double *res,
                   int ii, int jj, int kk, double *partials, double *sumsq, hipStream_t s) {
    const dim3 g = grid3(ii - 2, jj - 2, kk - 2);
    if (nstncl == 14)
        residual3_k<double, true, true><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, kk, partials);
    else
        residual3_k<double, false, true><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, kk, partials);
    k2::finalize_sum(partials, (std::size_t)g.x * g.y * g.z, sumsq, 1, s);
}

void residual_ir(const double *so, int nstncl, const double *qf, const double *x, const float *e,
                 double *xnew, float *r32, int ii, int jj, int kk, double *partials,
                 double *sumsq, hipStream_t s) {
    const dim3 g = grid3(ii - 2, jj - 2, kk - 2);
    if (nstncl == 14) {
        if (e)
            residual_ir3_k<true, true><<<g, 256, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, kk, partials);
        else
            residual_ir3_k<true, false><<<g, 256, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, kk, partials);
    } else {
        if (e)
            residual_ir3_k<false, true><<<g, 256, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, kk, partials);
        else
            residual_ir3_k<false, false><<<g, 256, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, kk, partials);
    }
    k2::finalize_sum(partials, (std::size_t)g.x * g.y * g.z, sumsq, 1, s);
}

void setup_recip(const double *so, int nstncl, double *sor, int ii, int jj, int kk,
                 hipStream_t s) {
    (void)nstncl;
    recip3_k<<<grid3(ii - 2, jj - 2, kk - 2), 256, 0, s>>>(so, sor, ii, jj, kk);
}

template <typename T>
void relax_gs(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj, int kk,
              int updown, hipStream_t s) {
    if (nstncl == 14) {
        // DOWN: colours pts = 8 .. 1, UP: 1 .. 8; colour pts starts at
        // (2 + mod(pts-1, 2), 2 + mod((pts-1)/2, 2), 2 + mod((pts-1)/4, 2))
        for (int t = 0; t < 8; ++t) {
            const int pts = updown == 0 ? 8 - t : 1 + t;
            const int pi = 2 + (pts - 1) % 2, pj = 2 + ((pts - 1) / 2) % 2, pk = 2 + ((pts - 1) / 4) % 2;
            const int ni = (ii - 1 - pi) / 2 + 1, nj = (jj - 1 - pj) / 2 + 1, nk = (kk - 1 - pk) / 2 + 1;
            if (ni <= 0 || nj <= 0 || nk <= 0) continue;
            gs27_colour_k<T><<<grid3(ni, nj, nk), 256, 0, s>>>(so, qf, q, sor, ii, jj, kk, pi, pj, pk,
                                                               t & 1);
        }
    } else {
        // DOWN: pts = 1 then 0; UP: 0 then 1
        for (int t = 0; t < 2; ++t) {
            const int pts = updown == 0 ? 1 - t : t;
            gs7_colour_k<T><<<grid3((ii - 1) / 2, jj - 2, kk - 2), 256, 0, s>>>(so, qf, q, sor, ii, jj, kk,
                                                                                pts, t & 1);
        }
    }
}

template <typename T>
void restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc,
               int kkc, hipStream_t s) {
    restrict3_k<T><<<grid3(iic - 2, jjc - 2, kkc - 2), 256, 0, s>>>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc);
}

template <typename T>
void interp_add(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                int jjc, int kkc, int iif, int jjf, int kkf, hipStream_t s) {
    (void)nstncl;
    interp3_k<T><<<grid3(iif - 1, jjf - 1, kkf - 1), 256, 0, s>>>(q, qc, res, so, ci, iic, jjc, kkc, iif, jjf, kkf);
}

void setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf, int iic,
                  int jjc, int kkc, hipStream_t s) {
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    const int mi = (iicf1 > iic - 1 ? iicf1 : iic - 1) - 1;
    const int mj = (jjcf1 > jjc - 1 ? jjcf1 : jjc - 1) - 1;
    const int mk = (kkcf1 > kkc - 1 ? kkcf1 : kkc - 1) - 1;
    const dim3 g = grid3(mi, mj, mk), gc = grid3(iicf1 - 2, jjcf1 - 2, kkcf1 - 2);
    if (nstncl == 14) {
        oi_edges_k<true><<<g, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
        oi_faces_k<true><<<g, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
        if (iicf1 >= 3 && jjcf1 >= 3 && kkcf1 >= 3)
            oi_cube_k<true><<<gc, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
    } else {
        oi_edges_k<false><<<g, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
        oi_faces_k<false><<<g, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
        if (iicf1 >= 3 && jjcf1 >= 3 && kkcf1 >= 3)
            oi_cube_k<false><<<gc, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc);
    }
}

void setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif, int jjf,
                    int kkf, int iic, int jjc, int kkc, hipStream_t s) {
    const dim3 g = grid3(iic - 2, jjc - 2, kkc - 2);
    if (nstncl == 14)
        galerkin3_k<true><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, kkf, iic, jjc, kkc);
    else
        galerkin3_k<false><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, kkf, iic, jjc, kkc);
}

int cg_bandwidth(int ii, int jj) { return (ii - 2) * (jj - 1) + 1; }

void setup_cg(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk,
              int *info, hipStream_t s) {
    setup_cg3_k<<<1, 256, 0, s>>>(so, nstncl, abd, ldab, ii, jj, kk, info);
}

template <typename T>
void solve_cg(T *q, const T *qf, const double *abd, int ldab, double *bbd, int ii, int jj, int kk,
              hipStream_t s) {
    solve_cg3_k<T><<<1, 256, 0, s>>>(q, qf, abd, ldab, bbd, ii, jj, kk);
}

// ---- periodic boundaries -------------------------------------------------------------------
int per_mask3(int jpn) {
    const int p = jpn < 0 ? -jpn : jpn;
    int m = 0;
    if (p == 2 || p == 3 || p == 6 || p == 8) m |= 1;
    if (p == 1 || p == 3 || p == 7 || p == 8) m |= 2;
    if (p == 5 || p == 6 || p == 7 || p == 8) m |= 4;
    return m;
}

int bc_from_mask3(int mask) {
    static const int map[8] = {0, 2, 1, 3, 5, 6, 7, 8};  // BMG_get_bc
    return mask >= 0 && mask < 8 ? map[mask] : -1;
}

namespace {
template <typename T>
void images_axis(T *q, int ii, int jj, int kk, int axis, int a0, int astep, int na, int b0,
                 int bstep, int nb, hipStream_t s) {
    if (na <= 0 || nb <= 0) return;
    images3_k<T><<<(unsigned)((na * nb + 255) / 256), 256, 0, s>>>(q, ii, jj, kk, axis, a0, astep, na,
                                                                  b0, bstep, nb);
}
}  // namespace

template <typename T>
void images(T *q, int ii, int jj, int kk, int jpn, int order_xyz, hipStream_t s) {
    const int m = per_mask3(jpn);
    const int seq[3] = {order_xyz ? 0 : 1, order_xyz ? 1 : 0, 2};
    for (int axis : seq) {
        if (!(m & (1 << axis))) continue;
        if (axis == 0) images_axis<T>(q, ii, jj, kk, 0, 1, 1, jj, 1, 1, kk, s);
        if (axis == 1) images_axis<T>(q, ii, jj, kk, 1, 1, 1, ii, 1, 1, kk, s);
        if (axis == 2) images_axis<T>(q, ii, jj, kk, 2, 1, 1, ii, 1, 1, jj, s);
    }
}

// src/3d/ftn/BMG3_SymStd_relax_GS.f90:189-352: after each colour, the x images of the rows and
// the y images of the planes that the colour visited; the z images after all colours
template <typename T>
void relax_gs_per(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                  int kk, int updown, int jpn, hipStream_t s) {
    const int m = per_mask3(jpn);
    if (nstncl == 14) {
        for (int t = 0; t < 8; ++t) {
            const int pts = updown == 0 ? 8 - t : 1 + t;
            const int pi = 2 + (pts - 1) % 2, pj = 2 + ((pts - 1) / 2) % 2, pk = 2 + ((pts - 1) / 4) % 2;
            auto count = [](int first, int last) { return last >= first ? (last - first) / 2 + 1 : 0; };
            const int ni = count(pi, ii - 1), nj = count(pj, jj - 1), nk = count(pk, kk - 1);
            if (ni > 0 && nj > 0 && nk > 0)
                gs27_colour_k<T><<<grid3(ni, nj, nk), 256, 0, s>>>(so, qf, q, sor, ii, jj, kk, pi, pj,
                                                                   pk, t & 1);
            if (m & 1) images_axis<T>(q, ii, jj, kk, 0, pj, 2, nj, pk, 2, nk, s);
            if (m & 2) images_axis<T>(q, ii, jj, kk, 1, 1, 1, ii, pk, 2, nk, s);
        }
    } else {
        for (int t = 0; t < 2; ++t) {
            const int pts = updown == 0 ? 1 - t : t;
            gs7_colour_k<T><<<grid3((ii - 1) / 2, jj - 2, kk - 2), 256, 0, s>>>(so, qf, q, sor, ii, jj, kk,
                                                                                pts, t & 1);
            if (m & 1) images_axis<T>(q, ii, jj, kk, 0, 2, 1, jj - 2, 2, 1, kk - 2, s);
            if (m & 2) images_axis<T>(q, ii, jj, kk, 1, 1, 1, ii, 2, 1, kk - 2, s);
        }
    }
    if (m & 4) images_axis<T>(q, ii, jj, kk, 2, 1, 1, ii, 1, 1, jj, s);
}

// src/3d/ftn/BMG3_SymStd_restrict.f90:79-141: the images of q first (y, x, z), in place
template <typename T>
void restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc, int kkc,
                  int jpn, hipStream_t s) {
    images<T>(q, iif, jjf, kkf, jpn, 0, s);
    restrict3_k<T><<<grid3(iic - 2, jjc - 2, kkc - 2), 256, 0, s>>>(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc);
}

// interp_add, then every image of the updated q (the reference's own image loops read an
// unassigned IPN and left-over loop indices, src/3d/ftn/BMG3_SymStd_interp_add.f90:237-280)
template <typename T>
void interp_add_per(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci,
                    int iic, int jjc, int kkc, int iif, int jjf, int kkf, int jpn, hipStream_t s) {
    interp_add<T>(q, qc, res, so, nstncl, ci, iic, jjc, kkc, iif, jjf, kkf, s);
    images<T>(q, iif, jjf, kkf, jpn, 0, s);
}

// src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:809-2784, loop nest by loop nest (see oi_nest_k);
// the 27- and 7-point branches differ in three statements (:1057-1065 / :2111-2119,
// :1274 / :2269, :1391 / :2359), followed as written
void setup_interp_per(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf, int iic,
                      int jjc, int kkc, int jpn, hipStream_t s) {
    const bool nine = nstncl == 14;
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
    const int iif1 = iif - 1, jjf1 = jjf - 1, kkf1 = kkf - 1;
    const int iicf = (iif - 2) / 2 + 3, jjcf = (jjf - 2) / 2 + 3, kkcf = (kkf - 2) / 2 + 3;
    const int iifc = 2 * (iicf - 2) + 2, jjfc = 2 * (jjcf - 2) + 2, kkfc = 2 * (kkcf - 2) + 2;
    const int m = per_mask3(jpn);
    const bool px = m & 1, py = m & 2, pz = m & 4;
    int ibeg = 2, ibegc = 3, ibeg_x = 2, iend_x = iif - 2, iendc = iicf - 1;
    int jbeg = 2, jbegc = 3, jbeg_y = 2, jend_y = jjf - 2, jendc = jjcf - 1;
    int kbeg = 2, kbegc = 3, kbeg_z = 2, kend_z = kkf - 2, kendc = kkcf - 1;
    if (px) {
        ibeg = 0;
        ibegc = 2;
        if (iic == iicf) { iendc = iicf; ibeg_x = 3; iend_x = iif1; }
    }
    if (py) {
        jbeg = 0;
        jbegc = 2;
        if (jjc == jjcf) { jendc = jjcf; jbeg_y = 3; jend_y = jjf1; }
    }
    if (pz) {
        kbeg = 0;
        kbegc = 2;
        if (kkc == kkcf) { kendc = kkcf; kbeg_z = 3; kend_z = kkf1; }
    }
    auto REG = [](int c0, int c1) { return oi_axis{0, c0, c1, 0, 1}; };
    auto WRP = [](int c0, int c1, int f0, int mod) { return oi_axis{1, c0, c1, f0, mod}; };
    auto FIX = [](int c, int f) { return oi_axis{2, c, c, f, 1}; };
    const oi_axis WX = WRP(ibegc, iendc, ibeg, iifc), WY = WRP(jbegc, jendc, jbeg, jjfc);
    const oi_axis WZ = WRP(kbegc, kendc, kbeg, kkfc);
    auto nest = [&](int type, oi_axis ax, oi_axis ay, oi_axis az) {
        auto cnt = [](const oi_axis &a) { return a.mode == 2 ? 1 : a.c1 - a.c0 + 1; };
        const int nx = cnt(ax), ny = cnt(ay), nz = cnt(az);
        if (nx <= 0 || ny <= 0 || nz <= 0) return;
        const unsigned g = (unsigned)(((long long)nx * ny * nz + 255) / 256);
#define OI_NEST(N, TY)                                                                              \
    oi_nest_k<N, TY><<<g, 256, 0, s>>>(so, ci, iif, jjf, kkf, iic, jjc, kkc, ax, ay, az, nx, ny, nz)
#define OI_NEST2(TY)          \
    if (nine)                 \
        OI_NEST(true, TY);    \
    else                      \
        OI_NEST(false, TY);
        switch (type) {
        case 0: OI_NEST2(0) break;
        case 1: OI_NEST2(1) break;
        case 2: OI_NEST2(2) break;
        case 3: OI_NEST2(3) break;
        case 4: OI_NEST2(4) break;
        case 5: OI_NEST2(5) break;
        default: OI_NEST2(6)
        }
#undef OI_NEST2
#undef OI_NEST
    };
    constexpr int V = -1;  // the loop variable
    struct stmt {
        int d[4], r[4];
    };
    auto copy = [&](int lo, int hi, std::initializer_list<stmt> st) {
        if (hi < lo) return;
        ci_copy cp{};
        cp.n = 0;
        cp.lo = lo;
        cp.hi = hi;
        for (const stmt &x : st) {
            for (int a = 0; a < 4; ++a) {
                cp.dst[cp.n][a] = x.d[a];
                cp.src[cp.n][a] = x.r[a];
            }
            ++cp.n;
        }
        ci_copy_k<<<(unsigned)((hi - lo + 1 + 255) / 256), 256, 0, s>>>(ci, iic, jjc, kkc, cp);
    };

    // fine points on coarse k-planes and fine-only x-lines (lxyl, lxyr): :930-1106
    nest(0, WX, REG(2, jjc1), REG(2, kkc1));
    if (py) {
        nest(0, WX, FIX(jjc, jbeg_y), REG(2, kkc1));
        nest(0, WX, FIX(1, jend_y), REG(2, kkc1));
    }
    if (pz) {
        nest(0, WX, REG(2, jjc1), FIX(kkc, kbeg_z));
        if (py) {
            if (nine) {  // jc is left at jjc by the loop before; one pass of the repeated copies
                if (ibegc <= iendc)
                    copy(0, 0, {{{1, jjc, kkc, lxyl}, {iic1, jjc, kkc, lxyl}},
                                {{1, jjc, kkc, lxyr}, {iic1, jjc, kkc, lxyr}},
                                {{iic, jjc, kkc, lxyl}, {2, jjc, kkc, lxyl}},
                                {{iic, jjc, kkc, lxyr}, {2, jjc, kkc, lxyr}}});
            } else {
                copy(ibegc, iendc, {{{V, 1, kkc, lxyl}, {iic, jjc1, kkc, lxyl}},
                                    {{V, 1, kkc, lxyr}, {iic, jjc1, kkc, lxyr}},
                                    {{V, jjc, kkc, lxyl}, {V, 2, kkc, lxyl}},
                                    {{V, jjc, kkc, lxyr}, {V, 2, kkc, lxyr}}});
            }
        }
        nest(0, WX, REG(2, jjc1), FIX(1, kend_z));
        if (py)
            copy(ibegc, iendc, {{{V, 1, 1, lxyl}, {V, jjc1, 1, lxyl}},
                                {{V, 1, 1, lxyr}, {V, jjc1, 1, lxyr}},
                                {{V, jjc, 1, lxyl}, {V, 2, 1, lxyl}},
                                {{V, jjc, 1, lxyr}, {V, 2, 1, lxyr}}});
    }

    // fine points on coarse k-planes and fine-only y-lines (lxya, lxyb): :1109-1277
    nest(1, REG(2, iic1), WY, REG(2, kkc1));
    if (px) {
        nest(1, FIX(iic, ibeg_x), WY, REG(2, kkc1));
        nest(1, FIX(1, iend_x), WY, REG(2, kkc1));
    }
    if (pz) {
        nest(1, REG(2, iic1), WY, FIX(kkc, kbeg_z));
        if (px)
            copy(jbegc, jendc, {{{1, V, kkc, lxya}, {iic1, V, kkc, lxya}},
                                {{1, V, kkc, lxyb}, {iic1, V, kkc, lxyb}},
                                {{iic, V, kkc, lxya}, {2, V, kkc, lxya}},
                                {{iic, V, kkc, lxyb}, {2, V, kkc, lxyb}}});
        nest(1, REG(2, iic1), WY, FIX(1, kend_z));
        if (px)
            copy(jbegc, jendc, {{{1, V, 1, lxya}, {iic1, V, 1, lxya}},
                                {{1, V, 1, lxyb}, {iic1, V, 1, lxyb}},
                                {{iic, V, 1, lxya}, {2, V, 1, lxya}},
                                {{iic, V, 1, lxyb}, {nine ? 1 : 2, V, 1, lxyb}}});
    }

    // fine points on coarse j-planes and fine-only z-lines (lxza, lxzb): :1279-1442
    nest(2, REG(2, iic1), REG(2, jjc1), WZ);
    if (px) {
        for (int side = 0; side < 2; ++side) {
            const int icx = side == 0 ? iic : 1;
            nest(2, FIX(icx, side == 0 ? ibeg_x : iend_x), REG(2, jjc1), WZ);
            if (py)
                copy(kbegc, kendc, {{{icx, 1, V, lxza}, {icx, jjc1, V, lxza}},
                                    {{icx, 1, V, lxzb}, {icx, jjc1, V, lxzb}},
                                    {{icx, jjc, V, lxza}, {icx, 2, V, lxza}},
                                    {{icx, jjc, V, lxzb}, {icx, 2, V, lxzb}}});
        }
    }
    if (py) {
        nest(2, REG(2, nine ? iic1 : jjc1), FIX(jjc, jbeg_y), WZ);
        nest(2, REG(2, iic1), FIX(1, jend_y), WZ);
    }

    // fine points on coarse k-planes, fine-only x- and y-lines: :1444-1583
    nest(3, WX, WY, REG(2, kkc1));
    if (pz) {
        nest(3, WX, WY, FIX(kkc, kbeg_z));
        nest(3, WX, WY, FIX(1, kend_z));
    }
    // fine points on coarse j-planes, fine-only x- and z-lines: :1587-1720
    nest(4, WX, REG(2, jjc1), WZ);
    if (py) {
        nest(4, WX, FIX(jjc, jbeg_y), WZ);
        nest(4, WX, FIX(1, jend_y), WZ);
    }
    // fine points on coarse i-planes, fine-only y- and z-lines, J advanced with the z
    // modulus: :1723-1859
    const oi_axis WYZ = WRP(jbegc, jendc, jbeg, kkfc);
    nest(5, REG(2, iic1), WYZ, WZ);
    if (px) {
        nest(5, FIX(iic, ibeg_x), WYZ, WZ);
        nest(5, FIX(1, iend_x), WYZ, WZ);
    }
    // cell centres: :1862-2013
    nest(6, WX, WY, WZ);
}

// Galerkin product over the wrapped geometry, then the coarse images (y, x, z) as
// src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90:1002-1113 copies them
void setup_galerkin_per(const double *so, int nstncl, const double *ci, double *soc, int iif,
                        int jjf, int kkf, int iic, int jjc, int kkc, int jpn, hipStream_t s) {
    const int m = per_mask3(jpn);
    const dim3 g = grid3(iic - 2, jjc - 2, kkc - 2);
    if (nstncl == 14)
        galerkin3_per_k<true><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, kkf, iic, jjc, kkc, m);
    else
        galerkin3_per_k<false><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, kkf, iic, jjc, kkc, m);
    for (int axis : {1, 0, 2}) {
        if (!(m & (1 << axis))) continue;
        const int na = axis == 0 ? jjc : iic, nb = axis == 2 ? jjc : kkc;
        soc_images_k<<<(unsigned)((na * nb + 255) / 256), 256, 0, s>>>(soc, iic, jjc, kkc, axis);
    }
}

void setup_cg_per(const double *so, int nstncl, double *abd, int ii, int jj, int kk, int jpn,
                  int *info, hipStream_t s) {
    setup_cg3_per_k<<<1, 256, 0, s>>>(so, nstncl, abd, ii, jj, kk, per_mask3(jpn), info);
}

template <typename T>
void solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int kk,
                  int jpn, hipStream_t s) {
    solve_cg3_per_k<T><<<1, 64, 0, s>>>(q, qf, abd, bbd, ii, jj, kk, per_mask3(jpn));
}

static dim3 plane_grid(int dir, int ii, int jj, int kk, int npl, bool interior) {
    const int n1 = dir == YZ ? jj : ii, n2 = dir == XY ? jj : kk;
    dim3 g = interior ? grid3(n1 - 2, n2 - 2, 1) : grid3(n1, n2, 1);
    g.z = (unsigned)npl;
    return g;
}

void copy_coeff(int dir, const double *so3, int nstncl3, double *so2, int first, int npl, int ii,
                int jj, int kk, hipStream_t s) {
    copy_coeff_k<<<plane_grid(dir, ii, jj, kk, npl, false), 256, 0, s>>>(dir, so3, nstncl3, so2, first, ii, jj, kk);
}

template <typename T>
void copy32(int dir, const T *x, T *x2, int first, int npl, int ii, int jj, int kk, hipStream_t s) {
    copy_plane_k<T, true><<<plane_grid(dir, ii, jj, kk, npl, false), 256, 0, s>>>(dir, const_cast<T *>(x), x2, first,
                                                                                ii, jj, kk);
}

template <typename T>
void copy23(int dir, const T *x2, T *x, int first, int npl, int ii, int jj, int kk, hipStream_t s) {
    copy_plane_k<T, false><<<plane_grid(dir, ii, jj, kk, npl, false), 256, 0, s>>>(dir, x, const_cast<T *>(x2),
                                                                                 first, ii, jj, kk);
}

template <typename T>
void copy_rhs(int dir, const T *so3, int nstncl3, const T *x, const T *b, T *b2, int first,
              int npl, int ii, int jj, int kk, hipStream_t s) {
    const dim3 g = plane_grid(dir, ii, jj, kk, npl, true);
    if (nstncl3 == 14)
        copy_rhs_k<T, true><<<g, 256, 0, s>>>(dir, so3, x, b, b2, first, ii, jj, kk);
    else
        copy_rhs_k<T, false><<<g, 256, 0, s>>>(dir, so3, x, b, b2, first, ii, jj, kk);
}

#define CDR3_INST(T)                                                                                \
    template void residual<T>(const T *, int, const T *, const T *, T *, int, int, int, hipStream_t); \
    template void relax_gs<T>(const T *, int, const T *, T *, const T *, int, int, int, int,        \
                              hipStream_t);                                                         \
    template void restrict_<T>(const T *, const T *, T *, int, int, int, int, int, int, hipStream_t); \
    template void interp_add<T>(T *, const T *, const T *, const T *, int, const T *, int, int, int, \
                                int, int, int, hipStream_t);                                        \
    template void solve_cg<T>(T *, const T *, const double *, int, double *, int, int, int,         \
                              hipStream_t);                                                         \
    template void copy32<T>(int, const T *, T *, int, int, int, int, int, hipStream_t);             \
    template void copy23<T>(int, const T *, T *, int, int, int, int, int, hipStream_t);             \
    template void copy_rhs<T>(int, const T *, int, const T *, const T *, T *, int, int, int, int, int, \
                              hipStream_t);                                                         \
    template void images<T>(T *, int, int, int, int, int, hipStream_t);                             \
    template void relax_gs_per<T>(const T *, int, const T *, T *, const T *, int, int, int, int, int, \
                                  hipStream_t);                                                     \
    template void restrict_per<T>(const T *, T *, T *, int, int, int, int, int, int, int, hipStream_t); \
    template void interp_add_per<T>(T *, const T *, const T *, const T *, int, const T *, int, int,  \
                                    int, int, int, int, int, hipStream_t);                           \
    template void solve_cg_per<T>(T *, const T *, const double *, double *, int, int, int, int,      \
                                  hipStream_t);
CDR3_INST(double)
CDR3_INST(float)
#undef CDR3_INST

}  // namespace k3
}  // namespace cdr

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
// device function below is one loop nest of the reference (27-point branch when NINE, 7-point
// branch otherwise) for the fine point tied to coarse (ic, jc, kc); i, j, k = 2 (ic-1), ...
// ------------------------------------------------------------------------------------------
#define OI_ARGS const double *__restrict__ so, double *__restrict__ ci, int ii, int jj, std::size_t P, \
                int iic, int jjc, std::size_t Pc, int ic, int jc, int kc
#define OI_PASS so, ci, ii, jj, P, iic, jjc, Pc, ic, jc, kc
#define OI_IJK const int i = 2 * (ic - 1), j = 2 * (jc - 1), k = 2 * (kc - 1); (void)i; (void)j; (void)k;

template <bool NINE>
__device__ inline void oi_xedge(OI_ARGS) {
    constexpr bool nine = NINE;
    OI_IJK
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
    OI_IJK
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
    OI_IJK
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
    OI_IJK
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
    OI_IJK
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
    OI_IJK
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
    OI_IJK
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

template <typename T>
__global__ void solve_cg3_k(T *__restrict__ q, const T *__restrict__ qf, const double *__restrict__ abd,
                            int ld, double *__restrict__ x, int ii, int jj, int kk) {
    if (threadIdx.x != 0) return;
    const int nx = ii - 2, ny = jj - 2, n = nx * ny * (kk - 2);
    const int kd = (ii - 2) * (jj - 1) + 1;
    for (int t = 0; t < n; ++t)
        x[t] = (double)Q3(qf, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2);
    band_solve_upper(abd, n, kd, ld, x);
    for (int t = 0; t < n; ++t) Q3(q, t % nx + 2, (t / nx) % ny + 2, t / (nx * ny) + 2) = (T)x[t];
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
    solve_cg3_k<T><<<1, 64, 0, s>>>(q, qf, abd, ldab, bbd, ii, jj, kk);
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
                              hipStream_t);
CDR3_INST(double)
CDR3_INST(float)
#undef CDR3_INST

}  // namespace k3
}  // namespace cdr

// 2D BoxMG kernels for CDNA4 (gfx950).
//
// Layout: Fortran order with one ghost layer; array index (i, j) with 0 <= i < ii, 0 <= j < jj,
// interior 1..ii-2 x 1..jj-2, point p = j*ii + i.  Stencil plane k of a level lives at
// so[k*ii*jj + p].  The off-diagonal couplings are stored negated; so[KNW] at (a, b) couples
// (a-1, b) with (a, b-1) and so[KSW] at (a, b) couples (a, b) with (a-1, b-1).
//
// Numerics: the FP64 kernels for residual, Gauss-Seidel, restriction, interpolation and
// operator-induced interpolation evaluate every point with the same operation order as Cedar's
// serial kernels and are compiled with -ffp-contract=off, so they reproduce them bit for bit.
// Line relaxation replaces the sequential LDL^T sweeps by a wavefront-wide affine scan and the
// Galerkin product is formed generically as P^T A P, so those two round differently (ulp level).
//
// All kernels are stencils on HBM-resident arrays (~0.2-0.3 flop/byte): no MFMA, coalesced
// accesses along i, 64-wide wavefronts, deterministic fixed-order reductions.
#include "../../include/cedar_hip.h"
#include "band_chol.hpp"
#include "bmg2_launch.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <vector>
#include <string>
#include <utility>

namespace cdr {
namespace k2 {

namespace {

constexpr int WAVE = 64;

__host__ __device__ inline std::size_t pidx(int i, int j, int ii) {
    return (std::size_t)j * (std::size_t)ii + (std::size_t)i;
}

inline dim3 grid2(int nx, int ny, int bx, int by) {
    return dim3((unsigned)((nx + bx - 1) / bx), (unsigned)((ny + by - 1) / by), 1);
}

// ------------------------------------------------------------------------------------------
// deterministic block reduction of one double per thread (256 threads = 4 waves)
// ------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// barrier that orders LDS only: prefetched global loads stay in flight across it
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS_ONLY: the barrier orders LDS alone (global loads issued before it stay in flight)
template <int NWAVES = 4, bool LDS_ONLY = false>
__device__ inline void block_sum_store(double v, double *dst) {
    __shared__ double part[NWAVES];
    v = wave_sum(v);
    int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    if (lane == 0) part[wid] = v;
    if (LDS_ONLY)
        lds_barrier();
    else
        __syncthreads();
    if (threadIdx.x == 0) {
        double t = part[0];
        for (int w = 1; w < NWAVES; ++w) t += part[w];
        *dst = t;
    }
}

// Batched problems: every kernel below takes the problem index from the last grid dimension
// (blockIdx.z for tiled kernels, blockIdx.y for line kernels, blockIdx.x for one-workgroup
// kernels); the problems of a batch are stored back to back, each with the layout above.
//
// Problems that share an operator come in groups of rep right-hand sides (problem b uses
// operator b / rep).  Such launches renumber their workgroups so that the rep workgroups of one
// tile of one operator run one after another on one XCD (workgroups are dealt round-robin over
// the 8 XCDs in launch order): the operator's tile comes from HBM once and from that XCD's L2
// for the other right-hand sides.  t (the workgroup's tile, tiles per problem) and b (its
// problem) are replaced by the logical pair; the map is a permutation of the launch, so every
// tile of every problem is still done once.  Launches whose tiles per problem are not a
// multiple of 8 keep the launch order.  Measured on 29 operators of 2049^2 with 4 right-hand
// sides each (mixed V(1,1) line-xy solve): 688 against 720 ms.
__device__ inline void shared_op_order(int tiles, int rep, int &t, int &b) {
    if (rep <= 1 || tiles % 8 != 0) return;
    const int op = b / rep;
    const int off = (b - op * rep) * tiles + t;  // position in the operator's range, which
    const int k = off / 8;                       // starts on an XCD-0 workgroup
    t = (k / rep) * 8 + off % 8;
    b = op * rep + k % rep;
}

// logical tile (x, y) and problem z of a tiled launch (problem index in blockIdx.z)
struct wg_pos {
    int x, y;
    std::size_t z;
};
__device__ inline wg_pos tiled_wg(int rep) {
    if (rep <= 1) return {(int)blockIdx.x, (int)blockIdx.y, (std::size_t)blockIdx.z};
    int t = (int)(blockIdx.y * gridDim.x + blockIdx.x), b = (int)blockIdx.z;
    shared_op_order((int)(gridDim.x * gridDim.y), rep, t, b);
    return {t % (int)gridDim.x, t / (int)gridDim.x, (std::size_t)b};
}

// one block of 1024 threads: sum of n partials in a fixed order -> *out.  Thread t adds
// partials t, t+1024, t+2048, ... in that order; loads are batched so they overlap.
__global__ __launch_bounds__(1024) void finalize_sum_k(const double *__restrict__ partials,
                                                       std::size_t n, double *__restrict__ out) {
    constexpr int T = 1024, U = 8;
    partials += (std::size_t)blockIdx.x * n;  // one workgroup per problem of a batch
    out += blockIdx.x;
    double v = 0.0;
    std::size_t k = threadIdx.x;
    for (; k + (U - 1) * T < n; k += U * T) {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = partials[k + (std::size_t)u * T];
#pragma unroll
        for (int u = 0; u < U; ++u) v += t[u];
    }
    for (; k < n; k += T) v += partials[k];
    block_sum_store<16>(v, out);
}

// ------------------------------------------------------------------------------------------
// residual  r = f + sum(couplings * q) - c q   (src/2d/ftn/BMG2_SymStd_residual.f90:85-119)
// block: 64 (i) x 4 (j)
// ------------------------------------------------------------------------------------------
template <typename T, bool NINE>
__device__ inline T apply_offdiag(const T *__restrict__ so, std::size_t P, const T *__restrict__ q,
                                  std::size_t p, int ii, T acc) {
    acc = acc + so[P * KW + p] * q[p - 1];
    acc = acc + so[P * KW + p + 1] * q[p + 1];
    acc = acc + so[P * KS + p] * q[p - ii];
    acc = acc + so[P * KS + p + ii] * q[p + ii];
    if (NINE) {
        acc = acc + so[P * KSW + p] * q[p - ii - 1];
        acc = acc + so[P * KNW + p + 1] * q[p - ii + 1];
        acc = acc + so[P * KNW + p + ii] * q[p + ii - 1];
        acc = acc + so[P * KSW + p + ii + 1] * q[p + ii + 1];
    }
    return acc;
}

template <typename T, bool NINE, bool NORM>
__global__ __launch_bounds__(256) void residual_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                  const T *__restrict__ q, T *__restrict__ res,
                                                  int ii, int jj, double *__restrict__ partials,
                                                  int rep) {
    const wg_pos g = tiled_wg(rep);
    const int i = g.x * 64 + (threadIdx.x % 64) + 1;
    const int j = g.y * 4 + (threadIdx.x / 64) + 1;
    const std::size_t P = (std::size_t)ii * jj, bz = g.z;
    so += (bz / rep) * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    res += bz * P;
    if (NORM) partials += bz * gridDim.x * gridDim.y;
    double sq = 0.0;
    if (i <= ii - 2 && j <= jj - 2) {
        const std::size_t p = pidx(i, j, ii);
        T r = apply_offdiag<T, NINE>(so, P, q, p, ii, qf[p]);
        r = r - so[p] * q[p];
        res[p] = r;
        if (NORM) sq = (double)r * (double)r;
    }
    if (NORM) block_sum_store(sq, &partials[g.y * gridDim.x + g.x]);
}

template <typename T>
__global__ __launch_bounds__(256) void sumsq_k(const T *__restrict__ x, int ii, int jj,
                                               double *__restrict__ partials) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    const std::size_t bz = blockIdx.z;
    x += bz * (std::size_t)ii * jj;
    partials += bz * gridDim.x * gridDim.y;
    double sq = 0.0;
    if (i <= ii - 2 && j <= jj - 2) {
        double v = (double)x[pidx(i, j, ii)];
        sq = v * v;
    }
    block_sum_store(sq, &partials[blockIdx.y * gridDim.x + blockIdx.x]);
}

// mixed precision outer step: xnew = x + e, r = f - A xnew (FP64), r32 = (float) r; blocks of
// 64 (i) x IR_ROWS (j) points, 64 * IR_ROWS threads.  One workgroup per tile of one operator
// (blockIdx.z) does the rep right-hand sides of that operator (problems z * rep .. z * rep +
// rep - 1) one after another: the operator's couplings of a point are loaded once into
// registers for all of them (with one workgroup per problem, each re-read the 40 bytes per point
// of the operator).  f is loaded with the staged iterate, and the barriers order LDS only, so
// that stores of one right-hand side do not hold up the next.  Measured on the mixed solve of 29
// operators of 2049^2 with 4 right-hand sides each: 658 ms, against 686 ms with one workgroup
// per problem and 702 ms with the next right-hand side's loads software-pipelined.
constexpr int IR_ROWS = 8;
template <bool NINE, bool HAS_E>
__global__ __launch_bounds__(64 * IR_ROWS) void residual_ir_k(const double *__restrict__ so,
                                                     const double *__restrict__ qf,
                                                     const double *__restrict__ x,
                                                     const float *__restrict__ e,
                                                     double *__restrict__ xnew,
                                                     float *__restrict__ r32, int ii, int jj,
                                                     double *__restrict__ partials,
                                                     const int *__restrict__ active, int rep,
                                                     float *__restrict__ r32T) {
    // the updated iterate of the block and its one-point halo, formed once per point
    __shared__ double xs[IR_ROWS + 2][66];
    __shared__ float rt[IR_ROWS][65];  // r32 of the block, for the transposed copy
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * IR_ROWS;  // halo corner of the block
    const int i = i0 + tx + 1, j = j0 + ty + 1;
    const std::size_t P = (std::size_t)ii * jj;
    const bool in = i <= ii - 2 && j <= jj - 2;
    const std::size_t p = in ? pidx(i, j, ii) : 0;
    so += (std::size_t)blockIdx.z * (NINE ? 5 : 3) * P;
    double cw0 = 0.0, cw1 = 0.0, cs0 = 0.0, cs1 = 0.0, co = 0.0;
    double csw0 = 0.0, cnw0 = 0.0, cnw1 = 0.0, csw1 = 0.0;
    if (in) {
        cw0 = so[P * KW + p];
        cw1 = so[P * KW + p + 1];
        cs0 = so[P * KS + p];
        cs1 = so[P * KS + p + ii];
        if (NINE) {
            csw0 = so[P * KSW + p];
            cnw0 = so[P * KNW + p + 1];
            cnw1 = so[P * KNW + p + ii];
            csw1 = so[P * KSW + p + ii + 1];
        }
        co = so[p];
    }
    auto X = [&](int di, int dj) { return xs[ty + 1 + dj][tx + 1 + di]; };
    for (int rh = 0; rh < rep; ++rh) {
        // xs, rt and the block sum's slots of the previous right-hand side are free: every
        // thread has passed the barrier of its block_sum_store (and, with r32T, the one before
        // its transposed copy), and thread 0 the staging barrier below
        const std::size_t bz = (std::size_t)blockIdx.z * rep + rh, off = bz * P;
        // a problem of a batch that has already converged keeps its iterate (active[b] == 0)
        const bool upd = HAS_E && (active == nullptr || active[bz] != 0);
        const double f = in ? qf[off + p] : 0.0;
        for (int t = threadIdx.x; t < (IR_ROWS + 2) * 66; t += 64 * IR_ROWS) {
            const int ti = i0 + t % 66, tj = j0 + t / 66;
            double v = 0.0;
            if (ti < ii && tj < jj) {
                const std::size_t k = off + pidx(ti, tj, ii);
                v = upd ? x[k] + (double)e[k] : x[k];
            }
            xs[t / 66][t % 66] = v;
        }
        lds_barrier();
        double sq = 0.0;
        if (in) {
            double r = f;
            r = r + cw0 * X(-1, 0);
            r = r + cw1 * X(1, 0);
            r = r + cs0 * X(0, -1);
            r = r + cs1 * X(0, 1);
            if (NINE) {
                r = r + csw0 * X(-1, -1);
                r = r + cnw0 * X(1, -1);
                r = r + cnw1 * X(-1, 1);
                r = r + csw1 * X(1, 1);
            }
            const double xc = X(0, 0);
            r = r - co * xc;
            xnew[off + p] = xc;
            r32[off + p] = (float)r;
            rt[ty][tx] = (float)r;
            sq = r * r;
        }
        block_sum_store<IR_ROWS, true>(
            sq, &partials[bz * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x]);
        if (r32T) {  // uniform over the launch
            lds_barrier();
            // consecutive threads write consecutive j of one row of the transposed grid
            const int a = threadIdx.x / IR_ROWS, c = threadIdx.x % IR_ROWS;
            const int i2 = i0 + a + 1, j2 = j0 + c + 1;
            if (i2 <= ii - 2 && j2 <= jj - 2) r32T[off + (std::size_t)i2 * jj + j2] = rt[c][a];
        }
    }
}

// ------------------------------------------------------------------------------------------
// point Gauss-Seidel, one launch per colour (src/2d/ftn/BMG2_SymStd_relax_GS.f90:78-137)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void recip_k(const double *__restrict__ so,
                                               double *__restrict__ sor, int ii, int jj, int ns) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    const std::size_t P = (std::size_t)ii * jj, bz = blockIdx.z;
    so += bz * ns * P;
    sor += bz * 2 * P;
    if (i <= ii - 2 && j <= jj - 2) {
        const std::size_t p = pidx(i, j, ii);
        sor[P + p] = 1.0 / so[p];
    }
}

// 5-point: colour = parity of (i + j); 9-point: colour = (i parity pi, j parity pj)
template <typename T, bool NINE>
// rev: the tiles are visited in reverse (a colour then starts where the previous one ended, in
// the memory-side cache); points of one colour are independent, so the result is unchanged
__global__ __launch_bounds__(256) void gs_colour_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                   T *__restrict__ q, const T *__restrict__ sor,
                                                   int ii, int jj, int pi, int pj, int rep,
                                                   int rev = 0) {
    const wg_pos g = tiled_wg(rep);
    const int bx = rev ? (int)gridDim.x - 1 - g.x : g.x;
    const int by = rev ? (int)gridDim.y - 1 - g.y : g.y;
    const int k = bx * 64 + (threadIdx.x % 64);
    const int m = by * 4 + (threadIdx.x / 64);
    int i, j;
    if (NINE) {
        i = 2 * k + pi;
        j = 2 * m + pj;
    } else {
        j = m + 1;
        i = 2 * k + ((pi + j) & 1);  // (i + j) % 2 == pi
    }
    if (i < 1 || i > ii - 2 || j < 1 || j > jj - 2) return;
    const std::size_t P = (std::size_t)ii * jj, bz = g.z, bo = bz / rep;
    so += bo * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    sor += bo * 2 * P;
    const std::size_t p = pidx(i, j, ii);
    T v = apply_offdiag<T, NINE>(so, P, q, p, ii, qf[p]);
    q[p] = v * sor[P + p];
}

// ------------------------------------------------------------------------------------------
// line relaxation
//
// setup (src/2d/ftn/BMG2_SymStd_SETUP_lines_{x,y}.f90): diag = c, off = -w (x) or -s (y), then
// LAPACK DPTTRF per line; one thread walks one line sequentially (setup-time only).
// x factors: sor[0][p] = D, sor[1][p] = E at (i, j) as in the reference; y factors stored
// transposed: plane k element (j, i) at k*ii*jj + i*jj + j.
// ------------------------------------------------------------------------------------------
// per: the line direction is periodic, so its wrap couplings are added to the end diagonals
// (src/2d/ftn/BMG2_SymStd_SETUP_lines_x.f90:75-82, _y.f90:77-84) before the factorisation
__global__ void setup_lines_x_k(const double *__restrict__ so, double *__restrict__ sor, int ii,
                                int jj, int ns, int *__restrict__ info, int per = 0) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > jj - 2) return;
    const std::size_t P = (std::size_t)ii * jj;
    so += (std::size_t)blockIdx.y * ns * P;
    sor += (std::size_t)blockIdx.y * 2 * P;
    const int n = ii - 2;
    double *D = sor + pidx(1, j, ii);      // D[k] at i = k + 1
    double *E = sor + P + pidx(2, j, ii);  // E[k] at i = k + 2
    for (int i = 1; i <= n; ++i) {
        const std::size_t p = pidx(i, j, ii);
        sor[P + p] = -so[P * KW + p];
        sor[p] = so[p];
    }
    if (per) {
        D[0] = D[0] + so[P * KW + pidx(1, j, ii)];
        D[n - 1] = D[n - 1] + so[P * KW + pidx(ii - 1, j, ii)];
    }
    for (int k = 0; k < n - 1; ++k) {
        const double d = D[k];
        if (!(d > 0.0)) {
            atomicMax(info, k + 1);
            return;
        }
        const double ei = E[k];
        E[k] = ei / d;
        D[k + 1] = D[k + 1] - E[k] * ei;
    }
    if (!(D[n - 1] > 0.0)) atomicMax(info, n);
}

__global__ void setup_lines_y_k(const double *__restrict__ so, double *__restrict__ sor, int ii,
                                int jj, int ns, int *__restrict__ info, int per = 0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > ii - 2) return;
    const std::size_t P = (std::size_t)ii * jj;
    so += (std::size_t)blockIdx.y * ns * P;
    sor += (std::size_t)blockIdx.y * 2 * P;
    const int n = jj - 2;
    double *D = sor + (std::size_t)i * jj + 1;
    double *E = sor + P + (std::size_t)i * jj + 2;
    for (int j = 1; j <= n; ++j) {
        const std::size_t p = pidx(i, j, ii);
        sor[P + (std::size_t)i * jj + j] = -so[P * KS + p];
        sor[(std::size_t)i * jj + j] = so[p];
    }
    if (per) {
        D[0] = D[0] + so[P * KS + pidx(i, 1, ii)];
        D[n - 1] = D[n - 1] + so[P * KS + pidx(i, jj - 1, ii)];
    }
    for (int k = 0; k < n - 1; ++k) {
        const double d = D[k];
        if (!(d > 0.0)) {
            atomicMax(info, k + 1);
            return;
        }
        const double ei = E[k];
        E[k] = ei / d;
        D[k + 1] = D[k + 1] - E[k] * ei;
    }
    if (!(D[n - 1] > 0.0)) atomicMax(info, n);
}

// Inclusive scans of affine maps y -> a*y + b across the 64 lanes of a wavefront.
// up:   lane l gets F_l o F_{l-1} o ... o F_0        (forward substitution)
// down: lane l gets G_l o G_{l+1} o ... o G_63       (backward substitution)
template <typename T>
__device__ inline void scan_affine_up(T &a, T &b, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const T ap = __shfl_up(a, d, WAVE);
        const T bp = __shfl_up(b, d, WAVE);
        if (lane >= d) {
            b = a * bp + b;
            a = a * ap;
        }
    }
}

template <typename T>
__device__ inline void scan_affine_down(T &a, T &b, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const T an = __shfl_down(a, d, WAVE);
        const T bn = __shfl_down(b, d, WAVE);
        if (lane + d < WAVE) {
            b = a * bn + b;
            a = a * an;
        }
    }
}

// Same inclusive scan with DPP lane moves (row_shr 1/2/4/8 inside 16-lane rows, then
// row_bcast 15 and 31 across rows) instead of LDS permutes: a few cycles per step, which keeps
// the carry chain of a line short.  Lanes whose source is out of range keep the identity (1, 0).
template <int CTRL, int RMASK>
__device__ inline float dpp_move(float v, float ident) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, ident),
                                                                 __builtin_bit_cast(int, v), CTRL,
                                                                 RMASK, 0xF, false));
}

template <int CTRL, int RMASK>
__device__ inline double dpp_move(double v, double ident) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned long long o = __builtin_bit_cast(unsigned long long, ident);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, CTRL, RMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), CTRL,
                                               RMASK, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// The affine step a * c + b of the line recurrences: a fused multiply-add in FP32 (the
// preconditioning V-cycle of the mixed solver), separate rounding in FP64.  Every line kernel
// evaluates its recurrences through this one function, so all of them (every launch shape, the
// direct and the transposed y sweeps, the fused coarse cycle) still round identically.
__device__ inline float affine(float a, float c, float b) { return __builtin_fmaf(a, c, b); }
__device__ inline double affine(double a, double c, double b) { return a * c + b; }
// y / d for a line pivot d: FP32 multiplies by the hardware reciprocal (about 1 ulp), FP64
// divides (correctly rounded, as the sequential substitution does).
__device__ inline float div_pivot(float y, float d) { return y * __builtin_amdgcn_rcpf(d); }
__device__ inline double div_pivot(double y, double d) { return y / d; }

// the source lane's value, 0 where the source lane is out of range or the row is masked off
template <int CTRL, int RMASK>
__device__ inline float dpp_zero(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 CTRL, RMASK, 0xF, true));
}
template <int CTRL, int RMASK>
__device__ inline double dpp_zero(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, RMASK, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, RMASK, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// lanes that receive a value in each step of the scan (row_shr:1/2/4/8 within 16-lane rows,
// row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3); the others keep (a, b)
__device__ inline int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
template <int STEP>
__device__ inline bool dpp_valid(int lane) {
    return STEP == 0 ? (lane & 15) >= 1 : STEP == 1 ? (lane & 15) >= 2 : STEP == 2 ? (lane & 15) >= 4
         : STEP == 3 ? (lane & 15) >= 8 : STEP == 4 ? ((lane >> 4) & 1) != 0 : lane >= 32;
}

// One step: (a, b) <- (a * a', a * b' + b) with (a', b') from the source lane.  Where no value
// arrives, b' reads 0 (a * 0 + b = b exactly) and a is kept: the identity map.  The product
// a * b' is rounded on its own in both precisions (a multiply that takes its lane move for free
// is cheaper here than a fused multiply-add, which cannot).
// The broadcast steps move data into every row (row 0 reads 0); the lanes outside the rows the
// step feeds are then restored by the select, as the row_shr steps restore a.
template <int CTRL, int STEP, typename T>
__device__ inline void dpp_scan_step(T &a, T &b, int lane) {
    const T ap = dpp_zero<CTRL, 0xF>(a);
    const T bp = dpp_zero<CTRL, 0xF>(b);
    const bool v = dpp_valid<STEP>(lane);
    const T bn = a * bp + b;
    b = (STEP < 4 || v) ? bn : b;
    a = v ? a * ap : a;
}

template <typename T>
__device__ inline void scan_affine_dpp(T &a, T &b) {
    const int lane = lane_id();
    dpp_scan_step<0x111, 0>(a, b, lane);  // row_shr:1
    dpp_scan_step<0x112, 1>(a, b, lane);  // row_shr:2
    dpp_scan_step<0x114, 2>(a, b, lane);  // row_shr:4
    dpp_scan_step<0x118, 3>(a, b, lane);  // row_shr:8
    dpp_scan_step<0x142, 4>(a, b, lane);  // row_bcast:15, kept in rows 1 and 3
    dpp_scan_step<0x143, 5>(a, b, lane);  // row_bcast:31, kept in rows 2 and 3
}

// the multiplier part of scan_affine_dpp alone (the same products in the same order)
template <int CTRL, int STEP, typename T>
__device__ inline void dpp_prod_step(T &a, int lane) {
    const T ap = dpp_zero<CTRL, 0xF>(a);
    a = dpp_valid<STEP>(lane) ? a * ap : a;
}

template <typename T>
__device__ inline void scan_prod_dpp(T &a) {
    const int lane = lane_id();
    dpp_prod_step<0x111, 0>(a, lane);
    dpp_prod_step<0x112, 1>(a, lane);
    dpp_prod_step<0x114, 2>(a, lane);
    dpp_prod_step<0x118, 3>(a, lane);
    dpp_prod_step<0x142, 4>(a, lane);
    dpp_prod_step<0x143, 5>(a, lane);
}

// value of lane 63, broadcast (v_readlane: no LDS round trip)
__device__ inline float last_lane(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), WAVE - 1));
}
__device__ inline double last_lane(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, WAVE - 1);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), WAVE - 1);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// One wavefront solves one line: L D L^T x = rhs with the stored factors.  Forward pass in
// 64-element chunks: y_k = rhs_k - E_{k-1} y_{k-1}, stores c_k = y_k / D_k in place of q;
// backward pass: x_k = c_k - E_k x_{k+1}.  A line of one colour only reads lines of the other
// colour, so the in-place staging is race free.
//
// DIR 0: x-lines (row j = first + 2*line, elements along i, unit stride)
// DIR 1: y-lines (column i = first + 2*line, elements along j, stride ii; factors transposed)
template <typename T, bool NINE, int DIR>
__global__ __launch_bounds__(256) void relax_lines_k(const T *__restrict__ so, const T *__restrict__ qf,
                                                     T *__restrict__ q, const T *__restrict__ sor,
                                                     int ii, int jj, int first, int nlines, int rep) {
    const int lane = threadIdx.x % WAVE;
    int bx = (int)blockIdx.x, prob = (int)blockIdx.y;
    shared_op_order((int)gridDim.x, rep, bx, prob);
    const int line = bx * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (line >= nlines) return;  // whole wavefront leaves together; no block barrier below
    const std::size_t P = (std::size_t)ii * jj, bz = prob, bo = bz / rep;
    so += bo * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    sor += bo * 2 * P;
    const int fixed = first + 2 * line;
    const int n = (DIR == 0) ? ii - 2 : jj - 2;
    const std::size_t stride = (DIR == 0) ? 1 : (std::size_t)ii;
    const std::size_t base = (DIR == 0) ? pidx(1, fixed, ii) : pidx(fixed, 1, ii);  // element 0
    const T *D = (DIR == 0) ? sor + pidx(1, fixed, ii) : sor + (std::size_t)fixed * jj + 1;
    const T *E = (DIR == 0) ? sor + P + pidx(2, fixed, ii) : sor + P + (std::size_t)fixed * jj + 2;
    const int nch = (n + WAVE - 1) / WAVE;

    T carry = T(0);
    for (int c = 0; c < nch; ++c) {
        const int k = c * WAVE + lane;
        const bool act = k < n;
        T a = T(0), b = T(0);
        if (act) {
            const std::size_t p = base + (std::size_t)k * stride;
            T rhs = qf[p];
            if (DIR == 0) {
                rhs = rhs + so[P * KS + p] * q[p - ii];
                rhs = rhs + so[P * KS + p + ii] * q[p + ii];
            } else {
                rhs = rhs + so[P * KW + p] * q[p - 1];
                rhs = rhs + so[P * KW + p + 1] * q[p + 1];
            }
            if (NINE) {
                rhs = rhs + so[P * KSW + p] * q[p - ii - 1];
                rhs = rhs + so[P * KNW + p + 1] * q[p - ii + 1];
                rhs = rhs + so[P * KNW + p + ii] * q[p + ii - 1];
                rhs = rhs + so[P * KSW + p + ii + 1] * q[p + ii + 1];
            }
            b = rhs;
            a = (k > 0) ? -E[k - 1] : T(0);
        }
        scan_affine_up(a, b, lane);
        const T y = affine(a, carry, b);
        carry = __shfl(y, WAVE - 1, WAVE);
        if (act) q[base + (std::size_t)k * stride] = div_pivot(y, D[k]);
    }
    carry = T(0);
    for (int c = nch - 1; c >= 0; --c) {
        const int k = c * WAVE + lane;
        const bool act = k < n;
        T a = T(1), b = T(0);
        if (act) {
            b = q[base + (std::size_t)k * stride];
            a = (k < n - 1) ? -E[k] : T(0);
        }
        scan_affine_down(a, b, lane);
        const T x = affine(a, carry, b);
        carry = __shfl(x, 0, WAVE);
        if (act) q[base + (std::size_t)k * stride] = x;
    }
}

// Workgroup index remapped so that consecutive logical tiles run on the same XCD (workgroups
// are dealt round-robin over the 8 XCDs): neighbouring lines share rows of q and of the
// operator, which then stay in one XCD's L2.  Launch 8 * ceil(ntiles / 8) workgroups.
// rev: each XCD walks its range of tiles from the end, so that a launch starts on the lines the
// previous launch touched last (still in the memory-side cache); the work is unchanged.
__device__ inline int xcd_tile(int bx, int ntiles, int rev = 0) {
    const int per = (ntiles + 7) / 8;
    const int k = bx / 8;
    return (bx % 8) * per + (rev ? per - 1 - k : k);
}

// raw operands of the right-hand side of one point of a line
template <typename T>
struct line_raw {
    T qf, s0, q0, s1, q1, sw0, q0m, nw0, q0p, nw1, q1m, sw1, q1p;
};

// Right-hand side of one point in the reference's order.  X-lines: s0/q0 = coupling to the row
// below, s1/q1 = to the row above; y-lines: the column to the left / right.
// SWAP adds the two NW terms in the opposite order: an x-line of the transposed grid then sums
// its terms in exactly the order of the corresponding y-line of the original grid.
template <typename T, bool NINE, bool SWAP = false>
__device__ inline T line_rhs(const line_raw<T> &r) {
    T rhs = r.qf;
    rhs = affine(r.s0, r.q0, rhs);
    rhs = affine(r.s1, r.q1, rhs);
    if (NINE) {
        rhs = affine(r.sw0, r.q0m, rhs);
        if (SWAP) {
            rhs = affine(r.nw1, r.q1m, rhs);
            rhs = affine(r.nw0, r.q0p, rhs);
        } else {
            rhs = affine(r.nw0, r.q0p, rhs);
            rhs = affine(r.nw1, r.q1m, rhs);
        }
        rhs = affine(r.sw1, r.q1p, rhs);
    }
    return rhs;
}

// loads the operands of point p (array index) of an x-line (DIR 0) or y-line (DIR 1)
template <typename T, bool NINE, int DIR>
__device__ inline void line_load(line_raw<T> &r, const T *__restrict__ so, const T *__restrict__ qf,
                                 const T *__restrict__ q, std::size_t P, int ii, std::size_t p,
                                 bool zq = false) {
    r.qf = qf[p];
    if (DIR == 0) {
        r.s0 = so[P * KS + p];
        r.q0 = zq ? T(0) : q[p - ii];
        r.s1 = so[P * KS + p + ii];
        r.q1 = zq ? T(0) : q[p + ii];
    } else {
        r.s0 = so[P * KW + p];
        r.q0 = zq ? T(0) : q[p - 1];
        r.s1 = so[P * KW + p + 1];
        r.q1 = zq ? T(0) : q[p + 1];
    }
    if (NINE) {
        r.sw0 = so[P * KSW + p];
        r.q0m = zq ? T(0) : q[p - ii - 1];
        r.nw0 = so[P * KNW + p + 1];
        r.q0p = zq ? T(0) : q[p - ii + 1];
        r.nw1 = so[P * KNW + p + ii];
        r.q1m = zq ? T(0) : q[p + ii - 1];
        r.sw1 = so[P * KSW + p + ii + 1];
        r.q1p = zq ? T(0) : q[p + ii + 1];
    }
}

// ------------------------------------------------------------------------------------------
// Segmented line solves.  W wavefronts share a line: wavefront w owns the CPW chunks
// [w*CPW, (w+1)*CPW) of 64 elements.  Every operand of a wavefront's chunks is loaded before the
// first scan, so no load latency sits inside the carry chain, and at most CPW chunks live in
// registers; W wavefronts per line put W times as many wavefronts in flight as there are lines.
//
// The carries between chunks are evaluated in exactly the single-wavefront order: chunk m's
// scanned map (a_m, b_m) at its last element goes to LDS, and the carry into chunk m is
// c_m = a_{m-1} c_{m-1} + b_{m-1} from c_0 = 0, the same operations the one-wavefront loop
// performs.  Results are therefore bitwise identical for every (CPW, W) and workgroup shape, so
// the shape can be picked by timing alone.
// ------------------------------------------------------------------------------------------
template <typename T>
struct seg_map {
    T a, b;  // carry_out = a * carry_in + b
};

// Forward recurrence y_k = a_k y_{k-1} + b_k over the wavefront's chunks.  On entry b[u] holds
// chunk u's right-hand sides; afun(u) returns its multipliers a_k (reloaded when needed, so that
// only b stays in registers) and dval(u) its pivots.  On exit b[u] = y_k / d_k.  xch: this
// line's LDS slots, one per chunk of the line.
template <typename T, int CPW, int W, typename AF, typename DF>
__device__ inline void seg_forward(T (&b)[CPW], AF afun, DF dval, seg_map<T> *xch, int w) {
    if (W > 1) {
#pragma unroll
        for (int u = 0; u < CPW; ++u) {
            T a = afun(u);
            scan_affine_dpp(a, b[u]);
            const seg_map<T> m{last_lane(a), last_lane(b[u])};
            if (w < W - 1 && threadIdx.x % WAVE == 0) xch[w * CPW + u] = m;
        }
        lds_barrier();
        // the multipliers are reloaded below rather than kept live across the exchange
        asm volatile("" ::: "memory");
    }
    T carry = T(0);
    for (int m = 0; m < w * CPW; ++m) carry = affine(xch[m].a, carry, xch[m].b);
#pragma unroll
    for (int u = 0; u < CPW; ++u) {
        T a = afun(u);
        if (W > 1)
            scan_prod_dpp(a);
        else
            scan_affine_dpp(a, b[u]);
        const T y = affine(a, carry, b[u]);
        carry = last_lane(y);
        b[u] = div_pivot(y, dval(u));
    }
}

// Backward recurrence x_k = a_k x_{k+1} + b_k, lanes in reversed element order (lane l of chunk
// c holds element 64c + 63 - l); the carry runs from the line's last chunk to its first.
// coef(u, a, b[u]) turns b[u] (the forward results) into the chunk's coefficients and returns
// the multipliers in a; afun(u) returns the multipliers alone.  On exit b[u] = x.
template <typename T, int CPW, int W, typename CF, typename AF>
__device__ inline void seg_backward(T (&b)[CPW], CF coef, AF afun, seg_map<T> *xch, int w) {
    if (W > 1) {
#pragma unroll
        for (int u = 0; u < CPW; ++u) {
            T a;
            coef(u, a, b[u]);
            scan_affine_dpp(a, b[u]);
            const seg_map<T> m{last_lane(a), last_lane(b[u])};
            if (w > 0 && threadIdx.x % WAVE == 0) xch[w * CPW + u] = m;
        }
        lds_barrier();
        // the multipliers are reloaded below rather than kept live across the exchange
        asm volatile("" ::: "memory");
    }
    T carry = T(0);
    for (int m = W * CPW - 1; m >= (w + 1) * CPW; --m) carry = affine(xch[m].a, carry, xch[m].b);
#pragma unroll
    for (int u = CPW - 1; u >= 0; --u) {
        T a;
        if (W > 1) {
            a = afun(u);
            scan_prod_dpp(a);
        } else {
            coef(u, a, b[u]);
            scan_affine_dpp(a, b[u]);
        }
        const T x = affine(a, carry, b[u]);
        carry = last_lane(x);
        b[u] = x;
    }
}

// Register-rich forms of the two recurrences (scanned multipliers kept in a[]), used by the
// x-line kernel: forward recurrence y_k = a_k y_{k-1} + b_k over the wavefront's chunks; on exit
// b[u] = y_k / D(u) with D(u) the pivots of chunk u (a functor, loaded after the scans).
// xch: this line's LDS slots, one per chunk of the line.
template <typename T, int CPW, int W, typename DF>
__device__ inline void seg_forward_ab(T (&a)[CPW], T (&b)[CPW], DF dval, seg_map<T> *xch, int w) {
#pragma unroll
    for (int u = 0; u < CPW; ++u) scan_affine_dpp(a[u], b[u]);
    T carry = T(0);
    if (W > 1) {
        if (w < W - 1) {  // chunk-end maps of all but the last segment
#pragma unroll
            for (int u = 0; u < CPW; ++u) {
                const seg_map<T> m{last_lane(a[u]), last_lane(b[u])};
                if (threadIdx.x % WAVE == 0) xch[w * CPW + u] = m;
            }
        }
        lds_barrier();
        for (int m = 0; m < w * CPW; ++m) carry = affine(xch[m].a, carry, xch[m].b);
    }
#pragma unroll
    for (int u = 0; u < CPW; ++u) {
        const T y = affine(a[u], carry, b[u]);
        carry = last_lane(y);
        b[u] = div_pivot(y, dval(u));
    }
}

// Backward recurrence x_k = a_k x_{k+1} + b_k, lanes in reversed element order (lane l of chunk
// c holds element 64c + 63 - l); the carry runs from the line's last chunk to its first.  On
// exit b[u] = x.
template <typename T, int CPW, int W>
__device__ inline void seg_backward_ab(T (&a)[CPW], T (&b)[CPW], seg_map<T> *xch, int w) {
#pragma unroll
    for (int u = 0; u < CPW; ++u) scan_affine_dpp(a[u], b[u]);
    T carry = T(0);
    if (W > 1) {
        if (w > 0) {
#pragma unroll
            for (int u = 0; u < CPW; ++u) {
                const seg_map<T> m{last_lane(a[u]), last_lane(b[u])};
                if (threadIdx.x % WAVE == 0) xch[w * CPW + u] = m;
            }
        }
        lds_barrier();
        for (int m = W * CPW - 1; m >= (w + 1) * CPW; --m) carry = affine(xch[m].a, carry, xch[m].b);
    }
#pragma unroll
    for (int u = CPW - 1; u >= 0; --u) {
        const T x = affine(a[u], carry, b[u]);
        carry = last_lane(x);
        b[u] = x;
    }
}

// x-lines of one colour, in place: L lines per workgroup, W wavefronts per line.  TR: the grid
// is the transpose of a y-line problem (right-hand side terms in the y-line order).
// rev: the tile walk of xcd_tile.  ZQ (zero start): q holds zeros outside the lines of this
// colour (a sweep that starts from a zero iterate), so the neighbour rows are not read and
// literal zeros enter the same arithmetic (a compile-time switch: a run-time select in front of
// the loads slowed the kernel by a quarter).
template <typename T, bool NINE, int CPW, int W, int L, bool TR = false, bool ZQ = false>
__global__ __launch_bounds__(L *W * 64) void relax_xlines_seg_k(const T *__restrict__ so,
                                                               const T *__restrict__ qf,
                                                               T *__restrict__ q,
                                                               const T *__restrict__ sor, int ii,
                                                               int jj, int first, int nlines,
                                                               int ntiles, int rev, int rep) {
    __shared__ seg_map<T> xch[2][L][W > 1 ? W * CPW : 1];
    int bx = (int)blockIdx.x, prob = (int)blockIdx.y;
    shared_op_order((int)gridDim.x, rep, bx, prob);  // gridDim.x is a multiple of 8
    const int tile = xcd_tile(bx, ntiles, rev);
    if (tile >= ntiles) return;  // uniform over the workgroup
    const int wl = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int slot = wl / W, w = wl % W;
    const int line = tile * L + slot;
    const bool valid = line < nlines;  // invalid wavefronts solve a copy of the last line
    const int j = first + 2 * min(line, nlines - 1);
    const std::size_t P = (std::size_t)ii * jj, bz = prob, bo = bz / rep;
    so += bo * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    sor += bo * 2 * P;
    const int n = ii - 2;
    const std::size_t row = (std::size_t)j * ii;
    const T *D = sor + row + 1;      // D[k]
    const T *E = sor + P + row + 2;  // E[k], k < n-1
    const int rl = WAVE - 1 - lane;
    auto bwd_coef = [&](int u, T &au, T &bu) {  // chunk u of the backward pass, reversed lanes
        const int k = (w * CPW + u) * WAVE + rl;
        const T c = __shfl(bu, rl, WAVE);
        const T e = E[min(k, max(n - 2, 0))];
        au = (k < n - 1) ? -e : T(0);
        au = (k < n) ? au : T(1);
        bu = (k < n) ? c : T(0);
    };
    T b[CPW];
    if constexpr (W == 1) {
        // one wavefront per line: scan and carry chunk by chunk, only the results kept
        T carry = T(0);
#pragma unroll
        for (int u = 0; u < CPW; ++u) {
            const int k = u * WAVE + lane;
            const int kc = min(k, n - 1);
            line_raw<T> r;
            line_load<T, NINE, 0>(r, so, qf, q, P, ii, row + kc + 1, ZQ);
            T bu = line_rhs<T, NINE, TR>(r);
            T au = (k > 0) ? -E[max(kc - 1, 0)] : T(0);
            scan_affine_dpp(au, bu);
            const T y = affine(au, carry, bu);
            carry = last_lane(y);
            b[u] = div_pivot(y, D[kc]);
        }
        carry = T(0);
#pragma unroll
        for (int u = CPW - 1; u >= 0; --u) {
            T au;
            bwd_coef(u, au, b[u]);
            scan_affine_dpp(au, b[u]);
            const T x = affine(au, carry, b[u]);
            carry = last_lane(x);
            b[u] = x;
        }
    } else {
        T a[CPW];
#pragma unroll
        for (int u = 0; u < CPW; ++u) {
            const int k = (w * CPW + u) * WAVE + lane;
            const int kc = min(k, n - 1);
            line_raw<T> r;
            line_load<T, NINE, 0>(r, so, qf, q, P, ii, row + kc + 1, ZQ);
            b[u] = line_rhs<T, NINE, TR>(r);
            a[u] = (k > 0) ? -E[max(kc - 1, 0)] : T(0);
        }
        seg_forward_ab<T, CPW, W>(
            a, b, [&](int u) { return D[min((w * CPW + u) * WAVE + lane, n - 1)]; },
            xch[0][slot], w);
#pragma unroll
        for (int u = 0; u < CPW; ++u) bwd_coef(u, a[u], b[u]);
        seg_backward_ab<T, CPW, W>(a, b, xch[1][slot], w);
    }
    if (valid) {
#pragma unroll
        for (int u = 0; u < CPW; ++u) {
            const int k = (w * CPW + u) * WAVE + rl;
            if (k < n) q[row + k + 1] = b[u];
        }
    }
}

// y-lines of one colour, in place: NC columns per workgroup, W wavefronts per column.  The
// right-hand sides of G chunk rows at a time are formed by the whole workgroup with
// row-contiguous loads (thread t: column t % NC, row t / NC of the W*64 rows) and handed to the
// column owners through LDS; results return through LDS for row-contiguous stores.
template <typename T, bool NINE, int CPW, int W, int NC, int G>
__global__ __launch_bounds__(NC *W * 64) void relax_ylines_seg_k(const T *__restrict__ so,
                                                                const T *__restrict__ qf,
                                                                T *__restrict__ q,
                                                                const T *__restrict__ sor,
                                                                int ii, int jj, int first,
                                                                int ncols, int ntiles, int rev,
                                                                int rep) {
    static_assert(CPW % G == 0, "G must divide CPW");
    __shared__ T tile_s[G][W][NC][WAVE + 1];
    __shared__ seg_map<T> xch[2][NC][W > 1 ? W * CPW : 1];
    int bx = (int)blockIdx.x, prob = (int)blockIdx.y;
    shared_op_order((int)gridDim.x, rep, bx, prob);  // gridDim.x is a multiple of 8
    const int tile = xcd_tile(bx, ntiles, rev);
    if (tile >= ntiles) return;  // uniform over the workgroup
    const int wl = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int cs = wl / W, w = wl % W;
    const std::size_t P = (std::size_t)ii * jj, bz = prob, bo = bz / rep;
    so += bo * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    sor += bo * 2 * P;
    const int n = jj - 2;
    const int col0 = tile * NC;
    const int nc = min(NC, ncols - col0);
    const int cb = first + 2 * col0;
    // the wavefront's own column (factors contiguous in j)
    const int ime = cb + 2 * min(cs, nc - 1);
    const T *D = sor + (std::size_t)ime * jj + 1;
    const T *E = sor + P + (std::size_t)ime * jj + 2;
    // this thread's tile position for the right-hand sides and the stores
    const int tc = threadIdx.x % NC, tr = threadIdx.x / NC;
    const int tseg = tr / WAVE, trow = tr % WAVE;
    const bool tok = tc < nc;
    const int ti = cb + 2 * min(tc, nc - 1);
    const int rl = WAVE - 1 - lane;
    auto fwd_a = [&](int u) {  // forward coefficient of chunk u
        const int k = (w * CPW + u) * WAVE + lane;
        return (k > 0) ? -E[max(min(k, n - 1) - 1, 0)] : T(0);
    };
    // Right-hand sides of chunk rows g .. g+G-1 go through the tile.  The raw operands of the
    // next group are loaded before the current group's barrier (software pipelining), so one
    // group of loads is always in flight across the LDS exchange.
    line_raw<T> raw[G];
    auto load_group = [&](int g) {
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int k = min((tseg * CPW + g + u) * WAVE + trow, n - 1);
            line_load<T, NINE, 1>(raw[u], so, qf, q, P, ii, pidx(ti, k + 1, ii));
        }
    };
    auto form_tile = [&](int g) {
#pragma unroll
        for (int u = 0; u < G; ++u) tile_s[u][tseg][tc][trow] = line_rhs<T, NINE>(raw[u]);
        if (g + G < CPW) load_group(g + G);
    };
    load_group(0);
    auto bwd_a = [&](int u) {  // backward multipliers of chunk u, reversed lanes
        const int k = (w * CPW + u) * WAVE + rl;
        const T e = E[min(k, max(n - 2, 0))];
        const T a = (k < n - 1) ? -e : T(0);
        return (k < n) ? a : T(1);
    };
    auto bwd_coef = [&](int u, T &au, T &bu) {
        const int k = (w * CPW + u) * WAVE + rl;
        const T c = __shfl(bu, rl, WAVE);
        au = bwd_a(u);
        bu = (k < n) ? c : T(0);
    };
    // all right-hand sides first (memory streaming only), then the scans
    T b[CPW];
#pragma unroll
    for (int g = 0; g < CPW; g += G) {
        form_tile(g);
        lds_barrier();
#pragma unroll
        for (int u = 0; u < G; ++u) b[g + u] = tile_s[u][w][cs][lane];
        lds_barrier();
    }
    seg_forward<T, CPW, W>(
        b, fwd_a, [&](int u) { return D[min((w * CPW + u) * WAVE + lane, n - 1)]; }, xch[0][cs],
        w);
    seg_backward<T, CPW, W>(b, bwd_coef, bwd_a, xch[1][cs], w);
#pragma unroll
    for (int g = 0; g < CPW; g += G) {
#pragma unroll
        for (int u = 0; u < G; ++u) tile_s[u][w][cs][rl] = b[g + u];
        lds_barrier();
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int k = (tseg * CPW + g + u) * WAVE + trow;
            if (tok && k < n) q[pidx(ti, k + 1, ii)] = tile_s[u][tseg][tc][trow];
        }
        lds_barrier();
    }
}

// ------------------------------------------------------------------------------------------
// restriction qc = P^T q (src/2d/ftn/BMG2_SymStd_restrict.f90:76-92); coarse (ic, jc) sits on
// fine (2ic-1, 2jc-1)
// ------------------------------------------------------------------------------------------
// coarse point (ic, jc) of qc = P^T q (the reference's operation order); ci, q, qc of one problem
template <typename T>
__device__ inline T restrict_point(const T *__restrict__ ci, const T *__restrict__ q, int iif,
                                   int iic, std::size_t Pc, int ic, int jc) {
    const int i = 2 * ic - 1, j = 2 * jc - 1;
    auto CI = [&](int a, int b, int k) { return ci[Pc * k + pidx(a, b, iic)]; };
    auto Q = [&](int a, int b) { return q[pidx(a, b, iif)]; };
    T v = CI(ic, jc, LNE) * Q(i - 1, j - 1);
    v = v + CI(ic, jc, LA) * Q(i, j - 1);
    v = v + CI(ic + 1, jc, LNW) * Q(i + 1, j - 1);
    v = v + CI(ic, jc, LR) * Q(i - 1, j);
    v = v + Q(i, j);
    v = v + CI(ic + 1, jc, LL) * Q(i + 1, j);
    v = v + CI(ic, jc + 1, LSE) * Q(i - 1, j + 1);
    v = v + CI(ic, jc + 1, LB) * Q(i, j + 1);
    v = v + CI(ic + 1, jc + 1, LSW) * Q(i + 1, j + 1);
    return v;
}

template <typename T>
__global__ __launch_bounds__(256) void restrict_k(const T *__restrict__ ci, const T *__restrict__ q,
                                                  T *__restrict__ qc, int iif, int jjf, int iic,
                                                  int jjc, int rep) {
    const wg_pos g = tiled_wg(rep);
    const int ic = g.x * 64 + (threadIdx.x % 64) + 1;
    const int jc = g.y * 4 + (threadIdx.x / 64) + 1;
    if (ic > iic - 2 || jc > jjc - 2) return;
    const std::size_t Pc = (std::size_t)iic * jjc, bz = g.z;
    ci += (bz / rep) * 8 * Pc;
    q += bz * iif * (std::size_t)jjf;
    qc += bz * Pc;
    qc[pidx(ic, jc, iic)] = restrict_point<T>(ci, q, iif, iic, Pc, ic, jc);
}

// Restriction for a V-cycle step: qc = P^T q as restrict_k (same operations), and in the same
// launch the coarse iterate xc is zeroed (ghosts included: the coarse cycle starts from 0) and,
// where qcT is given, qc is also written transposed (qcT[ic * jjc + jc], for a coarse level whose
// y sweeps run through the transposed grid).  Tiles of 64 (ic) x 16 (jc) coarse points over the
// ghost-inclusive coarse grid, 64 * rc_rows<T> threads.
// FP64: one thread per coarse point of the tile, so that every point's loads are in flight at
// once (with a loop over rows the FP64 build kept one row's loads in flight at a time).  FP32: 4
// rows per thread (one point per thread was 3.8x slower over the levels of a batch of 116
// problems of 2049^2, and no faster on 4095^2).
template <typename T>
constexpr int rc_rows = sizeof(T) == 8 ? 16 : 4;
template <typename T>
__global__ __launch_bounds__(64 * rc_rows<T>) void restrict_cycle_k(const T *__restrict__ ci,
                                                                    const T *__restrict__ q,
                                                                    T *__restrict__ qc,
                                                                    T *__restrict__ xc,
                                                                    T *__restrict__ qcT, int iif,
                                                                    int jjf, int iic, int jjc,
                                                                    int rep) {
    constexpr int NT = 64 * rc_rows<T>;
    __shared__ T tile[16][65];
    const wg_pos g = tiled_wg(rep);
    const int ic0 = g.x * 64, jc0 = g.y * 16;
    const std::size_t Pc = (std::size_t)iic * jjc, bz = g.z;
    ci += (bz / rep) * 8 * Pc;
    q += bz * iif * (std::size_t)jjf;
    qc += bz * Pc;
    xc += bz * Pc;
    const int tx = threadIdx.x % 64, ic = ic0 + tx;
#pragma unroll
    for (int u = 0; u < 16 / rc_rows<T>; ++u) {
        const int k = threadIdx.x / 64 + u * rc_rows<T>;
        const int jc = jc0 + k;
        if (ic >= iic || jc >= jjc) continue;
        const std::size_t pc = pidx(ic, jc, iic);
        xc[pc] = T(0);
        if (ic >= 1 && ic <= iic - 2 && jc >= 1 && jc <= jjc - 2) {
            const T v = restrict_point<T>(ci, q, iif, iic, Pc, ic, jc);
            qc[pc] = v;
            tile[k][tx] = v;
        }
    }
    if (!qcT) return;  // uniform over the launch
    qcT += bz * Pc;
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 16; e += NT) {
        const int a = e / 16, c = e % 16;  // a: ic offset, c: jc offset (contiguous in qcT)
        const int i2 = ic0 + a, j2 = jc0 + c;
        if (i2 >= 1 && i2 <= iic - 2 && j2 >= 1 && j2 <= jjc - 2)
            qcT[(std::size_t)i2 * jjc + j2] = tile[c][a];
    }
}

// ------------------------------------------------------------------------------------------
// interp_add q += P qc + res / c at F-points (src/2d/ftn/BMG2_SymStd_interp_add.f90:101-137);
// one thread per 2 x 2 block of fine points.  res is only read (the reference divides it in place; nothing
// reads it afterwards).  For an even fine extent the reference's loops also reach the last
// fine ghost column / row (coarse index IICF1 = IIC); those points are updated the same way,
// with the raw res value there since the reference divides res only in the interior.
// ------------------------------------------------------------------------------------------
// One 2 x 2 block of fine points (2m-1 .. 2m, 2n-1 .. 2n) of interp_add: the coarse point
// (m, n), the x-edge point (2m, 2n-1), the y-edge point (2m-1, 2n) and the cell point (2m, 2n),
// sharing the four coarse values around the cell.  Arrays of one problem.
// Two adjacent floats moved by one 8-byte access.  On this GPU an 8-byte global access needs only
// 4-byte alignment, so a pair may start at any float (rows of odd length included).
struct __attribute__((aligned(4))) float_pair {
    float a, b;
};
__device__ inline void load_pair(const float *p, float &a, float &b) {
    float_pair v;
    __builtin_memcpy(&v, p, sizeof(v));
    a = v.a;
    b = v.b;
}
__device__ inline void store_pair(float *p, float a, float b) {
    const float_pair v{a, b};
    __builtin_memcpy(p, &v, sizeof(v));
}

template <typename T>
__device__ inline void interp_block(T *__restrict__ q, const T *__restrict__ qc,
                                    const T *__restrict__ res, const T *__restrict__ so,
                                    const T *__restrict__ ci, int iic, int jjc, int iif, int jjf,
                                    int m, int n) {
    const int imax = iif - 2 + ((iif - 2) % 2 == 0 ? 1 : 0);
    const int jmax = jjf - 2 + ((jjf - 2) % 2 == 0 ? 1 : 0);
    const int i0 = 2 * m - 1, j0 = 2 * n - 1;
    if (i0 > imax || j0 > jmax) return;
    const bool xe = i0 + 1 <= imax, ye = j0 + 1 <= jmax;  // the even fine points exist
    const std::size_t Pc = (std::size_t)iic * jjc;
    auto CI = [&](int a, int b, int k) { return ci[Pc * k + pidx(a, b, iic)]; };
    auto QC = [&](int a, int b) { return qc[pidx(a, b, iic)]; };
    auto R = [&](int i, int j, std::size_t p) {
        return (i > iif - 2 || j > jjf - 2) ? res[p] : res[p] / so[p];
    };
    if constexpr (sizeof(T) == 4) {
        // FP32: the same operations, with horizontally adjacent values moved as pairs where the
        // point (2m, .) exists (xe): then i0 + 1 <= iif - 1 and m + 1 <= iic - 1, so every pair
        // lies inside its row (for even sizes i0 can be the ghost column, where xe is false)
        const std::size_t p = pidx(i0, j0, iif);
        T c00, c10 = T(0), c01 = T(0), c11 = T(0);
        if (xe) {
            load_pair(qc + pidx(m, n, iic), c00, c10);
            if (ye) load_pair(qc + pidx(m, n + 1, iic), c01, c11);
        } else {
            c00 = QC(m, n);
            if (ye) c01 = QC(m, n + 1);
        }
        if (xe) {  // coarse point and (2m, 2n-1)
            T q0, q1;
            load_pair(q + p, q0, q1);
            q0 = q0 + c00;
            const T a = CI(m + 1, n, LR) * c10 + CI(m + 1, n, LL) * c00;
            q1 = q1 + a + R(i0 + 1, j0, p + 1);
            store_pair(q + p, q0, q1);
        } else {
            q[p] = q[p] + c00;
        }
        if (ye) {  // (2m-1, 2n) and (2m, 2n)
            const std::size_t p2 = p + iif;
            const T a = CI(m, n + 1, LA) * c01 + CI(m, n + 1, LB) * c00;
            if (xe) {
                T q2, q3, r2, r3, s2, s3;
                load_pair(q + p2, q2, q3);
                load_pair(res + p2, r2, r3);
                load_pair(so + p2, s2, s3);
                q2 = q2 + a + ((i0 > iif - 2 || j0 + 1 > jjf - 2) ? r2 : r2 / s2);
                T b = CI(m + 1, n + 1, LSW) * c00;
                b = b + CI(m + 1, n + 1, LNW) * c01;
                b = b + CI(m + 1, n + 1, LNE) * c11;
                b = b + CI(m + 1, n + 1, LSE) * c10;
                q3 = q3 + b + ((i0 + 1 > iif - 2 || j0 + 1 > jjf - 2) ? r3 : r3 / s3);
                store_pair(q + p2, q2, q3);
            } else {
                q[p2] = q[p2] + a + R(i0, j0 + 1, p2);
            }
        }
        return;
    }
    const T q00 = QC(m, n);
    {  // coarse point
        const std::size_t p = pidx(i0, j0, iif);
        q[p] = q[p] + q00;
    }
    if (xe) {  // (2m, 2n-1): between coarse (m, n) and (m+1, n)
        const std::size_t p = pidx(i0 + 1, j0, iif);
        const T a = CI(m + 1, n, LR) * QC(m + 1, n) + CI(m + 1, n, LL) * q00;
        q[p] = q[p] + a + R(i0 + 1, j0, p);
    }
    if (ye) {  // (2m-1, 2n): between coarse (m, n) and (m, n+1)
        const std::size_t p = pidx(i0, j0 + 1, iif);
        const T a = CI(m, n + 1, LA) * QC(m, n + 1) + CI(m, n + 1, LB) * q00;
        q[p] = q[p] + a + R(i0, j0 + 1, p);
    }
    if (xe && ye) {  // (2m, 2n): the cell of coarse (m..m+1, n..n+1)
        const std::size_t p = pidx(i0 + 1, j0 + 1, iif);
        T a = CI(m + 1, n + 1, LSW) * q00;
        a = a + CI(m + 1, n + 1, LNW) * QC(m, n + 1);
        a = a + CI(m + 1, n + 1, LNE) * QC(m + 1, n + 1);
        a = a + CI(m + 1, n + 1, LSE) * QC(m + 1, n);
        q[p] = q[p] + a + R(i0 + 1, j0 + 1, p);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void interp_add_k(T *__restrict__ q, const T *__restrict__ qc,
                                                    const T *__restrict__ res,
                                                    const T *__restrict__ so,
                                                    const T *__restrict__ ci, int iic, int jjc,
                                                    int iif, int jjf, int ns, int rep) {
    const wg_pos g = tiled_wg(rep);
    const int m = g.x * 64 + (threadIdx.x % 64) + 1;
    const int n = g.y * 4 + (threadIdx.x / 64) + 1;
    const std::size_t Pc = (std::size_t)iic * jjc, Pf = (std::size_t)iif * jjf, bz = g.z;
    const std::size_t bo = bz / rep;
    interp_block<T>(q + bz * Pf, qc + bz * Pc, res + bz * Pf, so + bo * ns * Pf, ci + bo * 8 * Pc,
                    iic, jjc, iif, jjf, m, n);
}

// ------------------------------------------------------------------------------------------
// operator-induced interpolation (src/2d/ftn/BMG2_SymStd_SETUP_interp_OI.f90:109-257)
// ------------------------------------------------------------------------------------------
__device__ inline double lumped(double c, double sum, double ep) {
    const double emach = 0x1.0p-53;  // D1MACH(3)
    return (c - sum) * fmax(c - (1.0 + ep) * sum, 0.0) / (fabs(c - (1.0 + ep) * sum) + emach);
}

// edge weights: x-edges (LR, LL) for ic in [2, nxf/2+1], jc in [1, ncy]; y-edges (LA, LB) for
// ic in [1, ncx], jc in [2, nyf/2+1].  Threads cover the union of both ranges.
template <bool NINE>
__global__ __launch_bounds__(256) void oi_edges_k(const double *__restrict__ so, double *__restrict__ ci,
                                                  int iif, int jjf, int iic, int jjc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    auto S = [&](int a, int b, int k) { return so[P * k + pidx(a, b, iif)]; };
    const int icmax = (iif - 2) / 2 + 1, jcmax = (jjf - 2) / 2 + 1;
    const int i = 2 * ic - 1, j = 2 * jc - 1;
    const std::size_t pc = pidx(ic, jc, iic);
    if (ic >= 2 && ic <= icmax && jc <= jjc - 2) {
        double a, b;
        if (NINE) {
            a = S(i, j, KW) + S(i, j, KNW) + S(i, j + 1, KSW);
            b = S(i - 1, j, KW) + S(i - 1, j, KSW) + S(i - 1, j + 1, KNW);
        } else {
            a = S(i, j, KW);
            b = S(i - 1, j, KW);
        }
        const double c = S(i - 1, j, KO);
        const double ep = fmin(fabs(a / c), fabs(b / c));
        double sum = a + b + S(i - 1, j, KS) + S(i - 1, j + 1, KS);
        sum = a + b + lumped(c, sum, ep);
        sum = 1.0 / sum;
        ci[Pc * LR + pc] = a * sum;
        ci[Pc * LL + pc] = b * sum;
    }
    if (jc >= 2 && jc <= jcmax && ic <= iic - 2) {
        double a, b;
        if (NINE) {
            a = S(i, j, KS) + S(i, j, KNW) + S(i + 1, j, KSW);
            b = S(i, j - 1, KS) + S(i, j - 1, KSW) + S(i + 1, j - 1, KNW);
        } else {
            a = S(i, j, KS);
            b = S(i, j - 1, KS);
        }
        const double c = S(i, j - 1, KO);
        const double ep = fmin(fabs(a / c), fabs(b / c));
        double sum = a + b + S(i, j - 1, KW) + S(i + 1, j - 1, KW);
        sum = a + b + lumped(c, sum, ep);
        sum = 1.0 / sum;
        ci[Pc * LA + pc] = a * sum;
        ci[Pc * LB + pc] = b * sum;
    }
}

// corner weights for ic in [2, nxf/2+1], jc in [2, nyf/2+1], from the edge weights
template <bool NINE>
__global__ __launch_bounds__(256) void oi_corners_k(const double *__restrict__ so, double *__restrict__ ci,
                                                    int iif, int jjf, int iic, int jjc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    const int icmax = (iif - 2) / 2 + 1, jcmax = (jjf - 2) / 2 + 1;
    if (ic > icmax || jc > jcmax) return;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    auto S = [&](int a, int b, int k) { return so[P * k + pidx(a, b, iif)]; };
    auto CI = [&](int a, int b, int k) { return ci[Pc * k + pidx(a, b, iic)]; };
    const int i = 2 * ic - 1, j = 2 * jc - 1;
    const double c = S(i - 1, j - 1, KO);
    double sum, ep;
    if (NINE) {
        sum = S(i - 1, j - 1, KW) + S(i - 1, j, KNW) + S(i - 1, j, KS) + S(i, j, KSW)
            + S(i, j - 1, KW) + S(i, j - 1, KNW) + S(i - 1, j - 1, KS) + S(i - 1, j - 1, KSW);
        const double e1 = fabs((S(i - 1, j - 1, KSW) + S(i - 1, j - 1, KW) + S(i - 1, j, KNW)) / c);
        const double e2 = fabs((S(i - 1, j, KNW) + S(i - 1, j, KS) + S(i, j, KSW)) / c);
        const double e3 = fabs((S(i, j, KSW) + S(i, j - 1, KW) + S(i, j - 1, KNW)) / c);
        const double e4 = fabs((S(i, j - 1, KNW) + S(i - 1, j - 1, KS) + S(i - 1, j - 1, KSW)) / c);
        ep = fmin(fmin(fmin(e1, e2), e3), e4);
    } else {
        sum = S(i - 1, j - 1, KW) + S(i - 1, j, KS) + S(i, j - 1, KW) + S(i - 1, j - 1, KS);
        ep = fmin(fmin(fmin(fabs(S(i - 1, j - 1, KW) / c), fabs(S(i - 1, j, KS) / c)),
                       fabs(S(i, j - 1, KW) / c)),
                  fabs(S(i - 1, j - 1, KS) / c));
    }
    sum = sum + lumped(c, sum, ep);
    const double s = 1.0 / sum;
    const std::size_t pc = pidx(ic, jc, iic);
    double sw = S(i - 1, j - 1, KS) * CI(ic, jc - 1, LL) + S(i - 1, j - 1, KW) * CI(ic - 1, jc, LB);
    double se = S(i - 1, j - 1, KS) * CI(ic, jc - 1, LR) + S(i, j - 1, KW) * CI(ic, jc, LB);
    double nw = S(i - 1, j - 1, KW) * CI(ic - 1, jc, LA) + S(i - 1, j, KS) * CI(ic, jc, LL);
    double ne = S(i - 1, j, KS) * CI(ic, jc, LR) + S(i, j - 1, KW) * CI(ic, jc, LA);
    if (NINE) {
        sw = sw + S(i - 1, j - 1, KSW);
        se = se + S(i, j - 1, KNW);
        nw = nw + S(i - 1, j, KNW);
        ne = ne + S(i, j, KSW);
    }
    ci[Pc * LSW + pc] = sw * s;
    ci[Pc * LSE + pc] = se * s;
    ci[Pc * LNW + pc] = nw * s;
    ci[Pc * LNE + pc] = ne * s;
}

// ------------------------------------------------------------------------------------------
// Galerkin coarse operator, formed generically: K(X, Y) = sum_f sum_g P(f,X) A(f,g) P(g,Y)
// over the 3x3 fine supports of the two coarse points.  Coarse entries (lower half):
// c = K(C,C), w = -K(C,C-e_x), s = -K(C,C-e_y), sw = -K(C,C-e_x-e_y),
// nw(C) = -K(C-e_x, C-e_y).  Couplings to coarse ghosts are zero (Dirichlet).
// ------------------------------------------------------------------------------------------
// weight of fine point centre(ic,jc) + (di,dj) in the column of coarse point (ic, jc)
__device__ inline double pweight(const double *__restrict__ ci, std::size_t Pc, int iic, int ic,
                                 int jc, int di, int dj) {
    auto CI = [&](int a, int b, int k) { return ci[Pc * k + pidx(a, b, iic)]; };
    if (di == 0 && dj == 0) return 1.0;
    if (di == -1 && dj == 0) return CI(ic, jc, LR);
    if (di == 1 && dj == 0) return CI(ic + 1, jc, LL);
    if (di == 0 && dj == -1) return CI(ic, jc, LA);
    if (di == 0 && dj == 1) return CI(ic, jc + 1, LB);
    if (di == -1 && dj == -1) return CI(ic, jc, LNE);
    if (di == 1 && dj == -1) return CI(ic + 1, jc, LNW);
    if (di == -1 && dj == 1) return CI(ic, jc + 1, LSE);
    return CI(ic + 1, jc + 1, LSW);
}

// A(f, f + (dx, dy)) of the fine operator
template <bool NINE>
__device__ inline double aentry(const double *__restrict__ so, std::size_t P, int ii, int x, int y,
                                int dx, int dy) {
    const std::size_t p = pidx(x, y, ii);
    if (dx == 0 && dy == 0) return so[p];
    if (dx == -1 && dy == 0) return -so[P * KW + p];
    if (dx == 1 && dy == 0) return -so[P * KW + p + 1];
    if (dx == 0 && dy == -1) return -so[P * KS + p];
    if (dx == 0 && dy == 1) return -so[P * KS + p + ii];
    if (!NINE) return 0.0;
    if (dx == -1 && dy == -1) return -so[P * KSW + p];
    if (dx == 1 && dy == 1) return -so[P * KSW + p + ii + 1];
    if (dx == 1 && dy == -1) return -so[P * KNW + p + 1];
    return -so[P * KNW + p + ii];
}

template <bool NINE, int DXC, int DYC>
__device__ inline double kpair(const double *__restrict__ so, const double *__restrict__ ci,
                               int iif, int jjf, int iic, int jjc, int icx, int jcx) {
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc;
    const int nxf = iif - 2, nyf = jjf - 2;
    const int X0 = 2 * icx - 1, Y0 = 2 * jcx - 1;
    const int icy = icx + DXC, jcy = jcx + DYC;
    double sum = 0.0;
#pragma unroll
    for (int fy = -1; fy <= 1; ++fy) {
#pragma unroll
        for (int fx = -1; fx <= 1; ++fx) {
            const int x = X0 + fx, y = Y0 + fy;
            if (x < 1 || x > nxf || y < 1 || y > nyf) continue;
            double inner = 0.0;
            bool any = false;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!NINE && dx != 0 && dy != 0) continue;
                    const int gx = fx + dx - 2 * DXC, gy = fy + dy - 2 * DYC;
                    if (gx < -1 || gx > 1 || gy < -1 || gy > 1) continue;
                    const int xg = x + dx, yg = y + dy;
                    if (xg < 1 || xg > nxf || yg < 1 || yg > nyf) continue;
                    inner += aentry<NINE>(so, P, iif, x, y, dx, dy) *
                             pweight(ci, Pc, iic, icy, jcy, gx, gy);
                    any = true;
                }
            }
            if (any) sum += pweight(ci, Pc, iic, icx, jcx, fx, fy) * inner;
        }
    }
    return sum;
}

template <bool NINE>
__global__ __launch_bounds__(256) void galerkin_k(const double *__restrict__ so,
                                                  const double *__restrict__ ci,
                                                  double *__restrict__ soc, int iif, int jjf,
                                                  int iic, int jjc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    if (ic > iic - 2 || jc > jjc - 2) return;
    const std::size_t Pc = (std::size_t)iic * jjc, pc = pidx(ic, jc, iic), bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * (std::size_t)iif * jjf;
    ci += bz * 8 * Pc;
    soc += bz * 5 * Pc;
    soc[pc] = kpair<NINE, 0, 0>(so, ci, iif, jjf, iic, jjc, ic, jc);
    soc[Pc * KW + pc] = (ic >= 2) ? -kpair<NINE, -1, 0>(so, ci, iif, jjf, iic, jjc, ic, jc) : 0.0;
    soc[Pc * KS + pc] = (jc >= 2) ? -kpair<NINE, 0, -1>(so, ci, iif, jjf, iic, jjc, ic, jc) : 0.0;
    const bool corner = ic >= 2 && jc >= 2;
    soc[Pc * KSW + pc] = corner ? -kpair<NINE, -1, -1>(so, ci, iif, jjf, iic, jjc, ic, jc) : 0.0;
    soc[Pc * KNW + pc] = corner ? -kpair<NINE, 1, -1>(so, ci, iif, jjf, iic, jjc, ic - 1, jc) : 0.0;
}

// ------------------------------------------------------------------------------------------
// coarsest grid: banded Cholesky (DPBTF2 upper) and solve (DPBTRS = two DTBSV), FP64, one
// workgroup.  Band layout ABD(ii, n), diagonal in row kd+1 = ii, kd = ii-1
// (src/2d/ftn/BMG2_SymStd_SETUP_cg_LU.f90:92-118, src/2d/ftn/BMG2_SymStd_SOLVE_cg.f90:93-123).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void setup_cg_k(const double *__restrict__ so, int nstncl,
                                                  double *__restrict__ abd, int ii, int jj,
                                                  int *__restrict__ info) {
    const int nx = ii - 2, ny = jj - 2, n = nx * ny, kd = ii - 1, ld = ii;
    const std::size_t P = (std::size_t)ii * jj;
    so += (std::size_t)blockIdx.x * nstncl * P;  // one workgroup per problem of a batch
    abd += (std::size_t)blockIdx.x * ld * n;
    auto AB = [&](int r, int c) -> double & { return abd[(std::size_t)(c - 1) * ld + (r - 1)]; };
    for (int kk = threadIdx.x; kk < n; kk += blockDim.x) {
        const int i = kk % nx + 1, j = kk / nx + 1;
        const std::size_t p = pidx(i, j, ii);
        const int c = kk + 1;
        for (int r = 1; r <= ld; ++r) AB(r, c) = 0.0;
        AB(ii, c) = so[p];
        AB(ii - 1, c) = -so[P * KW + p];
        AB(3, c) = (nstncl == 5) ? -so[P * KNW + p + 1] : 0.0;
        AB(2, c) = -so[P * KS + p];
        AB(1, c) = (nstncl == 5) ? -so[P * KSW + p] : 0.0;
    }
    __syncthreads();
    __shared__ int bad;
    band_factor_upper(abd, n, kd, ld, bad);
    if (threadIdx.x == 0 && bad) atomicMax(info, bad);
}

// The coarsest level's band solve: one thread runs the two triangular solves, each step of which
// reads what the previous one wrote.  Where the factor (ld * n) and the iterate (n) fit in
// CG_LDS doubles, the workgroup first stages both into LDS, so that those steps wait on LDS and
// not on global-memory round trips (4x4 coarsest grid of the 4095^2 hierarchy: 25 us per solve
// through memory); otherwise the iterate lives in the global scratch x.  Same operations either
// way.  tid / nt: this thread and the workgroup's thread count; sh: the LDS staging area.
constexpr int CG_LDS = 1024;
template <typename T>
__device__ inline void coarse_band_solve(T *q, const T *qf, const double *abd, double *x, int ii,
                                         int jj, int tid, int nt, double *sh) {
    const int nx = ii - 2, ny = jj - 2, n = nx * ny, kd = ii - 1, ld = ii;
    const bool lds = (std::size_t)n * (ld + 1) <= (std::size_t)CG_LDS;
    double *xx = lds ? sh : x;
    const double *ab = abd;
    if (lds) {
        for (int k = tid; k < ld * n; k += nt) sh[n + k] = abd[k];
        ab = sh + n;
    }
    for (int kk = tid; kk < n; kk += nt) xx[kk] = (double)qf[pidx(kk % nx + 1, kk / nx + 1, ii)];
    __syncthreads();
    if (lds && kd < WAVE && nt >= WAVE) {  // bitwise the same sweeps, one wavefront
        if (tid < WAVE) band_solve_upper_wave(ab, n, kd, ld, xx, tid);
    } else if (tid == 0) {
        band_solve_upper(ab, n, kd, ld, xx);
    }
    __syncthreads();
    for (int kk = tid; kk < n; kk += nt) q[pidx(kk % nx + 1, kk / nx + 1, ii)] = (T)xx[kk];
}

template <typename T>
__global__ __launch_bounds__(64) void solve_cg_k(T *__restrict__ q, const T *__restrict__ qf,
                                                 const double *__restrict__ abd,
                                                 double *__restrict__ x, int ii, int jj, int rep) {
    __shared__ double sh[CG_LDS];
    const int n = (ii - 2) * (jj - 2);
    q += (std::size_t)blockIdx.x * ii * jj;  // one workgroup per problem of a batch
    qf += (std::size_t)blockIdx.x * ii * jj;
    abd += (std::size_t)(blockIdx.x / rep) * ii * n;
    x += (std::size_t)blockIdx.x * n;
    coarse_band_solve<T>(q, qf, abd, x, ii, jj, threadIdx.x, blockDim.x, sh);
}

// ------------------------------------------------------------------------------------------
// elementwise helpers
// ------------------------------------------------------------------------------------------
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void convert_k(const Tin *__restrict__ in, Tout *__restrict__ out,
                                                 std::size_t n) {
    for (std::size_t k = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (std::size_t)gridDim.x * blockDim.x)
        out[k] = (Tout)in[k];
}

template <typename T>
__global__ __launch_bounds__(256) void fill_k(T *__restrict__ x, T v, std::size_t n) {
    for (std::size_t k = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (std::size_t)gridDim.x * blockDim.x)
        x[k] = v;
}

// fill of a 16-byte aligned array: 16-byte stores (a quarter of the store instructions of
// fill_k; the zero start of the FP32 correction at 4095^2: 11.8 against 18.5 us), the tail in
// scalars
template <typename T>
__global__ __launch_bounds__(256) void fill16_k(T *__restrict__ x, T v, std::size_t n) {
    constexpr int W = 16 / sizeof(T);
    struct __attribute__((aligned(16))) vec {
        T e[W];
    };
    vec w;
#pragma unroll
    for (int u = 0; u < W; ++u) w.e[u] = v;
    vec *xv = reinterpret_cast<vec *>(x);
    const std::size_t nv = n / W;
    for (std::size_t k = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nv;
         k += (std::size_t)gridDim.x * blockDim.x)
        xv[k] = w;
    const std::size_t t = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n - nv * W) x[nv * W + t] = v;
}

__global__ __launch_bounds__(256) void add_correction_k(double *__restrict__ x, const float *__restrict__ e,
                                                        int ii, int jj) {
    const int i = blockIdx.x * 64 + (threadIdx.x % 64) + 1;
    const int j = blockIdx.y * 4 + (threadIdx.x / 64) + 1;
    if (i > ii - 2 || j > jj - 2) return;
    const std::size_t p = pidx(i, j, ii) + (std::size_t)blockIdx.z * ii * jj;
    x[p] = x[p] + (double)e[p];
}

// x[k] = xprev[k] for the problems of a batch with active[b] == 0 (whole arrays, ghosts included)
__global__ __launch_bounds__(256) void keep_inactive_k(double *__restrict__ x,
                                                       const double *__restrict__ xprev,
                                                       const int *__restrict__ active,
                                                       std::size_t per, std::size_t n) {
    for (std::size_t k = (std::size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (std::size_t)gridDim.x * blockDim.x)
        if (active[k / per] == 0) x[k] = xprev[k];
}

inline unsigned grid1(std::size_t n) {
    std::size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

}  // namespace

// ==========================================================================================
// launchers
// ==========================================================================================
std::size_t norm_partials(int ii, int jj) {
    const dim3 g = grid2(ii - 2, jj - 2, 64, 4);
    return (std::size_t)g.x * g.y;
}

// tiled grid over an nx x ny index range, one z-slice per problem of a batch
static dim3 gridb(int nx, int ny, int nb) {
    dim3 g = grid2(nx, ny, 64, 4);
    g.z = (unsigned)nb;
    return g;
}

template <typename T>
void residual(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj,
              hipStream_t s, int nb, int rep) {
    const dim3 g = gridb(ii - 2, jj - 2, nb);
    if (nstncl == 5)
        residual_k<T, true, false><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, nullptr, rep);
    else
        residual_k<T, false, false><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, nullptr, rep);
}

template <typename T>
void residual_norm(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj,
                   double *partials, double *sumsq_out, hipStream_t s, int nb, int rep) {
    const dim3 g = gridb(ii - 2, jj - 2, nb);
    if (nstncl == 5)
        residual_k<T, true, true><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, partials, rep);
    else
        residual_k<T, false, true><<<g, 256, 0, s>>>(so, qf, q, res, ii, jj, partials, rep);
    finalize_sum_k<<<nb, 1024, 0, s>>>(partials, (std::size_t)g.x * g.y, sumsq_out);
}

template <typename T>
void sumsq(const T *x, int ii, int jj, double *partials, double *out, hipStream_t s, int nb) {
    const dim3 g = gridb(ii - 2, jj - 2, nb);
    sumsq_k<T><<<g, 256, 0, s>>>(x, ii, jj, partials);
    finalize_sum_k<<<nb, 1024, 0, s>>>(partials, (std::size_t)g.x * g.y, out);
}

void residual_ir(const double *so, int nstncl, const double *qf, const double *x, const float *e,
                 double *xnew, float *r32, int ii, int jj, double *partials, double *sumsq_out,
                 hipStream_t s, int nb, const int *active, int rep, float *r32T) {
    // one workgroup per tile of an operator; nb = operators * rep
    const dim3 g((unsigned)((ii - 2 + 63) / 64), (unsigned)((jj - 2 + IR_ROWS - 1) / IR_ROWS),
                 (unsigned)(nb / rep));
    const unsigned nt = 64 * IR_ROWS;
    if (nstncl == 5) {
        if (e)
            residual_ir_k<true, true><<<g, nt, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, partials,
                                                       active, rep, r32T);
        else
            residual_ir_k<true, false><<<g, nt, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, partials,
                                                        active, rep, r32T);
    } else {
        if (e)
            residual_ir_k<false, true><<<g, nt, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, partials,
                                                        active, rep, r32T);
        else
            residual_ir_k<false, false><<<g, nt, 0, s>>>(so, qf, x, e, xnew, r32, ii, jj, partials,
                                                         active, rep, r32T);
    }
    finalize_sum_k<<<nb, 1024, 0, s>>>(partials, (std::size_t)g.x * g.y, sumsq_out);
}

void setup_recip(const double *so, int nstncl, double *sor, int ii, int jj, hipStream_t s,
                 int nb) {
    recip_k<<<gridb(ii - 2, jj - 2, nb), 256, 0, s>>>(so, sor, ii, jj, nstncl);
}

template <typename T>
void relax_gs(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
              int updown, hipStream_t s, int nb, int rep) {
    if (nstncl == 5) {
        // DOWN: (i odd, j odd), (even, odd), (odd, even), (even, even); UP: reverse
        static const int order[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
        const dim3 g = gridb(ii / 2 + 1, jj / 2 + 1, nb);
        for (int t = 0; t < 4; ++t) {
            const int c = updown == 0 ? t : 3 - t;
            gs_colour_k<T, true><<<g, 256, 0, s>>>(so, qf, q, sor, ii, jj, order[c][0], order[c][1],
                                                   rep, t & 1);
        }
    } else {
        // DOWN: (i + j) even first; UP: odd first
        const dim3 g = gridb(ii / 2 + 1, jj - 2, nb);
        for (int t = 0; t < 2; ++t) {
            const int parity = updown == 0 ? t : 1 - t;
            gs_colour_k<T, false><<<g, 256, 0, s>>>(so, qf, q, sor, ii, jj, parity, 0, rep, t & 1);
        }
    }
}

void setup_lines_x(const double *so, int nstncl, double *sor, int ii, int jj, int *info,
                   hipStream_t s, int nb) {
    const int n = jj - 2;
    setup_lines_x_k<<<dim3((n + 63) / 64, nb), 64, 0, s>>>(so, sor, ii, jj, nstncl, info);
}

void setup_lines_y(const double *so, int nstncl, double *sor, int ii, int jj, int *info,
                   hipStream_t s, int nb) {
    const int n = ii - 2;
    setup_lines_y_k<<<dim3((n + 63) / 64, nb), 64, 0, s>>>(so, sor, ii, jj, nstncl, info);
}

// dst (cols x rows) = transpose of src (rows x cols), row-major, for each problem of a batch;
// 64 x 64 tiles through LDS, 512 threads (8 rows per pass)
template <typename T>
__global__ __launch_bounds__(512) void transpose_k(const T *__restrict__ src, T *__restrict__ dst,
                                                  int rows, int cols) {
    __shared__ T tile[64][65];
    const std::size_t off = (std::size_t)blockIdx.z * rows * cols;
    src += off;
    dst += off;
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int k = ty; k < 64; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[(std::size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (c < cols && r < rows) dst[(std::size_t)c * rows + r] = tile[tx][k];
    }
}

// The operator of the transposed grid: plane k of problem b of soT is the transpose of plane
// from[k] of problem b of so (c <- c, w <- s, s <- w, sw <- sw, nw <- nw); z = b * nstncl + k.
template <typename T>
__global__ __launch_bounds__(512) void transpose_op_k(const T *__restrict__ so, T *__restrict__ soT,
                                                     int rows, int cols, int nstncl) {
    __shared__ T tile[64][65];
    const int k = (int)blockIdx.z % nstncl, b = (int)blockIdx.z / nstncl;
    const int from = k == 0 ? KO : k == 1 ? KS : k == 2 ? KW : k;
    const std::size_t P = (std::size_t)rows * cols;
    const T *src = so + ((std::size_t)b * nstncl + from) * P;
    T *dst = soT + ((std::size_t)b * nstncl + k) * P;
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int m = ty; m < 64; m += 8) {
        const int r = r0 + m, c = c0 + tx;
        if (r < rows && c < cols) tile[m][tx] = src[(std::size_t)r * cols + c];
    }
    __syncthreads();
    for (int m = ty; m < 64; m += 8) {
        const int c = c0 + m, r = r0 + tx;
        if (c < cols && r < rows) dst[(std::size_t)c * rows + r] = tile[tx][m];
    }
}

// FP32: 64 x 128 tiles moved as pairs of floats (see float_pair), twice the bytes per instruction
// of the one-float form (5.1 against 4.6 TB/s on a 4097^2 grid).
__global__ __launch_bounds__(512) void transpose_pairs_k(const float *__restrict__ src,
                                                         float *__restrict__ dst, int rows,
                                                         int cols) {
    __shared__ float tile[64][129];
    const std::size_t off = (std::size_t)blockIdx.z * rows * cols;
    src += off;
    dst += off;
    const int c0 = blockIdx.x * 128, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int c = c0 + 2 * tx;  // this lane's pair of source columns
    // all 8 loads of a thread in flight before the first LDS store
    float_pair v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = ty + 8 * u, r = r0 + k;
        const float *p = src + (std::size_t)r * cols + c;
        v[u] = float_pair{0.f, 0.f};
        if (r < rows) {
            if (c + 1 < cols)
                __builtin_memcpy(&v[u], p, sizeof(float_pair));
            else if (c < cols)
                v[u].a = p[0];
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int k = ty + 8 * u;
        if (r0 + k < rows && c < cols) {
            tile[k][2 * tx] = v[u].a;
            if (c + 1 < cols) tile[k][2 * tx + 1] = v[u].b;
        }
    }
    __syncthreads();
    // a half-wavefront writes 64 consecutive destination values (one source column) as 32 pairs
    const int h = tx % 32, r = r0 + 2 * h;
    for (int k = 2 * ty + tx / 32; k < 128; k += 16) {
        const int cc = c0 + k;
        if (cc >= cols) continue;
        float *p = dst + (std::size_t)cc * rows + r;
        if (r + 1 < rows) {
            const float_pair v{tile[2 * h][k], tile[2 * h + 1][k]};
            __builtin_memcpy(p, &v, sizeof(v));
        } else if (r < rows) {
            p[0] = tile[2 * h][k];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Kernels that take the iterate in the transposed layout (qT: ii rows of jj values,
// qT[i * jj + j] = q(i, j)), for levels whose y sweeps run as x sweeps of the transposed grid:
// after the last y sweep of the pre-smoothing the iterate stays transposed through the residual
// and interp_add, and the first y sweep of the post-smoothing starts from it, so the two grid
// transposes between them are not needed.  Every point is evaluated with the same operations in
// the same order as residual_k / interp_add_k, so results are bitwise identical.
// ------------------------------------------------------------------------------------------
constexpr int kTI = 64, kTJ = 32;  // residual_qT_k tile: 64 (i) x 32 (j) points, 256 threads

template <typename T, bool NINE>
__global__ __launch_bounds__(256) void residual_qT_k(const T *__restrict__ so,
                                                     const T *__restrict__ qf,
                                                     const T *__restrict__ qT,
                                                     T *__restrict__ res, int ii, int jj, int rep) {
    __shared__ T sq[kTJ + 2][kTI + 3];  // q(i0-1 .. i0+64, j0-1 .. j0+32) as [j][i]
    const wg_pos g = tiled_wg(rep);
    const int i0 = g.x * kTI + 1, j0 = g.y * kTJ + 1;
    const std::size_t P = (std::size_t)ii * jj, bz = g.z;
    so += (bz / rep) * (NINE ? 5 : 3) * P;
    qf += bz * P;
    qT += bz * P;
    res += bz * P;
    // stage: consecutive threads read consecutive j of one row of qT (a column of q); all loads
    // of a thread in flight before its LDS stores
    constexpr int NS = (kTI + 2) * (kTJ + 2), NU = (NS + 255) / 256;
    T v[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int e = threadIdx.x + 256 * u;
        const int a = e / (kTJ + 2), b = e % (kTJ + 2);
        const int i = i0 - 1 + a, j = j0 - 1 + b;
        v[u] = (e < NS && i < ii && j < jj) ? qT[(std::size_t)i * jj + j] : T(0);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int e = threadIdx.x + 256 * u;
        if (e < NS) sq[e % (kTJ + 2)][e / (kTJ + 2)] = v[u];
    }
    __syncthreads();
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int i = i0 + tx;
    if (i > ii - 2) return;
#pragma unroll
    for (int u = 0; u < kTJ / 4; ++u) {
        const int k = ty + 4 * u;
        const int j = j0 + k;
        if (j > jj - 2) break;
        const std::size_t p = pidx(i, j, ii);
        auto Q = [&](int di, int dj) { return sq[k + 1 + dj][tx + 1 + di]; };
        T acc = qf[p];
        acc = acc + so[P * KW + p] * Q(-1, 0);
        acc = acc + so[P * KW + p + 1] * Q(1, 0);
        acc = acc + so[P * KS + p] * Q(0, -1);
        acc = acc + so[P * KS + p + ii] * Q(0, 1);
        if (NINE) {
            acc = acc + so[P * KSW + p] * Q(-1, -1);
            acc = acc + so[P * KNW + p + 1] * Q(1, -1);
            acc = acc + so[P * KNW + p + ii] * Q(-1, 1);
            acc = acc + so[P * KSW + p + ii + 1] * Q(1, 1);
        }
        res[p] = acc - so[p] * Q(0, 0);
    }
}

// residual_qT_k for the nr (2..rq_maxr) right-hand sides of one operator in one workgroup
// (blockIdx.z: the operator): the iterate tiles of all of them are staged (36 KB of LDS), then
// each point's couplings are loaded once and applied to every right-hand side, in
// residual_qT_k's order.
template <typename T>
constexpr int rq_maxr = sizeof(T) == 8 ? 2 : 4;
template <typename T, bool NINE>
__global__ __launch_bounds__(256) void residual_qT_rhs_k(const T *__restrict__ so,
                                                         const T *__restrict__ qf,
                                                         const T *__restrict__ qT,
                                                         T *__restrict__ res, int ii, int jj,
                                                         int nr) {
    __shared__ T sq[rq_maxr<T>][kTJ + 2][kTI + 3];  // per right-hand side, as in residual_qT_k
    const int i0 = blockIdx.x * kTI + 1, j0 = blockIdx.y * kTJ + 1;
    const std::size_t P = (std::size_t)ii * jj, b0 = (std::size_t)blockIdx.z * nr;
    so += (std::size_t)blockIdx.z * (NINE ? 5 : 3) * P;
    for (int rh = 0; rh < nr; ++rh) {
        const T *q = qT + (b0 + rh) * P;
        for (int e = threadIdx.x; e < (kTI + 2) * (kTJ + 2); e += 256) {
            const int a = e / (kTJ + 2), b = e % (kTJ + 2);
            const int i = i0 - 1 + a, j = j0 - 1 + b;
            sq[rh][b][a] = (i < ii && j < jj) ? q[(std::size_t)i * jj + j] : T(0);
        }
    }
    __syncthreads();
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int i = i0 + tx;
    if (i > ii - 2) return;
    for (int u = 0; u < kTJ / 4; ++u) {
        const int k = ty + 4 * u;
        const int j = j0 + k;
        if (j > jj - 2) break;
        const std::size_t p = pidx(i, j, ii);
        const T cw0 = so[P * KW + p], cw1 = so[P * KW + p + 1];
        const T cs0 = so[P * KS + p], cs1 = so[P * KS + p + ii];
        T csw0 = T(0), cnw0 = T(0), cnw1 = T(0), csw1 = T(0);
        if (NINE) {
            csw0 = so[P * KSW + p];
            cnw0 = so[P * KNW + p + 1];
            cnw1 = so[P * KNW + p + ii];
            csw1 = so[P * KSW + p + ii + 1];
        }
        const T co = so[p];
        for (int rh = 0; rh < nr; ++rh) {
            auto Q = [&](int di, int dj) { return sq[rh][k + 1 + dj][tx + 1 + di]; };
            const std::size_t pb = (b0 + rh) * P + p;
            T acc = qf[pb];
            acc = acc + cw0 * Q(-1, 0);
            acc = acc + cw1 * Q(1, 0);
            acc = acc + cs0 * Q(0, -1);
            acc = acc + cs1 * Q(0, 1);
            if (NINE) {
                acc = acc + csw0 * Q(-1, -1);
                acc = acc + cnw0 * Q(1, -1);
                acc = acc + cnw1 * Q(-1, 1);
                acc = acc + csw1 * Q(1, 1);
            }
            res[pb] = acc - co * Q(0, 0);
        }
    }
}

// interp_add with the fine iterate transposed: phase 1 forms, per fine point of a 64 x 32 tile
// (32 x 16 blocks of 2 x 2 points, one thread each), the one (coarse point) or two (others) terms
// that interp_add_k adds, in its order; phase 2 adds them to qT with j-contiguous accesses.
template <typename T>
__global__ __launch_bounds__(512) void interp_add_qT_k(T *__restrict__ qT, const T *__restrict__ qc,
                                                       const T *__restrict__ res,
                                                       const T *__restrict__ so,
                                                       const T *__restrict__ ci, int iic, int jjc,
                                                       int iif, int jjf, int ns, int rep) {
    // one workgroup per tile of one operator (blockIdx.z) does the rep right-hand sides of that
    // operator (problems z * rep ..) one after another, with the block's interpolation weights
    // and operator diagonal loaded once into registers
    __shared__ T t1[32][65], t2[32][65];  // [j - jt][i - it]
    const int m0 = blockIdx.x * 32 + 1, n0 = blockIdx.y * 16 + 1;
    const int it = 2 * m0 - 1, jt = 2 * n0 - 1;  // first fine point of the tile
    const std::size_t Pc = (std::size_t)iic * jjc, Pf = (std::size_t)iif * jjf;
    const std::size_t b0 = (std::size_t)blockIdx.z * rep;
    so += (std::size_t)blockIdx.z * ns * Pf;
    ci += (std::size_t)blockIdx.z * 8 * Pc;
    const int imax = iif - 2 + ((iif - 2) % 2 == 0 ? 1 : 0);
    const int jmax = jjf - 2 + ((jjf - 2) % 2 == 0 ? 1 : 0);
    const int m = m0 + threadIdx.x % 32, n = n0 + threadIdx.x / 32;
    const int i0 = 2 * m - 1, j0 = 2 * n - 1;
    const bool act = i0 <= imax && j0 <= jmax;
    const bool xe = act && i0 + 1 <= imax, ye = act && j0 + 1 <= jmax;
    const int li = i0 - it, lj = j0 - jt;
    auto CI = [&](int a, int b, int k) { return ci[Pc * k + pidx(a, b, iic)]; };
    // res / c at the block's F-points (the raw res beyond the interior, as interp_add_k)
    auto fpt = [&](int i, int j, T &c) {
        const bool raw = i > iif - 2 || j > jjf - 2;
        c = raw ? T(1) : so[pidx(i, j, iif)];
        return raw;
    };
    T cLR = T(0), cLL = T(0), cLA = T(0), cLB = T(0);
    T cSW = T(0), cNW = T(0), cNE = T(0), cSE = T(0);
    T c10 = T(1), c01 = T(1), c11 = T(1);
    bool r10 = true, r01 = true, r11 = true;
    if (xe) {
        cLR = CI(m + 1, n, LR);
        cLL = CI(m + 1, n, LL);
        r10 = fpt(i0 + 1, j0, c10);
    }
    if (ye) {
        cLA = CI(m, n + 1, LA);
        cLB = CI(m, n + 1, LB);
        r01 = fpt(i0, j0 + 1, c01);
    }
    if (xe && ye) {
        cSW = CI(m + 1, n + 1, LSW);
        cNW = CI(m + 1, n + 1, LNW);
        cNE = CI(m + 1, n + 1, LNE);
        cSE = CI(m + 1, n + 1, LSE);
        r11 = fpt(i0 + 1, j0 + 1, c11);
    }
    for (int rh = 0; rh < rep; ++rh) {
        const std::size_t bz = b0 + rh;
        const T *qcb = qc + bz * Pc;
        const T *rsb = res + bz * Pf;
        T *qTb = qT + bz * Pf;
        if (rh > 0) __syncthreads();  // phase 2 of the previous right-hand side is done
        // phase 1: the one (coarse point) or two (others) terms interp_add_k adds, in its order
        if (act) {
            auto QC = [&](int a, int b) { return qcb[pidx(a, b, iic)]; };
            auto R = [&](int i, int j, bool raw, T c) {
                const T v = rsb[pidx(i, j, iif)];
                return raw ? v : v / c;
            };
            const T q00 = QC(m, n);
            t1[lj][li] = q00;
            if (xe) {
                t1[lj][li + 1] = cLR * QC(m + 1, n) + cLL * q00;
                t2[lj][li + 1] = R(i0 + 1, j0, r10, c10);
            }
            if (ye) {
                t1[lj + 1][li] = cLA * QC(m, n + 1) + cLB * q00;
                t2[lj + 1][li] = R(i0, j0 + 1, r01, c01);
            }
            if (xe && ye) {
                T a = cSW * q00;
                a = a + cNW * QC(m, n + 1);
                a = a + cNE * QC(m + 1, n + 1);
                a = a + cSE * QC(m + 1, n);
                t1[lj + 1][li + 1] = a;
                t2[lj + 1][li + 1] = R(i0 + 1, j0 + 1, r11, c11);
            }
        }
        __syncthreads();
        // phase 2: thread e -> (i offset e / 32, j offset e % 32): j-contiguous runs of 32 in
        // qT.  The thread's 4 points are distinct, so all 4 loads are issued before the first
        // store
        constexpr int NP = 64 * 32 / 512;
        T v[NP];
        bool ok[NP];
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int e = threadIdx.x + u * 512, pi = e / 32, pj = e % 32;
            const int i = it + pi, j = jt + pj;
            ok[u] = i <= imax && j <= jmax;
            v[u] = ok[u] ? qTb[(std::size_t)i * jjf + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < NP; ++u) {
            const int e = threadIdx.x + u * 512, pi = e / 32, pj = e % 32;
            if (!ok[u]) continue;
            T w = v[u] + t1[pj][pi];
            if (!((pi % 2 == 0) && (pj % 2 == 0))) w = w + t2[pj][pi];
            qTb[(std::size_t)(it + pi) * jjf + (jt + pj)] = w;
        }
    }
}

template <typename T>
void transpose(const T *src, T *dst, int rows, int cols, hipStream_t s, int nb) {
    if constexpr (sizeof(T) == 4) {
        const dim3 g((unsigned)((cols + 127) / 128), (unsigned)((rows + 63) / 64), (unsigned)nb);
        transpose_pairs_k<<<g, 512, 0, s>>>(src, dst, rows, cols);
    } else {
        const dim3 g((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64), (unsigned)nb);
        transpose_k<T><<<g, 512, 0, s>>>(src, dst, rows, cols);
    }
}

// ---- fused coarse-level V-cycle ---------------------------------------------------------------
// One workgroup per problem runs the whole V-cycle below a level of at most 64*CH points per side
// (smoothing, residual, restriction, coarsest band solve, interpolation) with workgroup barriers
// between the steps instead of one launch per step.  Every point is computed by the same device
// functions and in the same order as in the grid kernels, so the result is bitwise that of the
// unfused cycle.  Levels of a fused cycle are always 9-point (Galerkin operators).

// one x-line (row j) or y-line (column i) of at most 64*CH points, by one wavefront, in place
template <typename T, int DIR, int CH>
__device__ inline void wave_line(const T *__restrict__ so, const T *__restrict__ qf, T *q,
                                 const T *__restrict__ sor, int ii, int jj, int fixed) {
    const int lane = threadIdx.x % WAVE, rl = WAVE - 1 - lane;
    const std::size_t P = (std::size_t)ii * jj;
    const int n = DIR == 0 ? ii - 2 : jj - 2;
    const std::size_t f0 = DIR == 0 ? (std::size_t)fixed * ii : (std::size_t)fixed * jj;
    const T *D = sor + f0 + 1;
    const T *E = sor + P + f0 + 2;
    auto at = [&](int k) { return DIR == 0 ? pidx(k + 1, fixed, ii) : pidx(fixed, k + 1, ii); };
    T a[CH], b[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int k = u * WAVE + lane, kc = min(k, n - 1);
        line_raw<T> r;
        line_load<T, true, DIR>(r, so, qf, q, P, ii, at(kc));
        b[u] = line_rhs<T, true>(r);
        a[u] = (k > 0) ? -E[max(kc - 1, 0)] : T(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) scan_affine_dpp(a[u], b[u]);
    T carry = T(0);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const T y = affine(a[u], carry, b[u]);
        carry = last_lane(y);
        b[u] = div_pivot(y, D[min(u * WAVE + lane, n - 1)]);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int k = u * WAVE + rl;
        const T c = __shfl(b[u], rl, WAVE);
        const T e = E[min(k, max(n - 2, 0))];
        a[u] = (k < n - 1) ? -e : T(0);
        a[u] = (k < n) ? a[u] : T(1);
        b[u] = (k < n) ? c : T(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) scan_affine_dpp(a[u], b[u]);
    carry = T(0);
#pragma unroll
    for (int u = CH - 1; u >= 0; --u) {
        const T x = affine(a[u], carry, b[u]);
        carry = last_lane(x);
        b[u] = x;
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const int k = u * WAVE + rl;
        if (k < n) q[at(k)] = b[u];
    }
}

// LDS bytes the fused cycle needs to hold the arrays of one level on chip (per level: the
// operator, both relaxation factors, x, b, res and, below the first level, the interpolation),
// each array rounded up to 16 bytes.  first_mode (the first level of lds_mode 2 or 3): x, b
// and res (2) or x alone (3).
template <typename T>
__host__ __device__ inline std::size_t fused_level_bytes(int ii, int jj, bool has_ci,
                                                        int first_mode = 1) {
    const std::size_t plane = ((std::size_t)ii * jj * sizeof(T) + 15) / 16 * 16;
    if (first_mode == 2) return 3 * plane;
    if (first_mode == 3) return plane;
    return plane * (5 + 2 + 2 + 3 + (has_ci ? 8 : 0));
}

template <typename T, int CH>
__global__ __launch_bounds__(1024) void coarse_cycle_k(const coarse_level<T> *__restrict__ lv,
                                                       int nlev, int relax, int npre, int npost,
                                                       const double *__restrict__ abd,
                                                       double *__restrict__ bbd, int rep,
                                                       int lds_mode) {
    constexpr int NT = 1024, NWAVE = NT / WAVE;
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, wave = tid / WAVE;
    const std::size_t bz = blockIdx.x, bo = bz / rep;
    auto level_global = [&](int l) {  // the level's arrays of problem bz (operator bo)
        coarse_level<T> d = lv[l];
        const std::size_t P = (std::size_t)d.ii * d.jj;
        d.so += bo * 5 * P;
        d.x += bz * P;
        d.b += bz * P;
        d.res += bz * P;
        d.sor0 += bo * 2 * P;
        d.sor1 += bo * 2 * P;
        if (d.ci) d.ci += bo * 8 * P;
        return d;
    };
    // lds_mode 1: every array of every level lives in LDS for the whole cycle (the launcher
    // checked that they fit); the first level's b comes in from memory and its x goes back at the
    // end, the only values the caller exchanges with the fused levels.  For a first level too
    // large to hold whole, lds_mode 2 keeps only its x, b and res in LDS and lds_mode 3 its x
    // alone; its other arrays stay in memory, where the sweeps only read them, so that a sweep no
    // longer waits on memory for the previous sweep's writes.
    const bool on_chip = lds_mode != 0;
    const int m0 = lds_mode >= 2 ? lds_mode : 1;  // fused_level_bytes mode of the first level
    auto level_lds = [&](int l) {
        std::size_t off = 0;
        for (int m = 0; m < l; ++m)
            off += fused_level_bytes<T>(lv[m].ii, lv[m].jj, m > 0, m == 0 ? m0 : 1);
        coarse_level<T> d = lv[l];
        const std::size_t plane = ((std::size_t)d.ii * d.jj * sizeof(T) + 15) / 16 * 16;
        unsigned char *base = lds + off;
        if (l == 0 && m0 >= 2) {
            d = level_global(0);
            d.x = reinterpret_cast<T *>(base);
            if (m0 == 2) {
                d.b = reinterpret_cast<T *>(base + plane);
                d.res = reinterpret_cast<T *>(base + 2 * plane);
            }
            return d;
        }
        d.so = reinterpret_cast<T *>(base);
        d.sor0 = reinterpret_cast<T *>(base + 5 * plane);
        d.sor1 = reinterpret_cast<T *>(base + 7 * plane);
        d.x = reinterpret_cast<T *>(base + 9 * plane);
        d.b = reinterpret_cast<T *>(base + 10 * plane);
        d.res = reinterpret_cast<T *>(base + 11 * plane);
        d.ci = l > 0 ? reinterpret_cast<T *>(base + 12 * plane) : nullptr;
        return d;
    };
    auto level = [&](int l) { return on_chip ? level_lds(l) : level_global(l); };
    if (on_chip) {
        for (int l = 0; l < nlev; ++l) {
            const coarse_level<T> g = level_global(l), d = level_lds(l);
            const std::size_t P = (std::size_t)g.ii * g.jj;
            if (l == 0 && m0 == 3) {  // the first level's x alone
                for (std::size_t k = tid; k < P; k += NT) d.x[k] = g.x[k];
                continue;
            }
            if (!(l == 0 && m0 == 2)) {
                for (std::size_t k = tid; k < 5 * P; k += NT) const_cast<T *>(d.so)[k] = g.so[k];
                for (std::size_t k = tid; k < 2 * P; k += NT) {
                    const_cast<T *>(d.sor0)[k] = g.sor0[k];
                    const_cast<T *>(d.sor1)[k] = g.sor1[k];
                }
                if (l > 0)
                    for (std::size_t k = tid; k < 8 * P; k += NT)
                        const_cast<T *>(d.ci)[k] = g.ci[k];
            }
            // res and, below the first level, b start at zero as in memory: their ghosts are
            // never written, and interp_add adds the ghost res of an even-sized level to the
            // ghost of x
            for (std::size_t k = tid; k < P; k += NT) {
                d.res[k] = T(0);
                d.b[k] = l == 0 ? g.b[k] : T(0);
                if (l == 0) d.x[k] = g.x[k];
            }
        }
        __syncthreads();
    }
    auto lines = [&](const coarse_level<T> &d, const T *sor, int dir, int updown) {
        const int nfix = dir == 0 ? d.jj - 2 : d.ii - 2;
        for (int t = 0; t < 2; ++t) {
            const int first = (updown == 0) ? (t == 0 ? 2 : 1) : (t == 0 ? 1 : 2);
            if (first <= nfix) {
                const int nlines = (nfix - first) / 2 + 1;
                for (int line = wave; line < nlines; line += NWAVE) {
                    if (dir == 0)
                        wave_line<T, 0, CH>(d.so, d.b, d.x, sor, d.ii, d.jj, first + 2 * line);
                    else
                        wave_line<T, 1, CH>(d.so, d.b, d.x, sor, d.ii, d.jj, first + 2 * line);
                }
            }
            __syncthreads();
        }
    };
    auto gauss_seidel = [&](const coarse_level<T> &d, int updown) {
        static constexpr int order[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
        const std::size_t P = (std::size_t)d.ii * d.jj;
        const int nx = d.ii - 2, npts = nx * (d.jj - 2);
        for (int t = 0; t < 4; ++t) {
            const int c = updown == 0 ? t : 3 - t;
            for (int k = tid; k < npts; k += NT) {
                const int i = k % nx + 1, j = k / nx + 1;
                if ((i & 1) != order[c][0] || (j & 1) != order[c][1]) continue;
                const std::size_t p = pidx(i, j, d.ii);
                const T v = apply_offdiag<T, true>(d.so, P, d.x, p, d.ii, d.b[p]);
                d.x[p] = v * d.sor0[P + p];
            }
            __syncthreads();
        }
    };
    auto smooth = [&](const coarse_level<T> &d, int updown) {
        const int nrep = updown == 0 ? npre : npost;
        for (int r = 0; r < nrep; ++r) {
            if (relax == CDR_RELAX_POINT) {
                gauss_seidel(d, updown);
            } else if (relax == CDR_RELAX_LINE_X) {
                lines(d, d.sor0, 0, updown);
            } else if (relax == CDR_RELAX_LINE_Y) {
                lines(d, d.sor0, 1, updown);
            } else if (updown == 0) {
                lines(d, d.sor0, 0, updown);
                lines(d, d.sor1, 1, updown);
            } else {
                lines(d, d.sor1, 1, updown);
                lines(d, d.sor0, 0, updown);
            }
        }
    };
    // down: smooth, residual, restrict to the next level's b, zero its x
    for (int l = 0; l + 1 < nlev; ++l) {
        const coarse_level<T> F = level(l), C = level(l + 1);
        smooth(F, 0);
        const std::size_t P = (std::size_t)F.ii * F.jj;
        const int nx = F.ii - 2, npts = nx * (F.jj - 2);
        for (int k = tid; k < npts; k += NT) {
            const std::size_t p = pidx(k % nx + 1, k / nx + 1, F.ii);
            T r = apply_offdiag<T, true>(F.so, P, F.x, p, F.ii, F.b[p]);
            r = r - F.so[p] * F.x[p];
            F.res[p] = r;
        }
        __syncthreads();
        const std::size_t Pc = (std::size_t)C.ii * C.jj;
        const int nxc = C.ii - 2, nc = nxc * (C.jj - 2);
        for (int k = tid; k < nc; k += NT) {
            const int ic = k % nxc + 1, jc = k / nxc + 1;
            C.b[pidx(ic, jc, C.ii)] = restrict_point<T>(C.ci, F.res, F.ii, C.ii, Pc, ic, jc);
        }
        for (std::size_t k = tid; k < Pc; k += NT) C.x[k] = T(0);
        __syncthreads();
    }
    // coarsest level: band Cholesky solve (as solve_cg_k)
    {
        __shared__ double cg_sh[CG_LDS];
        const coarse_level<T> C = level(nlev - 1);
        const int n = (C.ii - 2) * (C.jj - 2);
        coarse_band_solve<T>(C.x, C.b, abd + bo * (std::size_t)C.ii * n, bbd + bz * n, C.ii, C.jj,
                             tid, NT, cg_sh);
        __syncthreads();
    }
    // up: interpolate and add, smooth
    for (int l = nlev - 2; l >= 0; --l) {
        const coarse_level<T> F = level(l), C = level(l + 1);
        const int mb = F.ii / 2, nblk = mb * (F.jj / 2);
        for (int k = tid; k < nblk; k += NT)
            interp_block<T>(F.x, C.x, F.res, F.so, C.ci, C.ii, C.jj, F.ii, F.jj, k % mb + 1,
                            k / mb + 1);
        __syncthreads();
        smooth(F, 1);
    }
    if (on_chip) {  // the smoothing ends with a workgroup barrier
        const coarse_level<T> g = level_global(0), d = level_lds(0);
        const std::size_t P = (std::size_t)g.ii * g.jj;
        for (std::size_t k = tid; k < P; k += NT) g.x[k] = d.x[k];
    }
}

// ---- line kernel shapes and their autotuner ---------------------------------------------------
// A shape is (CPW, W, L) for x-lines (chunks per wavefront, wavefronts per line, lines per
// workgroup) and (CPW, W, NC, G) for y-lines (..., columns per workgroup, chunk rows per LDS
// round).  All shapes of a chunk class give bitwise identical results (see the segmented
// kernels), so the shape of each (precision, stencil, direction, extents, batch) is picked by
// timing: tune_lines() times every candidate once on the caller's data and records the fastest in
// a process-wide table that the launchers consult; without an entry, candidate 0 runs.

template <typename T, bool NINE, int CPW, int W, int L, bool TR>
static void launch_x(const T *so, const T *qf, T *q, const T *sor, int ii, int jj, int first,
                     int nlines, hipStream_t s, int nb, int rev, int rep) {
    const int ntiles = (nlines + L - 1) / L;
    const dim3 blocks((unsigned)(8 * ((ntiles + 7) / 8)), (unsigned)nb);
    // rev bit 1: zero start (x-lines of the grid itself only)
    if (!TR && (rev & 2))
        relax_xlines_seg_k<T, NINE, CPW, W, L, TR, !TR><<<blocks, L * W * 64, 0, s>>>(
            so, qf, q, sor, ii, jj, first, nlines, ntiles, rev & 1, rep);
    else
        relax_xlines_seg_k<T, NINE, CPW, W, L, TR, false><<<blocks, L * W * 64, 0, s>>>(
            so, qf, q, sor, ii, jj, first, nlines, ntiles, rev & 1, rep);
}

template <typename T, bool NINE, int CPW, int W, int NC, int G>
static void launch_y(const T *so, const T *qf, T *q, const T *sor, int ii, int jj, int first,
                     int nlines, hipStream_t s, int nb, int rev, int rep) {
    const int ntiles = (nlines + NC - 1) / NC;
    const dim3 blocks((unsigned)(8 * ((ntiles + 7) / 8)), (unsigned)nb);
    relax_ylines_seg_k<T, NINE, CPW, W, NC, G><<<blocks, NC * W * 64, 0, s>>>(
        so, qf, q, sor, ii, jj, first, nlines, ntiles, rev, rep);
}

// Launches candidate c of chunk class NCH (lines of at most 64*NCH points); returns the number
// of candidates of the class (c outside [0, count) launches nothing).
static std::string shape_text(const char *k, int cpw, int w, const char *tn, int t, int g) {
    char buf[96];
    if (g > 0)
        std::snprintf(buf, sizeof(buf), "%s<CPW=%d,W=%d,%s=%d,G=%d>", k, cpw, w, tn, t, g);
    else
        std::snprintf(buf, sizeof(buf), "%s<CPW=%d,W=%d,%s=%d>", k, cpw, w, tn, t);
    return buf;
}

// With `name` non-null nothing is launched and *name receives candidate c's description.
// DIR 0: x-lines, 1: y-lines, 2: x-lines of a transposed y-line problem.
template <typename T, bool NINE, int DIR, int NCH>
static int launch_shape(int c, const T *so, const T *qf, T *q, const T *sor, int ii, int jj,
                        int first, int nlines, hipStream_t s, int nb, std::string *name = nullptr,
                        int rev = 0, int rep = 1) {
    constexpr bool TR = DIR == 2;
    constexpr bool F64 = sizeof(T) == 8;
#define CDR_X(k, C, W, L)                                                                       \
    if (c == k) {                                                                               \
        if (name)                                                                               \
            *name = shape_text("relax_xlines_seg_k", C, W, "L", L, 0);                          \
        else                                                                                    \
            launch_x<T, NINE, C, W, L, TR>(so, qf, q, sor, ii, jj, first, nlines, s, nb, rev,   \
                                           rep);                                                \
    }
#define CDR_Y(k, C, W, NC, G)                                                                   \
    if (c == k) {                                                                               \
        if (name)                                                                               \
            *name = shape_text("relax_ylines_seg_k", C, W, "NC", NC, G);                        \
        else                                                                                    \
            launch_y<T, NINE, C, W, NC, G>(so, qf, q, sor, ii, jj, first, nlines, s, nb, rev,   \
                                           rep);                                                \
    }
    // x: (CPW, W, L); y: (CPW, W, NC, G).  At most 16 wavefronts per workgroup.
    if constexpr (DIR != 1) {
        if constexpr (NCH <= 2) {
            CDR_X(0, NCH, 1, 4);
            return 1;
        } else if constexpr (NCH == 4) {
            CDR_X(0, 4, 1, 4);
            CDR_X(1, 1, 4, 1);
            return 2;
        } else if constexpr (NCH == 8) {
            CDR_X(0, 8, 1, 4);
            CDR_X(1, 4, 2, 2);
            CDR_X(2, 2, 4, 1);
            CDR_X(3, 1, 8, 1);  // one chunk per wavefront: the shortest chain per line
            return 4;
        } else if constexpr (NCH == 16) {
            CDR_X(0, 16, 1, 4);
            CDR_X(1, 8, 2, 2);
            CDR_X(2, 4, 4, 1);
            CDR_X(3, 2, 8, 1);
            CDR_X(4, 1, 16, 1);
            return 5;
        } else if constexpr (NCH == 32) {
            CDR_X(0, 16, 2, 2);
            CDR_X(1, 8, 4, 1);
            CDR_X(2, 4, 8, 1);
            CDR_X(3, F64 ? 2 : 32, F64 ? 16 : 1, F64 ? 1 : 4);
            // two neighbouring lines of a colour per workgroup share the row of q between them
            CDR_X(4, 8, 4, 2);
            CDR_X(5, 4, 8, 2);
            return 6;
        } else if constexpr (NCH == 3) {  // classes of 2^k + 1 chunks: W * CPW >= NCH
            CDR_X(0, 3, 1, 4);
            CDR_X(1, 1, 3, 2);
            return 2;
        } else if constexpr (NCH == 5) {
            CDR_X(0, 5, 1, 4);
            CDR_X(1, 3, 2, 2);
            CDR_X(2, 1, 5, 2);
            return 3;
        } else if constexpr (NCH == 9) {
            CDR_X(0, 9, 1, 4);
            CDR_X(1, 5, 2, 2);
            CDR_X(2, 3, 3, 2);
            CDR_X(3, 1, 9, 1);
            return 4;
        } else if constexpr (NCH == 17) {
            CDR_X(0, 17, 1, 4);
            CDR_X(1, 9, 2, 2);
            CDR_X(2, 6, 3, 1);
            CDR_X(3, 6, 3, 2);
            CDR_X(4, 3, 6, 1);
            return 5;
        } else if constexpr (NCH == 33) {
            CDR_X(0, 11, 3, 1);
            CDR_X(1, 11, 3, 2);
            CDR_X(2, 6, 6, 1);
            CDR_X(3, 3, 11, 1);
            CDR_X(4, 17, 2, 2);
            return 5;
        } else {
            CDR_X(0, 16, 4, 1);
            CDR_X(1, 8, 8, 1);
            CDR_X(2, 4, 16, 1);
            CDR_X(3, 64, 1, 4);
            CDR_X(4, 8, 8, 2);
            CDR_X(5, 16, 4, 2);
            return 6;
        }
    } else {
        // columns per workgroup for 64-byte (N1) and 128-byte (2 N1) row segments
        constexpr int N1 = F64 ? 4 : 8;
        // small levels: fewer columns per workgroup put more workgroups on the GPU (their rows
        // come from the caches at these sizes)
        if constexpr (NCH <= 2) {
            CDR_Y(0, NCH, 1, N1, NCH);
            CDR_Y(1, NCH, 1, 2 * N1, NCH);
            CDR_Y(2, NCH, 1, N1 / 4, NCH);
            return 3;
        } else if constexpr (NCH == 4) {
            CDR_Y(0, 4, 1, N1, 4);
            CDR_Y(1, 4, 1, 2 * N1, 2);
            CDR_Y(2, 1, 4, N1 / 2, 1);
            CDR_Y(3, 4, 1, N1 / 4, 4);
            CDR_Y(4, 2, 2, N1 / 4, 2);
            // FP32: one column per workgroup as well (FP64 has it as N1 / 4): the levels of a few
            // hundred columns are latency-bound, where more, smaller workgroups fill more CUs
            if constexpr (!F64) {
                CDR_Y(5, 2, 2, 1, 2);
                CDR_Y(6, 1, 4, 1, 1);
                return 7;
            }
            return 5;
        } else if constexpr (NCH == 8) {
            CDR_Y(0, 8, 1, N1, 4);
            CDR_Y(1, 8, 1, 2 * N1, 2);
            CDR_Y(2, 4, 2, N1, 4);
            CDR_Y(3, 2, 4, N1 / 2, 2);
            CDR_Y(4, 8, 1, N1 / 4, 4);
            CDR_Y(5, 4, 2, N1 / 4, 4);
            if constexpr (!F64) {
                CDR_Y(6, 4, 2, 1, 4);
                CDR_Y(7, 2, 4, 1, 2);
                return 8;
            }
            return 6;
        } else if constexpr (NCH == 16) {
            CDR_Y(0, 16, 1, N1, 4);
            CDR_Y(1, 8, 2, N1, 4);
            CDR_Y(2, 16, 1, 2 * N1, 2);
            CDR_Y(3, 4, 4, N1 / 2, 4);
            if constexpr (!F64) {
                CDR_Y(4, 4, 4, 1, 4);
                CDR_Y(5, 8, 2, 1, 4);
                return 6;
            }
            return 4;
        } else if constexpr (NCH == 32) {
            CDR_Y(0, 16, 2, N1, 2);
            CDR_Y(1, 32, 1, N1, 4);
            CDR_Y(2, 8, 4, N1 / 2, 2);
            CDR_Y(3, 4, 8, N1 / 4, 4);
            return 4;
        } else {  // 16-wavefront workgroups get 128 VGPRs: narrower tiles for long segments
            CDR_Y(0, 64, 1, N1, 2);
            CDR_Y(1, 16, 4, N1 / 2, 2);
            CDR_Y(2, 32, 2, N1 / 2, 2);
            CDR_Y(3, 64, 1, 2 * N1, 2);
            return 4;
        }
    }
#undef CDR_X
#undef CDR_Y
}

// chunk class of a line of n points: 1, 2, 4, ..., 64, or 0 beyond 4096 points.  With odd set
// (x-line shapes), lines of 3, 5, 9, 17 or 33 chunks get a class of their own: the lines of the
// 2^k + 1 grids (129 ... 2049 points) would otherwise leave almost half of the lanes of a
// power-of-two class idle.
static int chunk_class(int n, bool odd) {
    const int nch = (n + WAVE - 1) / WAVE;
    if (odd && (nch == 3 || nch == 5 || nch == 9 || nch == 17 || nch == 33)) return nch;
    int c = 1;
    while (c < nch && c < 64) c *= 2;
    return nch > 64 ? 0 : c;
}

template <typename T, bool NINE, int DIR>
static int launch_class(int c, const T *so, const T *qf, T *q, const T *sor, int ii, int jj,
                        int first, int nlines, hipStream_t s, int nb, std::string *name = nullptr,
                        int rev = 0, int rep = 1) {
    const int n = (DIR != 1) ? ii - 2 : jj - 2;
#define CDR_C(K)                                                                                  \
    case K:                                                                                       \
        return launch_shape<T, NINE, DIR, K>(c, so, qf, q, sor, ii, jj, first, nlines, s, nb, name, \
                                             rev, rep)
#define CDR_O(K)                                                                                  \
    case K:                                                                                       \
        if constexpr (DIR != 1)                                                                   \
            return launch_shape<T, NINE, DIR, K>(c, so, qf, q, sor, ii, jj, first, nlines, s, nb,  \
                                                 name, rev, rep);                                 \
        else                                                                                      \
            return 0;  /* not reached: y-lines use the power-of-two classes */
    switch (chunk_class(n, DIR != 1)) {
        CDR_C(1);
        CDR_C(2);
        CDR_O(3);
        CDR_C(4);
        CDR_O(5);
        CDR_C(8);
        CDR_O(9);
        CDR_C(16);
        CDR_O(17);
        CDR_C(32);
        CDR_O(33);
        CDR_C(64);
    default:  // lines longer than 4096: stage the forward results in q itself
        if constexpr (DIR == 2) {
            return 0;  // no transposed path for such lines
        } else {
            if (c == 0 && name) *name = "relax_lines_k";
            if (c == 0 && !name)
                relax_lines_k<T, NINE, DIR><<<dim3((unsigned)((nlines + 3) / 4), (unsigned)nb), 256,
                                              0, s>>>(so, qf, q, sor, ii, jj, first, nlines, rep);
            return 1;
        }
    }
#undef CDR_C
#undef CDR_O
}

namespace {
using tune_key = std::array<int, 6>;  // direction, word size, 9-point, ii, jj, batch
std::mutex g_tune_mu;
std::map<tune_key, int> g_tune;
std::atomic<int> g_autotune{1};
}  // namespace

template <typename T>
static tune_key key_of(int dir, int nstncl, int ii, int jj, int nb) {
    return tune_key{dir, (int)sizeof(T), nstncl == 5 ? 1 : 0, ii, jj, nb};
}

static int tuned_shape(const tune_key &k) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    const auto it = g_tune.find(k);
    return it == g_tune.end() ? 0 : it->second;
}

template <typename T, int DIR>
static int launch_colour(int c, const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii,
                         int jj, int first, int nlines, hipStream_t s, int nb,
                         std::string *name = nullptr, int rev = 0, int rep = 1) {
    if (nstncl == 5)
        return launch_class<T, true, DIR>(c, so, qf, q, sor, ii, jj, first, nlines, s, nb, name, rev,
                                          rep);
    return launch_class<T, false, DIR>(c, so, qf, q, sor, ii, jj, first, nlines, s, nb, name, rev,
                                       rep);
}

static std::string shape_name(const std::array<int, 6> &k, int c) {
    std::string name;
    const int ns = k[2] ? 5 : 3;
    // direction 2 (y-lines through the transposed grid): an x-line shape on the swapped extents
    if (k[0] == 2 && c < 0) return "direct y-lines";
    const int dir = k[0] == 1 ? 1 : 0;
    const int ii = k[0] == 2 ? k[4] : k[3], jj = k[0] == 2 ? k[3] : k[4];
    if (k[1] == 8)
        (void)(dir == 0 ? launch_colour<double, 0>(c, nullptr, ns, nullptr, nullptr, nullptr, ii,
                                                   jj, 1, 1, nullptr, 1, &name)
                        : launch_colour<double, 1>(c, nullptr, ns, nullptr, nullptr, nullptr, ii,
                                                   jj, 1, 1, nullptr, 1, &name));
    else
        (void)(dir == 0 ? launch_colour<float, 0>(c, nullptr, ns, nullptr, nullptr, nullptr, ii,
                                                  jj, 1, 1, nullptr, 1, &name)
                        : launch_colour<float, 1>(c, nullptr, ns, nullptr, nullptr, nullptr, ii,
                                                  jj, 1, 1, nullptr, 1, &name));
    return k[0] == 2 ? "transposed, " + name : name;
}

// zq (x-lines of at most 4096 points only): the sweep starts from q = 0 (see relax_xlines_seg_k)
template <typename T, int DIR>
static void relax_lines(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                        int updown, hipStream_t s, int nb, int rep = 1, int zq = 0) {
    const int c = tuned_shape(key_of<T>(DIR, nstncl, ii, jj, nb));
    // lines of one colour: DOWN relaxes the lines at even array index (odd Cedar index) first,
    // UP the odd ones first (src/2d/ftn/BMG2_SymStd_relax_lines_x.f90:82-97)
    const int nfix = (DIR == 0) ? jj - 2 : ii - 2;
    for (int t = 0; t < 2; ++t) {
        const int first = (updown == 0) ? (t == 0 ? 2 : 1) : (t == 0 ? 1 : 2);
        if (first > nfix) continue;
        const int nlines = (nfix - first) / 2 + 1;
        // the second colour walks the lines in the opposite order (see xcd_tile)
        const int z = (DIR == 0 && zq && t == 0) ? 2 : 0;
        launch_colour<T, DIR>(c, so, nstncl, qf, q, sor, ii, jj, first, nlines, s, nb, nullptr,
                              t | z, rep);
    }
}

template <typename T>
void relax_lines_x(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s, int nb, int rep, int zero_start) {
    relax_lines<T, 0>(so, nstncl, qf, q, sor, ii, jj, updown, s, nb, rep,
                      zero_start && chunk_class(ii - 2, true) != 0);
}

bool lines_zero_start(int ii) { return chunk_class(ii - 2, true) != 0; }

template <typename T>
void relax_lines_y(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s, int nb, int rep) {
    relax_lines<T, 1>(so, nstncl, qf, q, sor, ii, jj, updown, s, nb, rep);
}

int set_autotune(int on) { return g_autotune.exchange(on ? 1 : 0); }
bool autotune_enabled() { return g_autotune.load() != 0; }

int line_shape_count(int dir, int wbytes, int nstncl, int ii, int jj) {
    const hipStream_t s = nullptr;  // counting launches nothing
    if (dir == 2) {  // x-line shapes of the transposed grid; none for lines over 4096 points
        if (jj - 2 > 64 * WAVE) return 0;
        dir = 0;
        std::swap(ii, jj);
    }
    if (wbytes == 8)
        return dir == 0 ? launch_colour<double, 0>(-1, nullptr, nstncl, nullptr, nullptr, nullptr,
                                                   ii, jj, 1, 1, s, 1)
                        : launch_colour<double, 1>(-1, nullptr, nstncl, nullptr, nullptr, nullptr,
                                                   ii, jj, 1, 1, s, 1);
    return dir == 0 ? launch_colour<float, 0>(-1, nullptr, nstncl, nullptr, nullptr, nullptr, ii,
                                              jj, 1, 1, s, 1)
                    : launch_colour<float, 1>(-1, nullptr, nstncl, nullptr, nullptr, nullptr, ii,
                                              jj, 1, 1, s, 1);
}

void set_line_shape(int dir, int wbytes, int nstncl, int ii, int jj, int nb, int shape) {
    const tune_key k{dir, wbytes, nstncl == 5 ? 1 : 0, ii, jj, nb};
    std::lock_guard<std::mutex> g(g_tune_mu);
    if (shape < 0)
        g_tune.erase(k);
    else
        g_tune[k] = shape;
}

template <typename T>
int tune_lines(int dir, const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
               hipStream_t s, int nb, int rep) {
    const tune_key k = key_of<T>(dir, nstncl, ii, jj, nb);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        if (g_tune.count(k)) return g_tune[k];
    }
    auto run = [&](int c) {
        return dir == 0 ? launch_colour<T, 0>(c, so, nstncl, qf, q, sor, ii, jj, 1,
                                              (jj - 2 + 1) / 2, s, nb, nullptr, 0, rep)
                        : launch_colour<T, 1>(c, so, nstncl, qf, q, sor, ii, jj, 1,
                                              (ii - 2 + 1) / 2, s, nb, nullptr, 0, rep);
    };
    const int ncand = run(-1);  // count only
    if (!g_autotune.load() || ncand < 2) return 0;
    int best = 0;
    {
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess) return -1;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return -1;
        }
        // two rounds over the candidates, each candidate's time the smaller of its two: shapes
        // within a few percent of each other were otherwise picked by the noise of one timing
        std::vector<float> tmin(ncand, -1.f);
        for (int round = 0; round < 2; ++round)
            for (int c = 0; c < ncand; ++c) {
                run(c);  // warm-up
                (void)hipEventRecord(e0, s);
                for (int r = 0; r < 3; ++r) run(c);
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) break;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (tmin[c] < 0.f || ms < tmin[c]) tmin[c] = ms;
            }
        for (int c = 1; c < ncand; ++c)
            if (tmin[c] >= 0.f && tmin[c] < tmin[best]) best = c;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[k] = best;
    return best;
}

// ---- y-lines through the transposed grid ------------------------------------------------------
// With the operator transposed once at setup (soT: c, w, s, sw, nw of the transposed grid, i.e.
// the transposes of c, s, w, sw, nw), a y-line sweep is an x-line sweep of the transposed grid:
// q and f are transposed, both colours are relaxed with the x-line kernel (right-hand side terms in
// the y-line order, the y-line factors are already stored column-contiguous), and q is transposed
// back.  The result is bitwise that of relax_lines_y.

template <typename T>
void transpose_op(const T *so, T *soT, int nstncl, int ii, int jj, hipStream_t s, int nb) {
    // every stencil plane of every problem in one launch (z = problem * nstncl + plane)
    const dim3 g((unsigned)((ii + 63) / 64), (unsigned)((jj + 63) / 64), (unsigned)(nb * nstncl));
    transpose_op_k<T><<<g, 512, 0, s>>>(so, soT, jj, ii, nstncl);
}

template <typename T>
static void relax_lines_ytr(const T *soT, int nstncl, const T *qf, T *q, const T *sory, T *qT,
                            T *qfT, int ii, int jj, int updown, int shape, hipStream_t s, int nb,
                            bool qfT_ready = false, bool q_in_T = false, bool keep_T = false,
                            int rep = 1) {
    if (!q_in_T) transpose<T>(q, qT, jj, ii, s, nb);
    if (!qfT_ready) transpose<T>(qf, qfT, jj, ii, s, nb);
    const int nfix = ii - 2;  // columns of the original grid = rows of the transposed one
    for (int t = 0; t < 2; ++t) {
        const int first = (updown == 0) ? (t == 0 ? 2 : 1) : (t == 0 ? 1 : 2);
        if (first > nfix) continue;
        const int nlines = (nfix - first) / 2 + 1;
        if (nstncl == 5)
            launch_class<T, true, 2>(shape, soT, qfT, qT, sory, jj, ii, first, nlines, s, nb,
                                     nullptr, t, rep);
        else
            launch_class<T, false, 2>(shape, soT, qfT, qT, sory, jj, ii, first, nlines, s, nb,
                                      nullptr, t, rep);
    }
    if (!keep_T) transpose<T>(qT, q, ii, jj, s, nb);
}

static int tuned_transposed(const tune_key &k) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    const auto it = g_tune.find(k);
    return it == g_tune.end() ? -1 : it->second;
}

template <typename T>
bool relax_lines_y_auto(const T *so, const T *soT, int nstncl, const T *qf, T *q, const T *sory,
                        T *qT, T *qfT, int ii, int jj, int updown, hipStream_t s, int nb,
                        bool qfT_ready, bool q_in_T, bool keep_T, int rep) {
    const int c = uses_transposed_y<T>(soT, nstncl, ii, jj, nb);
    if (c < 0) {
        if (q_in_T) transpose<T>(qT, q, ii, jj, s, nb);
        relax_lines<T, 1>(so, nstncl, qf, q, sory, ii, jj, updown, s, nb, rep);
        return false;
    }
    relax_lines_ytr<T>(soT, nstncl, qf, q, sory, qT, qfT, ii, jj, updown, c, s, nb, qfT_ready,
                       q_in_T, keep_T, rep);
    return true;
}

template <typename T>
int uses_transposed_y(const T *soT, int nstncl, int ii, int jj, int nb) {
    return soT ? tuned_transposed(key_of<T>(2, nstncl, ii, jj, nb)) : -1;
}

// Milliseconds for 3 replays of what run(cs) launches, captured as a graph on the capture stream
// cs and replayed on s, the way the solvers execute their cycles (eager back-to-back launches
// would leave gaps that favour the variant with fewer launches); -1 if capture fails.
template <typename F>
static float time_captured(hipStream_t cs, hipStream_t s, hipEvent_t e0, hipEvent_t e1, F &&run) {
    if (!cs || hipStreamSynchronize(s) != hipSuccess) return -1.f;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1.f;
    run(cs);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(cs, &g) != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return -1.f;
    }
    hipGraphExec_t ge = nullptr;
    float ms = -1.f;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess) {
        (void)hipGraphLaunch(ge, s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 3; ++r) (void)hipGraphLaunch(ge, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipGraphExecDestroy(ge);
    }
    (void)hipGraphDestroy(g);
    return ms;
}

template <typename T>
int tune_lines_ytr(const T *so, const T *soT, int nstncl, const T *qf, T *q, const T *sory,
                   T *qT, T *qfT, int ii, int jj, hipStream_t s, int nb, int sweeps, int rep) {
    sweeps = sweeps < 1 ? 1 : sweeps;
    const tune_key k = key_of<T>(2, nstncl, ii, jj, nb);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        if (g_tune.count(k)) return g_tune[k];
    }
    if (!g_autotune.load()) return -1;
    if (tune_lines<T>(1, so, nstncl, qf, q, sory, ii, jj, s, nb, rep) < 0) return -2;
    const int ncand = nstncl == 5 ? launch_class<T, true, 2>(-1, soT, qf, q, sory, jj, ii, 1, 1, s, nb)
                                  : launch_class<T, false, 2>(-1, soT, qf, q, sory, jj, ii, 1, 1, s, nb);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -2;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -2;
    }
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) cs = nullptr;
    auto timed1 = [&](auto &&run) {  // run(stream) launches one timed unit on that stream
        const float tc = time_captured(cs, s, e0, e1, run);
        if (tc >= 0.f) return tc;
        (void)hipGetLastError();  // eager timing when capture is not available
        run(s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 3; ++r) run(s);
        (void)hipEventRecord(e1, s);
        float ms = -1.f;
        if (hipEventSynchronize(e1) == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    auto timed = [&](auto &&run) {  // the faster of two timings
        const float a = timed1(run), b = timed1(run);
        return a < 0.f ? b : (b < 0.f ? a : (b < a ? b : a));
    };
    // `sweeps` y sweeps share one right-hand side (one transpose of qf)
    const float tdirect = timed([&](hipStream_t st) {
        for (int k = 0; k < sweeps; ++k)
            relax_lines<T, 1>(so, nstncl, qf, q, sory, ii, jj, 0, st, nb, rep);
    });
    int best = -1;
    float tbest = tdirect;
    for (int c = 0; c < ncand; ++c) {
        const float t = timed([&](hipStream_t st) {
            for (int k = 0; k < sweeps; ++k)
                relax_lines_ytr<T>(soT, nstncl, qf, q, sory, qT, qfT, ii, jj, 0, c, st, nb, k > 0,
                                   k > 0, k < sweeps - 1, rep);
        });
        if (t >= 0.f && t < tbest) {
            tbest = t;
            best = c;
        }
    }
    if (cs) (void)hipStreamDestroy(cs);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (hipGetLastError() != hipSuccess) return -2;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[k] = best;
    return best;
}

std::string tune_report() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    std::string out = "[";
    for (const auto &kv : g_tune) {
        const tune_key &k = kv.first;
        char buf[160];
        std::snprintf(buf, sizeof(buf),
                      "%s{\"dir\":\"%c\",\"fp\":%d,\"stencil\":%d,\"ii\":%d,\"jj\":%d,\"nb\":%d,"
                      "\"shape\":%d,\"kernel\":\"",
                      out.size() > 1 ? "," : "", k[0] == 0 ? 'x' : (k[0] == 1 ? 'y' : 't'),
                      8 * k[1], k[2] ? 9 : 5,
                      k[3], k[4], k[5], kv.second);
        out += buf;
        out += shape_name(k, kv.second) + "\"}";
    }
    return out + "]";
}

template <typename T>
void relax_lines_xy(const T *so, int nstncl, const T *qf, T *q, const T *sorx, const T *sory,
                    T *work, int ii, int jj, int updown, hipStream_t s, int nb, int rep) {
    (void)work;
    if (updown == 0) {
        relax_lines<T, 0>(so, nstncl, qf, q, sorx, ii, jj, updown, s, nb, rep);
        relax_lines<T, 1>(so, nstncl, qf, q, sory, ii, jj, updown, s, nb, rep);
    } else {
        relax_lines<T, 1>(so, nstncl, qf, q, sory, ii, jj, updown, s, nb, rep);
        relax_lines<T, 0>(so, nstncl, qf, q, sorx, ii, jj, updown, s, nb, rep);
    }
}

template <typename T>
void restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, hipStream_t s,
               int nb, int rep) {
    restrict_k<T><<<gridb(iic - 2, jjc - 2, nb), 256, 0, s>>>(ci, q, qc, iif, jjf, iic, jjc, rep);
}

template <typename T>
void restrict_cycle(const T *ci, const T *q, T *qc, T *xc, T *qcT, int iif, int jjf, int iic,
                    int jjc, hipStream_t s, int nb, int rep) {
    const dim3 g((unsigned)((iic + 63) / 64), (unsigned)((jjc + 15) / 16), (unsigned)nb);
    restrict_cycle_k<T><<<g, 64 * rc_rows<T>, 0, s>>>(ci, q, qc, xc, qcT, iif, jjf, iic, jjc, rep);
}

template <typename T>
void interp_add(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                int jjc, int iif, int jjf, hipStream_t s, int nb, int rep) {
    // one thread per 2 x 2 fine block (the range includes the ghost point of even sizes)
    interp_add_k<T><<<gridb(iif / 2, jjf / 2, nb), 256, 0, s>>>(q, qc, res, so, ci, iic, jjc, iif,
                                                                jjf, nstncl, rep);
}

template <typename T>
void residual_qT(const T *so, int nstncl, const T *qf, const T *qT, T *res, int ii, int jj,
                 hipStream_t s, int nb, int rep) {
    if (rep >= 2 && rep <= rq_maxr<T>) {  // one workgroup per operator tile, all its rhs
        const dim3 g((unsigned)((ii - 2 + kTI - 1) / kTI), (unsigned)((jj - 2 + kTJ - 1) / kTJ),
                     (unsigned)(nb / rep));
        if (nstncl == 5)
            residual_qT_rhs_k<T, true><<<g, 256, 0, s>>>(so, qf, qT, res, ii, jj, rep);
        else
            residual_qT_rhs_k<T, false><<<g, 256, 0, s>>>(so, qf, qT, res, ii, jj, rep);
        return;
    }
    const dim3 g((unsigned)((ii - 2 + kTI - 1) / kTI), (unsigned)((jj - 2 + kTJ - 1) / kTJ),
                 (unsigned)nb);
    if (nstncl == 5)
        residual_qT_k<T, true><<<g, 256, 0, s>>>(so, qf, qT, res, ii, jj, rep);
    else
        residual_qT_k<T, false><<<g, 256, 0, s>>>(so, qf, qT, res, ii, jj, rep);
}

template <typename T>
void interp_add_qT(T *qT, const T *qc, const T *res, const T *so, int nstncl, const T *ci,
                   int iic, int jjc, int iif, int jjf, hipStream_t s, int nb, int rep) {
    // tiles of 32 x 16 blocks of 2 x 2 fine points over the blocks m <= iif/2, n <= jjf/2
    // (one workgroup per tile of an operator; nb = operators * rep)
    const dim3 g((unsigned)((iif / 2 + 31) / 32), (unsigned)((jjf / 2 + 15) / 16),
                 (unsigned)(nb / rep));
    interp_add_qT_k<T><<<g, 512, 0, s>>>(qT, qc, res, so, ci, iic, jjc, iif, jjf, nstncl, rep);
}

void setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int iic, int jjc,
                  hipStream_t s, int nb) {
    const int icmax = (iif - 2) / 2 + 1, jcmax = (jjf - 2) / 2 + 1;
    const int ne = icmax > iic - 2 ? icmax : iic - 2;
    const int me = jcmax > jjc - 2 ? jcmax : jjc - 2;
    const dim3 ge = gridb(ne, me, nb);
    const dim3 gc = gridb(icmax - 1, jcmax - 1, nb);
    if (nstncl == 5) {
        oi_edges_k<true><<<ge, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc);
        oi_corners_k<true><<<gc, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc);
    } else {
        oi_edges_k<false><<<ge, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc);
        oi_corners_k<false><<<gc, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc);
    }
}

void setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif, int jjf,
                    int iic, int jjc, hipStream_t s, int nb) {
    const dim3 g = gridb(iic - 2, jjc - 2, nb);
    if (nstncl == 5)
        galerkin_k<true><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, iic, jjc);
    else
        galerkin_k<false><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, iic, jjc);
}

void setup_cg(const double *so, int nstncl, double *abd, int ii, int jj, int *info, hipStream_t s,
              int nb) {
    setup_cg_k<<<nb, 256, 0, s>>>(so, nstncl, abd, ii, jj, info);
}

template <typename T>
void solve_cg(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, hipStream_t s,
              int nb, int rep) {
    solve_cg_k<T><<<nb, 64, 0, s>>>(q, qf, abd, bbd, ii, jj, rep);
}

std::size_t fused_cycle_lds_bytes(int wbytes, const int *ii, const int *jj, int nlev,
                                  int lds_mode) {
    std::size_t b = 0;
    for (int l = 0; l < nlev; ++l) {
        const int m = l == 0 ? lds_mode : 1;
        b += wbytes == 8 ? fused_level_bytes<double>(ii[l], jj[l], l > 0, m)
                         : fused_level_bytes<float>(ii[l], jj[l], l > 0, m);
    }
    return b;
}

template <typename T>
void coarse_cycle(const coarse_level<T> *lv, int nlev, int max_side, int relax, int npre,
                  int npost, const double *abd, double *bbd, hipStream_t s, int nb, int rep,
                  std::size_t lds_bytes, int lds_mode) {
    const int mode = lds_bytes > 0 ? lds_mode : 0;
    if (max_side <= 64) {
        if (mode)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(coarse_cycle_k<T, 1>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        coarse_cycle_k<T, 1><<<nb, 1024, lds_bytes, s>>>(lv, nlev, relax, npre, npost, abd, bbd,
                                                         rep, mode);
    } else {
        if (mode)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(coarse_cycle_k<T, 2>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        coarse_cycle_k<T, 2><<<nb, 1024, lds_bytes, s>>>(lv, nlev, relax, npre, npost, abd, bbd,
                                                         rep, mode);
    }
}

template <typename Tin, typename Tout>
void convert(const Tin *in, Tout *out, std::size_t n, hipStream_t s) {
    convert_k<Tin, Tout><<<grid1(n), 256, 0, s>>>(in, out, n);
}

template <typename T>
void fill(T *x, T v, std::size_t n, hipStream_t s) {
    if (reinterpret_cast<std::uintptr_t>(x) % 16 == 0 && (16 / sizeof(T)) > 1)
        fill16_k<T><<<grid1(n / (16 / sizeof(T)) + 1), 256, 0, s>>>(x, v, n);
    else
        fill_k<T><<<grid1(n), 256, 0, s>>>(x, v, n);
}

void finalize_sum(const double *partials, std::size_t n, double *out, int nb, hipStream_t s) {
    finalize_sum_k<<<nb, 1024, 0, s>>>(partials, n, out);
}

void keep_inactive(double *x, const double *xprev, const int *active, std::size_t per, int nb,
                   hipStream_t s) {
    const std::size_t n = per * (std::size_t)nb;
    keep_inactive_k<<<grid1(n), 256, 0, s>>>(x, xprev, active, per, n);
}

void add_correction(double *x, const float *e, int ii, int jj, hipStream_t s, int nb) {
    add_correction_k<<<gridb(ii - 2, jj - 2, nb), 256, 0, s>>>(x, e, ii, jj);
}

// ------------------------------------------------------------------------------------------
// Periodic boundaries (the JPN != 0 branches of the reference's 2D kernels).  jpn is the BoxMG
// boundary id of BMG_get_bc (src/2d/ftn/BMG_get_bc.f90): 2 periodic in x, 1 in y, 3 in both.
// The periodic images live in the ghost rows and columns: Q(1,J) = Q(II-1,J), Q(II,J) = Q(2,J)
// and Q(I,1) = Q(I,JJ-1), Q(I,JJ) = Q(I,2) in the reference's 1-based indices.
// ------------------------------------------------------------------------------------------
namespace {

__host__ __device__ inline bool per_x(int jpn) { return jpn == 2 || jpn == 3 || jpn == -2 || jpn == -3; }
__host__ __device__ inline bool per_y(int jpn) { return jpn == 1 || jpn == 3 || jpn == -1 || jpn == -3; }

// x images of rows j0, j0 + jstep, ... <= j1 (0-based rows)
template <typename T>
__global__ __launch_bounds__(256) void ghost_x_k(T *__restrict__ q, int ii, int jj, int j0, int j1,
                                                 int jstep) {
    const int j = j0 + (int)(blockIdx.x * 256 + threadIdx.x) * jstep;
    if (j > j1) return;
    q += (std::size_t)blockIdx.z * ii * jj;
    const std::size_t r = (std::size_t)j * ii;
    q[r] = q[r + ii - 2];
    q[r + ii - 1] = q[r + 1];
}

// y images of columns i0 .. i1
template <typename T>
__global__ __launch_bounds__(256) void ghost_y_k(T *__restrict__ q, int ii, int jj, int i0, int i1) {
    const int i = i0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (i > i1) return;
    q += (std::size_t)blockIdx.z * ii * jj;
    q[i] = q[(std::size_t)(jj - 2) * ii + i];
    q[(std::size_t)(jj - 1) * ii + i] = q[(std::size_t)ii + i];
}

template <typename T>
void ghost_x(T *q, int ii, int jj, int j0, int j1, int jstep, hipStream_t s, int nb) {
    if (j1 < j0) return;
    const int n = (j1 - j0) / jstep + 1;
    ghost_x_k<T><<<dim3((unsigned)((n + 255) / 256), 1, (unsigned)nb), 256, 0, s>>>(q, ii, jj, j0,
                                                                                   j1, jstep);
}
template <typename T>
void ghost_y(T *q, int ii, int jj, int i0, int i1, hipStream_t s, int nb) {
    if (i1 < i0) return;
    const int n = i1 - i0 + 1;
    ghost_y_k<T><<<dim3((unsigned)((n + 255) / 256), 1, (unsigned)nb), 256, 0, s>>>(q, ii, jj, i0,
                                                                                   i1);
}

// ---- operator-induced interpolation, periodic branch -------------------------------------
// (src/2d/ftn/BMG2_SymStd_SETUP_interp_OI.f90:259-616).  Loop bounds of the reference; the fine
// index of coarse index c follows I = MAX(MOD(I+2, N), MIN(I+2, 3)) from I = beg, which is
// 2(c-1) except at the periodic end, where it wraps to 3.
struct oi_bounds {
    int ibeg, ibegc, iendc, ibeg_x, iend_x, iifc;
    int jbeg, jbegc, jendc, jbeg_y, jend_y, jjfc;
};

__device__ inline int oi_fine(int c, int cbeg, int beg, int n) {
    const int prev = c == cbeg ? beg : 2 * (c - 2);
    const int a = (prev + 2) % n, b = prev + 2 < 3 ? prev + 2 : 3;
    return a > b ? a : b;
}

#define SOF(i, j, k) so[(std::size_t)((k) - 1) * P + (std::size_t)((j) - 1) * iif + ((i) - 1)]
#define CIF(i, j, k) ci[(std::size_t)((k) - 1) * Pc + (std::size_t)((j) - 1) * iic + ((i) - 1)]

__device__ inline double oi_lumped(double c, double sum, double ep) {
    const double emach = 0x1.0p-53;  // D1MACH(3)
    return (c - sum) * fmax(c - (1.0 + ep) * sum, 0.0) / (fabs(c - (1.0 + ep) * sum) + emach);
}

template <bool NINE>
__device__ void oi_xedge(const double *__restrict__ so, double *__restrict__ ci, std::size_t P,
                         std::size_t Pc, int iif, int iic, int i, int j, int ic, int jc) {
    constexpr int KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5, LL = 1, LR = 2;
    double a, b;
    if (NINE) {
        a = SOF(i, j, KW) + SOF(i, j, KNW) + SOF(i, j + 1, KSW);
        b = SOF(i - 1, j, KW) + SOF(i - 1, j, KSW) + SOF(i - 1, j + 1, KNW);
    } else {
        a = SOF(i, j, KW);
        b = SOF(i - 1, j, KW);
    }
    const double c = SOF(i - 1, j, KO);
    const double ep = fmin(fabs(a / c), fabs(b / c));
    double sum = a + b + SOF(i - 1, j, KS) + SOF(i - 1, j + 1, KS);
    sum = a + b + oi_lumped(c, sum, ep);
    sum = 1.0 / sum;
    CIF(ic, jc, LR) = a * sum;
    CIF(ic, jc, LL) = b * sum;
}

template <bool NINE>
__device__ void oi_yedge(const double *__restrict__ so, double *__restrict__ ci, std::size_t P,
                         std::size_t Pc, int iif, int iic, int i, int j, int ic, int jc) {
    constexpr int KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5, LA = 3, LB = 4;
    double a, b;
    if (NINE) {
        a = SOF(i, j, KS) + SOF(i, j, KNW) + SOF(i + 1, j, KSW);
        b = SOF(i, j - 1, KS) + SOF(i, j - 1, KSW) + SOF(i + 1, j - 1, KNW);
    } else {
        a = SOF(i, j, KS);
        b = SOF(i, j - 1, KS);
    }
    const double c = SOF(i, j - 1, KO);
    double sum = a + b + SOF(i, j - 1, KW) + SOF(i + 1, j - 1, KW);
    const double ep = fmin(fabs(a / c), fabs(b / c));
    sum = a + b + oi_lumped(c, sum, ep);
    sum = 1.0 / sum;
    CIF(ic, jc, LA) = a * sum;
    CIF(ic, jc, LB) = b * sum;
}

template <bool NINE>
__device__ void oi_corner(const double *__restrict__ so, double *__restrict__ ci, std::size_t P,
                          std::size_t Pc, int iif, int iic, int i, int j, int ic, int jc) {
    constexpr int KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5;
    constexpr int LL = 1, LR = 2, LA = 3, LB = 4, LSW = 5, LNW = 6, LNE = 7, LSE = 8;
    const double c = SOF(i - 1, j - 1, KO);
    double sum, ep;
    if (NINE) {
        sum = SOF(i - 1, j - 1, KW) + SOF(i - 1, j, KNW) + SOF(i - 1, j, KS) + SOF(i, j, KSW) +
              SOF(i, j - 1, KW) + SOF(i, j - 1, KNW) + SOF(i - 1, j - 1, KS) + SOF(i - 1, j - 1, KSW);
        const double e1 = fabs((SOF(i - 1, j - 1, KSW) + SOF(i - 1, j - 1, KW) + SOF(i - 1, j, KNW)) / c);
        const double e2 = fabs((SOF(i - 1, j, KNW) + SOF(i - 1, j, KS) + SOF(i, j, KSW)) / c);
        const double e3 = fabs((SOF(i, j, KSW) + SOF(i, j - 1, KW) + SOF(i, j - 1, KNW)) / c);
        const double e4 = fabs((SOF(i, j - 1, KNW) + SOF(i - 1, j - 1, KS) + SOF(i - 1, j - 1, KSW)) / c);
        ep = fmin(fmin(fmin(e1, e2), e3), e4);
    } else {
        sum = SOF(i - 1, j - 1, KW) + SOF(i - 1, j, KS) + SOF(i, j - 1, KW) + SOF(i - 1, j - 1, KS);
        const double e1 = fabs(SOF(i - 1, j - 1, KW) / c);
        const double e2 = fabs(SOF(i - 1, j, KS) / c);
        const double e3 = fabs(SOF(i, j - 1, KW) / c);
        const double e4 = fabs(SOF(i - 1, j - 1, KS) / c);
        ep = fmin(fmin(fmin(e1, e2), e3), e4);
    }
    sum = sum + oi_lumped(c, sum, ep);
    const double s = 1.0 / sum;
    if (NINE) {
        CIF(ic, jc, LSW) = (SOF(i - 1, j - 1, KS) * CIF(ic, jc - 1, LL) +
                            SOF(i - 1, j - 1, KW) * CIF(ic - 1, jc, LB) + SOF(i - 1, j - 1, KSW)) * s;
        CIF(ic, jc, LSE) = (SOF(i - 1, j - 1, KS) * CIF(ic, jc - 1, LR) +
                            SOF(i, j - 1, KW) * CIF(ic, jc, LB) + SOF(i, j - 1, KNW)) * s;
        CIF(ic, jc, LNW) = (SOF(i - 1, j - 1, KW) * CIF(ic - 1, jc, LA) +
                            SOF(i - 1, j, KS) * CIF(ic, jc, LL) + SOF(i - 1, j, KNW)) * s;
        CIF(ic, jc, LNE) = (SOF(i - 1, j, KS) * CIF(ic, jc, LR) +
                            SOF(i, j - 1, KW) * CIF(ic, jc, LA) + SOF(i, j, KSW)) * s;
    } else {
        CIF(ic, jc, LSW) = (SOF(i - 1, j - 1, KS) * CIF(ic, jc - 1, LL) +
                            SOF(i - 1, j - 1, KW) * CIF(ic - 1, jc, LB)) * s;
        CIF(ic, jc, LSE) = (SOF(i - 1, j - 1, KS) * CIF(ic, jc - 1, LR) +
                            SOF(i, j - 1, KW) * CIF(ic, jc, LB)) * s;
        CIF(ic, jc, LNW) = (SOF(i - 1, j - 1, KW) * CIF(ic - 1, jc, LA) +
                            SOF(i - 1, j, KS) * CIF(ic, jc, LL)) * s;
        CIF(ic, jc, LNE) = (SOF(i - 1, j, KS) * CIF(ic, jc, LR) +
                            SOF(i, j - 1, KW) * CIF(ic, jc, LA)) * s;
    }
}
#undef SOF
#undef CIF

// x-edge weights: rows jc = 2..jjc-1 (blockIdx.y rows of 4), and when periodic in y the two
// ghost rows jjc (from fine row jbeg_y) and 1 (from fine row jend_y) as rows jjc-1+1, jjc-1+2
template <bool NINE>
__global__ __launch_bounds__(256) void oi_per_xedge_k(const double *__restrict__ so,
                                                      double *__restrict__ ci, int iif, int jjf,
                                                      int iic, int jjc, oi_bounds B, int extra) {
    const int ic = B.ibegc + (int)(blockIdx.x * 64 + (threadIdx.x % 64));
    const int r = (int)(blockIdx.y * 4 + (threadIdx.x / 64));  // 0-based row item
    const int nrow = jjc - 2 + extra;
    if (ic > B.iendc || r >= nrow) return;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    const int i = oi_fine(ic, B.ibegc, B.ibeg, B.iifc);
    int jc, j;
    if (r < jjc - 2) {
        jc = r + 2;
        j = 2 * (jc - 1);
    } else if (r == jjc - 2) {
        jc = jjc;
        j = B.jbeg_y;
    } else {
        jc = 1;
        j = B.jend_y;
    }
    oi_xedge<NINE>(so, ci, P, Pc, iif, iic, i, j, ic, jc);
}

// y-edge weights: columns ic = 2..iic-1, and when periodic in x the ghost columns iic (from fine
// column ibeg_x) and 1 (from iend_x); rows jc = jbegc..jendc
template <bool NINE>
__global__ __launch_bounds__(256) void oi_per_yedge_k(const double *__restrict__ so,
                                                      double *__restrict__ ci, int iif, int jjf,
                                                      int iic, int jjc, oi_bounds B, int extra) {
    const int c = (int)(blockIdx.x * 64 + (threadIdx.x % 64));  // 0-based column item
    const int jc = B.jbegc + (int)(blockIdx.y * 4 + (threadIdx.x / 64));
    const int ncol = iic - 2 + extra;
    if (c >= ncol || jc > B.jendc) return;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    const int j = oi_fine(jc, B.jbegc, B.jbeg, B.jjfc);
    int ic, i;
    if (c < iic - 2) {
        ic = c + 2;
        i = 2 * (ic - 1);
    } else if (c == iic - 2) {
        ic = iic;
        i = B.ibeg_x;
    } else {
        ic = 1;
        i = B.iend_x;
    }
    oi_yedge<NINE>(so, ci, P, Pc, iif, iic, i, j, ic, jc);
}

template <bool NINE>
__global__ __launch_bounds__(256) void oi_per_corner_k(const double *__restrict__ so,
                                                       double *__restrict__ ci, int iif, int jjf,
                                                       int iic, int jjc, oi_bounds B) {
    const int ic = B.ibegc + (int)(blockIdx.x * 64 + (threadIdx.x % 64));
    const int jc = B.jbegc + (int)(blockIdx.y * 4 + (threadIdx.x / 64));
    if (ic > B.iendc || jc > B.jendc) return;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    const int i = oi_fine(ic, B.ibegc, B.ibeg, B.iifc);
    const int j = oi_fine(jc, B.jbegc, B.jbeg, B.jjfc);
    oi_corner<NINE>(so, ci, P, Pc, iif, iic, i, j, ic, jc);
}

// Galerkin coarse operator in the reference's hand-expanded form (the ITLI products of
// src/2d/ftn/BMG2_SymStd_SETUP_ITLI_ex.f90:94-331), one thread per coarse point.  With periodic
// boundaries these formulas read the periodic images in the ghost rows and columns of SO and CI,
// which the generic P^T A P of galerkin_k does not visit; Dirichlet problems keep galerkin_k.
template <bool NINE>
__global__ __launch_bounds__(256) void galerkin_itli_k(const double *__restrict__ so,
                                                       const double *__restrict__ ci,
                                                       double *__restrict__ soc, int iif, int jjf,
                                                       int iic, int jjc) {
    const int ic = blockIdx.x * 64 + (threadIdx.x % 64) + 2;  // Fortran indices
    const int jc = blockIdx.y * 4 + (threadIdx.x / 64) + 2;
    if (ic > iic - 1 || jc > jjc - 1) return;
    const std::size_t P = (std::size_t)iif * jjf, Pc = (std::size_t)iic * jjc, bz = blockIdx.z;
    so += bz * (NINE ? 5 : 3) * P;
    ci += bz * 8 * Pc;
    soc += bz * 5 * Pc;
#define SO(i, j, k) so[(std::size_t)((k) - 1) * P + (std::size_t)((j) - 1) * iif + ((i) - 1)]
#define CI(i, j, k) ci[(std::size_t)((k) - 1) * Pc + (std::size_t)((j) - 1) * iic + ((i) - 1)]
#define SOC(i, j, k) soc[(std::size_t)((k) - 1) * Pc + (std::size_t)((j) - 1) * iic + ((i) - 1)]
    constexpr int KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5;
    constexpr int LL = 1, LR = 2, LA = 3, LB = 4, LSW = 5, LNW = 6, LNE = 7, LSE = 8;
    const int j = 2 * jc - 2, i = 2 * ic - 2;
    if (NINE) {
        double CO = SO(i, j + 1, KNW) * CI(ic, jc + 1, LSW) + SO(i, j, KW) * CI(ic, jc, LL)
                  + SO(i, j, KSW) * CI(ic, jc, LNW);
        double CS = SO(i, j - 1, KW) * CI(ic, jc, LNW) + SO(i, j, KNW) * CI(ic, jc, LL);
        double CSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LNW) + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LA)
                   + SO(i - 1, j, KNW) + SO(i - 1, j, KS) * CI(ic, jc, LL);
        double CW = -SO(i - 1, j, KO) * CI(ic, jc, LL) + SO(i - 1, j, KS) * CI(ic, jc, LNW)
                  + SO(i - 1, j, KSW) * CI(ic - 1, jc, LA) + SO(i - 1, j, KW)
                  + SO(i - 1, j + 1, KNW) * CI(ic - 1, jc + 1, LB)
                  + SO(i - 1, j + 1, KS) * CI(ic, jc + 1, LSW);
        double CNW = -SO(i - 1, j + 1, KO) * CI(ic, jc + 1, LSW) + SO(i - 1, j + 1, KS) * CI(ic, jc, LL)
                   + SO(i - 1, j + 1, KSW) + SO(i - 1, j + 1, KW) * CI(ic - 1, jc + 1, LB);
        double CN = SO(i, j + 1, KSW) * CI(ic, jc, LL) + SO(i, j + 1, KW) * CI(ic, jc + 1, LSW);
        SOC(ic, jc, KW) = CO + CI(ic, jc, LA) * CS + CI(ic, jc, LNE) * CSW + CI(ic, jc, LR) * CW
                        + CI(ic, jc + 1, LSE) * CNW + CI(ic, jc + 1, LB) * CN;
        double COSW = SO(i, j, KSW) * CI(ic, jc, LSW);
        double CSSW = SO(i, j - 1, KSW) * CI(ic, jc - 1, LL) + SO(i, j - 1, KW) * CI(ic, jc, LSW);
        double CSWSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSW) + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LL)
                     + SO(i - 1, j - 1, KSW) + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LB);
        double CWSW = SO(i - 1, j, KS) * CI(ic, jc, LSW) + SO(i - 1, j, KSW) * CI(ic - 1, jc, LB);
        SOC(ic, jc, KSW) = COSW + CI(ic, jc, LA) * CSSW + CI(ic, jc, LNE) * CSWSW
                         + CI(ic, jc, LR) * CWSW;
        double COA = SO(i, j, KSW) * CI(ic, jc, LSE) + SO(i, j, KS) * CI(ic, jc, LB)
                   + SO(i + 1, j, KNW) * CI(ic + 1, jc, LSW);
        double CEA = SO(i + 1, j, KS) * CI(ic + 1, jc, LSW) + SO(i + 1, j, KSW) * CI(ic, jc, LB);
        double CSEA = -SO(i + 1, j - 1, KO) * CI(ic + 1, jc, LSW) + SO(i + 1, j - 1, KS) * CI(ic + 1, jc - 1, LL)
                    + SO(i + 1, j - 1, KSW) + SO(i + 1, j - 1, KW) * CI(ic, jc, LB);
        double CSA = -SO(i, j - 1, KO) * CI(ic, jc, LB) + SO(i + 1, j - 1, KW) * CI(ic + 1, jc, LSW)
                   + SO(i + 1, j - 1, KNW) * CI(ic + 1, jc - 1, LL) + SO(i, j - 1, KS)
                   + SO(i, j - 1, KSW) * CI(ic, jc - 1, LR) + SO(i, j - 1, KW) * CI(ic, jc, LSE);
        double CSWA = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSE) + SO(i, j - 1, KW) * CI(ic, jc, LB)
                    + SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR);
        double CWA = SO(i, j, KNW) * CI(ic, jc, LB) + SO(i - 1, j, KS) * CI(ic, jc, LSE);
        SOC(ic, jc, KS) = COA + CI(ic + 1, jc, LL) * CEA + CI(ic + 1, jc, LNW) * CSEA
                        + CI(ic, jc, LA) * CSA + CI(ic, jc, LNE) * CSWA + CI(ic, jc, LR) * CWA;
        double CONW = SO(i - 1, j, KNW) * CI(ic, jc, LSE);
        double CENW = SO(i, j, KNW) * CI(ic, jc, LB) + SO(i - 1, j, KS) * CI(ic, jc, LSE);
        double CSENW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSE) + SO(i, j - 1, KW) * CI(ic, jc, LB)
                     + SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR);
        double CFNW = SO(i - 1, j - 1, KW) * CI(ic, jc, LSE) + SO(i - 1, j - 1, KNW) * CI(ic, jc - 1, LR);
        SOC(ic, jc, KNW) = CONW + CI(ic, jc, LL) * CENW + CI(ic, jc, LNW) * CSENW
                         + CI(ic - 1, jc, LA) * CFNW;
        {
            double CO = SO(i, j, KW) * CI(ic, jc, LR) + SO(i, j + 1, KNW) * CI(ic, jc + 1, LSE)
                      + SO(i, j + 1, KS) * CI(ic, jc + 1, LB) + SO(i + 1, j + 1, KSW) * CI(ic + 1, jc + 1, LSW)
                      + SO(i + 1, j, KW) * CI(ic + 1, jc, LL) + SO(i + 1, j, KNW) * CI(ic + 1, jc, LNW)
                      + SO(i, j, KS) * CI(ic, jc, LA) + SO(i, j, KSW) * CI(ic, jc, LNE) - SO(i, j, KO);
            double CW = -SO(i - 1, j, KO) * CI(ic, jc, LR) + SO(i - 1, j + 1, KS) * CI(ic, jc + 1, LSE)
                      + SO(i, j + 1, KSW) * CI(ic, jc + 1, LB) + SO(i, j, KW) + SO(i, j, KNW) * CI(ic, jc, LA)
                      + SO(i - 1, j, KS) * CI(ic, jc, LNE);
            double CNW = -SO(i - 1, j + 1, KO) * CI(ic, jc + 1, LSE) + SO(i, j + 1, KW) * CI(ic, jc + 1, LB)
                       + SO(i, j + 1, KNW) + SO(i - 1, j + 1, KS) * CI(ic, jc, LR);
            double CN = -SO(i, j + 1, KO) * CI(ic, jc + 1, LB) + SO(i + 1, j + 1, KNW) * CI(ic + 1, jc, LL)
                      + SO(i, j + 1, KS) + SO(i, j + 1, KSW) * CI(ic, jc, LR)
                      + SO(i, j + 1, KW) * CI(ic, jc + 1, LSE) + SO(i + 1, j + 1, KW) * CI(ic + 1, jc + 1, LSW);
            double CNE = -SO(i + 1, j + 1, KO) * CI(ic + 1, jc + 1, LSW) + SO(i + 1, j + 1, KS) * CI(ic + 1, jc, LL)
                       + SO(i + 1, j + 1, KSW) + SO(i + 1, j + 1, KW) * CI(ic, jc + 1, LB);
            double CE = -SO(i + 1, j, KO) * CI(ic + 1, jc, LL) + SO(i + 1, j, KS) * CI(ic + 1, jc, LNW)
                      + SO(i + 1, j, KSW) * CI(ic, jc, LA) + SO(i + 1, j, KW)
                      + SO(i + 1, j + 1, KNW) * CI(ic, jc + 1, LB) + SO(i + 1, j + 1, KS) * CI(ic + 1, jc + 1, LSW);
            double CSE = -SO(i + 1, j - 1, KO) * CI(ic + 1, jc, LNW) + SO(i + 1, j - 1, KW) * CI(ic, jc, LA)
                       + SO(i + 1, j, KNW) + SO(i + 1, j, KS) * CI(ic + 1, jc, LL);
            double CS = -SO(i, j - 1, KO) * CI(ic, jc, LA) + SO(i, j - 1, KW) * CI(ic, jc, LNE)
                      + SO(i, j, KNW) * CI(ic, jc, LR) + SO(i, j, KS) + SO(i + 1, j, KSW) * CI(ic + 1, jc, LL)
                      + SO(i + 1, j - 1, KW) * CI(ic + 1, jc, LNW);
            double CSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LNE) + SO(i - 1, j, KS) * CI(ic, jc, LR)
                       + SO(i, j, KSW) + SO(i, j - 1, KW) * CI(ic, jc, LA);
            SOC(ic, jc, KO) = -CI(ic, jc + 1, LSE) * CNW - CI(ic, jc + 1, LB) * CN
                            - CI(ic + 1, jc + 1, LSW) * CNE - CI(ic, jc, LR) * CW - CO
                            - CI(ic + 1, jc, LL) * CE - CI(ic, jc, LNE) * CSW
                            - CI(ic, jc, LA) * CS - CI(ic + 1, jc, LNW) * CSE;
        }
    } else {
        double CO = SO(i, j, KW) * CI(ic, jc, LL);
        double CS = SO(i, j - 1, KW) * CI(ic, jc, LNW);
        double CSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LNW) + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LA)
                   + SO(i - 1, j, KS) * CI(ic, jc, LL);
        double CW = -SO(i - 1, j, KO) * CI(ic, jc, LL) + SO(i - 1, j, KS) * CI(ic, jc, LNW)
                  + SO(i - 1, j, KW) + SO(i - 1, j + 1, KS) * CI(ic, jc + 1, LSW);
        double CNW = -SO(i - 1, j + 1, KO) * CI(ic, jc + 1, LSW) + SO(i - 1, j + 1, KS) * CI(ic, jc, LL)
                   + SO(i - 1, j + 1, KW) * CI(ic - 1, jc + 1, LB);
        double CN = SO(i, j + 1, KW) * CI(ic, jc + 1, LSW);
        SOC(ic, jc, KW) = CO + CI(ic, jc, LA) * CS + CI(ic, jc, LNE) * CSW + CI(ic, jc, LR) * CW
                        + CI(ic, jc + 1, LSE) * CNW + CI(ic, jc + 1, LB) * CN;
        double CSSW = SO(i, j - 1, KW) * CI(ic, jc, LSW);
        double CSWSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSW) + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LL)
                     + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LB);
        double CWSW = SO(i - 1, j, KS) * CI(ic, jc, LSW);
        SOC(ic, jc, KSW) = CI(ic, jc, LA) * CSSW + CI(ic, jc, LNE) * CSWSW + CI(ic, jc, LR) * CWSW;
        double COA = SO(i, j, KS) * CI(ic, jc, LB);
        double CEA = SO(i + 1, j, KS) * CI(ic + 1, jc, LSW);
        double CSEA = -SO(i + 1, j - 1, KO) * CI(ic + 1, jc, LSW) + SO(i + 1, j - 1, KS) * CI(ic + 1, jc - 1, LL)
                    + SO(i + 1, j - 1, KW) * CI(ic, jc, LB);
        double CSA = -SO(i, j - 1, KO) * CI(ic, jc, LB) + SO(i + 1, j - 1, KW) * CI(ic + 1, jc, LSW)
                   + SO(i, j - 1, KS) + SO(i, j - 1, KW) * CI(ic, jc, LSE);
        double CSWA = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSE) + SO(i, j - 1, KW) * CI(ic, jc, LB)
                    + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR);
        double CWA = SO(i - 1, j, KS) * CI(ic, jc, LSE);
        SOC(ic, jc, KS) = COA + CI(ic + 1, jc, LL) * CEA + CI(ic + 1, jc, LNW) * CSEA
                        + CI(ic, jc, LA) * CSA + CI(ic, jc, LNE) * CSWA + CI(ic, jc, LR) * CWA;
        double CENW = SO(i - 1, j, KS) * CI(ic, jc, LSE);
        double CSENW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LSE) + SO(i, j - 1, KW) * CI(ic, jc, LB)
                     + SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR);
        double CSNW = SO(i - 1, j - 1, KW) * CI(ic, jc, LSE);
        SOC(ic, jc, KNW) = CI(ic, jc, LL) * CENW + CI(ic, jc, LNW) * CSENW + CI(ic - 1, jc, LA) * CSNW;
        CO = SO(i, j, KW) * CI(ic, jc, LR) + SO(i, j + 1, KS) * CI(ic, jc + 1, LB)
           + SO(i + 1, j, KW) * CI(ic + 1, jc, LL) + SO(i, j, KS) * CI(ic, jc, LA) - SO(i, j, KO);
        CW = -SO(i - 1, j, KO) * CI(ic, jc, LR) + SO(i - 1, j + 1, KS) * CI(ic, jc + 1, LSE)
           + SO(i, j, KW) + SO(i - 1, j, KS) * CI(ic, jc, LNE);
        CNW = -SO(i - 1, j + 1, KO) * CI(ic, jc + 1, LSE) + SO(i, j + 1, KW) * CI(ic, jc + 1, LB)
            + SO(i - 1, j + 1, KS) * CI(ic, jc, LR);
        CN = -SO(i, j + 1, KO) * CI(ic, jc + 1, LB) + SO(i, j + 1, KS)
           + SO(i, j + 1, KW) * CI(ic, jc + 1, LSE) + SO(i + 1, j + 1, KW) * CI(ic + 1, jc + 1, LSW);
        double CNE = -SO(i + 1, j + 1, KO) * CI(ic + 1, jc + 1, LSW) + SO(i + 1, j + 1, KS) * CI(ic + 1, jc, LL)
                   + SO(i + 1, j + 1, KW) * CI(ic, jc + 1, LB);
        double CE = -SO(i + 1, j, KO) * CI(ic + 1, jc, LL) + SO(i + 1, j, KS) * CI(ic + 1, jc, LNW)
                  + SO(i + 1, j, KW) + SO(i + 1, j + 1, KS) * CI(ic + 1, jc + 1, LSW);
        double CSE = -SO(i + 1, j - 1, KO) * CI(ic + 1, jc, LNW) + SO(i + 1, j - 1, KW) * CI(ic, jc, LA)
                   + SO(i + 1, j, KS) * CI(ic + 1, jc, LL);
        CS = -SO(i, j - 1, KO) * CI(ic, jc, LA) + SO(i, j - 1, KW) * CI(ic, jc, LNE) + SO(i, j, KS)
           + SO(i + 1, j - 1, KW) * CI(ic + 1, jc, LNW);
        CSW = -SO(i - 1, j - 1, KO) * CI(ic, jc, LNE) + SO(i - 1, j, KS) * CI(ic, jc, LR)
            + SO(i, j - 1, KW) * CI(ic, jc, LA);
        SOC(ic, jc, KO) = -CI(ic, jc + 1, LSE) * CNW - CI(ic, jc + 1, LB) * CN
                        - CI(ic + 1, jc + 1, LSW) * CNE - CI(ic, jc, LR) * CW - CO
                        - CI(ic + 1, jc, LL) * CE - CI(ic, jc, LNE) * CSW
                        - CI(ic, jc, LA) * CS - CI(ic + 1, jc, LNW) * CSE;
    }
#undef SO
#undef CI
#undef SOC
}


// periodic images of the coarse operator (all 5 planes): rows first, then columns
// (src/2d/ftn/BMG2_SymStd_SETUP_ITLI_ex.f90:219-246)
__global__ __launch_bounds__(256) void soc_ghost_y_k(double *__restrict__ soc, int iic, int jjc) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= iic) return;
    const std::size_t Pc = (std::size_t)iic * jjc;
    soc += (std::size_t)blockIdx.z * 5 * Pc;
    for (int k = 0; k < 5; ++k) {
        double *s = soc + k * Pc;
        s[(std::size_t)(jjc - 1) * iic + i] = s[(std::size_t)iic + i];
        s[i] = s[(std::size_t)(jjc - 2) * iic + i];
    }
}
__global__ __launch_bounds__(256) void soc_ghost_x_k(double *__restrict__ soc, int iic, int jjc) {
    const int j = (int)(blockIdx.x * 256 + threadIdx.x);
    if (j >= jjc) return;
    const std::size_t Pc = (std::size_t)iic * jjc;
    soc += (std::size_t)blockIdx.z * 5 * Pc;
    for (int k = 0; k < 5; ++k) {
        double *s = soc + k * Pc + (std::size_t)j * iic;
        s[iic - 1] = s[1];
        s[0] = s[iic - 2];
    }
}

// ---- coarsest grid, periodic: dense matrix, Cholesky, solve with mean removal ----------------
// Assembly of the upper triangle as src/2d/ftn/BMG2_SymStd_SETUP_cg_LU.f90:119-227 (9-point)
// and :251-335 (5-point) write it (one thread: the writes overlap by design), then an upper
// Cholesky in DPOTF2's order: each pivot by one lane, the row of U across the workgroup, every
// entry's dot product summed in index order.  ABD(n, n) column-major, n = nx*ny, per problem.
__global__ __launch_bounds__(256) void setup_cg_per_k(const double *__restrict__ so, int nstncl,
                                                      double *__restrict__ abd, int ii, int jj,
                                                      int jpn, int *__restrict__ info) {
    const int i1 = ii - 1, j1 = jj - 1, i2 = i1 - 1, n = i2 * (j1 - 1);
    const std::size_t P = (std::size_t)ii * jj;
    so += (std::size_t)blockIdx.x * nstncl * P;
    abd += (std::size_t)blockIdx.x * n * n;
    for (std::size_t t = threadIdx.x; t < (std::size_t)n * n; t += blockDim.x) abd[t] = 0.0;
    __syncthreads();
    auto SO = [&](int i, int j, int k) { return so[(std::size_t)(k - 1) * P + (std::size_t)(j - 1) * ii + (i - 1)]; };
    auto A = [&](int r, int c) -> double & { return abd[(std::size_t)(c - 1) * n + (r - 1)]; };
    constexpr int KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5;
    const bool nine = nstncl == 5, px = per_x(jpn), py = per_y(jpn), pxy = jpn == 3 || jpn == -3;
    if (threadIdx.x == 0) {
        int kk = 1;
        A(1, 1) = SO(2, 2, KO);
        for (int i = 3; i <= i1; ++i) {
            ++kk;
            A(kk, kk) = SO(i, 2, KO);
            A(kk - 1, kk) = -SO(i, 2, KW);
        }
        if (px) A(kk - i2 + 1, kk) = -SO(ii, 2, KW);
        for (int j = 3; j <= j1; ++j) {
            if (px) A(kk, kk + 1) = nine ? -SO(2, j, KSW) : 0.0;
            for (int i = 2; i <= i1; ++i) {
                ++kk;
                A(kk, kk) = SO(i, j, KO);
                if (i != 2) {
                    A(kk - 1, kk) = -SO(i, j, KW);
                    A(kk - i2 - 1, kk) = nine ? -SO(i, j, KSW) : 0.0;
                }
                A(kk - i2 + 1, kk) = nine ? -SO(i + 1, j, KNW) : 0.0;
                A(kk - i2, kk) = -SO(i, j, KS);
            }
            if (px) {
                A(kk - i2 + 1, kk) = -SO(ii, j, KW);
                A(kk - 2 * i2 + 1, kk) = nine ? -SO(ii, j, KNW) : 0.0;
            }
        }
        if (py) {
            kk = kk - i2;
            const int j2 = (j1 - 2) * i2;
            ++kk;
            A(kk - j2, kk) = -SO(2, jj, KS);
            A(kk - j2 + 1, kk) = nine ? -SO(3, jj, KSW) : 0.0;
            if (pxy) A(i2, kk) = nine ? -SO(2, jj, KNW) : 0.0;
            for (int i = 3; i <= i1; ++i) {
                ++kk;
                A(kk - j2, kk) = -SO(i, jj, KS);
                A(kk - j2 - 1, kk) = nine ? -SO(i, jj, KNW) : 0.0;
                A(kk - j2 + 1, kk) = nine ? -SO(i + 1, jj, KSW) : 0.0;
            }
            A(kk - j2 + 1, kk) = 0.0;
            if (pxy) A(1, kk) = nine ? -SO(ii, jj, KSW) : 0.0;
            if (px) A(kk - 2 * i2 + 1, kk) = nine ? -SO(ii, j1, KNW) : 0.0;
        }
    }
    __syncthreads();
    __shared__ double piv;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int j = 1; j <= n; ++j) {
        if (threadIdx.x == 0) {
            double ajj = A(j, j);
            for (int k = 1; k < j; ++k) ajj -= A(k, j) * A(k, j);
            if (!(ajj > 0.0)) {
                A(j, j) = ajj;
                bad = j;
            } else {
                ajj = sqrt(ajj);
                A(j, j) = ajj;
            }
            piv = ajj;
        }
        __syncthreads();
        if (bad) break;
        for (int c = j + 1 + (int)threadIdx.x; c <= n; c += (int)blockDim.x) {
            double s = A(j, c);
            for (int k = 1; k < j; ++k) s -= A(k, j) * A(k, c);
            A(j, c) = s / piv;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && bad) atomicMax(info, bad);
}

// DPOTRS on the factor, then the mean removal and the periodic images
// (src/2d/ftn/BMG2_SymStd_SOLVE_cg.f90:93-163); one lane per problem
template <typename T>
__global__ void solve_cg_per_k(T *__restrict__ q, const T *__restrict__ qf,
                               const double *__restrict__ abd, double *__restrict__ x, int ii,
                               int jj, int jpn) {
    if (threadIdx.x != 0) return;
    const int i1 = ii - 1, j1 = jj - 1, nx = ii - 2, n = nx * (jj - 2);
    q += (std::size_t)blockIdx.x * ii * jj;
    qf += (std::size_t)blockIdx.x * ii * jj;
    abd += (std::size_t)blockIdx.x * n * n;
    x += (std::size_t)blockIdx.x * n;
    auto A = [&](int r, int c) { return abd[(std::size_t)(c - 1) * n + (r - 1)]; };
    for (int kk = 0; kk < n; ++kk) x[kk] = (double)qf[pidx(kk % nx + 1, kk / nx + 1, ii)];
    for (int j = 1; j <= n; ++j) {
        double s = x[j - 1];
        for (int k = 1; k < j; ++k) s -= A(k, j) * x[k - 1];
        x[j - 1] = s / A(j, j);
    }
    for (int j = n; j >= 1; --j) {
        double s = x[j - 1];
        for (int k = j + 1; k <= n; ++k) s -= A(j, k) * x[k - 1];
        x[j - 1] = s / A(j, j);
    }
    // mean removal, sums in FP64 and in Fortran order (kk runs i fastest)
    double qint = 0.0, cint = 0.0;
    for (int kk = 0; kk < n; ++kk) {
        qint = qint + x[kk];
        cint = cint + 1.0;
    }
    const double c = -qint / cint;
    for (int kk = 0; kk < n; ++kk) q[pidx(kk % nx + 1, kk / nx + 1, ii)] = (T)(x[kk] + c);
    auto Q = [&](int i, int j) -> T & { return q[pidx(i - 1, j - 1, ii)]; };  // 1-based
    if (per_y(jpn))
        for (int i = 2; i <= i1; ++i) {
            Q(i, jj) = Q(i, 2);
            Q(i, 1) = Q(i, j1);
        }
    if (per_x(jpn))
        for (int j = 2; j <= j1; ++j) {
            Q(ii, j) = Q(2, j);
            Q(1, j) = Q(i1, j);
        }
    if (jpn == 3 || jpn == -3) {
        Q(1, 1) = Q(i1, j1);
        Q(ii, 1) = Q(2, j1);
        Q(1, jj) = Q(i1, 2);
        Q(ii, jj) = Q(2, 2);
    }
}

// ---- periodic line relaxation: one lane per line, the reference's sequential DPTTRS order ----
// (src/2d/ftn/BMG2_SymStd_relax_lines_x.f90:170-300, _y.f90:176-310).  The right-hand side is
// formed in place (it reads only the neighbouring lines), solved with the line's LDL^T factors,
// and, when the line direction is periodic (sm), corrected by Sherman-Morrison with the solve
// z = T^-1 B of the wrap couplings, kept in the scratch z at the line's positions.
template <typename T, bool NINE, int DIR>
__global__ __launch_bounds__(64) void relax_lines_per_k(const T *__restrict__ so,
                                                        const T *__restrict__ qf, T *__restrict__ q,
                                                        const T *__restrict__ sor, T *__restrict__ z,
                                                        int ii, int jj, int first, int sm) {
    const int L = first + 2 * (int)(blockIdx.x * 64 + threadIdx.x);
    if (L > (DIR == 0 ? jj - 2 : ii - 2)) return;
    const std::size_t P = (std::size_t)ii * jj, bz = blockIdx.y;
    so += bz * (NINE ? 5 : 3) * P;
    qf += bz * P;
    q += bz * P;
    z += bz * P;
    sor += bz * 2 * P;
    const int n = DIR == 0 ? ii - 2 : jj - 2;
    auto pos = [&](int k) { return DIR == 0 ? pidx(k + 1, L, ii) : pidx(L, k + 1, ii); };
    auto Dk = [&](int k) { return DIR == 0 ? sor[pidx(k + 1, L, ii)] : sor[(std::size_t)L * jj + k + 1]; };
    auto Ek = [&](int k) {
        return DIR == 0 ? sor[P + pidx(k + 2, L, ii)] : sor[P + (std::size_t)L * jj + k + 2];
    };
    for (int k = 0; k < n; ++k) {
        const int i = DIR == 0 ? k + 1 : L, j = DIR == 0 ? L : k + 1;
        const std::size_t p = pidx(i, j, ii);
        T v;
        if (DIR == 0) {
            v = qf[p] + so[P * KS + p] * q[p - ii] + so[P * KS + p + ii] * q[p + ii];
        } else {
            v = qf[p] + so[P * KW + p] * q[p - 1] + so[P * KW + p + 1] * q[p + 1];
        }
        if (NINE)
            v = v + so[P * KSW + p] * q[p - ii - 1] + so[P * KNW + p + 1] * q[p - ii + 1] +
                so[P * KNW + p + ii] * q[p + ii - 1] + so[P * KSW + p + ii + 1] * q[p + ii + 1];
        q[p] = v;
    }
    auto dpttrs = [&](T *b) {
        if (n <= 1) {
            b[pos(0)] = b[pos(0)] * (T(1) / Dk(0));
            return;
        }
        for (int k = 1; k < n; ++k) b[pos(k)] = b[pos(k)] - b[pos(k - 1)] * Ek(k - 1);
        b[pos(n - 1)] = b[pos(n - 1)] / Dk(n - 1);
        for (int k = n - 2; k >= 0; --k) b[pos(k)] = b[pos(k)] / Dk(k) - b[pos(k + 1)] * Ek(k);
    };
    dpttrs(q);
    if (!sm) return;
    for (int k = 0; k < n; ++k) z[pos(k)] = T(0);
    if (DIR == 0) {
        z[pos(0)] = -so[P * KW + pidx(1, L, ii)];
        z[pos(n - 1)] = -so[P * KW + pidx(ii - 1, L, ii)];
    } else {
        z[pos(0)] = -so[P * KS + pidx(L, 1, ii)];
        z[pos(n - 1)] = -so[P * KS + pidx(L, jj - 1, ii)];
    }
    dpttrs(z);
    const T alpha = z[pos(0)] + z[pos(n - 1)];
    T beta = q[pos(0)] + q[pos(n - 1)];
    beta = beta / (T(1) + alpha);
    for (int k = 0; k < n; ++k) q[pos(k)] = q[pos(k)] - beta * z[pos(k)];
}

}  // namespace

template <typename T>
void relax_gs_per(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                  int updown, int jpn, hipStream_t s, int nb) {
    if (nstncl == 5) {
        // per colour (i parity, j parity) as in relax_gs; after a colour the x images of the rows
        // it swept (the reference refreshes them row by row, src/2d/ftn/BMG2_SymStd_relax_GS.f90
        // :176-182)
        static const int order[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
        const dim3 g = gridb(ii / 2 + 1, jj / 2 + 1, nb);
        for (int t = 0; t < 4; ++t) {
            const int c = updown == 0 ? t : 3 - t;
            gs_colour_k<T, true><<<g, 256, 0, s>>>(so, qf, q, sor, ii, jj, order[c][0], order[c][1], 1);
            if (per_x(jpn)) ghost_x<T>(q, ii, jj, order[c][1] == 1 ? 1 : 2, jj - 2, 2, s, nb);
        }
    } else {
        const dim3 g = gridb(ii / 2 + 1, jj - 2, nb);
        for (int t = 0; t < 2; ++t) {
            const int parity = updown == 0 ? t : 1 - t;
            gs_colour_k<T, false><<<g, 256, 0, s>>>(so, qf, q, sor, ii, jj, parity, 0, 1);
            if (per_x(jpn)) ghost_x<T>(q, ii, jj, 1, jj - 2, 1, s, nb);
        }
    }
    if (per_y(jpn)) ghost_y<T>(q, ii, jj, 0, ii - 1, s, nb);
}

template <typename T>
void relax_lines_per(int dir, const T *so, int nstncl, const T *qf, T *q, const T *sor, T *work,
                     int ii, int jj, int updown, int jpn, hipStream_t s, int nb) {
    // DOWN relaxes the lines of odd Fortran index 3, 5, ... (C index 2, 4, ...) first
    const bool sm = dir == 0 ? per_x(jpn) : per_y(jpn);
    const int nl = dir == 0 ? jj - 2 : ii - 2;
    for (int pass = 0; pass < 2; ++pass) {
        const int first = (pass == 0) == (updown == 0) ? 2 : 1;
        const int cnt = first > nl ? 0 : (nl - first) / 2 + 1;
        if (cnt > 0) {
            const dim3 g((unsigned)((cnt + 63) / 64), (unsigned)nb);
            if (dir == 0) {
                if (nstncl == 5)
                    relax_lines_per_k<T, true, 0><<<g, 64, 0, s>>>(so, qf, q, sor, work, ii, jj, first, sm);
                else
                    relax_lines_per_k<T, false, 0><<<g, 64, 0, s>>>(so, qf, q, sor, work, ii, jj, first, sm);
            } else {
                if (nstncl == 5)
                    relax_lines_per_k<T, true, 1><<<g, 64, 0, s>>>(so, qf, q, sor, work, ii, jj, first, sm);
                else
                    relax_lines_per_k<T, false, 1><<<g, 64, 0, s>>>(so, qf, q, sor, work, ii, jj, first, sm);
            }
        }
        if (sm) {  // images refreshed after each colour, y first
            if (per_y(jpn)) ghost_y<T>(q, ii, jj, 0, ii - 1, s, nb);
            if (per_x(jpn)) ghost_x<T>(q, ii, jj, 0, jj - 1, 1, s, nb);
        }
    }
    if (!sm) {  // lines across the periodic direction: its images after the sweep
        if (per_y(jpn)) ghost_y<T>(q, ii, jj, 0, ii - 1, s, nb);
        if (per_x(jpn)) ghost_x<T>(q, ii, jj, 0, jj - 1, 1, s, nb);
    }
}

void setup_lines_per(int dir, const double *so, int nstncl, double *sor, int ii, int jj, int jpn,
                     int *info, hipStream_t s, int nb) {
    if (dir == 0) {
        const int n = jj - 2;
        setup_lines_x_k<<<dim3((n + 63) / 64, nb), 64, 0, s>>>(so, sor, ii, jj, nstncl, info,
                                                               per_x(jpn) ? 1 : 0);
    } else {
        const int n = ii - 2;
        setup_lines_y_k<<<dim3((n + 63) / 64, nb), 64, 0, s>>>(so, sor, ii, jj, nstncl, info,
                                                               per_y(jpn) ? 1 : 0);
    }
}

template <typename T>
void restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int iic, int jjc, int jpn,
                  hipStream_t s, int nb) {
    // src/2d/ftn/BMG2_SymStd_restrict.f90:100-111: images of the fine grid refreshed in place
    // when the fine extent is even
    if (per_y(jpn) && jjf / 2 + 1 == jjc) ghost_y<T>(q, iif, jjf, 0, iif - 1, s, nb);
    if (per_x(jpn) && iif / 2 + 1 == iic) ghost_x<T>(q, iif, jjf, 0, jjf - 1, 1, s, nb);
    restrict_<T>(ci, q, qc, iif, jjf, iic, jjc, s, nb, 1);
}

template <typename T>
void interp_add_per(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                    int jjc, int iif, int jjf, int jpn, hipStream_t s, int nb) {
    interp_add<T>(q, qc, res, so, nstncl, ci, iic, jjc, iif, jjf, s, nb, 1);
    // src/2d/ftn/BMG2_SymStd_interp_add.f90:145-156
    if (per_y(jpn)) ghost_y<T>(q, iif, jjf, 0, iif - 1, s, nb);
    if (per_x(jpn)) ghost_x<T>(q, iif, jjf, 0, jjf - 1, 1, s, nb);
}

void setup_interp_per(const double *so, int nstncl, double *ci, int iif, int jjf, int iic, int jjc,
                      int jpn, hipStream_t s, int nb) {
    oi_bounds B;
    const int iicf = (iif - 2) / 2 + 3, jjcf = (jjf - 2) / 2 + 3;
    B.iifc = 2 * (iicf - 2) + 2;
    B.jjfc = 2 * (jjcf - 2) + 2;
    B.ibeg = 2; B.ibegc = 3; B.ibeg_x = 2; B.iend_x = iif - 2; B.iendc = iicf - 1;
    B.jendc = jjcf - 1;
    if (per_x(jpn)) {
        B.ibeg = 0;
        B.ibegc = 2;
        if (iic == iicf) {
            B.iendc = iicf;
            B.ibeg_x = 3;
            B.iend_x = iif - 1;
        }
    }
    B.jbeg = 2; B.jbegc = 3; B.jbeg_y = 2; B.jend_y = jjf - 2;
    if (per_y(jpn)) {
        B.jbeg = 0;
        B.jbegc = 2;
        if (jjc == jjcf) {
            B.jendc = jjcf;
            B.jbeg_y = 3;
            B.jend_y = jjf - 1;
        }
    }
    const int ex = per_y(jpn) ? 2 : 0, ey = per_x(jpn) ? 2 : 0;
    const int nic = B.iendc - B.ibegc + 1, njc = B.jendc - B.jbegc + 1;
    const dim3 gx = gridb(nic, jjc - 2 + ex, nb), gy = gridb(iic - 2 + ey, njc, nb);
    const dim3 gc = gridb(nic, njc, nb);
    if (nstncl == 5) {
        if (nic > 0) oi_per_xedge_k<true><<<gx, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B, ex);
        if (njc > 0) oi_per_yedge_k<true><<<gy, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B, ey);
        if (nic > 0 && njc > 0)
            oi_per_corner_k<true><<<gc, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B);
    } else {
        if (nic > 0) oi_per_xedge_k<false><<<gx, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B, ex);
        if (njc > 0) oi_per_yedge_k<false><<<gy, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B, ey);
        if (nic > 0 && njc > 0)
            oi_per_corner_k<false><<<gc, 256, 0, s>>>(so, ci, iif, jjf, iic, jjc, B);
    }
}

void setup_galerkin_per(const double *so, int nstncl, const double *ci, double *soc, int iif,
                        int jjf, int iic, int jjc, int jpn, hipStream_t s, int nb) {
    const dim3 g = gridb(iic - 2, jjc - 2, nb);
    if (nstncl == 5)
        galerkin_itli_k<true><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, iic, jjc);
    else
        galerkin_itli_k<false><<<g, 256, 0, s>>>(so, ci, soc, iif, jjf, iic, jjc);
    if (per_y(jpn))
        soc_ghost_y_k<<<dim3((unsigned)((iic + 255) / 256), 1, (unsigned)nb), 256, 0, s>>>(soc, iic, jjc);
    if (per_x(jpn))
        soc_ghost_x_k<<<dim3((unsigned)((jjc + 255) / 256), 1, (unsigned)nb), 256, 0, s>>>(soc, iic, jjc);
}

void setup_cg_per(const double *so, int nstncl, double *abd, int ii, int jj, int jpn, int *info,
                  hipStream_t s, int nb) {
    setup_cg_per_k<<<nb, 256, 0, s>>>(so, nstncl, abd, ii, jj, jpn, info);
}

template <typename T>
void solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int jpn,
                  hipStream_t s, int nb) {
    solve_cg_per_k<T><<<nb, 64, 0, s>>>(q, qf, abd, bbd, ii, jj, jpn);
}

template <typename T>
void periodic_images(T *q, int ii, int jj, int jpn, hipStream_t s, int nb) {
    if (per_y(jpn)) ghost_y<T>(q, ii, jj, 0, ii - 1, s, nb);
    if (per_x(jpn)) ghost_x<T>(q, ii, jj, 0, jj - 1, 1, s, nb);
}

template <typename T>
void copy(const T *src, T *dst, std::size_t n, hipStream_t s) {
    convert_k<T, T><<<grid1(n), 256, 0, s>>>(src, dst, n);
}

// explicit instantiations
#define CDR_INST(T)                                                                               \
    template void residual<T>(const T *, int, const T *, const T *, T *, int, int, hipStream_t,   \
                              int, int);                                                          \
    template void residual_norm<T>(const T *, int, const T *, const T *, T *, int, int, double *, \
                                   double *, hipStream_t, int, int);                              \
    template void sumsq<T>(const T *, int, int, double *, double *, hipStream_t, int);            \
    template void relax_gs<T>(const T *, int, const T *, T *, const T *, int, int, int,           \
                              hipStream_t, int, int);                                             \
    template void relax_lines_x<T>(const T *, int, const T *, T *, const T *, int, int, int,      \
                                   hipStream_t, int, int, int);                                   \
    template void relax_lines_y<T>(const T *, int, const T *, T *, const T *, int, int, int,      \
                                   hipStream_t, int, int);                                        \
    template void relax_lines_xy<T>(const T *, int, const T *, T *, const T *, const T *, T *, int,\
                                    int, int, hipStream_t, int, int);                             \
    template int tune_lines<T>(int, const T *, int, const T *, T *, const T *, int, int,          \
                               hipStream_t, int, int);                                            \
    template void transpose<T>(const T *, T *, int, int, hipStream_t, int);                       \
    template void coarse_cycle<T>(const coarse_level<T> *, int, int, int, int, int, const double *, \
                                  double *, hipStream_t, int, int, std::size_t, int);             \
    template void transpose_op<T>(const T *, T *, int, int, int, hipStream_t, int);               \
    template bool relax_lines_y_auto<T>(const T *, const T *, int, const T *, T *, const T *, T *, \
                                        T *, int, int, int, hipStream_t, int, bool, bool, bool,   \
                                        int);                                                     \
    template int uses_transposed_y<T>(const T *, int, int, int, int);                           \
    template void residual_qT<T>(const T *, int, const T *, const T *, T *, int, int, hipStream_t, \
                                 int, int);                                                       \
    template void interp_add_qT<T>(T *, const T *, const T *, const T *, int, const T *, int, int, \
                                   int, int, hipStream_t, int, int);                              \
    template int tune_lines_ytr<T>(const T *, const T *, int, const T *, T *, const T *, T *, T *, \
                                   int, int, hipStream_t, int, int, int);                         \
    template void restrict_<T>(const T *, const T *, T *, int, int, int, int, hipStream_t, int,   \
                               int);                                                              \
    template void restrict_cycle<T>(const T *, const T *, T *, T *, T *, int, int, int, int,       \
                                    hipStream_t, int, int);                                       \
    template void interp_add<T>(T *, const T *, const T *, const T *, int, const T *, int, int,   \
                                int, int, hipStream_t, int, int);                                 \
    template void solve_cg<T>(T *, const T *, const double *, double *, int, int, hipStream_t,    \
                              int, int);                                                          \
    template void fill<T>(T *, T, std::size_t, hipStream_t);                                     \
    template void relax_gs_per<T>(const T *, int, const T *, T *, const T *, int, int, int, int,  \
                                  hipStream_t, int);                                              \
    template void restrict_per<T>(const T *, T *, T *, int, int, int, int, int, hipStream_t, int);\
    template void interp_add_per<T>(T *, const T *, const T *, const T *, int, const T *, int,    \
                                    int, int, int, int, hipStream_t, int);                        \
    template void solve_cg_per<T>(T *, const T *, const double *, double *, int, int, int,        \
                                  hipStream_t, int);                                              \
    template void copy<T>(const T *, T *, std::size_t, hipStream_t);                              \
    template void relax_lines_per<T>(int, const T *, int, const T *, T *, const T *, T *, int, int, \
                                     int, int, hipStream_t, int);                                 \
    template void periodic_images<T>(T *, int, int, int, hipStream_t, int);
CDR_INST(double)
CDR_INST(float)
#undef CDR_INST
template void fill<int>(int *, int, std::size_t, hipStream_t);
template void convert<double, float>(const double *, float *, std::size_t, hipStream_t);
template void convert<float, double>(const float *, double *, std::size_t, hipStream_t);

}  // namespace k2
}  // namespace cdr

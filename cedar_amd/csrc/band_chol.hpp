// Banded Cholesky (upper, LAPACK band storage) on the device, for the coarsest-grid solves.
//
// AB(r, c) (1-based) holds A(c - kd - 1 + r, c) for max(1, c - kd) <= row <= c, leading
// dimension ld >= kd + 1, diagonal in row kd + 1.  The factorisation follows DPBTF2 (column
// j: sqrt of the pivot, scale the row by 1/pivot, symmetric rank-1 update of the trailing band
// (DSYR, upper)); the solve follows DPBTRS (DTBSV with U^T, then with U).  LAPACK's blocked
// DPBTRF takes the same steps while kd < 32 (its block size), which covers every coarsest grid
// the level rules produce.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cdr {

// All threads of the workgroup call this; returns the first failing column (0 if SPD) to every
// thread.  `bad` must be a __shared__ int.
__device__ inline int band_factor_upper(double *__restrict__ abd, int n, int kd, int ld, int &bad) {
    auto AB = [&](int r, int c) -> double & { return abd[(std::size_t)(c - 1) * ld + (r - 1)]; };
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 1; j <= n; ++j) {
        if (threadIdx.x == 0) {
            double ajj = AB(kd + 1, j);
            if (!(ajj > 0.0)) {
                bad = j;
            } else {
                ajj = sqrt(ajj);
                AB(kd + 1, j) = ajj;
                const int kn = (kd < n - j) ? kd : n - j;
                const double r = 1.0 / ajj;
                for (int t = 1; t <= kn; ++t) AB(kd + 1 - t, j + t) *= r;
            }
        }
        __syncthreads();
        if (bad) break;
        const int kn = (kd < n - j) ? kd : n - j;
        // DSYR upper: M(p,q) += x(p) * (-x(q)), p <= q, for q = 1..kn
        const int npairs = kn * (kn + 1) / 2;
        for (int t = threadIdx.x; t < npairs; t += blockDim.x) {
            int qq = 1, rem = t;
            while (rem >= qq) {
                rem -= qq;
                ++qq;
            }
            const int pp = rem + 1;
            const double xq = AB(kd + 1 - qq, j + qq);
            if (xq != 0.0) {
                const double temp = -1.0 * xq;
                AB(kd + 1 + pp - qq, j + qq) =
                    AB(kd + 1 + pp - qq, j + qq) + AB(kd + 1 - pp, j + pp) * temp;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    return bad;
}

// One thread: x <- A^{-1} x with the factor from band_factor_upper.
__device__ inline void band_solve_upper(const double *__restrict__ abd, int n, int kd, int ld,
                                        double *__restrict__ x) {
    auto AB = [&](int r, int c) { return abd[(std::size_t)(c - 1) * ld + (r - 1)]; };
    const int kp1 = kd + 1;
    for (int j = 1; j <= n; ++j) {  // U^T y = x
        double temp = x[j - 1];
        const int l = kp1 - j;
        const int i0 = (j - kd > 1) ? j - kd : 1;
        for (int i = i0; i <= j - 1; ++i) temp = temp - AB(l + i, j) * x[i - 1];
        x[j - 1] = temp / AB(kp1, j);
    }
    for (int j = n; j >= 1; --j) {  // U x = y
        if (x[j - 1] != 0.0) {
            const int l = kp1 - j;
            x[j - 1] = x[j - 1] / AB(kp1, j);
            const double temp = x[j - 1];
            const int i0 = (j - kd > 1) ? j - kd : 1;
            for (int i = j - 1; i >= i0; --i) x[i - 1] = x[i - 1] - temp * AB(l + i, j);
        }
    }
}

// The two triangular sweeps of band_solve_upper by one wavefront (lane 0..63, kd <= 63), with x
// in LDS: as soon as an unknown is final, lanes t = 1..kd apply its column to the kd unknowns
// after it (U^T sweep) or before it (U sweep).  Each unknown receives the same terms in the same
// order as in band_solve_upper: the U^T sweep's row sums subtract U(i, c) y_i for increasing i,
// exactly the order in which the column updates arrive, and its U sweep is column-oriented
// already; so the result is bitwise that of band_solve_upper.  The dependency chain is one
// division per unknown instead of a row sum.  LDS accesses of one wavefront complete in issue
// order, which orders each step's reads after the previous step's writes.
__device__ inline void band_solve_upper_wave(const double *__restrict__ abd, int n, int kd, int ld,
                                             double *__restrict__ x, int lane) {
    auto AB = [&](int r, int c) { return abd[(std::size_t)(c - 1) * ld + (r - 1)]; };
    const int kp1 = kd + 1;
    for (int j = 1; j <= n; ++j) {  // U^T y = x
        const double y = x[j - 1] / AB(kp1, j);
        const int c = j + lane;
        if (lane == 0) x[j - 1] = y;
        if (lane >= 1 && lane <= kd && c <= n) x[c - 1] = x[c - 1] - AB(kp1 - lane, c) * y;
    }
    for (int j = n; j >= 1; --j) {  // U x = y
        const double xj = x[j - 1];
        if (xj != 0.0) {
            const double t = xj / AB(kp1, j);
            const int i = j - lane;
            if (lane == 0) x[j - 1] = t;
            if (lane >= 1 && lane <= kd && i >= 1) x[i - 1] = x[i - 1] - t * AB(kp1 - lane, j);
        }
    }
}

}  // namespace cdr

// Host-side launchers for the 2D BoxMG kernels (bmg2_kernels.hip).
//
// Every launcher takes device pointers in the solver's HBM layout -- Fortran order, one ghost
// layer, ghost-inclusive extents ii = nx+2, jj = ny+2, stencils as nstncl contiguous planes of
// ii*jj values {c,w,s[,sw,nw]} -- and enqueues work on `stream`.  No launcher allocates or
// synchronises; argument checking happens in the C ABI (capi2d.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cdr {
namespace k2 {

// stencil planes {c, w, s, sw, nw}
enum { KO = 0, KW = 1, KS = 2, KSW = 3, KNW = 4 };
// prolongation planes {L, R, A, B, SW, NW, NE, SE}
enum { LL = 0, LR = 1, LA = 2, LB = 3, LSW = 4, LNW = 5, LNE = 6, LSE = 7 };

constexpr int kNormBlock = 256;  // points per block in the fused residual/norm kernels

// number of per-block partial sums written by the norm-producing launchers for an ii x jj grid
std::size_t norm_partials(int ii, int jj);

template <typename T>
void residual(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj,
              hipStream_t s);
// residual + per-block partial sums of r^2 (FP64) + final reduction into *sumsq (device)
template <typename T>
void residual_norm(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj,
                   double *partials, double *sumsq, hipStream_t s);
// sum of squares of the interior of x into *sumsq (device), deterministic two-stage reduction
template <typename T>
void sumsq(const T *x, int ii, int jj, double *partials, double *out, hipStream_t s);

// Mixed-precision outer step: xnew = x + (double)e (e may be null: xnew = x),
// r = b - A xnew in FP64, r32 = (float) r, partial sums of r^2 into *sumsq.
void residual_ir(const double *so, int nstncl, const double *qf, const double *x, const float *e,
                 double *xnew, float *r32, int ii, int jj, double *partials, double *sumsq,
                 hipStream_t s);

void setup_recip(const double *so, double *sor, int ii, int jj, hipStream_t s);
template <typename T>
void relax_gs(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
              int updown, hipStream_t s);

// LDL^T factors of every line (x: rows, y: columns, stored transposed); info gets the first
// failing pivot (1-based line-local index), 0 if all lines are SPD.
void setup_lines_x(const double *so, double *sor, int ii, int jj, int *info, hipStream_t s);
void setup_lines_y(const double *so, double *sor, int ii, int jj, int *info, hipStream_t s);
template <typename T>
void relax_lines_x(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s);
template <typename T>
void relax_lines_y(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s);

template <typename T>
void restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, hipStream_t s);
template <typename T>
void interp_add(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                int jjc, int iif, int jjf, hipStream_t s);

void setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int iic, int jjc,
                  hipStream_t s);
void setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif, int jjf,
                    int iic, int jjc, hipStream_t s);

// coarsest grid: band ABD(ii, nx*ny) (upper, kd = ii-1), info on device
void setup_cg(const double *so, int nstncl, double *abd, int ii, int jj, int *info, hipStream_t s);
template <typename T>
void solve_cg(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, hipStream_t s);

// elementwise helpers
template <typename Tin, typename Tout>
void convert(const Tin *in, Tout *out, std::size_t n, hipStream_t s);
template <typename T>
void fill(T *x, T v, std::size_t n, hipStream_t s);
// x += (double) e over the interior
void add_correction(double *x, const float *e, int ii, int jj, hipStream_t s);

}  // namespace k2
}  // namespace cdr

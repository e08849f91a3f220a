// Host-side launchers for the 3D BoxMG kernels (bmg3_kernels.hip).
//
// Device pointers in the solver's HBM layout: Fortran order, one ghost layer, ghost-inclusive
// extents ii = nx+2, jj = ny+2, kk = nz+2; a stencil is nstncl contiguous planes of ii*jj*kk
// values, 4 for the 7-point operator {p, pw, ps, b} and 14 for the 27-point operator
// {p, pw, ps, b, psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw}; the prolongation has 26 planes
// on the coarse grid.  No launcher allocates or synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cdr {
namespace k3 {

enum { XY = 0, XZ = 1, YZ = 2 };  // plane families

// per-block partial sums written by residual_norm for an ii x jj x kk grid
std::size_t norm_partials(int ii, int jj, int kk);

template <typename T>
void residual(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj, int kk,
              hipStream_t s);
// residual + deterministic sum of r^2 over the interior into *sumsq (device)
void residual_norm(const double *so, int nstncl, const double *qf, const double *q, double *res,
                   int ii, int jj, int kk, double *partials, double *sumsq, hipStream_t s);

// mixed-precision outer step: xnew = x + (double) e (e may be null), r = b - A xnew in FP64,
// r32 = (float) r, sum of r^2 into *sumsq (device)
void residual_ir(const double *so, int nstncl, const double *qf, const double *x, const float *e,
                 double *xnew, float *r32, int ii, int jj, int kk, double *partials,
                 double *sumsq, hipStream_t s);

void setup_recip(const double *so, int nstncl, double *sor, int ii, int jj, int kk,
                 hipStream_t s);
template <typename T>
void relax_gs(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj, int kk,
              int updown, hipStream_t s);

template <typename T>
void restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc,
               int kkc, hipStream_t s);
// q += P qc + res / SO_p (res is read, not modified)
template <typename T>
void interp_add(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                int jjc, int kkc, int iif, int jjf, int kkf, hipStream_t s);

void setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf, int iic,
                  int jjc, int kkc, hipStream_t s);
void setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif, int jjf,
                    int kkf, int iic, int jjc, int kkc, hipStream_t s);

// coarsest grid: band ABD(ldab, nx*ny*nz) with kd = cg_bandwidth(ii, jj), ldab >= kd + 1;
// *info (zero on entry) is raised to the failing column when a pivot is not positive
int cg_bandwidth(int ii, int jj);
void setup_cg(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk,
              int *info, hipStream_t s);
template <typename T>
void solve_cg(T *q, const T *qf, const double *abd, int ldab, double *bbd, int ii, int jj, int kk,
              hipStream_t s);

// Periodic boundaries.  jpn is the BoxMG boundary id of the periodic mask (bc_from_mask3:
// bits x = 1, y = 2, z = 4 -> 2, 1, 3, 5, 6, 7, 8); per_mask3 maps it back.  Images live in the
// ghost layer of a periodic direction.  order_xyz: images refreshed x, y, z (else y, x, z).
int per_mask3(int jpn);
int bc_from_mask3(int mask);
template <typename T>
void images(T *q, int ii, int jj, int kk, int jpn, int order_xyz, hipStream_t s);
template <typename T>
void relax_gs_per(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                  int kk, int updown, int jpn, hipStream_t s);
// refreshes the images of q in place first
template <typename T>
void restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int kkf, int iic, int jjc, int kkc,
                  int jpn, hipStream_t s);
template <typename T>
void interp_add_per(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci,
                    int iic, int jjc, int kkc, int iif, int jjf, int kkf, int jpn, hipStream_t s);
// ci must be zero on entry (the periodic loops write only the entries the reference writes)
void setup_interp_per(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf, int iic,
                      int jjc, int kkc, int jpn, hipStream_t s);
void setup_galerkin_per(const double *so, int nstncl, const double *ci, double *soc, int iif,
                        int jjf, int kkf, int iic, int jjc, int kkc, int jpn, hipStream_t s);
// dense n x n factor (n = nx ny nz, column-major, upper triangle); *info as for setup_cg
void setup_cg_per(const double *so, int nstncl, double *abd, int ii, int jj, int kk, int jpn,
                  int *info, hipStream_t s);
template <typename T>
void solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int kk,
                  int jpn, hipStream_t s);

// Plane relaxation pieces for the planes first, first+2, ... (npl of them) of family dir; the
// npl 2D arrays of the batch are stored back to back, plane b holding plane first + 2b.
// Plane extents: XY (ii, jj), XZ (ii, kk), YZ (jj, kk).
void copy_coeff(int dir, const double *so3, int nstncl3, double *so2, int first, int npl, int ii,
                int jj, int kk, hipStream_t s);
template <typename T>
void copy32(int dir, const T *x, T *x2, int first, int npl, int ii, int jj, int kk, hipStream_t s);
template <typename T>
void copy23(int dir, const T *x2, T *x, int first, int npl, int ii, int jj, int kk, hipStream_t s);
template <typename T>
void copy_rhs(int dir, const T *so3, int nstncl3, const T *x, const T *b, T *b2, int first,
              int npl, int ii, int jj, int kk, hipStream_t s);

}  // namespace k3
}  // namespace cdr

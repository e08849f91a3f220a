// Checked, batch-capable forms of the 2D kernel entry points.
//
// Each function validates its arguments, enqueues the launcher(s) of bmg2_launch.hpp on
// `stream` for `nb` problems of identical extents stored back to back, and returns a cdr_status.
// The extern "C" functions of include/cedar_hip.h are these with nb = 1; the C++ solver layer
// calls them directly with its batch count (3D plane relaxation, batched solves).  `rep`:
// consecutive groups of rep problems share one operator (operator-side arrays hold nb / rep
// entries, problem b reads entry b / rep); nb must be a multiple of rep.
#pragma once

#include <hip/hip_runtime.h>

namespace cdr {
namespace api2 {

template <typename T>
int residual(const T *so, int ns, const T *qf, const T *q, T *res, int ii, int jj, hipStream_t s,
             int nb = 1, int rep = 1);
int residual_norm(const double *so, int ns, const double *qf, const double *q, double *res, int ii,
                  int jj, double *work, double *sumsq_dev, hipStream_t s, int nb = 1, int rep = 1);
int sumsq(const double *x, int ii, int jj, double *work, double *sumsq_dev, hipStream_t s,
          int nb = 1);
int setup_recip(const double *so, double *sor, int ii, int jj, int ns, hipStream_t s, int nb = 1);
template <typename T>
int relax_gs(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud,
             hipStream_t s, int nb = 1, int rep = 1);
// x-lines (dir 0) or y-lines (dir 1)
int setup_lines(int dir, const double *so, double *sor, int ii, int jj, int ns, int *info_dev,
                hipStream_t s, int nb = 1);
template <typename T>
int relax_lines(int dir, const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj,
                int ud, hipStream_t s, int nb = 1, int rep = 1);
// one line-xy smoothing step through the work grid (bmg2_launch.hpp relax_lines_xy)
template <typename T>
int relax_lines_xy(const T *so, int ns, const T *qf, T *q, const T *sorx, const T *sory, T *work,
                   int ii, int jj, int ud, hipStream_t s, int nb = 1, int rep = 1);
template <typename T>
int restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, hipStream_t s,
              int nb = 1, int rep = 1);
template <typename T>
int interp_add(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic, int jjc,
               int iif, int jjf, hipStream_t s, int nb = 1, int rep = 1);
int setup_interp(const double *so, int ns, double *ci, int iif, int jjf, int iic, int jjc,
                 hipStream_t s, int nb = 1);
int setup_galerkin(const double *so, int ns, const double *ci, double *soc, int iif, int jjf,
                   int iic, int jjc, hipStream_t s, int nb = 1);
int setup_cg(const double *so, int ns, double *abd, int ii, int jj, int *info_dev, hipStream_t s,
             int nb = 1);
template <typename T>
int solve_cg(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, hipStream_t s,
             int nb = 1, int rep = 1);

// periodic forms: jpn = BMG_get_bc(per_mask) in 0..3 (0 runs the Dirichlet kernel); the coarse
// matrix is dense, ABD(n, n) with n = nx*ny
template <typename T>
int relax_gs_per(const T *so, int ns, const T *qf, T *q, const T *sor, int ii, int jj, int ud,
                 int jpn, hipStream_t s, int nb = 1);
template <typename T>
int restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int iic, int jjc, int jpn,
                 hipStream_t s, int nb = 1);
template <typename T>
int interp_add_per(T *q, const T *qc, const T *res, const T *so, int ns, const T *ci, int iic,
                   int jjc, int iif, int jjf, int jpn, hipStream_t s, int nb = 1);
int setup_interp_per(const double *so, int ns, double *ci, int iif, int jjf, int iic, int jjc,
                     int jpn, hipStream_t s, int nb = 1);
int setup_galerkin_per(const double *so, int ns, const double *ci, double *soc, int iif, int jjf,
                       int iic, int jjc, int jpn, hipStream_t s, int nb = 1);
int setup_cg_per(const double *so, int ns, double *abd, int ii, int jj, int jpn, int *info_dev,
                 hipStream_t s, int nb = 1);
// lines (dir 0 x, 1 y); work: scratch grid, ii*jj values per problem
int setup_lines_per(int dir, const double *so, double *sor, int ii, int jj, int ns, int jpn,
                    int *info_dev, hipStream_t s, int nb = 1);
template <typename T>
int relax_lines_per(int dir, const T *so, int ns, const T *qf, T *q, const T *sor, T *work, int ii,
                    int jj, int ud, int jpn, hipStream_t s, int nb = 1);
template <typename T>
int solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int jpn,
                 hipStream_t s, int nb = 1);

}  // namespace api2
}  // namespace cdr

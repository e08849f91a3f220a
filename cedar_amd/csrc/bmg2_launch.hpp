// Host-side launchers for the 2D BoxMG kernels (bmg2_kernels.hip).
//
// Every launcher takes device pointers in the solver's HBM layout -- Fortran order, one ghost
// layer, ghost-inclusive extents ii = nx+2, jj = ny+2, stencils as nstncl contiguous planes of
// ii*jj values {c,w,s[,sw,nw]} -- and enqueues work on `stream`.  No launcher allocates or
// synchronises; argument checking happens in the C ABI (capi2d.hip).
//
// The trailing `nb` runs the same call on nb problems of identical extents stored back to back
// (each array of problem b starts at b times the size of one problem's array).  Where a launcher
// also takes `rep`, consecutive groups of rep problems share one operator: the operator-side
// arrays (stencil, interpolation, relaxation factors, coarse band factor) hold nb / rep entries
// and problem b reads entry b / rep, so several right-hand sides of one operator are relaxed,
// restricted and interpolated in the same launch while the operator is read once per group;
// per-problem
// outputs of the reductions (sumsq) are consecutive doubles, and `partials` holds
// nb * norm_partials(ii, jj) values.  Factorisation info words are shared by the batch: the
// kernels raise them with atomicMax, so they must start at 0 and end at 0 when every problem is
// SPD.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

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
              hipStream_t s, int nb = 1, int rep = 1);
// residual + per-block partial sums of r^2 (FP64) + final reduction into *sumsq (device)
template <typename T>
void residual_norm(const T *so, int nstncl, const T *qf, const T *q, T *res, int ii, int jj,
                   double *partials, double *sumsq, hipStream_t s, int nb = 1, int rep = 1);
// sum of squares of the interior of x into *sumsq (device), deterministic two-stage reduction
template <typename T>
void sumsq(const T *x, int ii, int jj, double *partials, double *out, hipStream_t s, int nb = 1);

// Mixed-precision outer step: xnew = x + (double)e (e may be null: xnew = x),
// r = b - A xnew in FP64, r32 = (float) r, partial sums of r^2 into *sumsq.  With `active`
// (nb ints, may be null) a problem b with active[b] == 0 keeps xnew = x.  With r32T non-null the
// interior of r32 is also written transposed (ii rows of jj values), for a finest FP32 level
// whose y sweeps run through the transposed grid.
void residual_ir(const double *so, int nstncl, const double *qf, const double *x, const float *e,
                 double *xnew, float *r32, int ii, int jj, double *partials, double *sumsq,
                 hipStream_t s, int nb = 1, const int *active = nullptr, int rep = 1,
                 float *r32T = nullptr);
// out[b] = sum of partials[b*n .. b*n+n-1] in a fixed order (deterministic), b < nb
void finalize_sum(const double *partials, std::size_t n, double *out, int nb, hipStream_t s);
// x[k] = xprev[k] over every problem b (per values each) with active[b] == 0
void keep_inactive(double *x, const double *xprev, const int *active, std::size_t per, int nb,
                   hipStream_t s);

void setup_recip(const double *so, int nstncl, double *sor, int ii, int jj, hipStream_t s,
                 int nb = 1);
template <typename T>
void relax_gs(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
              int updown, hipStream_t s, int nb = 1, int rep = 1);

// LDL^T factors of every line (x: rows, y: columns, stored transposed); info (zero on entry)
// is raised to the largest failing pivot index (1-based, line-local); it stays 0 when every
// line is SPD.
void setup_lines_x(const double *so, int nstncl, double *sor, int ii, int jj, int *info,
                   hipStream_t s, int nb = 1);
void setup_lines_y(const double *so, int nstncl, double *sor, int ii, int jj, int *info,
                   hipStream_t s, int nb = 1);
// zero_start: q is zero on entry (not read by the first colour; its zeros are the operands);
// honoured where lines_zero_start(ii), ignored otherwise
template <typename T>
void relax_lines_x(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s, int nb = 1, int rep = 1, int zero_start = 0);
bool lines_zero_start(int ii);
template <typename T>
void relax_lines_y(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                   int updown, hipStream_t s, int nb = 1, int rep = 1);

// One line-xy smoothing step (include/cedar/multilevel.h:175-178, 204-207): x-lines then y-lines
// on DOWN, y then x on UP.  `work` is unused (kept for callers that size it).
template <typename T>
void relax_lines_xy(const T *so, int nstncl, const T *qf, T *q, const T *sorx, const T *sory,
                    T *work, int ii, int jj, int updown, hipStream_t s, int nb = 1, int rep = 1);

// Line-kernel shape autotuner.  Every shape of the line kernels gives bitwise identical results;
// tune_lines times each candidate shape for one direction (0: x, 1: y) once on the given data
// (q is overwritten) and records the fastest for (precision, stencil, direction, ii, jj, nb) in a
// process-wide table that relax_lines_* consult.  It synchronises `s`; it returns the chosen
// shape index, or -1 on a HIP error.  With autotuning off (set_autotune(0), which returns the
// previous setting) nothing is timed or recorded and the default shape 0 runs.
template <typename T>
int tune_lines(int dir, const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
               hipStream_t s, int nb = 1, int rep = 1);
int set_autotune(int on);
bool autotune_enabled();

// One level of a fused coarse-level V-cycle: arrays of problem 0 of a batch (problem b at b times
// each array's size), 9-point operator, relaxation data as the solver stores it (point: sor0
// holds 1/c in its plane 1; lines: sor0 = x factors or the only direction's factors, sor1 = y
// factors of line-xy), ci = interpolation stored on this level (null on the first level).
template <typename T>
struct coarse_level {
    const T *so;
    T *x, *b, *res;
    const T *sor0, *sor1;
    const T *ci;
    int ii, jj;
};
// Runs one V-cycle over nlev levels (lv: device array, finest first; the last level is the
// coarsest, solved with the band factor abd) in one launch of one workgroup per problem, for x of
// the first level from its b (x zero on entry).  max_side: the largest nx or ny of the levels
// (at most 128).  relax: CDR_RELAX_*; npre / npost: smoothing steps.  Bitwise equal to the
// unfused V-cycle over the same levels.
// lds_bytes > 0 (at least fused_cycle_lds_bytes of the levels for lds_mode): lds_mode 1, every
// array of the fused levels is staged in LDS for the whole cycle (only the first level's b and x
// cross memory); lds_mode 2 or 3, the same except that of the first level only x, b and res (2)
// or x alone (3) are in LDS, its other arrays stay in memory.
template <typename T>
void coarse_cycle(const coarse_level<T> *lv, int nlev, int max_side, int relax, int npre,
                  int npost, const double *abd, double *bbd, hipStream_t s, int nb = 1,
                  int rep = 1, std::size_t lds_bytes = 0, int lds_mode = 1);
// LDS bytes that fused levels of these extents (first level first) need at word size wbytes
// with lds_mode 1, 2 or 3
std::size_t fused_cycle_lds_bytes(int wbytes, const int *ii, const int *jj, int nlev,
                                  int lds_mode = 1);

// dst (cols x rows) = transpose of src (rows x cols), row-major, per problem of a batch
template <typename T>
void transpose(const T *src, T *dst, int rows, int cols, hipStream_t s, int nb = 1);
// soT = the operator of the transposed grid: planes c, w, s, sw, nw = transposes of c, s, w, sw,
// nw of so (ii x jj per plane of so, jj x ii per plane of soT)
template <typename T>
void transpose_op(const T *so, T *soT, int nstncl, int ii, int jj, hipStream_t s, int nb = 1);
// One y-line sweep (relax_lines_y), run as an x-line sweep of the transposed grid (q and qf
// transposed into qT, qfT and q transposed back) when tune_lines_ytr found that faster for these
// extents; bitwise the same result either way.  soT may be null (direct y-lines).  With
// qfT_ready the caller guarantees that qfT already holds the transpose of qf.  Returns whether
// the transposed path ran (qfT then holds the transpose of qf).
// q_in_T: qT already holds the iterate (q is stale; no transpose in); keep_T: leave the result in
// qT only (no transpose back; q stays stale).  Both apply only where the transposed path runs.
template <typename T>
bool relax_lines_y_auto(const T *so, const T *soT, int nstncl, const T *qf, T *q, const T *sory,
                        T *qT, T *qfT, int ii, int jj, int updown, hipStream_t s, int nb = 1,
                        bool qfT_ready = false, bool q_in_T = false, bool keep_T = false,
                        int rep = 1);
// The shape the transposed y path runs with at these extents, or -1 where the direct y-line
// kernel runs (no transposed operator, or tuning chose it).
template <typename T>
int uses_transposed_y(const T *soT, int nstncl, int ii, int jj, int nb);
// residual and interp_add with the fine iterate in the transposed layout qT (ii rows of jj
// values); the other arrays in the normal layout.  Bitwise equal to residual / interp_add on the
// untransposed iterate.  interp_add_qT updates qT in place.
template <typename T>
void residual_qT(const T *so, int nstncl, const T *qf, const T *qT, T *res, int ii, int jj,
                 hipStream_t s, int nb = 1, int rep = 1);
template <typename T>
void interp_add_qT(T *qT, const T *qc, const T *res, const T *so, int nstncl, const T *ci,
                   int iic, int jjc, int iif, int jjf, hipStream_t s, int nb = 1, int rep = 1);
// Times the direct y-lines against the transposed path (each x-line shape) once and records the
// choice (-1: direct); `sweeps` y sweeps are timed per right-hand side (one transpose of qf), as
// a level visit of a V(nu1, nu2) cycle makes nu1 + nu2 of them.  Overwrites q, qT, qfT;
// synchronises s; returns the choice or -2 on a HIP error.  A no-op returning -1 with
// autotuning off.
template <typename T>
int tune_lines_ytr(const T *so, const T *soT, int nstncl, const T *qf, T *q, const T *sory, T *qT,
                   T *qfT, int ii, int jj, hipStream_t s, int nb = 1, int sweeps = 2, int rep = 1);
// number of candidate shapes for lines of the given direction, word size (4 or 8) and extents
int line_shape_count(int dir, int wbytes, int nstncl, int ii, int jj);
// records `shape` for (direction, word size, stencil, ii, jj, nb) (shape < 0: erases the entry)
void set_line_shape(int dir, int wbytes, int nstncl, int ii, int jj, int nb, int shape);
// JSON list of the recorded choices
std::string tune_report();

template <typename T>
void restrict_(const T *ci, const T *q, T *qc, int iif, int jjf, int iic, int jjc, hipStream_t s,
               int nb = 1, int rep = 1);
template <typename T>
void interp_add(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                int jjc, int iif, int jjf, hipStream_t s, int nb = 1, int rep = 1);
// qc = P^T q (bitwise restrict_), xc = 0 over the whole coarse grid, and, with qcT non-null, qc
// also written transposed (qcT: iic rows of jjc values, interior only): the restriction step of
// a V-cycle in one launch
template <typename T>
void restrict_cycle(const T *ci, const T *q, T *qc, T *xc, T *qcT, int iif, int jjf, int iic,
                    int jjc, hipStream_t s, int nb = 1, int rep = 1);

void setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int iic, int jjc,
                  hipStream_t s, int nb = 1);
void setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif, int jjf,
                    int iic, int jjc, hipStream_t s, int nb = 1);

// coarsest grid: band ABD(ii, nx*ny) (upper, kd = ii-1); info on device, zero on entry, gets
// the failing column when a pivot is not positive
void setup_cg(const double *so, int nstncl, double *abd, int ii, int jj, int *info, hipStream_t s, int nb = 1);
template <typename T>
void solve_cg(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, hipStream_t s,
              int nb = 1, int rep = 1);

// Periodic boundaries: jpn = BMG_get_bc(per_mask) (2 periodic in x, 1 in y, 3 both; 0 runs the
// Dirichlet launcher).  Each is the Dirichlet kernel plus the reference's periodic images in the
// ghost rows/columns; setup_interp_per and setup_galerkin_per are the reference's periodic loop
// structure and hand-expanded products; the coarse solve is a dense Cholesky of the n x n
// periodic matrix (ABD(n, n), n = nx*ny) with the mean of the solution removed.
template <typename T>
void relax_gs_per(const T *so, int nstncl, const T *qf, T *q, const T *sor, int ii, int jj,
                  int updown, int jpn, hipStream_t s, int nb = 1);
// lines (dir 0 x, 1 y): a periodic line direction gets the wrap couplings on its end diagonals
// and a Sherman-Morrison line solve (work: scratch grid of ii*jj per problem); lines across a
// periodic direction refresh its images after the sweep
void setup_lines_per(int dir, const double *so, int nstncl, double *sor, int ii, int jj, int jpn,
                     int *info, hipStream_t s, int nb = 1);
template <typename T>
void relax_lines_per(int dir, const T *so, int nstncl, const T *qf, T *q, const T *sor, T *work,
                     int ii, int jj, int updown, int jpn, hipStream_t s, int nb = 1);
// refreshes the images of q in place first when the fine extent is even (as the reference does)
template <typename T>
void restrict_per(const T *ci, T *q, T *qc, int iif, int jjf, int iic, int jjc, int jpn,
                  hipStream_t s, int nb = 1);
template <typename T>
void interp_add_per(T *q, const T *qc, const T *res, const T *so, int nstncl, const T *ci, int iic,
                    int jjc, int iif, int jjf, int jpn, hipStream_t s, int nb = 1);
void setup_interp_per(const double *so, int nstncl, double *ci, int iif, int jjf, int iic, int jjc,
                      int jpn, hipStream_t s, int nb = 1);
void setup_galerkin_per(const double *so, int nstncl, const double *ci, double *soc, int iif,
                        int jjf, int iic, int jjc, int jpn, hipStream_t s, int nb = 1);
void setup_cg_per(const double *so, int nstncl, double *abd, int ii, int jj, int jpn, int *info,
                  hipStream_t s, int nb = 1);
template <typename T>
void solve_cg_per(T *q, const T *qf, const double *abd, double *bbd, int ii, int jj, int jpn,
                  hipStream_t s, int nb = 1);

// periodic images of every ghost row (all columns) and then every ghost column (all rows)
template <typename T>
void periodic_images(T *q, int ii, int jj, int jpn, hipStream_t s, int nb = 1);

// elementwise helpers
template <typename T>
void copy(const T *src, T *dst, std::size_t n, hipStream_t s);
template <typename Tin, typename Tout>
void convert(const Tin *in, Tout *out, std::size_t n, hipStream_t s);
template <typename T>
void fill(T *x, T v, std::size_t n, hipStream_t s);
// x += (double) e over the interior
void add_correction(double *x, const float *e, int ii, int jj, hipStream_t s, int nb = 1);

}  // namespace k2
}  // namespace cdr

/*
 * cedar_hip.h -- C ABI of the MI355X (gfx950) BoxMG kernels and solver.
 *
 * Drop-in for Cedar's serial 2D kernel set: every kernel entry point below replaces one
 * BIND(C) Fortran subroutine that Cedar's C++ adapters call (cited per function), with the same
 * data layout and argument meaning, but operating on DEVICE pointers and enqueued on a HIP
 * stream.  Layout (unchanged from the reference):
 *   - grid functions: Fortran order, one ghost layer, ghost-inclusive extents ii = nx+2, jj = ny+2;
 *   - stencils: nstncl contiguous planes of ii*jj values, 3 planes {c,w,s} (5-point) or
 *     5 planes {c,w,s,sw,nw} (9-point), off-diagonals stored negated;
 *   - prolongation CI: 8 planes {L,R,A,B,SW,NW,NE,SE} on the coarse grid (iic x jjc);
 *   - relaxation stencil SOR: 2 planes of ii*jj; point GS keeps 1/c in plane 2 (msor); x-line
 *     factors (D in plane 1, L multipliers in plane 2) as in the reference, y-line factors
 *     stored transposed (jj x ii per plane).
 * Every function returns a cdr_status; none throws.  Kernel entry points do not allocate and
 * do not synchronise the stream unless stated ("synchronises").  `stream` is a hipStream_t
 * (NULL = the default stream).  Suffix _f64 computes in double, _f32 in float.
 *
 * Boundary conditions: Dirichlet (Cedar's BMG_BCs_definite, per_mask = 0).
 */
#ifndef CEDAR_HIP_H
#define CEDAR_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CDR_OK = 0,
    CDR_ERR_ARG = 1,         /* bad pointer, extent, stencil width or option */
    CDR_ERR_HIP = 2,         /* a HIP runtime call or kernel launch failed */
    CDR_ERR_NOT_SPD = 3,     /* a factorisation met a non-positive pivot (LAPACK INFO > 0) */
    CDR_ERR_NONFINITE = 4,   /* NaN or Inf in a residual norm */
    CDR_ERR_NO_DEVICE = 5,   /* no usable GPU */
    CDR_ERR_LEVELS = 6,      /* num_levels larger than the grid allows */
    CDR_ERR_STALLED = 7,     /* the solve stopped reducing the residual */
    CDR_ERR_TOO_LARGE = 8    /* the input needs more than a fixed-size library buffer holds */
} cdr_status;

enum { CDR_DOWN = 0, CDR_UP = 1 };  /* cycle::Dir, BMG_DOWN / BMG_UP */

const char *cdr_hip_status_string(int status);
/* Library version string ("cedar_amd <version> gfx950"). */
const char *cdr_hip_version(void);

/* Line-kernel shape autotuning (no counterpart in the reference, whose line solves run on one
 * CPU core).  The line relaxation kernels have several launch shapes (chunks per wavefront,
 * wavefronts per line, lines or columns per workgroup) that all give bitwise identical results.
 * With autotuning on (the default), solver setup times each candidate shape once per level and
 * direction and records the fastest in a process-wide table used by every later launch of the
 * same precision, stencil, direction, extents and batch size.  cdr_hip_autotune(0) turns it off
 * (shape 0 runs where nothing is recorded); returns the previous setting.
 * cdr_hip_autotune_report writes the recorded choices as a JSON list into buf (len bytes,
 * NUL-terminated, truncated if short) and returns the full length. */
int cdr_hip_autotune(int on);
int cdr_hip_autotune_report(char *buf, int len);
/* Number of candidate shapes for lines of direction dir (0: x, 1: y) at fp_bits (32 or 64)
 * precision, stencil nstncl and extents ii x jj; cdr_hip_set_line_shape2d pins one of them for
 * those extents and batch size nb (shape -1 removes the entry; returns a cdr_status).
 * dir 2 is the line-xy smoother's y sweep run as an x-line sweep of the transposed grid (used by
 * the solvers where tuning found it faster; the same result bit for bit): its shapes are x-line
 * shapes for the transposed extents, and a pinned shape makes the solvers take that path.
 * cdr_hip_line_shapes2d returns -CDR_ERR_ARG for invalid arguments. */
int cdr_hip_line_shapes2d(int dir, int fp_bits, int nstncl, int ii, int jj);
int cdr_hip_set_line_shape2d(int dir, int fp_bits, int nstncl, int ii, int jj, int nb, int shape);

/* ---------------------------------------------------------------------------------------
 * 2D kernels (FP64 and FP32 variants where the V-cycle needs them)
 * ------------------------------------------------------------------------------------- */

/* r = f + sum(off-diagonal couplings * q) - c q over the interior.
 * Replaces BMG2_SymStd_residual (include/cedar/2d/residual.h:9-10, src/2d/ftn/BMG2_SymStd_residual.f90:1-119). */
int cdr_hip_residual2d_f64(const double *so, int nstncl, const double *qf, const double *q,
                           double *res, int ii, int jj, void *stream);
int cdr_hip_residual2d_f32(const float *so, int nstncl, const float *qf, const float *q,
                           float *res, int ii, int jj, void *stream);

/* Residual fused with the squared l2 norm of r over the interior: *sumsq_dev (device double)
 * receives sum(r^2), reduced in a fixed order (deterministic).  work needs
 * cdr_hip_norm_work2d(ii, jj) doubles of device memory.
 * Replaces residual + grid_func::lp_norm<2> (include/cedar/2d/grid_func.h:43-54). */
size_t cdr_hip_norm_work2d(int ii, int jj);
int cdr_hip_residual_norm2d_f64(const double *so, int nstncl, const double *qf, const double *q,
                                double *res, int ii, int jj, double *work, double *sumsq_dev,
                                void *stream);
/* cdr_hip_residual_norm2d_f64 without the work argument, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The per-block partial sums
 * live in a library-owned device array of CDR_NOALLOC_NORM_PARTIALS doubles (4 MiB), one per GPU,
 * held until the process ends; every partial a call reads was written by that call, so the
 * array needs no clearing.  Extents with cdr_hip_norm_work2d(ii, jj) > CDR_NOALLOC_NORM_PARTIALS
 * (above about 8190 x 8190 interior points) return CDR_ERR_TOO_LARGE.  Calls that overlap on one
 * GPU (other streams or host threads) share the array and race: serialise them.  Results are
 * bitwise those of cdr_hip_residual_norm2d_f64. */
#define CDR_NOALLOC_NORM_PARTIALS 524288
int cdr_hip_residual_norm2d_f64_noalloc(const double *so, int nstncl, const double *qf,
                                        const double *q, double *res, int ii, int jj,
                                        double *sumsq_dev, void *stream);
/* sum of squares of the interior of x into *sumsq_dev (device). */
int cdr_hip_sumsq2d_f64(const double *x, int ii, int jj, double *work, double *sumsq_dev,
                        void *stream);

/* SOR(:,:,2) = 1/c.  Replaces BMG2_SymStd_SETUP_recip (src/2d/relax.cc:8, src/2d/ftn/BMG2_SymStd_SETUP_recip.f90). */
int cdr_hip_setup_recip2d_f64(const double *so, double *sor, int ii, int jj, int nstncl,
                              void *stream);

/* Red-black (5-point) or four-colour (9-point) point Gauss-Seidel, colours forward on DOWN and
 * reversed on UP.  Replaces BMG2_SymStd_relax_GS (src/2d/relax.cc:10-11). */
int cdr_hip_relax_gs2d_f64(const double *so, int nstncl, const double *qf, double *q,
                           const double *sor, int ii, int jj, int updown, void *stream);
int cdr_hip_relax_gs2d_f32(const float *so, int nstncl, const float *qf, float *q,
                           const float *sor, int ii, int jj, int updown, void *stream);

/* LDL^T factors of every x-line (y-line) of the operator.  Synchronises the stream to report a
 * non-SPD line as CDR_ERR_NOT_SPD.  info_dev: one device int of scratch.
 * Replace BMG2_SymStd_SETUP_lines_x / _y (src/2d/relax.cc:8-9). */
int cdr_hip_setup_lines_x2d_f64(const double *so, double *sor, int ii, int jj, int nstncl,
                                int *info_dev, void *stream);
int cdr_hip_setup_lines_y2d_f64(const double *so, double *sor, int ii, int jj, int nstncl,
                                int *info_dev, void *stream);

/* Zebra line relaxation with the stored factors; one wavefront per line.  The reference's B
 * scratch argument (level.res) is not needed and may be NULL.
 * Replace BMG2_SymStd_relax_lines_x / _y (src/2d/relax.cc:12-17). */
int cdr_hip_relax_lines_x2d_f64(const double *so, int nstncl, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int updown,
                                void *stream);
int cdr_hip_relax_lines_y2d_f64(const double *so, int nstncl, const double *qf, double *q,
                                const double *sor, double *b, int ii, int jj, int updown,
                                void *stream);
int cdr_hip_relax_lines_x2d_f32(const float *so, int nstncl, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int updown,
                                void *stream);
int cdr_hip_relax_lines_y2d_f32(const float *so, int nstncl, const float *qf, float *q,
                                const float *sor, float *b, int ii, int jj, int updown,
                                void *stream);
/* Batched forms of the two line sweeps above (no counterpart in the reference, which relaxes one
 * grid per call): nb problems of the same extents stored back to back -- so as nb blocks of
 * nstncl planes, qf and q as nb grids, sor as nb blocks of 2 planes (each problem's own line
 * factors) -- relaxed by one launch per colour.  dir 0: x-lines, 1: y-lines.  The plane
 * relaxation of the 3D solver and the batched 2D solver run this form; each problem gets
 * bitwise the result of its own single-problem call. */
int cdr_hip_relax_lines2d_batched_f64(int dir, const double *so, int nstncl, const double *qf,
                                      double *q, const double *sor, int ii, int jj, int nb,
                                      int updown, void *stream);
int cdr_hip_relax_lines2d_batched_f32(int dir, const float *so, int nstncl, const float *qf,
                                      float *q, const float *sor, int ii, int jj, int nb,
                                      int updown, void *stream);

/* One line-xy smoothing step: x-lines then y-lines (DOWN) or y then x (UP), the composition
 * Cedar's smoother makes of the two line kernels (include/cedar/multilevel.h:175-178, 204-207),
 * bit for bit cdr_hip_relax_lines_x2d followed by cdr_hip_relax_lines_y2d (or y then x).
 * `work` is not used and may be NULL. */
int cdr_hip_relax_lines_xy2d_f64(const double *so, int nstncl, const double *qf, double *q,
                                 const double *sorx, const double *sory, double *work, int ii,
                                 int jj, int updown, void *stream);
int cdr_hip_relax_lines_xy2d_f32(const float *so, int nstncl, const float *qf, float *q,
                                 const float *sorx, const float *sory, float *work, int ii, int jj,
                                 int updown, void *stream);
/* qc = P^T q.  Replaces BMG2_SymStd_restrict (src/2d/restrict.cc:6-7). */
int cdr_hip_restrict2d_f64(const double *ci, const double *q, double *qc, int iif, int jjf,
                           int iic, int jjc, void *stream);
int cdr_hip_restrict2d_f32(const float *ci, const float *q, float *qc, int iif, int jjf,
                           int iic, int jjc, void *stream);

/* q += P qc, plus res/c at the fine-only points.  Replaces BMG2_SymStd_interp_add
 * (src/2d/interp.cc:8-9).  Difference: the reference divides res by c in place; here res is
 * only read (Cedar's V-cycle never reads res after this call, include/cedar/cycle/vcycle.h:103-105). */
int cdr_hip_interp_add2d_f64(double *q, const double *qc, const double *res, const double *so,
                             int nstncl, const double *ci, int iic, int jjc, int iif, int jjf,
                             void *stream);
int cdr_hip_interp_add2d_f32(float *q, const float *qc, const float *res, const float *so,
                             int nstncl, const float *ci, int iic, int jjc, int iif, int jjf,
                             void *stream);

/* Operator-induced interpolation weights CI from the fine operator.
 * Replaces BMG2_SymStd_SETUP_interp_OI (src/2d/interp.cc:10-12). */
int cdr_hip_setup_interp2d_f64(const double *so, int nstncl, double *ci, int iif, int jjf,
                               int iic, int jjc, void *stream);

/* Galerkin coarse operator soc = P^T A P (9-point, 5 planes).
 * Replaces BMG2_SymStd_SETUP_ITLI_ex (include/cedar/2d/coarsen.h:10-12). */
int cdr_hip_setup_galerkin2d_f64(const double *so, int nstncl, const double *ci, double *soc,
                                 int iif, int jjf, int iic, int jjc, void *stream);

/* Coarsest grid: banded Cholesky of the operator into abd (ii x nx*ny doubles, kd = ii-1).
 * Synchronises the stream; CDR_ERR_NOT_SPD if a pivot is not positive.
 * Replaces BMG2_SymStd_SETUP_cg_LU (include/cedar/2d/solve_cg.h:10). */
int cdr_hip_setup_cg2d_f64(const double *so, int nstncl, double *abd, int ii, int jj,
                           int *info_dev, void *stream);
/* q = A^{-1} qf on the coarsest grid with the factors (FP64 arithmetic for both variants);
 * bbd: nx*ny doubles of device scratch.  Replaces BMG2_SymStd_SOLVE_cg (src/2d/solve_cg.cc:6). */
int cdr_hip_solve_cg2d_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                           int jj, void *stream);
int cdr_hip_solve_cg2d_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                           int jj, void *stream);

/* ---------------------------------------------------------------------------------------
 * Periodic boundaries (grid.periodic, include/cedar/kernel_params.h:35-42).  jpn is the BoxMG
 * boundary id BMG_get_bc returns for the periodic mask (src/2d/ftn/BMG_get_bc.f90:1-24): 0
 * Dirichlet (the kernels above), 2 periodic in x, 1 in y, 3 in both.  The periodic images live
 * in the ghost layer: a periodic direction of n points has period n, and the ghosts of the
 * operator, of q and of the right-hand side hold the values of the opposite interior points.
 * Each entry point is the kernel above with the reference's JPN != 0 branch:
 * ------------------------------------------------------------------------------------- */
/* src/2d/ftn/BMG2_SymStd_relax_GS.f90:139-226: images refreshed after each colour (x) and after
 * the sweep (y) */
int cdr_hip_relax_gs2d_per_f64(const double *so, int nstncl, const double *qf, double *q,
                               const double *sor, int ii, int jj, int updown, int jpn,
                               void *stream);
int cdr_hip_relax_gs2d_per_f32(const float *so, int nstncl, const float *qf, float *q,
                               const float *sor, int ii, int jj, int updown, int jpn,
                               void *stream);
/* Lines (src/2d/ftn/BMG2_SymStd_SETUP_lines_{x,y}.f90 and BMG2_SymStd_relax_lines_{x,y}.f90,
 * JPN != 0): a periodic line direction gets its wrap couplings on the end diagonals of the
 * factors and a Sherman-Morrison line solve (BETA/ALPHA correction); lines across a periodic
 * direction run the Dirichlet solve and refresh that direction's images after the sweep.  One
 * lane per line in the reference's sequential DPTTRS order.  work: device scratch of ii*jj. */
int cdr_hip_setup_lines_x2d_per_f64(const double *so, double *sor, int ii, int jj, int nstncl,
                                    int jpn, int *info_dev, void *stream);
int cdr_hip_setup_lines_y2d_per_f64(const double *so, double *sor, int ii, int jj, int nstncl,
                                    int jpn, int *info_dev, void *stream);
int cdr_hip_relax_lines_x2d_per_f64(const double *so, int nstncl, const double *qf, double *q,
                                    const double *sor, double *work, int ii, int jj, int updown,
                                    int jpn, void *stream);
int cdr_hip_relax_lines_y2d_per_f64(const double *so, int nstncl, const double *qf, double *q,
                                    const double *sor, double *work, int ii, int jj, int updown,
                                    int jpn, void *stream);
int cdr_hip_relax_lines_x2d_per_f32(const float *so, int nstncl, const float *qf, float *q,
                                    const float *sor, float *work, int ii, int jj, int updown,
                                    int jpn, void *stream);
int cdr_hip_relax_lines_y2d_per_f32(const float *so, int nstncl, const float *qf, float *q,
                                    const float *sor, float *work, int ii, int jj, int updown,
                                    int jpn, void *stream);
/* src/2d/ftn/BMG2_SymStd_restrict.f90:94-126: when the fine extent is even, the images of q are
 * refreshed in place before restricting (q is written) */
int cdr_hip_restrict2d_per_f64(const double *ci, double *q, double *qc, int iif, int jjf, int iic,
                               int jjc, int jpn, void *stream);
int cdr_hip_restrict2d_per_f32(const float *ci, float *q, float *qc, int iif, int jjf, int iic,
                               int jjc, int jpn, void *stream);
/* src/2d/ftn/BMG2_SymStd_interp_add.f90:139-157: images of q refreshed afterwards */
int cdr_hip_interp_add2d_per_f64(double *q, const double *qc, const double *res, const double *so,
                                 int nstncl, const double *ci, int iic, int jjc, int iif, int jjf,
                                 int jpn, void *stream);
int cdr_hip_interp_add2d_per_f32(float *q, const float *qc, const float *res, const float *so,
                                 int nstncl, const float *ci, int iic, int jjc, int iif, int jjf,
                                 int jpn, void *stream);
/* src/2d/ftn/BMG2_SymStd_SETUP_interp_OI.f90:259-616 (weights also in the ghost rows/columns of
 * CI that the periodic interpolation reads; ci must be zero on entry) */
int cdr_hip_setup_interp2d_per_f64(const double *so, int nstncl, double *ci, int iif, int jjf,
                                   int iic, int jjc, int jpn, void *stream);
/* src/2d/ftn/BMG2_SymStd_SETUP_ITLI_ex.f90:94-363: the hand-expanded products, then the images
 * of the coarse operator */
int cdr_hip_setup_galerkin2d_per_f64(const double *so, int nstncl, const double *ci, double *soc,
                                     int iif, int jjf, int iic, int jjc, int jpn, void *stream);
/* src/2d/ftn/BMG2_SymStd_SETUP_cg_LU.f90:119-227,251-335 and BMG2_SymStd_SOLVE_cg.f90:93-163:
 * the dense upper triangle abd(n, n), n = nx*ny (2 <= nx, ny; n <= 2048), factored by an upper
 * Cholesky (DPOTF2 order; CDR_ERR_NOT_SPD for a non-positive pivot); the solve removes the mean
 * of the coarse solution and refreshes its images.  bbd holds n doubles.  Synchronises (setup). */
int cdr_hip_setup_cg2d_per_f64(const double *so, int nstncl, double *abd, int ii, int jj, int jpn,
                               int *info_dev, void *stream);
int cdr_hip_solve_cg2d_per_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                               int jj, int jpn, void *stream);
int cdr_hip_solve_cg2d_per_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                               int jj, int jpn, void *stream);

/* ---------------------------------------------------------------------------------------
 * 2D solver (cdr2::solver<five_pt|nine_pt>: setup in create, solve(b, x))
 * ------------------------------------------------------------------------------------- */
enum {
    CDR_RELAX_POINT = 0,
    CDR_RELAX_LINE_X = 1,
    CDR_RELAX_LINE_Y = 2,
    CDR_RELAX_LINE_XY = 3,
    CDR_RELAX_PLANE_XY = 4,   /* 3D only */
    CDR_RELAX_PLANE_XZ = 5,
    CDR_RELAX_PLANE_YZ = 6,
    CDR_RELAX_PLANE_XYZ = 7
};
enum {
    CDR_CYCLE_V = 0,     /* solver.cycle.type "v" (include/cedar/cycle/vcycle.h) */
    CDR_CYCLE_F = 1      /* solver.cycle.type "f" (include/cedar/cycle/fcycle.h) */
};
enum {
    CDR_PREC_F64 = 0,    /* plain FP64 V-cycles (Cedar's solve) */
    CDR_PREC_MIXED = 1   /* FP32 V-cycle preconditioning FP64 iterative refinement */
};

typedef struct {
    int relaxation;   /* solver.relaxation: CDR_RELAX_* (default point) */
    int nrelax_pre;   /* solver.cycle.nrelax-pre (default 2) */
    int nrelax_post;  /* solver.cycle.nrelax-post (default 1) */
    int num_levels;   /* solver.num-levels (default -1: computed) */
    int maxiter;      /* solver.max-iter (default 10) */
    double tol;       /* solver.tol (default 1e-8) */
    int min_coarse;   /* solver.min_coarse (default 3) */
    int precision;    /* CDR_PREC_* (default CDR_PREC_F64) */
    int graph;        /* 1: each V-cycle is replayed from a captured hipGraph (default), 0: eager */
    int cycle;        /* solver.cycle.type: CDR_CYCLE_V (default) or CDR_CYCLE_F.  Like the
                         reference, an F-cycle recomputes x from b alone (include/cedar/cycle/
                         fcycle.h:75-81), so successive cycles of a solve repeat */
    int fused_coarse; /* 1 (default): in V-cycles, the levels of at most 32 points per side (below
                         the finest) run as one fused launch of one workgroup per problem; the
                         result is bitwise that of one launch per step.  0: one launch per step */
    int periodic;     /* grid.periodic as Cedar's per_mask (include/cedar/kernel_params.h:35-42):
                         bit 0 periodic in x, bit 1 in y; default 0 (Dirichlet).  The operator, b
                         and x carry the periodic images in their ghosts (the periodic kernels
                         above); line relaxation then runs one lane per line, and fused_coarse
                         and line-shape tuning do not apply */
    int stall_cycles; /* 0 (default, the reference's loop, include/cedar/multilevel.h:283-297):
                         cycle until rho < tol or max-iter.  k > 0: a problem whose rho has not
                         gone below its smallest earlier value (rho_0 = 1) for k consecutive
                         cycles stops there, and the solve returns CDR_ERR_STALLED (see
                         cdr_hip_stop_rule) */
} cdr_solver_settings;

typedef struct cdr_hip_solver2d cdr_hip_solver2d;

/* Cedar's defaults (src/multilevel_settings.cc:15-61). */
void cdr_hip_default_settings(cdr_solver_settings *s);

/* The stopping rule of every solve loop (2D, 3D, batched), on host data: feeds rho[0..n-1]
 * (the relative residuals after cycles 1..n) through the rule in order and stores in *ncycles
 * the number of cycles the loop runs (n when none stops it, as when max-iter ends the loop).
 * Returns CDR_OK (converged, or no stop within n), CDR_ERR_NONFINITE (a NaN/Inf rho) or
 * CDR_ERR_STALLED (stall_cycles > 0 cycles without a new smallest rho).  tol < 0 or
 * stall_cycles < 0 is CDR_ERR_ARG.  No GPU needed.  Host-side counterpart of the loop test in
 * include/cedar/multilevel.h:283-297 plus the two guards. */
int cdr_hip_stop_rule(const double *rho, int n, double tol, int stall_cycles, int *ncycles);

/* Builds the hierarchy on the GPU from the fine operator so_dev (device, held by reference as
 * Cedar holds the fine operator, include/cedar/level.h:25,30) of nx x ny interior points.
 * Synchronises.  Replaces cdr2::solver<fsten>(so, conf) (include/cedar/2d/solver.h:55-61). */
int cdr_hip_solver2d_create(const double *so_dev, int nstncl, int nx, int ny,
                            const cdr_solver_settings *s, void *stream, cdr_hip_solver2d **out);
/* Batched form (no counterpart in the reference, which solves one problem per solver): nb
 * independent problems of the same extents, so_dev holding their nb operators back to back
 * (problem b at so_dev + b * nstncl * (nx+2) * (ny+2)); every kernel launch covers the batch.
 * b_dev / x_dev of solve and vcycle then hold nb grids back to back. */
int cdr_hip_solver2d_create_batched(const double *so_dev, int nstncl, int nx, int ny, int nb,
                                    const cdr_solver_settings *s, void *stream,
                                    cdr_hip_solver2d **out);
/* Several right-hand sides per operator (no counterpart in the reference): nops operators of
 * the same extents (so_dev as for the batched form), each solved for nrhs right-hand sides, so
 * nops * nrhs problems (at most 65535).  b_dev / x_dev of solve and vcycle hold the problems back
 * to back, operator-major: right-hand side r of operator k is problem k * nrhs + r.  The
 * hierarchy is built once per operator and stores each operator's stencil, interpolation and
 * line factors once for its nrhs problems; every cycle kernel covers all problems in one launch
 * (the mixed outer step reads an operator point once for all its right-hand sides).  Each
 * problem gets bitwise the iterate and history of its own single solve.  Not
 * available with periodic boundaries (CDR_ERR_ARG). */
int cdr_hip_solver2d_create_shared(const double *so_dev, int nstncl, int nx, int ny, int nops,
                                   int nrhs, const cdr_solver_settings *s, void *stream,
                                   cdr_hip_solver2d **out);
/* solve(b, x) from the x given (x0 = 0 gives Cedar's solve(b)); rho_i = ||r_i|| / ||r_0|| go to
 * hist (host, up to maxhist per problem, problem b at hist[b * maxhist + i]), niter[b] gets the
 * cycles run for problem b.  A problem stops, keeping its iterate, once its rho < tol, so each
 * problem of a batch gets the result of solving it alone.  Synchronises once per cycle to read
 * the norms.  Replaces multilevel::solve (include/cedar/multilevel.h:278-298). */
int cdr_hip_solver2d_solve(cdr_hip_solver2d *slv, const double *b_dev, double *x_dev,
                           double *hist, int maxhist, int *niter);
/* One FP64 cycle of the configured type on (x, b) (include/cedar/cycle/vcycle.h:35-54,
 * include/cedar/cycle/fcycle.h:38-47). */
int cdr_hip_solver2d_vcycle(cdr_hip_solver2d *slv, const double *b_dev, double *x_dev);
int cdr_hip_solver2d_nlevels(const cdr_hip_solver2d *slv);
int cdr_hip_solver2d_nbatch(const cdr_hip_solver2d *slv);
/* Extents of level lvl (0 = finest) and device pointers to its FP64 data (NULL where absent);
 * for a batch, the data of problem 0 followed by the other problems'. */
int cdr_hip_solver2d_level(const cdr_hip_solver2d *slv, int lvl, int *nx, int *ny,
                           const double **so, const double **ci, const double **sor0,
                           const double **sor1);
/* Wall time (s, HIP events) of the last create (setup) and solve. */
int cdr_hip_solver2d_times(const cdr_hip_solver2d *slv, double *setup_s, double *solve_s);
/* Per-level timers (timer_begin / timer_end / timer_save of include/cedar/util/time_log.h:95-127):
 * with timers on, cycles run eagerly (no graph replay, no fused coarse levels; the same results)
 * and each phase of the V-cycle ("relaxation", "residual", "restrict", "coarse-solve",
 * "interp-add") is timed per level on the device; "setup" and "solve" are logged at level 0.
 * cdr_hip_solver2d_save_timers writes them to path in the reference's serial JSON layout
 * (src/util/time_log.cc:103-129: {"levels": [coarsest .. finest, label: {time, count}]}) and
 * resets them. */
int cdr_hip_solver2d_timers(cdr_hip_solver2d *slv, int on);
int cdr_hip_solver2d_save_timers(cdr_hip_solver2d *slv, const char *path);
/* Writes a device stencil (nstncl planes of ii x jj) as text in the layout of the reference's
 * operator<< (src/2d/stencil_op.cc:5-51). */
int cdr_hip_stencil2d_write(const double *so_dev, int nstncl, int ii, int jj, const char *path);
/* The solver's kernel registry (kernel_manager, include/cedar/kernel_manager.h:38-72):
 * cdr_hip_solver2d_kernels writes {"interface": {"impls": [...], "default": "..."}} as JSON into
 * buf (len bytes, NUL-terminated, truncated if short) and returns the full length;
 * cdr_hip_solver2d_set_kernel is kman->set<I>(impl) for the interface named iface (both
 * precisions), CDR_ERR_ARG for an unknown interface or implementation.  The interfaces are those
 * of ser_kernels (include/cedar/kernel_registry.h:23-35) that 2D solves use: "residual",
 * "point-relax", "line-relax-x", "line-relax-y", "setup-interp", "coarsen-op", "restriction",
 * "interp-add", "solve-cg"; each has the "hip" implementation. */
int cdr_hip_solver2d_kernels(cdr_hip_solver2d *slv, char *buf, int len);
int cdr_hip_solver2d_set_kernel(cdr_hip_solver2d *slv, const char *iface, const char *impl);
void cdr_hip_solver2d_destroy(cdr_hip_solver2d *slv);

/* ---------------------------------------------------------------------------------------
 * 3D kernels (FP64 unless noted).  Arrays are Fortran order with one ghost layer, extents
 * ii = nx+2, jj, kk; stencils are 4 planes {p, pw, ps, b} (7-point) or 14 planes {p, pw, ps, b,
 * psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw} (27-point) (include/cedar/3d/base_types.h:5-20),
 * the prolongation 26 planes on the coarse grid (src/3d/ftn/BMG_stencils_f90.h:59-65).
 * ------------------------------------------------------------------------------------- */
/* BMG3_SymStd_residual (include/cedar/3d/residual.h:9-13) */
int cdr_hip_residual3d_f64(const double *so, int nstncl, const double *qf, const double *q,
                           double *res, int ii, int jj, int kk, void *stream);
int cdr_hip_residual3d_f32(const float *so, int nstncl, const float *qf, const float *q,
                           float *res, int ii, int jj, int kk, void *stream);
/* residual plus the deterministic sum of r^2 over the interior into *sumsq_dev (device);
 * work holds cdr_hip_norm_work3d(ii, jj, kk) doubles (include/cedar/3d/grid_func.h lp_norm) */
size_t cdr_hip_norm_work3d(int ii, int jj, int kk);
int cdr_hip_residual_norm3d_f64(const double *so, int nstncl, const double *qf, const double *q,
                                double *res, int ii, int jj, int kk, double *work,
                                double *sumsq_dev, void *stream);
/* BMG3_SymStd_SETUP_recip / BMG3_SymStd_relax_GS (include/cedar/3d/relax.h:9-16); sor has 2
 * planes, the reciprocal in the second */
int cdr_hip_setup_recip3d_f64(const double *so, double *sor, int ii, int jj, int kk,
                              int nstncl, void *stream);
int cdr_hip_relax_gs3d_f64(const double *so, int nstncl, const double *qf, double *q,
                           const double *sor, int ii, int jj, int kk, int updown, void *stream);
int cdr_hip_relax_gs3d_f32(const float *so, int nstncl, const float *qf, float *q,
                           const float *sor, int ii, int jj, int kk, int updown, void *stream);
/* BMG3_SymStd_restrict (src/3d/restrict.cc:7-10) */
int cdr_hip_restrict3d_f64(const double *ci, const double *q, double *qc, int iif, int jjf,
                           int kkf, int iic, int jjc, int kkc, void *stream);
/* FP32 form (the mixed-precision 3D cycle's restriction; same operation order in float) */
int cdr_hip_restrict3d_f32(const float *ci, const float *q, float *qc, int iif, int jjf, int kkf,
                           int iic, int jjc, int kkc, void *stream);
/* BMG3_SymStd_interp_add (src/3d/interp.cc:7-11); res is not modified (the Fortran divides it
 * by SO_p in place) */
int cdr_hip_interp_add3d_f64(double *q, const double *qc, const double *res, const double *so,
                             int nstncl, const double *ci, int iic, int jjc, int kkc, int iif,
                             int jjf, int kkf, void *stream);
/* FP32 form (the mixed-precision 3D cycle's interpolation) */
int cdr_hip_interp_add3d_f32(float *q, const float *qc, const float *res, const float *so,
                             int nstncl, const float *ci, int iic, int jjc, int kkc, int iif,
                             int jjf, int kkf, void *stream);
/* BMG3_SymStd_SETUP_interp_OI (include/cedar/3d/interp.h:11-15); no scratch argument */
int cdr_hip_setup_interp3d_f64(const double *so, int nstncl, double *ci, int iif, int jjf,
                               int kkf, int iic, int jjc, int kkc, void *stream);
/* BMG3_SymStd_SETUP_ITLI07_ex / _ITLI27_ex (include/cedar/3d/coarsen.h:12-16): coarse 14-plane
 * operator */
int cdr_hip_setup_galerkin3d_f64(const double *so, int nstncl, const double *ci, double *soc,
                                 int iif, int jjf, int kkf, int iic, int jjc, int kkc,
                                 void *stream);
/* BMG3_SymStd_SETUP_cg_LU / BMG3_SymStd_SOLVE_cg (include/cedar/3d/solve_cg.h:10-11,
 * src/3d/solve_cg.cc:6-9): band ABD(ldab, nx*ny*nz), kd = cdr_hip_cg_bandwidth3d(ii, jj) =
 * nx*(ny+1)+1, ldab >= kd+1; bbd holds nx*ny*nz doubles (device) */
int cdr_hip_cg_bandwidth3d(int ii, int jj);
int cdr_hip_setup_cg3d_f64(const double *so, int nstncl, double *abd, int ldab, int ii, int jj,
                           int kk, int *info_dev, void *stream);
int cdr_hip_solve_cg3d_f64(double *q, const double *qf, const double *abd, int ldab, double *bbd,
                           int ii, int jj, int kk, void *stream);
/* FP32 right-hand side and solution with the FP64 band (the coarsest solve stays FP64: qf is
 * widened, the solve runs in FP64, q is rounded) */
int cdr_hip_solve_cg3d_f32(float *q, const float *qf, const double *abd, int ldab, double *bbd,
                           int ii, int jj, int kk, void *stream);

/* ---------------------------------------------------------------------------------------
 * 3D periodic boundaries (grid.periodic, include/cedar/kernel_params.h:35-42).  jpn is the
 * BoxMG boundary id BMG_get_bc returns for the periodic mask (bits x = 1, y = 2, z = 4;
 * src/2d/ftn/BMG_get_bc.f90:1-24): 2 x, 1 y, 3 xy, 5 z, 6 xz, 7 yz, 8 xyz
 * (include/cedar/2d/ftn/BMG_parameters_c.h:193-204); the Dirichlet id 0 is rejected here (use
 * the kernels above).  The images of a periodic direction live in its ghost layer: q(1) =
 * q(n-1), q(n) = q(2) for the operator, q and the right-hand side.  Each entry point is the
 * reference's JPN != 0 branch:
 * ------------------------------------------------------------------------------------- */
/* src/3d/ftn/BMG3_SymStd_relax_GS.f90:189-352 (include/cedar/3d/relax.h:94-98): x images of
 * each row and y images of each plane after its colour pass, z images after the sweep */
int cdr_hip_relax_gs3d_per_f64(const double *so, int nstncl, const double *qf, double *q,
                               const double *sor, int ii, int jj, int kk, int updown, int jpn,
                               void *stream);
int cdr_hip_relax_gs3d_per_f32(const float *so, int nstncl, const float *qf, float *q,
                               const float *sor, int ii, int jj, int kk, int updown, int jpn,
                               void *stream);
/* src/3d/ftn/BMG3_SymStd_restrict.f90:79-141 (src/3d/restrict.cc:27): the images of q are
 * refreshed in place (y, x, z) before restricting (q is written) */
int cdr_hip_restrict3d_per_f64(const double *ci, double *q, double *qc, int iif, int jjf, int kkf,
                               int iic, int jjc, int kkc, int jpn, void *stream);
int cdr_hip_restrict3d_per_f32(const float *ci, float *q, float *qc, int iif, int jjf, int kkf,
                               int iic, int jjc, int kkc, int jpn, void *stream);
/* src/3d/ftn/BMG3_SymStd_interp_add.f90:72-212 (src/3d/interp.cc:39), then every image of q
 * (y, x, z).  The reference's own image loops (:237-280) test IPN before assigning it and
 * index with left-over loop variables; this refresh is the well-defined one restriction uses. */
int cdr_hip_interp_add3d_per_f64(double *q, const double *qc, const double *res, const double *so,
                                 int nstncl, const double *ci, int iic, int jjc, int kkc, int iif,
                                 int jjf, int kkf, int jpn, void *stream);
int cdr_hip_interp_add3d_per_f32(float *q, const float *qc, const float *res, const float *so,
                                 int nstncl, const float *ci, int iic, int jjc, int kkc, int iif,
                                 int jjf, int kkf, int jpn, void *stream);
/* src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:809-2784 (include/cedar/3d/interp.h:65-73),
 * including the image weights it writes; ci must be zero on entry */
int cdr_hip_setup_interp3d_per_f64(const double *so, int nstncl, double *ci, int iif, int jjf,
                                   int kkf, int iic, int jjc, int kkc, int jpn, void *stream);
/* src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90, _ITLI27_ex.f90 (include/cedar/3d/coarsen.h:
 * 41-61) with IPN: the product P^T A P over the periodic geometry (equal to the reference's
 * hand-expanded products to rounding), then the images of the coarse operator (y, x, z).
 * Periodic extents must be even (n fine points coarsen to n/2). */
int cdr_hip_setup_galerkin3d_per_f64(const double *so, int nstncl, const double *ci, double *soc,
                                     int iif, int jjf, int kkf, int iic, int jjc, int kkc, int jpn,
                                     void *stream);
/* src/3d/ftn/BMG3_SymStd_SETUP_cg_LU.f90:190-612 and BMG3_SymStd_SOLVE_cg.f90:103-196
 * (include/cedar/3d/solver.h:123-126): the dense matrix abd(n, n), n = nx*ny*nz <= 2048, of the
 * operator with its periodic couplings, factored by an upper Cholesky (DPOTF2 order;
 * CDR_ERR_NOT_SPD for a non-positive pivot); the solve removes the mean of the coarse solution
 * and refreshes its images (x, y, z).  bbd holds n doubles.  Synchronises (setup). */
int cdr_hip_setup_cg3d_per_f64(const double *so, int nstncl, double *abd, int ii, int jj, int kk,
                               int jpn, int *info_dev, void *stream);
int cdr_hip_solve_cg3d_per_f64(double *q, const double *qf, const double *abd, double *bbd, int ii,
                               int jj, int kk, int jpn, void *stream);
int cdr_hip_solve_cg3d_per_f32(float *q, const float *qf, const double *abd, double *bbd, int ii,
                               int jj, int kk, int jpn, void *stream);

/* ---------------------------------------------------------------------------------------
 * 3D solver (cdr3::solver<seven_pt|xxvii_pt>), FP64, point or plane relaxation
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int relaxation;   /* CDR_RELAX_POINT or CDR_RELAX_PLANE_* (default point) */
    int nrelax_pre;   /* default 2 */
    int nrelax_post;  /* default 1 */
    int num_levels;   /* default -1: computed (include/cedar/3d/solver.h:59-77) */
    int maxiter;      /* default 10 */
    double tol;       /* default 1e-8 */
    int min_coarse;   /* default 3 */
    cdr_solver_settings plane; /* plane-config of the nested 2D plane solves: default line-xy,
                                  max-iter 1 (src/kernel_params.cc:83-89); its precision and
                                  graph fields are ignored, and its max-iter cycles always run
                                  (no per-plane stop test) */
    int cycle;        /* CDR_CYCLE_V (default) or CDR_CYCLE_F */
    int precision;    /* CDR_PREC_F64 (default) or CDR_PREC_MIXED: FP32 3D cycle and FP32 plane
                         solves inside FP64 iterative refinement */
    int stall_cycles; /* as in cdr_solver_settings (default 0: no stall test) */
    int periodic;     /* grid.periodic as Cedar's per_mask (include/cedar/kernel_params.h:35-42):
                         bit 0 x, bit 1 y, bit 2 z; default 0 (Dirichlet).  Point relaxation
                         only; the operator, b and x carry the periodic images in their ghosts,
                         and the periodic extents must be even on every coarsened level. */
} cdr_solver3_settings;

typedef struct cdr_hip_solver3d cdr_hip_solver3d;

void cdr_hip_default_settings3(cdr_solver3_settings *s);
/* Replaces cdr3::solver<fsten>(so, conf) (include/cedar/3d/solver.h:46-57); the fine operator
 * is held by reference.  Plane relaxation builds, per level and plane family, one batched 2D
 * hierarchy per colour (the reference's per-plane cdr2 solvers, include/cedar/3d/relax_planes.h
 * :166-213). */
int cdr_hip_solver3d_create(const double *so_dev, int nstncl, int nx, int ny, int nz,
                            const cdr_solver3_settings *s, void *stream, cdr_hip_solver3d **out);
/* multilevel::solve (include/cedar/multilevel.h:278-298) from the x given */
int cdr_hip_solver3d_solve(cdr_hip_solver3d *slv, const double *b_dev, double *x_dev,
                           double *hist, int maxhist, int *niter);
int cdr_hip_solver3d_vcycle(cdr_hip_solver3d *slv, const double *b_dev, double *x_dev);
/* One plane-relaxation sweep (both colours) of family 0 = xy, 1 = xz, 2 = yz at level lvl:
 * planes<rdir>::run (include/cedar/3d/relax_planes.h:217-228) */
int cdr_hip_solver3d_relax_planes(cdr_hip_solver3d *slv, int lvl, int family, double *x_dev,
                                  const double *b_dev, int updown);
int cdr_hip_solver3d_nlevels(const cdr_hip_solver3d *slv);
int cdr_hip_solver3d_level(const cdr_hip_solver3d *slv, int lvl, int *nx, int *ny, int *nz,
                           const double **so, const double **ci);
int cdr_hip_solver3d_times(const cdr_hip_solver3d *slv, double *setup_s, double *solve_s);
/* The 3D solver's kernel registry, as for 2D: interfaces "residual (3D)", "point-relax (3D)",
 * "plane-relax-xy", "plane-relax-xz", "plane-relax-yz" (include/cedar/kernels/plane_relax.h
 * :11-31), "setup-interp (3D)", "coarsen-op (3D)", "restriction (3D)", "interp-add (3D)",
 * "solve-cg (3D)"; the solver runs every kernel of its setup and cycle through it. */
int cdr_hip_solver3d_kernels(cdr_hip_solver3d *slv, char *buf, int len);
int cdr_hip_solver3d_set_kernel(cdr_hip_solver3d *slv, const char *iface, const char *impl);
void cdr_hip_solver3d_destroy(cdr_hip_solver3d *slv);

/* ---------------------------------------------------------------------------------------
 * y = A q over the interior (ghosts of y untouched), terms in the order of
 * BMG2_SymStd_UTILS_matvec (src/2d/ftn/mpi/BMG2_SymStd_UTILS_matvec.f90: 9-point branch for
 * nstncl 5, 5-point for 3) and BMG3_SymStd_UTILS_matvec (src/3d/ftn/mpi/
 * BMG3_SymStd_UTILS_matvec.f90:79-124, nstncl 14 or 4); the kernels behind cedar_matvec
 * (include/cedar/kernels/matvec.h).  kk - 2 <= 65535.
 * ------------------------------------------------------------------------------------- */
int cdr_hip_matvec2d_f64(const double *so, int nstncl, const double *q, double *y, int ii, int jj,
                         void *stream);
int cdr_hip_matvec3d_f64(const double *so, int nstncl, const double *q, double *y, int ii, int jj,
                         int kk, void *stream);

/* ---------------------------------------------------------------------------------------
 * Differentiated 2D diffusion solves (no counterpart in the reference).  kcell holds nb blocks
 * of (ny+1) x (nx+1) per-cell coefficients (cell (cj, ci) at cj*(nx+1) + ci, its corners the
 * nodes (ci, cj), (ci+1, cj), (ci+1, cj+1), (ci, cj+1) of the ghost-inclusive node grid).
 * ------------------------------------------------------------------------------------- */
/* 9-point bilinear finite-element operator of -div(k grad u), Dirichlet couplings dropped,
 * written as nb stencils {c, w, s, sw, nw} of (nx+2) x (ny+2) */
int cdr_hip_q1_assemble2d_f64(const double *kcell, double *so, int nx, int ny, int nb,
                              void *stream);
/* dL/dk per cell from the state u and the adjoint lambda = A^{-1} dL/du:
 * grad_k[e] = -lambda_e^T K_e u_e / k_e over the cell's corners (boundary corners count 0) */
int cdr_hip_q1_gradient2d_f64(const double *u, const double *lam, double *grad_k, int nx, int ny,
                              int nb, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CEDAR_HIP_H */

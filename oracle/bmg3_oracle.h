/*
 * FP64 CPU oracle for the 3D BoxMG solver path (point and plane relaxation).
 *
 * TEST INFRASTRUCTURE ONLY.  A plain-C restatement of Cedar's serial 3D kernels, of the plane
 * relaxation that nests 2D solves (built on the 2D oracle in bmg2_oracle.c) and of the 3D
 * multilevel driver, following the reference's floating-point operation order.  Used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 *
 * Conventions (same as the reference):
 *   - arrays are Fortran order, first index fastest, one ghost layer: ii = nx+2, jj, kk;
 *   - a 7-point stencil is 4 planes {p, pw, ps, b}, a 27-point stencil 14 planes
 *     {p, pw, ps, b, psw, pnw, bw, bnw, bn, bne, be, bse, bs, bsw}
 *     (include/cedar/3d/base_types.h:5-20), off-diagonals stored negated;
 *   - the prolongation CI has 26 planes on the coarse grid in the order of
 *     src/3d/ftn/BMG_stencils_f90.h:54-65 (xyl, xyr, xya, xyb, xza, xzb, xyne, xyse, xysw, xynw,
 *     xzsw, xznw, xzne, xzse, yzsw, yznw, yzne, yzse, bsw, bnw, bne, bse, tsw, tnw, tne, tse).
 */
#ifndef CEDAR_ORACLE_BMG3_H
#define CEDAR_ORACLE_BMG3_H

#include <stddef.h>

#include "bmg2_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC3_RELAX_POINT = 0,
    ORC3_RELAX_PLANE_XY = 4,
    ORC3_RELAX_PLANE_XZ = 5,
    ORC3_RELAX_PLANE_YZ = 6,
    ORC3_RELAX_PLANE_XYZ = 7
};
enum { ORC3_XY = 0, ORC3_XZ = 1, ORC3_YZ = 2 };

/* ---- kernels (3D, FP64, Dirichlet) ---- */
void orc3_residual(const double *so, int nstncl, const double *qf, const double *q, double *res,
                   int ii, int jj, int kk);
void orc3_setup_recip(const double *so, double *sor, int ii, int jj, int kk);
void orc3_relax_gs(const double *so, int nstncl, const double *qf, double *q, const double *sor,
                   int ii, int jj, int kk, int updown);
void orc3_restrict(const double *ci, const double *q, double *qc, int iif, int jjf, int kkf,
                   int iic, int jjc, int kkc);
/* res is divided by SO_p in the interior, in place, as the reference does */
void orc3_interp_add(double *q, const double *qc, double *res, const double *so, int nstncl,
                     const double *ci, int iic, int jjc, int kkc, int iif, int jjf, int kkf);
void orc3_setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf,
                       int iic, int jjc, int kkc);
/* One loop body of the interpolation setup: the weights of coarse point (ic,jc,kc) tied to the
 * fine anchor (i,j,k) (Dirichlet: i = 2(ic-1), ...), for the x-, y-, z-edge, xy-, xz-, yz-face
 * and cell-centre fine points; nine = 27-point fine operator.  Shared by the periodic branch. */
#define ORC3_OI_PT_ARGS                                                                      \
    const double *so, int nine, double *ci, int ii, int jj, size_t P, int iic, int jjc,      \
        size_t Pc, int i, int j, int k, int ic, int jc, int kc
void orc3_oi_xl(ORC3_OI_PT_ARGS);
void orc3_oi_ya(ORC3_OI_PT_ARGS);
void orc3_oi_za(ORC3_OI_PT_ARGS);
void orc3_oi_xyf(ORC3_OI_PT_ARGS);
void orc3_oi_xzf(ORC3_OI_PT_ARGS);
void orc3_oi_yzf(ORC3_OI_PT_ARGS);
void orc3_oi_cell(ORC3_OI_PT_ARGS);
/* Galerkin coarse operator P^T A P (14 planes) from the fine operator and CI */
void orc3_setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif,
                         int jjf, int kkf, int iic, int jjc, int kkc);
/* coarsest grid: band ABD(ldab, nx*ny*nz), kd = nx*(ny+1)+1, ldab >= kd+1 */
int  orc3_cg_bandwidth(int ii, int jj);
int  orc3_setup_cg(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk);
void orc3_solve_cg(double *q, const double *qf, const double *abd, int ldab, double *bbd,
                   int ii, int jj, int kk);
double orc3_norm(const double *x, int ii, int jj, int kk);
/* y = A q over the interior (ghosts of y untouched) */
void orc3_matvec(const double *so, int nstncl, const double *q, double *y, int ii, int jj,
                 int kk);

/* ---- periodic branches (bmg3_periodic.c): jpn is the BoxMG boundary id of the periodic
 * mask (orc3_bc_from_mask); jpn == 0 gives the Dirichlet kernels above ---- */
int  orc3_bc_from_mask(int mask);
int  orc3_per_x(int jpn);
int  orc3_per_y(int jpn);
int  orc3_per_z(int jpn);
void orc3_images(double *q, int ii, int jj, int kk, int jpn);
void orc3_relax_gs_bc(const double *so, int nstncl, const double *qf, double *q, const double *sor,
                      int ii, int jj, int kk, int updown, int jpn);
/* refreshes the images of q (the fine residual) in place first */
void orc3_restrict_bc(const double *ci, double *q, double *qc, int iif, int jjf, int kkf, int iic,
                      int jjc, int kkc, int jpn);
void orc3_interp_add_bc(double *q, const double *qc, double *res, const double *so, int nstncl,
                        const double *ci, int iic, int jjc, int kkc, int iif, int jjf, int kkf,
                        int jpn);
void orc3_setup_interp_bc(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf,
                          int iic, int jjc, int kkc, int jpn);
void orc3_setup_galerkin_bc(const double *so, int nstncl, const double *ci, double *soc, int iif,
                            int jjf, int kkf, int iic, int jjc, int kkc, int jpn);
/* periodic coarsest grid: dense n x n (n = nx ny nz) upper Cholesky factor; ldab is ignored */
void orc3_assemble_cg_periodic(const double *so, int nstncl, double *abd, int ii, int jj, int kk,
                               int jpn);
int  orc3_setup_cg_bc(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk,
                      int jpn);
void orc3_solve_cg_bc(double *q, const double *qf, const double *abd, int ldab, double *bbd,
                      int ii, int jj, int kk, int jpn);
/* shared pieces of the kernels above */
double orc3_gs_point(const double *so, int nstncl, const double *qf, const double *q, int i, int j,
                     int k, int ii, int jj, int kk);
double orc3_stencil_term(const double *so, int nstncl, int ii, int jj, int kk, int i, int j, int k,
                         int t, int *oi, int *oj, int *ok);

/* ---- plane relaxation pieces (include/cedar/3d/relax_planes.h, src/3d/relax_planes.cc) ---- */
/* 2D operator of plane ipl (1..n of the direction normal to the plane) */
void orc3_copy_coeff(int dir, const double *so3, int nstncl3, double *so2, int ipl, int ii, int jj,
                     int kk);
void orc3_copy_rhs(int dir, const double *so3, int nstncl3, const double *x, const double *b,
                   double *b2, int ipl, int ii, int jj, int kk);
void orc3_copy32(int dir, const double *x, double *x2, int ipl, int ii, int jj, int kk);
void orc3_copy23(int dir, const double *x2, double *x, int ipl, int ii, int jj, int kk);

/* ---- multilevel driver ---- */
typedef struct {
    int relax;        /* ORC3_RELAX_* */
    int nrelax_pre;   /* default 2 */
    int nrelax_post;  /* default 1 */
    int num_levels;   /* -1 = computed */
    int maxiter;
    double tol;
    int min_coarse;   /* default 3 */
    orc2_settings plane; /* nested plane solves: line-xy, max-iter 1 (src/kernel_params.cc:83-89) */
    int cycle;        /* 0 = V (default), 1 = F */
    int periodic;     /* per_mask: bits x = 1, y = 2, z = 4 (point relaxation) */
} orc3_settings;

typedef struct orc3_solver orc3_solver;

int  orc3_num_levels(int nx, int ny, int nz, int min_coarse);
orc3_solver *orc3_solver_create(const double *so_fine, int nstncl, int nx, int ny, int nz,
                                const orc3_settings *s);
int  orc3_solver_nlevels(const orc3_solver *slv);
void orc3_solver_level_dims(const orc3_solver *slv, int lvl, int *nx, int *ny, int *nz);
void orc3_solver_get_level(const orc3_solver *slv, int lvl, double *so, double *ci);
void orc3_solver_vcycle(orc3_solver *slv, double *x, const double *b);
void orc3_solver_cycle(orc3_solver *slv, double *x, const double *b);
int  orc3_solver_solve(orc3_solver *slv, const double *b, double *x, double *hist, int maxhist);
/* one plane-relaxation sweep (all planes of direction dir, both colours) at level lvl */
void orc3_solver_relax_planes(orc3_solver *slv, int lvl, int dir, double *x, const double *b,
                              int updown);
void orc3_solver_destroy(orc3_solver *slv);

#ifdef __cplusplus
}
#endif
#endif

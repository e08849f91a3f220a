/*
 * Periodic-boundary branches of the 2D BoxMG kernels -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatements of the JPN != 0 branches of Cedar's serial 2D Fortran kernels, in the
 * reference's operation order.  JPN is the BoxMG boundary id BMG_get_bc returns for the periodic
 * mask (src/2d/ftn/BMG_get_bc.f90:1-24): 0 Dirichlet, 2 periodic in x, 1 periodic in y, 3 both
 * (include/cedar/2d/ftn/BMG_parameters_c.h:193-196).  With jpn == 0 every function here is the
 * Dirichlet kernel of bmg2_oracle.c.
 *
 * Indices in the macros are the reference's 1-based Fortran indices (ghosts at 1 and ii).
 */
#include "bmg2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5 };
enum { LL = 1, LR = 2, LA = 3, LB = 4, LSW = 5, LNW = 6, LNE = 7, LSE = 8 };
enum { PER_Y = 1, PER_X = 2, PER_XY = 3 };

#define V2(a, n1, i, j) (a)[(size_t)((j) - 1) * (size_t)(n1) + (size_t)((i) - 1)]
#define P2(a, n1, n2, i, j, k) \
    (a)[((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1)]

static int per_x(int jpn) { int p = abs(jpn); return p == PER_X || p == PER_XY; }
static int per_y(int jpn) { int p = abs(jpn); return p == PER_Y || p == PER_XY; }

int orc2_bc_from_mask(int mask)
{
    /* BMG_get_bc: bcmap(1) = def_per_x = 2, bcmap(2) = def_per_y = 1, bcmap(3) = def_per_xy = 3 */
    static const int map[4] = {0, 2, 1, 3};
    return (mask >= 0 && mask < 4) ? map[mask] : 0;
}

/* Q(1,J) = Q(I1,J), Q(II,J) = Q(2,J) for rows j0..j1 */
static void ghost_x(double *q, int ii, int j0, int j1)
{
    for (int j = j0; j <= j1; ++j) {
        V2(q, ii, 1, j) = V2(q, ii, ii - 1, j);
        V2(q, ii, ii, j) = V2(q, ii, 2, j);
    }
}
/* Q(I,1) = Q(I,J1), Q(I,JJ) = Q(I,2) for columns 1..ii */
static void ghost_y(double *q, int ii, int jj)
{
    for (int i = 1; i <= ii; ++i) {
        V2(q, ii, i, 1) = V2(q, ii, i, jj - 1);
        V2(q, ii, i, jj) = V2(q, ii, i, 2);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* point Gauss-Seidel, periodic branch: src/2d/ftn/BMG2_SymStd_relax_GS.f90:139-226.  The x    */
/* ghosts of a row are refreshed after each colour pass over it, the y ghosts once at the end. */
/* ------------------------------------------------------------------------------------------ */
void orc2_relax_gs_bc(const double *so, int nstncl, const double *qf, double *q,
                      const double *sor, int ii, int jj, int updown, int jpn)
{
    if (jpn == 0) {
        orc2_relax_gs(so, nstncl, qf, q, sor, ii, jj, updown);
        return;
    }
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
#define SOR(i, j) P2(sor, ii, jj, i, j, 2)
    const int i1 = ii - 1, j1 = jj - 1;
    const int lstart = (updown == ORC_DOWN) ? 2 : 3;
    const int lend = (updown == ORC_DOWN) ? 3 : 2;
    const int lstride = (updown == ORC_DOWN) ? 1 : -1;
    if (nstncl == 5) {
        for (int jbeg = lstart; lstride > 0 ? jbeg <= lend : jbeg >= lend; jbeg += lstride) {
            const int jend = 2 * ((j1 - jbeg) / 2) + jbeg;
            for (int j = jbeg; j <= jend; j += 2)
                for (int ibeg = lstart; lstride > 0 ? ibeg <= lend : ibeg >= lend; ibeg += lstride) {
                    const int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                    for (int i = ibeg; i <= iend; i += 2)
                        Q(i, j) = (QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                                   + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1)
                                   + SO(i, j, KSW) * Q(i - 1, j - 1)
                                   + SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                                   + SO(i, j + 1, KNW) * Q(i - 1, j + 1)
                                   + SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1))
                                  * SOR(i, j);
                    if (per_x(jpn)) ghost_x(q, ii, j, j);
                }
        }
    } else {
        for (int jo = lstart; lstride > 0 ? jo <= lend : jo >= lend; jo += lstride)
            for (int j = 2; j <= j1; ++j) {
                const int ibeg = (j + jo) % 2 + 2;
                const int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                for (int i = ibeg; i <= iend; i += 2)
                    Q(i, j) = (QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                               + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1))
                              * SOR(i, j);
                if (per_x(jpn)) ghost_x(q, ii, j, j);
            }
    }
    if (per_y(jpn)) ghost_y(q, ii, jj);
#undef SO
#undef Q
#undef QF
#undef SOR
}

/* ------------------------------------------------------------------------------------------ */
/* restriction, periodic branch: src/2d/ftn/BMG2_SymStd_restrict.f90:94-126.  When the fine    */
/* extent is even the ghosts of Q (the fine residual) are refreshed first, in place.           */
/* ------------------------------------------------------------------------------------------ */
void orc2_restrict_bc(const double *ci, double *q, double *qc, int iif, int jjf, int iic, int jjc,
                      int jpn)
{
    if (jpn != 0) {
        if (per_y(jpn) && jjf / 2 + 1 == jjc)
            for (int i = 1; i <= iif; ++i) {
                V2(q, iif, i, 1) = V2(q, iif, i, jjf - 1);
                V2(q, iif, i, jjf) = V2(q, iif, i, 2);
            }
        if (per_x(jpn) && iif / 2 + 1 == iic)
            for (int j = 1; j <= jjf; ++j) {
                V2(q, iif, 1, j) = V2(q, iif, iif - 1, j);
                V2(q, iif, iif, j) = V2(q, iif, 2, j);
            }
    }
    orc2_restrict(ci, q, qc, iif, jjf, iic, jjc);
}

/* interp_add, periodic branch: src/2d/ftn/BMG2_SymStd_interp_add.f90:139-157 (ghosts of the
 * corrected fine iterate refreshed after the interpolation, y first) */
void orc2_interp_add_bc(double *q, const double *qc, double *res, const double *so, int nstncl,
                        const double *ci, int iic, int jjc, int iif, int jjf, int jpn)
{
    orc2_interp_add(q, qc, res, so, nstncl, ci, iic, jjc, iif, jjf);
    if (per_y(jpn)) ghost_y(q, iif, jjf);
    if (per_x(jpn)) ghost_x(q, iif, 1, jjf);
}

/* ------------------------------------------------------------------------------------------ */
/* operator-induced interpolation, periodic branch:                                             */
/* src/2d/ftn/BMG2_SymStd_SETUP_interp_OI.f90:259-616                                          */
/* ------------------------------------------------------------------------------------------ */
/* I = MAX(MOD(I+2, N), MIN(I+2, 3)): the next fine index, wrapping to 3 at the periodic end */
static int wrap_next(int prev, int n)
{
    const int a = (prev + 2) % n, b = prev + 2 < 3 ? prev + 2 : 3;
    return a > b ? a : b;
}

static double lumped(double c, double sum, double ep, double emach)
{
    return (c - sum) * fmax(c - (1.0 + ep) * sum, 0.0) / (fabs(c - (1.0 + ep) * sum) + emach);
}

#define SO(i, j, k) P2(so, iif, jjf, i, j, k)
#define CI(i, j, k) P2(ci, iic, jjc, i, j, k)

/* LR/LL of coarse point (ic, jc) from fine point (i, j) (an x-edge) */
static void oi_xedge(const double *so, double *ci, int iif, int jjf, int iic, int jjc, int nine,
                     int i, int j, int ic, int jc, double emach)
{
    double a, b;
    if (nine) {
        a = SO(i, j, KW) + SO(i, j, KNW) + SO(i, j + 1, KSW);
        b = SO(i - 1, j, KW) + SO(i - 1, j, KSW) + SO(i - 1, j + 1, KNW);
    } else {
        a = SO(i, j, KW);
        b = SO(i - 1, j, KW);
    }
    const double c = SO(i - 1, j, KO);
    const double ep = fmin(fabs(a / c), fabs(b / c));
    double sum = a + b + SO(i - 1, j, KS) + SO(i - 1, j + 1, KS);
    sum = a + b + lumped(c, sum, ep, emach);
    sum = 1.0 / sum;
    CI(ic, jc, LR) = a * sum;
    CI(ic, jc, LL) = b * sum;
}

/* LA/LB of coarse point (ic, jc) from fine point (i, j) (a y-edge) */
static void oi_yedge(const double *so, double *ci, int iif, int jjf, int iic, int jjc, int nine,
                     int i, int j, int ic, int jc, double emach)
{
    double a, b;
    if (nine) {
        a = SO(i, j, KS) + SO(i, j, KNW) + SO(i + 1, j, KSW);
        b = SO(i, j - 1, KS) + SO(i, j - 1, KSW) + SO(i + 1, j - 1, KNW);
    } else {
        a = SO(i, j, KS);
        b = SO(i, j - 1, KS);
    }
    const double c = SO(i, j - 1, KO);
    double sum = a + b + SO(i, j - 1, KW) + SO(i + 1, j - 1, KW);
    const double ep = fmin(fabs(a / c), fabs(b / c));
    sum = a + b + lumped(c, sum, ep, emach);
    sum = 1.0 / sum;
    CI(ic, jc, LA) = a * sum;
    CI(ic, jc, LB) = b * sum;
}

/* LSW/LSE/LNW/LNE of coarse point (ic, jc) from the cell-centre fine point (i-1, j-1) */
static void oi_corner(const double *so, double *ci, int iif, int jjf, int iic, int jjc, int nine,
                      int i, int j, int ic, int jc, double emach)
{
    const double c = SO(i - 1, j - 1, KO);
    double sum, ep;
    if (nine) {
        sum = SO(i - 1, j - 1, KW) + SO(i - 1, j, KNW) + SO(i - 1, j, KS) + SO(i, j, KSW)
            + SO(i, j - 1, KW) + SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS) + SO(i - 1, j - 1, KSW);
        const double e1 = fabs((SO(i - 1, j - 1, KSW) + SO(i - 1, j - 1, KW) + SO(i - 1, j, KNW)) / c);
        const double e2 = fabs((SO(i - 1, j, KNW) + SO(i - 1, j, KS) + SO(i, j, KSW)) / c);
        const double e3 = fabs((SO(i, j, KSW) + SO(i, j - 1, KW) + SO(i, j - 1, KNW)) / c);
        const double e4 = fabs((SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS) + SO(i - 1, j - 1, KSW)) / c);
        ep = fmin(fmin(fmin(e1, e2), e3), e4);
    } else {
        sum = SO(i - 1, j - 1, KW) + SO(i - 1, j, KS) + SO(i, j - 1, KW) + SO(i - 1, j - 1, KS);
        const double e1 = fabs(SO(i - 1, j - 1, KW) / c);
        const double e2 = fabs(SO(i - 1, j, KS) / c);
        const double e3 = fabs(SO(i, j - 1, KW) / c);
        const double e4 = fabs(SO(i - 1, j - 1, KS) / c);
        ep = fmin(fmin(fmin(e1, e2), e3), e4);
    }
    sum = sum + lumped(c, sum, ep, emach);
    const double s = 1.0 / sum;
    if (nine) {
        CI(ic, jc, LSW) = (SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LL)
                           + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LB) + SO(i - 1, j - 1, KSW)) * s;
        CI(ic, jc, LSE) = (SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR)
                           + SO(i, j - 1, KW) * CI(ic, jc, LB) + SO(i, j - 1, KNW)) * s;
        CI(ic, jc, LNW) = (SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LA)
                           + SO(i - 1, j, KS) * CI(ic, jc, LL) + SO(i - 1, j, KNW)) * s;
        CI(ic, jc, LNE) = (SO(i - 1, j, KS) * CI(ic, jc, LR)
                           + SO(i, j - 1, KW) * CI(ic, jc, LA) + SO(i, j, KSW)) * s;
    } else {
        CI(ic, jc, LSW) = (SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LL)
                           + SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LB)) * s;
        CI(ic, jc, LSE) = (SO(i - 1, j - 1, KS) * CI(ic, jc - 1, LR)
                           + SO(i, j - 1, KW) * CI(ic, jc, LB)) * s;
        CI(ic, jc, LNW) = (SO(i - 1, j - 1, KW) * CI(ic - 1, jc, LA)
                           + SO(i - 1, j, KS) * CI(ic, jc, LL)) * s;
        CI(ic, jc, LNE) = (SO(i - 1, j, KS) * CI(ic, jc, LR)
                           + SO(i, j - 1, KW) * CI(ic, jc, LA)) * s;
    }
}
#undef SO
#undef CI

void orc2_setup_interp_bc(const double *so, int nstncl, double *ci, int iif, int jjf, int iic,
                          int jjc, int jpn)
{
    if (jpn == 0) {
        orc2_setup_interp(so, nstncl, ci, iif, jjf, iic, jjc);
        return;
    }
    const double emach = ldexp(1.0, -53);
    const int nine = nstncl == 5;
    const int iic1 = iic - 1, jjc1 = jjc - 1, iif1 = iif - 1, iif2 = iif - 2;
    const int jjf1 = jjf - 1, jjf2 = jjf - 2;
    const int iicf = (iif - 2) / 2 + 3, jjcf = (jjf - 2) / 2 + 3;
    const int iicf1 = iicf - 1, jjcf1 = jjcf - 1;
    const int iifc = 2 * (iicf - 2) + 2, jjfc = 2 * (jjcf - 2) + 2;
    int ibeg = 2, ibegc = 3, ibeg_x = 2, iend_x = iif2, iendc = iicf1, jendc = jjcf1;
    if (per_x(jpn)) {
        ibeg = 0;
        ibegc = 2;
        if (iic == iicf) {
            iendc = iicf;
            ibeg_x = 3;
            iend_x = iif1;
        }
    }
    int jbeg = 2, jbegc = 3, jbeg_y = 2, jend_y = jjf2;
    if (per_y(jpn)) {
        jbeg = 0;
        jbegc = 2;
        if (jjc == jjcf) {
            jendc = jjcf;
            jbeg_y = 3;
            jend_y = jjf1;
        }
    }
    /* x-edges */
    int j = 0;
    for (int jc = 2; jc <= jjc1; ++jc) {
        j += 2;
        int i = ibeg;
        for (int ic = ibegc; ic <= iendc; ++ic) {
            i = wrap_next(i, iifc);
            oi_xedge(so, ci, iif, jjf, iic, jjc, nine, i, j, ic, jc, emach);
        }
    }
    if (per_y(jpn)) {
        int i = ibeg;
        for (int ic = ibegc; ic <= iendc; ++ic) {
            i = wrap_next(i, iifc);
            oi_xedge(so, ci, iif, jjf, iic, jjc, nine, i, jbeg_y, ic, jjc, emach);
        }
        i = ibeg;
        for (int ic = ibegc; ic <= iendc; ++ic) {
            i = wrap_next(i, iifc);
            oi_xedge(so, ci, iif, jjf, iic, jjc, nine, i, jend_y, ic, 1, emach);
        }
    }
    /* y-edges */
    j = jbeg;
    for (int jc = jbegc; jc <= jendc; ++jc) {
        j = wrap_next(j, jjfc);
        int i = 0;
        for (int ic = 2; ic <= iic1; ++ic) {
            i += 2;
            oi_yedge(so, ci, iif, jjf, iic, jjc, nine, i, j, ic, jc, emach);
        }
    }
    if (per_x(jpn)) {
        j = jbeg;
        for (int jc = jbegc; jc <= jendc; ++jc) {
            j = wrap_next(j, jjfc);
            oi_yedge(so, ci, iif, jjf, iic, jjc, nine, ibeg_x, j, iic, jc, emach);
        }
        j = jbeg;
        for (int jc = jbegc; jc <= jendc; ++jc) {
            j = wrap_next(j, jjfc);
            oi_yedge(so, ci, iif, jjf, iic, jjc, nine, iend_x, j, 1, jc, emach);
        }
    }
    /* corners */
    j = jbeg;
    for (int jc = jbegc; jc <= jendc; ++jc) {
        j = wrap_next(j, jjfc);
        int i = ibeg;
        for (int ic = ibegc; ic <= iendc; ++ic) {
            i = wrap_next(i, iifc);
            oi_corner(so, ci, iif, jjf, iic, jjc, nine, i, j, ic, jc, emach);
        }
    }
}

/* Galerkin coarse operator, periodic branch: the Dirichlet products followed by the periodic
 * copies of the coarse operator's ghost rows and columns (src/2d/ftn/BMG2_SymStd_SETUP_ITLI_ex.f90
 * :216-246 for a 9-point fine operator, :333-363 for a 5-point one; y first, then x) */
void orc2_setup_galerkin_bc(const double *so, int nstncl, const double *ci, double *soc, int iif,
                            int jjf, int iic, int jjc, int jpn)
{
    orc2_setup_galerkin(so, nstncl, ci, soc, iif, jjf, iic, jjc);
#define SOC(i, j, k) P2(soc, iic, jjc, i, j, k)
    static const int order[5] = {KS, KW, KNW, KSW, KO};
    if (per_y(jpn))
        for (int ic = 1; ic <= iic; ++ic) {
            for (int t = 0; t < 5; ++t) SOC(ic, jjc, order[t]) = SOC(ic, 2, order[t]);
            for (int t = 0; t < 5; ++t) SOC(ic, 1, order[t]) = SOC(ic, jjc - 1, order[t]);
        }
    if (per_x(jpn))
        for (int jc = 1; jc <= jjc; ++jc) {
            for (int t = 0; t < 5; ++t) SOC(iic, jc, order[t]) = SOC(2, jc, order[t]);
            for (int t = 0; t < 5; ++t) SOC(1, jc, order[t]) = SOC(iic - 1, jc, order[t]);
        }
#undef SOC
}

/* ------------------------------------------------------------------------------------------ */
/* coarsest grid, periodic: the dense upper triangle of A assembled as                          */
/* src/2d/ftn/BMG2_SymStd_SETUP_cg_LU.f90:119-227 (9-point) and :251-335 (5-point) write it,    */
/* factored by a Cholesky in DPOTF2's order; the solve (src/2d/ftn/BMG2_SymStd_SOLVE_cg.f90    */
/* :93-163) removes the mean of the coarse solution and refreshes its ghosts.                   */
/* ------------------------------------------------------------------------------------------ */
/* upper Cholesky A = U^T U of the n x n column-major array a (lda >= n), DPOTF2's order */
int orc_dpotrf_u(int n, double *a, int lda)
{
#define A(r, c) a[(size_t)((c) - 1) * (size_t)lda + (size_t)((r) - 1)]
    for (int j = 1; j <= n; ++j) {
        double ajj = A(j, j);
        for (int k = 1; k < j; ++k) ajj -= A(k, j) * A(k, j);
        if (ajj <= 0.0 || isnan(ajj)) {
            A(j, j) = ajj;
            return j;
        }
        ajj = sqrt(ajj);
        A(j, j) = ajj;
        for (int c = j + 1; c <= n; ++c) {
            double s = A(j, c);
            for (int k = 1; k < j; ++k) s -= A(k, j) * A(k, c);
            A(j, c) = s / ajj;
        }
    }
    return 0;
#undef A
}

/* solve U^T U x = b in place */
void orc_dpotrs_u(int n, const double *a, int lda, double *b)
{
#define A(r, c) a[(size_t)((c) - 1) * (size_t)lda + (size_t)((r) - 1)]
    for (int j = 1; j <= n; ++j) {
        double s = b[j - 1];
        for (int k = 1; k < j; ++k) s -= A(k, j) * b[k - 1];
        b[j - 1] = s / A(j, j);
    }
    for (int j = n; j >= 1; --j) {
        double s = b[j - 1];
        for (int k = j + 1; k <= n; ++k) s -= A(j, k) * b[k - 1];
        b[j - 1] = s / A(j, j);
    }
#undef A
}

/* the dense periodic matrix (ld = n = nx*ny), zero on entry */
void orc2_assemble_cg_periodic(const double *so, int nstncl, double *abd, int ii, int jj, int jpn)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
    const int i1 = ii - 1, j1 = jj - 1, i2 = i1 - 1, n = i2 * (j1 - 1);
#define ABD(r, c) abd[(size_t)((c) - 1) * (size_t)n + (size_t)((r) - 1)]
    const int nine = nstncl == 5;
    int kk = 1;
    ABD(1, 1) = SO(2, 2, KO);
    for (int i = 3; i <= i1; ++i) {
        ++kk;
        ABD(kk, kk) = SO(i, 2, KO);
        ABD(kk - 1, kk) = -SO(i, 2, KW);
    }
    if (per_x(jpn)) ABD(kk - i2 + 1, kk) = -SO(ii, 2, KW);
    for (int j = 3; j <= j1; ++j) {
        if (per_x(jpn)) ABD(kk, kk + 1) = nine ? -SO(2, j, KSW) : 0.0;
        for (int i = 2; i <= i1; ++i) {
            ++kk;
            ABD(kk, kk) = SO(i, j, KO);
            if (i != 2) {
                ABD(kk - 1, kk) = -SO(i, j, KW);
                ABD(kk - i2 - 1, kk) = nine ? -SO(i, j, KSW) : 0.0;
            }
            ABD(kk - i2 + 1, kk) = nine ? -SO(i + 1, j, KNW) : 0.0;
            ABD(kk - i2, kk) = -SO(i, j, KS);
        }
        if (per_x(jpn)) {
            ABD(kk - i2 + 1, kk) = -SO(ii, j, KW);
            ABD(kk - 2 * i2 + 1, kk) = nine ? -SO(ii, j, KNW) : 0.0;
        }
    }
    if (per_y(jpn)) {
        kk = kk - i2;
        const int j2 = (j1 - 2) * i2;
        ++kk;
        ABD(kk - j2, kk) = -SO(2, jj, KS);
        ABD(kk - j2 + 1, kk) = nine ? -SO(3, jj, KSW) : 0.0;
        if (abs(jpn) == PER_XY) ABD(i2, kk) = nine ? -SO(2, jj, KNW) : 0.0;
        for (int i = 3; i <= i1; ++i) {
            ++kk;
            ABD(kk - j2, kk) = -SO(i, jj, KS);
            ABD(kk - j2 - 1, kk) = nine ? -SO(i, jj, KNW) : 0.0;
            ABD(kk - j2 + 1, kk) = nine ? -SO(i + 1, jj, KSW) : 0.0;
        }
        ABD(kk - j2 + 1, kk) = 0.0;
        if (abs(jpn) == PER_XY) ABD(1, kk) = nine ? -SO(ii, jj, KSW) : 0.0;
        if (per_x(jpn)) ABD(kk - 2 * i2 + 1, kk) = nine ? -SO(ii, j1, KNW) : 0.0;
    }
#undef ABD
#undef SO
}

int orc2_setup_cg_bc(const double *so, int nstncl, double *abd, int ii, int jj, int jpn)
{
    if (jpn == 0) return orc2_setup_cg(so, nstncl, abd, ii, jj);
    const int n = (ii - 2) * (jj - 2);
    memset(abd, 0, sizeof(double) * (size_t)n * (size_t)n);
    orc2_assemble_cg_periodic(so, nstncl, abd, ii, jj, jpn);
    return orc_dpotrf_u(n, abd, n);
}

int orc2_solve_cg_bc(double *q, const double *qf, const double *abd, double *bbd, int ii, int jj,
                     int jpn)
{
    if (jpn == 0) return orc2_solve_cg(q, qf, abd, bbd, ii, jj);
    const int i1 = ii - 1, j1 = jj - 1, n = (ii - 2) * (jj - 2);
    int kk = 0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) bbd[kk++] = V2(qf, ii, i, j);
    orc_dpotrs_u(n, abd, n, bbd);
    kk = 0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) V2(q, ii, i, j) = bbd[kk++];
    double cint = 0.0, qint = 0.0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) {
            qint = qint + V2(q, ii, i, j);
            cint = cint + 1.0;
        }
    const double c = -qint / cint;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) V2(q, ii, i, j) = V2(q, ii, i, j) + c;
    if (per_y(jpn))
        for (int i = 2; i <= i1; ++i) {
            V2(q, ii, i, jj) = V2(q, ii, i, 2);
            V2(q, ii, i, 1) = V2(q, ii, i, j1);
        }
    if (per_x(jpn))
        for (int j = 2; j <= j1; ++j) {
            V2(q, ii, ii, j) = V2(q, ii, 2, j);
            V2(q, ii, 1, j) = V2(q, ii, i1, j);
        }
    if (abs(jpn) == PER_XY) {
        V2(q, ii, 1, 1) = V2(q, ii, i1, j1);
        V2(q, ii, ii, 1) = V2(q, ii, 2, j1);
        V2(q, ii, 1, jj) = V2(q, ii, i1, 2);
        V2(q, ii, ii, jj) = V2(q, ii, 2, 2);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* line relaxation, periodic: the factors of a periodic line get the wrap couplings added to   */
/* their end diagonals (src/2d/ftn/BMG2_SymStd_SETUP_lines_x.f90:75-82, _y.f90:77-84), and a    */
/* line solve becomes a Sherman-Morrison correction of the tridiagonal one                      */
/* (src/2d/ftn/BMG2_SymStd_relax_lines_x.f90:170-300, _y.f90:176-310).  Lines across the other  */
/* direction's periodicity run the Dirichlet branch and refresh that direction's images after   */
/* the sweep.                                                                                    */
/* ------------------------------------------------------------------------------------------ */
int orc2_setup_lines_x_bc(const double *so, double *sor, int ii, int jj, int jpn)
{
    if (!per_x(jpn)) return orc2_setup_lines_x(so, sor, ii, jj);
    int info = 0;
    for (int j = 2; j <= jj - 1; ++j)
        for (int i = 2; i <= ii - 1; ++i) {
            P2(sor, ii, jj, i, j, 2) = -P2(so, ii, jj, i, j, KW);
            P2(sor, ii, jj, i, j, 1) = P2(so, ii, jj, i, j, KO);
        }
    for (int j = 2; j <= jj - 1; ++j) {
        P2(sor, ii, jj, 2, j, 1) = P2(sor, ii, jj, 2, j, 1) + P2(so, ii, jj, 2, j, KW);
        P2(sor, ii, jj, ii - 1, j, 1) = P2(sor, ii, jj, ii - 1, j, 1) + P2(so, ii, jj, ii, j, KW);
    }
    for (int j = 2; j <= jj - 1; ++j) {
        int r = orc_dpttrf(ii - 2, &P2(sor, ii, jj, 2, j, 1), &P2(sor, ii, jj, 3, j, 2));
        if (r && !info) info = r;
    }
    return info;
}

int orc2_setup_lines_y_bc(const double *so, double *sor, int ii, int jj, int jpn)
{
    if (!per_y(jpn)) return orc2_setup_lines_y(so, sor, ii, jj);
    int info = 0;
    for (int i = 2; i <= ii - 1; ++i)
        for (int j = 2; j <= jj - 1; ++j) {
            P2(sor, jj, ii, j, i, 2) = -P2(so, ii, jj, i, j, KS);
            P2(sor, jj, ii, j, i, 1) = P2(so, ii, jj, i, j, KO);
        }
    for (int i = 2; i <= ii - 1; ++i) {
        P2(sor, jj, ii, 2, i, 1) = P2(sor, jj, ii, 2, i, 1) + P2(so, ii, jj, i, 2, KS);
        P2(sor, jj, ii, jj - 1, i, 1) = P2(sor, jj, ii, jj - 1, i, 1) + P2(so, ii, jj, i, jj, KS);
    }
    for (int i = 2; i <= ii - 1; ++i) {
        int r = orc_dpttrf(jj - 2, &P2(sor, jj, ii, 2, i, 1), &P2(sor, jj, ii, 3, i, 2));
        if (r && !info) info = r;
    }
    return info;
}

/* b: scratch of >= 2*max(ii, jj) doubles */
void orc2_relax_lines_x_bc(const double *so, int nstncl, const double *qf, double *q,
                           const double *sor, double *b, int ii, int jj, int updown, int jpn)
{
    if (!per_x(jpn)) {
        orc2_relax_lines_x(so, nstncl, qf, q, sor, ii, jj, updown);
        if (per_y(jpn)) ghost_y(q, ii, jj);
        return;
    }
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
    const int i1 = ii - 1, j1 = jj - 1;
    const int first = (updown == ORC_DOWN) ? 3 : 2;
    for (int pass = 0; pass < 2; ++pass) {
        const int jbeg = (pass == 0) ? first : 5 - first;
        const int jend = 2 * ((j1 - jbeg) / 2) + jbeg;
        for (int j = jbeg; j <= jend; j += 2) {
            for (int i = 2; i <= i1; ++i) {
                if (nstncl == 5)
                    Q(i, j) = QF(i, j) + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1)
                            + SO(i, j, KSW) * Q(i - 1, j - 1) + SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                            + SO(i, j + 1, KNW) * Q(i - 1, j + 1)
                            + SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1);
                else
                    Q(i, j) = QF(i, j) + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1);
            }
            const double *d = &P2(sor, ii, jj, 2, j, 1), *e = &P2(sor, ii, jj, 3, j, 2);
            orc_dpttrs(i1 - 1, d, e, &Q(2, j));
            for (int i = 2; i <= i1; ++i) b[i - 1] = 0.0;
            b[1] = -SO(2, j, KW);
            b[i1 - 1] = -SO(ii, j, KW);
            orc_dpttrs(i1 - 1, d, e, &b[1]);
            const double alpha = b[1] + b[i1 - 1];
            double beta = Q(2, j) + Q(i1, j);
            beta = beta / (1.0 + alpha);
            for (int i = 2; i <= i1; ++i) Q(i, j) = Q(i, j) - beta * b[i - 1];
        }
        if (per_y(jpn)) ghost_y(q, ii, jj);
        ghost_x(q, ii, 1, jj);
    }
#undef SO
#undef Q
#undef QF
}

void orc2_relax_lines_y_bc(const double *so, int nstncl, const double *qf, double *q,
                           const double *sor, double *b, int ii, int jj, int updown, int jpn)
{
    if (!per_y(jpn)) {
        orc2_relax_lines_y(so, nstncl, qf, q, sor, b, ii, jj, updown);
        if (per_x(jpn)) ghost_x(q, ii, 1, jj);
        return;
    }
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
    const int i1 = ii - 1, j1 = jj - 1;
    double *z = b + jj; /* B(JJ+J) */
    const int first = (updown == ORC_DOWN) ? 3 : 2;
    for (int pass = 0; pass < 2; ++pass) {
        const int ibeg = (pass == 0) ? first : 5 - first;
        const int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
        for (int i = ibeg; i <= iend; i += 2) {
            for (int j = 2; j <= j1; ++j) {
                if (nstncl == 5)
                    b[j - 1] = QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                             + SO(i, j, KSW) * Q(i - 1, j - 1) + SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                             + SO(i, j + 1, KNW) * Q(i - 1, j + 1)
                             + SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1);
                else
                    b[j - 1] = QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j);
            }
            const double *d = &P2(sor, jj, ii, 2, i, 1), *e = &P2(sor, jj, ii, 3, i, 2);
            orc_dpttrs(j1 - 1, d, e, &b[1]);
            for (int j = 2; j <= j1; ++j) z[j - 1] = 0.0;
            z[1] = -SO(i, 2, KS);
            z[j1 - 1] = -SO(i, jj, KS);
            orc_dpttrs(j1 - 1, d, e, &z[1]);
            const double alpha = z[1] + z[j1 - 1];
            double beta = b[1] + b[j1 - 1];
            beta = beta / (1.0 + alpha);
            for (int j = 2; j <= j1; ++j) Q(i, j) = b[j - 1] - beta * z[j - 1];
        }
        ghost_y(q, ii, jj);
        if (per_x(jpn)) ghost_x(q, ii, 1, jj);
    }
#undef SO
#undef Q
#undef QF
}

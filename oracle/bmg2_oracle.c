/*
 * FP64 CPU oracle for the 2D BoxMG solver path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Cedar's serial 2D kernels (the .f90 files of src/2d/ftn) and of the multilevel
 * driver (include/cedar/multilevel.h, include/cedar/cycle/vcycle.h, include/cedar/2d/solver.h).
 * Every kernel keeps the reference's floating-point operation order so that, built with
 * -ffp-contract=off, it is bitwise identical to the reference kernels compiled from their own
 * sources.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Index convention: the macros below take the reference's 1-based (Fortran) indices, where index
 * 1 is the ghost layer and the first interior point is 2.
 */
#include "bmg2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* stencil planes (src/2d/ftn/BMG_stencils_f90.h:29-30) */
enum { KO = 1, KW = 2, KS = 3, KSW = 4, KNW = 5 };
/* interpolation planes (src/2d/ftn/BMG_stencils_f90.h:35-36) */
enum { LL = 1, LR = 2, LA = 3, LB = 4, LSW = 5, LNW = 6, LNE = 7, LSE = 8 };

#define V2(a, n1, i, j) (a)[(size_t)((j) - 1) * (size_t)(n1) + (size_t)((i) - 1)]
#define P2(a, n1, n2, i, j, k) \
    (a)[((size_t)((k) - 1) * (size_t)(n2) + (size_t)((j) - 1)) * (size_t)(n1) + (size_t)((i) - 1)]

/* ------------------------------------------------------------------------------------------ */
/* residual: src/2d/ftn/BMG2_SymStd_residual.f90:85-119                                         */
/* ------------------------------------------------------------------------------------------ */
void orc2_residual(const double *so, int nstncl, const double *qf, const double *q,
                   double *res, int ii, int jj)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
#define RES(i, j) V2(res, ii, i, j)
    int i1 = ii - 1, j1 = jj - 1;
    if (nstncl == 5) {
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i)
                RES(i, j) = QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                          + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1)
                          + SO(i, j, KSW) * Q(i - 1, j - 1) + SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                          + SO(i, j + 1, KNW) * Q(i - 1, j + 1)
                          + SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1) - SO(i, j, KO) * Q(i, j);
    } else {
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i)
                RES(i, j) = QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                          + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1)
                          - SO(i, j, KO) * Q(i, j);
    }
#undef SO
#undef Q
#undef QF
#undef RES
}

/* l2 norm over the interior, sequential in storage order: include/cedar/2d/grid_func.h:43-54 */
double orc2_norm(const double *x, int ii, int jj)
{
    double s = 0.0;
    for (int j = 2; j <= jj - 1; ++j)
        for (int i = 2; i <= ii - 1; ++i) {
            double v = V2(x, ii, i, j);
            s += v * v;
        }
    return sqrt(s);
}

/* y = A q: src/2d/ftn/mpi/BMG2_SymStd_UTILS_matvec.f90 (9-point branch for a 5-plane
 * operator, 5-point branch for a 3-plane one), terms summed left to right as written there */
void orc2_matvec(const double *so, int nstncl, const double *q, double *y, int ii, int jj)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
    for (int j = 2; j <= jj - 1; ++j)
        for (int i = 2; i <= ii - 1; ++i) {
            double v = SO(i, j, KO) * Q(i, j) - SO(i, j, KW) * Q(i - 1, j)
                     - SO(i + 1, j, KW) * Q(i + 1, j) - SO(i, j, KS) * Q(i, j - 1)
                     - SO(i, j + 1, KS) * Q(i, j + 1);
            if (nstncl == 5)
                v = v - SO(i, j, KSW) * Q(i - 1, j - 1) - SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                      - SO(i, j + 1, KNW) * Q(i - 1, j + 1) - SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1);
            V2(y, ii, i, j) = v;
        }
#undef SO
#undef Q
}

/* SOR(:,:,msor=2) = 1/SO_c: src/2d/ftn/BMG2_SymStd_SETUP_recip.f90:63-67 */
void orc2_setup_recip(const double *so, double *sor, int ii, int jj)
{
    for (int j = 2; j <= jj - 1; ++j)
        for (int i = 2; i <= ii - 1; ++i)
            P2(sor, ii, jj, i, j, 2) = 1.0 / P2(so, ii, jj, i, j, KO);
}

/* ------------------------------------------------------------------------------------------ */
/* point Gauss-Seidel: src/2d/ftn/BMG2_SymStd_relax_GS.f90:78-137 (non-periodic branch)        */
/* ------------------------------------------------------------------------------------------ */
void orc2_relax_gs(const double *so, int nstncl, const double *qf, double *q,
                   const double *sor, int ii, int jj, int updown)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
#define SOR(i, j) P2(sor, ii, jj, i, j, 2)
    int i1 = ii - 1, j1 = jj - 1;
    /* symmetric relaxation: DOWN runs colours forward, UP in reverse */
    int lstart = (updown == ORC_DOWN) ? 2 : 3;
    int lend = (updown == ORC_DOWN) ? 3 : 2;
    int lstride = (updown == ORC_DOWN) ? 1 : -1;
    if (nstncl == 5) {
        for (int jbeg = lstart; lstride > 0 ? jbeg <= lend : jbeg >= lend; jbeg += lstride) {
            int jend = 2 * ((j1 - jbeg) / 2) + jbeg;
            for (int j = jbeg; j <= jend; j += 2)
                for (int ibeg = lstart; lstride > 0 ? ibeg <= lend : ibeg >= lend; ibeg += lstride) {
                    int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                    for (int i = ibeg; i <= iend; i += 2)
                        Q(i, j) = (QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                                   + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1)
                                   + SO(i, j, KSW) * Q(i - 1, j - 1)
                                   + SO(i + 1, j, KNW) * Q(i + 1, j - 1)
                                   + SO(i, j + 1, KNW) * Q(i - 1, j + 1)
                                   + SO(i + 1, j + 1, KSW) * Q(i + 1, j + 1))
                                  * SOR(i, j);
                }
        }
    } else {
        for (int jo = lstart; lstride > 0 ? jo <= lend : jo >= lend; jo += lstride)
            for (int j = 2; j <= j1; ++j) {
                int ibeg = (j + jo) % 2 + 2;
                int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                for (int i = ibeg; i <= iend; i += 2)
                    Q(i, j) = (QF(i, j) + SO(i, j, KW) * Q(i - 1, j) + SO(i + 1, j, KW) * Q(i + 1, j)
                               + SO(i, j, KS) * Q(i, j - 1) + SO(i, j + 1, KS) * Q(i, j + 1))
                              * SOR(i, j);
            }
    }
#undef SO
#undef Q
#undef QF
#undef SOR
}

/* ------------------------------------------------------------------------------------------ */
/* LAPACK restatements (reference LAPACK 3.x, published algorithms)                            */
/* ------------------------------------------------------------------------------------------ */

/* DPTTRF: L*D*L^T of an SPD tridiagonal matrix; d[0..n-1] diagonal, e[0..n-2] off-diagonal.
   On exit e holds the unit-lower multipliers and d the pivots.  Returns INFO. */
int orc_dpttrf(int n, double *d, double *e)
{
    for (int i = 0; i < n - 1; ++i) {
        if (d[i] <= 0.0) return i + 1;
        double ei = e[i];
        e[i] = ei / d[i];
        d[i + 1] = d[i + 1] - e[i] * ei;
    }
    if (n > 0 && d[n - 1] <= 0.0) return n;
    return 0;
}

/* DPTTRS -> DPTTS2 with one right-hand side: forward L, then D*L^T backward. */
void orc_dpttrs(int n, const double *d, const double *e, double *b)
{
    if (n <= 1) {
        if (n == 1) b[0] = b[0] * (1.0 / d[0]); /* DSCAL by 1/D(1) */
        return;
    }
    for (int i = 1; i < n; ++i) b[i] = b[i] - b[i - 1] * e[i - 1];
    b[n - 1] = b[n - 1] / d[n - 1];
    for (int i = n - 2; i >= 0; --i) b[i] = b[i] / d[i] - b[i + 1] * e[i];
}

/* DPBTRF('U') for KD smaller than the block size: the unblocked DPBTF2 (DSCAL + DSYR). */
int orc_dpbtrf_u(int n, int kd, double *ab, int ldab)
{
#define AB(r, c) ab[(size_t)((c) - 1) * (size_t)ldab + (size_t)((r) - 1)]
    for (int j = 1; j <= n; ++j) {
        double ajj = AB(kd + 1, j);
        if (ajj <= 0.0) return j;
        ajj = sqrt(ajj);
        AB(kd + 1, j) = ajj;
        int kn = (kd < n - j) ? kd : n - j;
        if (kn > 0) {
            double r = 1.0 / ajj;
            for (int t = 1; t <= kn; ++t) AB(kd + 1 - t, j + t) *= r;
            /* DSYR upper, alpha = -1, x(t) = A(j, j+t), M(p,q) = A(j+p, j+q) */
            for (int qq = 1; qq <= kn; ++qq) {
                double xq = AB(kd + 1 - qq, j + qq);
                if (xq != 0.0) {
                    double temp = -1.0 * xq;
                    for (int p = 1; p <= qq; ++p)
                        AB(kd + 1 + p - qq, j + qq) = AB(kd + 1 + p - qq, j + qq)
                                                      + AB(kd + 1 - p, j + p) * temp;
                }
            }
        }
    }
    return 0;
#undef AB
}

/* DPBTRS('U', NRHS=1): DTBSV(U,T,N) then DTBSV(U,N,N). */
void orc_dpbtrs_u(int n, int kd, const double *ab, int ldab, double *x)
{
#define AB(r, c) ab[(size_t)((c) - 1) * (size_t)ldab + (size_t)((r) - 1)]
#define X(i) x[(i) - 1]
    int kp1 = kd + 1;
    for (int j = 1; j <= n; ++j) {
        double temp = X(j);
        int l = kp1 - j;
        int i0 = (j - kd > 1) ? j - kd : 1;
        for (int i = i0; i <= j - 1; ++i) temp = temp - AB(l + i, j) * X(i);
        temp = temp / AB(kp1, j);
        X(j) = temp;
    }
    for (int j = n; j >= 1; --j) {
        if (X(j) != 0.0) {
            int l = kp1 - j;
            X(j) = X(j) / AB(kp1, j);
            double temp = X(j);
            int i0 = (j - kd > 1) ? j - kd : 1;
            for (int i = j - 1; i >= i0; --i) X(i) = X(i) - temp * AB(l + i, j);
        }
    }
#undef AB
#undef X
}

/* ------------------------------------------------------------------------------------------ */
/* line relaxation                                                                             */
/* ------------------------------------------------------------------------------------------ */

/* src/2d/ftn/BMG2_SymStd_SETUP_lines_x.f90:66-87: SOR(:,:,1)=SO_c, SOR(:,:,2)=-SO_w, DPTTRF/row */
int orc2_setup_lines_x(const double *so, double *sor, int ii, int jj)
{
    int info = 0;
    for (int j = 2; j <= jj - 1; ++j)
        for (int i = 2; i <= ii - 1; ++i) {
            P2(sor, ii, jj, i, j, 2) = -P2(so, ii, jj, i, j, KW);
            P2(sor, ii, jj, i, j, 1) = P2(so, ii, jj, i, j, KO);
        }
    for (int j = 2; j <= jj - 1; ++j) {
        int r = orc_dpttrf(ii - 2, &P2(sor, ii, jj, 2, j, 1), &P2(sor, ii, jj, 3, j, 2));
        if (r && !info) info = r;
    }
    return info;
}

/* src/2d/ftn/BMG2_SymStd_SETUP_lines_y.f90:69-88: factors stored transposed, SOR(jj, ii, 2) */
int orc2_setup_lines_y(const double *so, double *sor, int ii, int jj)
{
    int info = 0;
    for (int i = 2; i <= ii - 1; ++i)
        for (int j = 2; j <= jj - 1; ++j) {
            P2(sor, jj, ii, j, i, 2) = -P2(so, ii, jj, i, j, KS);
            P2(sor, jj, ii, j, i, 1) = P2(so, ii, jj, i, j, KO);
        }
    for (int i = 2; i <= ii - 1; ++i) {
        int r = orc_dpttrf(jj - 2, &P2(sor, jj, ii, 2, i, 1), &P2(sor, jj, ii, 3, i, 2));
        if (r && !info) info = r;
    }
    return info;
}

/* zebra x-lines: src/2d/ftn/BMG2_SymStd_relax_lines_x.f90:82-160 */
void orc2_relax_lines_x(const double *so, int nstncl, const double *qf, double *q,
                        const double *sor, int ii, int jj, int updown)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
    int i1 = ii - 1, j1 = jj - 1;
    /* DOWN: odd J first (3,5,...), UP: even J first */
    int first = (updown == ORC_DOWN) ? 3 : 2;
    for (int pass = 0; pass < 2; ++pass) {
        int jbeg = (pass == 0) ? first : 5 - first;
        int jend = 2 * ((j1 - jbeg) / 2) + jbeg;
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
        }
        for (int j = jbeg; j <= jend; j += 2)
            orc_dpttrs(i1 - 1, &P2(sor, ii, jj, 2, j, 1), &P2(sor, ii, jj, 3, j, 2), &Q(2, j));
    }
#undef SO
#undef Q
#undef QF
}

/* zebra y-lines: src/2d/ftn/BMG2_SymStd_relax_lines_y.f90:77-175; b is a scratch of >= jj */
void orc2_relax_lines_y(const double *so, int nstncl, const double *qf, double *q,
                        const double *sor, double *b, int ii, int jj, int updown)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define Q(i, j) V2(q, ii, i, j)
#define QF(i, j) V2(qf, ii, i, j)
    int i1 = ii - 1, j1 = jj - 1;
    int first = (updown == ORC_DOWN) ? 3 : 2;
    for (int pass = 0; pass < 2; ++pass) {
        int ibeg = (pass == 0) ? first : 5 - first;
        int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
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
            orc_dpttrs(j1 - 1, &P2(sor, jj, ii, 2, i, 1), &P2(sor, jj, ii, 3, i, 2), &b[1]);
            for (int j = 2; j <= j1; ++j) Q(i, j) = b[j - 1];
        }
    }
#undef SO
#undef Q
#undef QF
}

/* ------------------------------------------------------------------------------------------ */
/* restriction QC = P^T Q: src/2d/ftn/BMG2_SymStd_restrict.f90:76-92                          */
/* ------------------------------------------------------------------------------------------ */
void orc2_restrict(const double *ci, const double *q, double *qc,
                   int iif, int jjf, int iic, int jjc)
{
#define CI(i, j, k) P2(ci, iic, jjc, i, j, k)
#define Q(i, j) V2(q, iif, i, j)
#define QC(i, j) V2(qc, iic, i, j)
    (void)jjf;
    for (int jc = 2; jc <= jjc - 1; ++jc) {
        int j = 2 * jc - 2;
        for (int ic = 2; ic <= iic - 1; ++ic) {
            int i = 2 * ic - 2;
            QC(ic, jc) = CI(ic, jc, LNE) * Q(i - 1, j - 1) + CI(ic, jc, LA) * Q(i, j - 1)
                       + CI(ic + 1, jc, LNW) * Q(i + 1, j - 1) + CI(ic, jc, LR) * Q(i - 1, j)
                       + Q(i, j) + CI(ic + 1, jc, LL) * Q(i + 1, j)
                       + CI(ic, jc + 1, LSE) * Q(i - 1, j + 1) + CI(ic, jc + 1, LB) * Q(i, j + 1)
                       + CI(ic + 1, jc + 1, LSW) * Q(i + 1, j + 1);
        }
    }
#undef CI
#undef Q
#undef QC
}

/* ------------------------------------------------------------------------------------------ */
/* interp_add: src/2d/ftn/BMG2_SymStd_interp_add.f90:101-137 (RES is divided in place)        */
/* ------------------------------------------------------------------------------------------ */
void orc2_interp_add(double *q, const double *qc, double *res, const double *so, int nstncl,
                     const double *ci, int iic, int jjc, int iif, int jjf)
{
#define CI(i, j, k) P2(ci, iic, jjc, i, j, k)
#define Q(i, j) V2(q, iif, i, j)
#define QC(i, j) V2(qc, iic, i, j)
#define RES(i, j) V2(res, iif, i, j)
    (void)nstncl;
    int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2;
    for (int j = 2; j <= jjf - 1; ++j)
        for (int i = 2; i <= iif - 1; ++i) RES(i, j) = RES(i, j) / P2(so, iif, jjf, i, j, KO);
    int j = 2, i = 2;
    Q(2, j) = Q(2, j) + QC(2, 2);
    for (int ic = 3; ic <= iicf1; ++ic) {
        i += 2;
        Q(i, j) = Q(i, j) + QC(ic, 2);
        double a = CI(ic, 2, LR) * QC(ic, 2) + CI(ic, 2, LL) * QC(ic - 1, 2);
        Q(i - 1, j) = Q(i - 1, j) + a + RES(i - 1, j);
    }
    for (int jc = 3; jc <= jjcf1; ++jc) {
        j += 2;
        i = 2;
        Q(2, j) = Q(2, j) + QC(2, jc);
        double aq = CI(2, jc, LA) * QC(2, jc) + CI(2, jc, LB) * QC(2, jc - 1);
        Q(2, j - 1) = Q(2, j - 1) + aq + RES(2, j - 1);
        for (int ic = 3; ic <= iicf1; ++ic) {
            i += 2;
            Q(i, j) = Q(i, j) + QC(ic, jc);
            double a = CI(ic, jc, LR) * QC(ic, jc) + CI(ic, jc, LL) * QC(ic - 1, jc);
            Q(i - 1, j) = Q(i - 1, j) + a + RES(i - 1, j);
            aq = CI(ic, jc, LA) * QC(ic, jc) + CI(ic, jc, LB) * QC(ic, jc - 1);
            Q(i, j - 1) = Q(i, j - 1) + aq + RES(i, j - 1);
            a = CI(ic, jc, LSW) * QC(ic - 1, jc - 1) + CI(ic, jc, LNW) * QC(ic - 1, jc)
              + CI(ic, jc, LNE) * QC(ic, jc) + CI(ic, jc, LSE) * QC(ic, jc - 1);
            Q(i - 1, j - 1) = Q(i - 1, j - 1) + a + RES(i - 1, j - 1);
        }
    }
#undef CI
#undef Q
#undef QC
#undef RES
}

/* ------------------------------------------------------------------------------------------ */
/* operator-induced interpolation: src/2d/ftn/BMG2_SymStd_SETUP_interp_OI.f90:109-257         */
/* ------------------------------------------------------------------------------------------ */
static double orc_lumped(double c, double sum, double ep, double emach)
{
    /* SUM + (c - SUM)*max(c - (1+ep)*SUM, 0)/(|c - (1+ep)*SUM| + emach) with the reference's
       left-to-right evaluation */
    return (c - sum) * fmax(c - (1.0 + ep) * sum, 0.0) / (fabs(c - (1.0 + ep) * sum) + emach);
}

void orc2_setup_interp(const double *so, int nstncl, double *ci,
                       int iif, int jjf, int iic, int jjc)
{
#define SO(i, j, k) P2(so, iif, jjf, i, j, k)
#define CI(i, j, k) P2(ci, iic, jjc, i, j, k)
    const double emach = ldexp(1.0, -53); /* D1MACH(3): src/2d/ftn/util.f90:65-66 */
    int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2;
    int nine = (nstncl == 5);
    /* x-edge weights LR/LL */
    for (int jc = 2; jc <= jjc - 1; ++jc) {
        int j = 2 * jc - 2;
        for (int ic = 3; ic <= iicf1; ++ic) {
            int i = 2 * ic - 2;
            double a, b;
            if (nine) {
                a = SO(i, j, KW) + SO(i, j, KNW) + SO(i, j + 1, KSW);
                b = SO(i - 1, j, KW) + SO(i - 1, j, KSW) + SO(i - 1, j + 1, KNW);
            } else {
                a = SO(i, j, KW);
                b = SO(i - 1, j, KW);
            }
            double c = SO(i - 1, j, KO);
            double ep = fmin(fabs(a / c), fabs(b / c));
            double sum = a + b + SO(i - 1, j, KS) + SO(i - 1, j + 1, KS);
            sum = a + b + orc_lumped(c, sum, ep, emach);
            sum = 1.0 / sum;
            CI(ic, jc, LR) = a * sum;
            CI(ic, jc, LL) = b * sum;
        }
    }
    /* y-edge weights LA/LB */
    for (int jc = 3; jc <= jjcf1; ++jc) {
        int j = 2 * jc - 2;
        for (int ic = 2; ic <= iic - 1; ++ic) {
            int i = 2 * ic - 2;
            double a, b;
            if (nine) {
                a = SO(i, j, KS) + SO(i, j, KNW) + SO(i + 1, j, KSW);
                b = SO(i, j - 1, KS) + SO(i, j - 1, KSW) + SO(i + 1, j - 1, KNW);
            } else {
                a = SO(i, j, KS);
                b = SO(i, j - 1, KS);
            }
            double c = SO(i, j - 1, KO);
            double ep = fmin(fabs(a / c), fabs(b / c));
            double sum = a + b + SO(i, j - 1, KW) + SO(i + 1, j - 1, KW);
            sum = a + b + orc_lumped(c, sum, ep, emach);
            sum = 1.0 / sum;
            CI(ic, jc, LA) = a * sum;
            CI(ic, jc, LB) = b * sum;
        }
    }
    /* corner weights from the edge weights */
    for (int jc = 3; jc <= jjcf1; ++jc) {
        int j = 2 * jc - 2;
        for (int ic = 3; ic <= iicf1; ++ic) {
            int i = 2 * ic - 2;
            double c = SO(i - 1, j - 1, KO);
            double sum, ep;
            if (nine) {
                sum = SO(i - 1, j - 1, KW) + SO(i - 1, j, KNW) + SO(i - 1, j, KS) + SO(i, j, KSW)
                    + SO(i, j - 1, KW) + SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS)
                    + SO(i - 1, j - 1, KSW);
                double e1 = fabs((SO(i - 1, j - 1, KSW) + SO(i - 1, j - 1, KW) + SO(i - 1, j, KNW)) / c);
                double e2 = fabs((SO(i - 1, j, KNW) + SO(i - 1, j, KS) + SO(i, j, KSW)) / c);
                double e3 = fabs((SO(i, j, KSW) + SO(i, j - 1, KW) + SO(i, j - 1, KNW)) / c);
                double e4 = fabs((SO(i, j - 1, KNW) + SO(i - 1, j - 1, KS) + SO(i - 1, j - 1, KSW)) / c);
                ep = fmin(fmin(fmin(e1, e2), e3), e4);
            } else {
                sum = SO(i - 1, j - 1, KW) + SO(i - 1, j, KS) + SO(i, j - 1, KW) + SO(i - 1, j - 1, KS);
                double e1 = fabs(SO(i - 1, j - 1, KW) / c);
                double e2 = fabs(SO(i - 1, j, KS) / c);
                double e3 = fabs(SO(i, j - 1, KW) / c);
                double e4 = fabs(SO(i - 1, j - 1, KS) / c);
                ep = fmin(fmin(fmin(e1, e2), e3), e4);
            }
            sum = sum + orc_lumped(c, sum, ep, emach);
            double s = 1.0 / sum;
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
    }
#undef SO
#undef CI
}

/* ------------------------------------------------------------------------------------------ */
/* Galerkin coarse operator SOC = P^T A P: src/2d/ftn/BMG2_SymStd_SETUP_ITLI_ex.f90:94-331     */
/* ------------------------------------------------------------------------------------------ */
void orc2_setup_galerkin(const double *so, int nstncl, const double *ci, double *soc,
                         int iif, int jjf, int iic, int jjc)
{
#define SO(i, j, k) P2(so, iif, jjf, i, j, k)
#define CI(i, j, k) P2(ci, iic, jjc, i, j, k)
#define SOC(i, j, k) P2(soc, iic, jjc, i, j, k)
    if (nstncl == 5) {
        for (int jc = 2; jc <= jjc - 1; ++jc) {
            int j = 2 * jc - 2;
            for (int ic = 2; ic <= iic - 1; ++ic) {
                int i = 2 * ic - 2;
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
            }
        }
        for (int jc = 2; jc <= jjc - 1; ++jc) {
            int j = 2 * jc - 2;
            for (int ic = 2; ic <= iic - 1; ++ic) {
                int i = 2 * ic - 2;
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
        }
    } else {
        for (int jc = 2; jc <= jjc - 1; ++jc) {
            int j = 2 * jc - 2;
            for (int ic = 2; ic <= iic - 1; ++ic) {
                int i = 2 * ic - 2;
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
        }
    }
#undef SO
#undef CI
#undef SOC
}

/* ------------------------------------------------------------------------------------------ */
/* coarsest grid: banded Cholesky. src/2d/ftn/BMG2_SymStd_SETUP_cg_LU.f90:92-118,229-255 and   */
/* src/2d/ftn/BMG2_SymStd_SOLVE_cg.f90:93-123 (Dirichlet branch)                                */
/* ------------------------------------------------------------------------------------------ */
int orc2_setup_cg(const double *so, int nstncl, double *abd, int ii, int jj)
{
#define SO(i, j, k) P2(so, ii, jj, i, j, k)
#define ABD(r, c) abd[(size_t)((c) - 1) * (size_t)ii + (size_t)((r) - 1)]
    int i1 = ii - 1, j1 = jj - 1, i2 = i1 - 1;
    int n = i2 * (j1 - 1), kk = 0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) {
            ++kk;
            ABD(ii, kk) = SO(i, j, KO);
            ABD(i1, kk) = -SO(i, j, KW);
            ABD(3, kk) = (nstncl == 5) ? -SO(i + 1, j, KNW) : 0.0;
            ABD(2, kk) = -SO(i, j, KS);
            ABD(1, kk) = (nstncl == 5) ? -SO(i, j, KSW) : 0.0;
        }
    return orc_dpbtrf_u(n, i1, abd, ii);
#undef SO
#undef ABD
}

int orc2_solve_cg(double *q, const double *qf, const double *abd, double *bbd, int ii, int jj)
{
    int i1 = ii - 1, j1 = jj - 1, i2 = i1 - 1;
    int n = i2 * (j1 - 1), kk = 0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) bbd[kk++] = V2(qf, ii, i, j);
    orc_dpbtrs_u(n, i1, abd, ii, bbd);
    kk = 0;
    for (int j = 2; j <= j1; ++j)
        for (int i = 2; i <= i1; ++i) V2(q, ii, i, j) = bbd[kk++];
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* multilevel driver: include/cedar/multilevel.h:95-298, include/cedar/cycle/vcycle.h:35-115,   */
/* include/cedar/2d/solver.h:63-123                                                             */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, ii, jj, nstncl;
    double *so;       /* owned except on level 0 */
    double *ci;       /* P of this level (stored on the coarse level), NULL on level 0 */
    double *x, *b, *res;
    double *sor[2];
} orc2_level;

struct orc2_solver {
    orc2_settings s;
    int jpn;          /* BMG_get_bc(per_mask) */
    int nlevels;
    orc2_level *lv;
    double *abd, *bbd;
    int nabd1, nabd2;
    double *scratch;  /* y-line scratch */
};

int orc2_num_levels(int nx, int ny, int min_coarse)
{
    /* include/cedar/2d/solver.h:63-79 (float coarse sizes, unsigned division) */
    float nxc, nyc;
    int ng = 0;
    do {
        ng++;
        nxc = (float)((unsigned)(nx - 1) / (unsigned)(1 << ng) + 1);
        nyc = (float)((unsigned)(ny - 1) / (unsigned)(1 << ng) + 1);
    } while ((nxc < nyc ? nxc : nyc) >= (float)min_coarse);
    return ng;
}

static double *orc_zeros(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void orc2_smooth(orc2_solver *slv, orc2_level *L, double *x, const double *b, int dir)
{
    int nrep = (dir == ORC_DOWN) ? slv->s.nrelax_pre : slv->s.nrelax_post;
    for (int r = 0; r < nrep; ++r) {
        switch (slv->s.relax) {
        case ORC_RELAX_POINT:
            orc2_relax_gs_bc(L->so, L->nstncl, b, x, L->sor[0], L->ii, L->jj, dir, slv->jpn);
            break;
        case ORC_RELAX_LINE_X:
            orc2_relax_lines_x_bc(L->so, L->nstncl, b, x, L->sor[0], slv->scratch, L->ii, L->jj,
                                  dir, slv->jpn);
            break;
        case ORC_RELAX_LINE_Y:
            orc2_relax_lines_y_bc(L->so, L->nstncl, b, x, L->sor[0], slv->scratch, L->ii, L->jj,
                                  dir, slv->jpn);
            break;
        default: /* line-xy: x then y on DOWN, y then x on UP (include/cedar/multilevel.h:175-207) */
            if (dir == ORC_DOWN) {
                orc2_relax_lines_x_bc(L->so, L->nstncl, b, x, L->sor[0], slv->scratch, L->ii,
                                      L->jj, dir, slv->jpn);
                orc2_relax_lines_y_bc(L->so, L->nstncl, b, x, L->sor[1], slv->scratch, L->ii,
                                      L->jj, dir, slv->jpn);
            } else {
                orc2_relax_lines_y_bc(L->so, L->nstncl, b, x, L->sor[1], slv->scratch, L->ii,
                                      L->jj, dir, slv->jpn);
                orc2_relax_lines_x_bc(L->so, L->nstncl, b, x, L->sor[0], slv->scratch, L->ii,
                                      L->jj, dir, slv->jpn);
            }
        }
    }
}

orc2_solver *orc2_solver_create(const double *so_fine, int nstncl, int nx, int ny,
                                const orc2_settings *s)
{
    const int jpn = orc2_bc_from_mask(s->periodic);
    orc2_solver *slv = (orc2_solver *)calloc(1, sizeof(orc2_solver));
    slv->s = *s;
    slv->jpn = jpn;
    int nl = orc2_num_levels(nx, ny, s->min_coarse);
    if (s->num_levels > 0 && s->num_levels <= nl) nl = s->num_levels;
    slv->nlevels = nl;
    slv->lv = (orc2_level *)calloc((size_t)nl, sizeof(orc2_level));
    int cx = nx, cy = ny, maxjj = 0;
    for (int l = 0; l < nl; ++l) {
        orc2_level *L = &slv->lv[l];
        L->nx = cx; L->ny = cy; L->ii = cx + 2; L->jj = cy + 2;
        size_t np = (size_t)L->ii * L->jj;
        if (L->jj > maxjj) maxjj = L->jj;
        if (l == 0) {
            L->nstncl = nstncl;
            L->so = (double *)so_fine; /* held by reference, as the reference does */
        } else {
            L->nstncl = 5;
            L->so = orc_zeros(np * 5);
            L->ci = orc_zeros(np * 8);
            L->x = orc_zeros(np);
            L->b = orc_zeros(np);
        }
        L->res = orc_zeros(np);
        L->sor[0] = orc_zeros(np * 2);
        L->sor[1] = orc_zeros(np * 2);
        /* coarse size: len_t nxc = (nx-1)/2. + 1 */
        cx = (int)((cx - 1) / 2.0 + 1);
        cy = (int)((cy - 1) / 2.0 + 1);
    }
    slv->scratch = orc_zeros((size_t)(2 * (maxjj > nx + 2 ? maxjj : nx + 2)));
    orc2_level *C = &slv->lv[nl - 1];
    /* include/cedar/2d/solver.h:116-121: a dense n x n array when periodic */
    slv->nabd1 = jpn ? C->nx * C->ny : C->nx + 2;
    slv->nabd2 = C->nx * C->ny;
    slv->abd = orc_zeros((size_t)slv->nabd1 * slv->nabd2);
    slv->bbd = orc_zeros((size_t)slv->nabd2);

    /* setup: interp -> galerkin -> relax per level (include/cedar/multilevel.h:243-265) */
    for (int l = 0; l < nl - 1; ++l) {
        orc2_level *F = &slv->lv[l], *K = &slv->lv[l + 1];
        orc2_setup_interp_bc(F->so, F->nstncl, K->ci, F->ii, F->jj, K->ii, K->jj, jpn);
        orc2_setup_galerkin_bc(F->so, F->nstncl, K->ci, K->so, F->ii, F->jj, K->ii, K->jj, jpn);
        switch (s->relax) {
        case ORC_RELAX_POINT: orc2_setup_recip(F->so, F->sor[0], F->ii, F->jj); break;
        case ORC_RELAX_LINE_X: orc2_setup_lines_x_bc(F->so, F->sor[0], F->ii, F->jj, jpn); break;
        case ORC_RELAX_LINE_Y: orc2_setup_lines_y_bc(F->so, F->sor[0], F->ii, F->jj, jpn); break;
        default:
            orc2_setup_lines_x_bc(F->so, F->sor[0], F->ii, F->jj, jpn);
            orc2_setup_lines_y_bc(F->so, F->sor[1], F->ii, F->jj, jpn);
        }
    }
    orc2_setup_cg_bc(C->so, C->nstncl, slv->abd, C->ii, C->jj, jpn);
    return slv;
}

int orc2_solver_nlevels(const orc2_solver *slv) { return slv->nlevels; }

void orc2_solver_level_dims(const orc2_solver *slv, int lvl, int *nx, int *ny)
{
    *nx = slv->lv[lvl].nx;
    *ny = slv->lv[lvl].ny;
}

void orc2_solver_get_level(const orc2_solver *slv, int lvl, double *so, double *ci,
                           double *sor0, double *sor1)
{
    const orc2_level *L = &slv->lv[lvl];
    size_t np = (size_t)L->ii * L->jj;
    if (so) memcpy(so, L->so, np * (size_t)L->nstncl * sizeof(double));
    if (ci && L->ci) memcpy(ci, L->ci, np * 8 * sizeof(double));
    if (sor0) memcpy(sor0, L->sor[0], np * 2 * sizeof(double));
    if (sor1) memcpy(sor1, L->sor[1], np * 2 * sizeof(double));
}

const double *orc2_solver_abd(const orc2_solver *slv, int *nabd1, int *nabd2)
{
    *nabd1 = slv->nabd1;
    *nabd2 = slv->nabd2;
    return slv->abd;
}

static void orc2_ncycle(orc2_solver *slv, int lvl, double *x, const double *b)
{
    orc2_level *L = &slv->lv[lvl], *K = &slv->lv[lvl + 1];
    orc2_smooth(slv, L, x, b, ORC_DOWN);
    orc2_residual(L->so, L->nstncl, b, x, L->res, L->ii, L->jj);
    orc2_restrict_bc(K->ci, L->res, K->b, L->ii, L->jj, K->ii, K->jj, slv->jpn);
    memset(K->x, 0, (size_t)K->ii * K->jj * sizeof(double));
    if (lvl + 1 == slv->nlevels - 1)
        orc2_solve_cg_bc(K->x, K->b, slv->abd, slv->bbd, K->ii, K->jj, slv->jpn);
    else
        orc2_ncycle(slv, lvl + 1, K->x, K->b);
    orc2_interp_add_bc(x, K->x, L->res, L->so, L->nstncl, K->ci, K->ii, K->jj, L->ii, L->jj,
                       slv->jpn);
    orc2_smooth(slv, L, x, b, ORC_UP);
}

void orc2_solver_vcycle(orc2_solver *slv, double *x, const double *b)
{
    if (slv->nlevels == 1) {
        orc2_level *L = &slv->lv[0];
        orc2_solve_cg_bc(x, b, slv->abd, slv->bbd, L->ii, L->jj, slv->jpn);
    } else {
        orc2_ncycle(slv, 0, x, b);
    }
}

/* fcycle::fmg_cycle (include/cedar/cycle/fcycle.h:49-83): restrict b itself to the next level,
 * recurse, then x = P x_coarse (x and res zeroed first) and one V-cycle from this level.  The
 * finest x is overwritten, so every F-cycle of a solve starts from zero. */
static void orc2_fmg(orc2_solver *slv, int lvl, double *x, const double *b)
{
    if (lvl == slv->nlevels - 1) {
        orc2_level *L = &slv->lv[lvl];
        orc2_solve_cg_bc(x, b, slv->abd, slv->bbd, L->ii, L->jj, slv->jpn);
        return;
    }
    orc2_level *L = &slv->lv[lvl], *K = &slv->lv[lvl + 1];
    /* the reference's restriction refreshes the ghosts of b in place when periodic */
    orc2_restrict_bc(K->ci, (double *)b, K->b, L->ii, L->jj, K->ii, K->jj, slv->jpn);
    orc2_fmg(slv, lvl + 1, K->x, K->b);
    const size_t np = (size_t)L->ii * L->jj;
    memset(x, 0, np * sizeof(double));
    memset(L->res, 0, np * sizeof(double));
    orc2_interp_add_bc(x, K->x, L->res, L->so, L->nstncl, K->ci, K->ii, K->jj, L->ii, L->jj,
                       slv->jpn);
    orc2_ncycle(slv, lvl, x, b);
}

void orc2_solver_cycle(orc2_solver *slv, double *x, const double *b)
{
    if (slv->s.cycle == 1 && slv->nlevels > 1)
        orc2_fmg(slv, 0, x, b);
    else
        orc2_solver_vcycle(slv, x, b);
}

int orc2_solver_solve(orc2_solver *slv, const double *b, double *x, double *hist, int maxhist)
{
    orc2_level *L = &slv->lv[0];
    orc2_residual(L->so, L->nstncl, b, x, L->res, L->ii, L->jj);
    double res0 = orc2_norm(L->res, L->ii, L->jj);
    int it = 0;
    for (int i = 0; i < slv->s.maxiter; ++i) {
        orc2_solver_cycle(slv, x, b);
        orc2_residual(L->so, L->nstncl, b, x, L->res, L->ii, L->jj);
        double rel = orc2_norm(L->res, L->ii, L->jj) / res0;
        if (i < maxhist) hist[i] = rel;
        it = i + 1;
        if (rel < slv->s.tol) break;
    }
    return it;
}

void orc2_solver_destroy(orc2_solver *slv)
{
    if (!slv) return;
    for (int l = 0; l < slv->nlevels; ++l) {
        orc2_level *L = &slv->lv[l];
        if (l > 0) {
            free(L->so);
            free(L->ci);
            free(L->x);
            free(L->b);
        }
        free(L->res);
        free(L->sor[0]);
        free(L->sor[1]);
    }
    free(slv->lv);
    free(slv->abd);
    free(slv->bbd);
    free(slv->scratch);
    free(slv);
}

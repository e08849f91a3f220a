/*
 * FP64 CPU oracle for the 3D BoxMG path -- TEST INFRASTRUCTURE ONLY (see bmg3_oracle.h).
 *
 * Indices in this file are the reference's Fortran indices (1-based, ghosts at 1 and ii), so
 * each loop can be checked against the cited Fortran lines; sums are written in the
 * reference's left-to-right order so that results round identically.
 */
#include "bmg3_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* stencil planes, 1-based (src/3d/ftn/BMG_stencils_f90.h:46-48) */
enum { kp = 1, kpw, kps, kb, kpsw, kpnw, kbw, kbnw, kbn, kbne, kbe, kbse, kbs, kbsw };
/* prolongation planes, 1-based (src/3d/ftn/BMG_stencils_f90.h:59-65) */
enum {
    lxyl = 1, lxyr, lxya, lxyb, lxza, lxzb, lxyne, lxyse, lxysw, lxynw, lxzsw, lxznw, lxzne,
    lxzse, lyzsw, lyznw, lyzne, lyzse, lbsw, lbnw, lbne, lbse, ltsw, ltnw, ltne, ltse
};

static const double EMACH = 0x1.0p-53; /* D1MACH(3), src/2d/ftn/util.f90:2-77 */

static inline size_t ix3(int i, int j, int k, int ii, int jj)
{
    return (size_t)(i - 1) + (size_t)ii * ((size_t)(j - 1) + (size_t)jj * (size_t)(k - 1));
}

/* fine-grid views (ii, jj, kk) and coarse-grid views (iic, jjc, kkc) */
#define Q3(a, i, j, k) (a)[ix3((i), (j), (k), ii, jj)]
#define S3(i, j, k, d) so[(size_t)((d)-1) * P + ix3((i), (j), (k), ii, jj)]
#define QC3(a, i, j, k) (a)[ix3((i), (j), (k), iic, jjc)]
#define CI3(i, j, k, d) ci[(size_t)((d)-1) * Pc + ix3((i), (j), (k), iic, jjc)]

static double *zeros(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static double dmin2(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------------------------
 * The 26 neighbour terms of the 27-point operator in the reference's summation order
 * (src/3d/ftn/BMG3_SymStd_residual.f90:71-97): coefficient stored at (i,j,k)+s in plane d,
 * multiplying q at (i,j,k)+o.  The 7-point operator uses entries 0, 2, 4, 6, 8, 17 in this
 * order (src/3d/ftn/BMG3_SymStd_residual.f90:106-112).
 * ---------------------------------------------------------------------------------------- */
typedef struct { int si, sj, sk, d, oi, oj, ok; } nbr3;
static const nbr3 NB27[26] = {
    {0, 0, 0, kpw, -1, 0, 0},   {0, 1, 0, kpnw, -1, 1, 0}, {0, 1, 0, kps, 0, 1, 0},
    {1, 1, 0, kpsw, 1, 1, 0},   {1, 0, 0, kpw, 1, 0, 0},   {1, 0, 0, kpnw, 1, -1, 0},
    {0, 0, 0, kps, 0, -1, 0},   {0, 0, 0, kpsw, -1, -1, 0}, {0, 0, 0, kb, 0, 0, -1},
    {0, 0, 0, kbw, -1, 0, -1},  {0, 1, 0, kbnw, -1, 1, -1}, {0, 1, 0, kbn, 0, 1, -1},
    {1, 1, 0, kbne, 1, 1, -1},  {1, 0, 0, kbe, 1, 0, -1},  {1, 0, 0, kbse, 1, -1, -1},
    {0, 0, 0, kbs, 0, -1, -1},  {0, 0, 0, kbsw, -1, -1, -1}, {0, 0, 1, kb, 0, 0, 1},
    {0, 0, 1, kbe, -1, 0, 1},   {0, 1, 1, kbse, -1, 1, 1}, {0, 1, 1, kbs, 0, 1, 1},
    {1, 1, 1, kbsw, 1, 1, 1},   {1, 0, 1, kbw, 1, 0, 1},   {1, 0, 1, kbnw, 1, -1, 1},
    {0, 0, 1, kbn, 0, -1, 1},   {0, 0, 1, kbne, -1, -1, 1}};
static const int NB7[6] = {0, 2, 4, 6, 8, 17};

/* qf + sum of the off-diagonal terms at (i,j,k) */
static double offdiag3(const double *so, int nstncl, const double *qf, const double *q, int i,
                       int j, int k, int ii, int jj, size_t P)
{
    double v = Q3(qf, i, j, k);
    if (nstncl == 14) {
        for (int t = 0; t < 26; ++t) {
            const nbr3 *n = &NB27[t];
            v = v + S3(i + n->si, j + n->sj, k + n->sk, n->d) * Q3(q, i + n->oi, j + n->oj, k + n->ok);
        }
    } else {
        for (int u = 0; u < 6; ++u) {
            const nbr3 *n = &NB27[NB7[u]];
            v = v + S3(i + n->si, j + n->sj, k + n->sk, n->d) * Q3(q, i + n->oi, j + n->oj, k + n->ok);
        }
    }
    return v;
}

/* src/3d/ftn/BMG3_SymStd_residual.f90:63-121 */
void orc3_residual(const double *so, int nstncl, const double *qf, const double *q, double *res,
                   int ii, int jj, int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i)
                Q3(res, i, j, k) = offdiag3(so, nstncl, qf, q, i, j, k, ii, jj, P) -
                                   S3(i, j, k, kp) * Q3(q, i, j, k);
}

/* src/3d/ftn/BMG3_SymStd_SETUP_recip.f90:63-69: SOR(:,:,:,msor=2) = 1/SO_p */
void orc3_setup_recip(const double *so, double *sor, int ii, int jj, int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i)
                sor[P + ix3(i, j, k, ii, jj)] = 1.0 / S3(i, j, k, kp);
}

/* src/3d/ftn/BMG3_SymStd_relax_GS.f90:74-181 (definite branch): 27-point with 8 colours
 * (DOWN pts 8..1), 7-point with 2 colours (DOWN pts 1, 0) */
void orc3_relax_gs(const double *so, int nstncl, const double *qf, double *q, const double *sor,
                   int ii, int jj, int kk, int updown)
{
    const size_t P = (size_t)ii * jj * kk;
    const int i1 = ii - 1, j1 = jj - 1, k1 = kk - 1;
    if (nstncl == 14) {
        const int s = updown == ORC_UP ? 1 : 8, e = updown == ORC_UP ? 8 : 1;
        const int st = updown == ORC_UP ? 1 : -1;
        for (int pts = s;; pts += st) {
            for (int k = 2 + ((pts - 1) / 4) % 2; k <= k1; k += 2)
                for (int j = 2 + ((pts - 1) / 2) % 2; j <= j1; j += 2)
                    for (int i = 2 + (pts - 1) % 2; i <= i1; i += 2)
                        Q3(q, i, j, k) = offdiag3(so, 14, qf, q, i, j, k, ii, jj, P) *
                                         sor[P + ix3(i, j, k, ii, jj)];
            if (pts == e) break;
        }
    } else {
        const int s = updown == ORC_UP ? 0 : 1, e = updown == ORC_UP ? 1 : 0;
        const int st = updown == ORC_UP ? 1 : -1;
        for (int pts = s;; pts += st) {
            for (int k = 2; k <= k1; ++k)
                for (int j = 2; j <= j1; ++j) {
                    const int ibeg = (j + k + pts) % 2 + 2;
                    const int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                    for (int i = ibeg; i <= iend; i += 2)
                        Q3(q, i, j, k) = offdiag3(so, 4, qf, q, i, j, k, ii, jj, P) *
                                         sor[P + ix3(i, j, k, ii, jj)];
                }
            if (pts == e) break;
        }
    }
}

/* src/3d/ftn/BMG3_SymStd_restrict.f90:110-141 */
void orc3_restrict(const double *ci, const double *q, double *qc, int iif, int jjf, int kkf,
                   int iic, int jjc, int kkc)
{
    const int ii = iif, jj = jjf;
    const size_t Pc = (size_t)iic * jjc * kkc;
    (void)kkf;
    for (int kc = 2; kc <= kkc - 1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 2; jc <= jjc - 1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 2; ic <= iic - 1; ++ic) {
                const int i = 2 * (ic - 1);
                QC3(qc, ic, jc, kc) =
                    CI3(ic, jc, kc, lxyne) * Q3(q, i - 1, j - 1, k) +
                    CI3(ic, jc, kc, lxya) * Q3(q, i, j - 1, k) +
                    CI3(ic + 1, jc, kc, lxynw) * Q3(q, i + 1, j - 1, k) +
                    CI3(ic, jc, kc, lxyr) * Q3(q, i - 1, j, k) + Q3(q, i, j, k) +
                    CI3(ic + 1, jc, kc, lxyl) * Q3(q, i + 1, j, k) +
                    CI3(ic, jc + 1, kc, lxyse) * Q3(q, i - 1, j + 1, k) +
                    CI3(ic, jc + 1, kc, lxyb) * Q3(q, i, j + 1, k) +
                    CI3(ic + 1, jc + 1, kc, lxysw) * Q3(q, i + 1, j + 1, k) +
                    CI3(ic, jc, kc, ltne) * Q3(q, i - 1, j - 1, k - 1) +
                    CI3(ic, jc, kc, lyznw) * Q3(q, i, j - 1, k - 1) +
                    CI3(ic + 1, jc, kc, ltnw) * Q3(q, i + 1, j - 1, k - 1) +
                    CI3(ic, jc, kc, lxzne) * Q3(q, i - 1, j, k - 1) +
                    CI3(ic, jc, kc, lxza) * Q3(q, i, j, k - 1) +
                    CI3(ic + 1, jc, kc, lxznw) * Q3(q, i + 1, j, k - 1) +
                    CI3(ic, jc + 1, kc, ltse) * Q3(q, i - 1, j + 1, k - 1) +
                    CI3(ic, jc + 1, kc, lyzne) * Q3(q, i, j + 1, k - 1) +
                    CI3(ic + 1, jc + 1, kc, ltsw) * Q3(q, i + 1, j + 1, k - 1) +
                    CI3(ic, jc, kc + 1, lbne) * Q3(q, i - 1, j - 1, k + 1) +
                    CI3(ic, jc, kc + 1, lyzsw) * Q3(q, i, j - 1, k + 1) +
                    CI3(ic + 1, jc, kc + 1, lbnw) * Q3(q, i + 1, j - 1, k + 1) +
                    CI3(ic, jc, kc + 1, lxzse) * Q3(q, i - 1, j, k + 1) +
                    CI3(ic, jc, kc + 1, lxzb) * Q3(q, i, j, k + 1) +
                    CI3(ic + 1, jc, kc + 1, lxzsw) * Q3(q, i + 1, j, k + 1) +
                    CI3(ic, jc + 1, kc + 1, lbse) * Q3(q, i - 1, j + 1, k + 1) +
                    CI3(ic, jc + 1, kc + 1, lyzse) * Q3(q, i, j + 1, k + 1) +
                    CI3(ic + 1, jc + 1, kc + 1, lbsw) * Q3(q, i + 1, j + 1, k + 1);
            }
        }
    }
}

/* src/3d/ftn/BMG3_SymStd_interp_add.f90:72-212 (definite branch) */
void orc3_interp_add(double *q, const double *qc, double *res, const double *so, int nstncl,
                     const double *ci, int iic, int jjc, int kkc, int iif, int jjf, int kkf)
{
    const int ii = iif, jj = jjf;
    const size_t P = (size_t)iif * jjf * kkf, Pc = (size_t)iic * jjc * kkc;
    const int iic1 = iic - 1, kkc1 = kkc - 1;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    (void)nstncl;
    for (int k = 2; k <= kkf - 1; ++k)
        for (int j = 2; j <= jjf - 1; ++j)
            for (int i = 2; i <= iif - 1; ++i) Q3(res, i, j, k) = Q3(res, i, j, k) / S3(i, j, k, kp);

    /* fine planes k = 2(kc-1) that carry coarse points: the 2D pattern */
    for (int kc = 2; kc <= kkc1; ++kc) {
        const int k = 2 * (kc - 1);
        int j = 2, i = 2;
        Q3(q, 2, 2, k) = Q3(q, 2, 2, k) + QC3(qc, 2, 2, kc);
        for (int ic = 3; ic <= iicf1; ++ic) {
            i += 2;
            Q3(q, i, 2, k) = Q3(q, i, 2, k) + QC3(qc, ic, 2, kc);
            const double a = CI3(ic, 2, kc, lxyr) * QC3(qc, ic, 2, kc) +
                             CI3(ic, 2, kc, lxyl) * QC3(qc, ic - 1, 2, kc);
            Q3(q, i - 1, j, k) = Q3(q, i - 1, j, k) + a + Q3(res, i - 1, j, k);
        }
        for (int jc = 3; jc <= jjcf1; ++jc) {
            j += 2;
            i = 2;
            Q3(q, 2, j, k) = Q3(q, 2, j, k) + QC3(qc, 2, jc, kc);
            const double aq = CI3(2, jc, kc, lxya) * QC3(qc, 2, jc, kc) +
                              CI3(2, jc, kc, lxyb) * QC3(qc, 2, jc - 1, kc);
            Q3(q, 2, j - 1, k) = Q3(q, 2, j - 1, k) + aq + Q3(res, 2, j - 1, k);
            for (int ic = 3; ic <= iicf1; ++ic) {
                i += 2;
                Q3(q, i, j, k) = Q3(q, i, j, k) + QC3(qc, ic, jc, kc);
                double a = CI3(ic, jc, kc, lxyr) * QC3(qc, ic, jc, kc) +
                           CI3(ic, jc, kc, lxyl) * QC3(qc, ic - 1, jc, kc);
                Q3(q, i - 1, j, k) = Q3(q, i - 1, j, k) + a + Q3(res, i - 1, j, k);
                const double aq2 = CI3(ic, jc, kc, lxya) * QC3(qc, ic, jc, kc) +
                                   CI3(ic, jc, kc, lxyb) * QC3(qc, ic, jc - 1, kc);
                Q3(q, i, j - 1, k) = Q3(q, i, j - 1, k) + aq2 + Q3(res, i, j - 1, k);
                a = CI3(ic, jc, kc, lxysw) * QC3(qc, ic - 1, jc - 1, kc) +
                    CI3(ic, jc, kc, lxynw) * QC3(qc, ic - 1, jc, kc) +
                    CI3(ic, jc, kc, lxyne) * QC3(qc, ic, jc, kc) +
                    CI3(ic, jc, kc, lxyse) * QC3(qc, ic, jc - 1, kc);
                Q3(q, i - 1, j - 1, k) = Q3(q, i - 1, j - 1, k) + a + Q3(res, i - 1, j - 1, k);
            }
        }
    }
    /* fine-only planes k = 2kc-3, x-coarse columns i = 2(ic-1) */
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * kc - 3;
        for (int ic = 2; ic <= iic1; ++ic) {
            const int i = 2 * (ic - 1);
            Q3(q, i, 2, k) = Q3(q, i, 2, k) + CI3(ic, 2, kc, lxza) * QC3(qc, ic, 2, kc) +
                             CI3(ic, 2, kc, lxzb) * QC3(qc, ic, 2, kc - 1) + Q3(res, i, 2, k);
        }
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 2; ic <= iic1; ++ic) {
                const int i = 2 * (ic - 1);
                Q3(q, i, j, k) = Q3(q, i, j, k) + CI3(ic, jc, kc, lxza) * QC3(qc, ic, jc, kc) +
                                 CI3(ic, jc, kc, lxzb) * QC3(qc, ic, jc, kc - 1) + Q3(res, i, j, k);
                Q3(q, i, j - 1, k) = Q3(q, i, j - 1, k) +
                                     CI3(ic, jc, kc, lyznw) * QC3(qc, ic, jc, kc) +
                                     CI3(ic, jc, kc, lyzne) * QC3(qc, ic, jc - 1, kc) +
                                     CI3(ic, jc, kc, lyzsw) * QC3(qc, ic, jc, kc - 1) +
                                     CI3(ic, jc, kc, lyzse) * QC3(qc, ic, jc - 1, kc - 1) +
                                     Q3(res, i, j - 1, k);
            }
        }
    }
    /* fine-only planes, x-fine columns i = 2ic-3 */
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * kc - 3;
        for (int ic = 3; ic <= iicf1; ++ic) {
            const int i = 2 * ic - 3;
            Q3(q, i, 2, k) = Q3(q, i, 2, k) + CI3(ic, 2, kc, lxznw) * QC3(qc, ic - 1, 2, kc) +
                             CI3(ic, 2, kc, lxzne) * QC3(qc, ic, 2, kc) +
                             CI3(ic, 2, kc, lxzsw) * QC3(qc, ic - 1, 2, kc - 1) +
                             CI3(ic, 2, kc, lxzse) * QC3(qc, ic, 2, kc - 1) + Q3(res, i, 2, k);
        }
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 3; ic <= iicf1; ++ic) {
                const int i = 2 * ic - 3;
                Q3(q, i, j, k) = Q3(q, i, j, k) + CI3(ic, jc, kc, lxznw) * QC3(qc, ic - 1, jc, kc) +
                                 CI3(ic, jc, kc, lxzne) * QC3(qc, ic, jc, kc) +
                                 CI3(ic, jc, kc, lxzsw) * QC3(qc, ic - 1, jc, kc - 1) +
                                 CI3(ic, jc, kc, lxzse) * QC3(qc, ic, jc, kc - 1) + Q3(res, i, j, k);
                Q3(q, i, j - 1, k) = Q3(q, i, j - 1, k) +
                                     CI3(ic, jc, kc, ltnw) * QC3(qc, ic - 1, jc, kc) +
                                     CI3(ic, jc, kc, ltne) * QC3(qc, ic, jc, kc) +
                                     CI3(ic, jc, kc, ltsw) * QC3(qc, ic - 1, jc - 1, kc) +
                                     CI3(ic, jc, kc, ltse) * QC3(qc, ic, jc - 1, kc) +
                                     CI3(ic, jc, kc, lbnw) * QC3(qc, ic - 1, jc, kc - 1) +
                                     CI3(ic, jc, kc, lbne) * QC3(qc, ic, jc, kc - 1) +
                                     CI3(ic, jc, kc, lbsw) * QC3(qc, ic - 1, jc - 1, kc - 1) +
                                     CI3(ic, jc, kc, lbse) * QC3(qc, ic, jc - 1, kc - 1) +
                                     Q3(res, i, j - 1, k);
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * operator-induced interpolation (src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:86-808)
 * ---------------------------------------------------------------------------------------- */

/* c <- a + b + (p - c) max(p - (1+ep) c, 0) / (|p - (1+ep) c| + emach)  (lumped denominator) */
static double lump3(double a, double b, double c, double p, double ep)
{
    return a + b + (p - c) * dmax2(p - (1.0 + ep) * c, 0.0) / (fabs(p - (1.0 + ep) * c) + EMACH);
}
/* dp <- dp + (s - dp) max(s - (1+ep) dp, 0) / (|s - (1+ep) dp| + emach) */
static double lumpd(double dp, double s, double ep)
{
    return dp + (s - dp) * dmax2(s - (1.0 + ep) * dp, 0.0) / (fabs(s - (1.0 + ep) * dp) + EMACH);
}
static double min4(double a, double b, double c, double d)
{
    return dmin2(dmin2(dmin2(a, b), c), d);
}

#define OI_PT_ARGS ORC3_OI_PT_ARGS

/* lxyl, lxyr of coarse (ic,jc,kc): the fine x-edge point (i-1, j, k) */
void orc3_oi_xl(OI_PT_ARGS)
{
    const double p0 = S3(i - 1, j, k, kp);
    double a, b, c, ep;
    if (nine) {
        a = S3(i - 1, j + 1, k, kpnw) + S3(i - 1, j, k, kpw) + S3(i - 1, j, k, kpsw) +
            S3(i - 1, j + 1, k, kbnw) + S3(i - 1, j, k, kbw) + S3(i - 1, j, k, kbsw) +
            S3(i - 1, j + 1, k + 1, kbse) + S3(i - 1, j, k + 1, kbe) +
            S3(i - 1, j, k + 1, kbne);
        b = S3(i, j + 1, k, kpsw) + S3(i, j, k, kpw) + S3(i, j, k, kpnw) +
            S3(i, j + 1, k, kbne) + S3(i, j, k, kbe) + S3(i, j, k, kbse) +
            S3(i, j + 1, k + 1, kbsw) + S3(i, j, k + 1, kbw) + S3(i, j, k + 1, kbnw);
        c = a + b + S3(i - 1, j, k, kps) + S3(i - 1, j + 1, k, kps) +
            S3(i - 1, j + 1, k, kbn) + S3(i - 1, j, k, kb) + S3(i - 1, j, k, kbs) +
            S3(i - 1, j + 1, k + 1, kbs) + S3(i - 1, j, k + 1, kb) +
            S3(i - 1, j, k + 1, kbn);
        ep = dmin2(fabs(a / p0), fabs(b / p0));
    } else {
        a = S3(i - 1, j, k, kpw);
        b = S3(i, j, k, kpw);
        ep = dmin2(fabs(a / p0), fabs(b) / p0);
        c = a + b + S3(i - 1, j, k, kps) + S3(i - 1, j + 1, k, kps) + S3(i - 1, j, k, kb) +
            S3(i - 1, j, k + 1, kb);
    }
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxyl) = a / c;
    CI3(ic, jc, kc, lxyr) = b / c;
}

/* lxya, lxyb of coarse (ic,jc,kc): the fine y-edge point (i, j-1, k) */
void orc3_oi_ya(OI_PT_ARGS)
{
    const double p0 = S3(i, j - 1, k, kp);
    double a, b, c;
    if (nine) {
        a = S3(i, j, k, kpnw) + S3(i, j, k, kps) + S3(i + 1, j, k, kpsw) +
            S3(i, j, k, kbnw) + S3(i, j, k, kbn) + S3(i + 1, j, k, kbne) +
            S3(i, j, k + 1, kbse) + S3(i, j, k + 1, kbs) + S3(i + 1, j, k + 1, kbsw);
        b = S3(i, j - 1, k, kpsw) + S3(i, j - 1, k, kps) + S3(i + 1, j - 1, k, kpnw) +
            S3(i, j - 1, k, kbsw) + S3(i, j - 1, k, kbs) + S3(i + 1, j - 1, k, kbse) +
            S3(i, j - 1, k + 1, kbne) + S3(i, j - 1, k + 1, kbn) +
            S3(i + 1, j - 1, k + 1, kbnw);
        c = a + b + S3(i, j - 1, k, kpw) + S3(i + 1, j - 1, k, kpw) +
            S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kb) + S3(i + 1, j - 1, k, kbe) +
            S3(i, j - 1, k + 1, kbe) + S3(i, j - 1, k + 1, kb) +
            S3(i + 1, j - 1, k + 1, kbw);
    } else {
        a = S3(i, j, k, kps);
        b = S3(i, j - 1, k, kps);
        c = a + b + S3(i, j - 1, k, kpw) + S3(i + 1, j - 1, k, kpw) + S3(i, j - 1, k, kb) +
            S3(i, j - 1, k + 1, kb);
    }
    const double ep = dmin2(fabs(a / p0), fabs(b / p0));
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxya) = a / c;
    CI3(ic, jc, kc, lxyb) = b / c;
}

/* lxza, lxzb of coarse (ic,jc,kc): the fine z-edge point (i, j, k-1) */
void orc3_oi_za(OI_PT_ARGS)
{
    const double p0 = S3(i, j, k - 1, kp);
    double a, b, c;
    if (nine) {
        a = S3(i, j + 1, k, kbse) + S3(i, j + 1, k, kbs) + S3(i + 1, j + 1, k, kbsw) +
            S3(i, j, k, kbe) + S3(i, j, k, kb) + S3(i + 1, j, k, kbw) +
            S3(i, j, k, kbne) + S3(i, j, k, kbn) + S3(i + 1, j, k, kbnw);
        b = S3(i, j + 1, k - 1, kbnw) + S3(i, j + 1, k - 1, kbn) +
            S3(i + 1, j + 1, k - 1, kbne) + S3(i, j, k - 1, kbw) + S3(i, j, k - 1, kb) +
            S3(i + 1, j, k - 1, kbe) + S3(i, j, k - 1, kbsw) + S3(i, j, k - 1, kbs) +
            S3(i + 1, j, k - 1, kbse);
        c = a + b + S3(i, j, k - 1, kpw) + S3(i + 1, j, k - 1, kpw) +
            S3(i, j + 1, k - 1, kpnw) + S3(i, j + 1, k - 1, kps) +
            S3(i + 1, j + 1, k - 1, kpsw) + S3(i, j, k - 1, kpsw) +
            S3(i, j, k - 1, kps) + S3(i + 1, j, k - 1, kpnw);
    } else {
        a = S3(i, j, k, kb);
        b = S3(i, j, k - 1, kb);
        c = a + b + S3(i, j, k - 1, kpw) + S3(i + 1, j, k - 1, kpw) +
            S3(i, j + 1, k - 1, kps) + S3(i, j, k - 1, kps);
    }
    const double ep = dmin2(fabs(a / p0), fabs(b / p0));
    c = lump3(a, b, c, p0, ep);
    CI3(ic, jc, kc, lxza) = a / c;
    CI3(ic, jc, kc, lxzb) = b / c;
}

/* lxy corners of coarse (ic,jc,kc): the fine xy-face point (i-1, j-1, k) */
void orc3_oi_xyf(OI_PT_ARGS)
{
    const double p0 = S3(i - 1, j - 1, k, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i - 1, j, k, kpnw) + S3(i - 1, j, k, kbnw) + S3(i - 1, j, k + 1, kbse);
        dn = S3(i - 1, j, k, kps) + S3(i - 1, j, k, kbn) + S3(i - 1, j, k + 1, kbs);
        dne = S3(i, j, k, kpsw) + S3(i, j, k, kbne) + S3(i, j, k + 1, kbsw);
        dw = S3(i - 1, j - 1, k, kpw) + S3(i - 1, j - 1, k, kbw) + S3(i - 1, j - 1, k + 1, kbe);
        de = S3(i, j - 1, k, kpw) + S3(i, j - 1, k, kbe) + S3(i, j - 1, k + 1, kbw);
        dsw = S3(i - 1, j - 1, k, kpsw) + S3(i - 1, j - 1, k, kbsw) +
              S3(i - 1, j - 1, k + 1, kbne);
        ds = S3(i - 1, j - 1, k, kps) + S3(i - 1, j - 1, k, kbs) + S3(i - 1, j - 1, k + 1, kbn);
        dse = S3(i, j - 1, k, kpnw) + S3(i, j - 1, k, kbse) + S3(i, j - 1, k + 1, kbnw);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i - 1, j, k, kps);
        dw = S3(i - 1, j - 1, k, kpw);
        de = S3(i, j - 1, k, kpw);
        ds = S3(i - 1, j - 1, k, kps);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i - 1, j - 1, k, kb) - S3(i - 1, j - 1, k + 1, kb);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lxynw) = dp * (dnw + CI3(ic - 1, jc, kc, lxya) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxyne) = dp * (dne + CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxya) * de);
        CI3(ic, jc, kc, lxyse) = dp * (dse + CI3(ic, jc, kc, lxyb) * de + CI3(ic, jc - 1, kc, lxyr) * ds);
        CI3(ic, jc, kc, lxysw) = dp * (dsw + CI3(ic, jc - 1, kc, lxyl) * ds + CI3(ic - 1, jc, kc, lxyb) * dw);
    } else {
        CI3(ic, jc, kc, lxynw) = dp * (CI3(ic - 1, jc, kc, lxya) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxyne) = dp * (CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxya) * de);
        CI3(ic, jc, kc, lxyse) = dp * (CI3(ic, jc, kc, lxyb) * de + CI3(ic, jc - 1, kc, lxyr) * ds);
        CI3(ic, jc, kc, lxysw) = dp * (CI3(ic, jc - 1, kc, lxyl) * ds + CI3(ic - 1, jc, kc, lxyb) * dw);
    }
}

/* lxz corners of coarse (ic,jc,kc): the fine xz-face point (i-1, j, k-1) */
void orc3_oi_xzf(OI_PT_ARGS)
{
    const double p0 = S3(i - 1, j, k - 1, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i - 1, j + 1, k, kbse) + S3(i - 1, j, k, kbe) + S3(i - 1, j, k, kbne);
        dn = S3(i - 1, j + 1, k, kbs) + S3(i - 1, j, k, kb) + S3(i - 1, j, k, kbn);
        dne = S3(i, j + 1, k, kbsw) + S3(i, j, k, kbw) + S3(i, j, k, kbnw);
        dw = S3(i - 1, j + 1, k - 1, kpnw) + S3(i - 1, j, k - 1, kpw) + S3(i - 1, j, k - 1, kpsw);
        de = S3(i, j + 1, k - 1, kpsw) + S3(i, j, k - 1, kpw) + S3(i, j, k - 1, kpnw);
        dsw = S3(i - 1, j + 1, k - 1, kbnw) + S3(i - 1, j, k - 1, kbw) +
              S3(i - 1, j, k - 1, kbsw);
        ds = S3(i - 1, j + 1, k - 1, kbn) + S3(i - 1, j, k - 1, kb) + S3(i - 1, j, k - 1, kbs);
        dse = S3(i, j + 1, k - 1, kbne) + S3(i, j, k - 1, kbe) + S3(i, j, k - 1, kbse);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i - 1, j, k, kb);
        dw = S3(i - 1, j, k - 1, kpw);
        de = S3(i, j, k - 1, kpw);
        ds = S3(i - 1, j, k - 1, kb);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i - 1, j + 1, k - 1, kps) - S3(i - 1, j, k - 1, kps);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lxznw) = dp * (dnw + CI3(ic - 1, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxzne) = dp * (dne + CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxza) * de);
        CI3(ic, jc, kc, lxzse) = dp * (dse + CI3(ic, jc, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyr) * ds);
        CI3(ic, jc, kc, lxzsw) = dp * (dsw + CI3(ic, jc, kc - 1, lxyl) * ds + CI3(ic - 1, jc, kc, lxzb) * dw);
    } else {
        CI3(ic, jc, kc, lxznw) = dp * (CI3(ic - 1, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxyl) * dn);
        CI3(ic, jc, kc, lxzne) = dp * (CI3(ic, jc, kc, lxyr) * dn + CI3(ic, jc, kc, lxza) * de);
        CI3(ic, jc, kc, lxzse) = dp * (CI3(ic, jc, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyr) * ds);
        CI3(ic, jc, kc, lxzsw) = dp * (CI3(ic, jc, kc - 1, lxyl) * ds + CI3(ic - 1, jc, kc, lxzb) * dw);
    }
}

/* lyz corners of coarse (ic,jc,kc): the fine yz-face point (i, j-1, k-1) */
void orc3_oi_yzf(OI_PT_ARGS)
{
    const double p0 = S3(i, j - 1, k - 1, kp);
    double dnw, dn, dne, dw, de, dsw, ds, dse, ep;
    if (nine) {
        dnw = S3(i, j, k, kbse) + S3(i, j, k, kbs) + S3(i + 1, j, k, kbsw);
        dn = S3(i, j - 1, k, kbe) + S3(i, j - 1, k, kb) + S3(i + 1, j - 1, k, kbw);
        dne = S3(i, j - 1, k, kbne) + S3(i, j - 1, k, kbn) + S3(i + 1, j - 1, k, kbnw);
        dw = S3(i, j, k - 1, kpnw) + S3(i, j, k - 1, kps) + S3(i + 1, j, k - 1, kpsw);
        de = S3(i, j - 1, k - 1, kpsw) + S3(i, j - 1, k - 1, kps) + S3(i + 1, j - 1, k - 1, kpnw);
        dsw = S3(i, j, k - 1, kbnw) + S3(i, j, k - 1, kbn) + S3(i + 1, j, k - 1, kbne);
        ds = S3(i, j - 1, k - 1, kbw) + S3(i, j - 1, k - 1, kb) + S3(i + 1, j - 1, k - 1, kbe);
        dse = S3(i, j - 1, k - 1, kbsw) + S3(i, j - 1, k - 1, kbs) +
              S3(i + 1, j - 1, k - 1, kbse);
        ep = min4(fabs((dsw + dw + dnw) / p0), fabs((dnw + dn + dne) / p0),
                  fabs((dne + de + dse) / p0), fabs((dse + ds + dsw) / p0));
    } else {
        dn = S3(i, j - 1, k, kb);
        dw = S3(i, j, k - 1, kps);
        de = S3(i, j - 1, k - 1, kps);
        ds = S3(i, j - 1, k - 1, kb);
        dnw = dne = dsw = dse = 0.0;
        ep = min4(fabs(dw / p0), fabs(dn / p0), fabs(de / p0), fabs(ds / p0));
    }
    double dp = nine ? dw + dnw + dn + dne + de + dse + ds + dsw : dw + dn + de + ds;
    const double sum = p0 - S3(i, j - 1, k - 1, kpw) - S3(i + 1, j - 1, k - 1, kpw);
    dp = 1.0 / lumpd(dp, sum, ep);
    if (nine) {
        CI3(ic, jc, kc, lyznw) = dp * (dnw + CI3(ic, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxya) * dn);
        CI3(ic, jc, kc, lyzne) = dp * (dne + CI3(ic, jc, kc, lxyb) * dn + CI3(ic, jc - 1, kc, lxza) * de);
        CI3(ic, jc, kc, lyzse) = dp * (dse + CI3(ic, jc - 1, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyb) * ds);
        CI3(ic, jc, kc, lyzsw) = dp * (dsw + CI3(ic, jc, kc - 1, lxya) * ds + CI3(ic, jc, kc, lxzb) * dw);
    } else {
        CI3(ic, jc, kc, lyznw) = dp * (CI3(ic, jc, kc, lxza) * dw + CI3(ic, jc, kc, lxya) * dn);
        CI3(ic, jc, kc, lyzne) = dp * (CI3(ic, jc, kc, lxyb) * dn + CI3(ic, jc - 1, kc, lxza) * de);
        CI3(ic, jc, kc, lyzse) = dp * (CI3(ic, jc - 1, kc, lxzb) * de + CI3(ic, jc, kc - 1, lxyb) * ds);
        CI3(ic, jc, kc, lyzsw) = dp * (CI3(ic, jc, kc - 1, lxya) * ds + CI3(ic, jc, kc, lxzb) * dw);
    }
}

/* lt*, lb* of coarse (ic,jc,kc): the fine cell-centre point (i-1, j-1, k-1) */
void orc3_oi_cell(OI_PT_ARGS)
{
    const double p0 = S3(i - 1, j - 1, k - 1, kp);
    if (nine) {
        double yp = S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kpnw) +
                    S3(i - 1, j, k - 1, kps) + S3(i, j, k - 1, kpsw) +
                    S3(i, j - 1, k - 1, kpw) + S3(i, j - 1, k - 1, kpnw) +
                    S3(i - 1, j - 1, k - 1, kps) + S3(i - 1, j - 1, k - 1, kpsw) +
                    S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k - 1, kbw) +
                    S3(i - 1, j, k - 1, kbnw) + S3(i - 1, j, k - 1, kbn) +
                    S3(i, j, k - 1, kbne) + S3(i, j - 1, k - 1, kbe) +
                    S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                    S3(i - 1, j - 1, k - 1, kbsw) + S3(i - 1, j - 1, k, kb) +
                    S3(i - 1, j - 1, k, kbe) + S3(i - 1, j, k, kbse) +
                    S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) + S3(i, j - 1, k, kbw) +
                    S3(i, j - 1, k, kbnw) + S3(i - 1, j - 1, k, kbn) +
                    S3(i - 1, j - 1, k, kbne);
        /* as in the reference, the third and fifth sums are not divided by SO_p,
           and kbse enters the fourth sum twice */
        double ye = min4(
            fabs(S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kpnw) +
                 S3(i - 1, j, k, kbse) + S3(i - 1, j - 1, k, kbe) +
                 S3(i - 1, j - 1, k, kbne) + S3(i - 1, j - 1, k - 1, kpsw) +
                 S3(i - 1, j - 1, k - 1, kbsw) + S3(i - 1, j - 1, k - 1, kbw) +
                 S3(i - 1, j, k - 1, kbnw)) / p0,
            fabs(S3(i, j - 1, k - 1, kpw) + S3(i, j, k - 1, kpsw) + S3(i, j, k, kbsw) +
                 S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kbnw) + S3(i, j - 1, k - 1, kpnw) +
                 S3(i, j - 1, k - 1, kbse) + S3(i, j - 1, k - 1, kbe) +
                 S3(i, j, k - 1, kbne)) / p0,
            fabs(S3(i - 1, j, k - 1, kps) + S3(i - 1, j, k - 1, kpnw) +
                 S3(i - 1, j, k, kbse) + S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) +
                 S3(i, j, k - 1, kpsw) + S3(i, j, k - 1, kbne) + S3(i - 1, j, k - 1, kbn) +
                 S3(i - 1, j, k - 1, kbnw)),
            fabs(S3(i - 1, j - 1, k - 1, kps) + S3(i - 1, j - 1, k - 1, kpsw) +
                 S3(i - 1, j - 1, k, kbne) + S3(i - 1, j - 1, k, kbn) +
                 S3(i, j - 1, k, kbnw) + S3(i, j - 1, k - 1, kpnw) +
                 S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                 S3(i, j - 1, k - 1, kbse)) / p0);
        ye = dmin2(dmin2(ye,
                         fabs(S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k - 1, kbw) +
                              S3(i - 1, j, k - 1, kbnw) + S3(i - 1, j, k - 1, kbn) +
                              S3(i, j, k - 1, kbne) + S3(i, j - 1, k - 1, kbe) +
                              S3(i, j - 1, k - 1, kbse) + S3(i - 1, j - 1, k - 1, kbs) +
                              S3(i - 1, j - 1, k - 1, kbsw))),
                   fabs(S3(i - 1, j - 1, k, kb) + S3(i - 1, j - 1, k, kbe) +
                        S3(i - 1, j, k, kbse) + S3(i - 1, j, k, kbs) + S3(i, j, k, kbsw) +
                        S3(i, j - 1, k, kbw) + S3(i, j - 1, k, kbnw) +
                        S3(i - 1, j - 1, k, kbn) + S3(i - 1, j - 1, k, kbne)) / p0);
        yp = lumpd(yp, p0, ye);
        yp = 1.0 / yp;
        CI3(ic, jc, kc, ltnw) = yp * (S3(i - 1, j, k, kbse) +
            CI3(ic - 1, jc, kc, lyznw) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic - 1, jc, kc, lxza) * S3(i - 1, j, k - 1, kpnw) +
            CI3(ic, jc, kc, lxznw) * S3(i - 1, j, k - 1, kps) +
            CI3(ic - 1, jc, kc, lxya) * S3(i - 1, j - 1, k, kbe) +
            CI3(ic, jc, kc, lxyl) * S3(i - 1, j, k, kbs) +
            CI3(ic, jc, kc, lxynw) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, ltne) = yp * (S3(i, j, k, kbsw) +
            CI3(ic, jc, kc, lxzne) * S3(i - 1, j, k - 1, kps) +
            CI3(ic, jc, kc, lxza) * S3(i, j, k - 1, kpsw) +
            CI3(ic, jc, kc, lyznw) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc, kc, lxyr) * S3(i - 1, j, k, kbs) +
            CI3(ic, jc, kc, lxya) * S3(i, j - 1, k, kbw) +
            CI3(ic, jc, kc, lxyne) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, lbnw) = yp * (S3(i - 1, j, k - 1, kbnw) +
            CI3(ic - 1, jc, kc - 1, lxya) * S3(i - 1, j - 1, k - 1, kbw) +
            CI3(ic, jc, kc - 1, lxyl) * S3(i - 1, j, k - 1, kbn) +
            CI3(ic, jc, kc - 1, lxynw) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic - 1, jc, kc, lyzsw) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic - 1, jc, kc, lxzb) * S3(i - 1, j, k - 1, kpnw) +
            CI3(ic, jc, kc, lxzsw) * S3(i - 1, j, k - 1, kps));
        CI3(ic, jc, kc, lbne) = yp * (S3(i, j, k - 1, kbne) +
            CI3(ic, jc, kc - 1, lxyne) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc, kc - 1, lxyr) * S3(i - 1, j, k - 1, kbn) +
            CI3(ic, jc, kc - 1, lxya) * S3(i, j - 1, k - 1, kbe) +
            CI3(ic, jc, kc, lxzse) * S3(i - 1, j, k - 1, kps) +
            CI3(ic, jc, kc, lxzb) * S3(i, j, k - 1, kpsw) +
            CI3(ic, jc, kc, lyzsw) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbsw) = yp * (S3(i - 1, j - 1, k - 1, kbsw) +
            CI3(ic - 1, jc, kc - 1, lxyb) * S3(i - 1, j - 1, k - 1, kbw) +
            CI3(ic, jc, kc - 1, lxysw) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc - 1, kc - 1, lxyl) * S3(i - 1, j - 1, k - 1, kbs) +
            CI3(ic - 1, jc, kc, lyzse) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxzsw) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic - 1, jc - 1, kc, lxzb) * S3(i - 1, j - 1, k - 1, kpsw));
        CI3(ic, jc, kc, ltsw) = yp * (S3(i - 1, j - 1, k, kbne) +
            CI3(ic - 1, jc, kc, lxyb) * S3(i - 1, j - 1, k, kbe) +
            CI3(ic, jc, kc, lxysw) * S3(i - 1, j - 1, k, kb) +
            CI3(ic, jc - 1, kc, lxyl) * S3(i - 1, j - 1, k, kbn) +
            CI3(ic - 1, jc, kc, lyzne) * S3(i - 1, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxznw) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic - 1, jc - 1, kc, lxza) * S3(i - 1, j - 1, k - 1, kpsw));
        CI3(ic, jc, kc, ltse) = yp * (S3(i, j - 1, k, kbnw) +
            CI3(ic, jc - 1, kc, lxyr) * S3(i - 1, j - 1, k, kbn) +
            CI3(ic, jc, kc, lxyse) * S3(i - 1, j - 1, k, kb) +
            CI3(ic, jc, kc, lxyb) * S3(i, j - 1, k, kbw) +
            CI3(ic, jc - 1, kc, lxzne) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic, jc, kc, lyzne) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxza) * S3(i, j - 1, k - 1, kpnw));
        CI3(ic, jc, kc, lbse) = yp * (S3(i, j - 1, k - 1, kbse) +
            CI3(ic, jc - 1, kc - 1, lxyr) * S3(i - 1, j - 1, k - 1, kbs) +
            CI3(ic, jc, kc - 1, lxyse) * S3(i - 1, j - 1, k - 1, kb) +
            CI3(ic, jc, kc - 1, lxyb) * S3(i, j - 1, k - 1, kbe) +
            CI3(ic, jc - 1, kc, lxzse) * S3(i - 1, j - 1, k - 1, kps) +
            CI3(ic, jc, kc, lyzse) * S3(i, j - 1, k - 1, kpw) +
            CI3(ic, jc - 1, kc, lxzb) * S3(i, j - 1, k - 1, kpnw));
    } else {
        double dp = S3(i - 1, j - 1, k - 1, kpw) + S3(i - 1, j, k - 1, kps) +
                    S3(i, j - 1, k - 1, kpw) + S3(i - 1, j - 1, k - 1, kps) +
                    S3(i - 1, j - 1, k - 1, kb) + S3(i - 1, j - 1, k, kb);
        const double ep = dmin2(
            dmin2(min4(fabs(S3(i - 1, j - 1, k - 1, kpw) / p0), fabs(S3(i - 1, j, k - 1, kps) / p0),
                       fabs(S3(i, j - 1, k - 1, kpw) / p0), fabs(S3(i - 1, j - 1, k - 1, kps) / p0)),
                  fabs(S3(i - 1, j - 1, k - 1, kb) / p0)),
            fabs(S3(i - 1, j - 1, k, kb) / p0));
        /* here the reference adds the correction before dp */
        dp = (p0 - dp) * dmax2(p0 - (1.0 + ep) * dp, 0.0) /
                 (fabs(p0 - (1.0 + ep) * dp) + EMACH) + dp;
        dp = 1.0 / dp;
        CI3(ic, jc, kc, ltnw) = dp * (CI3(ic - 1, jc, kc, lyznw) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxznw) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lxynw) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, ltne) = dp * (CI3(ic, jc, kc, lxzne) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lyznw) * S3(i, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxyne) * S3(i - 1, j - 1, k, kb));
        CI3(ic, jc, kc, lbnw) = dp * (CI3(ic, jc, kc - 1, lxynw) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic - 1, jc, kc, lyzsw) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc, kc, lxzsw) * S3(i - 1, j, k - 1, kps));
        CI3(ic, jc, kc, lbne) = dp * (CI3(ic, jc, kc - 1, lxyne) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic, jc, kc, lxzse) * S3(i - 1, j, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzsw) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbsw) = dp * (CI3(ic, jc, kc - 1, lxysw) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic - 1, jc, kc, lyzse) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc - 1, kc, lxzsw) * S3(i - 1, j - 1, k - 1, kps));
        CI3(ic, jc, kc, ltsw) = dp * (CI3(ic, jc, kc, lxysw) * S3(i - 1, j - 1, k, kb) +
                                      CI3(ic - 1, jc, kc, lyzne) * S3(i - 1, j - 1, k - 1, kpw) +
                                      CI3(ic, jc - 1, kc, lxznw) * S3(i - 1, j - 1, k - 1, kps));
        CI3(ic, jc, kc, ltse) = dp * (CI3(ic, jc, kc, lxyse) * S3(i - 1, j - 1, k, kb) +
                                      CI3(ic, jc - 1, kc, lxzne) * S3(i - 1, j - 1, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzne) * S3(i, j - 1, k - 1, kpw));
        CI3(ic, jc, kc, lbse) = dp * (CI3(ic, jc, kc - 1, lxyse) * S3(i - 1, j - 1, k - 1, kb) +
                                      CI3(ic, jc - 1, kc, lxzse) * S3(i - 1, j - 1, k - 1, kps) +
                                      CI3(ic, jc, kc, lyzse) * S3(i, j - 1, k - 1, kpw));
    }
}

void orc3_setup_interp(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf,
                       int iic, int jjc, int kkc)
{
    const int ii = iif, jj = jjf;
    const size_t P = (size_t)iif * jjf * kkf, Pc = (size_t)iic * jjc * kkc;
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
    const int iicf1 = (iif - 2) / 2 + 2, jjcf1 = (jjf - 2) / 2 + 2, kkcf1 = (kkf - 2) / 2 + 2;
    const int nine = nstncl == 14;
    for (int kc = 2; kc <= kkc1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 2; jc <= jjc1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 3; ic <= iicf1; ++ic) {
                orc3_oi_xl(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 2; kc <= kkc1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 2; ic <= iic1; ++ic) {
                orc3_oi_ya(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 2; jc <= jjc1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 2; ic <= iic1; ++ic) {
                orc3_oi_za(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 2; kc <= kkc1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 3; ic <= iicf1; ++ic) {
                orc3_oi_xyf(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 2; jc <= jjc1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 3; ic <= iicf1; ++ic) {
                orc3_oi_xzf(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 2; ic <= iic1; ++ic) {
                orc3_oi_yzf(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
    for (int kc = 3; kc <= kkcf1; ++kc) {
        const int k = 2 * (kc - 1);
        for (int jc = 3; jc <= jjcf1; ++jc) {
            const int j = 2 * (jc - 1);
            for (int ic = 3; ic <= iicf1; ++ic) {
                orc3_oi_cell(so, nine, ci, ii, jj, P, iic, jjc, Pc, 2 * (ic - 1), j, k, ic, jc, kc);
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Galerkin coarse operator, computed as the triple product P^T A P.  P is the prolongation
 * that interp_add applies (src/3d/ftn/BMG3_SymStd_interp_add.f90:96-212): fine point
 * (2ic-1+di, 2jc-1+dj, 2kc-1+dk) (0-based) receives weight w(d) of coarse point (ic,jc,kc),
 * the weight being stored at coarse index (ic,jc,kc) + max(d, 0) under the plane named in
 * PW3 below; w(0) = 1.  A is the fine operator: A(f, f+o) = -so(f+s, d) for the 26 terms of
 * NB27, A(f,f) = so(f, p).  Fine points outside the interior carry no unknown and are skipped.
 * The reference forms the same product with explicit per-direction formulas
 * (src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90, BMG3_SymStd_SETUP_ITLI27_ex.f90); both agree
 * to rounding.
 * ---------------------------------------------------------------------------------------- */
static int pw3_plane(int di, int dj, int dk)
{
    /* CI plane of the weight for offset d (from interp_add's use of each plane) */
    static const int T[3][3][3] = {
        /* dk = -1 */ {{ltne, lyznw, ltnw}, {lxzne, lxza, lxznw}, {ltse, lyzne, ltsw}},
        /* dk =  0 */ {{lxyne, lxya, lxynw}, {lxyr, 0, lxyl}, {lxyse, lxyb, lxysw}},
        /* dk = +1 */ {{lbne, lyzsw, lbnw}, {lxzse, lxzb, lxzsw}, {lbse, lyzse, lbsw}}};
    return T[dk + 1][dj + 1][di + 1];
}

/* weight of fine offset (di,dj,dk) from coarse point (ic,jc,kc), Fortran coarse indices */
static double pweight3(const double *ci, size_t Pc, int iic, int jjc, int ic, int jc, int kc,
                       int di, int dj, int dk)
{
    if (di == 0 && dj == 0 && dk == 0) return 1.0;
    const int d = pw3_plane(di, dj, dk);
    return CI3(ic + (di > 0), jc + (dj > 0), kc + (dk > 0), d);
}

/* fine matrix entry A(f, f+o) at Fortran fine point f */
static double aentry3(const double *so, int nstncl, size_t P, int ii, int jj, int i, int j, int k,
                      int oi, int oj, int ok)
{
    if (oi == 0 && oj == 0 && ok == 0) return S3(i, j, k, kp);
    for (int t = 0; t < 26; ++t) {
        const nbr3 *n = &NB27[t];
        if (n->oi == oi && n->oj == oj && n->ok == ok) {
            if (nstncl == 4 && n->d > kb) return 0.0;
            return -S3(i + n->si, j + n->sj, k + n->sk, n->d);
        }
    }
    return 0.0;
}

/* index n of a periodic direction with interior 2..m+1, mapped into the interior */
static int wrap_in(int n, int m) { return ((n - 2) % m + m) % m + 2; }

/* (P^T A P)(X, Y) for coarse X = (icx,jcx,kcx), Y = X + (DX,DY,DZ) (Fortran indices).  In a
 * periodic direction (per[d]) the fine points of the product are taken in the unwrapped
 * geometry between X and Y (so each stored coupling is the one across its own side, as the
 * reference's hand-expanded products form it) and only the array reads wrap: coarse indices
 * into 2..n+1, fine points into the interior. */
static double kpair3w(const double *so, int nstncl, const double *ci, int iif, int jjf, int kkf,
                      int iic, int jjc, int kkc, int icx, int jcx, int kcx, int DX, int DY, int DZ,
                      const int per[3])
{
    const int ii = iif, jj = jjf;
    const size_t P = (size_t)iif * jjf * kkf, Pc = (size_t)iic * jjc * kkc;
    (void)ii;
    (void)jj;
    const int mf[3] = {iif - 2, jjf - 2, kkf - 2}, mc[3] = {iic - 2, jjc - 2, kkc - 2};
    if (per[0]) icx = wrap_in(icx, mc[0]);
    if (per[1]) jcx = wrap_in(jcx, mc[1]);
    if (per[2]) kcx = wrap_in(kcx, mc[2]);
    int icy = icx + DX, jcy = jcx + DY, kcy = kcx + DZ;
    if (per[0]) icy = wrap_in(icy, mc[0]);
    if (per[1]) jcy = wrap_in(jcy, mc[1]);
    if (per[2]) kcy = wrap_in(kcy, mc[2]);
    const int X0 = 2 * (icx - 1), Y0 = 2 * (jcx - 1), Z0 = 2 * (kcx - 1);
    double sum = 0.0;
    for (int fz = -1; fz <= 1; ++fz)
        for (int fy = -1; fy <= 1; ++fy)
            for (int fx = -1; fx <= 1; ++fx) {
                int x = X0 + fx, y = Y0 + fy, z = Z0 + fz;
                if (x < 2 || x > iif - 1) { if (!per[0]) continue; x = wrap_in(x, mf[0]); }
                if (y < 2 || y > jjf - 1) { if (!per[1]) continue; y = wrap_in(y, mf[1]); }
                if (z < 2 || z > kkf - 1) { if (!per[2]) continue; z = wrap_in(z, mf[2]); }
                double inner = 0.0;
                int any = 0;
                for (int oz = -1; oz <= 1; ++oz)
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int ox = -1; ox <= 1; ++ox) {
                            const int gx = fx + ox - 2 * DX, gy = fy + oy - 2 * DY,
                                      gz = fz + oz - 2 * DZ;
                            if (gx < -1 || gx > 1 || gy < -1 || gy > 1 || gz < -1 || gz > 1) continue;
                            const int xg = x + ox, yg = y + oy, zg = z + oz;
                            if ((!per[0] && (xg < 2 || xg > iif - 1)) ||
                                (!per[1] && (yg < 2 || yg > jjf - 1)) ||
                                (!per[2] && (zg < 2 || zg > kkf - 1)))
                                continue;
                            if (nstncl == 4 && (ox != 0) + (oy != 0) + (oz != 0) > 1) continue;
                            inner += aentry3(so, nstncl, P, iif, jjf, x, y, z, ox, oy, oz) *
                                     pweight3(ci, Pc, iic, jjc, icy, jcy, kcy, gx, gy, gz);
                            any = 1;
                        }
                if (any) sum += pweight3(ci, Pc, iic, jjc, icx, jcx, kcx, fx, fy, fz) * inner;
            }
    return sum;
}

/* (P^T A P)(X, Y) for coarse X = (icx,jcx,kcx), Y = X + (DX,DY,DZ) (Fortran indices) */
static double kpair3(const double *so, int nstncl, const double *ci, int iif, int jjf, int kkf,
                     int iic, int jjc, int kkc, int icx, int jcx, int kcx, int DX, int DY, int DZ)
{
    const int ii = iif, jj = jjf;
    const size_t P = (size_t)iif * jjf * kkf, Pc = (size_t)iic * jjc * kkc;
    (void)ii;
    (void)jj;
    const int X0 = 2 * (icx - 1), Y0 = 2 * (jcx - 1), Z0 = 2 * (kcx - 1);
    double sum = 0.0;
    for (int fz = -1; fz <= 1; ++fz)
        for (int fy = -1; fy <= 1; ++fy)
            for (int fx = -1; fx <= 1; ++fx) {
                const int x = X0 + fx, y = Y0 + fy, z = Z0 + fz;
                if (x < 2 || x > iif - 1 || y < 2 || y > jjf - 1 || z < 2 || z > kkf - 1) continue;
                double inner = 0.0;
                int any = 0;
                for (int oz = -1; oz <= 1; ++oz)
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int ox = -1; ox <= 1; ++ox) {
                            const int gx = fx + ox - 2 * DX, gy = fy + oy - 2 * DY,
                                      gz = fz + oz - 2 * DZ;
                            if (gx < -1 || gx > 1 || gy < -1 || gy > 1 || gz < -1 || gz > 1) continue;
                            const int xg = x + ox, yg = y + oy, zg = z + oz;
                            if (xg < 2 || xg > iif - 1 || yg < 2 || yg > jjf - 1 || zg < 2 ||
                                zg > kkf - 1)
                                continue;
                            if (nstncl == 4 && (ox != 0) + (oy != 0) + (oz != 0) > 1) continue;
                            inner += aentry3(so, nstncl, P, iif, jjf, x, y, z, ox, oy, oz) *
                                     pweight3(ci, Pc, iic, jjc, icx + DX, jcx + DY, kcx + DZ, gx, gy, gz);
                            any = 1;
                        }
                if (any) sum += pweight3(ci, Pc, iic, jjc, icx, jcx, kcx, fx, fy, fz) * inner;
            }
    return sum;
}

/* coarse plane d at point g couples g+a with g+b (offsets from the stored-position table) */
typedef struct { int d; int ax, ay, az, bx, by, bz; } cpair3;
static const cpair3 CPAIRS[13] = {
    {kpw, 0, 0, 0, -1, 0, 0},   {kps, 0, 0, 0, 0, -1, 0},   {kb, 0, 0, 0, 0, 0, -1},
    {kpsw, 0, 0, 0, -1, -1, 0}, {kpnw, 0, -1, 0, -1, 0, 0}, {kbw, 0, 0, 0, -1, 0, -1},
    {kbnw, 0, -1, 0, -1, 0, -1}, {kbn, 0, -1, 0, 0, 0, -1}, {kbne, -1, -1, 0, 0, 0, -1},
    {kbe, -1, 0, 0, 0, 0, -1},  {kbse, -1, 0, 0, 0, -1, -1}, {kbs, 0, 0, 0, 0, -1, -1},
    {kbsw, 0, 0, 0, -1, -1, -1}};

void orc3_setup_galerkin(const double *so, int nstncl, const double *ci, double *soc, int iif,
                         int jjf, int kkf, int iic, int jjc, int kkc)
{
    const size_t Pc = (size_t)iic * jjc * kkc;
    for (int kc = 2; kc <= kkc - 1; ++kc)
        for (int jc = 2; jc <= jjc - 1; ++jc)
            for (int ic = 2; ic <= iic - 1; ++ic) {
                const size_t pc = ix3(ic, jc, kc, iic, jjc);
                soc[pc] = kpair3(so, nstncl, ci, iif, jjf, kkf, iic, jjc, kkc, ic, jc, kc, 0, 0, 0);
                for (int t = 0; t < 13; ++t) {
                    const cpair3 *c = &CPAIRS[t];
                    const int xa = ic + c->ax, ya = jc + c->ay, za = kc + c->az;
                    const int xb = ic + c->bx, yb = jc + c->by, zb = kc + c->bz;
                    double v = 0.0;
                    if (xa >= 2 && ya >= 2 && za >= 2 && xb >= 2 && yb >= 2 && zb >= 2)
                        v = -kpair3(so, nstncl, ci, iif, jjf, kkf, iic, jjc, kkc, xa, ya, za,
                                    xb - xa, yb - ya, zb - za);
                    soc[(size_t)(c->d - 1) * Pc + pc] = v;
                }
            }
}

/* Galerkin coarse operator with periodic directions: P^T A P with the wrapped geometry of
 * kpair3w, then the images of the coarse operator copied as the reference's hand-expanded
 * products end (src/3d/ftn/BMG3_SymStd_SETUP_ITLI07_ex.f90:1002-1113, identical in
 * BMG3_SymStd_SETUP_ITLI27_ex.f90:1884-1995): y images, then x, then z, all 14 planes. */
void orc3_setup_galerkin_bc(const double *so, int nstncl, const double *ci, double *soc, int iif,
                            int jjf, int kkf, int iic, int jjc, int kkc, int jpn)
{
    const int per[3] = {orc3_per_x(jpn), orc3_per_y(jpn), orc3_per_z(jpn)};
    if (!per[0] && !per[1] && !per[2]) {
        orc3_setup_galerkin(so, nstncl, ci, soc, iif, jjf, kkf, iic, jjc, kkc);
        return;
    }
    const size_t Pc = (size_t)iic * jjc * kkc;
    for (int kc = 2; kc <= kkc - 1; ++kc)
        for (int jc = 2; jc <= jjc - 1; ++jc)
            for (int ic = 2; ic <= iic - 1; ++ic) {
                const size_t pc = ix3(ic, jc, kc, iic, jjc);
                soc[pc] = kpair3w(so, nstncl, ci, iif, jjf, kkf, iic, jjc, kkc, ic, jc, kc, 0, 0,
                                  0, per);
                for (int t = 0; t < 13; ++t) {
                    const cpair3 *c = &CPAIRS[t];
                    const int xa = ic + c->ax, ya = jc + c->ay, za = kc + c->az;
                    const int xb = ic + c->bx, yb = jc + c->by, zb = kc + c->bz;
                    double v = 0.0;
                    if ((per[0] || (xa >= 2 && xb >= 2)) && (per[1] || (ya >= 2 && yb >= 2)) &&
                        (per[2] || (za >= 2 && zb >= 2)))
                        v = -kpair3w(so, nstncl, ci, iif, jjf, kkf, iic, jjc, kkc, xa, ya, za,
                                     xb - xa, yb - ya, zb - za, per);
                    soc[(size_t)(c->d - 1) * Pc + pc] = v;
                }
            }
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
#define SOC(i, j, k, d) soc[(size_t)((d)-1) * Pc + ix3((i), (j), (k), iic, jjc)]
    if (per[1])
        for (int kc = 1; kc <= kkc; ++kc)
            for (int ic = 1; ic <= iic; ++ic) {
                for (int d = 1; d <= 14; ++d) SOC(ic, jjc, kc, d) = SOC(ic, 2, kc, d);
                for (int d = 1; d <= 14; ++d) SOC(ic, 1, kc, d) = SOC(ic, jjc1, kc, d);
            }
    if (per[0])
        for (int kc = 1; kc <= kkc; ++kc)
            for (int jc = 1; jc <= jjc; ++jc) {
                for (int d = 1; d <= 14; ++d) SOC(iic, jc, kc, d) = SOC(2, jc, kc, d);
                for (int d = 1; d <= 14; ++d) SOC(1, jc, kc, d) = SOC(iic1, jc, kc, d);
            }
    if (per[2])
        for (int jc = 1; jc <= jjc; ++jc)
            for (int ic = 1; ic <= iic; ++ic) {
                for (int d = 1; d <= 14; ++d) SOC(ic, jc, kkc, d) = SOC(ic, jc, 2, d);
                for (int d = 1; d <= 14; ++d) SOC(ic, jc, 1, d) = SOC(ic, jc, kkc1, d);
            }
#undef SOC
}

/* qf + the off-diagonal terms of the Gauss-Seidel update at (i,j,k) (shared with the periodic
 * branch) */
double orc3_gs_point(const double *so, int nstncl, const double *qf, const double *q, int i, int j,
                     int k, int ii, int jj, int kk)
{
    return offdiag3(so, nstncl, qf, q, i, j, k, ii, jj, (size_t)ii * jj * kk);
}

/* matrix entry A(r, r+o) of term t of the stencil at r = (i,j,k) in the residual's term order
 * (t = -1: the diagonal, o = 0) */
double orc3_stencil_term(const double *so, int nstncl, int ii, int jj, int kk, int i, int j, int k,
                         int t, int *oi, int *oj, int *ok)
{
    const size_t P = (size_t)ii * jj * kk;
    if (t < 0) {
        *oi = *oj = *ok = 0;
        return S3(i, j, k, kp);
    }
    const nbr3 *n = &NB27[nstncl == 14 ? t : NB7[t]];
    *oi = n->oi;
    *oj = n->oj;
    *ok = n->ok;
    return -S3(i + n->si, j + n->sj, k + n->sk, n->d);
}

/* ------------------------------------------------------------------------------------------
 * coarsest grid (src/3d/ftn/BMG3_SymStd_SETUP_cg_LU.f90:100-193,
 * src/3d/ftn/BMG3_SymStd_SOLVE_cg.f90:78-160, non-periodic): band ABD with kd = nx*(ny+1)+1
 * ---------------------------------------------------------------------------------------- */
int orc3_cg_bandwidth(int ii, int jj) { return (ii - 2) * (jj - 1) + 1; }

int orc3_setup_cg(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    const int i1 = ii - 1, j1 = jj - 1, k1 = kk - 1, i2 = i1 - 1;
    const int ibw = i2 * j1 + 1;
    const int n = (ii - 2) * (jj - 2) * (kk - 2);
#define AB(r, c) abd[(size_t)((c)-1) * ldab + (size_t)((r)-1)]
    memset(abd, 0, sizeof(double) * (size_t)ldab * (size_t)n);
    int kl = 0;
    for (int k = 2; k <= k1; ++k)
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i) {
                ++kl;
                /* the reference's assignment order (later rows win where rows coincide) */
                const int nine = nstncl == 14;
                AB(ibw + 1, kl) = S3(i, j, k, kp);
                AB(ibw, kl) = -S3(i, j, k, kpw);
                AB(ibw - i1 + 3, kl) = nine ? -S3(i + 1, j, k, kpnw) : 0.0;
                AB(ibw - i1 + 2, kl) = -S3(i, j, k, kps);
                AB(ibw - i1 + 1, kl) = nine ? -S3(i, j, k, kpsw) : 0.0;
                AB(ibw - (j1 - 2) * i2 + 2, kl) = nine ? -S3(i + 1, j + 1, k, kbne) : 0.0;
                AB(ibw - (j1 - 2) * i2 + 1, kl) = nine ? -S3(i, j + 1, k, kbn) : 0.0;
                AB(ibw - (j1 - 2) * i2, kl) = nine ? -S3(i, j + 1, k, kbnw) : 0.0;
                AB(ibw - (j1 - 1) * i2 + 2, kl) = nine ? -S3(i + 1, j, k, kbe) : 0.0;
                AB(ibw - (j1 - 1) * i2 + 1, kl) = -S3(i, j, k, kb);
                AB(ibw - (j1 - 1) * i2, kl) = nine ? -S3(i, j, k, kbw) : 0.0;
                if (nine) {
                    AB(3, kl) = -S3(i + 1, j, k, kbse);
                    AB(2, kl) = -S3(i, j, k, kbs);
                    AB(1, kl) = -S3(i, j, k, kbsw);
                } else {
                    AB(3, kl) = 0.0;
                    AB(2, kl) = 0.0;
                    AB(1, kl) = 0.0;
                }
            }
#undef AB
    return orc_dpbtrf_u(n, ibw, abd, ldab);
}

void orc3_solve_cg(double *q, const double *qf, const double *abd, int ldab, double *bbd, int ii,
                   int jj, int kk)
{
    const int ibw = orc3_cg_bandwidth(ii, jj);
    const int n = (ii - 2) * (jj - 2) * (kk - 2);
    int kt = 0;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i) bbd[kt++] = Q3(qf, i, j, k);
    orc_dpbtrs_u(n, ibw, abd, ldab, bbd);
    kt = 0;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i) Q3(q, i, j, k) = bbd[kt++];
}

/* lp_norm<2>: sequential sum of squares in Fortran order (include/cedar/3d/grid_func.h) */
double orc3_norm(const double *x, int ii, int jj, int kk)
{
    double s = 0.0;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i) s += pow(Q3(x, i, j, k), 2);
    return pow(s, 0.5);
}

/* y = A q: src/3d/ftn/mpi/BMG3_SymStd_UTILS_matvec.f90:79-124, terms summed left to right in
 * the order written there (27-point: the 26 neighbours plane by plane; 7-point: w, n, e, s, b, t) */
void orc3_matvec(const double *so, int nstncl, const double *q, double *y, int ii, int jj, int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i) {
                double v;
                if (nstncl == 14) {
                    v = S3(i, j, k, kp) * Q3(q, i, j, k) - S3(i, j, k, kpw) * Q3(q, i - 1, j, k)
                      - S3(i, j + 1, k, kpnw) * Q3(q, i - 1, j + 1, k)
                      - S3(i, j + 1, k, kps) * Q3(q, i, j + 1, k)
                      - S3(i + 1, j + 1, k, kpsw) * Q3(q, i + 1, j + 1, k)
                      - S3(i + 1, j, k, kpw) * Q3(q, i + 1, j, k)
                      - S3(i + 1, j, k, kpnw) * Q3(q, i + 1, j - 1, k)
                      - S3(i, j, k, kps) * Q3(q, i, j - 1, k)
                      - S3(i, j, k, kpsw) * Q3(q, i - 1, j - 1, k)
                      - S3(i, j, k, kb) * Q3(q, i, j, k - 1)
                      - S3(i, j, k, kbw) * Q3(q, i - 1, j, k - 1)
                      - S3(i, j + 1, k, kbnw) * Q3(q, i - 1, j + 1, k - 1)
                      - S3(i, j + 1, k, kbn) * Q3(q, i, j + 1, k - 1)
                      - S3(i + 1, j + 1, k, kbne) * Q3(q, i + 1, j + 1, k - 1)
                      - S3(i + 1, j, k, kbe) * Q3(q, i + 1, j, k - 1)
                      - S3(i + 1, j, k, kbse) * Q3(q, i + 1, j - 1, k - 1)
                      - S3(i, j, k, kbs) * Q3(q, i, j - 1, k - 1)
                      - S3(i, j, k, kbsw) * Q3(q, i - 1, j - 1, k - 1)
                      - S3(i, j, k + 1, kb) * Q3(q, i, j, k + 1)
                      - S3(i, j, k + 1, kbe) * Q3(q, i - 1, j, k + 1)
                      - S3(i, j + 1, k + 1, kbse) * Q3(q, i - 1, j + 1, k + 1)
                      - S3(i, j + 1, k + 1, kbs) * Q3(q, i, j + 1, k + 1)
                      - S3(i + 1, j + 1, k + 1, kbsw) * Q3(q, i + 1, j + 1, k + 1)
                      - S3(i + 1, j, k + 1, kbw) * Q3(q, i + 1, j, k + 1)
                      - S3(i + 1, j, k + 1, kbnw) * Q3(q, i + 1, j - 1, k + 1)
                      - S3(i, j, k + 1, kbn) * Q3(q, i, j - 1, k + 1)
                      - S3(i, j, k + 1, kbne) * Q3(q, i - 1, j - 1, k + 1);
                } else {
                    v = S3(i, j, k, kp) * Q3(q, i, j, k) - S3(i, j, k, kpw) * Q3(q, i - 1, j, k)
                      - S3(i, j + 1, k, kps) * Q3(q, i, j + 1, k)
                      - S3(i + 1, j, k, kpw) * Q3(q, i + 1, j, k)
                      - S3(i, j, k, kps) * Q3(q, i, j - 1, k)
                      - S3(i, j, k, kb) * Q3(q, i, j, k - 1)
                      - S3(i, j, k + 1, kb) * Q3(q, i, j, k + 1);
                }
                Q3(y, i, j, k) = v;
            }
}

/* ------------------------------------------------------------------------------------------
 * plane relaxation (include/cedar/3d/relax_planes.h:35-163, src/3d/relax_planes.cc:25-238).
 * Here ipl and the 2D/3D indices are the reference's C++ indices (ghost = 0); plane operators
 * of a 7-point operator are 5-point {c,w,s}, of a 27-point operator 9-point {c,w,s,sw,nw}.
 * ---------------------------------------------------------------------------------------- */
#define X3C(a, i, j, k) (a)[(size_t)(i) + (size_t)ii * ((size_t)(j) + (size_t)jj * (size_t)(k))]
#define S3C(i, j, k, d) so3[(size_t)((d)-1) * P + (size_t)(i) + (size_t)ii * ((size_t)(j) + (size_t)jj * (size_t)(k))]

void orc3_copy_coeff(int dir, const double *so3, int nstncl3, double *so2, int ipl, int ii, int jj,
                     int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    const int nine = nstncl3 == 14;
    if (dir == ORC3_XY) {
        const size_t P2 = (size_t)ii * jj;
        for (int j = 0; j < jj; ++j)
            for (int i = 0; i < ii; ++i) {
                const size_t p = (size_t)i + (size_t)ii * j;
                so2[p] = S3C(i, j, ipl, kp);
                so2[P2 + p] = S3C(i, j, ipl, kpw);
                so2[2 * P2 + p] = S3C(i, j, ipl, kps);
                if (nine) {
                    so2[3 * P2 + p] = S3C(i, j, ipl, kpsw);
                    so2[4 * P2 + p] = S3C(i, j, ipl, kpnw);
                }
            }
    } else if (dir == ORC3_XZ) {
        const size_t P2 = (size_t)ii * kk;
        for (int k = 1; k <= kk - 2; ++k) /* interior k only, as the reference */
            for (int i = 0; i < ii; ++i) {
                const size_t p = (size_t)i + (size_t)ii * k;
                so2[p] = S3C(i, ipl, k, kp);
                so2[P2 + p] = S3C(i, ipl, k, kpw);
                so2[2 * P2 + p] = S3C(i, ipl, k, kb);
                if (nine) {
                    so2[3 * P2 + p] = S3C(i, ipl, k, kbw);
                    so2[4 * P2 + p] = S3C(i, ipl, k, kbe);
                }
            }
    } else {
        const size_t P2 = (size_t)jj * kk;
        for (int k = 0; k < kk; ++k)
            for (int j = 0; j < jj; ++j) {
                const size_t p = (size_t)j + (size_t)jj * k;
                so2[p] = S3C(ipl, j, k, kp);
                so2[P2 + p] = S3C(ipl, j, k, kps);
                so2[2 * P2 + p] = S3C(ipl, j, k, kb);
                if (nine) {
                    so2[3 * P2 + p] = S3C(ipl, j, k, kbs);
                    so2[4 * P2 + p] = S3C(ipl, j, k, kbn);
                }
            }
    }
}

void orc3_copy_rhs(int dir, const double *so3, int nstncl3, const double *x, const double *b,
                   double *b2, int ipl, int ii, int jj, int kk)
{
    const size_t P = (size_t)ii * jj * kk;
    const int nine = nstncl3 == 14;
#define X(i, j, k) X3C(x, i, j, k)
#define S(i, j, k, d) S3C(i, j, k, d)
    if (dir == ORC3_XY) {
        for (int j = 1; j <= jj - 2; ++j)
            for (int i = 1; i <= ii - 2; ++i) {
                const int k = ipl;
                double v;
                if (!nine) {
                    v = X3C(b, i, j, k) + S(i, j, k, kb) * X(i, j, k - 1) + S(i, j, k + 1, kb) * X(i, j, k + 1);
                } else {
                    v = X3C(b, i, j, k) + S(i, j, k, kb) * X(i, j, k - 1) +
                        S(i, j, k, kbw) * X(i - 1, j, k - 1) +
                        S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) +
                        S(i, j + 1, k, kbn) * X(i, j + 1, k - 1) +
                        S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) +
                        S(i + 1, j, k, kbe) * X(i + 1, j, k - 1) +
                        S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) +
                        S(i, j, k, kbs) * X(i, j - 1, k - 1) +
                        S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) +
                        S(i, j, k + 1, kbe) * X(i - 1, j, k + 1) +
                        S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) +
                        S(i, j + 1, k + 1, kbs) * X(i, j + 1, k + 1) +
                        S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) +
                        S(i + 1, j, k + 1, kbw) * X(i + 1, j, k + 1) +
                        S(i, j, k + 1, kb) * X(i, j, k + 1) +
                        S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1) +
                        S(i, j, k + 1, kbn) * X(i, j - 1, k + 1) +
                        S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1);
                }
                b2[(size_t)i + (size_t)ii * j] = v;
            }
    } else if (dir == ORC3_XZ) {
        for (int k = 1; k <= kk - 2; ++k)
            for (int i = 1; i <= ii - 2; ++i) {
                const int j = ipl;
                double v;
                if (!nine) {
                    v = X3C(b, i, j, k) + S(i, j, k, kps) * X(i, j - 1, k) + S(i, j + 1, k, kps) * X(i, j + 1, k);
                } else {
                    v = X3C(b, i, j, k) + S(i, j + 1, k, kpnw) * X(i - 1, j + 1, k) +
                        S(i, j + 1, k, kps) * X(i, j + 1, k) +
                        S(i + 1, j + 1, k, kpsw) * X(i + 1, j + 1, k) +
                        S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) +
                        S(i, j + 1, k, kbn) * X(i, j + 1, k - 1) +
                        S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) +
                        S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) +
                        S(i, j + 1, k + 1, kbs) * X(i, j + 1, k + 1) +
                        S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) +
                        S(i, j, k, kpsw) * X(i - 1, j - 1, k) + S(i, j, k, kps) * X(i, j - 1, k) +
                        S(i + 1, j, k, kpnw) * X(i + 1, j - 1, k) +
                        S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) +
                        S(i, j, k, kbs) * X(i, j - 1, k - 1) +
                        S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) +
                        S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1) +
                        S(i, j, k + 1, kbn) * X(i, j - 1, k + 1) +
                        S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1);
                }
                b2[(size_t)i + (size_t)ii * k] = v;
            }
    } else {
        for (int k = 1; k <= kk - 2; ++k)
            for (int j = 1; j <= jj - 2; ++j) {
                const int i = ipl;
                double v;
                if (!nine) {
                    v = X3C(b, i, j, k) + S(i, j, k, kpw) * X(i - 1, j, k) + S(i + 1, j, k, kpw) * X(i + 1, j, k);
                } else {
                    v = X3C(b, i, j, k) + S(i, j + 1, k, kpnw) * X(i - 1, j + 1, k) +
                        S(i, j, k, kpw) * X(i - 1, j, k) + S(i, j, k, kpsw) * X(i - 1, j - 1, k) +
                        S(i, j + 1, k, kbnw) * X(i - 1, j + 1, k - 1) +
                        S(i, j, k, kbw) * X(i - 1, j, k - 1) +
                        S(i, j, k, kbsw) * X(i - 1, j - 1, k - 1) +
                        S(i, j + 1, k + 1, kbse) * X(i - 1, j + 1, k + 1) +
                        S(i, j, k + 1, kbe) * X(i - 1, j, k + 1) +
                        S(i, j, k + 1, kbne) * X(i - 1, j - 1, k + 1) +
                        S(i + 1, j + 1, k, kpsw) * X(i + 1, j + 1, k) +
                        S(i + 1, j, k, kpw) * X(i + 1, j, k) +
                        S(i + 1, j, k, kpnw) * X(i + 1, j - 1, k) +
                        S(i + 1, j + 1, k, kbne) * X(i + 1, j + 1, k - 1) +
                        S(i + 1, j, k, kbe) * X(i + 1, j, k - 1) +
                        S(i + 1, j, k, kbse) * X(i + 1, j - 1, k - 1) +
                        S(i + 1, j + 1, k + 1, kbsw) * X(i + 1, j + 1, k + 1) +
                        S(i + 1, j, k + 1, kbw) * X(i + 1, j, k + 1) +
                        S(i + 1, j, k + 1, kbnw) * X(i + 1, j - 1, k + 1);
                }
                b2[(size_t)j + (size_t)jj * k] = v;
            }
    }
#undef X
#undef S
}

void orc3_copy32(int dir, const double *x, double *x2, int ipl, int ii, int jj, int kk)
{
    if (dir == ORC3_XY) {
        for (int j = 0; j < jj; ++j)
            for (int i = 0; i < ii; ++i) x2[(size_t)i + (size_t)ii * j] = X3C(x, i, j, ipl);
    } else if (dir == ORC3_XZ) {
        for (int k = 0; k < kk; ++k)
            for (int i = 0; i < ii; ++i) x2[(size_t)i + (size_t)ii * k] = X3C(x, i, ipl, k);
    } else {
        for (int k = 0; k < kk; ++k)
            for (int j = 0; j < jj; ++j) x2[(size_t)j + (size_t)jj * k] = X3C(x, ipl, j, k);
    }
}

void orc3_copy23(int dir, const double *x2, double *x, int ipl, int ii, int jj, int kk)
{
    if (dir == ORC3_XY) {
        for (int j = 0; j < jj; ++j)
            for (int i = 0; i < ii; ++i) X3C(x, i, j, ipl) = x2[(size_t)i + (size_t)ii * j];
    } else if (dir == ORC3_XZ) {
        for (int k = 0; k < kk; ++k)
            for (int i = 0; i < ii; ++i) X3C(x, i, ipl, k) = x2[(size_t)i + (size_t)ii * k];
    } else {
        for (int k = 0; k < kk; ++k)
            for (int j = 0; j < jj; ++j) X3C(x, ipl, j, k) = x2[(size_t)j + (size_t)jj * k];
    }
}

/* ------------------------------------------------------------------------------------------
 * multilevel driver (include/cedar/3d/solver.h:17-128, include/cedar/multilevel.h)
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int n;                /* planes of this direction (interior extent normal to them) */
    int nx2, ny2;         /* interior extents of each plane */
    double **so2;         /* plane operators (owned; the 2D solvers hold them by reference) */
    orc2_solver **slv;    /* one 2D solver per plane */
    double *x2, *b2;      /* plane work arrays ((nx2+2)*(ny2+2)) */
} orc3_planes;

typedef struct {
    int nx, ny, nz, ii, jj, kk, nstncl;
    double *so, *ci, *x, *b, *res, *sor;
    orc3_planes pl[3];
} orc3_level;

struct orc3_solver {
    orc3_settings s;
    int jpn; /* BoxMG boundary id of s.periodic */
    int nlevels;
    orc3_level *lv;
    double *abd, *bbd;
    int ldab;
};

int orc3_num_levels(int nx, int ny, int nz, int min_coarse)
{
    /* include/cedar/3d/solver.h:59-77 (integer division, float coarse sizes) */
    float nxc, nyc, nzc;
    int ng = 0;
    do {
        ng++;
        nxc = (float)((unsigned)(nx - 1) / (unsigned)(1 << ng) + 1);
        nyc = (float)((unsigned)(ny - 1) / (unsigned)(1 << ng) + 1);
        nzc = (float)((unsigned)(nz - 1) / (unsigned)(1 << ng) + 1);
        float m = nxc < nyc ? nxc : nyc;
        if (nzc < m) m = nzc;
        if (!(m >= (float)min_coarse)) break;
    } while (1);
    return ng;
}

static void planes_setup(orc3_level *L, int dir, const orc2_settings *ps)
{
    orc3_planes *pl = &L->pl[dir];
    const int ns2 = L->nstncl == 14 ? 5 : 3;
    if (dir == ORC3_XY) { pl->n = L->nz; pl->nx2 = L->nx; pl->ny2 = L->ny; }
    else if (dir == ORC3_XZ) { pl->n = L->ny; pl->nx2 = L->nx; pl->ny2 = L->nz; }
    else { pl->n = L->nx; pl->nx2 = L->ny; pl->ny2 = L->nz; }
    const size_t np2 = (size_t)(pl->nx2 + 2) * (pl->ny2 + 2);
    pl->so2 = (double **)calloc((size_t)pl->n, sizeof(double *));
    pl->slv = (orc2_solver **)calloc((size_t)pl->n, sizeof(orc2_solver *));
    pl->x2 = zeros(np2);
    pl->b2 = zeros(np2);
    for (int ipl = 1; ipl <= pl->n; ++ipl) {
        pl->so2[ipl - 1] = zeros(np2 * ns2);
        orc3_copy_coeff(dir, L->so, L->nstncl, pl->so2[ipl - 1], ipl, L->ii, L->jj, L->kk);
        pl->slv[ipl - 1] = orc2_solver_create(pl->so2[ipl - 1], ns2, pl->nx2, pl->ny2, ps);
    }
}

static void planes_free(orc3_planes *pl)
{
    for (int t = 0; t < pl->n; ++t) {
        orc2_solver_destroy(pl->slv[t]);
        free(pl->so2[t]);
    }
    free(pl->slv);
    free(pl->so2);
    free(pl->x2);
    free(pl->b2);
    memset(pl, 0, sizeof(*pl));
}

/* relax_planes (include/cedar/3d/relax_planes.h:35-70): per colour, per plane: copy32,
 * copy_rhs, nested 2D solve (residual, max-iter cycles, residual), copy23 */
static void relax_planes(orc3_solver *slv, orc3_level *L, int dir, double *x, const double *b,
                         int updown)
{
    orc3_planes *pl = &L->pl[dir];
    double hist[64];
    const int beg0 = updown == ORC_DOWN ? 1 : 2, beg1 = updown == ORC_DOWN ? 2 : 1;
    for (int c = 0; c < 2; ++c) {
        const int beg = c == 0 ? beg0 : beg1;
        for (int ipl = beg; ipl <= pl->n; ipl += 2) {
            orc3_copy32(dir, x, pl->x2, ipl, L->ii, L->jj, L->kk);
            orc3_copy_rhs(dir, L->so, L->nstncl, x, b, pl->b2, ipl, L->ii, L->jj, L->kk);
            orc2_solver_solve(pl->slv[ipl - 1], pl->b2, pl->x2, hist,
                              slv->s.plane.maxiter < 64 ? slv->s.plane.maxiter : 64);
            orc3_copy23(dir, pl->x2, x, ipl, L->ii, L->jj, L->kk);
        }
    }
}

static void smooth(orc3_solver *slv, orc3_level *L, double *x, const double *b, int updown)
{
    const int nrep = updown == ORC_DOWN ? slv->s.nrelax_pre : slv->s.nrelax_post;
    for (int r = 0; r < nrep; ++r) {
        switch (slv->s.relax) {
        case ORC3_RELAX_POINT:
            orc3_relax_gs_bc(L->so, L->nstncl, b, x, L->sor, L->ii, L->jj, L->kk, updown, slv->jpn);
            break;
        case ORC3_RELAX_PLANE_XY: relax_planes(slv, L, ORC3_XY, x, b, updown); break;
        case ORC3_RELAX_PLANE_XZ: relax_planes(slv, L, ORC3_XZ, x, b, updown); break;
        case ORC3_RELAX_PLANE_YZ: relax_planes(slv, L, ORC3_YZ, x, b, updown); break;
        default: /* plane-xyz: xy, yz, xz going down; xz, yz, xy going up */
            if (updown == ORC_DOWN) {
                relax_planes(slv, L, ORC3_XY, x, b, updown);
                relax_planes(slv, L, ORC3_YZ, x, b, updown);
                relax_planes(slv, L, ORC3_XZ, x, b, updown);
            } else {
                relax_planes(slv, L, ORC3_XZ, x, b, updown);
                relax_planes(slv, L, ORC3_YZ, x, b, updown);
                relax_planes(slv, L, ORC3_XY, x, b, updown);
            }
        }
    }
}

void orc3_solver_relax_planes(orc3_solver *slv, int lvl, int dir, double *x, const double *b,
                              int updown)
{
    relax_planes(slv, &slv->lv[lvl], dir, x, b, updown);
}

static void coarse_solve(orc3_solver *slv, double *x, const double *b)
{
    orc3_level *C = &slv->lv[slv->nlevels - 1];
    orc3_solve_cg_bc(x, b, slv->abd, slv->ldab, slv->bbd, C->ii, C->jj, C->kk, slv->jpn);
}

/* vcycle::ncycle_helper (include/cedar/cycle/vcycle.h:57-115) */
static void ncycle(orc3_solver *slv, int l, double *x, const double *b)
{
    orc3_level *F = &slv->lv[l], *C = &slv->lv[l + 1];
    smooth(slv, F, x, b, ORC_DOWN);
    orc3_residual(F->so, F->nstncl, b, x, F->res, F->ii, F->jj, F->kk);
    orc3_restrict_bc(C->ci, F->res, C->b, F->ii, F->jj, F->kk, C->ii, C->jj, C->kk, slv->jpn);
    memset(C->x, 0, sizeof(double) * (size_t)C->ii * C->jj * C->kk);
    if (l + 1 == slv->nlevels - 1)
        coarse_solve(slv, C->x, C->b);
    else
        ncycle(slv, l + 1, C->x, C->b);
    orc3_interp_add_bc(x, C->x, F->res, F->so, F->nstncl, C->ci, C->ii, C->jj, C->kk, F->ii,
                       F->jj, F->kk, slv->jpn);
    smooth(slv, F, x, b, ORC_UP);
}

void orc3_solver_vcycle(orc3_solver *slv, double *x, const double *b)
{
    if (slv->nlevels == 1)
        coarse_solve(slv, x, b);
    else
        ncycle(slv, 0, x, b);
}

/* fcycle::fmg_cycle (include/cedar/cycle/fcycle.h:49-83) */
static void fmg3(orc3_solver *slv, int l, double *x, const double *b)
{
    if (l == slv->nlevels - 1) {
        coarse_solve(slv, x, b);
        return;
    }
    orc3_level *F = &slv->lv[l], *C = &slv->lv[l + 1];
    orc3_restrict_bc(C->ci, (double *)b, C->b, F->ii, F->jj, F->kk, C->ii, C->jj, C->kk, slv->jpn);
    fmg3(slv, l + 1, C->x, C->b);
    const size_t np = (size_t)F->ii * F->jj * F->kk;
    memset(x, 0, np * sizeof(double));
    memset(F->res, 0, np * sizeof(double));
    orc3_interp_add_bc(x, C->x, F->res, F->so, F->nstncl, C->ci, C->ii, C->jj, C->kk, F->ii,
                       F->jj, F->kk, slv->jpn);
    ncycle(slv, l, x, b);
}

void orc3_solver_cycle(orc3_solver *slv, double *x, const double *b)
{
    if (slv->s.cycle == 1 && slv->nlevels > 1)
        fmg3(slv, 0, x, b);
    else
        orc3_solver_vcycle(slv, x, b);
}

static int plane_relax(int r) { return r >= ORC3_RELAX_PLANE_XY && r <= ORC3_RELAX_PLANE_XYZ; }

orc3_solver *orc3_solver_create(const double *so_fine, int nstncl, int nx, int ny, int nz,
                                const orc3_settings *s)
{
    orc3_solver *slv = (orc3_solver *)calloc(1, sizeof(orc3_solver));
    slv->s = *s;
    slv->jpn = orc3_bc_from_mask(s->periodic);
    int nl = orc3_num_levels(nx, ny, nz, s->min_coarse);
    if (s->num_levels > 0 && s->num_levels <= nl) nl = s->num_levels;
    slv->nlevels = nl;
    slv->lv = (orc3_level *)calloc((size_t)nl, sizeof(orc3_level));
    int cx = nx, cy = ny, cz = nz;
    for (int l = 0; l < nl; ++l) {
        orc3_level *L = &slv->lv[l];
        L->nx = cx; L->ny = cy; L->nz = cz;
        L->ii = cx + 2; L->jj = cy + 2; L->kk = cz + 2;
        const size_t np = (size_t)L->ii * L->jj * L->kk;
        if (l == 0) {
            L->nstncl = nstncl;
            L->so = (double *)so_fine; /* held by reference */
        } else {
            L->nstncl = 14;
            L->so = zeros(np * 14);
            L->ci = zeros(np * 26);
            L->x = zeros(np);
            L->b = zeros(np);
        }
        L->res = zeros(np);
        L->sor = zeros(np * 2);
        cx = (int)((cx - 1) / 2.0 + 1);
        cy = (int)((cy - 1) / 2.0 + 1);
        cz = (int)((cz - 1) / 2.0 + 1);
    }
    orc3_level *C = &slv->lv[nl - 1];
    const int kd = orc3_cg_bandwidth(C->ii, C->jj);
    const size_t n = (size_t)C->nx * C->ny * C->nz;
    /* periodic: a dense n x n factor (include/cedar/3d/solver.h:123-126) */
    slv->ldab = slv->jpn != 0 ? (int)n : kd + 1;
    slv->abd = zeros((size_t)slv->ldab * n);
    slv->bbd = zeros(n);
    for (int l = 0; l < nl - 1; ++l) {
        orc3_level *F = &slv->lv[l], *Cn = &slv->lv[l + 1];
        orc3_setup_interp_bc(F->so, F->nstncl, Cn->ci, F->ii, F->jj, F->kk, Cn->ii, Cn->jj, Cn->kk,
                             slv->jpn);
        orc3_setup_galerkin_bc(F->so, F->nstncl, Cn->ci, Cn->so, F->ii, F->jj, F->kk, Cn->ii,
                               Cn->jj, Cn->kk, slv->jpn);
        if (s->relax == ORC3_RELAX_POINT) {
            orc3_setup_recip(F->so, F->sor, F->ii, F->jj, F->kk);
        } else if (plane_relax(s->relax)) {
            /* multilevel.h:150-154 sets up xy, xz, yz in that order for plane-xyz */
            if (s->relax == ORC3_RELAX_PLANE_XY || s->relax == ORC3_RELAX_PLANE_XYZ)
                planes_setup(F, ORC3_XY, &s->plane);
            if (s->relax == ORC3_RELAX_PLANE_XZ || s->relax == ORC3_RELAX_PLANE_XYZ)
                planes_setup(F, ORC3_XZ, &s->plane);
            if (s->relax == ORC3_RELAX_PLANE_YZ || s->relax == ORC3_RELAX_PLANE_XYZ)
                planes_setup(F, ORC3_YZ, &s->plane);
        }
    }
    orc3_setup_cg_bc(C->so, C->nstncl, slv->abd, slv->ldab, C->ii, C->jj, C->kk, slv->jpn);
    return slv;
}

int orc3_solver_nlevels(const orc3_solver *slv) { return slv->nlevels; }

void orc3_solver_level_dims(const orc3_solver *slv, int lvl, int *nx, int *ny, int *nz)
{
    *nx = slv->lv[lvl].nx;
    *ny = slv->lv[lvl].ny;
    *nz = slv->lv[lvl].nz;
}

void orc3_solver_get_level(const orc3_solver *slv, int lvl, double *so, double *ci)
{
    const orc3_level *L = &slv->lv[lvl];
    const size_t np = (size_t)L->ii * L->jj * L->kk;
    if (so) memcpy(so, L->so, sizeof(double) * np * (size_t)L->nstncl);
    if (ci && L->ci) memcpy(ci, L->ci, sizeof(double) * np * 26);
}

/* multilevel::solve (include/cedar/multilevel.h:278-298) */
int orc3_solver_solve(orc3_solver *slv, const double *b, double *x, double *hist, int maxhist)
{
    orc3_level *F = &slv->lv[0];
    orc3_residual(F->so, F->nstncl, b, x, F->res, F->ii, F->jj, F->kk);
    const double res0 = orc3_norm(F->res, F->ii, F->jj, F->kk);
    int it = 0;
    for (int i = 0; i < slv->s.maxiter; ++i) {
        orc3_solver_cycle(slv, x, b);
        orc3_residual(F->so, F->nstncl, b, x, F->res, F->ii, F->jj, F->kk);
        const double rel = orc3_norm(F->res, F->ii, F->jj, F->kk) / res0;
        if (i < maxhist) hist[i] = rel;
        it = i + 1;
        if (rel < slv->s.tol) break;
    }
    return it;
}

void orc3_solver_destroy(orc3_solver *slv)
{
    if (!slv) return;
    for (int l = 0; l < slv->nlevels; ++l) {
        orc3_level *L = &slv->lv[l];
        for (int d = 0; d < 3; ++d)
            if (L->pl[d].n) planes_free(&L->pl[d]);
        if (l > 0) {
            free(L->so);
            free(L->ci);
            free(L->x);
            free(L->b);
        }
        free(L->res);
        free(L->sor);
    }
    free(slv->lv);
    free(slv->abd);
    free(slv->bbd);
    free(slv);
}

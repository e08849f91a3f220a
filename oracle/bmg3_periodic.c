/*
 * Periodic-boundary branches of the 3D BoxMG kernels -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatements of the JPN != 0 branches of Cedar's serial 3D Fortran kernels, in the
 * reference's operation order.  JPN is the BoxMG boundary id BMG_get_bc returns for the periodic
 * mask (src/2d/ftn/BMG_get_bc.f90:1-24; bits x = 1, y = 2, z = 4): 0 Dirichlet, 2 x, 1 y, 3 xy,
 * 5 z, 6 xz, 7 yz, 8 xyz (include/cedar/2d/ftn/BMG_parameters_c.h:193-204).  With jpn == 0
 * every function here is the Dirichlet kernel of bmg3_oracle.c.
 *
 * Indices are the reference's 1-based Fortran indices (ghosts at 1 and ii).
 */
#include "bmg3_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { kp = 1, kpw, kps, kb, kpsw, kpnw, kbw, kbnw, kbn, kbne, kbe, kbse, kbs, kbsw };
enum {
    lxyl = 1, lxyr, lxya, lxyb, lxza, lxzb, lxyne, lxyse, lxysw, lxynw, lxzsw, lxznw, lxzne,
    lxzse, lyzsw, lyznw, lyzne, lyzse, lbsw, lbnw, lbne, lbse, ltsw, ltnw, ltne, ltse
};

static inline size_t ix3(int i, int j, int k, int ii, int jj)
{
    return (size_t)(i - 1) + (size_t)ii * ((size_t)(j - 1) + (size_t)jj * (size_t)(k - 1));
}
#define Q3(a, i, j, k) (a)[ix3((i), (j), (k), ii, jj)]
#define CI3(i, j, k, d) ci[(size_t)((d)-1) * Pc + ix3((i), (j), (k), iic, jjc)]

int orc3_per_x(int jpn) { const int p = abs(jpn); return p == 2 || p == 3 || p == 6 || p == 8; }
int orc3_per_y(int jpn) { const int p = abs(jpn); return p == 1 || p == 3 || p == 7 || p == 8; }
int orc3_per_z(int jpn) { const int p = abs(jpn); return p == 5 || p == 6 || p == 7 || p == 8; }
/* the definite branch: BMG_BCs_definite (0) or BMG_BCs_indef_nonper (-4) */
static int nonper(int jpn) { return jpn == 0 || jpn == -4; }

int orc3_bc_from_mask(int mask)
{
    /* BMG_get_bc: bcmap(0..7) = definite, def_per_x, def_per_y, def_per_xy, def_per_z,
       def_per_xz, def_per_yz, def_per_xyz */
    static const int map[8] = {0, 2, 1, 3, 5, 6, 7, 8};
    return (mask >= 0 && mask < 8) ? map[mask] : 0;
}

/* periodic images of q on every ghost of the periodic directions, refreshed in the order
 * y, x, z over the full extents (src/3d/ftn/BMG3_SymStd_restrict.f90:79-106) */
void orc3_images(double *q, int ii, int jj, int kk, int jpn)
{
    if (orc3_per_y(jpn))
        for (int k = 1; k <= kk; ++k)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, 1, k) = Q3(q, i, jj - 1, k);
                Q3(q, i, jj, k) = Q3(q, i, 2, k);
            }
    if (orc3_per_x(jpn))
        for (int k = 1; k <= kk; ++k)
            for (int j = 1; j <= jj; ++j) {
                Q3(q, 1, j, k) = Q3(q, ii - 1, j, k);
                Q3(q, ii, j, k) = Q3(q, 2, j, k);
            }
    if (orc3_per_z(jpn))
        for (int j = 1; j <= jj; ++j)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, j, 1) = Q3(q, i, j, kk - 1);
                Q3(q, i, j, kk) = Q3(q, i, j, 2);
            }
}

/* ------------------------------------------------------------------------------------------ */
/* Gauss-Seidel, periodic branch: src/3d/ftn/BMG3_SymStd_relax_GS.f90:189-352.  The x images of */
/* a row are refreshed after each colour pass over it, the y images of a plane after the pass  */
/* over the plane, the z images once after all colours.                                        */
/* ------------------------------------------------------------------------------------------ */
void orc3_relax_gs_bc(const double *so, int nstncl, const double *qf, double *q, const double *sor,
                      int ii, int jj, int kk, int updown, int jpn)
{
    if (nonper(jpn)) {
        orc3_relax_gs(so, nstncl, qf, q, sor, ii, jj, kk, updown);
        return;
    }
    const int px = orc3_per_x(jpn), py = orc3_per_y(jpn), pz = orc3_per_z(jpn);
    const int i1 = ii - 1, j1 = jj - 1, k1 = kk - 1;
    const size_t P = (size_t)ii * jj * kk;
    if (nstncl == 14) {
        const int s = updown == ORC_UP ? 1 : 8, e = updown == ORC_UP ? 8 : 1;
        const int st = updown == ORC_UP ? 1 : -1;
        for (int pts = s;; pts += st) {
            for (int k = 2 + ((pts - 1) / 4) % 2; k <= k1; k += 2) {
                for (int j = 2 + ((pts - 1) / 2) % 2; j <= j1; j += 2) {
                    for (int i = 2 + (pts - 1) % 2; i <= i1; i += 2)
                        Q3(q, i, j, k) = orc3_gs_point(so, 14, qf, q, i, j, k, ii, jj, kk) *
                                         sor[P + ix3(i, j, k, ii, jj)];
                    if (px) {
                        Q3(q, 1, j, k) = Q3(q, i1, j, k);
                        Q3(q, ii, j, k) = Q3(q, 2, j, k);
                    }
                }
                if (py)
                    for (int i = 1; i <= ii; ++i) {
                        Q3(q, i, 1, k) = Q3(q, i, j1, k);
                        Q3(q, i, jj, k) = Q3(q, i, 2, k);
                    }
            }
            if (pts == e) break;
        }
    } else {
        const int s = updown == ORC_UP ? 0 : 1, e = updown == ORC_UP ? 1 : 0;
        const int st = updown == ORC_UP ? 1 : -1;
        for (int pts = s;; pts += st) {
            for (int k = 2; k <= k1; ++k) {
                for (int j = 2; j <= j1; ++j) {
                    const int ibeg = (j + k + pts) % 2 + 2;
                    const int iend = 2 * ((i1 - ibeg) / 2) + ibeg;
                    for (int i = ibeg; i <= iend; i += 2)
                        Q3(q, i, j, k) = orc3_gs_point(so, 4, qf, q, i, j, k, ii, jj, kk) *
                                         sor[P + ix3(i, j, k, ii, jj)];
                    if (px) {
                        Q3(q, 1, j, k) = Q3(q, i1, j, k);
                        Q3(q, ii, j, k) = Q3(q, 2, j, k);
                    }
                }
                if (py)
                    for (int i = 1; i <= ii; ++i) {
                        Q3(q, i, 1, k) = Q3(q, i, j1, k);
                        Q3(q, i, jj, k) = Q3(q, i, 2, k);
                    }
            }
            if (pts == e) break;
        }
    }
    if (pz)
        for (int j = 1; j <= jj; ++j)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, j, 1) = Q3(q, i, j, k1);
                Q3(q, i, j, kk) = Q3(q, i, j, 2);
            }
}

/* restriction: src/3d/ftn/BMG3_SymStd_restrict.f90:79-141 -- the images of Q (the fine
 * residual) are refreshed first, in place, then the definite formula */
void orc3_restrict_bc(const double *ci, double *q, double *qc, int iif, int jjf, int kkf, int iic,
                      int jjc, int kkc, int jpn)
{
    if (!nonper(jpn)) orc3_images(q, iif, jjf, kkf, jpn);
    orc3_restrict(ci, q, qc, iif, jjf, kkf, iic, jjc, kkc);
}

/* interp_add: src/3d/ftn/BMG3_SymStd_interp_add.f90:72-212 (the definite loops, which are the
 * same for every boundary id), then the images of the updated Q.  The reference's own image
 * loops (:237-280) test IPN before assigning it and index with loop variables left over from
 * other loops (column I for the y images, row KKF+1 for the x images); here every image of the
 * periodic directions is refreshed, as restriction does. */
void orc3_interp_add_bc(double *q, const double *qc, double *res, const double *so, int nstncl,
                        const double *ci, int iic, int jjc, int kkc, int iif, int jjf, int kkf,
                        int jpn)
{
    orc3_interp_add(q, qc, res, so, nstncl, ci, iic, jjc, kkc, iif, jjf, kkf);
    if (!nonper(jpn)) orc3_images(q, iif, jjf, kkf, jpn);
}

/* ------------------------------------------------------------------------------------------ */
/* operator-induced interpolation, periodic branch:                                            */
/* src/3d/ftn/BMG3_SymStd_SETUP_interp_OI.f90:809-2784.  The loop bodies are those of the      */
/* definite branch (orc3_oi_*); what differs is where the loops run: the fine index of a       */
/* periodic direction advances by I = MAX(MOD(I+2, N), MIN(I+2, 3)) from 0, and the image rows */
/* of the coarse grid get their own loops and copies.  The 27- and 7-point branches differ in */
/* three statements (marked below); both are followed as written.                             */
/* ------------------------------------------------------------------------------------------ */
static int wnext(int prev, int n)
{
    const int a = (prev + 2) % n, b = prev + 2 < 3 ? prev + 2 : 3;
    return a > b ? a : b;
}

void orc3_setup_interp_bc(const double *so, int nstncl, double *ci, int iif, int jjf, int kkf,
                          int iic, int jjc, int kkc, int jpn)
{
    if (nonper(jpn)) {
        orc3_setup_interp(so, nstncl, ci, iif, jjf, kkf, iic, jjc, kkc);
        return;
    }
    const int ii = iif, jj = jjf, nine = nstncl == 14;
    const size_t P = (size_t)iif * jjf * kkf, Pc = (size_t)iic * jjc * kkc;
#define PT(fn, I, J, K, IC, JC, KC) \
    orc3_oi_##fn(so, nine, ci, ii, jj, P, iic, jjc, Pc, (I), (J), (K), (IC), (JC), (KC))
    const int iic1 = iic - 1, jjc1 = jjc - 1, kkc1 = kkc - 1;
    const int iif1 = iif - 1, jjf1 = jjf - 1, kkf1 = kkf - 1;
    const int iif2 = iif - 2, jjf2 = jjf - 2, kkf2 = kkf - 2;
    const int iicf = (iif - 2) / 2 + 3, jjcf = (jjf - 2) / 2 + 3, kkcf = (kkf - 2) / 2 + 3;
    const int iicf1 = iicf - 1, jjcf1 = jjcf - 1, kkcf1 = kkcf - 1;
    const int iifc = 2 * (iicf - 2) + 2, jjfc = 2 * (jjcf - 2) + 2, kkfc = 2 * (kkcf - 2) + 2;
    const int px = orc3_per_x(jpn), py = orc3_per_y(jpn), pz = orc3_per_z(jpn);
    int ibeg = 2, ibegc = 3, ibeg_x = 2, iend_x = iif2, iendc = iicf1;
    int jbeg = 2, jbegc = 3, jbeg_y = 2, jend_y = jjf2, jendc = jjcf1;
    int kbeg = 2, kbegc = 3, kbeg_z = 2, kend_z = kkf2, kendc = kkcf1;
    if (px) {
        ibeg = 0;
        ibegc = 2;
        if (iic == iicf) { iendc = iicf; ibeg_x = 3; iend_x = iif1; }
    }
    if (py) {
        jbeg = 0;
        jbegc = 2;
        if (jjc == jjcf) { jendc = jjcf; jbeg_y = 3; jend_y = jjf1; }
    }
    if (pz) {
        kbeg = 0;
        kbegc = 2;
        if (kkc == kkcf) { kendc = kkcf; kbeg_z = 3; kend_z = kkf1; }
    }
    int i, j, k, ic, jc, kc;

    /* fine points on coarse k-planes and fine-only x-lines (lxyl, lxyr): :930-1106 */
    for (kc = 2; kc <= kkc1; ++kc) {
        k = 2 * (kc - 1);
        for (jc = 2; jc <= jjc1; ++jc) {
            j = 2 * (jc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xl, i, j, k, ic, jc, kc); }
        }
    }
    if (py) {
        j = jbeg_y;
        for (kc = 2; kc <= kkc1; ++kc) {
            k = 2 * (kc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xl, i, j, k, ic, jjc, kc); }
        }
        j = jend_y;
        for (kc = 2; kc <= kkc1; ++kc) {
            k = 2 * (kc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xl, i, j, k, ic, 1, kc); }
        }
    }
    if (pz) {
        k = kbeg_z;
        for (jc = 2; jc <= jjc1; ++jc) {
            j = 2 * (jc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xl, i, j, k, ic, jc, kkc); }
        }
        if (py) {
            /* jc is left at jjc by the loop above */
            for (ic = ibegc; ic <= iendc; ++ic) {
                if (nine) { /* :1061-1064 */
                    CI3(1, jc, kkc, lxyl) = CI3(iic1, jc, kkc, lxyl);
                    CI3(1, jc, kkc, lxyr) = CI3(iic1, jc, kkc, lxyr);
                    CI3(iic, jc, kkc, lxyl) = CI3(2, jc, kkc, lxyl);
                    CI3(iic, jc, kkc, lxyr) = CI3(2, jc, kkc, lxyr);
                } else { /* :2115-2118 */
                    CI3(ic, 1, kkc, lxyl) = CI3(iic, jjc1, kkc, lxyl);
                    CI3(ic, 1, kkc, lxyr) = CI3(iic, jjc1, kkc, lxyr);
                    CI3(ic, jjc, kkc, lxyl) = CI3(ic, 2, kkc, lxyl);
                    CI3(ic, jjc, kkc, lxyr) = CI3(ic, 2, kkc, lxyr);
                }
            }
        }
        k = kend_z;
        for (jc = 2; jc <= jjc1; ++jc) {
            j = 2 * (jc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xl, i, j, k, ic, jc, 1); }
        }
        if (py)
            for (ic = ibegc; ic <= iendc; ++ic) {
                CI3(ic, 1, 1, lxyl) = CI3(ic, jjc1, 1, lxyl);
                CI3(ic, 1, 1, lxyr) = CI3(ic, jjc1, 1, lxyr);
                CI3(ic, jjc, 1, lxyl) = CI3(ic, 2, 1, lxyl);
                CI3(ic, jjc, 1, lxyr) = CI3(ic, 2, 1, lxyr);
            }
    }

    /* fine points on coarse k-planes and fine-only y-lines (lxya, lxyb): :1109-1277 */
    for (kc = 2; kc <= kkc1; ++kc) {
        k = 2 * (kc - 1);
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, jjfc);
            for (ic = 2; ic <= iic1; ++ic) PT(ya, 2 * (ic - 1), j, k, ic, jc, kc);
        }
    }
    if (px) {
        i = ibeg_x;
        for (kc = 2; kc <= kkc1; ++kc) {
            k = 2 * (kc - 1);
            j = jbeg;
            for (jc = jbegc; jc <= jendc; ++jc) { j = wnext(j, jjfc); PT(ya, i, j, k, iic, jc, kc); }
        }
        i = iend_x;
        for (kc = 2; kc <= kkc1; ++kc) {
            k = 2 * (kc - 1);
            j = jbeg;
            for (jc = jbegc; jc <= jendc; ++jc) { j = wnext(j, jjfc); PT(ya, i, j, k, 1, jc, kc); }
        }
    }
    if (pz) {
        k = kbeg_z;
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, jjfc);
            for (ic = 2; ic <= iic1; ++ic) PT(ya, 2 * (ic - 1), j, k, ic, jc, kkc);
            if (px) {
                CI3(1, jc, kkc, lxya) = CI3(iic1, jc, kkc, lxya);
                CI3(1, jc, kkc, lxyb) = CI3(iic1, jc, kkc, lxyb);
                CI3(iic, jc, kkc, lxya) = CI3(2, jc, kkc, lxya);
                CI3(iic, jc, kkc, lxyb) = CI3(2, jc, kkc, lxyb);
            }
        }
        k = kend_z;
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, jjfc);
            for (ic = 2; ic <= iic1; ++ic) PT(ya, 2 * (ic - 1), j, k, ic, jc, 1);
            if (px) {
                CI3(1, jc, 1, lxya) = CI3(iic1, jc, 1, lxya);
                CI3(1, jc, 1, lxyb) = CI3(iic1, jc, 1, lxyb);
                CI3(iic, jc, 1, lxya) = CI3(2, jc, 1, lxya);
                /* :1274 (27-point) copies from column 1, :2269 (7-point) from column 2 */
                CI3(iic, jc, 1, lxyb) = CI3(nine ? 1 : 2, jc, 1, lxyb);
            }
        }
    }

    /* fine points on coarse j-planes and fine-only z-lines (lxza, lxzb): :1279-1442 */
    k = kbeg;
    for (kc = kbegc; kc <= kendc; ++kc) {
        k = wnext(k, kkfc);
        for (jc = 2; jc <= jjc1; ++jc)
            for (ic = 2; ic <= iic1; ++ic) PT(za, 2 * (ic - 1), 2 * (jc - 1), k, ic, jc, kc);
    }
    if (px) {
        for (int side = 0; side < 2; ++side) {
            const int icx = side == 0 ? iic : 1;
            i = side == 0 ? ibeg_x : iend_x;
            k = kbeg;
            for (kc = kbegc; kc <= kendc; ++kc) {
                k = wnext(k, kkfc);
                for (jc = 2; jc <= jjc1; ++jc) PT(za, i, 2 * (jc - 1), k, icx, jc, kc);
                if (py) {
                    CI3(icx, 1, kc, lxza) = CI3(icx, jjc1, kc, lxza);
                    CI3(icx, 1, kc, lxzb) = CI3(icx, jjc1, kc, lxzb);
                    CI3(icx, jjc, kc, lxza) = CI3(icx, 2, kc, lxza);
                    CI3(icx, jjc, kc, lxzb) = CI3(icx, 2, kc, lxzb);
                }
            }
        }
    }
    if (py) {
        /* the 7-point branch runs this loop to jjc1 (:2359), the 27-point one to iic1 (:1391) */
        const int iend_q = nine ? iic1 : jjc1;
        j = jbeg_y;
        k = kbeg;
        for (kc = kbegc; kc <= kendc; ++kc) {
            k = wnext(k, kkfc);
            for (ic = 2; ic <= iend_q; ++ic) PT(za, 2 * (ic - 1), j, k, ic, jjc, kc);
        }
        j = jend_y;
        k = kbeg;
        for (kc = kbegc; kc <= kendc; ++kc) {
            k = wnext(k, kkfc);
            for (ic = 2; ic <= iic1; ++ic) PT(za, 2 * (ic - 1), j, k, ic, 1, kc);
        }
    }

    /* fine points on coarse k-planes, fine-only x- and y-lines (lxy corners): :1444-1583 */
    i = ibeg;
    for (ic = ibegc; ic <= iendc; ++ic) {
        i = wnext(i, iifc);
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, jjfc);
            for (kc = 2; kc <= kkc1; ++kc) PT(xyf, i, j, 2 * (kc - 1), ic, jc, kc);
        }
    }
    if (pz) {
        for (int side = 0; side < 2; ++side) {
            const int kcx = side == 0 ? kkc : 1;
            k = side == 0 ? kbeg_z : kend_z;
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) {
                i = wnext(i, iifc);
                j = jbeg;
                for (jc = jbegc; jc <= jendc; ++jc) { j = wnext(j, jjfc); PT(xyf, i, j, k, ic, jc, kcx); }
            }
        }
    }

    /* fine points on coarse j-planes, fine-only x- and z-lines (lxz corners): :1587-1720 */
    k = kbeg;
    for (kc = kbegc; kc <= kendc; ++kc) {
        k = wnext(k, kkfc);
        for (jc = 2; jc <= jjc1; ++jc) {
            j = 2 * (jc - 1);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xzf, i, j, k, ic, jc, kc); }
        }
    }
    if (py) {
        for (int side = 0; side < 2; ++side) {
            const int jcx = side == 0 ? jjc : 1;
            k = kbeg;
            for (kc = kbegc; kc <= kendc; ++kc) {
                k = wnext(k, kkfc);
                j = side == 0 ? jbeg_y : jend_y;
                i = ibeg;
                for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(xzf, i, j, k, ic, jcx, kc); }
            }
        }
    }

    /* fine points on coarse i-planes, fine-only y- and z-lines (lyz corners): :1723-1859; the
     * reference advances J with the z modulus KKFC here */
    k = kbeg;
    for (kc = kbegc; kc <= kendc; ++kc) {
        k = wnext(k, kkfc);
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, kkfc);
            for (ic = 2; ic <= iic1; ++ic) PT(yzf, 2 * (ic - 1), j, k, ic, jc, kc);
        }
    }
    if (px) {
        for (int side = 0; side < 2; ++side) {
            const int icx = side == 0 ? iic : 1;
            k = kbeg;
            for (kc = kbegc; kc <= kendc; ++kc) {
                k = wnext(k, kkfc);
                j = jbeg;
                for (jc = jbegc; jc <= jendc; ++jc) {
                    j = wnext(j, kkfc);
                    PT(yzf, side == 0 ? ibeg_x : iend_x, j, k, icx, jc, kc);
                }
            }
        }
    }

    /* fine points on fine-only x-, y- and z-lines (cell centres): :1862-2013 */
    k = kbeg;
    for (kc = kbegc; kc <= kendc; ++kc) {
        k = wnext(k, kkfc);
        j = jbeg;
        for (jc = jbegc; jc <= jendc; ++jc) {
            j = wnext(j, jjfc);
            i = ibeg;
            for (ic = ibegc; ic <= iendc; ++ic) { i = wnext(i, iifc); PT(cell, i, j, k, ic, jc, kc); }
        }
    }
#undef PT
}

/* ------------------------------------------------------------------------------------------ */
/* coarsest grid, periodic: the dense matrix of the operator with its periodic couplings (the   */
/* coefficient of q(c) in the equation of row r, every term of the 27-point stencil whose       */
/* neighbour wraps onto c added in the residual's term order), factored by a Cholesky in        */
/* DPOTF2's order (src/3d/ftn/BMG3_SymStd_SETUP_cg_LU.f90:190-612 writes the same matrix term   */
/* by term and calls DPOTRF); the solve (src/3d/ftn/BMG3_SymStd_SOLVE_cg.f90:103-196) removes   */
/* the mean of the coarse solution and refreshes its images in the order x, y, z.              */
/* ------------------------------------------------------------------------------------------ */
void orc3_assemble_cg_periodic(const double *so, int nstncl, double *abd, int ii, int jj, int kk,
                               int jpn)
{
    const int nx = ii - 2, ny = jj - 2, nz = kk - 2;
    const size_t n = (size_t)nx * ny * nz;
    const int px = orc3_per_x(jpn), py = orc3_per_y(jpn), pz = orc3_per_z(jpn);
    memset(abd, 0, sizeof(double) * n * n);
    for (int k = 2; k <= kk - 1; ++k)
        for (int j = 2; j <= jj - 1; ++j)
            for (int i = 2; i <= ii - 1; ++i) {
                const size_t r = (size_t)(i - 2) + (size_t)nx * ((size_t)(j - 2) + (size_t)ny * (k - 2));
                int oi, oj, ok;
                abd[r * n + r] += orc3_stencil_term(so, nstncl, ii, jj, kk, i, j, k, -1, &oi, &oj, &ok);
                const int nt = nstncl == 14 ? 26 : 6;
                for (int t = 0; t < nt; ++t) {
                    const double a = orc3_stencil_term(so, nstncl, ii, jj, kk, i, j, k, t, &oi,
                                                       &oj, &ok);
                    int ci_ = i + oi, cj = j + oj, ck = k + ok;
                    if (ci_ < 2 || ci_ > ii - 1) {
                        if (!px) continue;
                        ci_ = ci_ < 2 ? ci_ + nx : ci_ - nx;
                    }
                    if (cj < 2 || cj > jj - 1) {
                        if (!py) continue;
                        cj = cj < 2 ? cj + ny : cj - ny;
                    }
                    if (ck < 2 || ck > kk - 1) {
                        if (!pz) continue;
                        ck = ck < 2 ? ck + nz : ck - nz;
                    }
                    const size_t c = (size_t)(ci_ - 2) + (size_t)nx * ((size_t)(cj - 2) + (size_t)ny * (ck - 2));
                    if (r <= c) abd[c * n + r] += a;
                }
            }
}

int orc3_setup_cg_bc(const double *so, int nstncl, double *abd, int ldab, int ii, int jj, int kk,
                     int jpn)
{
    if (nonper(jpn)) return orc3_setup_cg(so, nstncl, abd, ldab, ii, jj, kk);
    const int n = (ii - 2) * (jj - 2) * (kk - 2);
    orc3_assemble_cg_periodic(so, nstncl, abd, ii, jj, kk, jpn);
    return orc_dpotrf_u(n, abd, n);
}

void orc3_solve_cg_bc(double *q, const double *qf, const double *abd, int ldab, double *bbd,
                      int ii, int jj, int kk, int jpn)
{
    if (nonper(jpn)) {
        orc3_solve_cg(q, qf, abd, ldab, bbd, ii, jj, kk);
        return;
    }
    const int i1 = ii - 1, j1 = jj - 1, k1 = kk - 1, n = (ii - 2) * (jj - 2) * (kk - 2);
    int kt = 0;
    for (int k = 2; k <= k1; ++k)
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i) bbd[kt++] = Q3(qf, i, j, k);
    orc_dpotrs_u(n, abd, n, bbd);
    kt = 0;
    for (int k = 2; k <= k1; ++k)
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i) Q3(q, i, j, k) = bbd[kt++];
    double cint = 0.0, qint = 0.0;
    for (int k = 2; k <= k1; ++k)
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i) {
                qint = qint + Q3(q, i, j, k);
                cint = cint + 1.0;
            }
    const double c = -qint / cint;
    for (int k = 2; k <= k1; ++k)
        for (int j = 2; j <= j1; ++j)
            for (int i = 2; i <= i1; ++i) Q3(q, i, j, k) = Q3(q, i, j, k) + c;
    if (orc3_per_x(jpn))
        for (int k = 1; k <= kk; ++k)
            for (int j = 1; j <= jj; ++j) {
                Q3(q, 1, j, k) = Q3(q, i1, j, k);
                Q3(q, ii, j, k) = Q3(q, 2, j, k);
            }
    if (orc3_per_y(jpn))
        for (int k = 1; k <= kk; ++k)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, 1, k) = Q3(q, i, j1, k);
                Q3(q, i, jj, k) = Q3(q, i, 2, k);
            }
    if (orc3_per_z(jpn))
        for (int j = 1; j <= jj; ++j)
            for (int i = 1; i <= ii; ++i) {
                Q3(q, i, j, 1) = Q3(q, i, j, k1);
                Q3(q, i, j, kk) = Q3(q, i, j, 2);
            }
}

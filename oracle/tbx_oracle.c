/*
 * tbx_oracle.c -- THE ORACLE.  TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C (C99 + OpenMP) restatement of tblite's GFN2-xTB one-electron
 * integral path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or, in
 * bench.py, as the timed CPU baseline).  The HIP product never calls it.
 *
 * Every function follows the reference routine it names (file:line relative
 * to the tblite source tree, version 0.7.0):
 *
 *   orc_overlap_1d            integral/native/integrals.f90:91-103
 *   orc_shift_poly            integral/native/integrals.f90:105-130 (horizontal_shift)
 *   orc_poly_product          integral/native/integrals.f90:132-222 (form_product)
 *   orc_multipole_3d          integral/native/integrals.f90:759-804
 *   orc_multipole_grad_3d     integral/native/integrals.f90:806-900
 *   orc_shift_operator_grad   integral/native/integrals.f90:902-936
 *   orc_transform0            integral/trafo.f90:104-257 (Cartesian -> spherical, l <= 2)
 *   orc_overlap_cgto          integral/native/integrals.f90:315-360
 *   orc_multipole_cgto        integral/native/integrals.f90:1182-1253
 *   orc_multipole_grad_cgto   integral/native/integrals.f90:1256-1391
 *   orc_shift_operator        xtb/h0.f90:704-747
 *   orc_get_selfenergy        xtb/h0.f90:113-217 (CN branch; GFN2 has no cn_en / q terms)
 *   orc_new_adjacency_list    adjlist.f90:96-143
 *   orc_get_hamiltonian       xtb/h0.f90:220-429 (do_diat_scale = .false. for GFN2)
 *   orc_get_hamiltonian_diat  the same with do_diat_scale = .true. (the CEH
 *                             parametrisation, ceh/ceh.f90:1207-1242)
 *   orc_harmtr, orc_diat_trafo  integral/diat_trafo.f90:221-335 and :99-156
 *                             with scale_diatomic_frame :642-689 (no gradient)
 *   orc_get_overlap_diat_lat  integral/native/integrals.f90:473-582
 *   orc_get_hamiltonian_gradient  xtb/h0.f90:432-666
 *   orc_get_hamiltonian_adjoint   xtb/h0.f90:220-429 differentiated as in :432-666
 *                             (the adjoint of get_hamiltonian for arbitrary
 *                             output adjoints; test-only, no reference routine)
 *   orc_get_multipole_integrals_lat  integral/native/integrals.f90:1394-1465
 *                             (its S and S, D are get_overlap_lat :411-470 and
 *                             get_dipole_integrals_lat :994-1059)
 *   orc_ncoord_dexp           mctc-lib ncoord "dexp" as called at xtb/gfn2.f90:640,
 *                             xtb/singlepoint.f90:224 (third-party, absent:
 *                             restated from its published formula; unpinned)
 *
 * Conventions kept from the reference: all arrays are Fortran column-major,
 * indices are 0-based on this side; Cartesian components in CCA order;
 * p spherical order (y, z, x); d spherical order m = -2..2; quadrupoles
 * packed xx, xy, yy, xz, yz, zz and made traceless as 1.5 q - tr/2; the
 * multipole operator sits on the ket (column) function.
 *
 * Angular momenta: Cartesian integrals are general up to l = 4 (g); the
 * spherical transformation is implemented for l <= 2, the GFN2 maximum
 * (xtb/gfn2.f90:928).  Callers asking for a spherical block with l > 2 get
 * status -1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXG 12
#define ORC_MAXL 4
#define ORC_MAXL2 (2 * ORC_MAXL + 4)
#define ORC_NCART_MAX 15

typedef struct {
    int ang;
    int nprim;
    double alpha[ORC_MAXG];
    double coeff[ORC_MAXG];
} orc_cgto;

static const double ORC_PI = 3.14159265358979323846264338327950288;

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int nsph(int l) { return 2 * l + 1; }

/* Cartesian exponents in CCA order: decreasing lx, then decreasing ly. */
static void cart_exponents(int l, int idx, int e[3])
{
    int k = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            if (k == idx) {
                e[0] = lx;
                e[1] = ly;
                e[2] = l - lx - ly;
                return;
            }
            ++k;
        }
}

/* (2n-1)!! for n = 0..7 */
static double dfact(int n)
{
    double r = 1.0;
    for (int k = 1; k <= n; ++k) r *= (double)(2 * k - 1);
    return r;
}

/* Gaussian moment int x^m exp(-a x^2) dx / sqrt(pi/a): integrals.f90:91-103 */
double orc_overlap_1d(int moment, double alpha)
{
    if (moment % 2 != 0) return 0.0;
    double base = 0.5 / alpha, p = 1.0;
    for (int k = 0; k < moment / 2; ++k) p *= base;
    return p * dfact(moment / 2);
}

/* Re-expand the top coefficient c[l] of a polynomial in (x - A) about the
 * product centre: (y + a)^l = sum_t binom(l, t) a^(l-t) y^t.  The other
 * coefficients are added to, as horizontal_shift does (integrals.f90:105-130). */
static void orc_shift_poly(double a, int l, double *c)
{
    if (l <= 0) return;
    double top = c[l];
    double apow = 1.0, binom = 1.0;
    /* binom(l, t) a^(l-t) for t = l-1 .. 0 */
    for (int t = l - 1; t >= 0; --t) {
        apow *= a;
        binom = binom * (double)(t + 1) / (double)(l - t);
        c[t] += binom * apow * top;
    }
}

/* Polynomial product d = a * b for degrees la, lb (form_product). */
static void orc_poly_product(const double *a, const double *b, int la, int lb, double *d)
{
    for (int k = 0; k <= la + lb; ++k) d[k] = 0.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) d[i + j] += a[i] * b[j];
}

/* One primitive pair, one Cartesian pair: overlap, dipole (3) and raw
 * quadrupole (6) with the operator about centre i (rpi = P - R_i):
 * integrals.f90:759-804. */
static void orc_multipole_3d(const double rpj[3], const double rpi[3], const int lj[3],
                             const int li[3], const double *s1d, double *s3d, double d3d[3],
                             double q3d[6])
{
    double v1d[3][3];
    for (int k = 0; k < 3; ++k) {
        double vi[ORC_MAXL2], vj[ORC_MAXL2], vv[ORC_MAXL2];
        memset(vi, 0, sizeof(vi));
        memset(vj, 0, sizeof(vj));
        vi[li[k]] = 1.0;
        vj[lj[k]] = 1.0;
        orc_shift_poly(rpi[k], li[k], vi);
        orc_shift_poly(rpj[k], lj[k], vj);
        orc_poly_product(vi, vj, li[k], lj[k], vv);
        double r = rpi[k];
        v1d[k][0] = v1d[k][1] = v1d[k][2] = 0.0;
        for (int l = 0; l <= li[k] + lj[k]; ++l) {
            v1d[k][0] += s1d[l] * vv[l];
            v1d[k][1] += (s1d[l + 1] + r * s1d[l]) * vv[l];
            v1d[k][2] += (s1d[l + 2] + 2 * r * s1d[l + 1] + r * r * s1d[l]) * vv[l];
        }
    }
    *s3d = v1d[0][0] * v1d[1][0] * v1d[2][0];
    d3d[0] = v1d[0][1] * v1d[1][0] * v1d[2][0];
    d3d[1] = v1d[0][0] * v1d[1][1] * v1d[2][0];
    d3d[2] = v1d[0][0] * v1d[1][0] * v1d[2][1];
    q3d[0] = v1d[0][2] * v1d[1][0] * v1d[2][0];
    q3d[1] = v1d[0][1] * v1d[1][1] * v1d[2][0];
    q3d[2] = v1d[0][0] * v1d[1][2] * v1d[2][0];
    q3d[3] = v1d[0][1] * v1d[1][0] * v1d[2][1];
    q3d[4] = v1d[0][0] * v1d[1][1] * v1d[2][1];
    q3d[5] = v1d[0][0] * v1d[1][0] * v1d[2][2];
}

/* Same with derivatives with respect to R_i; moments about centre j
 * (rpc = rpj, integrals.f90:835).  ds[3], dd[3][3] (dd[dir][comp]) and
 * dq[3][6] are laid out like the Fortran (3,3) / (3,6) arrays, first index
 * = derivative direction: element (a,b) at a + 3*b. */
static void orc_multipole_grad_3d(const double rpj[3], const double rpi[3], double ai,
                                  const int lj[3], const int li[3], const double *s1d,
                                  double *s3d, double d3d[3], double q3d[6], double ds3d[3],
                                  double dd3d[9], double dq3d[18])
{
    double v1d[3][3], g1d[3][3];
    for (int k = 0; k < 3; ++k) {
        double vi[ORC_MAXL2], vj[ORC_MAXL2], vv[ORC_MAXL2], gi[ORC_MAXL2], gg[ORC_MAXL2];
        memset(vi, 0, sizeof(vi));
        memset(vj, 0, sizeof(vj));
        memset(gi, 0, sizeof(gi));
        memset(vv, 0, sizeof(vv));
        memset(gg, 0, sizeof(gg));
        double rpc = rpj[k];
        vi[li[k]] = 1.0;
        vj[lj[k]] = 1.0;
        gi[li[k] + 1] = 2 * ai;
        if (li[k] > 0) gi[li[k] - 1] = -li[k];
        orc_shift_poly(rpi[k], li[k] - 1, gi);
        orc_shift_poly(rpi[k], li[k] + 1, gi);
        orc_shift_poly(rpi[k], li[k], vi);
        orc_shift_poly(rpj[k], lj[k], vj);
        orc_poly_product(vi, vj, li[k], lj[k], vv);
        orc_poly_product(gi, vj, li[k] + 1, lj[k], gg);
        for (int m = 0; m < 3; ++m) v1d[k][m] = g1d[k][m] = 0.0;
        for (int l = 0; l <= li[k] + lj[k] + 1; ++l) {
            double m0 = s1d[l];
            double m1 = s1d[l + 1] + rpc * s1d[l];
            double m2 = s1d[l + 2] + 2 * rpc * s1d[l + 1] + rpc * rpc * s1d[l];
            v1d[k][0] += m0 * vv[l];
            v1d[k][1] += m1 * vv[l];
            v1d[k][2] += m2 * vv[l];
            g1d[k][0] += m0 * gg[l];
            g1d[k][1] += m1 * gg[l];
            g1d[k][2] += m2 * gg[l];
        }
    }
    /* which moment (0,1,2) each direction carries for s, d(3), q(6) */
    static const int mom_s[3] = {0, 0, 0};
    static const int mom_d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    static const int mom_q[6][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0},
                                    {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};
    *s3d = v1d[0][0] * v1d[1][0] * v1d[2][0];
    for (int c = 0; c < 3; ++c)
        d3d[c] = v1d[0][mom_d[c][0]] * v1d[1][mom_d[c][1]] * v1d[2][mom_d[c][2]];
    for (int c = 0; c < 6; ++c)
        q3d[c] = v1d[0][mom_q[c][0]] * v1d[1][mom_q[c][1]] * v1d[2][mom_q[c][2]];
    for (int a = 0; a < 3; ++a) {
        /* derivative along direction a replaces v1d by g1d in that direction */
        double f[3];
        for (int k = 0; k < 3; ++k) f[k] = (k == a) ? g1d[k][mom_s[k]] : v1d[k][mom_s[k]];
        ds3d[a] = f[0] * f[1] * f[2];
        for (int c = 0; c < 3; ++c) {
            double p = 1.0;
            for (int k = 0; k < 3; ++k) p *= (k == a) ? g1d[k][mom_d[c][k]] : v1d[k][mom_d[c][k]];
            dd3d[a + 3 * c] = p;
        }
        for (int c = 0; c < 6; ++c) {
            double p = 1.0;
            for (int k = 0; k < 3; ++k) p *= (k == a) ? g1d[k][mom_q[c][k]] : v1d[k][mom_q[c][k]];
            dq3d[a + 3 * c] = p;
        }
    }
}

/* Shift the operator of the derivative integrals from centre j to centre i
 * (integrals.f90:902-936; called there with the dummy names swapped, :1337-1339).
 * In: s, d (op on j), ds, dd (op on j); out: ddo, dqo (op on i). */
static void orc_shift_operator_grad(const double vec[3], double s, const double *d,
                                    const double *ds, const double *dd, const double *dq,
                                    double *ddo, double *dqo)
{
#define DD(a, c) dd[(a) + 3 * (c)]
#define DQ(a, c) dq[(a) + 3 * (c)]
#define DDO(a, c) ddo[(a) + 3 * (c)]
#define DQO(a, c) dqo[(a) + 3 * (c)]
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) DDO(a, c) = DD(a, c) - vec[c] * ds[a];
    DDO(0, 0) -= s;
    DDO(1, 1) -= s;
    DDO(2, 2) -= s;
    for (int a = 0; a < 3; ++a) {
        DQO(a, 0) = DQ(a, 0) - 2 * vec[0] * DD(a, 0) + vec[0] * vec[0] * ds[a];
        DQO(a, 2) = DQ(a, 2) - 2 * vec[1] * DD(a, 1) + vec[1] * vec[1] * ds[a];
        DQO(a, 5) = DQ(a, 5) - 2 * vec[2] * DD(a, 2) + vec[2] * vec[2] * ds[a];
        DQO(a, 1) = DQ(a, 1) - vec[0] * DD(a, 1) - vec[1] * DD(a, 0) + vec[0] * vec[1] * ds[a];
        DQO(a, 3) = DQ(a, 3) - vec[0] * DD(a, 2) - vec[2] * DD(a, 0) + vec[0] * vec[2] * ds[a];
        DQO(a, 4) = DQ(a, 4) - vec[1] * DD(a, 2) - vec[2] * DD(a, 1) + vec[1] * vec[2] * ds[a];
    }
    DQO(0, 0) += -2 * d[0] + 2 * vec[0] * s;
    DQO(1, 2) += -2 * d[1] + 2 * vec[1] * s;
    DQO(2, 5) += -2 * d[2] + 2 * vec[2] * s;
    DQO(0, 1) += -d[1] + vec[1] * s;
    DQO(1, 1) += -d[0] + vec[0] * s;
    DQO(0, 3) += -d[2] + vec[2] * s;
    DQO(2, 3) += -d[0] + vec[0] * s;
    DQO(1, 4) += -d[2] + vec[2] * s;
    DQO(2, 4) += -d[1] + vec[1] * s;
#undef DD
#undef DQ
#undef DDO
#undef DQO
}

/* Cartesian -> spherical coefficient matrix T(l): sph[m] = sum_c T[m][c] cart[c].
 * trafo.f90:34-44 (d) and :252-253 (p order y, z, x). */
static int sph_matrix(int l, double T[5][6])
{
    memset(T, 0, sizeof(double) * 30);
    const double s3 = sqrt(3.0);
    switch (l) {
    case 0:
        T[0][0] = 1.0;
        return 0;
    case 1:
        T[0][1] = 1.0; /* y */
        T[1][2] = 1.0; /* z */
        T[2][0] = 1.0; /* x */
        return 0;
    case 2: /* cart order xx xy xz yy yz zz */
        T[0][1] = s3;            /* m=-2: sqrt3 xy */
        T[1][4] = s3;            /* m=-1: sqrt3 yz */
        T[2][5] = 1.0;           /* m= 0: zz - (xx+yy)/2 */
        T[2][0] = -0.5;
        T[2][3] = -0.5;
        T[3][2] = s3;            /* m= 1: sqrt3 xz */
        T[4][0] = 0.5 * s3;      /* m= 2: sqrt3/2 (xx - yy) */
        T[4][3] = -0.5 * s3;
        return 0;
    default:
        return -1;
    }
}

/* sph(nsj, nsi) = T(lj) cart(ncj, nci) T(li)^T, one component: trafo.f90:104-257 */
static int orc_transform0(int lj, int li, const double *cart, int cstride, double *sphr,
                          int sstride)
{
    double Tj[5][6], Ti[5][6];
    if (sph_matrix(lj, Tj) || sph_matrix(li, Ti)) return -1;
    int ncj = ncart(lj), nci = ncart(li), nsj = nsph(lj), nsi = nsph(li);
    for (int mi = 0; mi < nsi; ++mi)
        for (int mj = 0; mj < nsj; ++mj) {
            double acc = 0.0;
            for (int ci = 0; ci < nci; ++ci) {
                if (Ti[mi][ci] == 0.0) continue;
                double t = 0.0;
                for (int cj = 0; cj < ncj; ++cj) t += Tj[mj][cj] * cart[cstride * (cj + ncj * ci)];
                acc += t * Ti[mi][ci];
            }
            sphr[sstride * (mj + nsj * mi)] = acc;
        }
    return 0;
}

static void prim_pair(const orc_cgto *cj, const orc_cgto *ci, int ip, int jp, double r2,
                      const double vec[3], double *oab, double *est, double rpi[3],
                      double rpj[3])
{
    double eab = ci->alpha[ip] + cj->alpha[jp];
    *oab = 1.0 / eab;
    *est = ci->alpha[ip] * cj->alpha[jp] * r2 * *oab;
    for (int k = 0; k < 3; ++k) {
        rpi[k] = -vec[k] * cj->alpha[jp] * *oab;
        rpj[k] = +vec[k] * ci->alpha[ip] * *oab;
    }
}

static double prim_prefactor(double est, double oab)
{
    double so = sqrt(oab);
    return exp(-est) * (sqrt(ORC_PI) * ORC_PI) * so * so * so;
}

/* Cartesian overlap block s3d(ncj, nci) (accumulated from zero). */
static void cart_overlap(const orc_cgto *cj, const orc_cgto *ci, double r2, const double vec[3],
                         double intcut, double *s3d)
{
    int ncj = ncart(cj->ang), nci = ncart(ci->ang);
    memset(s3d, 0, sizeof(double) * ncj * nci);
    for (int ip = 0; ip < ci->nprim; ++ip)
        for (int jp = 0; jp < cj->nprim; ++jp) {
            double oab, est, rpi[3], rpj[3], s1d[ORC_MAXL2];
            prim_pair(cj, ci, ip, jp, r2, vec, &oab, &est, rpi, rpj);
            if (est > intcut) continue;
            double pre = prim_prefactor(est, oab);
            double eab = ci->alpha[ip] + cj->alpha[jp];
            for (int l = 0; l <= ci->ang + cj->ang + 2; ++l) s1d[l] = orc_overlap_1d(l, eab);
            double cc = ci->coeff[ip] * cj->coeff[jp] * pre;
            for (int mi = 0; mi < nci; ++mi)
                for (int mj = 0; mj < ncj; ++mj) {
                    int li[3], lj[3];
                    double s, d[3], q[6];
                    cart_exponents(ci->ang, mi, li);
                    cart_exponents(cj->ang, mj, lj);
                    orc_multipole_3d(rpj, rpi, lj, li, s1d, &s, d, q);
                    s3d[mj + ncj * mi] += cc * s;
                }
        }
}

/* Cartesian overlap block (test hook for l > 2 normalisation checks). */
int orc_cart_overlap(const orc_cgto *cj, const orc_cgto *ci, double r2, const double *vec,
                     double intcut, double *s3d)
{
    if (cj->ang > ORC_MAXL || ci->ang > ORC_MAXL) return -1;
    cart_overlap(cj, ci, r2, vec, intcut, s3d);
    return 0;
}

/* overlap_cgto: integrals.f90:315-360 */
int orc_overlap_cgto(const orc_cgto *cj, const orc_cgto *ci, double r2, const double *vec,
                     double intcut, double *overlap)
{
    double s3d[ORC_NCART_MAX * ORC_NCART_MAX];
    if (cj->ang > 2 || ci->ang > 2) return -1;
    cart_overlap(cj, ci, r2, vec, intcut, s3d);
    return orc_transform0(cj->ang, ci->ang, s3d, 1, overlap, 1);
}

static void remove_trace(double *q)
{
    double tr = 0.5 * (q[0] + q[2] + q[5]);
    q[0] = 1.5 * q[0] - tr;
    q[1] = 1.5 * q[1];
    q[2] = 1.5 * q[2] - tr;
    q[3] = 1.5 * q[3];
    q[4] = 1.5 * q[4];
    q[5] = 1.5 * q[5] - tr;
}

/* multipole_cgto: integrals.f90:1182-1253.
 * overlap(nsj,nsi), dpint(3,nsj,nsi), qpint(6,nsj,nsi); operator on centre i. */
int orc_multipole_cgto(const orc_cgto *cj, const orc_cgto *ci, double r2, const double *vec,
                       double intcut, double *overlap, double *dpint, double *qpint)
{
    if (cj->ang > 2 || ci->ang > 2) return -1;
    int ncj = ncart(cj->ang), nci = ncart(ci->ang);
    int nsj = nsph(cj->ang), nsi = nsph(ci->ang);
    double s3d[36], d3d[3 * 36], q3d[6 * 36];
    memset(s3d, 0, sizeof(s3d));
    memset(d3d, 0, sizeof(d3d));
    memset(q3d, 0, sizeof(q3d));
    for (int ip = 0; ip < ci->nprim; ++ip)
        for (int jp = 0; jp < cj->nprim; ++jp) {
            double oab, est, rpi[3], rpj[3], s1d[ORC_MAXL2];
            prim_pair(cj, ci, ip, jp, r2, vec, &oab, &est, rpi, rpj);
            if (est > intcut) continue;
            double pre = prim_prefactor(est, oab);
            double eab = ci->alpha[ip] + cj->alpha[jp];
            for (int l = 0; l <= ci->ang + cj->ang + 2; ++l) s1d[l] = orc_overlap_1d(l, eab);
            double cc = ci->coeff[ip] * cj->coeff[jp] * pre;
            for (int mi = 0; mi < nci; ++mi)
                for (int mj = 0; mj < ncj; ++mj) {
                    int li[3], lj[3];
                    double s, d[3], q[6];
                    cart_exponents(ci->ang, mi, li);
                    cart_exponents(cj->ang, mj, lj);
                    orc_multipole_3d(rpj, rpi, lj, li, s1d, &s, d, q);
                    int ij = mj + ncj * mi;
                    s3d[ij] += cc * s;
                    for (int k = 0; k < 3; ++k) d3d[k + 3 * ij] += cc * d[k];
                    for (int k = 0; k < 6; ++k) q3d[k + 6 * ij] += cc * q[k];
                }
        }
    orc_transform0(cj->ang, ci->ang, s3d, 1, overlap, 1);
    for (int k = 0; k < 3; ++k) orc_transform0(cj->ang, ci->ang, d3d + k, 3, dpint + k, 3);
    for (int k = 0; k < 6; ++k) orc_transform0(cj->ang, ci->ang, q3d + k, 6, qpint + k, 6);
    for (int ij = 0; ij < nsj * nsi; ++ij) remove_trace(qpint + 6 * ij);
    return 0;
}

static void remove_trace_grad(double *dq /* (3,6) */)
{
    for (int a = 0; a < 3; ++a) {
        double dtr = dq[a + 0] + dq[a + 6] + dq[a + 15];
        dq[a + 0] = 1.5 * dq[a + 0] - 0.5 * dtr;
        dq[a + 3] = 1.5 * dq[a + 3];
        dq[a + 6] = 1.5 * dq[a + 6] - 0.5 * dtr;
        dq[a + 9] = 1.5 * dq[a + 9];
        dq[a + 12] = 1.5 * dq[a + 12];
        dq[a + 15] = 1.5 * dq[a + 15] - 0.5 * dtr;
    }
}

/* multipole_grad_cgto: integrals.f90:1256-1391.  Argument order as the
 * Fortran dummy list: overlap, dpint (op on j), qpint (op on j), doverlap
 * (d/dR_i), ddpintj, dqpintj, ddpinti, dqpinti. */
int orc_multipole_grad_cgto(const orc_cgto *cj, const orc_cgto *ci, double r2, const double *vec,
                            double intcut, double *overlap, double *dpint, double *qpint,
                            double *doverlap, double *ddpintj, double *dqpintj, double *ddpinti,
                            double *dqpinti)
{
    if (cj->ang > 2 || ci->ang > 2) return -1;
    int ncj = ncart(cj->ang), nci = ncart(ci->ang);
    int nsj = nsph(cj->ang), nsi = nsph(ci->ang);
    enum { NC = 36 };
    double s3d[NC], d3dj[3 * NC], q3dj[6 * NC], ds3d[3 * NC];
    double dd3dj[9 * NC], dq3dj[18 * NC], dd3di[9 * NC], dq3di[18 * NC];
    memset(s3d, 0, sizeof(s3d));
    memset(d3dj, 0, sizeof(d3dj));
    memset(q3dj, 0, sizeof(q3dj));
    memset(ds3d, 0, sizeof(ds3d));
    memset(dd3dj, 0, sizeof(dd3dj));
    memset(dq3dj, 0, sizeof(dq3dj));
    for (int ip = 0; ip < ci->nprim; ++ip)
        for (int jp = 0; jp < cj->nprim; ++jp) {
            double oab, est, rpi[3], rpj[3], s1d[ORC_MAXL2];
            prim_pair(cj, ci, ip, jp, r2, vec, &oab, &est, rpi, rpj);
            if (est > intcut) continue;
            double pre = prim_prefactor(est, oab);
            double eab = ci->alpha[ip] + cj->alpha[jp];
            for (int l = 0; l <= ci->ang + cj->ang + 3; ++l) s1d[l] = orc_overlap_1d(l, eab);
            double cc = ci->coeff[ip] * cj->coeff[jp] * pre;
            for (int mi = 0; mi < nci; ++mi)
                for (int mj = 0; mj < ncj; ++mj) {
                    int li[3], lj[3];
                    double s, d[3], q[6], ds[3], dd[9], dq[18];
                    cart_exponents(ci->ang, mi, li);
                    cart_exponents(cj->ang, mj, lj);
                    orc_multipole_grad_3d(rpj, rpi, ci->alpha[ip], lj, li, s1d, &s, d, q, ds, dd,
                                          dq);
                    int ij = mj + ncj * mi;
                    s3d[ij] += cc * s;
                    for (int k = 0; k < 3; ++k) d3dj[k + 3 * ij] += cc * d[k];
                    for (int k = 0; k < 6; ++k) q3dj[k + 6 * ij] += cc * q[k];
                    for (int k = 0; k < 3; ++k) ds3d[k + 3 * ij] += cc * ds[k];
                    for (int k = 0; k < 9; ++k) dd3dj[k + 9 * ij] += cc * dd[k];
                    for (int k = 0; k < 18; ++k) dq3dj[k + 18 * ij] += cc * dq[k];
                }
        }
    for (int ij = 0; ij < ncj * nci; ++ij)
        orc_shift_operator_grad(vec, s3d[ij], d3dj + 3 * ij, ds3d + 3 * ij, dd3dj + 9 * ij,
                                dq3dj + 18 * ij, dd3di + 9 * ij, dq3di + 18 * ij);
    orc_transform0(cj->ang, ci->ang, s3d, 1, overlap, 1);
    for (int k = 0; k < 3; ++k) orc_transform0(cj->ang, ci->ang, d3dj + k, 3, dpint + k, 3);
    for (int k = 0; k < 6; ++k) orc_transform0(cj->ang, ci->ang, q3dj + k, 6, qpint + k, 6);
    for (int k = 0; k < 3; ++k) orc_transform0(cj->ang, ci->ang, ds3d + k, 3, doverlap + k, 3);
    for (int k = 0; k < 9; ++k) {
        orc_transform0(cj->ang, ci->ang, dd3dj + k, 9, ddpintj + k, 9);
        orc_transform0(cj->ang, ci->ang, dd3di + k, 9, ddpinti + k, 9);
    }
    for (int k = 0; k < 18; ++k) {
        orc_transform0(cj->ang, ci->ang, dq3dj + k, 18, dqpintj + k, 18);
        orc_transform0(cj->ang, ci->ang, dq3di + k, 18, dqpinti + k, 18);
    }
    for (int ij = 0; ij < nsj * nsi; ++ij) {
        remove_trace(qpint + 6 * ij);
        remove_trace_grad(dqpinti + 18 * ij);
        remove_trace_grad(dqpintj + 18 * ij);
    }
    return 0;
}

/* shift_operator of the forward path: xtb/h0.f90:704-747 (op on i -> op on j). */
void orc_shift_operator(const double *vec, double s, const double *di, const double *qi,
                        double *dj, double *qj)
{
    for (int k = 0; k < 3; ++k) dj[k] = di[k] + vec[k] * s;
    qj[0] = 2 * vec[0] * di[0] + vec[0] * vec[0] * s;
    qj[2] = 2 * vec[1] * di[1] + vec[1] * vec[1] * s;
    qj[5] = 2 * vec[2] * di[2] + vec[2] * vec[2] * s;
    qj[1] = vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * s;
    qj[3] = vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * s;
    qj[4] = vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * s;
    double tr = 0.5 * (qj[0] + qj[2] + qj[5]);
    qj[0] = qi[0] + 1.5 * qj[0] - tr;
    qj[1] = qi[1] + 1.5 * qj[1];
    qj[2] = qi[2] + 1.5 * qj[2] - tr;
    qj[3] = qi[3] + 1.5 * qj[3];
    qj[4] = qi[4] + 1.5 * qj[4];
    qj[5] = qi[5] + 1.5 * qj[5] - tr;
}

/* ------------------------------------------------------------------------ */
/* System-level routines                                                     */
/* ------------------------------------------------------------------------ */

/* Flattened basis: shell ish (0-based) of species isp is cgto[ish + msh*isp];
 * nsh_id[isp]; ish_at[iat] first shell of atom; iao_sh[sh] first AO of shell. */
typedef struct {
    int nat, nid, msh, nsh, nao;
    const int *id;      /* [nat], 0-based species */
    const double *xyz;  /* [3*nat] */
    const int *nsh_id;  /* [nid] */
    const int *ish_at;  /* [nat] */
    const int *iao_sh;  /* [nsh] */
    const orc_cgto *cgto;
    double intcut;
    const double *hscale; /* hscale(jsh, ish, jsp, isp): jsh + msh*(ish + msh*(jsp + nid*isp)) */
    const double *shpoly; /* shpoly(ish, isp) */
    const double *rad;    /* rad(isp) */
} orc_system;

/* get_selfenergy (h0.f90:113-217), cn branch with derivative. */
void orc_get_selfenergy(int nat, const int *id, const int *ish_at, const int *nsh_id, int msh,
                        const double *selfenergy_sp, const double *kcn_sp, const double *cn,
                        double *selfenergy, double *dsedcn)
{
    for (int iat = 0; iat < nat; ++iat) {
        int izp = id[iat], ii = ish_at[iat];
        for (int ish = 0; ish < nsh_id[izp]; ++ish) {
            selfenergy[ii + ish] = selfenergy_sp[ish + msh * izp];
            if (cn) selfenergy[ii + ish] -= kcn_sp[ish + msh * izp] * cn[iat];
            if (dsedcn) dsedcn[ii + ish] = cn ? -kcn_sp[ish + msh * izp] : 0.0;
        }
    }
}

/* new_adjacency_list / generate (adjlist.f90:96-143), half list (complete = .false.).
 * First call with nlat == NULL returns the number of entries; second call fills. */
int orc_new_adjacency_list(int nat, const double *xyz, int ntr, const double *trans,
                           double cutoff, int *inl, int *nnl, int *nlat, int *nltr)
{
    int img = 0;
    double cutoff2 = cutoff * cutoff;
    const double eps = 2.220446049250313e-16;
    for (int iat = 0; iat < nat; ++iat) {
        if (inl) inl[iat] = img;
        for (int jat = 0; jat <= iat; ++jat)
            for (int itr = 0; itr < ntr; ++itr) {
                double v[3], r2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    v[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
                    r2 += v[k] * v[k];
                }
                if (r2 < eps || r2 > cutoff2) continue;
                if (nlat) {
                    nlat[img] = jat;
                    nltr[img] = itr;
                }
                ++img;
            }
        if (nnl) nnl[iat] = img - (inl ? inl[iat] : 0);
    }
    return img;
}

/* --- diatomic-frame scaling (integral/diat_trafo.f90) --------------------
 * Slots of an atom's block: 0 s; 1-3 p (y, z, x); 4-8 d (m = -2..2), as
 * smap/sdim place the shells s, p, d of an atom (integrals.f90:56-58).
 * Matrices are 9 x 9, column-major. */

/* harmtr (diat_trafo.f90:221-335): R[row + 9 col] takes the real spherical
 * functions of the molecular frame (rows) to those of the diatomic frame
 * whose z axis is vec (columns), up to angular momentum maxl. */
void orc_harmtr(int maxl, const double *vec, double *R)
{
#define T_(r, c) R[((r) - 1) + 9 * ((c) - 1)]
    for (int k = 0; k < 81; ++k) R[k] = 0.0;
    T_(1, 1) = 1.0;
    if (maxl == 0) return;
    const double len = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    const double n[3] = {vec[0] / len, vec[1] / len, vec[2] / len};
    const double cost = n[2];
    double sint, cosp, sinp;
    if (fabs(cost) == 1.0) {  /* phi arbitrary on the z axis: x chosen */
        sint = 0.0;
        cosp = 1.0;
        sinp = 0.0;
    } else if (fabs(cost) == 0.0) {
        sint = 1.0;
        cosp = n[0];
        sinp = n[1];
    } else {
        sint = sqrt(n[0] * n[0] + n[1] * n[1]);
        cosp = n[0] / sint;
        sinp = n[1] / sint;
    }
    T_(2, 2) = cosp;        T_(3, 2) = 0.0;   T_(4, 2) = -sinp;
    T_(2, 3) = sint * sinp; T_(3, 3) = cost;  T_(4, 3) = sint * cosp;
    T_(2, 4) = cost * sinp; T_(3, 4) = -sint; T_(4, 4) = cost * cosp;
    if (maxl <= 1) return;
    const double cos2t = cost * cost - sint * sint, sin2t = 2.0 * sint * cost;
    const double cos2p = cosp * cosp - sinp * sinp, sin2p = 2.0 * sinp * cosp;
    const double sqrt3 = sqrt(3.0);
    T_(5, 5) = cost * cos2p;
    T_(6, 5) = -sint * cosp;
    T_(7, 5) = 0.0;
    T_(8, 5) = sint * sinp;
    T_(9, 5) = -cost * sin2p;
    T_(5, 6) = sint * cos2p;
    T_(6, 6) = cost * cosp;
    T_(7, 6) = 0.0;
    T_(8, 6) = -cost * sinp;
    T_(9, 6) = -sint * sin2p;
    T_(5, 7) = sqrt3 * sint * sint * sin2p * 0.5;
    T_(6, 7) = sqrt3 * sin2t * sinp * 0.5;
    T_(7, 7) = (3.0 * cost * cost - 1.0) * 0.5;
    T_(8, 7) = sqrt3 * sin2t * cosp * 0.5;
    T_(9, 7) = sqrt3 * sint * sint * cos2p * 0.5;
    T_(5, 8) = sin2t * sin2p * 0.5;
    T_(6, 8) = cos2t * sinp;
    T_(7, 8) = -sqrt3 * sin2t * 0.5;
    T_(8, 8) = cos2t * cosp;
    T_(9, 8) = sin2t * cos2p * 0.5;
    T_(5, 9) = (1.0 + cost * cost) * sin2p * 0.5;
    T_(6, 9) = -sin2t * sinp * 0.5;
    T_(7, 9) = sqrt3 * sint * sint * 0.5;
    T_(8, 9) = -sin2t * cosp * 0.5;
    T_(9, 9) = (1.0 + cost * cost) * cos2p * 0.5;
#undef T_
}

/* scale_diatomic_frame (diat_trafo.f90:642-689): sigma, pi and delta
 * elements of a diatomic-frame block M (rows: atom j up to maxlj, columns:
 * atom i up to maxli). */
static void scale_diat(double *M, double ksig, double kpi, double kdel, int maxlj, int maxli)
{
#define M_(r, c) M[((r) - 1) + 9 * ((c) - 1)]
    M_(1, 1) *= ksig;
    if (maxlj > 0) M_(3, 1) *= ksig;
    if (maxli > 0) M_(1, 3) *= ksig;
    if (maxlj > 0 && maxli > 0) {
        M_(3, 3) *= ksig;
        M_(4, 4) *= kpi;
        M_(2, 2) *= kpi;
        if (maxlj > 1) {
            M_(7, 3) *= ksig;
            M_(8, 4) *= kpi;
            M_(6, 2) *= kpi;
        }
        if (maxli > 1) {
            M_(3, 7) *= ksig;
            M_(4, 8) *= kpi;
            M_(2, 6) *= kpi;
        }
    }
    if (maxlj > 1) M_(7, 1) *= ksig;
    if (maxli > 1) M_(1, 7) *= ksig;
    if (maxlj > 1 && maxli > 1) {
        M_(7, 7) *= ksig;
        M_(8, 8) *= kpi;
        M_(6, 6) *= kpi;
        M_(9, 9) *= kdel;
        M_(5, 5) *= kdel;
    }
#undef M_
}

/* setup_diat_trafo + diat_trafo without derivatives (diat_trafo.f90:61-156):
 * the atom-pair overlap block B (9 x 9, rows atom j, columns atom i) is
 * taken to the diatomic frame, S' = R_j^T B R_i, scaled, and taken back,
 * B = R_j S'' R_i^T.  maxlj / maxli: highest angular momentum of the atoms
 * (their shells are s, p, d in this order). */
void orc_diat_trafo(const double *vec, int maxlj, int maxli, double ksig, double kpi,
                    double kdel, double *B)
{
    if (ksig == 1.0 && kpi == 1.0 && kdel == 1.0) return;
    static const int sdim[3] = {1, 4, 9};
    maxlj = maxlj < 0 ? 0 : (maxlj > 2 ? 2 : maxlj);
    maxli = maxli < 0 ? 0 : (maxli > 2 ? 2 : maxli);
    const int dj = sdim[maxlj], di = sdim[maxli];
    if (dj == 1 && di == 1) {
        B[0] *= ksig;
        return;
    }
    double R[81], tmp[81], dm[81];
    orc_harmtr(maxlj > maxli ? maxlj : maxli, vec, R);
    for (int b = 0; b < di; ++b)  /* tmp = R^T B (dj x di) */
        for (int a = 0; a < dj; ++a) {
            double t = 0.0;
            for (int c = 0; c < dj; ++c) t += R[c + 9 * a] * B[c + 9 * b];
            tmp[a + 9 * b] = t;
        }
    for (int b = 0; b < 81; ++b) dm[b] = 0.0;
    for (int b = 0; b < di; ++b)  /* S' = tmp R */
        for (int a = 0; a < dj; ++a) {
            double t = 0.0;
            for (int d = 0; d < di; ++d) t += tmp[a + 9 * d] * R[d + 9 * b];
            dm[a + 9 * b] = t;
        }
    scale_diat(dm, ksig, kpi, kdel, maxlj, maxli);
    for (int b = 0; b < di; ++b)  /* tmp = R S'' */
        for (int a = 0; a < dj; ++a) {
            double t = 0.0;
            for (int c = 0; c < dj; ++c) t += R[a + 9 * c] * dm[c + 9 * b];
            tmp[a + 9 * b] = t;
        }
    for (int b = 0; b < di; ++b)  /* B = tmp R^T */
        for (int a = 0; a < dj; ++a) {
            double t = 0.0;
            for (int d = 0; d < di; ++d) t += tmp[a + 9 * d] * R[b + 9 * d];
            B[a + 9 * b] = t;
        }
}

/* The shells of every species are s, p, d in this order (shell k has l = k),
 * as the diatomic-frame slots assume (smap(ish - 1), h0.f90:288-294). */
static int diat_shells_ok(int nid, int msh, const int *nsh_id, const orc_cgto *cgto)
{
    for (int isp = 0; isp < nid; ++isp)
        for (int ish = 0; ish < nsh_id[isp]; ++ish)
            if (cgto[ish + msh * isp].ang != ish) return 0;
    return 1;
}

#define HSCALE(sys, jsh, ish, jsp, isp) \
    (sys)->hscale[(jsh) + (sys)->msh * ((ish) + (sys)->msh * ((jsp) + (sys)->nid * (isp)))]

/* get_hamiltonian: xtb/h0.f90:220-429.  ksig, kpi, kdel [nid*nid] (izp +
 * nid*jzp) switch the diatomic-frame scaling of the overlap in H0 on
 * (do_diat_scale, :257-259, :345-390), or are NULL. */
static int ham_impl(int nat, int nid, int msh, int nsh, int nao, const int *id,
                    const double *xyz, int ntr, const double *trans, const int *inl,
                    const int *nnl, const int *nlat, const int *nltr, const int *nsh_id,
                    const int *ish_at, const int *iao_sh, const orc_cgto *cgto, double intcut,
                    const double *hscale, const double *shpoly, const double *rad,
                    const double *ksig, const double *kpi, const double *kdel,
                    const double *selfenergy, double *overlap, double *dpint, double *qpint,
                    double *hamiltonian)
{
    const int diat = ksig != NULL;
    /* with the scaling, H0 is made of the scaled overlap alone */
    const double mod_h0_fraction = diat ? 1.0 : 0.0;
    orc_system sys = {nat, nid, msh, nsh, nao, id, xyz, nsh_id, ish_at, iao_sh, cgto,
                      intcut, hscale, shpoly, rad};
    size_t n2 = (size_t)nao * nao;
    memset(overlap, 0, sizeof(double) * n2);
    memset(dpint, 0, sizeof(double) * 3 * n2);
    memset(qpint, 0, sizeof(double) * 6 * n2);
    memset(hamiltonian, 0, sizeof(double) * n2);
    for (int isp = 0; isp < nid; ++isp)
        for (int ish = 0; ish < nsh_id[isp]; ++ish)
            if (cgto[ish + msh * isp].ang > 2) return -1;
    if (diat && !diat_shells_ok(nid, msh, nsh_id, cgto)) return -2;

#define S_(r, c) overlap[(size_t)(r) + (size_t)nao * (c)]
#define H_(r, c) hamiltonian[(size_t)(r) + (size_t)nao * (c)]
#define D_(k, r, c) dpint[(k) + 3 * ((size_t)(r) + (size_t)nao * (c))]
#define Q_(k, r, c) qpint[(k) + 6 * ((size_t)(r) + (size_t)nao * (c))]

#pragma omp parallel for schedule(dynamic)
    for (int iat = 0; iat < nat; ++iat) {
        double stmp[25], dtmpi[75], qtmpi[150], block[81];
        int izp = id[iat], is = ish_at[iat], nsi = nsh_id[izp];
        for (int img = 0; img < nnl[iat]; ++img) {
            int jat = nlat[inl[iat] + img], itr = nltr[inl[iat] + img];
            int jzp = id[jat], js = ish_at[jat], nsj = nsh_id[jzp];
            double vec[3], r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
            }
            r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
            double rr = sqrt(sqrt(r2) / (rad[jzp] + rad[izp]));
            if (diat)
                for (int k = 0; k < 81; ++k) block[k] = 0.0;
            for (int ish = 0; ish < nsi; ++ish) {
                int ii = iao_sh[is + ish];
                const orc_cgto *ci = &cgto[ish + msh * izp];
                double shpolyi = 1.0 + shpoly[ish + msh * izp] * rr;
                for (int jsh = 0; jsh < nsj; ++jsh) {
                    int jj = iao_sh[js + jsh];
                    const orc_cgto *cj = &cgto[jsh + msh * jzp];
                    orc_multipole_cgto(cj, ci, r2, vec, intcut, stmp, dtmpi, qtmpi);
                    double shp = shpolyi * (1.0 + shpoly[jsh + msh * jzp] * rr);
                    double hij = 0.5 * (1.0 - mod_h0_fraction) *
                                 (selfenergy[is + ish] + selfenergy[js + jsh]) *
                                 HSCALE(&sys, jsh, ish, jzp, izp) * shp;
                    int nj = nsph(cj->ang), ni = nsph(ci->ang);
                    for (int iao = 0; iao < ni; ++iao)
                        for (int jao = 0; jao < nj; ++jao) {
                            int ij = jao + nj * iao;
                            double dtmpj[3], qtmpj[6];
                            orc_shift_operator(vec, stmp[ij], dtmpi + 3 * ij, qtmpi + 6 * ij, dtmpj,
                                               qtmpj);
                            /* smap: shell k of the atom starts at slot k * k */
                            if (diat) block[jsh * jsh + jao + 9 * (ish * ish + iao)] = stmp[ij];
                            S_(jj + jao, ii + iao) += stmp[ij];
                            for (int k = 0; k < 3; ++k) D_(k, jj + jao, ii + iao) += dtmpi[3 * ij + k];
                            for (int k = 0; k < 6; ++k) Q_(k, jj + jao, ii + iao) += qtmpi[6 * ij + k];
                            H_(jj + jao, ii + iao) += stmp[ij] * hij;
                            if (iat != jat) {
                                S_(ii + iao, jj + jao) += stmp[ij];
                                for (int k = 0; k < 3; ++k) D_(k, ii + iao, jj + jao) += dtmpj[k];
                                for (int k = 0; k < 6; ++k) Q_(k, ii + iao, jj + jao) += qtmpj[k];
                                H_(ii + iao, jj + jao) += stmp[ij] * hij;
                            }
                        }
                }
            }
            if (!diat) continue;
            orc_diat_trafo(vec, nsj - 1, nsi - 1, ksig[izp + nid * jzp], kpi[izp + nid * jzp],
                           kdel[izp + nid * jzp], block);
            for (int ish = 0; ish < nsi; ++ish) {
                int ii = iao_sh[is + ish];
                double shpolyi = 1.0 + shpoly[ish + msh * izp] * rr;
                for (int jsh = 0; jsh < nsj; ++jsh) {
                    int jj = iao_sh[js + jsh];
                    double shp = shpolyi * (1.0 + shpoly[jsh + msh * jzp] * rr);
                    double hij = 0.5 * mod_h0_fraction *
                                 (selfenergy[is + ish] + selfenergy[js + jsh]) *
                                 HSCALE(&sys, jsh, ish, jzp, izp) * shp;
                    int nj = nsph(cgto[jsh + msh * jzp].ang), ni = nsph(cgto[ish + msh * izp].ang);
                    for (int iao = 0; iao < ni; ++iao)
                        for (int jao = 0; jao < nj; ++jao) {
                            double b = block[jsh * jsh + jao + 9 * (ish * ish + iao)];
                            H_(jj + jao, ii + iao) += b * hij;
                            if (iat != jat) H_(ii + iao, jj + jao) += b * hij;
                        }
                }
            }
        }
        /* onsite block */
        double vec0[3] = {0.0, 0.0, 0.0};
        for (int ish = 0; ish < nsi; ++ish) {
            int ii = iao_sh[is + ish];
            const orc_cgto *ci = &cgto[ish + msh * izp];
            for (int jsh = 0; jsh < nsi; ++jsh) {
                int jj = iao_sh[is + jsh];
                const orc_cgto *cj = &cgto[jsh + msh * izp];
                orc_multipole_cgto(cj, ci, 0.0, vec0, intcut, stmp, dtmpi, qtmpi);
                double hij = 0.5 * (selfenergy[is + ish] + selfenergy[is + jsh]);
                int nj = nsph(cj->ang), ni = nsph(ci->ang);
                for (int iao = 0; iao < ni; ++iao)
                    for (int jao = 0; jao < nj; ++jao) {
                        int ij = jao + nj * iao;
                        S_(jj + jao, ii + iao) += stmp[ij];
                        for (int k = 0; k < 3; ++k) D_(k, jj + jao, ii + iao) += dtmpi[3 * ij + k];
                        for (int k = 0; k < 6; ++k) Q_(k, jj + jao, ii + iao) += qtmpi[6 * ij + k];
                        H_(jj + jao, ii + iao) += stmp[ij] * hij;
                    }
            }
        }
    }
#undef S_
#undef H_
#undef D_
#undef Q_
    return 0;
}

int orc_get_hamiltonian(int nat, int nid, int msh, int nsh, int nao, const int *id,
                        const double *xyz, int ntr, const double *trans, const int *inl,
                        const int *nnl, const int *nlat, const int *nltr, const int *nsh_id,
                        const int *ish_at, const int *iao_sh, const orc_cgto *cgto, double intcut,
                        const double *hscale, const double *shpoly, const double *rad,
                        const double *selfenergy, double *overlap, double *dpint, double *qpint,
                        double *hamiltonian)
{
    return ham_impl(nat, nid, msh, nsh, nao, id, xyz, ntr, trans, inl, nnl, nlat, nltr, nsh_id,
                    ish_at, iao_sh, cgto, intcut, hscale, shpoly, rad, NULL, NULL, NULL,
                    selfenergy, overlap, dpint, qpint, hamiltonian);
}

/* Returns -2 if a species' shells are not s, p, d in this order. */
int orc_get_hamiltonian_diat(int nat, int nid, int msh, int nsh, int nao, const int *id,
                             const double *xyz, int ntr, const double *trans, const int *inl,
                             const int *nnl, const int *nlat, const int *nltr, const int *nsh_id,
                             const int *ish_at, const int *iao_sh, const orc_cgto *cgto,
                             double intcut, const double *hscale, const double *shpoly,
                             const double *rad, const double *ksig, const double *kpi,
                             const double *kdel, const double *selfenergy, double *overlap,
                             double *dpint, double *qpint, double *hamiltonian)
{
    if (!ksig || !kpi || !kdel) return -3;
    return ham_impl(nat, nid, msh, nsh, nao, id, xyz, ntr, trans, inl, nnl, nlat, nltr, nsh_id,
                    ish_at, iao_sh, cgto, intcut, hscale, shpoly, rad, ksig, kpi, kdel,
                    selfenergy, overlap, dpint, qpint, hamiltonian);
}

/* get_overlap_diat_lat (integrals.f90:473-582): the overlap and the overlap
 * with every atom-pair block (r2 > tiny) scaled in its diatomic frame, full
 * double loop over (iat, jat, itr) within the cutoff. */
int orc_get_overlap_diat_lat(int nat, int nid, int msh, int nao, const int *id,
                             const double *xyz, int ntr, const double *trans, double cutoff,
                             const int *nsh_id, const int *ish_at, const int *iao_sh,
                             const orc_cgto *cgto, double intcut, const double *ksig,
                             const double *kpi, const double *kdel, double *overlap,
                             double *overlap_diat)
{
    size_t n2 = (size_t)nao * nao;
    memset(overlap, 0, sizeof(double) * n2);
    memset(overlap_diat, 0, sizeof(double) * n2);
    for (int isp = 0; isp < nid; ++isp)
        for (int ish = 0; ish < nsh_id[isp]; ++ish)
            if (cgto[ish + msh * isp].ang > 2) return -1;
    if (!diat_shells_ok(nid, msh, nsh_id, cgto)) return -2;
    const double cutoff2 = cutoff * cutoff;
    const double tiny = 2.2250738585072014e-308;
#pragma omp parallel for schedule(dynamic)
    for (int iat = 0; iat < nat; ++iat) {
        double stmp[25], dtmp[75], qtmp[150], block[81];
        int izp = id[iat], is = ish_at[iat], nsi = nsh_id[izp];
        for (int jat = 0; jat < nat; ++jat) {
            int jzp = id[jat], js = ish_at[jat], nsj = nsh_id[jzp];
            for (int itr = 0; itr < ntr; ++itr) {
                double vec[3];
                for (int k = 0; k < 3; ++k)
                    vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
                double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
                if (r2 > cutoff2) continue;
                for (int k = 0; k < 81; ++k) block[k] = 0.0;
                for (int ish = 0; ish < nsi; ++ish) {
                    int ii = iao_sh[is + ish];
                    const orc_cgto *ci = &cgto[ish + msh * izp];
                    for (int jsh = 0; jsh < nsj; ++jsh) {
                        int jj = iao_sh[js + jsh];
                        const orc_cgto *cj = &cgto[jsh + msh * jzp];
                        orc_multipole_cgto(cj, ci, r2, vec, intcut, stmp, dtmp, qtmp);
                        int nj = nsph(cj->ang), ni = nsph(ci->ang);
                        for (int iao = 0; iao < ni; ++iao)
                            for (int jao = 0; jao < nj; ++jao) {
                                int ij = jao + nj * iao;
                                block[jsh * jsh + jao + 9 * (ish * ish + iao)] = stmp[ij];
                                overlap[(size_t)(jj + jao) + (size_t)nao * (ii + iao)] += stmp[ij];
                            }
                    }
                }
                if (r2 > tiny)
                    orc_diat_trafo(vec, nsj - 1, nsi - 1, ksig[izp + nid * jzp],
                                   kpi[izp + nid * jzp], kdel[izp + nid * jzp], block);
                for (int ish = 0; ish < nsi; ++ish) {
                    int ii = iao_sh[is + ish], ni = nsph(cgto[ish + msh * izp].ang);
                    for (int jsh = 0; jsh < nsj; ++jsh) {
                        int jj = iao_sh[js + jsh], nj = nsph(cgto[jsh + msh * jzp].ang);
                        for (int iao = 0; iao < ni; ++iao)
                            for (int jao = 0; jao < nj; ++jao)
                                overlap_diat[(size_t)(jj + jao) + (size_t)nao * (ii + iao)] +=
                                    block[jsh * jsh + jao + 9 * (ish * ish + iao)];
                    }
                }
            }
        }
    }
    return 0;
}

/* get_multipole_integrals_lat: integral/native/integrals.f90:1394-1465.
 * Full double loop over (iat, jat, itr) within the cutoff, every block
 * computed directly with the operator on its column atom (no half list, no
 * shift_operator).  overlap(nao,nao); dpint(3,nao,nao) and qpint(6,nao,nao)
 * may be NULL (get_overlap_lat, get_dipole_integrals_lat). */
int orc_get_multipole_integrals_lat(int nat, int msh, int nao, const int *id, const double *xyz,
                                    int ntr, const double *trans, double cutoff,
                                    const int *nsh_id, const int *ish_at, const int *iao_sh,
                                    const orc_cgto *cgto, double intcut, double *overlap,
                                    double *dpint, double *qpint)
{
    size_t n2 = (size_t)nao * nao;
    memset(overlap, 0, sizeof(double) * n2);
    if (dpint) memset(dpint, 0, sizeof(double) * 3 * n2);
    if (qpint) memset(qpint, 0, sizeof(double) * 6 * n2);
    for (int iat = 0; iat < nat; ++iat)
        for (int ish = 0; ish < nsh_id[id[iat]]; ++ish)
            if (cgto[ish + msh * id[iat]].ang > 2) return -1;
    const double cutoff2 = cutoff * cutoff;
#pragma omp parallel for schedule(dynamic)
    for (int iat = 0; iat < nat; ++iat) {
        double stmp[25], dtmp[75], qtmp[150];
        int izp = id[iat], is = ish_at[iat];
        for (int jat = 0; jat < nat; ++jat) {
            int jzp = id[jat], js = ish_at[jat];
            for (int itr = 0; itr < ntr; ++itr) {
                double vec[3];
                for (int k = 0; k < 3; ++k)
                    vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
                double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
                if (r2 > cutoff2) continue;
                for (int ish = 0; ish < nsh_id[izp]; ++ish) {
                    int ii = iao_sh[is + ish];
                    const orc_cgto *ci = &cgto[ish + msh * izp];
                    for (int jsh = 0; jsh < nsh_id[jzp]; ++jsh) {
                        int jj = iao_sh[js + jsh];
                        const orc_cgto *cj = &cgto[jsh + msh * jzp];
                        orc_multipole_cgto(cj, ci, r2, vec, intcut, stmp, dtmp, qtmp);
                        int nj = nsph(cj->ang), ni = nsph(ci->ang);
                        /* column ii + iao belongs to iat: this thread owns it */
                        for (int iao = 0; iao < ni; ++iao)
                            for (int jao = 0; jao < nj; ++jao) {
                                int ij = jao + nj * iao;
                                size_t e = (size_t)(jj + jao) + (size_t)nao * (ii + iao);
                                overlap[e] += stmp[ij];
                                if (dpint)
                                    for (int k = 0; k < 3; ++k) dpint[k + 3 * e] += dtmp[3 * ij + k];
                                if (qpint)
                                    for (int k = 0; k < 6; ++k) qpint[k + 6 * e] += qtmp[6 * ij + k];
                            }
                    }
                }
            }
        }
    }
    return 0;
}

/* get_hamiltonian_gradient: xtb/h0.f90:432-666.
 * pmat(nao,nao,nspin), xmat(nao,nao,nspin), vao(nao,nspin), vdp(3,nat,nspin),
 * vqp(6,nat,nspin); dEdcn(nat), gradient(3,nat), sigma(3,3) are accumulated. */
int orc_get_hamiltonian_gradient(int nat, int nid, int msh, int nsh, int nao, const int *id,
                                 const double *xyz, int ntr, const double *trans, const int *inl,
                                 const int *nnl, const int *nlat, const int *nltr,
                                 const int *nsh_id, const int *ish_at, const int *iao_sh,
                                 const orc_cgto *cgto, double intcut, const double *hscale,
                                 const double *shpoly, const double *rad,
                                 const double *selfenergy, const double *dsedcn, int nspin,
                                 const double *vao, const double *vdp, const double *vqp,
                                 const double *pmat, const double *xmat, double *dEdcn,
                                 double *gradient, double *sigma)
{
    orc_system sys = {nat, nid, msh, nsh, nao, id, xyz, nsh_id, ish_at, iao_sh, cgto,
                      intcut, hscale, shpoly, rad};
    size_t n2 = (size_t)nao * nao;
    const double eps = 2.220446049250313e-16;
    for (int isp = 0; isp < nid; ++isp)
        for (int ish = 0; ish < nsh_id[isp]; ++ish)
            if (cgto[ish + msh * isp].ang > 2) return -1;

#pragma omp parallel
    {
        double *g_loc = calloc((size_t)3 * nat, sizeof(double));
        double *cn_loc = calloc((size_t)nat, sizeof(double));
        double sig_loc[9] = {0};
        double stmp[25], dtmp[75], qtmp[150], dstmp[75];
        double ddtmpi[225], dqtmpi[450], ddtmpj[225], dqtmpj[450];
#pragma omp for schedule(dynamic)
        for (int iat = 0; iat < nat; ++iat) {
            int izp = id[iat], is = ish_at[iat], nsi = nsh_id[izp];
            for (int img = 0; img < nnl[iat]; ++img) {
                int jat = nlat[inl[iat] + img], itr = nltr[inl[iat] + img];
                int jzp = id[jat], js = ish_at[jat], nsj = nsh_id[jzp];
                double vec[3];
                for (int k = 0; k < 3; ++k)
                    vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
                double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
                if (r2 <= eps) continue;
                double rr = sqrt(sqrt(r2) / (rad[jzp] + rad[izp]));
                double dG[3] = {0, 0, 0}, dcni = 0.0, dcnj = 0.0;
                for (int ish = 0; ish < nsi; ++ish) {
                    int ii = iao_sh[is + ish];
                    const orc_cgto *ci = &cgto[ish + msh * izp];
                    double shpolyi = 1.0 + shpoly[ish + msh * izp] * rr;
                    for (int jsh = 0; jsh < nsj; ++jsh) {
                        int jj = iao_sh[js + jsh];
                        const orc_cgto *cj = &cgto[jsh + msh * jzp];
                        orc_multipole_grad_cgto(cj, ci, r2, vec, intcut, stmp, dtmp, qtmp, dstmp,
                                                ddtmpj, dqtmpj, ddtmpi, dqtmpi);
                        double shpolyj = 1.0 + shpoly[jsh + msh * jzp] * rr;
                        double shp = shpolyi * shpolyj;
                        double dshpoly = (shpolyi * shpoly[jsh + msh * jzp] +
                                          shpolyj * shpoly[ish + msh * izp]) *
                                         0.5 * rr / r2;
                        double dsv[3];
                        for (int k = 0; k < 3; ++k) dsv[k] = dshpoly / shp * vec[k];
                        double hs = HSCALE(&sys, jsh, ish, jzp, izp) * shp;
                        double hij = 0.5 * (selfenergy[is + ish] + selfenergy[js + jsh]) * hs;
                        double dhdcni = dsedcn[is + ish] * hs;
                        double dhdcnj = dsedcn[js + jsh] * hs;
                        int nj = nsph(cj->ang), ni = nsph(ci->ang);
                        for (int iao = 0; iao < ni; ++iao)
                            for (int jao = 0; jao < nj; ++jao) {
                                int ij = jao + nj * iao;
                                size_t pidx = (size_t)(jj + jao) + (size_t)nao * (ii + iao);
                                for (int spin = 0; spin < nspin; ++spin) {
                                    double pij = pmat[pidx + n2 * spin];
                                    double sval = -pij * (vao[jj + jao + (size_t)nao * spin] +
                                                          vao[ii + iao + (size_t)nao * spin]);
                                    for (int k = 0; k < 3; ++k) dG[k] += sval * dstmp[3 * ij + k];
                                }
                                double pij = pmat[pidx];
                                double hpij = pij * hij;
                                double sval = 2 * hpij - 2 * xmat[pidx];
                                for (int a = 0; a < 3; ++a) {
                                    double t = sval * dstmp[3 * ij + a] +
                                               2 * hpij * stmp[ij] * dsv[a];
                                    double md = 0.0;
                                    for (int c = 0; c < 3; ++c) {
                                        md += ddtmpi[9 * ij + a + 3 * c] * vdp[c + 3 * iat];
                                        md += ddtmpj[9 * ij + a + 3 * c] * vdp[c + 3 * jat];
                                    }
                                    for (int c = 0; c < 6; ++c) {
                                        md += dqtmpi[18 * ij + a + 3 * c] * vqp[c + 6 * iat];
                                        md += dqtmpj[18 * ij + a + 3 * c] * vqp[c + 6 * jat];
                                    }
                                    dG[a] += t - pij * md;
                                }
                                dcni += dhdcni * pmat[pidx] * stmp[ij];
                                dcnj += dhdcnj * pmat[pidx] * stmp[ij];
                            }
                    }
                }
                if (iat == jat) {
                    cn_loc[iat] += 0.5 * (dcni + dcnj);
                } else {
                    cn_loc[iat] += dcni;
                    cn_loc[jat] += dcnj;
                }
                for (int k = 0; k < 3; ++k) {
                    g_loc[3 * iat + k] += dG[k];
                    g_loc[3 * jat + k] -= dG[k];
                }
                double f = (iat == jat) ? 0.25 : 0.5;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        sig_loc[a + 3 * b] += f * (vec[a] * dG[b] + dG[a] * vec[b]);
            }
            /* onsite: diagonal of P only (h0.f90:654-663) */
            for (int ish = 0; ish < nsi; ++ish) {
                int ii = iao_sh[is + ish];
                double dhdcni = dsedcn[is + ish], dcni = 0.0;
                int ni = nsph(cgto[ish + msh * izp].ang);
                for (int iao = 0; iao < ni; ++iao)
                    dcni += dhdcni * pmat[(size_t)(ii + iao) + (size_t)nao * (ii + iao)];
                cn_loc[iat] += dcni;
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < 3 * nat; ++k) gradient[k] += g_loc[k];
            for (int k = 0; k < nat; ++k) dEdcn[k] += cn_loc[k];
            for (int k = 0; k < 9; ++k) sigma[k] += sig_loc[k];
        }
        free(g_loc);
        free(cn_loc);
    }
    return 0;
}

/* Adjoint of get_hamiltonian (xtb/h0.f90:220-429, differentiated the way
 * get_hamiltonian_gradient differentiates it, h0.f90:432-666).  Given the
 * adjoints sbar(nao,nao), hbar(nao,nao), dbar(3,nao,nao), qbar(6,nao,nao) of
 * the four outputs, E = <sbar,S> + <hbar,H> + <dbar,D> + <qbar,Q> and this
 * accumulates dE/dxyz into gradient(3,nat), dE/dselfenergy into dEdse(nsh)
 * and the strain derivative into sigma(3,3).  Every half-list pair contributes through the blocks
 * get_hamiltonian writes for it: block A (rows of jat, columns of iat,
 * operator on iat) and, for iat /= jat, its mirror (operator on jat). */
int orc_get_hamiltonian_adjoint(int nat, int nid, int msh, int nsh, int nao, const int *id,
                                const double *xyz, int ntr, const double *trans, const int *inl,
                                const int *nnl, const int *nlat, const int *nltr,
                                const int *nsh_id, const int *ish_at, const int *iao_sh,
                                const orc_cgto *cgto, double intcut, const double *hscale,
                                const double *shpoly, const double *rad,
                                const double *selfenergy, const double *sbar, const double *hbar,
                                const double *dbar, const double *qbar, double *gradient,
                                double *dEdse, double *sigma)
{
    orc_system sys = {nat, nid, msh, nsh, nao, id, xyz, nsh_id, ish_at, iao_sh, cgto,
                      intcut, hscale, shpoly, rad};
    const double eps = 2.220446049250313e-16;
    for (int isp = 0; isp < nid; ++isp)
        for (int ish = 0; ish < nsh_id[isp]; ++ish)
            if (cgto[ish + msh * isp].ang > 2) return -1;
#define E_(r, c) ((size_t)(r) + (size_t)nao * (c))

#pragma omp parallel
    {
        double *g_loc = calloc((size_t)3 * nat, sizeof(double));
        double *se_loc = calloc((size_t)nsh, sizeof(double));
        double sig_loc[9] = {0};
        double stmp[25], dtmp[75], qtmp[150], dstmp[75];
        double ddtmpi[225], dqtmpi[450], ddtmpj[225], dqtmpj[450];
        double stmp0[25], dtmp0[75], qtmp0[150];
#pragma omp for schedule(dynamic)
        for (int iat = 0; iat < nat; ++iat) {
            int izp = id[iat], is = ish_at[iat], nsi = nsh_id[izp];
            for (int img = 0; img < nnl[iat]; ++img) {
                int jat = nlat[inl[iat] + img], itr = nltr[inl[iat] + img];
                int jzp = id[jat], js = ish_at[jat], nsj = nsh_id[jzp];
                double vec[3];
                for (int k = 0; k < 3; ++k)
                    vec[k] = xyz[3 * iat + k] - xyz[3 * jat + k] - trans[3 * itr + k];
                double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
                if (r2 <= eps) continue;
                double rr = sqrt(sqrt(r2) / (rad[jzp] + rad[izp]));
                int mirror = iat != jat;
                double dG[3] = {0, 0, 0};
                for (int ish = 0; ish < nsi; ++ish) {
                    int ii = iao_sh[is + ish];
                    const orc_cgto *ci = &cgto[ish + msh * izp];
                    double shpolyi = 1.0 + shpoly[ish + msh * izp] * rr;
                    for (int jsh = 0; jsh < nsj; ++jsh) {
                        int jj = iao_sh[js + jsh];
                        const orc_cgto *cj = &cgto[jsh + msh * jzp];
                        orc_multipole_grad_cgto(cj, ci, r2, vec, intcut, stmp, dtmp, qtmp, dstmp,
                                                ddtmpj, dqtmpj, ddtmpi, dqtmpi);
                        double shpolyj = 1.0 + shpoly[jsh + msh * jzp] * rr;
                        double shp = shpolyi * shpolyj;
                        double dshpoly = (shpolyi * shpoly[jsh + msh * jzp] +
                                          shpolyj * shpoly[ish + msh * izp]) *
                                         0.5 * rr / r2;
                        double hs = HSCALE(&sys, jsh, ish, jzp, izp) * shp;
                        double hij = 0.5 * (selfenergy[is + ish] + selfenergy[js + jsh]) * hs;
                        int nj = nsph(cj->ang), ni = nsph(ci->ang);
                        double hsum = 0.0;  /* sum of hbar * S over the block (and mirror) */
                        for (int iao = 0; iao < ni; ++iao)
                            for (int jao = 0; jao < nj; ++jao) {
                                int ij = jao + nj * iao;
                                size_t ea = E_(jj + jao, ii + iao), em = E_(ii + iao, jj + jao);
                                double sb = sbar[ea], hb = hbar[ea];
                                if (mirror) {
                                    sb += sbar[em];
                                    hb += hbar[em];
                                }
                                hsum += hb * stmp[ij];
                                for (int a = 0; a < 3; ++a) {
                                    /* dS/dR_i and d(S shpoly)/dR_i */
                                    double t = (sb + hb * hij) * dstmp[3 * ij + a] +
                                               hb * hij * stmp[ij] * dshpoly / shp * vec[a];
                                    for (int c = 0; c < 3; ++c) {
                                        t += dbar[c + 3 * ea] * ddtmpi[9 * ij + a + 3 * c];
                                        if (mirror)
                                            t += dbar[c + 3 * em] * ddtmpj[9 * ij + a + 3 * c];
                                    }
                                    for (int c = 0; c < 6; ++c) {
                                        t += qbar[c + 6 * ea] * dqtmpi[18 * ij + a + 3 * c];
                                        if (mirror)
                                            t += qbar[c + 6 * em] * dqtmpj[18 * ij + a + 3 * c];
                                    }
                                    dG[a] += t;
                                }
                            }
                        se_loc[is + ish] += 0.5 * hs * hsum;
                        se_loc[js + jsh] += 0.5 * hs * hsum;
                    }
                }
                for (int k = 0; k < 3; ++k) {
                    g_loc[3 * iat + k] += dG[k];
                    g_loc[3 * jat + k] -= dG[k];
                }
                /* every block is counted once here (a self-image pair iat == jat has no
                 * mirror), so each pair adds the symmetrised vec x dE/dvec */
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        sig_loc[a + 3 * b] += 0.5 * (vec[a] * dG[b] + dG[a] * vec[b]);
            }
            /* onsite block: H = S (se_i + se_j)/2, no geometry dependence */
            double vec0[3] = {0.0, 0.0, 0.0};
            for (int ish = 0; ish < nsi; ++ish) {
                int ii = iao_sh[is + ish];
                const orc_cgto *ci = &cgto[ish + msh * izp];
                for (int jsh = 0; jsh < nsi; ++jsh) {
                    int jj = iao_sh[is + jsh];
                    const orc_cgto *cj = &cgto[jsh + msh * izp];
                    orc_multipole_cgto(cj, ci, 0.0, vec0, intcut, stmp0, dtmp0, qtmp0);
                    int nj = nsph(cj->ang), ni = nsph(ci->ang);
                    double hsum = 0.0;
                    for (int iao = 0; iao < ni; ++iao)
                        for (int jao = 0; jao < nj; ++jao)
                            hsum += hbar[E_(jj + jao, ii + iao)] * stmp0[jao + nj * iao];
                    se_loc[is + ish] += 0.5 * hsum;
                    se_loc[is + jsh] += 0.5 * hsum;
                }
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < 3 * nat; ++k) gradient[k] += g_loc[k];
            for (int k = 0; k < nsh; ++k) dEdse[k] += se_loc[k];
            for (int k = 0; k < 9; ++k) sigma[k] += sig_loc[k];
        }
        free(g_loc);
        free(se_loc);
    }
#undef E_
    return 0;
}

/* Double-exponential coordination number of GFN2 (mctc-lib v0.5 ncoord
 * "dexp", created at xtb/gfn2.f90:640 and evaluated at
 * xtb/singlepoint.f90:224):
 *   cn(i) = sum_{j, T; 0 < r_ijT <= cutoff} f(r_ijT),
 *   f(r) = 1/(1+exp(-ka (r0/r - 1))) * 1/(1+exp(-kb ((r0+rshift)/r - 1))),
 *   r0 = rcov(id(i)) + rcov(id(j)),
 * with its derivatives dcndr(3, nat, nat) (d cn(i) / d xyz(:, j), stored
 * dcndr[c + 3*(j + nat*i)]) and dcndL(3, 3, nat) (strain derivative), the
 * arrays singlepoint.f90:375-378 contracts with dEdcn.  dcndr / dcndL may
 * be NULL.  mctc-lib is not part of the reference tree: its formula and the
 * GFN2 constants (ka = 10, kb = 20, rshift = 2 bohr, cutoff 25 bohr) are
 * restated from its published source; parity with it is unpinned. */
static void dexp_count(double ka, double kb, double rshift, double r, double r0, double *f,
                       double *df)
{
    double e1 = exp(-ka * (r0 / r - 1.0)), e2 = exp(-kb * ((r0 + rshift) / r - 1.0));
    double c1 = 1.0 / (1.0 + e1), c2 = 1.0 / (1.0 + e2);
    /* d/dr 1/(1+exp(-k (a/r - 1))) = -k a e / (r^2 (1+e)^2) */
    double d1 = -ka * r0 * e1 / (r * r * (1.0 + e1) * (1.0 + e1));
    double d2 = -kb * (r0 + rshift) * e2 / (r * r * (1.0 + e2) * (1.0 + e2));
    *f = c1 * c2;
    *df = d1 * c2 + c1 * d2;
}

void orc_ncoord_dexp(int nat, const int *id, const double *xyz, int ntr, const double *trans,
                     const double *rcov, double ka, double kb, double rshift, double cutoff,
                     double *cn, double *dcndr, double *dcndL)
{
    const double eps = 2.220446049250313e-16, cut2 = cutoff * cutoff;
    memset(cn, 0, sizeof(double) * nat);
    if (dcndr) memset(dcndr, 0, sizeof(double) * 3 * (size_t)nat * nat);
    if (dcndL) memset(dcndL, 0, sizeof(double) * 9 * (size_t)nat);
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j)
            for (int t = 0; t < ntr; ++t) {
                double v[3], r2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    v[k] = xyz[3 * i + k] - xyz[3 * j + k] - trans[3 * t + k];
                    r2 += v[k] * v[k];
                }
                if (r2 > cut2 || r2 < eps) continue;
                double r = sqrt(r2), f, df;
                dexp_count(ka, kb, rshift, r, rcov[id[i]] + rcov[id[j]], &f, &df);
                cn[i] += f;
                if (i != j) cn[j] += f;
                double g[3];
                for (int k = 0; k < 3; ++k) g[k] = df * v[k] / r;
                if (dcndr) {
                    for (int k = 0; k < 3; ++k) {
                        dcndr[k + 3 * (i + (size_t)nat * i)] += g[k];
                        dcndr[k + 3 * (j + (size_t)nat * j)] -= g[k];
                        dcndr[k + 3 * (j + (size_t)nat * i)] -= g[k];
                        dcndr[k + 3 * (i + (size_t)nat * j)] += g[k];
                    }
                }
                if (dcndL) {
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            dcndL[a + 3 * b + 9 * i] += g[a] * v[b];
                            if (i != j) dcndL[a + 3 * b + 9 * j] += g[a] * v[b];
                        }
                }
            }
}

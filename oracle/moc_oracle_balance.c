/*
 * moc_oracle_balance.c -- ORACLE (test infrastructure only): FP64 restatement
 * of MOCASSIN's per-cell ionisation / thermal balance, update_mod.f90:
 *   updateCell :15-272, iterateT :277-482, thermBalance :887-1215,
 *   forLines :1221-1262, ionBalance :1575-1992, calcAlpha :2298-2413,
 *   RRbadnell :2415, DRbadnell :2435, radRecFit :3094-3202,
 *   dielectronic :3204-3305, collisionalIonisation :3552-3582,
 *   photoionisation :3584-3832, compton :3834-3897, getPhotoRates :3971-4040;
 * with equilibrium / luSlv (emission_mod.f90:2505-3106) and spline / splint
 * (interpolation_mod.f90:207-340) for the forbidden-line cooling.
 * Written one cell at a time, in the order and with the loops of the
 * reference, everything in FP64.  Deliberate differences (also in the GPU
 * path): an element that is switched off has no He terms in the H-only
 * convergence test (the reference reads ionDen(elementXref(2)=0, ...)), and a
 * shell whose threshold lies below the first mesh point starts at bin 1.
 *
 * Pinned by tests/test_balance.py: calcAlpha for H and He II (known values,
 * hydrogenic scaling, T slope), the n-level solve (closed-form two-level
 * atom, 5-level linear solve, low-density limit), and the H photoionisation
 * equilibrium of ionBalance.  Parity unpinned (no independent check here):
 * the heavy-element ionBalance chain with Auger spreading and charge
 * exchange, the recombination fits for ions with more than 11 electrons
 * (radRecFit, dielectronic, Aldrovandi), the thermBalance cooling fits and
 * the iterateT root-find path.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "moc_oracle.h"
#include "moc_oracle_balance.h"

#define NEL 30
#define HCRYD 2.1799153e-11   /* update_mod.f90:64 */
#define RYD2ERG 2.1799e-11     /* constants_mod.f90:20 */
#define KBOLTZ 1.380658e-16    /* constants_mod.f90:17 */
#define XMAX 0.99999           /* constants_mod.f90:55 */

/* charge-exchange fits (a[1e-9 cm^3/s], b, c, d) and deltaE/k, as listed in
 * photoionisation / ionBalance (update_mod.f90:3648-3714, 1624-1690) */
typedef struct { int z, ion; double a, b, c, d; } chex_row;
static const chex_row CHEX[] = {
    {2, 1, 7.47e-6, 2.06, 9.93, -3.89}, {2, 2, 1.e-5, 0., 0., 0.},
    {3, 2, 1.26, 0.96, 3.02, -0.65}, {3, 3, 1.e-5, 0., 0., 0.},
    {4, 2, 1.e-5, 0., 0., 0.}, {4, 3, 1.e-5, 0., 0., 0.}, {4, 4, 5.17, 0.82, -.69, -1.12},
    {5, 2, 2.e-2, 0., 0., 0.}, {5, 3, 1.e-5, 0., 0., 0.}, {5, 4, 5.27e-1, 0.76, -0.63, -1.17},
    {6, 1, 1.76e-9, 8.33, 4278.78, -6.41}, {6, 2, 1.67e-4, 2.79, 304.72, -4.07},
    {6, 3, 3.25, 0.21, 0.19, -3.29}, {6, 4, 332.46, -0.11, -0.995, -1.58e-3},
    {7, 1, 1.01e-3, -0.29, -0.92, -8.38}, {7, 2, 3.05e-1, 0.60, 2.65, -0.93},
    {7, 3, 4.54, 0.57, -0.65, -0.89}, {7, 4, 3.28, 0.52, -0.52, -0.19},
    {8, 1, 1.04, 3.15e-2, -0.61, -9.73}, {8, 2, 1.04, 0.27, 2.02, -5.92},
    {8, 3, 3.98, 0.26, 0.56, -2.62}, {8, 4, 2.52e-1, 0.63, 2.08, -4.16},
    {9, 2, 1.e-5, 0., 0., 0.}, {9, 3, 9.86, 0.29, -0.21, -1.15}, {9, 4, 7.15e-1, 1.21, -0.70, -0.85},
    {10, 2, 1.e-5, 0., 0., 0.}, {10, 3, 14.73, 4.52e-2, -0.84, -0.31}, {10, 4, 6.47, 0.54, 3.59, -5.22},
    {11, 2, 1.e-5, 0., 0., 0.}, {11, 3, 1.33, 1.15, 1.20, -0.32}, {11, 4, 1.01e-1, 1.34, 10.05, -6.41},
    {12, 2, 8.58e-5, 2.49e-3, 2.93e-2, -4.33}, {12, 3, 6.49, 0.53, 2.82, -7.63},
    {12, 4, 6.36, 0.55, 3.86, -5.19}, {13, 2, 1.e-5, 0., 0., 0.},
    {13, 3, 7.11e-5, 4.12, 1.72e4, -22.24}, {13, 4, 7.52e-1, 0.77, 6.24, -5.67},
    {14, 2, 1.23, 0.24, 3.17, 4.18e-3}, {14, 3, 4.900e-1, -8.74e-2, -0.36, -0.79},
    {14, 4, 7.58, 0.37, 1.06, -4.09}, {16, 1, 3.82e-7, 11.10, 2.57e4, -8.22},
    {16, 2, 1.e-5, 0., 0., 0.}, {16, 3, 2.29, 4.02e-2, 1.59, -6.06}, {16, 4, 6.44, 0.13, 2.69, -5.69},
    {18, 2, 1.e-5, 0., 0., 0.}, {18, 3, 4.57, 0.27, -0.18, -1.57}, {18, 4, 6.37, 2.12, 10.21, -6.22},
    /* the reference's "Ca" rows are written into Ar's slots and overwrite them */
    {18, 3, 3.17e-2, 2.12, 12.06, -0.40}, {18, 4, 2.68, 0.69, -0.68, -4.47},
    {26, 2, 1.26, 7.72e-2, -0.41, -7.31}, {26, 3, 3.42, 0.51, -2.06, -8.99},
};

/* outer shell and statistical weights (getOuterShell, hydro_mod.f90:89-182) */
static void outer_shell(int z, int nel, int* out, int* g0, int* g1) {
    static const int ss[30] = {1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6,
                               6, 6, 6, 6, 6, 6, 6, 6, 7, 7};
    static const int gl[30] = {2, 1, 2, 1, 6, 9, 4, 9, 6, 1, 2, 1, 6, 9, 4, 9, 6, 1, 2, 1,
                               10, 21, 28, 7, 6, 25, 28, 21, 2, 1};
    static const int glh[12] = {10, 21, 28, 25, 6, 25, 28, 21, 10, 1, 2, 1}; /* 19..30 */
    *out = ss[nel - 1];
    if (z == nel && z > 18) *out = 7;
    if (z == nel + 1 && (z == 20 || z == 21 || z == 22 || z == 25 || z == 26)) *out = 7;
    *g0 = gl[nel - 1];
    *g1 = nel == 1 ? 1 : gl[nel - 2];
    if (z > 20 && nel >= 19 && z - nel >= 1) {
        *g0 = glh[nel - 19];
        *g1 = nel == 19 ? gl[nel - 2] : glh[nel - 20];
    }
    if (z == nel) {
        if (z > 20 && nel >= 21) *g1 = glh[nel - 20];
        if (z == 21) *g1 = 15;
        if (z == 22) *g1 = 28;
        if (z == 25) *g1 = 7;
        if (z == 26) *g1 = 30;
    }
    if (z - nel == 1) {
        if (z == 21) *g0 = 15;
        if (z == 22) *g0 = 28;
        if (z == 25) *g0 = 7;
        if (z == 26) *g0 = 30;
    }
}

/* natural cubic spline (interpolation_mod.f90:207-283) */
static void spline_nat(const double* x, const double* y, int n, double* y2) {
    double u[64], sig, p;
    int i, k;
    y2[0] = 0.;
    u[0] = 0.;
    for (i = 1; i < n - 1; i++) {
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        p = sig * y2[i - 1] + 2.;
        y2[i] = (sig - 1.) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6. * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.;
    for (k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

static double splint_(const double* xa, const double* ya, const double* y2, int n, double x) {
    int klo = 0, khi = n - 1, k;
    double h, a, b;
    if (x <= xa[0]) return ya[0];
    if (x >= xa[n - 1]) return ya[n - 1];
    while (khi - klo > 1) {
        k = (khi + klo) / 2;
        if (xa[k] > x) khi = k;
        else klo = k;
    }
    h = xa[khi] - xa[klo];
    a = (xa[khi] - x) / h;
    b = (x - xa[klo]) / h;
    return a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2[klo] + (b * b * b - b) * y2[khi]) * (h * h) / 6.;
}

/* equilibrium(rec=.false.) (emission_mod.f90:2505-3028) for one ion, summed
 * as forLines/thermBalance do: every line emissivity A_ji dE_ij n_j, scaled
 * by abundance * ionDen * 1.9865e-16, that exceeds 1e-35 [erg/s per n_H] */
static double nlevel_lines(const oracle_balance_tables* t, int atom, double Te, double Ne,
                           double ionDenUp, int rec, double scale, double unit, double* lineEm,
                           int ld) {
    const int64_t* m = t->atom_meta + 7 * atom;
    const int nfile = (int)m[0], ntemp = (int)m[1];
    const double* lt = t->atom_logt + m[2];
    const double* g = t->atom_g + m[3];
    const double* e = t->atom_e + m[3];
    const double* A = t->atom_A + m[4];
    const float* q = t->atom_qom + m[5];
    const int iup2p = (int)m[6];
    const double atp = t->atom_atp[atom];
    const int n = nfile < t->maxLev ? nfile : t->maxLev;
    double qeff[32][32], x[32][32], y[32], qq[64], qq2[64], nn[32];
    const double sqrTe = sqrt(Te), log10Te = log10(Te), constant = 1.4388463;
    double sumN = 0., cool = 0.;
    int i, j, k, it;
    memset(qeff, 0, sizeof(qeff));
    memset(x, 0, sizeof(x));
    for (i = 2; i <= n; i++)
        for (j = i; j <= n; j++) {
            double qom, delTeK, expFac;
            for (it = 0; it < ntemp; it++)
                qq[it] = (double)q[((size_t)it * nfile + (i - 2)) * nfile + (j - 1)];
            if (ntemp == 1) qom = qq[0];
            else if (ntemp == 2) qom = qq[0] + (qq[1] - qq[0]) / (lt[1] - lt[0]) * (log10Te - lt[0]);
            else {
                spline_nat(lt, qq, ntemp, qq2);
                qom = splint_(lt, qq, qq2, ntemp, log10Te);
            }
            if (e[i - 2] <= e[j - 1]) {
                delTeK = (e[i - 2] - e[j - 1]) * constant;
                expFac = exp(delTeK / Te);
                qeff[i - 1][j] = 8.63e-6 * qom * expFac / (g[i - 2] * sqrTe);
                qeff[j][i - 1] = 8.63e-6 * qom / (g[j - 1] * sqrTe);
            } else {
                delTeK = (e[j - 1] - e[i - 2]) * constant;
                expFac = exp(delTeK / Te);
                qeff[j][i - 1] = 8.63e-6 * qom * expFac / (g[i - 2] * sqrTe);
                qeff[i - 1][j] = 8.63e-6 * qom / (g[j - 1] * sqrTe);
            }
        }
#define AA(jj, ii) A[((size_t)((jj)-1)) * nfile + ((ii)-1)]
    for (j = 1; j <= n; j++) x[1][j] = 1.;
    y[1] = 1.;
    for (i = 2; i <= n; i++) {
        for (j = 1; j <= n; j++) {
            if (j != i) {
                x[i][j] += Ne * qeff[j][i];
                x[i][i] -= Ne * qeff[i][j];
                if (j > i) x[i][j] += AA(j, i);
                else x[i][i] -= AA(i, j);
            }
        }
        y[i] = 0.;
    }
    /* recombination feed of the levels (:2828-2877, :2961) */
    if (rec && t->atom_rec_kind && t->atom_rec_kind[atom] != 0) {
        const int kind = t->atom_rec_kind[atom];
        const double* rd = t->atom_rec + 11 * m[3];
        const int32_t* rv = t->atom_rec_valid + m[3];
        for (i = 2; i <= n; i++) {
            double al = 0.;
            const double* r = rd + 11 * (i - 1);
            if (rv[i - 1]) {
                if (kind == 1) {
                    double z = r[1], tz = Te * 1.e-4 / (z * z), t4 = Te * 1.e-4;
                    al = 1.e-13 * r[0] * z * r[2] * pow(tz, r[3]) / (1. + r[4] * pow(tz, r[5])) +
                         1.0e-12 * r[0] * (r[6] / t4 + r[7] + r[8] * t4 + r[9] * t4 * t4) * pow(t4, -1.5) *
                             exp(-r[10] * 1.e4 / Te);
                } else {
                    al = pow(10., r[0] + r[1] * log10Te + r[2] * log10Te * log10Te +
                                      r[3] * log10Te * log10Te * log10Te);
                }
            }
            y[i] = -Ne * ionDenUp * al;
        }
    }
    /* luSlv: Gaussian elimination without pivoting (:3031-3106) */
    for (i = 1; i <= n - 1; i++)
        for (k = i + 1; k <= n; k++) {
            double f = x[k][i] / x[i][i];
            for (j = i + 1; j <= n; j++) x[k][j] -= x[i][j] * f;
        }
    for (i = 1; i <= n - 1; i++)
        for (j = i + 1; j <= n; j++) y[j] -= y[i] * x[j][i] / x[i][i];
    y[n] = y[n] / x[n][n];
    for (i = 1; i <= n - 1; i++) {
        k = n - i;
        for (j = k + 1; j <= n; j++) y[k] -= y[j] * x[k][j];
        y[k] = y[k] / x[k][k];
    }
    for (i = 1; i <= n; i++) sumN += y[i];
    for (i = 1; i <= n; i++) nn[i] = y[i] / sumN;
    for (i = 1; i <= n - 1; i++)
        for (j = i + 1; j <= n; j++) {
            double a = AA(j, i);
            if (a != 0.) {
                double em;
                if (j == iup2p && i == 1 && atp > 0.) a -= atp;
                em = a * (e[j - 1] - e[i - 1]) * nn[j] * scale * unit;
                if (em > 1.e-35) cool += em;
                if (lineEm) lineEm[(i - 1) * ld + (j - 1)] = em; /* fLineEm(i, j) (:2999) */
            }
        }
#undef AA
    return cool;
}

static double nlevel_cooling(const oracle_balance_tables* t, int atom, double Te, double Ne,
                             double scale) {
    return nlevel_lines(t, atom, Te, Ne, 0., 0, scale, 1.9865e-16, NULL, 0);
}

double oracle_nlevel_lines(const oracle_balance_tables* t, int atom, double Te, double Ne,
                           double ionDenUp, int rec, double scale, double unit) {
    return nlevel_lines(t, atom, Te, Ne, ionDenUp, rec, scale, unit, NULL, 0);
}

double oracle_nlevel_line_em(const oracle_balance_tables* t, int atom, double Te, double Ne,
                             double ionDenUp, int rec, double scale, double unit, double* lineEm,
                             int ld) {
    return nlevel_lines(t, atom, Te, Ne, ionDenUp, rec, scale, unit, lineEm, ld);
}

/* ---- per-cell state ------------------------------------------------------ */
typedef struct {
    double Te, Ne, H;
    double X[31][32];      /* ionDen by (Z, stage) */
    double X0[31][32];     /* ionDen at the start of the cell update (grid%ionDen) */
    double ab[31];         /* abundance by Z */
    int on[31];
    double alpha[31][32];  /* alphaTot */
    double cI1[31][32], cI2[31][32]; /* collIon(1|2, Z, ion) */
    double pI1[11][31][32], pI2[11][31][32]; /* photoIon(1|2, shell, Z, ion) */
    double cHeat, cCool, cRecoil;   /* Compton */
    int nIterX;
} ocell;

#define RRB(p, z, n) t->rrb[((size_t)(p)*31 + (z)) * 31 + (n)]
#define DRC(i, z, n) t->drb_c[((size_t)(i)*31 + (z)) * 31 + (n)]
#define DRE(i, z, n) t->drb_e[((size_t)(i)*31 + (z)) * 31 + (n)]
#define RREC(i, z, n) t->rrec[((size_t)(i)*31 + (z)) * 31 + (n)]
#define RNEW(i, z, n) t->rnew[((size_t)(i)*31 + (z)) * 31 + (n)]
#define CF(i, z, n) t->cion_cf[((size_t)(i)*31 + (z)) * 31 + (n)]
#define CTH(z, ion) t->cion_th[(size_t)(z)*32 + (ion)]
#define NAUG(z, ion, s) t->nauger[((size_t)(z)*32 + (ion)) * 11 + (s)]
#define AUG(z, ion, s, k) t->auger[(((size_t)(z)*32 + (ion)) * 11 + (s)) * 11 + (k)]

static int min_i(int a, int b) { return a < b ? a : b; }
static int max_i(int a, int b) { return a > b ? a : b; }

/* calcAlpha with lgBadnell = .true., lgNahar = .false. (:2298-2362) */
static void calc_alpha(const oracle_balance_tables* t, ocell* c) {
    const double T = c->Te;
    static const double fe[3][10] = {
        {4.33e-10, 3.91e-10, 3.49e-10, 3.16e-10, 2.96e-10, 2.59e-10, 2.24e-10, 1.91e-10, 1.68e-10, 1.46e-10},
        {0.531, 0.523, 0.521, 0.534, 0.557, 0.567, 0.579, 0.601, 0.602, 0.597},
        {5.77e-02, 6.15e-02, 6.22e-02, 6.02e-02, 5.79e-02, 5.65e-02, 5.49e-02, 5.10e-02, 5.07e-02, 5.22e-02}};
    double diRec[31][32], aldro[31][32], tt;
    int z, ion, i;
    memset(diRec, 0, sizeof(diRec));
    memset(aldro, 0, sizeof(aldro));
    memset(c->alpha, 0, sizeof(c->alpha));
    /* dielectronic (:3204-3305) */
    tt = T > 1000. ? T / 10000. : 0.1;
    for (i = 0; i < t->nDiel; i++) {
        const double* r = t->diel + 8 * i;
        int el = (int)r[0], n = (int)r[1], g = (int)r[7];
        double v = 1.e-12 * (r[2] / tt + r[3] + r[4] * tt + r[5] * tt * tt) * pow(tt, -1.5) * exp(-r[6] / tt);
        ion = el + 1 - n;
        if (ion > t->nstages || ion < 1) continue;
        if (g == 0 || (g == 1 && T < 20000.) || (g == 2 && T >= 20000.)) diRec[el][ion] = v;
    }
    for (i = 0; i < t->nAldro; i++) {
        const double* r = t->aldro + 6 * i;
        int el = (int)r[0], n = (int)r[1];
        double a = r[2] * pow(T, -1.5) * exp(-r[4] / T) * (1. + r[3] * exp(-r[5] / T));
        ion = el - n + 1;
        if (ion <= t->nstages && ion >= 1) aldro[el][ion] = a;
    }
    if (T > 60000.) memcpy(diRec, aldro, sizeof(diRec));
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(t->nstages - 1, z); ion++) {
            int nel = z - ion;
            if (nel <= 11) {
                double bb = RRB(1, z, nel) + RRB(4, z, nel) * exp(-RRB(5, z, nel) / T);
                double s0 = sqrt(T / RRB(2, z, nel)), s1 = sqrt(T / RRB(3, z, nel)), dr = 0.;
                int k;
                c->alpha[z][ion] = RRB(0, z, nel) / (s0 * pow(1. + s0, 1. - bb) * pow(1. + s1, 1. + bb));
                for (k = 0; k < 9; k++)
                    if (DRE(k, z, nel) / T < 87.) dr += DRC(k, z, nel) * exp(-DRE(k, z, nel) / T);
                c->alpha[z][ion] += dr * pow(T, -1.5);
            } else {
                int n = z - ion + 1;
                double rr;
                if (n <= 3 || n == 11 || (z > 5 && z < 9) || z == 10 || (z == 26 && n > 11)) {
                    double q = sqrt(T / RNEW(2, z, n));
                    rr = RNEW(0, z, n) / (q * pow(q + 1., 1. - RNEW(1, z, n)) *
                                          pow(1. + sqrt(T / RNEW(3, z, n)), 1. + RNEW(1, z, n)));
                } else {
                    double q = T * 1.e-4;
                    if (z == 26 && n <= 13) rr = fe[0][n - 4] / pow(q, fe[1][n - 4] + fe[2][n - 4] * log10(q));
                    else rr = RREC(0, z, n) / pow(q, RREC(1, z, n));
                }
                c->alpha[z][ion] = rr;
                if (diRec[z][ion] == 0. && ion < 7 && c->on[8]) diRec[z][ion] = diRec[8][ion];
                c->alpha[z][ion] += diRec[z][ion] > 0. ? diRec[z][ion] : 0.;
            }
        }
    }
}

/* collisionalIonisation (:3552-3582), including its early return */
static void coll_ion(const oracle_balance_tables* t, ocell* c) {
    const double te = c->Te * 8.617385e-05;
    int z, ion;
    memset(c->cI1, 0, sizeof(c->cI1));
    memset(c->cI2, 0, sizeof(c->cI2));
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(t->nstages - 1, z - 1); ion++) {
            int elec = z - ion + 1;
            double u = CF(0, z, elec) / te;
            if (u > 80.0) return;
            c->cI1[z][ion] = CF(2, z, elec) * (1.0 + CF(1, z, elec) * sqrt(u)) / (CF(3, z, elec) + u) *
                             pow(u, CF(4, z, elec)) * exp(-u);
            c->cI2[z][ion] = CTH(z, ion) * c->cI1[z][ion] * KBOLTZ;
        }
    }
}

static double chex_scaled(int z, int ion, double T, int scale_all, int* found) {
    size_t k;
    double a = 0., b = 0., cc = 0., d = 0.;
    *found = 0;
    for (k = 0; k < sizeof(CHEX) / sizeof(CHEX[0]); k++)
        if (CHEX[k].z == z && CHEX[k].ion == ion) { a = CHEX[k].a; b = CHEX[k].b; cc = CHEX[k].c; d = CHEX[k].d; *found = 1; }
    a *= 1.e-9;
    if (!scale_all) return a;
    {
        double t4 = T / 10000., v;
        if (T < 6000. || T > 5.e4) v = 0.;
        else v = a * pow(t4, b) * (1. + cc * exp(d * t4));
        return v < 0. ? 0. : v;
    }
}

static double deltaE_k(int z, int ion) {
    if (z == 7 && ion == 1) return 10863.;
    if (z == 8 && ion == 1) return 2205.;
    return 0.;
}

/* photoionisation + getPhotoRates (:3584-3832, 3971-4040) */
static void photo_ion(const oracle_balance_tables* t, const double* J, int C1, int cell, ocell* c) {
    const int B = t->nbins;
    double xe = 0., heatef;
    int z, ion, s;
#define JJ(b) J[(size_t)((b)-1) * C1 + cell] /* 1-based bin */
#define NU(b) t->nuArray[(b)-1]
    memset(c->pI1, 0, sizeof(c->pI1));
    memset(c->pI2, 0, sizeof(c->pI2));
    for (z = 1; z <= NEL; z++)
        for (ion = 1; ion < 32; ion++)
            for (s = 1; s <= 10; s++) c->pI1[s][z][ion] = 1.e-20;
    for (z = 1; z <= NEL; z++)
        if (c->on[z]) xe += c->X[z][1] * c->ab[z] * c->H;
    xe = c->Ne / xe;
    if (xe > 0.95) heatef = 1.;
    else {
        if (xe < 1.e-4) xe = 1.e-4;
        heatef = 0.9971 * (1. - pow(1. - pow(xe, 0.2663), 1.3163));
    }
    for (s = 0; s < t->nShells; s++) {
        const int sz = t->shZ[s], si = t->shIon[s], sn = t->shNum[s];
        const int lo = t->shLo[s], hi = t->shHi[s] < B ? t->shHi[s] : B;
        const double* xs = t->xs + t->shOff[s];
        if (!c->on[sz] || si > min_i(sz, t->nstages - 1)) continue;
        if (sz < 3) {
            double g = 1.e-20, h = 0.;
            int j;
            for (j = lo; j <= B; j++) {
                double ph = xs[j - lo];
                if (ph < 1.e-35) ph = 0.;
                if (JJ(j) > 0.) {
                    g += JJ(j) * ph / (HCRYD * NU(j));
                    h += ph * JJ(j) * (NU(j) - NU(lo)) / NU(j);
                }
            }
            c->pI1[sn][sz][si] = g;
            c->pI2[sn][sz][si] = h;
        } else {
            const double y1 = AUG(sz, si, sn, 1);
            const double augerE = HCRYD * NU(lo) * y1;
            double g = 0., h = 0., gH = 0., hH = 0., ph;
            int i, iup = hi;
            for (i = lo + 1; i <= min_i(iup, t->secIonP - 1); i++) {
                if (JJ(i) > 0.) {
                    ph = xs[i - lo];
                    if (ph < 1.e-35) ph = 0.;
                    g += JJ(i) * ph / (HCRYD * NU(i));
                    h += JJ(i) * ph;
                }
            }
            h -= g * augerE;
            ph = xs[0];
            g += JJ(lo) * ph / (HCRYD * NU(lo));
            h += JJ(lo) * ph * (HCRYD * NU(lo) - augerE) / (HCRYD * NU(lo));
            for (i = max_i(lo + 1, t->secIonP); i <= iup; i++) {
                ph = xs[i - lo];
                gH += JJ(lo) * ph / (HCRYD * NU(i)); /* radField(IPnuP), as the reference has it */
                hH += JJ(lo) * ph;
            }
            g += gH;
            hH -= gH * augerE;
            h += hH * heatef;
            c->pI1[sn][sz][si] = g;
            c->pI2[sn][sz][si] = h;
        }
    }
    /* reverse charge exchange onto the outer shell (:3715-3830) */
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(z, t->nstages - 1); ion++) {
            int out, g0, g1, found;
            double chx = chex_scaled(z, ion, c->Te, 1, &found), rev = 0.;
            outer_shell(z, z - ion + 1, &out, &g0, &g1);
            if (deltaE_k(z, ion) > 1.)
                rev = chx * (1. - c->X[1][1]) * c->H * 2. * exp(-deltaE_k(z, ion) / c->Te) / ((double)g0 / (double)g1);
            c->pI1[out][z][ion] += rev;
        }
    }
#undef JJ
#undef NU
}

/* compton (:3834-3897) */
static void compton_rates(const oracle_balance_tables* t, const double* J, int C1, int cell, ocell* c) {
    double s1 = 0., s2 = 0., rH = 0., rHe = 0.;
    int i;
    for (i = 1; i <= t->nbins; i++) {
        double rf = J[(size_t)(i - 1) * C1 + cell], nu = t->nuArray[i - 1];
        s1 += rf;
        s2 += rf / nu;
        if (i > t->cRecoilHP) rH += t->xSecRecoil[i - 1] * rf * (2.66e-5 * nu * nu - 1.);
        if (i > t->cRecoilHeP) rHe += t->xSecRecoil[i - 1] * rf * (2.66e-5 * nu * nu - 1.8);
    }
    s1 *= RYD2ERG;
    c->cCool = (6.65e-25 * 6.75e-10 * c->Te * s2) / c->H;
    c->cHeat = (6.65e-25 * s1 / 8.184e-7) / c->H;
    c->cRecoil = (rH * c->X[1][1] * c->ab[1] + rHe * (c->on[2] ? c->X[2][1] * c->ab[2] : 0.) * 2.) * RYD2ERG;
}

/* the chain of one heavy element Z >= 3 with multi-electron Auger ejection
 * (ionBalance :1826-1929), at the cell's Ne, with damping on the first pass */
static void heavy_chain(const oracle_balance_tables* t, ocell* c, int z) {
    const int NS = t->nstages;
    double out_[32], in_[32], mat[33], sumMat;
    int top = min_i(NS, z + 1), ion;
    if (!c->on[z]) return;
    memset(out_, 0, sizeof(out_));
    memset(in_, 0, sizeof(in_));
    for (ion = 1; ion <= min_i(z, NS - 1); ion++) {
        int osh, g0, g1, sh;
        outer_shell(z, z - ion + 1, &osh, &g0, &g1);
        out_[ion] += c->cI1[z][ion] * c->Ne;
        for (sh = 1; sh <= osh; sh++) {
            int na = NAUG(z, ion, sh), ne;
            out_[ion] += c->pI1[sh][z][ion];
            for (ne = 2; ne <= na; ne++) {
                int maxim = min_i(ion + ne - 1, min_i(NS - 1, z));
                double ratio = c->X[z][maxim] > 1e-30 ? c->X[z][ion] / c->X[z][maxim] : 1.;
                out_[maxim] += c->pI1[sh][z][ion] * AUG(z, ion, sh, ne) * ratio;
            }
        }
    }
    for (ion = 1; ion <= min_i(z, NS - 1); ion++) {
        int found;
        double chx = chex_scaled(z, ion, c->Te, 0, &found); /* not T-scaled here */
        in_[ion] = c->Ne * c->alpha[z][ion] + chx * c->H * c->X[1][1];
    }
    mat[1] = 1.;
    sumMat = 1.;
    for (ion = 2; ion <= top; ion++) {
        mat[ion] = mat[ion - 1] * out_[ion - 1] / in_[ion - 1];
        sumMat += mat[ion];
    }
    sumMat = 1. / sumMat;
    c->X[z][top] = mat[top] * sumMat;
    for (ion = 1; ion <= min_i(z, NS - 1); ion++) {
        double v;
        if (c->nIterX < 2 && mat[ion] * sumMat > 0.) {
            double old = c->X[z][ion] <= 0. ? 1.e-5 : c->X[z][ion];
            v = pow(10., 0.5 * (log10(old) + log10(mat[ion] * sumMat)));
        } else {
            v = mat[ion] * sumMat;
        }
        if (v > XMAX) v = XMAX;
        if (v < 1.e-20) v = 1.e-20;
        c->X[z][ion] = v;
    }
}

/* ionBalance (:1575-1992); returns lgConv */
static int ion_balance(const oracle_balance_tables* t, ocell* c) {
    const int NS = t->nstages;
    int conv = 1;
    for (;;) {
        double expFact = 157893.94 / c->Te, collIonH, HIOld, HeIOld = 0., HeIIOld = 0., Ne;
        double ionRatio[3][3], ionProd[3][3], den[3];
        int z, ion, i;
        if (expFact > 75.) expFact = 75.;
        collIonH = 2.75E-16 * c->Te * sqrt(c->Te) * (157893.94 / c->Te + 2.) * exp(-expFact);
        HIOld = c->X[1][1];
        if (c->on[2]) { HeIOld = c->X[2][1]; HeIIOld = c->X[2][2]; }
        for (z = 0; z < 3; z++) for (i = 0; i < 3; i++) { ionRatio[z][i] = 0.; ionProd[z][i] = 1.; }
        den[1] = den[2] = 1.;
        /* H and He: ratio chain with charge exchange (:1715-1790) */
        for (z = 1; z <= 2; z++) {
            if (!c->on[z]) continue;
            for (ion = 1; ion <= min_i(z, NS - 1); ion++) {
                int out, g0, g1, found;
                double chx = chex_scaled(z, ion, c->Te, 1, &found), rev = 0.;
                outer_shell(z, z - ion + 1, &out, &g0, &g1);
                if (deltaE_k(z, ion) > 1.)
                    rev = chx * (1. - c->X[1][1]) * c->H * 2. * exp(-deltaE_k(z, ion) / c->Te) / ((double)g0 / (double)g1);
                if (z > 1 || ion > 1) collIonH = 0.;
                ionRatio[z][ion] = (c->pI1[out][z][ion] + collIonH * c->Ne + rev) /
                                   (c->Ne * c->alpha[z][ion] + chx * c->H * c->X[1][1]);
            }
        }
        for (z = 1; z <= 2; z++) {
            if (!c->on[z]) continue;
            for (ion = 1; ion <= min_i(z, NS - 1); ion++)
                for (i = 1; i <= ion; i++) ionProd[z][ion] *= ionRatio[z][i];
            for (ion = 1; ion <= min_i(z, NS - 1); ion++) den[z] += ionProd[z][ion];
            for (ion = 1; ion <= min_i(z + 1, NS); ion++) {
                double v = ion == 1 ? 1. / den[z] : ionProd[z][ion - 1] / den[z];
                if (v > XMAX) v = XMAX;
                if (v < 1.e-20) v = 1.e-20;
                c->X[z][ion] = v;
            }
        }
        /* heavy elements: chain with multi-electron Auger ejection (:1826-1929) */
        for (z = 3; z <= NEL; z++) heavy_chain(t, c, z);
        /* electron density (:1932-1963) */
        Ne = 0.;
        for (z = 1; z <= NEL; z++) {
            if (!c->on[z]) continue;
            for (ion = 2; ion <= min_i(z + 1, NS); ion++)
                if (c->X[z][ion] >= 1.e-10) Ne += (ion - 1) * c->ab[z] * c->X[z][ion];
        }
        Ne *= c->H;
        if (Ne == 0.) Ne = 1.;
        c->Ne = Ne;
        {
            double dHI = (c->X[1][1] - HIOld) / HIOld;
            double dHeI = c->on[2] ? (c->X[2][1] - HeIOld) / HeIOld : 0.;
            double dHeII = c->on[2] ? (c->X[2][2] - HeIIOld) / HeIIOld : 0.;
            int big = fabs(dHI) > 0.01 || fabs(dHeI) > 0.01 || fabs(dHeII) > 0.01;
            if (big && c->nIterX < 15) { c->nIterX++; continue; }
            if (big && c->nIterX == 15) conv = 0;
        }
        return conv;
    }
}

/* thermBalance (:887-1215): heating and cooling integrals */
/* ---- grains: charge, photoelectric heating, gas-grain collisions ---------- */
#define ORC_AMU 1.66053e-24    /* constants_mod.f90:11 */
#define ORC_ME 9.109382e-28    /* constants_mod.f90:17 */
#define ORC_PIB 3.141592654    /* constants_mod.f90:45 */
#define GY0 0.5                /* updateCell Y0, Y1 (Baldwin et al. 1991) */
#define GY1 0.2
static const double AWEIGHT[30] = {1.0080, 4.0026, 6.941, 9.0122, 10.811, 12.0111, 14.0067, 15.9994,
                                   18.9984, 20.179, 22.9898, 24.305, 26.9815, 28.086, 30.9738, 32.06,
                                   35.453, 39.948, 39.102, 40.08, 44.956, 47.90, 50.9414, 51.996,
                                   54.9380, 55.847, 58.9332, 58.71, 63.546, 65.37};

/* getGrainEmission (update_mod.f90:557-604) */
static double grain_emission(const oracle_balance_tables* t, const oracle_balance_state* st, int cell,
                             int g, int s, int a, double Vg) {
    const int C1 = st->nCells + 1, B = t->nbins;
    const double Vn = st->gVn[s];
    const double thres = Vn + Vg > Vn ? Vn + Vg : Vn;
    const double* cabs = st->gCabs + (size_t)g * B;
    const double r = st->gRadius[a];
    double sum = 0.;
    int i, ip = oracle_locate(t->nuArray, B, thres) + 1;
    for (i = ip; i <= B; i++) {
        const double nu = t->nuArray[i - 1];
        double Yn = GY0 * (1. - Vn / nu), Yhat, photFlux, Qa, f;
        if (Yn > GY1) Yn = GY1;
        f = 1. - Vg / (nu - Vn);
        if (f < 0.) f = 0.;
        if (f > 1.) f = 1.;
        Yhat = Yn * f;
        photFlux = st->J[(size_t)(i - 1) * C1 + cell] / (HCRYD * nu);
        photFlux = photFlux * (4. * ORC_PIB);
        Qa = cabs[i - 1] / (1.e-8 * (r * r));
        sum = sum + Yhat * photFlux * Qa;
    }
    return sum;
}

/* getGrainRecombination (update_mod.f90:609-682): electrons minus the ions of
 * the grid state at the start of the update */
static double grain_recombination(const oracle_balance_tables* t, const ocell* c, double Vg) {
    const double kT = 6.336e-6 * c->Te;
    const double e8 = (1.1045e-15) * c->Te / ORC_PIB;
    double vmean = sqrt(e8 / ORC_ME), eta = -Vg / kT, rec;
    int z, istage;
    eta = eta <= 0. ? 1. - eta : exp(-eta);
    rec = 0. + c->Ne * vmean * 1. * eta;
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (istage = 2; istage <= min_i(z + 1, t->nstages); istage++) {
            const double cpDen = c->X0[z][istage] * c->ab[z] * c->H;
            const double mcp = AWEIGHT[z - 1] * ORC_AMU;
            const double Z = (double)(istage - 1);
            vmean = sqrt(e8 / mcp);
            eta = Z * Vg / kT;
            eta = eta <= 0. ? 1. - eta : exp(-eta);
            rec = rec - cpDen * vmean * 1. * eta;
        }
    }
    return rec;
}

/* setGrainPotential (update_mod.f90:484-553): secant iteration on grain
 * emission = recombination, started at 0 and 0.05 Ryd, steps capped at 5 % of
 * the threshold, 100 passes at most.  The secant slope starts at 0 for each
 * grain (the reference's SAVEd slope carries over between grains). */
static double grain_potential(const oracle_balance_tables* t, const oracle_balance_state* st,
                              const ocell* c, int cell, int g, int s, int a) {
    double Vg = 0., VgOld = 0., EmiOld = 0., RecOld = 0., slope = 0.;
    int n;
    for (n = 1; n <= 100; n++) {
        double thr, Emi, Rec, delta, d1, mx;
        if (n == 1) {
            Vg = 0.;
            VgOld = Vg;
            EmiOld = grain_emission(t, st, cell, g, s, a, Vg);
            RecOld = grain_recombination(t, c, Vg);
            Vg = VgOld + 0.05;
        }
        thr = st->gVn[s] + Vg > st->gVn[s] ? st->gVn[s] + Vg : st->gVn[s];
        Emi = grain_emission(t, st, cell, g, s, a, Vg);
        Rec = grain_recombination(t, c, Vg);
        delta = Emi - Rec;
        mx = Emi + Rec > 1.e-35 ? Emi + Rec : 1.e-35;
        d1 = fabs(delta / (0.5 * mx));
        if (d1 < 0.005) return Vg;
        if (Vg != VgOld) {
            const double fac = (Emi - EmiOld) - (Rec - RecOld);
            if (fac != 0.) slope = fac / (Vg - VgOld);
        }
        VgOld = Vg;
        RecOld = Rec;
        EmiOld = Emi;
        d1 = -delta / slope;
        delta = fabs(d1);
        if (0.05 * thr < delta) delta = 0.05 * thr;
        delta = copysign(delta, d1);
        Vg = Vg + delta;
    }
    return Vg;
}

/* ionisation potential [erg] released when stage istage of z recombines: the
 * first bin of the outermost shell of stage istage-1 */
static double ip_erg(const oracle_balance_tables* t, int z, int ion) {
    int k, best = -1, bestNum = -1;
    for (k = 0; k < t->nShells; k++)
        if (t->shZ[k] == z && t->shIon[k] == ion && t->shNum[k] > bestNum) {
            bestNum = t->shNum[k];
            best = k;
        }
    return best < 0 ? 0. : t->nuArray[t->shLo[best] - 1] * RYD2ERG;
}

/* setPhotoelHeatCool and setDustGasCollHeatCool (update_mod.f90:686-884) */
static void grain_terms(const oracle_balance_tables* t, const oracle_balance_state* st, const ocell* c,
                        int cell, const double* Vg, double* peD, double* collD, double* peG,
                        double* collG) {
    const int C1 = st->nCells + 1, B = t->nbins, nS = st->gS, nA = st->gA;
    const double T = c->Te, Ne = c->Ne;
    const double kT = 6.336e-6 * T, e8 = (1.1045e-15) * T / ORC_PIB;
    const double* Td = st->Tdust + (size_t)cell * (nA + 1) * (nS + 1);
    double pg = 0., cg = 0., IPerg = 0., vmean;
    int s, a, i, z, istage;
    for (s = 0; s < nS; s++)
        for (a = 0; a < nA; a++) {
            const int g = s * nA + a;
            const double Vn = st->gVn[s], V = Vg[g];
            const double th = Vn + V > Vn ? Vn + V : Vn;
            const double* cabs = st->gCabs + (size_t)g * B;
            int thP = oracle_locate(t->nuArray, B, th);
            if (thP <= 0) thP = 1;
            peD[g] = 0.;
            for (i = thP; i <= B; i++) {
                const double nu = t->nuArray[i - 1], dn = nu - Vn;
                double Yn = GY0 * (1. - Vn / nu), f, Yhat, photFlux, EY, m2;
                if (Yn > GY1) Yn = GY1;
                f = 1. - V / dn;
                if (f < 0.) f = 0.;
                if (f > 1.) f = 1.;
                Yhat = Yn * f;
                photFlux = st->J[(size_t)(i - 1) * C1 + cell] / (HCRYD * nu);
                m2 = (dn * dn - V * V) / dn;
                if (m2 < 0.) m2 = 0.;
                EY = Yn * 0.5 * (dn < m2 ? dn : m2);
                peD[g] = peD[g] + cabs[i - 1] * photFlux * EY * HCRYD / ORC_PIB;
                pg = pg + cabs[i - 1] * photFlux * (EY - Yhat * V) * st->gAbun[s] * st->gWeight[a];
            }
        }
    *peG = pg * st->Ndust[cell] * HCRYD / c->H;
    for (s = 0; s < nS * nA; s++) collD[s] = 0.;
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (istage = 1; istage <= min_i(t->nstages, z + 1); istage++) {
            const double Z = (double)(istage - 1);
            const double mcp = AWEIGHT[z - 1] * ORC_AMU;
            vmean = sqrt(e8 / mcp);
            if (istage > 1) IPerg = ip_erg(t, z, istage - 1);
            for (s = 0; s < nS; s++) {
                const double S = istage == 1 ? 2. * mcp * st->gMsurf[s] / pow(mcp + st->gMsurf[s], 2.) : 1.;
                for (a = 0; a < nA; a++) {
                    const int g = s * nA + a;
                    const double r = st->gRadius[a], V = Vg[g];
                    const double psi = Z * V / kT;
                    const double Tg = Td[(size_t)(a + 1) * (nS + 1) + s + 1];
                    double eta, xi;
                    if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
                    else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
                    collD[g] = collD[g] + c->X[z][istage] * c->ab[z] * c->H * ORC_PIB * r * r * 1.e-8 * S *
                                              vmean *
                                              (2. * kT * RYD2ERG * xi -
                                               eta * (Z * V * RYD2ERG - IPerg + 2. * KBOLTZ * Tg));
                    cg = cg + c->X[z][istage] * st->gWeight[a] * st->gAbun[s] * st->Ndust[cell] * c->ab[z] *
                                  ORC_PIB * r * r * 1.e-8 * S * vmean *
                                  (2. * kT * RYD2ERG * xi - eta * 2. * KBOLTZ * Tg);
                }
            }
        }
    }
    vmean = sqrt(e8 / ORC_ME);
    for (s = 0; s < nS; s++)
        for (a = 0; a < nA; a++) {
            const int g = s * nA + a;
            const double r = st->gRadius[a], V = Vg[g], S = 1., Z = -1.;
            const double psi = Z * V / kT;
            double eta, xi;
            if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
            else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
            collD[g] = collD[g] + Ne * ORC_PIB * r * r * 1.e-8 * S * vmean *
                                      (2. * kT * RYD2ERG * xi - eta * (Z * V * RYD2ERG));
            cg = cg + Ne * st->gWeight[a] * st->gAbun[s] * st->Ndust[cell] * ORC_PIB * r * r * 1.e-8 * S * vmean *
                          (2. * kT * RYD2ERG * xi) / c->H;
        }
    for (s = 0; s < nS * nA; s++) collD[s] = collD[s] / (4. * ORC_PIB);
    *collG = cg / (4. * ORC_PIB);
}

/* parts (may be NULL): free-free, recombination, H Ly-alpha/beta plus
 * forbidden-line, collisional-ionisation cooling */
static void therm_balance_parts(const oracle_balance_tables* t, ocell* c, double* heat, double* cool,
                                double* parts) {
    const int NS = t->nstages;
    const double T = c->Te, Ne = c->Ne;
    double coolFF = 0., coolRec = 0., coolColl, coolCollH = 0., heatSte = 0., Te4 = T / 1.e4, lt;
    int z, ion;
    for (z = 1; z <= NEL; z++) {
        if (!(c->on[z] && NS > z)) continue;
        double Np = c->X[z][z + 1] * c->ab[z], bFF, bRec;
        lt = log10(T / (double)(z * z));
        Te4 = T / (z * z * 1.e4);
        bFF = z * (1.0108464E-11 + 9.7930778E-13 * lt - 6.6433144E-13 * lt * lt + 2.4793747E-13 * lt * lt * lt -
                   2.3938215E-14 * lt * lt * lt * lt);
        coolFF += Np * Ne * bFF * KBOLTZ * T / sqrt(T);
        bRec = z * (9.4255985E-11 - 4.04794384E-12 * lt - 1.0055237E-11 * lt * lt + 1.99266862E-12 * lt * lt * lt -
                    1.06681387E-13 * lt * lt * lt * lt);
        coolRec += Np * Ne * bRec * KBOLTZ * T / sqrt(T);
    }
    Te4 = T / 1.e4;
    for (z = 2; z <= NEL; z++) {
        if (!(c->on[z] && NS > z)) continue;
        double Np = c->X[z][z] * c->ab[z], bFF, bRec;
        lt = log10(4. * T / (double)(z * z));
        Te4 = 4. * T / (z * z * 1.e4);
        bFF = (z / 2.) * (1.070073e-11 - 2.5730207e-13 * lt + 2.109134e-13 * lt * lt);
        coolFF += Np * Ne * bFF * KBOLTZ * T / sqrt(T);
        bRec = (z / 2.) * (1.06926e-10 - 2.41756e-11 * lt + 1.14542e-12 * lt * lt - 5.06535e-13 * lt * lt * lt +
                           9.16745e-14 * lt * lt * lt * lt);
        coolRec += Np * Ne * bRec * KBOLTZ * T / sqrt(T);
    }
    /* H Ly-alpha/beta collisional excitation; Te4 is whatever the loop above left */
    if (T > 5000.)
        coolColl = (2.47e-8 * exp(-118338. / T) * pow(Te4, -0.228) + 1.32e-8 * exp(-140252. / T) * pow(Te4, -0.460)) *
                   HCRYD * c->X[1][1] * Ne;
    else
        coolColl = 0.;
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(z, NS - 1); ion++) coolCollH += c->ab[z] * c->X[z][ion] * c->cI2[z][ion] * Ne;
    }
    /* forbidden-line cooling (forLines :1221-1262) */
    for (z = 3; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(z + 1, NS); ion++) {
            int a = t->atom_idx[z * 32 + ion];
            if (a < 0) continue;
            coolColl += nlevel_cooling(t, a, T, Ne, c->ab[z] * c->X[z][ion]);
        }
    }
    for (z = 1; z <= NEL; z++) {
        if (!c->on[z]) continue;
        for (ion = 1; ion <= min_i(z, NS - 1); ion++) {
            int osh, g0, g1, sh;
            outer_shell(z, z - ion + 1, &osh, &g0, &g1);
            if (z <= 2) osh = 1;
            for (sh = 1; sh <= osh; sh++) heatSte += c->pI2[sh][z][ion] * c->X[z][ion] * c->ab[z];
        }
    }
    *heat = heatSte;
    *cool = coolFF + coolRec + coolColl + coolCollH;
    if (parts) {
        parts[0] = coolFF;
        parts[1] = coolRec;
        parts[2] = coolColl;
        parts[3] = coolCollH;
    }
    if (t->lgCompton) {
        *heat += c->cHeat * Ne + c->cRecoil;
        *cool += c->cCool * Ne;
    }
}

static void therm_balance(const oracle_balance_tables* t, ocell* c, double* heat, double* cool) {
    therm_balance_parts(t, c, heat, cool, NULL);
}

/* updateCell + iterateT for one cell */
static void update_one(const oracle_balance_tables* t, oracle_balance_state* st, int cell,
                       double* nHit, double* nIon, double* nTe) {
    const int C1 = st->nCells + 1, NE = t->nElemUsed, NS = t->nstages;
    ocell c;
    double XOld, Thigh = 0., Tlow = 0., resHigh = 0., resLow = 0.;
    double Vg[4096], peD[4096], collD[4096];
    int e, ion, b, hit = 0, nT = 0, ibConv = 1, g;
    const int nG = st->gS * st->gA;
    for (b = 0; b < t->nbins; b++)
        if (st->J[(size_t)b * C1 + cell] > 0.) { hit = 1; break; }
    st->lgBlack[cell] = 0;
    if (!hit) {
        st->lgBlack[cell] = 1;
        *nHit += 1.;
        if (st->nIterT) st->nIterT[cell] = 0;
        return;
    }
    memset(&c, 0, sizeof(c));
    c.Te = st->Te[cell];
    c.Ne = st->Ne[cell];
    c.H = st->Hden[cell];
    for (e = 0; e < NE; e++) {
        int z = t->elemZ[e];
        c.on[z] = 1;
        c.ab[z] = t->abund[e];
        for (ion = 1; ion <= NS; ion++) c.X[z][ion] = st->ionDen[((size_t)(ion - 1) * NE + e) * C1 + cell];
    }
    memcpy(c.X0, c.X, sizeof(c.X));
    XOld = c.X[1][1];
    for (;;) {
        double heat, cool, res;
        nT++;
        calc_alpha(t, &c);
        photo_ion(t, st->J, C1, cell, &c);
        coll_ion(t, &c);
        c.nIterX = 1;
        if (t->lgCompton) compton_rates(t, st->J, C1, cell, &c);
        ibConv = ion_balance(t, &c);
        if (nG > 0) {  /* grain charges, photoelectric and collision terms (:318-333) */
            double peG = 0., collG = 0.;
            for (g = 0; g < nG; g++) Vg[g] = grain_potential(t, st, &c, cell, g, g / st->gA, g % st->gA);
            grain_terms(t, st, &c, cell, Vg, peD, collD, &peG, &collG);
            therm_balance(t, &c, &heat, &cool);
            cool = cool + collG;  /* thermBalance :1205-1208 */
            heat = heat + peG;
        } else {
            therm_balance(t, &c, &heat, &cool);
        }
        res = heat - cool;
        if (fabs(res) >= 0.02 * (fabs(cool) + fabs(heat)) / 2.) {
            if (nT < 20) {
                if (res < 0.) {
                    resLow = res;
                    Tlow = c.Te;
                    if (Thigh != 0.) c.Te = Tlow - (Thigh - Tlow) * resLow / (resHigh - resLow);
                    else c.Te = c.Te / 1.2;
                } else {
                    resHigh = res;
                    Thigh = c.Te;
                    if (Tlow != 0.) c.Te = Tlow - (Thigh - Tlow) * resLow / (resHigh - resLow);
                    else c.Te = c.Te * 1.2;
                }
                if (nT >= 6 && fabs(Thigh - Tlow) <= 0.002 * (Thigh + Tlow)) {
                    if (res < 0.) { c.Te = Thigh - 100.; Thigh = 0.; }
                    else { c.Te = Tlow + 100.; Tlow = 0.; }
                }
                if (c.Te <= 0.) c.Te = 1.;
                continue;
            }
            if (Thigh == 0.) c.Te = Tlow;
            else if (Tlow == 0.) c.Te = Thigh;
            else c.Te = (Thigh + Tlow) * 0.5;
            *nTe += 1.;
        }
        break;
    }
    if (!ibConv) *nIon += 1.;
    for (g = 0; g < nG; g++) {
        st->collD[(size_t)g * C1 + cell] = collD[g];
        st->peD[(size_t)g * C1 + cell] = peD[g];
        if (st->grainPot) st->grainPot[(size_t)g * C1 + cell] = Vg[g];
    }
    st->Te[cell] = c.Te;
    st->Ne[cell] = c.Ne;
    for (e = 0; e < NE; e++) {
        int z = t->elemZ[e];
        for (ion = 1; ion <= NS; ion++) st->ionDen[((size_t)(ion - 1) * NE + e) * C1 + cell] = c.X[z][ion];
    }
    st->lgConverged[cell] = fabs((c.X[1][1] - XOld) / XOld) <= st->XHILimit ? 1 : 0;
    if (st->nIterT) st->nIterT[cell] = nT;
}

int oracle_update(const oracle_balance_tables* t, oracle_balance_state* s, int nthreads) {
    double nHit = 0., nIon = 0., nTe = 0.;
    int cell;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 1 ? nthreads : 1) reduction(+ : nHit, nIon, nTe)
#endif
    for (cell = 1; cell <= s->nCells; cell++) update_one(t, s, cell, &nHit, &nIon, &nTe);
    (void)nthreads;
    s->noHit = nHit;
    s->noIonBal = nIon;
    s->noTeBal = nTe;
    return 0;
}

/* ---- exported pieces, for the known-answer tests ------------------------ */
void oracle_alpha(const oracle_balance_tables* t, double Te, double* out /* [31*32] */) {
    ocell c;
    int e, z, ion;
    memset(&c, 0, sizeof(c));
    c.Te = Te;
    for (e = 0; e < t->nElemUsed; e++) c.on[t->elemZ[e]] = 1;
    calc_alpha(t, &c);
    for (z = 0; z < 31; z++)
        for (ion = 0; ion < 32; ion++) out[z * 32 + ion] = c.alpha[z][ion];
}

double oracle_nlevel_cooling(const oracle_balance_tables* t, int atom, double Te, double Ne,
                             double scale) {
    return nlevel_cooling(t, atom, Te, Ne, scale);
}

/* The heavy-element chain of element Z alone (ionBalance :1826-1929) at the
 * given Te, Ne, H and X(H0), from the per-stage alpha, collisional rate cI1
 * and per-shell photoionisation rates pI1[shell][stage] (shells 1..10):
 * X[stage] (1..31) in, out; damp != 0 takes the first-pass geometric
 * damping. */
void oracle_heavy_chain(const oracle_balance_tables* t, int z, double Te, double Ne, double H, double XH0,
                        const double* alpha /* [32] */, const double* cI1 /* [32] */,
                        const double* pI1 /* [11][32] */, int damp, double* X /* [32] */) {
    ocell* c = (ocell*)calloc(1, sizeof(ocell));
    int ion, sh;
    if (!c) return;
    c->Te = Te;
    c->Ne = Ne;
    c->H = H;
    c->on[1] = 1;
    c->on[z] = 1;
    c->X[1][1] = XH0;
    c->nIterX = damp ? 1 : 2;
    for (ion = 0; ion < 32; ion++) {
        c->X[z][ion] = X[ion];
        c->alpha[z][ion] = alpha[ion];
        c->cI1[z][ion] = cI1[ion];
        for (sh = 0; sh < 11; sh++) c->pI1[sh][z][ion] = pI1[sh * 32 + ion];
    }
    heavy_chain(t, c, z);
    for (ion = 0; ion < 32; ion++) X[ion] = c->X[z][ion];
    free(c);
}

/* One T iteration of updateCell (update_mod.f90:15-482) for a cell at a given
 * Te, from the state in s (J, Ne, Hden, ionDen): calcAlpha, the photo-
 * ionisation and collisional rates, ionBalance and thermBalance; returns the
 * heating and cooling.  No grains. */
void oracle_cell_heat_cool(const oracle_balance_tables* t, const oracle_balance_state* st, int cell,
                           double Te, double* heat, double* cool) {
    const int C1 = st->nCells + 1, NE = t->nElemUsed, NS = t->nstages;
    ocell* c = (ocell*)calloc(1, sizeof(ocell));
    int e, ion;
    if (!c) return;
    c->Te = Te;
    c->Ne = st->Ne[cell];
    c->H = st->Hden[cell];
    for (e = 0; e < NE; e++) {
        int z = t->elemZ[e];
        c->on[z] = 1;
        c->ab[z] = t->abund[e];
        for (ion = 1; ion <= NS; ion++) c->X[z][ion] = st->ionDen[((size_t)(ion - 1) * NE + e) * C1 + cell];
    }
    memcpy(c->X0, c->X, sizeof(c->X));
    calc_alpha(t, c);
    photo_ion(t, st->J, C1, cell, c);
    coll_ion(t, c);
    c->nIterX = 1;
    if (t->lgCompton) compton_rates(t, st->J, C1, cell, c);
    (void)ion_balance(t, c);
    therm_balance(t, c, heat, cool);
    free(c);
}

/* The cooling terms of thermBalance (:887-1215) of a cell at Te and Ne with
 * ion fractions X[31*32] (Z, stage) of the elements on, no radiation field:
 * parts = free-free, recombination, H Ly-alpha/beta plus forbidden lines,
 * collisional ionisation [erg/s per n_H]. */
void oracle_cooling_parts(const oracle_balance_tables* t, double Te, double Ne, const double* X,
                          double* parts) {
    ocell* c = (ocell*)calloc(1, sizeof(ocell));
    double heat, cool;
    int e, z, ion;
    if (!c) return;
    c->Te = Te;
    c->Ne = Ne;
    c->H = 1.;
    for (e = 0; e < t->nElemUsed; e++) {
        z = t->elemZ[e];
        c->on[z] = 1;
        c->ab[z] = t->abund[e];
        for (ion = 0; ion < 32; ion++) c->X[z][ion] = X[z * 32 + ion];
    }
    coll_ion(t, c);
    therm_balance_parts(t, c, &heat, &cool, parts);
    free(c);
}

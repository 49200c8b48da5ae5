/*
 * moc_oracle_emission.c -- ORACLE (test infrastructure only): FP64
 * restatement of MOCASSIN's per-cell diffuse emission, emission_mod.f90:
 *   emissionDriver :36-252, fb_ff :336-618, twoPhoton / hydro2phot /
 *   HeI2photSub :620-817, free_free / ffCoeff :818-855, RecLinesEmission
 *   :856-1153, forLines :1481-1526 (equilibrium with rec=.true.),
 *   setDiffusePDF :1528-1919 (gas only), with the edge pointers of
 *   initGammaCont (ph_mod.f90:31-202); and the line luminosities of
 *   outputGas (output_mod.f90:397-432): per-transition n-level lines and
 *   Hbeta, times Hden dV, summed over the reached cells.
 * Output per cell: the diffuse-continuum CDF recPDF(cell, 1:nbins) and the
 * line fraction totalLines(cell).  Differences from the reference, each
 * deliberate:
 *  - the free-free Gaunt factors are evaluated at each cell's own T_e (the
 *    reference reuses the arrays the last opacity evaluation left behind);
 *  - a switched-off He contributes nothing (the reference reads its ion
 *    fractions through elementXref(2) = 0, outside the array, and multiplies
 *    them by a zero abundance);
 *  - a He II Lyman line below the mesh is dropped (the reference would write
 *    bin 0), and the He II Hbeta table column past its 18 entries is clamped
 *    to the last one (the reference reads past r2aEdge when T_e > 1e5 K and
 *    the density index reaches 9);
 *  - the He I / He II Milne terms take a zero cross-section in the bins
 *    between the last gamma-table pointer and the threshold pointer (the
 *    reference indexes xSecArray before the start of that ion's block there,
 *    reading the preceding table's entries);
 *  - all arithmetic is FP64 (the reference mixes single-precision locals).
 *
 * Pinned by tests/test_emission.py: the Milne free-bound of H (recombination
 * to 1s), its hydrogenic scaling to He II and C VI, the two-photon energy per
 * decay, the free-free constant, the Hbeta emissivity, the gamma-fit
 * interpolation at a tabulated T, the recombination feed of the n-level
 * solve and the CDF assembly; by tests/test_output.py: the line
 * luminosities (per-ion sums against the pinned n-level emissivity, ratios
 * of lines from one upper level against A dE, Hbeta against case B, and
 * additivity over cells).  Parity unpinned: the He I recombination-line
 * fits, the hydrogenic line tables above 3e4 K (rbEdge / r2aEdge branches)
 * and the free-bound of non-hydrogenic heavy ions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "moc_oracle.h"
#include "moc_oracle_balance.h"
#include "moc_oracle_emission.h"
#include "moc_oracle_dust.h"
#include "moc_oracle_output.h"

#define PI_REF 3.141592654   /* constants_mod.f90 */
#define HCRYD 2.1799153e-11  /* continuum_mod.f90:15 */
#define HCRYD_K 157893.94    /* continuum_mod.f90:14 */
#define CRYD 3.2898423e15    /* continuum_mod.f90:16 */
#define HPLANCK 6.6262e-27   /* constants_mod.f90:14 */

/* hydroLinesTemps, grid_mod.f90:76-77 */
static const double HL_TEMPS[12] = {0.05, 0.10, 0.30, 0.50, 0.75, 1.0, 1.25, 1.50, 2.0, 3.0, 5.0, 10.0};

typedef struct {
    int* edge[3];      /* [nlim] 1-based bin pointers */
    double* lg[3];     /* [ntk][nlim] log10 gamma, 0 where gamma <= 0 */
    int j2TsP;         /* bin of the He I 2^3S line at 1.45673 Ryd */
    int lymP[4];       /* bins of the He II Lyman lines, 0 = below the mesh */
} prep_t;

static void prep_free(prep_t* p) {
    int k;
    for (k = 0; k < 3; k++) {
        free(p->edge[k]);
        free(p->lg[k]);
    }
}

/* initGammaCont (ph_mod.f90:96-196): locate each table frequency, move to
 * the upper bin when it lies above the bin midpoint, force the first pointer
 * to 1; fb_ff stops when the two pointers of a pair coincide (:436-442). */
static int prep_make(const oracle_emission_tables* t, prep_t* p) {
    const double* nu = t->bal->nuArray;
    const int B = t->bal->nbins;
    int k, i;
    memset(p, 0, sizeof(*p));
    for (k = 0; k < 3; k++) {
        const int nl = t->nlim[k];
        p->edge[k] = (int*)calloc(nl, sizeof(int));
        p->lg[k] = (double*)calloc((size_t)t->ntk * nl, sizeof(double));
        for (i = 0; i < nl; i++) {
            const double g = (double)(float)t->gnu[k][i];
            int e = oracle_locate(nu, B, g);
            if (e >= 1 && e < B && g > (nu[e - 1] + nu[e]) / 2.) e++;
            p->edge[k][i] = e;
        }
        p->edge[k][0] = 1;
        for (i = 0; i < t->ntk * nl; i++) p->lg[k][i] = t->gval[k][i] > 0. ? log10(t->gval[k][i]) : 0.;
        for (i = 1; i <= nl / 2; i++)
            if (p->edge[k][2 * i - 1] == p->edge[k][2 * i - 2]) {
                prep_free(p);
                return -1;
            }
    }
    p->j2TsP = oracle_locate(nu, B, (double)1.45673f);
    for (i = 0; i < 4; i++) p->lymP[i] = oracle_locate(nu, B, (double)(float)t->heiiLyman[2 * i + 1]);
    return 0;
}

/* fb_ff :389-507: log-log interpolation of the gamma fits in T, then the
 * exponential interpolation in nu between the two pointers of each pair */
static void gamma_fill(const oracle_emission_tables* t, const prep_t* p, int k, double Te, double* gam) {
    const double* nu = t->bal->nuArray;
    const int nl = t->nlim[k], ntk = t->ntk;
    const double* lg = p->lg[k];
    const int* edge = p->edge[k];
    double loc[128];
    int i, j;
    if (Te <= t->tk[0]) {
        for (i = 0; i < nl; i++) loc[i] = lg[i];
    } else if (Te >= t->tk[ntk - 1]) {
        for (i = 0; i < nl; i++) loc[i] = lg[(size_t)(ntk - 1) * nl + i];
    } else {
        const int nT = oracle_locate(t->tk, ntk, Te); /* 1-based */
        const double factor = log10(Te / t->tk[nT - 1]);
        for (i = 0; i < nl; i++) {
            const double a = (lg[(size_t)nT * nl + i] - lg[(size_t)(nT - 1) * nl + i]) /
                             log10(t->tk[nT] / t->tk[nT - 1]);
            loc[i] = lg[(size_t)(nT - 1) * nl + i] + a * factor;
        }
    }
    for (i = 1; i <= nl / 2; i++) {
        const int iup = edge[2 * i - 1], ilow = edge[2 * i - 2];
        const double aFit = (loc[2 * i - 1] - loc[2 * i - 2]) / (nu[iup - 1] - nu[ilow - 1]);
        for (j = ilow; j <= iup - 1; j++) gam[j - 1] = pow(10., loc[2 * i - 1] + aFit * (nu[j - 1] - nu[iup - 1]));
    }
}

/* hydro2phot :640-760 (Nussbaumer & Schmutz two-photon shape) */
static double hydro2phot(double nu, int Z, double Te, double Ne) {
    double nu0, y, aEff, Ay, gNu, Q;
    if (Z == 1) {
        nu0 = 0.7496;
        y = nu / nu0;
        if (nu >= nu0) return 0.;
        aEff = 0.8368 * pow(Te * 1.e-4, -0.723);
        Ay = 202.0 * (y * (1 - y) * (1 - pow(4 * y * (1 - y), 0.8)) +
                      0.88 * pow(y * (1 - y), 1.53) * pow(4 * y * (1 - y), 0.8));
        gNu = HPLANCK * y * Ay / 8.2249 * 1e27;
        if (Te <= 10000.) Q = 5.31e-4;
        else if (Te >= 20000.) Q = 4.71e-4;
        else Q = pow(10., log10(5.31e-4) + log10(4.71e-4 / 5.31e-4) / log10(2.) * log10(Te / 10000.));
        return aEff * gNu / (1. + (Ne * Q) / 8.23);
    }
    nu0 = 3.00;
    y = nu / nu0;
    if (nu >= nu0) return 0.;
    if (Te <= 5000.) aEff = 6.161;
    else if (Te >= 30000.) aEff = 2.035;
    else if (Te <= 10000.) aEff = pow(10., log10(6.161) + log10(4.091 / 6.161) / log10(2.) * log10(Te / 5000.));
    else if (Te <= 15000.) aEff = pow(10., log10(4.091) + log10(3.189 / 4.091) / log10(1.5) * log10(Te / 10000.));
    else aEff = pow(10., log10(3.189) + log10(2.035 / 3.189) / log10(2.) * log10(Te / 15000.));
    Ay = 64. * 0.9994667 * 202.0 * (y * (1 - y) * (1 - pow(4 * y * (1 - y), 0.8)) +
                                     0.88 * pow(y * (1 - y), 1.53) * pow(4 * y * (1 - y), 0.8));
    gNu = HPLANCK * y * Ay / (8.226 * 64.) * 1e27;
    return aEff * gNu;
}

/* RecLinesEmission :877-888: nearest temperature bin of the line tables */
static int hl_itemp(double T4) {
    int itemp = 1, i;
    for (i = 1; i <= 12; i++)
        if (T4 > HL_TEMPS[i - 1]) itemp = i;
    if (itemp < 12 && T4 > (HL_TEMPS[itemp] + HL_TEMPS[itemp - 1]) / 2.) itemp++;
    return itemp;
}

/* :915-920 (Ne > dens); outputGas's copy takes Ne >= dens (ge,
 * output_mod.f90:2054-2057) */
static int hl_idenp(const oracle_emission_tables* t, int f, double Ne, int ge) {
    const double* dens = t->hlDens + (size_t)f * 14;
    int idenp = 1, iden;
    for (iden = 1; iden <= 13; iden++)
        if ((ge ? Ne >= dens[iden - 1] : Ne > dens[iden - 1]) && dens[iden - 1] > 0.) idenp = iden;
    return idenp;
}

/* :915-938: the (density-interpolated) lines of table file f, each times Hb,
 * summed in the reference's (iup, ilow) order */
static double hl_lines(const oracle_emission_tables* t, int f, double Ne, double Hb) {
    const double* dens = t->hlDens + (size_t)f * 14;
    const double* L = t->hlLines + (size_t)f * 14 * 16 * 9;
    const int idenp = hl_idenp(t, f, Ne, 0);
    double acc = 0.;
    int iup, ilow;
#define HL(id, u, l) L[(((size_t)((id)-1)) * 16 + (u)) * 9 + (l)]
    if ((dens[12] > 0. && idenp < 13) || (dens[12] == 0. && idenp < 9)) {
        const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
        for (iup = 2; iup <= 15; iup++)
            for (ilow = 1; ilow <= (iup - 1 < 8 ? iup - 1 : 8); ilow++)
                acc += (HL(idenp, iup, ilow) + (HL(idenp + 1, iup, ilow) - HL(idenp, iup, ilow)) * w) * Hb;
    } else {
        for (iup = 2; iup <= 15; iup++)
            for (ilow = 1; ilow <= (iup - 1 < 8 ? iup - 1 : 8); ilow++) acc += HL(idenp, iup, ilow) * Hb;
    }
#undef HL
    return acc;
}

typedef struct {
    double X[31][32], ab[31];
    int on[31];
} cellcomp_t;

/* fb_ff, free_free, twoPhoton and the sums of emissionDriver :187-212.
 * w holds 12 x B doubles: eHI eHeI eHeII gHI gHeI gHeII gHv ff1 ff2 t2HI t2HeI t2HeII */
static void continuum(const oracle_emission_tables* t, const prep_t* p, const cellcomp_t* c, double Te,
                      double Ne, double* w) {
    const int B = t->bal->nbins, NS = t->bal->nstages;
    const double* nu = t->bal->nuArray;
    double *eHI = w, *eHeI = w + B, *eHeII = w + 2 * B, *gHI = w + 3 * B, *gHeI = w + 4 * B,
           *gHeII = w + 5 * B, *gHv = w + 6 * B, *ff1 = w + 7 * B, *ff2 = w + 8 * B, *t2HI = w + 9 * B,
           *t2HeI = w + 10 * B, *t2HeII = w + 11 * B;
    double *lnu = (double*)calloc(B > 0 ? B : 1, sizeof(double)), *g1 = (double*)malloc(sizeof(double) * B),
           *g2 = (double*)malloc(sizeof(double) * B);
    const double sqrTe = sqrt(Te), fourPi = 4. * PI_REF, constant = 4.9874105e-6;
    const int HIP = t->HIPnuP, HeIP = t->HeIPnuP, HeIIP = t->HeIIPnuP;
    const int heOn = c->on[2] && t->xsHeI && t->xsHeII;
    int i, j, z, ion;
    memset(w, 0, sizeof(double) * 12 * B);
    for (i = 0; i < B; i++) lnu[i] = log10(nu[i]);
    oracle_gaunt_ff(0., log10(Te), lnu, B, g1);
    oracle_gaunt_ff(0.30103, log10(Te), lnu, B, g2);
    gamma_fill(t, p, 0, Te, gHI);
    gamma_fill(t, p, 1, Te, gHeI);
    gamma_fill(t, p, 2, Te, gHeII);
    for (i = 0; i < B; i++) { /* free_free / ffCoeff :818-855 */
        const double ex = exp(-nu[i] * HCRYD_K / Te);
        ff1[i] = fourPi * 54.43 * 1. * g1[i] * ex / sqrTe;
        ff2[i] = fourPi * 54.43 * 4. * g2[i] * ex / sqrTe;
    }
    /* Milne relation above the H and He edges (:515-550) */
    for (i = HIP; i <= B; i++) {
        const double factor = nu[i - 1] * nu[i - 1] * nu[i - 1] / (Te * sqrTe);
        double ex = exp((-nu[i - 1] + nu[HIP - 1]) * HCRYD_K / Te);
        gHI[i - 1] = fourPi * t->xsHI[i - HIP] * 2. * HCRYD * constant * factor * ex * 1.e20 * 1.e20 + ff1[i - 1];
        if (heOn && i >= p->edge[1][t->nlim[1] - 1]) {
            const double sI = i >= HeIP ? t->xsHeI[i - HeIP] : 0.;
            ex = exp((-nu[i - 1] + nu[HeIP - 1]) * HCRYD_K / Te);
            gHeI[i - 1] = fourPi * sI * 0.5 * HCRYD * constant * factor * ex * 1.e20 * 1.e20 + ff1[i - 1];
            if (i >= p->edge[2][t->nlim[2] - 1]) {
                const double sII = i >= HeIIP ? t->xsHeII[i - HeIIP] : 0.;
                ex = exp((-nu[i - 1] + nu[HeIIP - 1]) * HCRYD_K / Te);
                gHeII[i - 1] = fourPi * sII * 2. * HCRYD * constant * factor * ex * 1.e20 * 1.e20 + ff2[i - 1];
            }
        }
    }
    /* heavy-ion free-bound, Z = 3..19 (:552-616) */
    for (i = 1; i <= B; i++) {
        double factor;
        for (z = 3; z <= 19; z++) {
            if (!c->on[z]) continue;
            for (ion = 1; ion <= NS - 1 && ion <= z; ion++) {
                const int k = z * 32 + ion, lo = t->fbLo[k], hi = t->fbHi[k];
                if (lo >= 1 && i >= lo && i < hi) {
                    const double ex = exp((-nu[i - 1] + nu[lo - 1]) * HCRYD_K / Te);
                    const double gr = (double)((float)t->fbG0[k] / (float)t->fbG1[k]);
                    gHv[i - 1] += ex * t->fbXs[t->fbOff[k] + (i - lo)] * gr * c->X[z][ion + 1] * c->ab[z];
                }
            }
        }
        factor = nu[i - 1] * nu[i - 1] * nu[i - 1] / (Te * sqrTe);
        gHv[i - 1] = gHv[i - 1] * fourPi * constant * HCRYD * factor * 1.e20 * 1.e20;
    }
    /* two-photon continua (:620-817) */
    for (i = 0; i < B; i++) {
        t2HI[i] = hydro2phot(nu[i], 1, Te, Ne);
        t2HeII[i] = hydro2phot(nu[i], 2, Te, Ne);
    }
    {
        const double aEff = 6.23 * pow(Te / 10000., -0.827);
        const double* yd = t->hei2phot; /* [41][2] */
        j = 1;
        for (i = 1; i <= B; i++) {
            const double y = nu[i - 1] / 1.514;
            double f1, f2, Ay;
            if (y > yd[2 * (j - 1)]) j = j + 1;
            if (j >= 41) break;
            f1 = (yd[2 * j + 1] - yd[2 * (j - 1) + 1]) / (yd[2 * j] - yd[2 * (j - 1)]);
            f2 = yd[2 * (j - 1) + 1] - yd[2 * (j - 1)] * f1;
            Ay = f1 * y + f2;
            if (Ay < 9.2163086E-03) Ay = 0.;
            t2HeI[i - 1] = aEff * Ay * 0.66262 * y / 51.3;
        }
    }
    for (i = 0; i < B; i++) { /* emissionDriver :187-212 */
        eHI[i] = (gHI[i] + t2HI[i]) * c->ab[1] * c->X[1][2] * Ne + gHv[i] * Ne;
        eHeI[i] = heOn ? (gHeI[i] + t2HeI[i]) * c->ab[2] * c->X[2][2] * Ne : 0.;
        eHeII[i] = heOn ? (gHeII[i] + t2HeII[i]) * c->ab[2] * c->X[2][3] * Ne : 0.;
        if (eHI[i] < 1.e-15) eHI[i] = 0.;
        if (eHeI[i] < 1.e-15) eHeI[i] = 0.;
        if (eHeII[i] < 1.e-15) eHeII[i] = 0.;
    }
    free(lnu);
    free(g1);
    free(g2);
}

/* RecLinesEmission :856-1153, summed as setDiffusePDF :1772-1782 */
/* one line of table file f interpolated in density, times Hb */
static double hl_line(const oracle_emission_tables* t, int f, double Ne, double Hb, int iup, int ilow,
                      int ge) {
    const double* dens = t->hlDens + (size_t)f * 14;
    const double* L = t->hlLines + (size_t)f * 14 * 16 * 9;
    const int idenp = hl_idenp(t, f, Ne, ge);
#define HL(id, u, l) L[(((size_t)((id)-1)) * 16 + (u)) * 9 + (l)]
    if ((dens[12] > 0. && idenp < 13) || (dens[12] == 0. && idenp < 9)) {
        const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
        return (HL(idenp, iup, ilow) + (HL(idenp + 1, iup, ilow) - HL(idenp, iup, ilow)) * w) * Hb;
    }
    return HL(idenp, iup, ilow) * Hb;
#undef HL
}

/* He I recombination line i (0-based, the order of HeIrecLines.dat) per
 * n_e n(He+) [erg cm^3/s] (:1110-1152; Benjamin, Skillman & Smits 1999 fits),
 * the T4 clamps as written there, the density bin from the grid's n_e */
static double hei_line_coeff(const oracle_emission_tables* t, int i, double Te, double Ne, double NeGrid) {
    const double* c0 = t->heiRec + (size_t)i * 12;
    double T4h = Te / 10000.;
    int denint;
    if (T4h < 0.5) T4h = 0.5;
    if (T4h < 2.0) T4h = 2.0;
    if (NeGrid <= 100.) denint = 0;
    else if (NeGrid <= 1.e4) denint = 1;
    else if (NeGrid <= 1.e6) denint = 2;
    else denint = 3;
#define HC(d, k) c0[((d)-1) * 4 + (k)-1]
    if (denint > 0 && denint < 3) {
        const double x1 = HC(denint, 1) * pow(T4h, HC(denint, 2)) * exp(HC(denint, 3) / T4h);
        const double x2 = HC(denint + 1, 1) * pow(T4h, HC(denint + 1, 2)) * exp(HC(denint + 1, 3) / T4h);
        return x1 + ((x2 - x1) * (Ne - pow(100., denint)) / (pow(100., denint + 1) - pow(100., denint)));
    }
    if (denint == 0) return HC(1, 1) * pow(T4h, HC(1, 2)) * exp(HC(1, 3) / T4h);
    return HC(3, 1) * pow(T4h, HC(3, 2)) * exp(HC(3, 3) / T4h);
#undef HC
}

double oracle_hei_line(const oracle_emission_tables* t, int i, double Te, double Ne) {
    return hei_line_coeff(t, i, Te, Ne, Ne);
}

/* RecLinesEmission (:856-1153) summed; hbeta (may be NULL) receives
 * hydroLines(1,4,2) */
static double rec_lines(const oracle_emission_tables* t, const cellcomp_t* c, double Te, double Ne,
                        double NeGrid, double log10Ne, double* hbeta, int ge) {
    const int NS = t->bal->nstages;
    const double T4 = Te / 10000.;
    const int elUp = T4 > HL_TEMPS[11] ? 2 : 8;
    double normRec = 0.;
    int izp, i;
    for (izp = 1; izp <= elUp; izp++) {
        int f, idenp;
        double Hb;
        if (!(c->on[izp] && NS > izp)) continue;
        f = t->hlMap[(izp - 1) * 12 + (hl_itemp(T4) - 1)];
        idenp = hl_idenp(t, f, Ne, ge);
        if ((izp == 1 && T4 <= 3.) || (izp > 1 && T4 <= HL_TEMPS[11])) {
            const double* A = t->hbA + 4 * (izp - 1);
            const double* Bc = t->hbB + 4 * (izp - 1);
            double zf = (log10Ne - A[1]) / A[2], Af, Bf;
            Af = A[0] * exp(-zf * zf / 2.) + A[3];
            zf = (log10Ne - Bc[1]) / Bc[2];
            Bf = Bc[0] * exp(-zf * zf / 2.) + Bc[3];
            Hb = pow(10., Af + Bf * log10(Te));
        } else {
            const double* dens = t->hlDens + (size_t)f * 14;
            const double xx = log10(T4 * 1.e4);
            double yy1, yy2, x1, x2;
            if (izp == 1) {
                const double* r = t->rbEdge + 2 * 26; /* rbEdge(1, 3, :) */
                if (idenp < 13) {
                    const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
                    yy1 = r[idenp - 1] + (r[idenp - 1] - r[idenp]) * w;
                    yy2 = r[13 + idenp - 1] + (r[13 + idenp - 1] - r[13 + idenp]) * w;
                } else {
                    yy1 = r[idenp - 1];
                    yy2 = r[13 + idenp - 1];
                }
                x1 = t->rbEdge[0];
                x2 = t->rbEdge[13];
            } else {
                const double* r = t->r2aEdge + 2 * 18; /* r2aEdge(3, :) */
                const int a9 = 9 + idenp < 18 ? 9 + idenp : 18, b9 = 10 + idenp < 18 ? 10 + idenp : 18;
                if (idenp < 13) {
                    const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
                    yy1 = r[idenp - 1] + (r[idenp - 1] - r[idenp]) * w;
                    yy2 = r[a9 - 1] + (r[a9 - 1] - r[b9 - 1]) * w;
                } else {
                    yy1 = r[idenp - 1];
                    yy2 = r[a9 - 1];
                }
                x1 = t->r2aEdge[0];
                x2 = t->r2aEdge[9];
            }
            Hb = pow(10., yy1 - (yy1 - yy2) * (xx - x1) / (x2 - x1));
        }
        Hb = Hb * Ne * c->X[izp][izp + 1] * c->ab[izp];
        normRec += hl_lines(t, f, Ne, Hb);
        if (izp == 1 && hbeta) *hbeta = hl_line(t, f, Ne, Hb, 4, 2, ge);
    }
    for (izp = elUp + 1; izp <= 30; izp++) {
        double T4z, lTeZ, Nez, lNez, Hb, zf, Af, Bf;
        int f;
        if (!(c->on[izp] && NS > izp)) continue;
        T4z = T4 * 4. / (double)(izp * izp);
        lTeZ = log10(Te * 4. / (double)(izp * izp));
        Nez = Ne * 4. / (double)(izp * izp);
        lNez = log10(Ne * 4. / (double)(izp * izp));
        f = t->hlMap[1 * 12 + (hl_itemp(T4z) - 1)];
        zf = (lNez - t->hbA[4 + 1]) / t->hbA[4 + 2];
        Af = t->hbA[4 + 0] * exp(-zf * zf / 2.) + t->hbA[4 + 3];
        zf = (lNez - t->hbB[4 + 1]) / t->hbB[4 + 2];
        Bf = t->hbB[4 + 0] * exp(-zf * zf / 2.) + t->hbB[4 + 3];
        Hb = ((double)(izp * izp * izp) / 8.) * pow(10., Af + Bf * lTeZ);
        Hb = Hb * Ne * c->X[izp][izp + 1] * c->ab[izp];
        normRec += hl_lines(t, f, Nez, Hb);
    }
    /* He I recombination lines (:1110-1152) */
    if (c->on[2])
        for (i = 0; i < 34; i++) normRec += hei_line_coeff(t, i, Te, Ne, NeGrid) * Ne * c->ab[2] * c->X[2][2];
    return normRec;
}

/* forLines :1481-1526 with equilibrium(rec=.true.) summed as setDiffusePDF
 * :1784-1799 (lines above 1e-35 in units of 1e-25 erg/s per n_H) */
static double for_lines(const oracle_emission_tables* t, const cellcomp_t* c, double Te, double Ne) {
    const oracle_balance_tables* b = t->bal;
    const int NS = b->nstages;
    double normFor = 0.;
    int z, ion;
    for (z = 3; z <= 30; z++) {
        const int top = z + 1 < NS ? z + 1 : NS;
        if (!c->on[z]) continue;
        for (ion = 1; ion <= top; ion++) {
            const int a = b->atom_idx[z * 32 + ion];
            double up = 0.;
            if (a < 0) continue;
            if (ion < top && c->X[z][ion] > 0.) up = c->X[z][ion + 1] / c->X[z][ion];
            normFor += oracle_nlevel_lines(b, a, Te, Ne, up, 1, c->ab[z] * c->X[z][ion], 1.9865e9);
        }
    }
    return normFor;
}

static void cell_comp(const oracle_balance_tables* b, const double* ionDen, size_t stride, size_t cell,
                      cellcomp_t* c) {
    int e, ion;
    memset(c, 0, sizeof(*c));
    for (e = 0; e < b->nElemUsed; e++) {
        const int z = b->elemZ[e];
        c->on[z] = 1;
        c->ab[z] = b->abund[e];
        for (ion = 1; ion <= b->nstages; ion++)
            c->X[z][ion] = ionDen[((size_t)(ion - 1) * b->nElemUsed + e) * stride + cell];
    }
}

static void emission_one(const oracle_emission_tables* t, const prep_t* p, oracle_emission_state* s,
                         int cell, double* w) {
    const int B = t->bal->nbins, C1 = s->nCells + 1;
    const double NeGrid = s->Ne[cell], Te = s->Te[cell];
    const double log10Ne = log10(NeGrid);
    const double Ne = NeGrid < 1.e-5 ? 1.e-5 : NeGrid;
    const double *eHI = w, *eHeI = w + B, *eHeII = w + 2 * B;
    double *sHI = w + 12 * B, *sHeI = w + 13 * B, *sHeII = w + 14 * B, *sDust = w + 15 * B;
    double normDust = 0.;
    double normHI = 0., normHeI = 0., normHeII = 0., normRec, normFor, normalize, lines;
    double* row = s->recPDF + (size_t)cell * B;
    cellcomp_t c;
    int i;
    cell_comp(t->bal, s->ionDen, (size_t)C1, (size_t)cell, &c);
    continuum(t, p, &c, Te, Ne, w);
    normRec = rec_lines(t, &c, Te, Ne, NeGrid, log10Ne, NULL, 0);
    normFor = for_lines(t, &c, Te, Ne);
    /* setDiffusePDF :1593-1693 */
    for (i = 0; i < B; i++) {
        sHI[i] = CRYD * t->widFlx[i] * eHI[i] / 1.e15;
        normHI += sHI[i];
        sHeI[i] = CRYD * t->widFlx[i] * eHeI[i] / 1.e15;
        normHeI += sHeI[i];
        sHeII[i] = CRYD * t->widFlx[i] * eHeII[i] / 1.e15;
        normHeII += sHeII[i];
    }
    if (c.on[2]) {
        const double T4 = Te * 1.e-4;
        const double corr = 27.2 * pow(T4, -0.678) * Ne * c.ab[2] * c.X[2][2] * 1.45673 * HCRYD * 1e11;
        double ly[4];
        int k;
        sHeI[p->j2TsP - 1] += corr;
        normHeI += corr;
        for (k = 0; k < 4; k++) ly[k] = t->heiiLyman[2 * k];
        ly[3] = pow(10., -0.792 * log10(Te) + 6.01) * Ne * c.X[2][3] * c.ab[2];
        for (k = 0; k < 3; k++) ly[k] = ly[k] * ly[3];
        for (k = 0; k < 4; k++) {
            if (p->lymP[k] < 1) continue;
            sHeII[p->lymP[k] - 1] += ly[k];
            normHeII += ly[k];
        }
    }
    if (s->dustS > 0) { /* dust emission (:1685-1752): sizes outer, species inner */
        const double cst = 6.6262e-27 * 4. * 3.141592654 * 1.e25 * CRYD;
        const int nS = s->dustS, nA = s->dustA;
        int ai, sp;
        for (i = 0; i < B; i++) sDust[i] = 0.;
        for (ai = 1; ai <= nA; ai++)
            for (sp = 1; sp <= nS; sp++) {
                const double T = s->Tdust[((size_t)cell * (nA + 1) + ai) * (nS + 1) + sp];
                const double* c = s->dustCabs + ((size_t)(sp - 1) * nA + (ai - 1)) * B;
                const int q = s->qIdx ? s->qIdx[(sp - 1) * nA + ai - 1] : -1;
                if (!(T < s->dustTsub[sp - 1])) continue;
                for (i = 0; i < B; i++) {
                    if (q >= 0) {  /* quantum heating (:1701-1730) */
                        const double* qt = s->qT + ((size_t)cell * s->nQ + q) * s->nTb;
                        const double* qp = s->qP + ((size_t)cell * s->nQ + q) * s->nTb;
                        int iT;
                        for (iT = 0; iT < s->nTb; iT++) {
                            const double v = c[i] * oracle_planck_h(t->bal->nuArray[i], qt[iT]) * t->widFlx[i] *
                                             s->dustWeight[ai - 1] * s->dustAbun[sp - 1] * qp[iT];
                            sDust[i] = sDust[i] + v;
                            normDust = normDust + v;
                        }
                        continue;
                    }
                    {
                        const double v = c[i] * oracle_planck_h(t->bal->nuArray[i], T) * t->widFlx[i] *
                                         s->dustWeight[ai - 1] * s->dustAbun[sp - 1];
                        sDust[i] = sDust[i] + v;
                        normDust = normDust + v;
                    }
                }
            }
        for (i = 0; i < B; i++) sDust[i] = sDust[i] * cst * s->Ndust[cell] / s->Hden[cell];
        normDust = normDust * cst * s->Ndust[cell] / s->Hden[cell];
    }
    normalize = normHI + normHeI + normHeII;
    if (s->dustS > 0) normalize = normalize + normDust;
    /* :1804-1835 cumulative sum, normalised, last entry 1 */
    if (s->dustS > 0) {
        row[0] = sHI[0] + sHeI[0] + sHeII[0] + sDust[0];
        for (i = 1; i < B; i++) row[i] = row[i - 1] + sHI[i] + sHeI[i] + sHeII[i] + sDust[i];
    } else {
        row[0] = sHI[0] + sHeI[0] + sHeII[0];
        for (i = 1; i < B; i++) row[i] = row[i - 1] + (sHI[i] + sHeI[i] + sHeII[i]);
    }
    {
        const double last = row[B - 1];
        for (i = 0; i < B; i++) row[i] = row[i] / last;
    }
    row[B - 1] = 1.;
    lines = normRec + normFor;
    s->totalLines[cell] = lines / (normalize + lines);
    if (s->normCont) s->normCont[cell] = normalize;
    if (s->normRec) s->normRec[cell] = normRec;
    if (s->normFor) s->normFor[cell] = normFor;
}

int oracle_emission(const oracle_emission_tables* t, oracle_emission_state* s, int nthreads) {
    prep_t p;
    if (prep_make(t, &p) != 0) return -1;
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 1 ? nthreads : 1)
#endif
    {
        double* w = (double*)malloc(sizeof(double) * 16 * (size_t)t->bal->nbins);
        int cell;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (cell = 1; cell <= s->nCells; cell++) emission_one(t, &p, s, cell, w);
        free(w);
    }
    prep_free(&p);
    return 0;
}

int oracle_emission_coeffs(const oracle_emission_tables* t, double Te, double Ne, const double* X,
                           double* out) {
    prep_t p;
    cellcomp_t c;
    if (prep_make(t, &p) != 0) return -1;
    cell_comp(t->bal, X, 1, 0, &c);
    continuum(t, &p, &c, Te, Ne < 1.e-5 ? 1.e-5 : Ne, out);
    prep_free(&p);
    return 0;
}

/* ---- line luminosities of outputGas (output_mod.f90:397-432) ------------- */
int oracle_line_luminosities(const oracle_emission_tables* t, const oracle_emission_state* s,
                             const oracle_output_in* g, int L, double* lines, double* hbeta) {
    const oracle_balance_tables* b = t->bal;
    const int NS = b->nstages;
    const size_t C1 = (size_t)s->nCells + 1;
    double* em = (double*)calloc((size_t)L * L, sizeof(double));
    int i, j, k, z, ion, q;
    if (!em) return -1;
    for (q = 0; q < 31 * 32 * L * L; q++) lines[q] = 0.;
    *hbeta = 0.;
    for (i = 0; i < g->nx; i++)
        for (j = 0; j < g->ny; j++)
            for (k = 0; k < g->nz; k++) {
                const int cell = g->active[(size_t)i + (size_t)g->nx * ((size_t)j + (size_t)g->ny * k)];
                cellcomp_t c;
                double Te, NeGrid, Ne, dV, H, hb = 0.;
                if (cell <= 0) continue;
                if (g->lgBlack && g->lgBlack[cell] >= 1) continue;
                cell_comp(b, s->ionDen, C1, (size_t)cell, &c);
                Te = s->Te[cell];
                NeGrid = s->Ne[cell];
                Ne = NeGrid; /* outputGas uses the grid's n_e unfloored (output_mod.f90:321) */
                H = g->Hden[cell];
                dV = oracle_cell_width(g->xAxis, g->nx, i, g->sym, g->plane) *
                     oracle_cell_width(g->yAxis, g->ny, j, g->sym, g->plane) *
                     oracle_cell_width(g->zAxis, g->nz, k, g->sym, g->plane);
                (void)rec_lines(t, &c, Te, Ne, NeGrid, log10(NeGrid), &hb, 1);
                if (hb * H * dV > 1.e-35) *hbeta += hb * H * dV;
                for (z = 3; z <= 30; z++) {
                    const int top = z + 1 < NS ? z + 1 : NS;
                    if (!c.on[z]) continue;
                    for (ion = 1; ion <= top; ion++) {
                        const int a = b->atom_idx[z * 32 + ion];
                        double up = 0.;
                        double* out = lines + (size_t)(z * 32 + ion) * L * L;
                        if (a < 0) continue;
                        if (ion < top && c.X[z][ion] > 0.) up = c.X[z][ion + 1] / c.X[z][ion];
                        for (q = 0; q < L * L; q++) em[q] = 0.;
                        oracle_nlevel_line_em(b, a, Te, Ne, up, 1, c.ab[z] * c.X[z][ion], 1.9865e9, em, L);
                        for (q = 0; q < L * L; q++) out[q] += em[q] * H * dV;
                    }
                }
            }
    free(em);
    return 0;
}

/*
 * moc_oracle_opacity.c -- ORACLE (test infrastructure only): FP64 restatement
 * of the per-cell opacity of MOCASSIN, ionization_mod.f90:
 *   ionizationDriver :26-131 (ion densities), BoltGaunt :136-176,
 *   getGauntFF :196-336 (Hummer 1988 Chebyshev fit), eDenSum :338-351,
 *   addOpacity :353-477 (free-free, bound-free per shell, Klein-Nishina).
 * One deliberate divergence: the free-free Gaunt factors and Boltzmann
 * factors are evaluated at each cell's own T_e (the reference caches them
 * across cells while T_e stays within 10 %, ionization_mod.f90:140,172-177).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "moc_oracle.h"

#define TE1RYD 1.578866e5 /* constants_mod.f90:25 */

/* Hummer (1988) coefficients d(8,11), column-major as in getGauntFF */
static const double HUMMER_D[88] = {
    8.986940175e+00, -4.009515855e+00, 8.808871266e-01, 2.640245111e-02, -4.580645915e-02,
    -3.568055702e-03, 2.827798067e-03, 3.365860195e-04, -8.006936989e-01, 9.466021705e-01,
    9.043402532e-02, -9.608451450e-02, -1.885629865e-02, 1.050313890e-02, 2.800889961e-03,
    -1.078209202e-03, -3.781305103e-01, 1.102726332e-01, -1.543619180e-02, 8.310561114e-03,
    2.179620525e-02, 4.259726289e-03, -4.181588794e-03, -1.770208330e-03, 1.877213132e-02,
    -1.004885705e-01, -5.483366378e-02, -4.520154409e-03, 8.366530426e-03, 3.700273930e-03,
    6.889320423e-04, 9.460313195e-05, 7.300158392e-02, 3.576785497e-03, -4.545307025e-03,
    -1.017965604e-02, -9.530211924e-03, -3.450186162e-03, 1.040482914e-03, 1.407073544e-03,
    -1.744671550e-03, 2.864013856e-02, 1.903394837e-02, 7.091074494e-03, -9.668371391e-04,
    -2.999107465e-03, -1.820642230e-03, -3.874082085e-04, -1.707268366e-02, -4.694254776e-03,
    1.311691517e-03, 5.316703136e-03, 5.178193095e-03, 2.451228935e-03, -2.277321615e-05,
    -8.182359057e-04, 2.567331664e-04, -9.155339970e-03, -6.997479192e-03, -3.571518641e-03,
    -2.096101038e-04, 1.553822487e-03, 1.509584686e-03, 6.212627837e-04, 4.098322531e-03,
    1.635218463e-03, -5.918883504e-04, -2.333091048e-03, -2.484138313e-03, -1.359996060e-03,
    -5.371426147e-05, 5.553549563e-04, 3.837562402e-05, 2.938325230e-03, 2.393747064e-03,
    1.328839809e-03, 9.135013312e-05, -7.137252303e-04, -7.656848158e-04, -3.504683798e-04,
    -8.491991820e-04, -3.615327726e-04, 3.148015257e-04, 8.909207650e-04, 9.869737522e-04,
    6.134671184e-04, 1.068883394e-04, -2.046080100e-04};

#define D(j, k) HUMMER_D[((k)-1) * 8 + ((j)-1)]

void oracle_gaunt_ff(double z, double log10Te, const double* xlf, int n, double* g) {
    double b[12], c[9], xlrkt, txg, gamma2, con;
    int i, j, ir;
    if (log10Te < 2.5) xlrkt = 5.1983649 - 2.5;
    else if (log10Te > 8.) xlrkt = 5.1983649 - 8.;
    else xlrkt = 5.1983649 - log10Te;
    txg = 0.66666667 * (2.0 * z + xlrkt);
    gamma2 = pow(10., txg * 1.5);
    con = 0.72727273 * xlrkt + 0.90909091;
    for (j = 1; j <= 8; j++) {
        ir = 9;
        b[11] = D(j, 11);
        b[10] = txg * b[11] + D(j, 10);
        for (i = 1; i <= 9; i++) {
            b[ir] = txg * b[ir + 1] - b[ir + 2] + D(j, ir);
            ir = ir - 1;
        }
        c[j] = 0.25 * (b[1] - b[3]);
    }
    for (i = 0; i < n; i++) {
        double txu = 0.72727273 * xlf[i] + con;
        double u = pow(10., (txu - .90909091) / .72727273);
        if (fabs(txu) <= 2.0) {
            ir = 6;
            b[8] = c[8];
            b[7] = txu * b[8] + c[7];
            for (j = 1; j <= 6; j++) {
                b[ir] = txu * b[ir + 1] - b[ir + 2] + c[ir];
                ir = ir - 1;
            }
            g[i] = b[1] - b[3];
        } else if (log10(u) < -4.0 && gamma2 < 0.3) {
            g[i] = 0.551329 * log(2.24593 / u);
        } else if (log10(u) < -4.0 && gamma2 >= 0.3) {
            g[i] = 0.551329 * log(0.944931 / u / sqrt(gamma2));
        } else if (txu > 2.0) {
            if (log10(gamma2) < -3.) {
                g[i] = sqrt(0.9549297 / u);
            } else if (log10(gamma2) < 0.0) {
                double slope = (sqrt(12 * gamma2 / u) - sqrt(0.9549297 / u)) / 3.0;
                g[i] = sqrt(12 * gamma2 / u) + slope * log10(gamma2);
            } else {
                g[i] = sqrt(12. * gamma2 / u);
                if (g[i] > 1.05) g[i] = 1.05;
            }
        } else {
            g[i] = 0.0; /* not reached: |txu| > 2 with txu < -2 means log10(u) < -4 */
        }
    }
}

int oracle_opacity(const oracle_opacity_in* in, double* opacity) {
    const int B = in->nbins, C1 = in->nCells + 1, NE = in->nElemUsed, NS = in->nstages;
    double *lognu = malloc(sizeof(double) * B), *gff = malloc(sizeof(double) * B),
           *boltz = malloc(sizeof(double) * B), *dens = malloc(sizeof(double) * 31 * 32);
    int cell, b, e, i, s;
    for (b = 0; b < B; b++) lognu[b] = log10(in->nuArray[b]);
    for (b = 0; b < B; b++) opacity[(size_t)b * C1] = 0.0; /* dummy cell 0 */
    for (cell = 1; cell < C1; cell++) {
        const double Te = in->Te[cell], Ne = in->Ne[cell], H = in->Hden[cell];
        double eDenFFSum = 0., fac1, fac2;
#define DENS(z, ion) dens[(z)*32 + (ion)]
        memset(dens, 0, sizeof(double) * 31 * 32);
        for (e = 0; e < NE; e++) {
            int z = in->elemZ[e];
            /* density(n, i) for i <= min(n, nstages) only (:63-75): the bare
             * nucleus (H+, He++, ...) never enters, so the free-free factor
             * below sees density(1,2) = density(2,3) = 0 as in the reference */
            for (i = 1; i <= (z < NS ? z : NS); i++)
                DENS(z, i) = in->ionDen[((size_t)(i - 1) * NE + e) * C1 + cell] * in->abund[e] * H;
        }
        /* BoltGaunt (:136-176): contBoltz = exp(-(Te1Ryd/Te + nu)) while <= 84 */
        {
            int stop = 0;
            for (b = 0; b < B; b++) {
                double ex = TE1RYD / Te + in->nuArray[b];
                if (ex > 84.) stop = 1;
                boltz[b] = stop ? 0. : exp(-ex);
            }
        }
        oracle_gaunt_ff(0., log10(Te), lognu, B, gff);
        /* eDenSum (:338-351) */
        for (i = 3; i <= 30; i++) {
            int k;
            for (k = 1; k <= (i < NS - 1 ? i : NS - 1); k++) eDenFFSum += (double)(k * k) * DENS(i, k + 1);
        }
        fac1 = (Ne / 1.e10) * ((DENS(1, 2) + DENS(2, 2) + 4. * DENS(2, 3) + eDenFFSum) / 1.e10) / sqrt(Te);
        fac2 = fac1 * TE1RYD / Te;
        for (b = 0; b < B; b++) {
            double ff;
            if (boltz[b] < 0.995) ff = in->brems[b] * (1. - boltz[b]) * gff[b] * fac1;
            else ff = in->brems[b] * in->nuArray[b] * gff[b] * fac2;
            opacity[(size_t)b * C1 + cell] = ff;
        }
        /* bound-free per shell (inOpacity with b = 0) */
        for (s = 0; s < in->nShells; s++) {
            double den = DENS(in->shZ[s], in->shIon[s]);
            int lo = in->shLo[s], hi = in->shHi[s] < B ? in->shHi[s] : B;
            if (hi < lo) hi = lo;
            for (b = lo; b <= hi; b++)
                opacity[(size_t)(b - 1) * C1 + cell] += in->xs[in->shOff[s] + (b - lo)] * den;
        }
        if (in->lgCompton)
            for (b = 0; b < B; b++) opacity[(size_t)b * C1 + cell] += in->KNsigmaT[b] * Ne;
#undef DENS
    }
    free(lognu);
    free(gff);
    free(boltz);
    free(dens);
    return 0;
}

/*
 * ORACLE -- test infrastructure only.  Plain-C FP64 restatement of MOCASSIN's
 * grain routines, used by tests/ as the checker of the device path; nothing
 * in the product links it.
 *
 *   oracle_planck_h         getFlux('blackbody')     continuum_mod.f90:358-400
 *   oracle_dust_em_integral dustEmissionInt          dust_mod.f90:146-184
 *   oracle_dust_opacity     dust block of iterateMC  iteration_mod.f90:194-246
 *   oracle_dust_pdf         setDustPDF               emission_mod.f90:1921-1986
 *   oracle_dust_temperature getDustT                 update_mod.f90:3346-3453
 *                           (+ dust-only convergence :240-267)
 *
 * Pins: the Planck function against its closed-form limits (Wien and
 * Rayleigh-Jeans) and the Stefan-Boltzmann integral; the emission integral
 * of a grey grain against 4 pi Cabs sigma T^4 / pi (tests/test_dust.py);
 * getDustT against the inverse of the emission-integral table on a field
 * of known T (a dilute blackbody for which Kirchhoff's law fixes T); the
 * opacity and PDF sums against their definitions on hand-made inputs, and
 * against the sublimation switch.  The quantum-heating branches (lgQHeat)
 * are not restated.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "moc_oracle_dust.h"

#define ORC_PI 3.141592654
#define ORC_FR1RYD 3.28984e15
#define ORC_HPLANCK 6.6262e-27

double oracle_planck_h(double nu, double T) {
    const double constant = 0.5250229 / ORC_HPLANCK;
    const double x = 157893.94 * nu / T;
    double den;
    if (x > 86.) return constant * nu * nu * nu * exp(-x);
    den = exp(x) - 1.;
    if (den <= 0.) return 3.32154e-6 * nu * nu * T / ORC_HPLANCK;
    return constant * nu * nu * nu / den;
}

void oracle_dust_em_integral(const double* Cabs, const double* nu, const double* widFlx, int B,
                             const double* temps, int nT, double* out) {
    int t, i;
    for (t = 0; t < nT; t++) {
        double s = 0.;
        for (i = 0; i < B; i++) s += Cabs[i] * oracle_planck_h(nu[i], temps[t]) * ORC_FR1RYD * widFlx[i];
        out[t] = s * ORC_HPLANCK * 4.;
    }
}

static size_t tdust_at(const oracle_dust_in* in, int cell, int s, int ai) {
    return ((size_t)cell * (in->nA + 1) + ai) * (in->nS + 1) + s;
}

void oracle_dust_opacity(const oracle_dust_in* in, const double* Tdust, const double* gasOpac,
                         double* opacity, double* absOpac, double* scaOpac) {
    const size_t C1 = (size_t)in->nCells + 1;
    int cell, b, s, ai;
    for (b = 0; b < in->B; b++) {
        for (cell = 1; cell <= in->nCells; cell++) {
            double sca = 0., ab = 0.;
            const size_t t = (size_t)b * C1 + cell;
            for (s = 1; s <= in->nS; s++)
                for (ai = 1; ai <= in->nA; ai++) {
                    size_t k;
                    if (!(Tdust[tdust_at(in, cell, s, ai)] < in->tsub[s - 1])) continue;
                    k = ((size_t)(s - 1) * in->nA + (ai - 1)) * in->B + b;
                    sca = sca + in->abun[s - 1] * in->weight[ai - 1] * in->Ndust[cell] * in->Csca[k];
                    ab = ab + in->abun[s - 1] * in->weight[ai - 1] * in->Ndust[cell] * in->Cabs[k];
                }
            scaOpac[t] = sca;
            absOpac[t] = ab;
            opacity[t] = (gasOpac ? gasOpac[t] : 0.) + (sca + ab);
        }
        if (!gasOpac) opacity[(size_t)b * C1] = 0.;
    }
}

void oracle_dust_pdf(const oracle_dust_in* in, const double* Tdust, double* pdf) {
    int cell, i, s, ai;
    const int B = in->B;
    for (cell = 1; cell <= in->nCells; cell++) {
        double* row = pdf + (size_t)cell * B;
        for (i = 0; i < B; i++) row[i] = 0.;
        for (s = 1; s <= in->nS; s++)
            for (ai = 1; ai <= in->nA; ai++) {
                const double T = Tdust[tdust_at(in, cell, s, ai)];
                const double* c = in->Cabs + ((size_t)(s - 1) * in->nA + (ai - 1)) * B;
                if (!(T > 0. && T < in->tsub[s - 1])) continue;
                for (i = 0; i < B; i++)
                    row[i] = row[i] + c[i] * oracle_planck_h(in->nu[i], T) * in->widFlx[i] *
                                          in->weight[ai - 1] * in->abun[s - 1];
            }
        for (i = 1; i < B; i++) row[i] = row[i - 1] + row[i];
        {
            const double total = row[B - 1];
            for (i = 0; i < B; i++) row[i] = total > 0. ? row[i] / total : 0.;
            row[B - 1] = 1.;
        }
    }
}

/* locate (interpolation_mod.f90:49-106) on an increasing table, 1-based */
static int locate_inc(const double* xa, int n, double x) {
    int lo = 0, hi = n;
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (xa[m] > x) hi = m; else lo = m + 1;
    }
    return lo;
}

int oracle_dust_temperature(const oracle_dust_in* in, const double* J, double* Tdust,
                            int32_t* lgConverged, int32_t* lgBlack, double counters[2]) {
    const size_t C1 = (size_t)in->nCells + 1;
    int cell, s, ai, i;
    counters[0] = counters[1] = 0.;
    for (cell = 1; cell <= in->nCells; cell++) {
        double* Td = Tdust + tdust_at(in, cell, 0, 0);
        double old;
        int hit = 0;
        for (i = 0; i < in->B; i++)
            if (J[(size_t)i * C1 + cell] > 0.) { hit = 1; break; }
        if (!hit) {  /* updateCell returns with the grain temperatures kept (:158-171) */
            if (in->dustOnly) {
                lgBlack[cell] = 1;
                lgConverged[cell] = 0;
                counters[0] += 1.;
            }
            continue;
        }
        old = Td[0];
        for (i = 0; i < (in->nS + 1) * (in->nA + 1); i++) Td[i] = 0.;
        for (s = 1; s <= in->nS; s++) {
            for (ai = 1; ai <= in->nA; ai++) {
                const size_t g = (size_t)(s - 1) * in->nA + (ai - 1);
                const double* c = in->Cabs + g * in->B;
                const double* E = in->emInt + g * in->nT;
                double x = 0., T;
                int iT;
                for (i = 0; i < in->B; i++) x = x + c[i] * (J[(size_t)i * C1 + cell] / ORC_PI);
                if (in->collD)  /* gas-grain collisions and photoelectric cooling (:3390-3393) */
                    x = x + in->collD[g * C1 + cell] - in->peD[g * C1 + cell];
                iT = locate_inc(E, in->nT, x);
                if (iT <= 0) T = 1.;
                else if (iT >= in->nT) T = (double)in->nT;
                else T = (double)iT + (x - E[iT - 1]) * 1. / (E[iT] - E[iT - 1]);
                Td[(size_t)ai * (in->nS + 1) + s] = T;
                Td[s] = Td[s] + T * in->weight[ai - 1];
            }
            Td[0] = Td[0] + Td[s] * in->abun[s - 1];
        }
        if (in->dustOnly) {
            const int conv = fabs((Td[0] - old) / old) <= in->XHILimit;
            lgBlack[cell] = 0;
            lgConverged[cell] = conv;
            if (conv) counters[1] += 1.;
        }
    }
    return 0;
}

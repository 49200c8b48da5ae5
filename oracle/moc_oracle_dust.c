/*
 * ORACLE -- test infrastructure only.  Plain-C FP64 restatement of MOCASSIN's
 * grain routines, used by tests/ to check the device path; nothing
 * in the product links it.
 *
 *   oracle_planck_h         getFlux('blackbody')     continuum_mod.f90:358-400
 *   oracle_dust_em_integral dustEmissionInt          dust_mod.f90:146-184
 *   oracle_dust_opacity     dust block of iterateMC  iteration_mod.f90:194-246
 *   oracle_dust_pdf         setDustPDF               emission_mod.f90:1921-1986
 *   oracle_dust_temperature getDustT                 update_mod.f90:3346-3453
 *                           (+ dust-only convergence :240-267)
 *   oracle_qheat            qHeat, getTmin, enthalpy, bbody, clrate, cheat
 *                                                    emission_mod.f90:1993-2504
 *
 * Pins: the Planck function against its closed-form limits (Wien and
 * Rayleigh-Jeans) and the Stefan-Boltzmann integral; the emission integral
 * of a grey grain against 4 pi Cabs sigma T^4 / pi (tests/test_dust.py);
 * getDustT against the inverse of the emission-integral table on a field
 * of known T (a dilute blackbody for which Kirchhoff's law fixes T); the
 * opacity and PDF sums against their definitions on hand-made inputs, and
 * against the sublimation switch.  qHeat: the distributions sum to 1, a
 * large grain in a thermal field sits at the field's temperature, the energy
 * absorbed equals the energy emitted by the distribution (to the few per
 * cent of the binning) in thermal and dilute hot fields, and the mean
 * temperature rises with the field (tests/test_qheat.py).
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
                const int q = in->qIdx ? in->qIdx[(s - 1) * in->nA + ai - 1] : -1;
                if (!(T > 0. && T < in->tsub[s - 1])) continue;
                if (q >= 0) {  /* quantum heating (:1946-1960) */
                    const double* qt = in->qT + ((size_t)cell * in->nQ + q) * in->nTb;
                    const double* qp = in->qP + ((size_t)cell * in->nQ + q) * in->nTb;
                    int iT;
                    for (i = 0; i < B; i++)
                        for (iT = 0; iT < in->nTb; iT++)
                            row[i] = row[i] + c[i] * oracle_planck_h(in->nu[i], qt[iT]) * in->widFlx[i] *
                                                  in->weight[ai - 1] * in->abun[s - 1] * qp[iT];
                    continue;
                }
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

/* ---------------------------------------------------------------- qHeat
 * Quantum heating of a small grain (qHeat, emission_mod.f90:1993-2264, after
 * Guhathakurta & Draine 1989), with getTmin (:2267-2320), enthalpy
 * (:2326-2375), bbody (:2381-2396), clrate (:2398-2433) and cheat
 * (:2436-2500).  All arithmetic in FP64 with the reference's constants.
 * Kept from the reference: getTmin zeroes the field below its 9.11e-5 Ryd
 * cutoff before it integrates over those same bins, so the lowest
 * temperature is always the emission table's first, 1 K; the last bin's
 * probability is left 0; the 1e250 rescaling of the recurrence.  Guarded
 * where the reference reads past its arrays: the wavelength searches clamp to
 * the last mesh point, and the heating of bin if+2 from the cooling of bin
 * if is skipped when if+2 lies past the last bin. */
#define QH_C 2.9979250e10
#define QH_CUT 0.1

static double qh_enthalpy(double tg, int natom, double radius, int kind) {
    const double enth1 = (1.4e3 / 3.) * pow(tg, 3.);
    const double enth1_50 = (1.4e3 / 3.) * pow(50., 3.);
    const double enth2 = (2.1647e4 / 2.3) * (pow(tg, 2.3) - pow(50., 2.3));
    const double enth2_150 = (2.1647e4 / 2.3) * (pow(150., 2.3) - pow(50., 2.3));
    const double enth3 = (4.8375e5 / 1.68) * (pow(tg, 1.68) - pow(150., 1.68));
    const double enth3_500 = (4.8375e5 / 1.68) * (pow(500., 1.68) - pow(150., 1.68));
    const double enth4 = 3.3107e7 * (tg - 500.);
    const double vg = 4.18878 * 1.e-12 * radius * radius * radius;
    double e;
    if (kind == 1) {
        if (tg <= 50.) e = enth1;
        else if (tg <= 150.) e = enth1_50 + enth2;
        else if (tg <= 500.) e = enth1_50 + enth2_150 + enth3;
        else e = enth1_50 + enth2_150 + enth3_500 + enth4;
        return (1. - (2. / (double)natom)) * vg * e;
    }
    e = (double)natom * (4.15e-22 * pow(tg, 3.3));
    e = e / (1. + 6.51e-3 * tg + 1.5e-6 * tg * tg + 8.3e-7 * pow(tg, 2.3));
    return (1. - (2. / (double)natom)) * e;
}

static double qh_bbody(double lam, double T) {
    const double C1 = 3.74185e-5, C2 = 1.43883;
    const double x = C2 / (lam * T);
    if (x > 400.) return 0.;
    return C1 * (1. / pow(lam, 5.)) / (exp(x) - 1.);
}

/* last 0-based index with lam < x, or -1 */
static int qh_below(const double* lam, int B, double x) {
    int k = B - 1;
    while (k >= 0 && !(lam[k] < x)) --k;
    return k;
}

static double qh_clrate(double tem, const double* lam, const double* fl, const double* sQ, int B) {
    double engout = 0., prev = 0.;
    int k, cut = qh_below(lam, B, QH_CUT);
    if (cut < 0) cut = 0;
    for (k = 0; k < B; k++) {
        const double e = sQ[k] * qh_bbody(lam[k], tem);
        if (k >= 1) engout = engout + 4. * ((e + prev) / 2.) * (lam[k] - lam[k - 1]);
        prev = e;
    }
    return engout - 0.25 * fl[cut] * sQ[cut] * QH_CUT;
}

/* it: 1-based bin */
static double qh_cheat(int it, int n, const double* lam, const double* fl, const double* sQ, int B,
                       const double* U, double hc) {
    double einmax, shwav, engin = 0.;
    int k, nshort, nlong;
    if (it >= n) return 0.;
    einmax = 0.5 * (U[it] - U[it - 1]);
    shwav = hc / einmax;
    if (shwav >= QH_CUT) return 0.;
    nshort = qh_below(lam, B, shwav) + 2;  /* 1-based, the first with lam >= shwav */
    nlong = qh_below(lam, B, QH_CUT) + 1;  /* 1-based, the last with lam < cutoff */
    if (nshort >= nlong) return 0.;
    for (k = nshort; k <= nlong; k++)
        if (k >= 2) engin = engin + ((sQ[k - 1] * fl[k - 1] + sQ[k - 2] * fl[k - 2]) / 2.) * (lam[k - 1] - lam[k - 2]);
    return engin;
}

int oracle_qheat(const oracle_qheat_in* q, double tg, double* temp, double* pss, double* work) {
    const int B = q->B, n = q->nTbins;
    const double hc = 6.6262e-27 * QH_C;
    double *lam = work, *rf = lam + B, *sQ = rf + B, *U = sQ + B, *dU = U + n, *x = dU + n;
    double *A = x + n, *Bm = A + (size_t)n * n;  /* A[f][i], Bm[f][i], 0-based */
    int natom, i, f, ii, j;
    double sumx = 0.;
#define AA(r, c) A[(size_t)(r - 1) * n + (c - 1)]
#define BB(r, c) Bm[(size_t)(r - 1) * n + (c - 1)]
    for (i = 0; i < n; i++) temp[i] = pss[i] = 0.;
    /* the field as F_lambda, the mesh reversed to increasing wavelength */
    for (i = 0; i < B; i++) {
        const int k = B - 1 - i;
        const double r = q->J[k] / (q->widFlx[k] * 3.28984e15);
        rf[i] = r * pow(q->nu[k] * 3.28984e15, 2.) / QH_C;
        lam[i] = QH_C / (q->nu[k] * 3.28984e15);
        sQ[i] = q->Cabs[k];
    }
    if (q->kind == 1) natom = (int)(0.44 * 1.e12 * pow(q->radius, 3.));
    else if (q->kind == 2) natom = (int)(0.454 * 3. * 1.e12 * pow(q->radius, 3.));
    else return 0;  /* neither silicate nor carbon: no distribution */
    temp[0] = 1.;   /* getTmin */
    U[0] = qh_enthalpy(temp[0], natom, q->radius, q->kind);
    for (i = 2; i <= 10 && i <= n; i++) {
        temp[i - 1] = temp[0] + (double)(i - 1) / 9. * (tg - temp[0]);
        U[i - 1] = qh_enthalpy(temp[i - 1], natom, q->radius, q->kind);
    }
    for (i = 11; i <= n; i++) {
        temp[i - 1] = tg + (double)(i - 10) / (double)(n - 10) * (q->Tmax - tg);
        U[i - 1] = qh_enthalpy(temp[i - 1], natom, q->radius, q->kind);
    }
    dU[0] = 0.5 * (U[1] - U[0]);
    for (i = 2; i <= n - 1; i++) dU[i - 1] = 0.5 * (U[i] - U[i - 2]);
    dU[n - 1] = 0.5 * (U[n - 1] - U[n - 2]);
    /* heating: from bin ii to bin f by the photons of hc / (U(f) - U(ii)) */
    for (ii = 1; ii <= n - 1; ii++) {
        double UiTrun, wlim, qint = 0.;
        for (f = ii + 1; f <= n; f++) {
            const double wl = hc / (U[f - 1] - U[ii - 1]);
            int wlp = qh_below(lam, B, wl) + 2;
            if (wlp > B) wlp = B;
            if (dU[f - 1] >= 0.1 * (U[f - 1] - U[ii - 1]) && f < n) {
                double wls = hc / ((U[f - 1] + U[f]) / 2. - U[ii - 1]);
                double wll = hc / ((U[f - 1] + U[f - 2]) / 2. - U[ii - 1]);
                if (wl >= 0.1 || wl < lam[0]) {
                    AA(f, ii) = 0.;
                } else {
                    int wllp, wlsp;
                    if (wll > 0.1) wll = .1;
                    if (wls < lam[0]) wls = lam[0];
                    wllp = qh_below(lam, B, wll) + 2;
                    if (wllp > B) wllp = B;
                    wlsp = qh_below(lam, B, wls) + 2;
                    if (wlsp > B) wlsp = B;
                    AA(f, ii) = 0.5 * (sQ[wlsp - 1] * rf[wlsp - 1] * (wls * wls * wls) + sQ[wlp - 1] * rf[wlp - 1] * (wl * wl * wl)) *
                                (U[f] - U[f - 1]) / 2.;
                    AA(f, ii) = AA(f, ii) + 0.5 * (sQ[wlp - 1] * rf[wlp - 1] * (wl * wl * wl) + sQ[wllp - 1] * rf[wllp - 1] * (wll * wll * wll)) *
                                                (U[f - 1] - U[f - 2]) / 2.;
                    AA(f, ii) = AA(f, ii) / pow(hc, 2.);
                }
            } else {
                if (wl >= 0.1 || wl < lam[0]) AA(f, ii) = 0.;
                else AA(f, ii) = sQ[wlp - 1] * rf[wlp - 1] * dU[f - 1] * (wl * wl * wl) / pow(hc, 2.);
            }
        }
        /* heating beyond the last bin (G&D p232) */
        UiTrun = U[n - 1] + dU[n - 1] / 2. - U[ii - 1];
        wlim = hc / UiTrun;
        for (j = 1; j <= B - 1; j++) {
            const double q1 = lam[j - 1] * sQ[j - 1] * rf[j - 1], q2 = lam[j] * sQ[j] * rf[j];
            const double del = lam[j] - lam[j - 1];
            if (lam[j] < wlim) qint = qint + 0.5 * (q1 + q2) * del;
            else if (lam[j - 1] < wlim) qint = qint + (wlim - lam[j - 1]) * (q1 + 0.5 * (wlim - lam[j - 1]) * (q2 - q1) / del);
        }
        qint = qint / hc;
        AA(n, ii) = AA(n, ii) + qint;
    }
    /* cooling: from bin if+1 to bin if */
    for (f = 1; f <= n - 1; f++) {
        const double mdUdt = qh_clrate(temp[f], lam, rf, sQ, B);
        const double ch = qh_cheat(f + 1, n, lam, rf, sQ, B, U, hc);
        AA(f, f + 1) = mdUdt / dU[f];
        if (mdUdt <= ch) {
            AA(f, f + 1) = AA(f, f + 1) * 1.e-20;
            if (f + 2 <= n) AA(f + 2, f + 1) = AA(f + 2, f + 1) + (ch - mdUdt) / dU[f];
        } else {
            AA(f, f + 1) = AA(f, f + 1) - ch / dU[f];
        }
    }
    if (AA(2, 1) == 0.) AA(2, 1) = 2. * AA(1, 2);
    /* B(f, ii) = sum of A(f', ii) over f' >= f (G&D p233) */
    for (ii = 1; ii <= n; ii++) {
        if (ii < n) BB(n, ii) = AA(n, ii);
        for (f = n - 1; f >= ii + 1; f--) BB(f, ii) = BB(f + 1, ii) + AA(f, ii);
    }
    x[0] = 1.e-100;
    for (f = 2; f <= n; f++) {
        double s = 0.;
        for (j = 1; j <= f - 1; j++) s = s + BB(f, j) * x[j - 1];
        x[f - 1] = s / AA(f - 1, f);
        if (x[f - 1] > 1.e250)
            for (i = 1; i <= f; i++) x[i - 1] = x[i - 1] < 1. ? 0. : x[i - 1] * 1.e-250;
    }
    for (j = 1; j <= n; j++) sumx = sumx + x[j - 1];
    for (i = 1; i <= n - 1; i++) {
        if (x[i - 1] > 0. && log10(x[i - 1]) - log10(sumx) <= -250.) x[i - 1] = 0.;
        pss[i - 1] = x[i - 1] / sumx;
    }
#undef AA
#undef BB
    return 1;
}

size_t oracle_qheat_work(int B, int nTbins) {
    return (size_t)3 * B + (size_t)3 * nTbins + (size_t)2 * nTbins * nTbins;
}

// Grains on gfx950: dust opacity, the dust emission CDF and the grain
// temperatures.
//
//  k_dust_opacity -- the dust block of iterateMC (iteration_mod.f90:194-246):
//                    for every (bin, cell), scaOpac and absOpac summed over
//                    the (species, size) pairs below their sublimation
//                    temperature, in the reference's species-then-size order,
//                    and opacity = gas kappa + (scaOpac + absOpac).  One lane
//                    per cell walks the bins: the cell's grain temperatures
//                    are read once into a sublimation mask, the cross-sections
//                    of a bin are the same for the whole wave (scalar loads),
//                    and each bin's planes are written in coalesced rows.
//                    More than 64 grains: one lane per (bin, cell).
//  k_dust_pdf     -- setDustPDF (emission_mod.f90:1921-1986), dust-only runs:
//                    one workgroup per cell; each lane forms whole bins
//                    sum Cabs B_nu(Tdust)/h widFlx grainWeight grainAbun over
//                    the grains with 0 < Tdust < Tsub, the workgroup scans
//                    them (wave shuffles, 256 bins per step) and normalises
//                    the row, whose last entry is set to 1.
//  k_dust_temp    -- getDustT (update_mod.f90:3346-3453): the absorption
//                    integrals sum_nu Cabs J/pi of every grain come from the
//                    k_rates contraction with columns Cabs/pi (and a column
//                    of ones for the hit test), a row per cell; one lane per
//                    cell then
//                    inverts the Kirchhoff table (locate + linear
//                    interpolation in T) and forms the size- and
//                    abundance-weighted means.
// All arithmetic is FP64.
#include <hip/hip_runtime.h>

#include "moc_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr double kPiRef = 3.141592654;  // constants_mod.f90:45

__device__ __forceinline__ size_t tdust_at(const MocDustArgs& a, int cell, int s, int ai) {
    return ((size_t)cell * (a.nA + 1) + ai) * (a.nS + 1) + s;
}

__global__ void __launch_bounds__(kThreads) k_dust_opacity_cells(MocDustArgs a) {
    const int C1 = a.nCells + 1;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= C1) return;
    const int B = a.B, nA = a.nA, nS = a.nS;
    if (cell == 0) {
        if (a.dustOnly)
            for (int b = 0; b < B; ++b) a.opacity[(size_t)b * C1] = 0.;
        return;
    }
    uint64_t live = 0;  // grains below their sublimation temperature
    for (int s = 1; s <= nS; ++s)
        for (int ai = 1; ai <= nA; ++ai)
            if (a.Tdust[tdust_at(a, cell, s, ai)] < a.tsub[s - 1])
                live |= 1ull << ((s - 1) * nA + ai - 1);
    const double nd = a.Ndust[cell];
    for (int b = 0; b < B; ++b) {
        double sca = 0., ab = 0.;
        for (int g = 0; g < nS * nA; ++g) {
            if (!((live >> g) & 1ull)) continue;
            const int s = g / nA, ai = g - s * nA;
            const size_t k = (size_t)g * B + b;
            const double f = a.abun[s] * a.weight[ai] * nd;
            sca = sca + f * a.Csca[k];
            ab = ab + f * a.Cabs[k];
        }
        const size_t t = (size_t)b * C1 + cell;
        a.scaOpac[t] = sca;
        a.absOpac[t] = ab;
        const double base = a.gasOpac ? a.gasOpac[t] : 0.;
        a.opacity[t] = base + (sca + ab);
    }
}

__global__ void __launch_bounds__(kThreads) k_dust_opacity(MocDustArgs a) {
    const int64_t C1 = a.nCells + 1;
    const int64_t n = C1 * a.B;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(t / C1);
        const int cell = (int)(t - (int64_t)b * C1);
        if (cell == 0) {
            if (a.dustOnly) a.opacity[t] = 0.;
            continue;
        }
        double sca = 0., ab = 0.;
        const double nd = a.Ndust[cell];
        for (int s = 1; s <= a.nS; ++s)
            for (int ai = 1; ai <= a.nA; ++ai) {
                if (!(a.Tdust[tdust_at(a, cell, s, ai)] < a.tsub[s - 1])) continue;
                const size_t k = ((size_t)(s - 1) * a.nA + (ai - 1)) * a.B + b;
                const double f = a.abun[s - 1] * a.weight[ai - 1] * nd;
                sca = sca + f * a.Csca[k];
                ab = ab + f * a.Cabs[k];
            }
        a.scaOpac[t] = sca;
        a.absOpac[t] = ab;
        const double base = a.gasOpac ? a.gasOpac[t] : 0.;
        a.opacity[t] = base + (sca + ab);
    }
}

__device__ __forceinline__ double wave_sum_incl(double v, int lane) {
    for (int off = 1; off < 64; off <<= 1) {
        const double t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

__global__ void __launch_bounds__(kThreads) k_dust_pdf(MocDustArgs a) {
    extern __shared__ double sT[];  // [nS][nA] grain temperatures of the cell
    __shared__ double sWave[kThreads / 64];
    const int cell = blockIdx.x + 1;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int B = a.B, nG = a.nS * a.nA;
    for (int k = tid; k < nG; k += kThreads) {
        const int s = k / a.nA + 1, ai = k % a.nA + 1;
        sT[k] = a.Tdust[tdust_at(a, cell, s, ai)];
    }
    __syncthreads();
    double* row = a.pdf + (size_t)cell * B;
    double carry = 0.;
    for (int b0 = 0; b0 < B; b0 += kThreads) {
        const int b = b0 + tid;
        double v = 0.;
        if (b < B) {
            const double nu = a.nu[b], w = a.widFlx[b];
            for (int s = 1; s <= a.nS; ++s)
                for (int ai = 1; ai <= a.nA; ++ai) {
                    const int g = (s - 1) * a.nA + ai - 1;
                    const double T = sT[g];
                    if (!(T > 0. && T < a.tsub[s - 1])) continue;
                    const double c = a.Cabs[(size_t)g * B + b];
                    const int q = a.qActive ? a.qIdx[g] : -1;
                    if (q >= 0) {  // quantum heating: the grain's temperature distribution
                        const double* qt = a.qT + ((size_t)cell * a.nQ + q) * a.nTb;
                        const double* qp = a.qP + ((size_t)cell * a.nQ + q) * a.nTb;
                        for (int iT = 0; iT < a.nTb; ++iT)
                            v = v + c * moc_planck_h(nu, qt[iT]) * w * a.weight[ai - 1] * a.abun[s - 1] * qp[iT];
                    } else {
                        v = v + c * moc_planck_h(nu, T) * w * a.weight[ai - 1] * a.abun[s - 1];
                    }
                }
        }
        const double sc = wave_sum_incl(v, lane);
        if (lane == 63) sWave[wv] = sc;
        __syncthreads();
        double pre = carry;
        for (int k = 0; k < wv; ++k) pre += sWave[k];
        if (b < B) row[b] = pre + sc;
        double next = carry;
        for (int k = 0; k < kThreads / 64; ++k) next += sWave[k];
        __syncthreads();
        carry = next;
    }
    // a row with no emitting grain stays 0 up to its last entry
    const double total = carry;
    for (int b = tid; b < B; b += kThreads)
        row[b] = b == B - 1 ? 1. : (total > 0. ? row[b] / total : 0.);
}

// locate (interpolation_mod.f90:49-106) on an increasing table: the largest
// 1-based index with xa <= x, 0 below the table, n above it
__device__ int locate_inc(const double* xa, int n, double x) {
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int lo = 0, hi = n;  // first index with xa > x
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (xa[m] > x) hi = m; else lo = m + 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kThreads) k_dust_temp(MocDustArgs a) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (cell > a.nCells) return;
    const int C1 = a.nCells + 1, nG = a.nS * a.nA;
    const double* Rc = a.R + (size_t)cell * (nG + 1);  // the cell's absorption integrals, hit last
    if (!(Rc[nG] > 0.)) {  // no packet reached the cell: Tdust kept
        if (a.dustOnly) {
            a.lgBlack[cell] = 1;
            a.lgConverged[cell] = 0;
            atomicAdd(&a.counters[0], 1ull);
        }
        return;
    }
    double* Td = a.Tdust + tdust_at(a, cell, 0, 0);
    const double old = Td[0];
    for (int k = 0; k < (a.nS + 1) * (a.nA + 1); ++k) Td[k] = 0.;
    for (int s = 1; s <= a.nS; ++s) {
        for (int ai = 1; ai <= a.nA; ++ai) {
            const int g = (s - 1) * a.nA + (ai - 1);
            double x = Rc[g];
            if (a.collD)  // gas-grain collisions and photoelectric cooling (:3390-3393)
                x = x + a.collD[(size_t)g * C1 + cell] - a.peD[(size_t)g * C1 + cell];
            const double* E = a.emInt + (size_t)g * a.nT;
            const int iT = locate_inc(E, a.nT, x);
            double T;
            if (iT <= 0) T = 1.;
            else if (iT >= a.nT) T = (double)a.nT;
            else T = (double)iT + (x - E[iT - 1]) * 1. / (E[iT] - E[iT - 1]);
            Td[(size_t)ai * (a.nS + 1) + s] = T;
            Td[s] = Td[s] + T * a.weight[ai - 1];
        }
        Td[0] = Td[0] + Td[s] * a.abun[s - 1];
    }
    if (a.dustOnly) {
        a.lgBlack[cell] = 0;
        const bool conv = fabs((Td[0] - old) / old) <= a.XHILimit;
        a.lgConverged[cell] = conv ? 1 : 0;
        if (conv) atomicAdd(&a.counters[1], 1ull);
    }
}

// ------------------------------------------------------------------ qHeat
// k_qheat: the temperature distribution of each grain smaller than minaQHeat
// in each cell (qHeat, emission_mod.f90:1993-2264, after Guhathakurta &
// Draine 1989), one workgroup per (cell, grain), as the oracle restates it:
// the field as F_lambda on the reversed mesh, temperature bins from 1 K
// (getTmin) through Tdust to Tmax, the enthalpies (enthalpy :2326-2375), the
// heating rates A(f, ii) of every upward transition (a lane per initial bin,
// summed at once into B(f, ii) = sum over f' >= f of A(f', ii), held in the
// workgroup's row of qWork), the cooling rates (clrate, cheat: a lane per
// bin), then the recurrence x(f) = sum_j<f B(f, j) x(j) / A(f-1, f) with the
// sums over j taken by the workgroup.  All in FP64.
constexpr double kQhC = 2.9979250e10;     // constants_mod.f90:12
constexpr double kQhCut = 0.1;            // cm
constexpr double kFr1Ryd = 3.28984e15;    // constants_mod.f90:13

__device__ double qh_enthalpy(double tg, int natom, double radius, int kind) {
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

// last 0-based index with lam < x on the increasing mesh, or -1
__device__ int qh_below(const double* lam, int B, double x) {
    int lo = 0, hi = B;  // first index with lam >= x
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (lam[m] < x) lo = m + 1; else hi = m;
    }
    return lo - 1;
}

__device__ double qh_bbody(double lam5, double lam, double T) {
    const double C1 = 3.74185e-5, C2 = 1.43883;
    const double x = C2 / (lam * T);
    if (x > 400.) return 0.;
    return C1 * (1. / lam5) / (exp(x) - 1.);
}

__device__ double qh_block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kThreads) k_qheat(MocDustArgs a) {
    extern __shared__ double qs[];
    __shared__ double red[kThreads];
    const int B = a.B, n = a.nTb, tid = threadIdx.x;
    const int C1 = a.nCells + 1;
    double *lam = qs, *rf = lam + B, *sQ = rf + B, *lam5 = sQ + B;
    double *temp = lam5 + B, *U = temp + n, *dU = U + n, *sup = dU + n, *corr = sup + n + 1, *x = corr + n + 2;
    double* Bm = a.qWork + (size_t)blockIdx.x * n * n;  // B(f, ii) at [(f - 1) n + ii - 1]
    const double hc = 6.6262e-27 * kQhC;
    const double hc2 = pow(hc, 2.);
    for (int t = blockIdx.x; t < a.nCells * a.nQ; t += gridDim.x) {
        const int cell = t / a.nQ + 1, q = t % a.nQ;
        const int g = a.qG[q], s = g / a.nA, ai = g % a.nA;
        double* oT = a.qT + ((size_t)cell * a.nQ + q) * n;
        double* oP = a.qP + ((size_t)cell * a.nQ + q) * n;
        const double tg = a.Tdust[tdust_at(a, cell, s + 1, ai + 1)];
        const bool on = a.dustOnly ? (tg > 0. && tg < a.tsub[s]) : tg < a.tsub[s];
        const int kind = a.kind[s];
        if (!on) continue;  // the emission skips this grain
        if (kind != 1 && kind != 2) {  // no distribution: the grain emits nothing
            for (int i = tid; i < n; i += kThreads) oT[i] = oP[i] = 0.;
            continue;
        }
        const double radius = a.radius[ai];
        const double* cab = a.Cabs + (size_t)g * B;
        for (int i = tid; i < B; i += kThreads) {
            const int k = B - 1 - i;
            const double r = a.J[(size_t)k * C1 + cell] / (a.widFlx[k] * kFr1Ryd);
            rf[i] = r * pow(a.nu[k] * kFr1Ryd, 2.) / kQhC;
            lam[i] = kQhC / (a.nu[k] * kFr1Ryd);
            lam5[i] = pow(lam[i], 5.);
            sQ[i] = cab[k];
        }
        const int natom = kind == 1 ? (int)(0.44 * 1.e12 * pow(radius, 3.))
                                    : (int)(0.454 * 3. * 1.e12 * pow(radius, 3.));
        for (int i = tid + 1; i <= n; i += kThreads) {
            double T;
            if (i == 1) T = 1.;  // getTmin
            else if (i <= 10) T = 1. + (double)(i - 1) / 9. * (tg - 1.);
            else T = tg + (double)(i - 10) / (double)(n - 10) * (a.Tmax - tg);
            temp[i - 1] = T;
            U[i - 1] = qh_enthalpy(T, natom, radius, kind);
        }
        __syncthreads();
        for (int i = tid + 1; i <= n; i += kThreads)
            dU[i - 1] = i == 1 ? 0.5 * (U[1] - U[0]) : (i == n ? 0.5 * (U[n - 1] - U[n - 2]) : 0.5 * (U[i] - U[i - 2]));
        __syncthreads();
        // cooling from bin f+1 to bin f (clrate, cheat): A(f, f+1) -> sup[f],
        // and the heating of bin f+2 from bin f+1 where the continuous
        // heating within the bin exceeds the cooling -> corr[f+1]
        const int cutP = max(qh_below(lam, B, kQhCut), 0);
        for (int f = tid + 1; f <= n - 1; f += kThreads) {
            const double tem = temp[f];
            double engout = 0., prev = 0.;
            for (int k = 0; k < B; ++k) {
                const double e = sQ[k] * qh_bbody(lam5[k], lam[k], tem);
                if (k >= 1) engout = engout + 4. * ((e + prev) / 2.) * (lam[k] - lam[k - 1]);
                prev = e;
            }
            const double mdUdt = engout - 0.25 * rf[cutP] * sQ[cutP] * kQhCut;
            double ch = 0.;
            const int it = f + 1;
            if (it < n) {
                const double shwav = hc / (0.5 * (U[it] - U[it - 1]));
                if (shwav < kQhCut) {
                    const int nshort = qh_below(lam, B, shwav) + 2, nlong = qh_below(lam, B, kQhCut) + 1;
                    if (nshort < nlong)
                        for (int k = nshort; k <= nlong; ++k)
                            if (k >= 2)
                                ch = ch + ((sQ[k - 1] * rf[k - 1] + sQ[k - 2] * rf[k - 2]) / 2.) *
                                              (lam[k - 1] - lam[k - 2]);
                }
            }
            double A = mdUdt / dU[f];
            double cr = 0.;
            if (mdUdt <= ch) {
                A = A * 1.e-20;
                if (f + 2 <= n) cr = (ch - mdUdt) / dU[f];
            } else {
                A = A - ch / dU[f];
            }
            sup[f] = A;
            corr[f + 1] = cr;
        }
        if (tid == 0) corr[1] = 0.;
        __syncthreads();
        // heating from bin ii: A(f, ii) for f = n .. ii+1, summed into B(f, ii)
        for (int ii = tid + 1; ii <= n - 1; ii += kThreads) {
            const double Ui = U[ii - 1];
            double bsum = 0.;
            for (int f = n; f >= ii + 1; --f) {
                const double wl = hc / (U[f - 1] - Ui);
                int wlp = qh_below(lam, B, wl) + 2;
                if (wlp > B) wlp = B;
                double A;
                if (dU[f - 1] >= 0.1 * (U[f - 1] - Ui) && f < n) {
                    double wls = hc / ((U[f - 1] + U[f]) / 2. - Ui);
                    double wll = hc / ((U[f - 1] + U[f - 2]) / 2. - Ui);
                    if (wl >= 0.1 || wl < lam[0]) {
                        A = 0.;
                    } else {
                        if (wll > 0.1) wll = .1;
                        if (wls < lam[0]) wls = lam[0];
                        int wllp = qh_below(lam, B, wll) + 2, wlsp = qh_below(lam, B, wls) + 2;
                        if (wllp > B) wllp = B;
                        if (wlsp > B) wlsp = B;
                        A = 0.5 * (sQ[wlsp - 1] * rf[wlsp - 1] * (wls * wls * wls) + sQ[wlp - 1] * rf[wlp - 1] * (wl * wl * wl)) *
                            (U[f] - U[f - 1]) / 2.;
                        A = A + 0.5 * (sQ[wlp - 1] * rf[wlp - 1] * (wl * wl * wl) + sQ[wllp - 1] * rf[wllp - 1] * (wll * wll * wll)) *
                                    (U[f - 1] - U[f - 2]) / 2.;
                        A = A / hc2;
                    }
                } else {
                    A = (wl >= 0.1 || wl < lam[0]) ? 0. : sQ[wlp - 1] * rf[wlp - 1] * dU[f - 1] * (wl * wl * wl) / hc2;
                }
                if (f == n) {  // the heating beyond the last bin (G&D p232)
                    const double wlim = hc / (U[n - 1] + dU[n - 1] / 2. - Ui);
                    double qint = 0.;
                    for (int j = 1; j <= B - 1; ++j) {
                        const double q1 = lam[j - 1] * sQ[j - 1] * rf[j - 1], q2 = lam[j] * sQ[j] * rf[j];
                        const double del = lam[j] - lam[j - 1];
                        if (lam[j] < wlim) qint = qint + 0.5 * (q1 + q2) * del;
                        else if (lam[j - 1] < wlim)
                            qint = qint + (wlim - lam[j - 1]) * (q1 + 0.5 * (wlim - lam[j - 1]) * (q2 - q1) / del);
                    }
                    A = A + qint / hc;
                }
                if (f == ii + 1) {
                    A = A + corr[ii];
                    if (ii == 1 && A == 0.) A = 2. * sup[1];
                }
                bsum = f == n ? A : bsum + A;
                Bm[(size_t)(f - 1) * n + ii - 1] = bsum;
            }
        }
        __syncthreads();
        // the recurrence over the bins (G&D eqn 2.16) with the 1e250 rescaling
        if (tid == 0) x[0] = 1.e-100;
        __syncthreads();
        for (int f = 2; f <= n; ++f) {
            double p = 0.;
            for (int j = tid + 1; j <= f - 1; j += kThreads) p += Bm[(size_t)(f - 1) * n + j - 1] * x[j - 1];
            const double sum = qh_block_sum(p, red);
            const double xf = sum / sup[f - 1];
            if (xf > 1.e250) {
                for (int i = tid + 1; i <= f - 1; i += kThreads) x[i - 1] = x[i - 1] < 1. ? 0. : x[i - 1] * 1.e-250;
                if (tid == 0) x[f - 1] = xf * 1.e-250;
            } else if (tid == 0) {
                x[f - 1] = xf;
            }
            __syncthreads();
        }
        double p = 0.;
        for (int i = tid; i < n; i += kThreads) p += x[i];
        const double sumx = qh_block_sum(p, red);
        for (int i = tid + 1; i <= n; i += kThreads) {
            double xi = x[i - 1];
            if (i < n && xi > 0. && log10(xi) - log10(sumx) <= -250.) xi = 0.;
            oT[i - 1] = temp[i - 1];
            oP[i - 1] = i < n ? xi / sumx : 0.;
        }
        __syncthreads();  // the LDS rows are refilled for the next task
    }
}

int check(moc_ctx* ctx, const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string(what) + ": " + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

}  // namespace

int moc_launch_dust_opacity(moc_ctx* ctx, const MocDustArgs& a) {
    if (a.nS * a.nA <= 64) {
        const unsigned blocks = (unsigned)((a.nCells + 1 + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(k_dust_opacity_cells, dim3(blocks), dim3(kThreads), 0, ctx->stream, a);
        return check(ctx, "dust opacity launch");
    }
    const int64_t n = (int64_t)(a.nCells + 1) * a.B;
    int64_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_dust_opacity, dim3((unsigned)blocks), dim3(kThreads), 0, ctx->stream, a);
    return check(ctx, "dust opacity launch");
}

int moc_launch_dust_pdf(moc_ctx* ctx, const MocDustArgs& a) {
    if (a.nCells < 1) return MOC_OK;
    const size_t lds = (size_t)a.nS * a.nA * sizeof(double);
    hipLaunchKernelGGL(k_dust_pdf, dim3((unsigned)a.nCells), dim3(kThreads), lds, ctx->stream, a);
    return check(ctx, "dust emission launch");
}

size_t moc_qheat_lds(int B, int nTb) {
    return ((size_t)4 * B + (size_t)6 * nTb + 3) * sizeof(double);
}

int moc_launch_qheat(moc_ctx* ctx, const MocDustArgs& a) {
    if (a.nCells < 1 || a.nQ < 1) return MOC_OK;
    if (a.nTb < 11 || a.qBlocks < 1) {
        ctx->err = "qHeat: at least 11 temperature bins and a work buffer";
        return MOC_ERR_ARG;
    }
    hipLaunchKernelGGL(k_qheat, dim3((unsigned)a.qBlocks), dim3(kThreads), moc_qheat_lds(a.B, a.nTb), ctx->stream, a);
    return check(ctx, "qHeat launch");
}

int moc_launch_dust_temp(moc_ctx* ctx, const MocDustArgs& a, const double* J, const double* W,
                         int ncol, double* R) {
    moc_launch_rates(ctx, J, a.nCells + 1, W, ncol, R);
    if (a.nCells < 1) return check(ctx, "dust absorption integrals");
    hipLaunchKernelGGL(k_dust_temp, dim3((unsigned)((a.nCells + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, ctx->stream, a);
    return check(ctx, "dust temperature launch");
}

// Per-cell diffuse emission on gfx950 (emissionDriver, emission_mod.f90:36-252).
//
// Two kernels per cell.  k_emission, one workgroup of four waves per cell,
// takes the continuum, data-parallel over frequency: each lane evaluates
// whole bins -- the H I / He I / He II
// free-bound fits between their edge pointers (fb_ff :389-507), the Milne
// relation above the edges (:515-550), the heavy-ion free-bound sum
// (:552-616), free-free with the Hummer Gaunt factor at the cell's T_e
// (free_free :818-855) and the two-photon continua (:620-817) -- and the
// workgroup turns the bin energies c*dnu*(e_HI+e_HeI+e_HeII) into the
// normalised CDF recPDF(cell, :) with a wave-shuffle scan (setDiffusePDF
// :1593-1835), 256 bins per step, the running total carried across steps.
// k_em_lines, one wave per cell and run first, takes the lines --
// hydrogenic recombination lines per element (RecLinesEmission :856-1108),
// He I recombination lines (:1110-1152) and the n-level forbidden lines with
// the recombination feed (forLines :1481-1526) -- one task per lane; their
// sum against the continuum gives totalLines (:1904-1910).
// All arithmetic is FP64.  The row of recPDF is written once unnormalised and
// rescaled in place by the lanes that wrote it.
#include <hip/hip_runtime.h>

#include "gaunt.h"
#include "moc_internal.h"
#include "nlevel.h"

namespace {

constexpr double kPi = 3.141592654;        // constants_mod.f90
constexpr double kHcRyd = 2.1799153e-11;   // continuum_mod.f90:15
constexpr double kHcRydK = 157893.94;      // continuum_mod.f90:14
constexpr double kCRyd = 3.2898423e15;     // continuum_mod.f90:16
constexpr double kHPlanck = 6.6262e-27;    // constants_mod.f90:14
constexpr double kMilne = 4.9874105e-6;    // fb_ff :515
constexpr int kThreads = 256;
constexpr int kLineLumLds = 49152;  // line accumulators of one ion group

// hydroLinesTemps (grid_mod.f90:76-77)
__constant__ double c_hlT[12] = {0.05, 0.10, 0.30, 0.50, 0.75, 1.0, 1.25, 1.50, 2.0, 3.0, 5.0, 10.0};

// hydro2phot (:640-760): two-photon emission of H I (Z=1) and He II (Z=2).
// The temperature factors depend on the cell only and are formed once.
struct TwoPhotT {
    double aEffH, QH, aEffHeII;
};

__device__ TwoPhotT two_phot_setup(double Te) {
    TwoPhotT t;
    t.aEffH = 0.8368 * pow(Te * 1.e-4, -0.723);
    if (Te <= 10000.) t.QH = 5.31e-4;
    else if (Te >= 20000.) t.QH = 4.71e-4;
    else t.QH = pow(10., log10(5.31e-4) + log10(4.71e-4 / 5.31e-4) / log10(2.) * log10(Te / 10000.));
    if (Te <= 5000.) t.aEffHeII = 6.161;
    else if (Te >= 30000.) t.aEffHeII = 2.035;
    else if (Te <= 10000.) t.aEffHeII = pow(10., log10(6.161) + log10(4.091 / 6.161) / log10(2.) * log10(Te / 5000.));
    else if (Te <= 15000.) t.aEffHeII = pow(10., log10(4.091) + log10(3.189 / 4.091) / log10(1.5) * log10(Te / 10000.));
    else t.aEffHeII = pow(10., log10(3.189) + log10(2.035 / 3.189) / log10(2.) * log10(Te / 15000.));
    return t;
}

__device__ double hydro2phot(double nu, int Z, const TwoPhotT& t, double Ne) {
    const double nu0 = Z == 1 ? 0.7496 : 3.00;
    if (nu >= nu0) return 0.;
    const double y = nu / nu0;
    const double shape = y * (1 - y) * (1 - pow(4 * y * (1 - y), 0.8)) +
                         0.88 * pow(y * (1 - y), 1.53) * pow(4 * y * (1 - y), 0.8);
    if (Z == 1) {
        const double Ay = 202.0 * shape;
        const double gNu = kHPlanck * y * Ay / 8.2249 * 1e27;
        return t.aEffH * gNu / (1. + (Ne * t.QH) / 8.23);
    }
    const double Ay = 64. * 0.9994667 * 202.0 * shape;
    const double gNu = kHPlanck * y * Ay / (8.226 * 64.) * 1e27;
    return t.aEffHeII * gNu;
}

// nearest temperature bin of the hydrogenic line tables (:877-888)
__device__ int hl_itemp(double T4) {
    int itemp = 1;
    for (int i = 1; i <= 12; ++i)
        if (T4 > c_hlT[i - 1]) itemp = i;
    if (itemp < 12 && T4 > (c_hlT[itemp] + c_hlT[itemp - 1]) / 2.) ++itemp;
    return itemp;
}

// density bin of table file f: the last tabulated density below Ne
// (RecLinesEmission :915-920); outputGas takes the last one at or below Ne
// (ge, output_mod.f90:2054-2057)
__device__ int hl_idenp(const MocEmissionArgs& a, int f, double Ne, bool ge = false) {
    const double* dens = a.hlDens + (size_t)f * 14;
    int idenp = 1;
    for (int iden = 1; iden <= 13; ++iden)
        if ((ge ? Ne >= dens[iden - 1] : Ne > dens[iden - 1]) && dens[iden - 1] > 0.) idenp = iden;
    return idenp;
}

// the lines of table file f interpolated in density, each times Hb (:915-938)
__device__ double hl_lines(const MocEmissionArgs& a, int f, double Ne, double Hb) {
    const double* dens = a.hlDens + (size_t)f * 14;
    const double* L = a.hlLines + (size_t)f * 14 * 16 * 9;
    const int idenp = hl_idenp(a, f, Ne);
    const double* L0 = L + (size_t)(idenp - 1) * 16 * 9;
    double acc = 0.;
    if ((dens[12] > 0. && idenp < 13) || (dens[12] == 0. && idenp < 9)) {
        const double* L1 = L0 + 16 * 9;
        const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
        for (int iup = 2; iup <= 15; ++iup)
            for (int ilow = 1; ilow <= (iup - 1 < 8 ? iup - 1 : 8); ++ilow) {
                const double l0 = L0[iup * 9 + ilow];
                acc += (l0 + (L1[iup * 9 + ilow] - l0) * w) * Hb;
            }
    } else {
        for (int iup = 2; iup <= 15; ++iup)
            for (int ilow = 1; ilow <= (iup - 1 < 8 ? iup - 1 : 8); ++ilow) acc += L0[iup * 9 + ilow] * Hb;
    }
    return acc;
}

// one line (iup -> ilow) of table file f interpolated in density, times Hb
__device__ double hl_line(const MocEmissionArgs& a, int f, double Ne, double Hb, int iup, int ilow,
                          bool ge = false) {
    const double* dens = a.hlDens + (size_t)f * 14;
    const double* L = a.hlLines + (size_t)f * 14 * 16 * 9;
    const int idenp = hl_idenp(a, f, Ne, ge);
    const double* L0 = L + (size_t)(idenp - 1) * 16 * 9;
    if ((dens[12] > 0. && idenp < 13) || (dens[12] == 0. && idenp < 9)) {
        const double* L1 = L0 + 16 * 9;
        const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
        const double l0 = L0[iup * 9 + ilow];
        return (l0 + (L1[iup * 9 + ilow] - l0) * w) * Hb;
    }
    return L0[iup * 9 + ilow] * Hb;
}

struct HydroLines {
    int f;       // table file
    double Ne;   // density the table is read at
    double Hb;   // Hbeta emissivity the table is scaled by
};

// hydrogenic recombination lines of element izp (RecLinesEmission :890-1099):
// the table and the scale; hl_lines sums them, hl_line picks one
__device__ HydroLines hydro_setup(const MocEmissionArgs& a, int izp, double Xup, double ab, double Te,
                                  double Ne, double log10Ne, bool ge = false) {
    const double T4 = Te / 10000.;
    const int elUp = T4 > c_hlT[11] ? 2 : 8;
    double Hb;
    if (izp <= elUp) {
        const int f = a.hlMap[(izp - 1) * 12 + hl_itemp(T4) - 1];
        if ((izp == 1 && T4 <= 3.) || (izp > 1 && T4 <= c_hlT[11])) {
            const double* A = a.hbA + 4 * (izp - 1);
            const double* Bc = a.hbB + 4 * (izp - 1);
            double zf = (log10Ne - A[1]) / A[2];
            const double Af = A[0] * exp(-zf * zf / 2.) + A[3];
            zf = (log10Ne - Bc[1]) / Bc[2];
            const double Bf = Bc[0] * exp(-zf * zf / 2.) + Bc[3];
            Hb = pow(10., Af + Bf * log10(Te));
        } else {
            // H above 3e4 K, He above 1e5 K: the edge tables in density and T
            const int idenp = hl_idenp(a, f, Ne, ge);
            const double* dens = a.hlDens + (size_t)f * 14;
            const double xx = log10(T4 * 1.e4);
            double yy1, yy2, x1, x2;
            if (izp == 1) {
                const double* r = a.rbEdge + 2 * 26;
                if (idenp < 13) {
                    const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
                    yy1 = r[idenp - 1] + (r[idenp - 1] - r[idenp]) * w;
                    yy2 = r[13 + idenp - 1] + (r[13 + idenp - 1] - r[13 + idenp]) * w;
                } else {
                    yy1 = r[idenp - 1];
                    yy2 = r[13 + idenp - 1];
                }
                x1 = a.rbEdge[0];
                x2 = a.rbEdge[13];
            } else {
                // column 9+idenp(+1) of r2aEdge clamped to its 18 entries
                const double* r = a.r2aEdge + 2 * 18;
                const int a9 = 9 + idenp < 18 ? 9 + idenp : 18, b9 = 10 + idenp < 18 ? 10 + idenp : 18;
                if (idenp < 13) {
                    const double w = (Ne - dens[idenp - 1]) / (dens[idenp] - dens[idenp - 1]);
                    yy1 = r[idenp - 1] + (r[idenp - 1] - r[idenp]) * w;
                    yy2 = r[a9 - 1] + (r[a9 - 1] - r[b9 - 1]) * w;
                } else {
                    yy1 = r[idenp - 1];
                    yy2 = r[a9 - 1];
                }
                x1 = a.r2aEdge[0];
                x2 = a.r2aEdge[9];
            }
            Hb = pow(10., yy1 - (yy1 - yy2) * (xx - x1) / (x2 - x1));
        }
        Hb = Hb * Ne * Xup * ab;
        return HydroLines{f, Ne, Hb};
    }
    // heavier hydrogenic ions: He II tables at the Z-scaled T and n_e
    const double z2 = (double)(izp * izp);
    const double T4z = T4 * 4. / z2, lTeZ = log10(Te * 4. / z2);
    const double Nez = Ne * 4. / z2, lNez = log10(Ne * 4. / z2);
    const int f = a.hlMap[1 * 12 + hl_itemp(T4z) - 1];
    double zf = (lNez - a.hbA[5]) / a.hbA[6];
    const double Af = a.hbA[4] * exp(-zf * zf / 2.) + a.hbA[7];
    zf = (lNez - a.hbB[5]) / a.hbB[6];
    const double Bf = a.hbB[4] * exp(-zf * zf / 2.) + a.hbB[7];
    double Hb2 = ((double)(izp * izp * izp) / 8.) * pow(10., Af + Bf * lTeZ);
    Hb2 = Hb2 * Ne * Xup * ab;
    return HydroLines{f, Nez, Hb2};
}

__device__ double hydro_element(const MocEmissionArgs& a, int izp, double Xup, double ab, double Te,
                                double Ne, double log10Ne) {
    const HydroLines h = hydro_setup(a, izp, Xup, ab, Te, Ne, log10Ne);
    return hl_lines(a, h.f, h.Ne, h.Hb);
}

// He I recombination line i of the 34 (:1110-1152); the T4 clamps are those
// written there.  Each line is its own task; the workgroup sums them.
constexpr int kHeILines = 34;
__device__ double hei_line(const MocEmissionArgs& a, int i, double Te, double Ne, double NeGrid,
                           double X2, double ab) {
    double T4 = Te / 10000.;
    if (T4 < 0.5) T4 = 0.5;
    if (T4 < 2.0) T4 = 2.0;
    int d;
    if (NeGrid <= 100.) d = 0;
    else if (NeGrid <= 1.e4) d = 1;
    else if (NeGrid <= 1.e6) d = 2;
    else d = 3;
    const double* c = a.heiRec + (size_t)i * 12;
    double v;
    if (d > 0 && d < 3) {
        const double* c1 = c + (d - 1) * 4;
        const double* c2 = c + d * 4;
        const double x1 = c1[0] * pow(T4, c1[1]) * exp(c1[2] / T4);
        const double x2 = c2[0] * pow(T4, c2[1]) * exp(c2[2] / T4);
        v = x1 + ((x2 - x1) * (Ne - pow(100., d)) / (pow(100., d + 1) - pow(100., d)));
    } else {
        const double* c1 = c + (d == 0 ? 0 : 8);
        v = c1[0] * pow(T4, c1[1]) * exp(c1[2] / T4);
    }
    return v * Ne * ab * X2;
}

struct CellCtx {
    int cell;
    double Te, Ne, sqrTe;
    double abH, XH2, abHe, XHe2, XHe3;
    bool heOn;
    double aEffHeI;    // He I 2^1S effective recombination (HeI2photSub :781)
    GauntT g1, g2;     // Z = 1 and Z = 2 Gaunt fits at Te
    TwoPhotT tp;       // two-photon temperature factors
    double dustScale;  // hPlanck 4 pi 1e25 cRyd Ndust / Hden (setDiffusePDF :1749-1752)
    double corr;       // He I 2^3S line energy (setDiffusePDF :1611-1622)
    double ly[4];      // He II Lyman line energies (:1646-1673)
};

// c * dnu * (e_HI + e_HeI + e_HeII) of bin i (1-based), with the He I / He II
// line energies of setDiffusePDF added in their bins
// Milne factors of the heavy ions: exp(-(nu - nu_lo) h/kT) for the bins of a
// 256-bin pass that starts at bin b0 (0-based) is exp(-(nu - nu_b0) h/kT)
// (ePass, one per lane) times exp(-(nu_b0 - nu_lo) h/kT), for the ions whose
// edge is at or below nu_b0; both factors lie in (0, 1].  sF holds the second
// factor times the ion's Gaunt-free weight g X abundance, once per pass.  An
// edge inside the pass takes the direct exponential.
__device__ double bin_energy(const MocEmissionArgs& a, const CellCtx& c, const double (*sLoc)[MOC_EM_MAXLIM],
                             const double (*sFit)[MOC_EM_MAXLIM / 2], const double (*sX)[MOC_MAXS + 1],
                             const double* sTd, int i, int b0, double ePass, const double* sF) {
    const double nu = a.nu[i - 1];
    const double Te = c.Te, Ne = c.Ne;
    double gam[3] = {0., 0., 0.};
    // free-bound fits between the edge pointers: the pair that covers the bin
    // (the last one, where pairs overlap) comes from a per-bin table
    for (int k = 0; k < 3; ++k) {
        const int p = a.gpair[k * a.B + i - 1];
        if (p > 0) {
            const int iup = a.edge[k][2 * p - 1];
            gam[k] = exp10(sLoc[k][2 * p - 1] + sFit[k][p - 1] * (nu - a.nu[iup - 1]));
        }
    }
    const double xlf = a.log10nu[i - 1];
    const double ex0 = exp(-nu * kHcRydK / Te);
    const double fourPi = 4. * kPi;
    const double ff1 = fourPi * 54.43 * 1. * gaunt_eval(c.g1, xlf) * ex0 / c.sqrTe;
    const double ff2 = fourPi * 54.43 * 4. * gaunt_eval(c.g2, xlf) * ex0 / c.sqrTe;
    const double factor = nu * nu * nu / (Te * c.sqrTe);
    if (i >= a.HIP) {
        const double ex = exp((-nu + a.nu[a.HIP - 1]) * kHcRydK / Te);
        gam[0] = fourPi * a.xsHI[i - a.HIP] * 2. * kHcRyd * kMilne * factor * ex * 1.e20 * 1.e20 + ff1;
        if (c.heOn && i >= a.edge[1][a.nlim[1] - 1]) {
            const double sI = i >= a.HeIP ? a.xsHeI[i - a.HeIP] : 0.;
            const double exI = exp((-nu + a.nu[a.HeIP - 1]) * kHcRydK / Te);
            gam[1] = fourPi * sI * 0.5 * kHcRyd * kMilne * factor * exI * 1.e20 * 1.e20 + ff1;
            if (i >= a.edge[2][a.nlim[2] - 1]) {
                const double sII = i >= a.HeIIP ? a.xsHeII[i - a.HeIIP] : 0.;
                const double exII = exp((-nu + a.nu[a.HeIIP - 1]) * kHcRydK / Te);
                gam[2] = fourPi * sII * 2. * kHcRyd * kMilne * factor * exII * 1.e20 * 1.e20 + ff2;
            }
        }
    }
    // heavy-ion free-bound (Z = 3..19 entries in Z, ion order): an ion whose
    // edge is at or below the pass start contributes xs * sF (its weight
    // exp(-(nu_b0 - nu_lo) h/kT) g X abundance), all times ePass; an edge
    // inside the pass takes the direct exponential
    double gFac = 0., gDir = 0.;
    for (int n = 0; n < a.nFb; ++n) {
        const int lo = a.fbLo[n];
        if (i < lo || i >= a.fbHi[n]) continue;
        const double xs = a.fbXs[a.fbOff[n] + (i - lo)];
        if (lo - 1 <= b0) {
            gFac += xs * sF[n];
        } else {
            const int s = a.fbSlot[n];
            const double ex = exp((-nu + a.nu[lo - 1]) * kHcRydK / Te);
            gDir += ex * xs * a.fbG[n] * sX[s][a.fbIon[n] + 1] * a.abund[s];
        }
    }
    double gHv = ePass * gFac + gDir;
    gHv = gHv * fourPi * kMilne * kHcRyd * factor * 1.e20 * 1.e20;
    // two-photon continua
    const double t2HI = hydro2phot(nu, 1, c.tp, Ne);
    const double t2HeII = hydro2phot(nu, 2, c.tp, Ne);
    double t2HeI = 0.;
    const int j = a.he2Row[i - 1];
    if (j > 0) {
        const double* yd = a.hei2phot;
        const double y = nu / 1.514;
        const double f1 = (yd[2 * j + 1] - yd[2 * (j - 1) + 1]) / (yd[2 * j] - yd[2 * (j - 1)]);
        const double f2 = yd[2 * (j - 1) + 1] - yd[2 * (j - 1)] * f1;
        double Ay = f1 * y + f2;
        if (Ay < 9.2163086E-03) Ay = 0.;
        t2HeI = c.aEffHeI * Ay * 0.66262 * y / 51.3;
    }
    double eHI = (gam[0] + t2HI) * c.abH * c.XH2 * Ne + gHv * Ne;
    double eHeI = c.heOn ? (gam[1] + t2HeI) * c.abHe * c.XHe2 * Ne : 0.;
    double eHeII = c.heOn ? (gam[2] + t2HeII) * c.abHe * c.XHe3 * Ne : 0.;
    if (eHI < 1.e-15) eHI = 0.;
    if (eHeI < 1.e-15) eHeI = 0.;
    if (eHeII < 1.e-15) eHeII = 0.;
    // grains below sublimation radiate Cabs B_nu(Tdust) (sizes outer, species inner)
    double sDust = 0.;
    if (a.nDS > 0) {
        double acc = 0.;
        const double wf = a.widFlx[i - 1];
        for (int ai = 0; ai < a.nDA; ++ai)
            for (int s = 0; s < a.nDS; ++s) {
                const double T = sTd[s * a.nDA + ai];
                if (!(T < a.dTsub[s])) continue;
                const double cab = a.dCabs[((size_t)s * a.nDA + ai) * a.B + (i - 1)];
                const int q = a.qActive ? a.qIdx[s * a.nDA + ai] : -1;
                if (q >= 0) {  // quantum heating: the grain's temperature distribution
                    const double* qt = a.qT + ((size_t)c.cell * a.nQ + q) * a.nTb;
                    const double* qp = a.qP + ((size_t)c.cell * a.nQ + q) * a.nTb;
                    for (int iT = 0; iT < a.nTb; ++iT)
                        acc = acc + cab * moc_planck_h(nu, qt[iT]) * wf * a.dWeight[ai] * a.dAbun[s] * qp[iT];
                } else {
                    acc = acc + cab * moc_planck_h(nu, T) * wf * a.dWeight[ai] * a.dAbun[s];
                }
            }
        sDust = acc * c.dustScale;
    }
    const double w = a.widFlx[i - 1];
    const double sHI = kCRyd * w * eHI / 1.e15;
    double sHeI = kCRyd * w * eHeI / 1.e15;
    double sHeII = kCRyd * w * eHeII / 1.e15;
    if (c.heOn) {
        if (i == a.j2TsP) sHeI += c.corr;
        for (int k = 0; k < 4; ++k)
            if (i == a.lymP[k]) sHeII += c.ly[k];
    }
    if (a.nDS > 0) return sHI + sHeI + sHeII + sDust;
    return sHI + sHeI + sHeII;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The diffuse continuum and its CDF, one workgroup of four waves per cell;
// the cell's line emissivity comes from k_em_lines.  Registers capped at 128
// (four waves per SIMD, a few hundred bytes of spills): 3 % faster on the
// 32^3 nebula than the compiler's 226.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_emission(MocEmissionArgs a) {
    __shared__ double sX[MOC_MAXE][MOC_MAXS + 1];
    __shared__ double sLoc[3][MOC_EM_MAXLIM];
    __shared__ double sFit[3][MOC_EM_MAXLIM / 2];
    __shared__ double sWave[kThreads / 64];
    __shared__ double sTd[MOC_EM_MAXGRAIN];   // grain temperatures of the cell [species][size]
    __shared__ double sF[MOC_EM_MAXFB];       // heavy-ion Milne factors of a pass
    const int cell = blockIdx.x + 1;
    if (moc_unchanged(a.TeOld, a.NeOld, a.Te, a.Ne, cell)) return;  // uniform per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int C1 = a.C1, B = a.B;
    const double NeGrid = a.Ne[cell], Te = a.Te[cell];
    const double log10Ne = log10(NeGrid);
    const double Ne = NeGrid < 1.e-5 ? 1.e-5 : NeGrid;
    for (int k = tid; k < a.NE * a.NS; k += kThreads) {
        const int e = k % a.NE, ion = k / a.NE + 1;
        sX[e][ion] = a.ionDen[(size_t)k * C1 + cell];
    }
    // gamma fits at Te: log-log interpolation in T (fb_ff :389-432)
    for (int k = 0; k < 3; ++k) {
        const int nl = a.nlim[k];
        const double* lg = a.lg[k];
        for (int i = tid; i < nl; i += kThreads) {
            double v;
            if (Te <= a.tk[0]) v = lg[i];
            else if (Te >= a.tk[a.ntk - 1]) v = lg[(size_t)(a.ntk - 1) * nl + i];
            else {
                int lo = 0, hi = a.ntk - 1;   // tk[lo] <= Te < tk[hi]
                while (hi - lo > 1) {
                    const int m = (lo + hi) >> 1;
                    if (a.tk[m] <= Te) lo = m; else hi = m;
                }
                const double factor = log10(Te / a.tk[lo]);
                const double s = (lg[(size_t)hi * nl + i] - lg[(size_t)lo * nl + i]) / log10(a.tk[hi] / a.tk[lo]);
                v = lg[(size_t)lo * nl + i] + s * factor;
            }
            sLoc[k][i] = v;
        }
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k)
        for (int p = tid + 1; p <= a.nlim[k] / 2; p += kThreads) {
            const int ilow = a.edge[k][2 * p - 2], iup = a.edge[k][2 * p - 1];
            sFit[k][p - 1] = (sLoc[k][2 * p - 1] - sLoc[k][2 * p - 2]) / (a.nu[iup - 1] - a.nu[ilow - 1]);
        }
    CellCtx c;
    c.Te = Te;
    c.Ne = Ne;
    c.sqrTe = sqrt(Te);
    c.abH = a.abund[a.hSlot];
    c.XH2 = sX[a.hSlot][2];
    c.heOn = a.heSlot >= 0 && a.xsHeI && a.xsHeII;
    c.abHe = c.heOn ? a.abund[a.heSlot] : 0.;
    c.XHe2 = c.heOn ? sX[a.heSlot][2] : 0.;
    c.XHe3 = c.heOn ? sX[a.heSlot][3] : 0.;
    c.aEffHeI = 6.23 * pow(Te / 10000., -0.827);
    c.g1 = gaunt_setup(0., log10(Te));
    c.g2 = gaunt_setup(0.30103, log10(Te));
    c.tp = two_phot_setup(Te);
    c.cell = cell;
    c.dustScale = 0.;
    if (a.nDS > 0) {
        for (int k = tid; k < a.nDS * a.nDA; k += kThreads) {
            const int s = k / a.nDA, ai = k % a.nDA;
            sTd[k] = a.Tdust[((size_t)cell * (a.nDA + 1) + ai + 1) * (a.nDS + 1) + s + 1];
        }
        c.dustScale = 6.6262e-27 * 4. * kPi * 1.e25 * kCRyd;
        c.dustScale = c.dustScale * a.Ndust[cell] / a.Hden[cell];
    }
    c.corr = 0.;
    for (int k = 0; k < 4; ++k) c.ly[k] = 0.;
    if (c.heOn) {
        const double T4 = Te * 1.e-4;
        c.corr = 27.2 * pow(T4, -0.678) * Ne * c.abHe * c.XHe2 * 1.45673 * kHcRyd * 1e11;
        c.ly[3] = pow(10., -0.792 * log10(Te) + 6.01) * Ne * c.XHe3 * c.abHe;
        for (int k = 0; k < 3; ++k) c.ly[k] = a.heiiLyman[2 * k] * c.ly[3];
    }
    // ---- continuum CDF: scan of the bin energies, 256 bins per step
    double* row = a.recPDF + (size_t)cell * B;
    double carry = 0.;
    for (int b0 = 0; b0 < B; b0 += kThreads) {
        const int b = b0 + tid;
        const double nuA = a.nu[b0];
        for (int n = tid; n < a.nFb; n += kThreads) {
            const int lo = a.fbLo[n], sl = a.fbSlot[n];
            sF[n] = lo - 1 <= b0 ? exp((-nuA + a.nu[lo - 1]) * kHcRydK / Te) * a.fbG[n] *
                                       sX[sl][a.fbIon[n] + 1] * a.abund[sl]
                                 : 0.;
        }
        __syncthreads();
        const double ePass = b < B ? exp((-a.nu[b] + nuA) * kHcRydK / Te) : 0.;
        const double v = b < B ? bin_energy(a, c, sLoc, sFit, sX, sTd, b + 1, b0, ePass, sF) : 0.;
        double sc = v;
        for (int off = 1; off < 64; off <<= 1) {
            const double t = __shfl_up(sc, off, 64);
            if (lane >= off) sc += t;
        }
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
    const double total = carry;
    for (int b = tid; b < B; b += kThreads) row[b] = b == B - 1 ? 1. : row[b] / total;
    if (tid == 0) {
        const double L = a.lineSum[cell];
        a.totalLines[cell] = L / (total + L);
    }
}

// ---------------------------------------------------------------------------
// The continuum of the mixed-precision form: the per-bin energies in FP32 (the
// fits, exponentials, Gaunt factors and sums over ions), the CDF scan and its
// normalisation in FP64, the row written in FP32 (the transport's copy).  The
// constant factors of each term are folded so that no intermediate product
// leaves the FP32 range while its term can still matter: the Milne terms take
// 4 pi kMilne hcRyd 1e40 (times 2 or 0.5) as one factor, and the heavy-ion
// weights carry 1e25 that their common factor takes back.  Per-cell set-up
// (the gamma fits at T_e, the Gaunt coefficients, the two-photon and He line
// factors) stays in FP64 and is rounded once.
constexpr double kScaleFb = 1.e25;

// hydro2phot in FP32; hPlanck 1e27 folded into 6.6262
__device__ float hydro2phot_f32(float nu, int Z, float aEffH, float QH, float aEffHeII, float Ne) {
    const float nu0 = Z == 1 ? 0.7496f : 3.00f;
    if (nu >= nu0) return 0.f;
    const float y = nu / nu0;
    const float shape = y * (1.f - y) * (1.f - powf(4.f * y * (1.f - y), 0.8f)) +
                        0.88f * powf(y * (1.f - y), 1.53f) * powf(4.f * y * (1.f - y), 0.8f);
    if (Z == 1) {
        const float Ay = 202.0f * shape;
        const float gNu = 6.6262f * y * Ay / 8.2249f;
        return aEffH * gNu / (1.f + (Ne * QH) / 8.23f);
    }
    const float Ay = 64.f * 0.9994667f * 202.0f * shape;
    const float gNu = 6.6262f * y * Ay / (8.226f * 64.f);
    return aEffHeII * gNu;
}

struct CellCtxF {
    float Te, Ne, sqrTe, kT;      // kT: hcRyd/k / Te
    float abHXH2, abHeXHe2, abHeXHe3;
    bool heOn;
    float aEffHeI;
    GauntF g1, g2;
    float aEffH, QH, aEffHeII;
    float corr, ly[4];
};

// bin_energy in FP32 (gas only: no grain term)
__device__ float bin_energy_f32(const MocEmissionArgs& a, const CellCtxF& c, const float (*sLoc)[MOC_EM_MAXLIM],
                                const float (*sFit)[MOC_EM_MAXLIM / 2], const double (*sX)[MOC_MAXS + 1],
                                int i, int b0, float ePass, const float* sF) {
    constexpr float kFourPi = (float)(4. * kPi);
    constexpr float kM2 = (float)(4. * kPi * 2. * kHcRyd * kMilne * 1.e40);
    constexpr float kM05 = (float)(4. * kPi * 0.5 * kHcRyd * kMilne * 1.e40);
    constexpr float kMH = (float)(4. * kPi * kMilne * kHcRyd * 1.e40 / kScaleFb);
    constexpr float kFF = (float)(4. * kPi * 54.43);
    const double nuD = a.nu[i - 1];  // frequency differences are taken in FP64 (no cancellation)
    const float nu = (float)nuD;
    const float Ne = c.Ne;
    float gam[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 3; ++k) {
        const int p = a.gpair[k * a.B + i - 1];
        if (p > 0) {
            const int iup = a.edge[k][2 * p - 1];
            gam[k] = exp10f(sLoc[k][2 * p - 1] + sFit[k][p - 1] * (float)(nuD - a.nu[iup - 1]));
        }
    }
    const float xlf = (float)a.log10nu[i - 1];
    const float ex0 = expf(-nu * c.kT);
    const float ff1 = kFF * 1.f * gaunt_eval_f32(c.g1, xlf) * ex0 / c.sqrTe;
    const float ff2 = kFF * 4.f * gaunt_eval_f32(c.g2, xlf) * ex0 / c.sqrTe;
    const float factor = nu * nu * nu / (c.Te * c.sqrTe);
    if (i >= a.HIP) {
        const float ex = expf((float)(a.nu[a.HIP - 1] - nuD) * c.kT);
        gam[0] = kM2 * (float)a.xsHI[i - a.HIP] * factor * ex + ff1;
        if (c.heOn && i >= a.edge[1][a.nlim[1] - 1]) {
            const float sI = i >= a.HeIP ? (float)a.xsHeI[i - a.HeIP] : 0.f;
            const float exI = expf((float)(a.nu[a.HeIP - 1] - nuD) * c.kT);
            gam[1] = kM05 * sI * factor * exI + ff1;
            if (i >= a.edge[2][a.nlim[2] - 1]) {
                const float sII = i >= a.HeIIP ? (float)a.xsHeII[i - a.HeIIP] : 0.f;
                const float exII = expf((float)(a.nu[a.HeIIP - 1] - nuD) * c.kT);
                gam[2] = kM2 * sII * factor * exII + ff2;
            }
        }
    }
    float gFac = 0.f, gDir = 0.f;
    for (int n = 0; n < a.nFb; ++n) {
        const int lo = a.fbLo[n];
        if (i < lo || i >= a.fbHi[n]) continue;
        const float xs = (float)a.fbXs[a.fbOff[n] + (i - lo)];
        if (lo - 1 <= b0) {
            gFac += xs * sF[n];
        } else {
            const int s = a.fbSlot[n];
            const float ex = expf((float)(a.nu[lo - 1] - nuD) * c.kT);
            gDir += ex * xs * (float)(a.fbG[n] * sX[s][a.fbIon[n] + 1] * a.abund[s] * kScaleFb);
        }
    }
    const float gHv = (ePass * gFac + gDir) * kMH * factor;
    const float t2HI = hydro2phot_f32(nu, 1, c.aEffH, c.QH, c.aEffHeII, Ne);
    const float t2HeII = hydro2phot_f32(nu, 2, c.aEffH, c.QH, c.aEffHeII, Ne);
    float t2HeI = 0.f;
    const int j = a.he2Row[i - 1];
    if (j > 0) {
        const double* yd = a.hei2phot;
        const float y = nu / 1.514f;
        const float f1 = (float)((yd[2 * j + 1] - yd[2 * (j - 1) + 1]) / (yd[2 * j] - yd[2 * (j - 1)]));
        const float f2 = (float)yd[2 * (j - 1) + 1] - (float)yd[2 * (j - 1)] * f1;
        float Ay = f1 * y + f2;
        if (Ay < 9.2163086E-03f) Ay = 0.f;
        t2HeI = c.aEffHeI * Ay * 0.66262f * y / 51.3f;
    }
    float eHI = (gam[0] + t2HI) * c.abHXH2 * Ne + gHv * Ne;
    float eHeI = c.heOn ? (gam[1] + t2HeI) * c.abHeXHe2 * Ne : 0.f;
    float eHeII = c.heOn ? (gam[2] + t2HeII) * c.abHeXHe3 * Ne : 0.f;
    if (eHI < 1.e-15f) eHI = 0.f;
    if (eHeI < 1.e-15f) eHeI = 0.f;
    if (eHeII < 1.e-15f) eHeII = 0.f;
    const float w = (float)(a.widFlx[i - 1] * (kCRyd / 1.e15));
    const float sHI = w * eHI;
    float sHeI = w * eHeI;
    float sHeII = w * eHeII;
    if (c.heOn) {
        if (i == a.j2TsP) sHeI += c.corr;
        for (int k = 0; k < 4; ++k)
            if (i == a.lymP[k]) sHeII += c.ly[k];
    }
    (void)kFourPi;
    return sHI + sHeI + sHeII;
}

// The continuum and its CDF in the mixed form (gas runs), one workgroup of
// four waves per cell, as k_emission; the row goes to the FP32 recPDF.
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_emission_f32(MocEmissionArgs a) {
    __shared__ double sX[MOC_MAXE][MOC_MAXS + 1];
    __shared__ double sLocD[3][MOC_EM_MAXLIM];
    __shared__ float sLoc[3][MOC_EM_MAXLIM];
    __shared__ float sFit[3][MOC_EM_MAXLIM / 2];
    __shared__ double sWave[kThreads / 64];
    __shared__ float sF[MOC_EM_MAXFB];
    const int cell = blockIdx.x + 1;
    if (moc_unchanged(a.TeOld, a.NeOld, a.Te, a.Ne, cell)) return;  // uniform per workgroup
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int C1 = a.C1, B = a.B;
    const double NeGrid = a.Ne[cell], Te = a.Te[cell];
    const double Ne = NeGrid < 1.e-5 ? 1.e-5 : NeGrid;
    for (int k = tid; k < a.NE * a.NS; k += kThreads) {
        const int e = k % a.NE, ion = k / a.NE + 1;
        sX[e][ion] = a.ionDen[(size_t)k * C1 + cell];
    }
    for (int k = 0; k < 3; ++k) {  // gamma fits at Te (fb_ff :389-432), in FP64
        const int nl = a.nlim[k];
        const double* lg = a.lg[k];
        for (int i = tid; i < nl; i += kThreads) {
            double v;
            if (Te <= a.tk[0]) v = lg[i];
            else if (Te >= a.tk[a.ntk - 1]) v = lg[(size_t)(a.ntk - 1) * nl + i];
            else {
                int lo = 0, hi = a.ntk - 1;
                while (hi - lo > 1) {
                    const int m = (lo + hi) >> 1;
                    if (a.tk[m] <= Te) lo = m; else hi = m;
                }
                const double factor = log10(Te / a.tk[lo]);
                const double s = (lg[(size_t)hi * nl + i] - lg[(size_t)lo * nl + i]) / log10(a.tk[hi] / a.tk[lo]);
                v = lg[(size_t)lo * nl + i] + s * factor;
            }
            sLocD[k][i] = v;
            sLoc[k][i] = (float)v;
        }
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k)
        for (int p = tid + 1; p <= a.nlim[k] / 2; p += kThreads) {
            const int ilow = a.edge[k][2 * p - 2], iup = a.edge[k][2 * p - 1];
            sFit[k][p - 1] = (float)((sLocD[k][2 * p - 1] - sLocD[k][2 * p - 2]) / (a.nu[iup - 1] - a.nu[ilow - 1]));
        }
    CellCtxF c;
    c.Te = (float)Te;
    c.Ne = (float)Ne;
    c.sqrTe = (float)sqrt(Te);
    c.kT = (float)(kHcRydK / Te);
    c.abHXH2 = (float)(a.abund[a.hSlot] * sX[a.hSlot][2]);
    c.heOn = a.heSlot >= 0 && a.xsHeI && a.xsHeII;
    const double abHe = c.heOn ? a.abund[a.heSlot] : 0.;
    const double XHe2 = c.heOn ? sX[a.heSlot][2] : 0., XHe3 = c.heOn ? sX[a.heSlot][3] : 0.;
    c.abHeXHe2 = (float)(abHe * XHe2);
    c.abHeXHe3 = (float)(abHe * XHe3);
    c.aEffHeI = (float)(6.23 * pow(Te / 10000., -0.827));
    c.g1 = gaunt_f32(gaunt_setup(0., log10(Te)));
    c.g2 = gaunt_f32(gaunt_setup(0.30103, log10(Te)));
    const TwoPhotT tp = two_phot_setup(Te);
    c.aEffH = (float)tp.aEffH;
    c.QH = (float)tp.QH;
    c.aEffHeII = (float)tp.aEffHeII;
    c.corr = 0.f;
    for (int k = 0; k < 4; ++k) c.ly[k] = 0.f;
    if (c.heOn) {
        const double T4 = Te * 1.e-4;
        c.corr = (float)(27.2 * pow(T4, -0.678) * Ne * abHe * XHe2 * 1.45673 * kHcRyd * 1e11);
        const double ly3 = pow(10., -0.792 * log10(Te) + 6.01) * Ne * XHe3 * abHe;
        c.ly[3] = (float)ly3;
        for (int k = 0; k < 3; ++k) c.ly[k] = (float)(a.heiiLyman[2 * k] * ly3);
    }
    float* row = a.recPDF32 + (size_t)cell * B;
    double carry = 0.;
    for (int b0 = 0; b0 < B; b0 += kThreads) {
        const int b = b0 + tid;
        const double nuA = a.nu[b0];
        for (int n = tid; n < a.nFb; n += kThreads) {
            const int lo = a.fbLo[n], sl = a.fbSlot[n];
            sF[n] = lo - 1 <= b0 ? (float)(exp((-nuA + a.nu[lo - 1]) * kHcRydK / Te) * a.fbG[n] *
                                           sX[sl][a.fbIon[n] + 1] * a.abund[sl] * kScaleFb)
                                 : 0.f;
        }
        __syncthreads();
        const float ePass = b < B ? expf((float)((-a.nu[b] + nuA) * kHcRydK / Te)) : 0.f;
        const double v = b < B ? (double)bin_energy_f32(a, c, sLoc, sFit, sX, b + 1, b0, ePass, sF) : 0.;
        double sc = v;
        for (int off = 1; off < 64; off <<= 1) {
            const double t = __shfl_up(sc, off, 64);
            if (lane >= off) sc += t;
        }
        if (lane == 63) sWave[wv] = sc;
        __syncthreads();
        double pre = carry;
        for (int k = 0; k < wv; ++k) pre += sWave[k];
        if (b < B) row[b] = (float)(pre + sc);
        double next = carry;
        for (int k = 0; k < kThreads / 64; ++k) next += sWave[k];
        __syncthreads();
        carry = next;
    }
    const double total = carry;
    for (int b = tid; b < B; b += kThreads) row[b] = b == B - 1 ? 1.f : (float)((double)row[b] / total);
    if (tid == 0) {
        const double L = a.lineSum[cell];
        a.totalLines[cell] = L / (total + L);
    }
}

// n-level line emission of the ions of team class c (nlevel_team_class), P =
// 64 / W ions at a time, one per team of W lanes; with lineAcc (the line
// luminosities) only the ions k0 .. k0 + nk - 1 of the launch's group add
// their lines to acc.  Returns this lane's share of the line sum (the team's
// sum on the team's first lane).
template <int NL>
__device__ double team_lines(const MocEmissionArgs& a, const double (*sX)[MOC_MAXS + 1], int cnt,
                             const int* list, double Te, double Ne, int lane, double* acc = nullptr,
                             double w = 0., int L = 0, int k0 = 0, int nk = 0) {
    constexpr int W = NL <= 8 ? 8 : (NL <= 16 ? 16 : 32), P = 64 / W;
    double lines = 0.;
    for (int b0 = 0; b0 < cnt; b0 += P) {
        const int bt = b0 + lane / W;
        const int k = list[bt < cnt ? bt : b0];  // idle teams repeat one
        // the line luminosities take one ion group per launch: skip rounds
        // without an ion of the group (the lists ascend in k)
        if (acc && !__any(bt < cnt && k >= k0 && k < k0 + nk)) continue;
        const int s = a.forSlot[k], ion = a.forIon[k];
        const double X = sX[s][ion];
        const double up = (ion < a.forTop[k] && X > 0.) ? sX[s][ion + 1] / X : 0.;
        double* la = nullptr;
        if (acc && bt < cnt && k >= k0 && k < k0 + nk) la = acc + (size_t)(k - k0) * L * L;
        const double v = nlevel_team<NL>(a.bal, a.forAtom[k], Te, Ne, up, true, a.abund[s] * X, 1.9865e9, la, w, L);
        if ((lane & (W - 1)) == 0 && bt < cnt) lines += v;
    }
    return lines;
}

// The lines of a cell (RecLinesEmission :856-1153, forLines :1481-1526): the
// hydrogenic elements, the 34 He I lines and the n-level ions of up to
// kNlevLane levels are separate tasks, one per lane of a single wave per
// cell; the larger n-level ions are then solved by teams of 8, 16 or 32
// lanes.  Their sum goes to lineSum for the CDF kernel.
template <bool TEAMS>
__global__ void __launch_bounds__(64) k_em_lines(MocEmissionArgs a) {
    __shared__ double sX[MOC_MAXE][MOC_MAXS + 1];
    const int cell = blockIdx.x + 1;
    if (moc_unchanged(a.TeOld, a.NeOld, a.Te, a.Ne, cell)) return;
    const int lane = threadIdx.x;
    const int C1 = a.C1;
    const double NeGrid = a.Ne[cell], Te = a.Te[cell];
    const double log10Ne = log10(NeGrid);
    const double Ne = NeGrid < 1.e-5 ? 1.e-5 : NeGrid;
    for (int k = lane; k < a.NE * a.NS; k += 64) {
        const int e = k % a.NE, ion = k / a.NE + 1;
        sX[e][ion] = a.ionDen[(size_t)k * C1 + cell];
    }
    __syncthreads();
    const bool heOn = a.heSlot >= 0 && a.xsHeI && a.xsHeII;
    const double abHe = heOn ? a.abund[a.heSlot] : 0.;
    const double XHe2 = heOn ? sX[a.heSlot][2] : 0.;
    double lines = 0.;
    const int nTasks = a.nHyd + kHeILines + a.nFor;
    for (int t = lane; t < nTasks; t += 64) {
        if (t < a.nHyd) {
            const int s = a.hydSlot[t], z = a.hydZ[t];
            lines += hydro_element(a, z, sX[s][z + 1], a.abund[s], Te, Ne, log10Ne);
        } else if (t < a.nHyd + kHeILines) {
            if (heOn) lines += hei_line(a, t - a.nHyd, Te, Ne, NeGrid, XHe2, abHe);
        } else {
            const int k = t - a.nHyd - kHeILines;
            const int atom = a.forAtom[k];
            if (nlevel_count(a.bal, atom) > kNlevLane) continue;
            const int s = a.forSlot[k], ion = a.forIon[k];
            const double X = sX[s][ion];
            const double up = (ion < a.forTop[k] && X > 0.) ? sX[s][ion + 1] / X : 0.;
            lines += nlevel_fixed<kNlevLane>(a.bal, atom, Te, Ne, up, true, a.abund[s] * X, 1.9865e9);
        }
    }
    if constexpr (TEAMS) {
        lines += team_lines<8>(a, sX, a.nForT[0], a.forT[0], Te, Ne, lane);
        lines += team_lines<10>(a, sX, a.nForT[1], a.forT[1], Te, Ne, lane);
        lines += team_lines<13>(a, sX, a.nForT[2], a.forT[2], Te, Ne, lane);
        lines += team_lines<16>(a, sX, a.nForT[3], a.forT[3], Te, Ne, lane);
        lines += team_lines<17>(a, sX, a.nForT[4], a.forT[4], Te, Ne, lane);
        lines += team_lines<27>(a, sX, a.nForT[5], a.forT[5], Te, Ne, lane);
    }
    lines = wave_sum(lines);
    if (lane == 0) a.lineSum[cell] = lines;
}

// Line luminosities of outputGas (output_mod.f90:397-432): over the cells
// packets reached, sum of each n-level line's emissivity (forLines, per
// transition) times H dV, and of the Hbeta emissivity times H dV where that
// exceeds 1e-35.  One wave per block walks a chunk of cells; lane k owns
// n-level ion k0 + k (and k + 64, ...) of the launch's ion group and
// accumulates its lines in LDS; the block writes its sums, which the host adds
// block by block.  The first group also carries Hbeta.
template <bool TEAMS>
__global__ void __launch_bounds__(64) k_line_lum(MocEmissionArgs a, const double* __restrict__ vol,
                                                 const int* __restrict__ black,
                                                 const double* __restrict__ H, int nChunks, int L,
                                                 int k0, int nk, double* __restrict__ part) {
    __shared__ double sX[MOC_MAXE][MOC_MAXS + 1];
    extern __shared__ double acc[];  // [nk][L][L]: ions k0 .. k0+nk-1
    const int lane = threadIdx.x;
    const int C1 = a.C1, nL = nk * L * L;
    for (int k = lane; k < nL; k += 64) acc[k] = 0.;
    double hb = 0.;
    const int per = (C1 - 1 + nChunks - 1) / nChunks;
    const int c0 = 1 + blockIdx.x * per;
    const int c1 = c0 + per < C1 ? c0 + per : C1;
    const int hs = a.hSlot;
    for (int cell = c0; cell < c1; ++cell) {
        if (black && black[cell] > 0) continue;
        __syncthreads();  // sX of the previous cell is no longer read
        for (int k = lane; k < a.NE * a.NS; k += 64) {
            const int e = k % a.NE, ion = k / a.NE + 1;
            sX[e][ion] = a.ionDen[(size_t)k * C1 + cell];
        }
        __syncthreads();
        // outputGas takes the grid's n_e as it is (output_mod.f90:321), where
        // the emission floors it at 1e-5
        const double Ne = a.Ne[cell], Te = a.Te[cell];
        const double log10Ne = log10(Ne);
        const double w = H[cell] * vol[cell];
        for (int kk = lane; kk < nk; kk += 64) {
            const int k = k0 + kk;
            if (nlevel_count(a.bal, a.forAtom[k]) > kNlevLane) continue;
            const int s = a.forSlot[k], ion = a.forIon[k];
            const double X = sX[s][ion];
            const double up = (ion < a.forTop[k] && X > 0.) ? sX[s][ion + 1] / X : 0.;
            nlevel_fixed<kNlevLane>(a.bal, a.forAtom[k], Te, Ne, up, true, a.abund[s] * X, 1.9865e9,
                             acc + (size_t)kk * L * L, w, L);
        }
        if constexpr (TEAMS) {  // ions of more than kNlevLane levels, by teams of lanes
            team_lines<8>(a, sX, a.nForT[0], a.forT[0], Te, Ne, lane, acc, w, L, k0, nk);
            team_lines<10>(a, sX, a.nForT[1], a.forT[1], Te, Ne, lane, acc, w, L, k0, nk);
            team_lines<13>(a, sX, a.nForT[2], a.forT[2], Te, Ne, lane, acc, w, L, k0, nk);
            team_lines<16>(a, sX, a.nForT[3], a.forT[3], Te, Ne, lane, acc, w, L, k0, nk);
            team_lines<17>(a, sX, a.nForT[4], a.forT[4], Te, Ne, lane, acc, w, L, k0, nk);
            team_lines<27>(a, sX, a.nForT[5], a.forT[5], Te, Ne, lane, acc, w, L, k0, nk);
        }
        if (lane == 0 && k0 == 0 && hs >= 0 && a.NS > 1) {  // hydroLines(1,4,2) (RecLinesEmission)
            const HydroLines h = hydro_setup(a, 1, sX[hs][2], a.abund[hs], Te, Ne, log10Ne, true);
            const double v = hl_line(a, h.f, h.Ne, h.Hb, 4, 2, true) * H[cell] * vol[cell];
            if (v > 1.e-35) hb += v;
        }
    }
    __syncthreads();
    double* out = part + (size_t)blockIdx.x * (nL + 1);
    for (int k = lane; k < nL; k += 64) out[k] = acc[k];
    if (lane == 0) out[nL] = hb;
}

}  // namespace

int moc_line_lum_group(const MocEmissionArgs& a, int L) {
    (void)a;
    const int g = (int)(kLineLumLds / ((size_t)L * L * sizeof(double)));
    return g < 1 ? 1 : g;
}

int moc_launch_line_lum(moc_ctx* ctx, const MocEmissionArgs& a, const double* vol, const int* black,
                        const double* H, int nChunks, int L, int k0, int nk, double* part) {
    const size_t lds = (size_t)nk * L * L * sizeof(double);
    if (nk < (a.nFor > 0 ? 1 : 0) || k0 < 0 || k0 + nk > a.nFor || lds > (size_t)kLineLumLds) {
        ctx->err = "line luminosities: ion group " + std::to_string(k0) + "+" + std::to_string(nk) +
                   " of " + std::to_string(a.nFor) + " does not fit";
        return MOC_ERR_ARG;
    }
    const int maxLev = a.bal.maxLev;
    const dim3 grid((unsigned)nChunks), block(64);
    if (maxLev <= kNlevLane) hipLaunchKernelGGL(k_line_lum<false>, grid, block, lds, ctx->stream, a, vol, black, H, nChunks, L, k0, nk, part);
    else hipLaunchKernelGGL(k_line_lum<true>, grid, block, lds, ctx->stream, a, vol, black, H, nChunks, L, k0, nk, part);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("line luminosities launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

int moc_launch_emission(moc_ctx* ctx, const MocEmissionArgs& a) {
    if (a.nCells < 1) return MOC_OK;
    const int maxLev = a.bal.maxLev;
    const dim3 grid((unsigned)a.nCells);
    if (maxLev <= kNlevLane) hipLaunchKernelGGL(k_em_lines<false>, grid, dim3(64), 0, ctx->stream, a);
    else hipLaunchKernelGGL(k_em_lines<true>, grid, dim3(64), 0, ctx->stream, a);
    if (a.recPDF32) hipLaunchKernelGGL(k_emission_f32, grid, dim3(kThreads), 0, ctx->stream, a);
    else hipLaunchKernelGGL(k_emission, grid, dim3(kThreads), 0, ctx->stream, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("emission launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

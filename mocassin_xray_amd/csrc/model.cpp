// Native model set-up (host C++): the frequency mesh, the source, the
// cross-section stack, the balance and emission tables and the starting
// state of one gas model, from the reference's data files, behind the C ABI
// of mocmodel.h.  Each builder follows the reference routine it restates
// (cited at the function); the arithmetic is FP64 except where the
// reference rounds a value to single precision, which is kept.

#include "../../include/mocmodel.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <tuple>

#include "model_internal.h"

using namespace mocmodel;

namespace {

constexpr double kRydToEV = 13.6056981;    // constants_mod.f90:51
constexpr double kHIIonPot = 0.99946;      // constants_mod.f90:15
constexpr double kKShellLimit = 7.35e4;    // constants_mod.f90:40 [Ryd]
constexpr double kHcRyd = 2.1799153e-11;   // update_mod.f90:64 [erg]
constexpr double kPiRef = 3.141592654;     // constants_mod.f90:45
constexpr double kHPlanck = 6.6262e-27;    // constants_mod.f90:14
constexpr double kHcRydK = 157893.94;      // continuum_mod.f90:14
constexpr double kRadio = 1.489434e-6;     // constants_mod.f90:19 [Ryd]
constexpr double kMeC2Ryd = 37557.66267;   // ph_mod.f90:1933
constexpr double kHalfRe2 = 3.970393838e-26;  // ph_mod.f90:1958
constexpr double kHEdge = 0.999434;        // grid_mod.f90:194
constexpr int kNAngles = 180;

const int kNTot[30] = {1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5,
                       5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7};
const int kNInn[30] = {0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3,
                       3, 3, 3, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5};
const int kLevel[7] = {0, 0, 1, 0, 1, 2, 0};
// setPointers calls setShells in this element order (hydro_mod.f90:584-641)
const int kShellOrder[28] = {6,  8,  7,  10, 11, 12, 13, 14, 15, 16, 17, 26, 18, 19,
                             20, 21, 22, 23, 24, 25, 9,  3,  4,  5,  27, 28, 29, 30};
// H-series edges of the gas mesh (grid_mod.f90:169-171)
const double kSeriesEdges[17] = {0.0069,  0.0083,  0.01,    0.0123,  0.0156,  0.0204,
                                 0.0278,  0.04,    0.0625,  0.11117, 0.11610, 0.12248,
                                 0.13732, 0.24763, 0.24994, 0.26630, 0.29189};

// Charge-exchange fits (a [1e-9 cm^3/s], b, c, d) by (Z, stage), in the
// order photoionisation / ionBalance list them (update_mod.f90:1624-1690);
// the last two "Ca" rows go into argon's slots, as the reference writes them.
struct ChexRow {
    int z, ion;
    double a, b, c, d;
};
const ChexRow kChex[] = {
    {2, 1, 7.47e-6, 2.06, 9.93, -3.89},     {2, 2, 1.e-5, 0., 0., 0.},
    {3, 2, 1.26, 0.96, 3.02, -0.65},        {3, 3, 1.e-5, 0., 0., 0.},
    {4, 2, 1.e-5, 0., 0., 0.},              {4, 3, 1.e-5, 0., 0., 0.},
    {4, 4, 5.17, 0.82, -.69, -1.12},        {5, 2, 2.e-2, 0., 0., 0.},
    {5, 3, 1.e-5, 0., 0., 0.},              {5, 4, 5.27e-1, 0.76, -0.63, -1.17},
    {6, 1, 1.76e-9, 8.33, 4278.78, -6.41},  {6, 2, 1.67e-4, 2.79, 304.72, -4.07},
    {6, 3, 3.25, 0.21, 0.19, -3.29},        {6, 4, 332.46, -0.11, -0.995, -1.58e-3},
    {7, 1, 1.01e-3, -0.29, -0.92, -8.38},   {7, 2, 3.05e-1, 0.60, 2.65, -0.93},
    {7, 3, 4.54, 0.57, -0.65, -0.89},       {7, 4, 3.28, 0.52, -0.52, -0.19},
    {8, 1, 1.04, 3.15e-2, -0.61, -9.73},    {8, 2, 1.04, 0.27, 2.02, -5.92},
    {8, 3, 3.98, 0.26, 0.56, -2.62},        {8, 4, 2.52e-1, 0.63, 2.08, -4.16},
    {9, 2, 1.e-5, 0., 0., 0.},              {9, 3, 9.86, 0.29, -0.21, -1.15},
    {9, 4, 7.15e-1, 1.21, -0.70, -0.85},    {10, 2, 1.e-5, 0., 0., 0.},
    {10, 3, 14.73, 4.52e-2, -0.84, -0.31},  {10, 4, 6.47, 0.54, 3.59, -5.22},
    {11, 2, 1.e-5, 0., 0., 0.},             {11, 3, 1.33, 1.15, 1.20, -0.32},
    {11, 4, 1.01e-1, 1.34, 10.05, -6.41},   {12, 2, 8.58e-5, 2.49e-3, 2.93e-2, -4.33},
    {12, 3, 6.49, 0.53, 2.82, -7.63},       {12, 4, 6.36, 0.55, 3.86, -5.19},
    {13, 2, 1.e-5, 0., 0., 0.},             {13, 3, 7.11e-5, 4.12, 1.72e4, -22.24},
    {13, 4, 7.52e-1, 0.77, 6.24, -5.67},    {14, 2, 1.23, 0.24, 3.17, 4.18e-3},
    {14, 3, 4.900e-1, -8.74e-2, -0.36, -0.79}, {14, 4, 7.58, 0.37, 1.06, -4.09},
    {16, 1, 3.82e-7, 11.10, 2.57e4, -8.22}, {16, 2, 1.e-5, 0., 0., 0.},
    {16, 3, 2.29, 4.02e-2, 1.59, -6.06},    {16, 4, 6.44, 0.13, 2.69, -5.69},
    {18, 2, 1.e-5, 0., 0., 0.},             {18, 3, 4.57, 0.27, -0.18, -1.57},
    {18, 4, 6.37, 2.12, 10.21, -6.22},      {18, 3, 3.17e-2, 2.12, 12.06, -0.40},
    {18, 4, 2.68, 0.69, -0.68, -4.47},      {26, 2, 1.26, 7.72e-2, -0.41, -7.31},
    {26, 3, 3.42, 0.51, -2.06, -8.99},
};

thread_local std::string g_error;

// ------------------------------------------------------------------ helpers
// locate (interpolation_mod.f90:49-106): 1-based bracket index, 0 below the
// first point, n above the last, the index of an exact hit
int locate(const std::vector<double>& xa, double x) {
    const int n = (int)xa.size();
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    auto lb = std::lower_bound(xa.begin(), xa.end(), x);
    if (lb != xa.end() && *lb == x) return (int)(lb - xa.begin()) + 1;
    return (int)(std::upper_bound(xa.begin(), xa.end(), x) - xa.begin());
}

// the summation order of numpy's add.reduce over a contiguous row
double pairwise_sum(const double* a, size_t n) {
    if (n < 8) {
        double r = 0.0;
        for (size_t i = 0; i < n; ++i) r += a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        size_t i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    size_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

// getOuterShell (hydro_mod.f90:89-132)
int outer_shell(int z, int nelec) {
    int out = kNTot[nelec - 1];
    if (z == nelec && z > 18) out = 7;
    if (z == nelec + 1 && (z == 20 || z == 21 || z == 22 || z == 25 || z == 26)) out = 7;
    return out;
}

// the statistical weights of getOuterShell (hydro_mod.f90:133-179)
void stat_weights(int z, int nel, int* g0o, int* g1o) {
    static const int gl[30] = {2, 1, 2, 1, 6, 9, 4, 9, 6, 1, 2,  1,  6,  9, 4,
                               9, 6, 1, 2, 1, 10, 21, 28, 7, 6, 25, 28, 21, 2, 1};
    static const int glh[12] = {10, 21, 28, 25, 6, 25, 28, 21, 10, 1, 2, 1};  // nel 19..30
    auto GLH = [&](int n) { return glh[n - 19]; };
    auto special = [](int zz, int dflt) {
        switch (zz) {
            case 21: return 15;
            case 22: return 28;
            case 25: return 7;
            case 26: return 30;
            default: return dflt;
        }
    };
    int g0 = gl[nel - 1];
    int g1 = nel == 1 ? 1 : gl[nel - 2];
    if (z > 20 && nel >= 19 && z - nel >= 1) {
        g0 = GLH(nel);
        g1 = nel == 19 ? gl[nel - 2] : GLH(nel - 1);
    }
    if (z == nel) {
        if (z > 20 && nel >= 21) g1 = GLH(nel - 1);
        g1 = special(z, g1);
    }
    if (z - nel == 1) g0 = special(z, g0);
    *g0o = g0;
    *g1o = g1;
}

// phFitEl (ph_mod.f90:592-668): Verner partial cross-section [Mb] at E [eV]
double phfit_el(const AtomicData& d, int nz, int ne, int shell, double E) {
    if (nz < 1 || nz > 30 || ne < 1 || ne > nz) return 0.0;
    int nout = kNTot[ne - 1];
    if (nz == ne && nz > 18) nout = 7;
    if (nz == ne + 1 && (nz == 20 || nz == 21 || nz == 22 || nz == 25 || nz == 26)) nout = 7;
    if (shell > nout) return 0.0;
    auto p = [&](int i) { return d.ph1[PH1(i, nz - 1, ne - 1, shell - 1)]; };
    if (E < p(0)) return 0.0;
    const int nint = kNInn[ne - 1];
    double einn;
    if (nz == 15 || nz == 17 || nz == 19 || (nz > 20 && nz != 26)) einn = 0.0;
    else einn = ne < 3 ? 1.e30 : d.ph1[PH1(0, nz - 1, ne - 1, nint - 1)];
    if (shell < nout && shell > nint && E < einn) return 0.0;
    if (shell <= nint || E >= einn) {
        const double p1 = -p(4);
        const double y = E / p(1);
        const double q = -0.5 * p1 - kLevel[shell - 1] - 5.5;
        const double a = p(2) * (std::pow(y - 1.0, 2.0) + p(5) * p(5));
        const double b = std::sqrt(y / p(3)) + 1.0;
        return a * std::pow(y, q) * std::pow(b, p1);
    }
    auto p2 = [&](int i) { return d.ph2[((size_t)i * 30 + (nz - 1)) * 30 + (ne - 1)]; };
    const double p1 = -p2(3);
    const double q = -0.5 * p1 - 5.5;
    const double x = E / p2(0) - p2(5);
    const double zz = std::sqrt(x * x + p2(6) * p2(6));
    const double a = p2(1) * (std::pow(x - 1.0, 2.0) + p2(4) * p2(4));
    const double b = std::sqrt(zz / p2(2)) + 1.0;
    return a * std::pow(zz, q) * std::pow(b, p1);
}

// ------------------------------------------------------------------ mesh
// the ionisation edges the gas mesh resolves (initCartesianGrid,
// grid_mod.f90:170-209), each held in single precision there
std::vector<double> ion_edges(const AtomicData& d, const std::set<int>& on, int nstages,
                              double nuMax) {
    std::vector<double> out;
    for (int z : on)
        for (int ion = 1; ion <= std::min(z, nstages - 1); ++ion) {
            double e;
            if (z > 2) {
                const int nel = z - ion + 1;
                e = d.ph1[PH1(0, z - 1, nel - 1, outer_shell(z, nel) - 1)] / kRydToEV;
            } else if (z == 1) {
                e = 0.999434;
            } else {
                e = ion == 1 ? 1.80804 : 4.0;
            }
            e = (double)(float)e;
            if (e <= nuMax) out.push_back(e);
        }
    std::sort(out.begin(), out.end());
    return out;
}

// the gas log mesh of initCartesianGrid (grid_mod.f90:251-375): each H-series
// and ionisation edge adds (e - 0.0003, e, e + 0.0003), edges below nuMin pull
// the log part down to e - 0.0006, the rest log-spaced by repeated
// multiplication, sorted ascending with repeated values kept once (sortUp)
std::vector<double> gas_mesh(int nbins, double nuMin, double nuMax,
                             const std::vector<double>& edges, bool strict) {
    std::vector<double> pts;
    double lo = nuMin, hi = nuMax;
    if (nuMin < kRadio) pts.push_back(kRadio);
    for (double e : kSeriesEdges) {
        pts.push_back(e);
        pts.push_back(e - 0.0003);
        pts.push_back(e + 0.0003);
        if (e < lo) lo = e - 0.0006;
        if (e > hi) hi = e;
    }
    for (double e : edges)
        if (e < hi) {
            pts.push_back(e);
            pts.push_back(e - 0.0003);
            pts.push_back(e + 0.0003);
            if (e < lo) lo = e - 0.0006;
        }
    const int nCount = (int)pts.size() + 1;
    const int iCount = nbins - nCount + 1;
    if (iCount < 2)
        throw Error("nbins = " + std::to_string(nbins) + " leaves no room for the log mesh beside " +
                    std::to_string(nCount - 1) + " edge points");
    const double step = (std::log10(hi) - std::log10(lo)) / (iCount - 1);
    double prev = lo;
    pts.push_back(prev);
    for (int i = 0; i < iCount - 1; ++i) {
        prev = std::pow(10.0, std::log10(prev) + step);
        pts.push_back(prev);
    }
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    if (strict && (int)pts.size() != nbins)
        throw Error("mesh has " + std::to_string(pts.size()) + " distinct points, expected " +
                    std::to_string(nbins));
    return pts;
}

// widFlx (grid_mod.f90:378-382)
std::vector<double> widths_of(const std::vector<double>& nu) {
    const size_t n = nu.size();
    std::vector<double> w(n);
    w[0] = nu[1] - nu[0];
    for (size_t i = 1; i + 1 < n; ++i) w[i] = (nu[i + 1] - nu[i - 1]) / 2.0;
    w[n - 1] = nu[n - 1] - nu[n - 2];
    return w;
}

// getFlux('blackbody') (continuum_mod.f90:358-392)
std::vector<double> blackbody_flux(const std::vector<double>& nu, double T) {
    const double c = 0.5250229 / kHPlanck;
    std::vector<double> out(nu.size());
    for (size_t i = 0; i < nu.size(); ++i) {
        const double x = kHcRydK * nu[i] / T;
        const double n3 = std::pow(nu[i], 3.0);
        if (x > 86.0) {
            out[i] = c * n3 * std::exp(-x);
        } else {
            const double den = std::exp(x) - 1.0;
            out[i] = den > 0 ? c * n3 / den : 3.32154e-6 * (nu[i] * nu[i]) * T / kHPlanck;
        }
    }
    return out;
}

// ------------------------------------------------------------------ xsec stack
struct Shell {
    int elem, ion, shell, lo, hi;  // bins 1-based, [lo, hi]
    std::vector<double> sigma;     // cm^2
};

struct XSec {
    std::vector<std::pair<int, int>> species;  // (Z, stage) in opacity order
    std::vector<double> sigma;                 // [S][B]
    std::vector<Shell> shells;
    std::vector<double> brems;
    int clamped = 0;
};

// initXSecArray / makeOpacity (ph_mod.f90:204-517, 861-921), the shell bin
// ranges of setShells / setPointers (hydro_mod.f90:208-677)
XSec build_xsec(const AtomicData& d, const std::vector<double>& nu, const std::set<int>& on,
                int nstages) {
    const int B = (int)nu.size();
    const int kshellP = locate(nu, kKShellLimit);
    XSec x;
    auto lohi = [&](int& lo, int& hi) {
        if (lo < 1) {
            ++x.clamped;
            lo = 1;
        }
        hi = std::min(hi, B);
    };
    auto shell_of = [&](int z, int ion, int sh, int lo, int hi, int nz, int ne, bool clampE) {
        Shell s{z, ion, sh, lo, hi, {}};
        const double emin = d.ph1[PH1(0, nz - 1, ne - 1, sh - 1)];
        for (int b = lo; b <= hi; ++b) {
            double E = nu[b - 1] * kRydToEV;
            if (clampE) E = std::max(E, emin);
            s.sigma.push_back(phfit_el(d, nz, ne, sh, E) * 1e-18);
        }
        x.shells.push_back(std::move(s));
    };
    if (on.count(1)) {  // H I ground continuum (ph_mod.f90:389-394)
        int lo = locate(nu, kHIIonPot), hi = B;
        lohi(lo, hi);
        shell_of(1, 1, 1, lo, hi, 1, 1, true);
    }
    if (on.count(2)) {
        int lo = locate(nu, 1.80802), hi = B;  // He I singlet ground (:441-446)
        lohi(lo, hi);
        shell_of(2, 1, 1, lo, hi, 2, 2, false);
        lo = locate(nu, 4.0);  // He II (:456-461)
        hi = B;
        lohi(lo, hi);
        shell_of(2, 2, 1, lo, hi, 2, 1, true);
    }
    std::map<std::tuple<int, int, int>, std::pair<int, int>> eP;
    for (int z : kShellOrder) {
        if (!on.count(z)) continue;
        for (int ion = 1; ion <= z; ++ion) {
            const int nel = z - ion + 1;
            const int out = outer_shell(z, nel);
            double thres = 0.0;
            for (int sh = 1; sh <= out; ++sh) {
                thres = d.ph1[PH1(0, z - 1, nel - 1, sh - 1)] / kRydToEV;
                if (thres <= 0.1) {
                    eP[{z, ion, sh}] = {2, 1};
                } else {
                    const int lo = locate(nu, thres);
                    int hi;
                    if (sh <= 3) hi = kshellP;
                    else if (sh <= 6) hi = eP[{z, ion, 1}].first - 1;
                    else
                        hi = (eP[{z, ion, 6}].first < 3 ? eP[{z, ion, 5}].first
                                                        : eP[{z, ion, 6}].first) - 1;
                    eP[{z, ion, sh}] = {lo, hi};
                }
            }
            // the outer shell's lower pointer is re-located with the last threshold
            eP[{z, ion, out}].first = locate(nu, thres);
        }
    }
    for (int z = 3; z <= 30; ++z) {
        if (!on.count(z)) continue;
        for (int ion = 1; ion <= std::min(z, nstages); ++ion) {
            const int nel = z - ion + 1;
            for (int sh = 1; sh <= outer_shell(z, nel); ++sh) {
                int lo = eP[{z, ion, sh}].first, hi = eP[{z, ion, sh}].second;
                if (lo > hi) continue;
                lohi(lo, hi);
                shell_of(z, ion, sh, lo, hi, z, nel, true);
            }
        }
    }
    for (int z : on)
        for (int ion = 1; ion <= std::min(z, nstages); ++ion) x.species.push_back({z, ion});
    std::map<std::pair<int, int>, int> sidx;
    for (size_t i = 0; i < x.species.size(); ++i) sidx[x.species[i]] = (int)i;
    x.sigma.assign(x.species.size() * B, 0.0);
    for (const auto& s : x.shells) {
        const int i = sidx.at({s.elem, s.ion});
        for (int b = s.lo; b <= s.hi; ++b) x.sigma[(size_t)i * B + b - 1] += s.sigma[b - s.lo];
    }
    x.brems.resize(B);
    for (int b = 0; b < B; ++b) x.brems[b] = 1.03680e-18 / std::pow(nu[b], 3.0);
    return x;
}

// ------------------------------------------------------------------ balance
// natural cubic spline second derivatives (spline, interpolation_mod.f90:207-283)
void natural_spline_y2(const double* x, const double* y, int n, double* y2) {
    std::vector<double> u(n, 0.0);
    for (int i = 0; i < n; ++i) y2[i] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        const double ui = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * ui / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

// xSecRecoil (ph_mod.f90:369-385), as the reference writes it
std::vector<double> recoil_xsec(const std::vector<double>& nu) {
    const double thom = 6.65e-25;
    const int xrayP = locate(nu, 20.6);
    std::vector<double> out(nu.size(), thom);
    for (size_t i = xrayP; i < nu.size(); ++i) {
        const double dx = nu[i] / 3.7573e4;
        out[i] = thom * 3. / 4. *
                 ((1. + dx) / std::pow(dx, 3.0) *
                      (2. * dx * (1. + dx) / (1. + 2. * dx) - std::log(1. + 2. * dx)) +
                  1. / 2. / dx * std::log(1. + 2. * dx) - (1. + 3. * dx) / std::pow(1. + 2. * dx, 3.0));
    }
    return out;
}

}  // namespace

// ------------------------------------------------------------------ the model
struct moc_model {
    int32_t precision = 0;
    AtomicData d;
    // mesh and source
    std::vector<double> nu, wid, KNt, KNcdf, Pcomp, srcCDF, Lstar, nPhot, starPos;
    std::vector<int32_t> starIdx;
    int32_t lgCompton = 1;
    // grid
    int32_t nx = 0, nCells = 0;
    std::vector<double> axis;
    std::vector<int32_t> active;
    // atomic
    std::vector<int32_t> elemZ, spE, spI;
    std::vector<double> abund;
    int32_t nstages = 0;
    XSec xs;
    // balance
    int32_t maxLev = 27;
    double XHILimit = 0.05;
    std::vector<int32_t> shellNum, speciesShell, outShell, statW0, statW1, chexZ, chexIon;
    std::vector<double> W, chexFit, qom, q2;
    std::vector<int64_t> meta8;
    // emission
    std::vector<int32_t> edgeP[3];
    std::vector<double> logGamma[3];
    int32_t HIPnuP = 0, HeIPnuP = 0, HeIIPnuP = 0;
    const Shell *sHI = nullptr, *sHeI = nullptr, *sHeII = nullptr;
    std::vector<int32_t> fbZ, fbIon, fbLo, fbHi, he2Row;
    std::vector<int64_t> fbOff;
    std::vector<double> fbXs, fbGratio;
    int32_t j2TsP = 0, lymP[4] = {0, 0, 0, 0};
    // state
    std::vector<double> Te, Ne, Hden, ionDen;
};

namespace {

void build_source(moc_model& m, const moc_model_spec& s) {
    const auto& nu = m.nu;
    const size_t B = nu.size();
    m.wid = widths_of(nu);
    std::vector<double> f = blackbody_flux(nu, s.Tstellar);
    if (s.plFraction > 0.0) {  // a tabulated F_nu: power law plus blackbody, each normalised
        const double rydKeV = 13.6057e-3;
        std::vector<double> pl(B), plw(B), bbw(B);
        for (size_t i = 0; i < B; ++i) {
            pl[i] = (nu[i] >= s.plLoKeV / rydKeV && nu[i] <= s.plHiKeV / rydKeV) ? 1.0 / nu[i] : 0.0;
            plw[i] = pl[i] * m.wid[i];
            bbw[i] = f[i] * m.wid[i];
        }
        const double sp = pairwise_sum(plw.data(), B), sb = pairwise_sum(bbw.data(), B);
        if (!(sp > 0.0)) throw Error("the power law has no bin inside the mesh");
        for (size_t i = 0; i < B; ++i)
            f[i] = s.plFraction * pl[i] / sp + (1.0 - s.plFraction) * f[i] / sb;
    }
    // inSpectrumProbDen (setProbDen, continuum_mod.f90:415-504)
    m.srcCDF.assign(B, 0.0);
    double acc = 0.0;
    for (size_t i = 0; i < B; ++i) m.srcCDF[i] = acc += f[i] * m.wid[i];
    const double last = m.srcCDF[B - 1];
    if (!(last > 0.0)) throw Error("the source spectrum is zero on the mesh");
    double mx = 0.0;
    for (size_t i = 0; i < B; ++i) mx = std::max(mx, m.srcCDF[i] /= last);
    for (size_t i = 0; i < B; ++i)
        if (m.srcCDF[i] >= mx) m.srcCDF[i] = 1.0;
    // setCompton / Pcomp / KNsigma (ph_mod.f90:1883-1963), 1-degree steps
    m.KNt.assign(B, 0.0);
    m.KNcdf.assign(B * kNAngles, 0.0);
    m.Pcomp.assign(B * kNAngles, 0.0);
    std::vector<double> contrib(kNAngles), th(kNAngles), wgt(kNAngles);
    for (int a = 0; a < kNAngles; ++a) {
        th[a] = (a + 1) * kPiRef / 180.0;
        wgt[a] = 2.0 * kPiRef * std::sin((a + 1) * kPiRef / 180.0) * kPiRef / 180.0;
    }
    for (size_t b = 0; b < B; ++b) {
        for (int a = 0; a < kNAngles; ++a) {
            const double P = 1.0 / (1.0 + (nu[b] / kMeC2Ryd) * (1.0 - std::cos(th[a])));
            const double sn = std::sin(th[a]);
            const double ks = kHalfRe2 * (P - P * P * (sn * sn) + std::pow(P, 3.0));
            m.Pcomp[b * kNAngles + a] = P;
            contrib[a] = ks * wgt[a];
        }
        const double st = pairwise_sum(contrib.data(), kNAngles);
        m.KNt[b] = st;
        double c = 0.0;
        for (int a = 0; a < kNAngles; ++a) m.KNcdf[b * kNAngles + a] = (c += contrib[a]) / st;
        m.KNcdf[b * kNAngles + kNAngles - 1] = 1.0;
    }
    m.Lstar = {s.LStar};
    m.nPhot = {s.nPhotons};
    m.starPos = {0.0, 0.0, 0.0};
    m.starIdx = {1, 1, 1};  // the octant's star sits in its corner cell
    m.lgCompton = s.lgCompton ? 1 : 0;
}

// fillGrid / setMotherGrid (grid_mod.f90:698-711, 1242-1295): the octant axis
// (i-1)/(n-1) R and the cells with n_H > 0 numbered with x slowest, z fastest
void build_grid(moc_model& m, const moc_model_spec& s) {
    const int n = s.nx;
    m.nx = n;
    m.axis.resize(n);
    for (int i = 0; i < n; ++i) m.axis[i] = (double)i / (n - 1) * s.Rout;
    m.active.assign((size_t)n * n * n, 0);
    std::vector<double> hcell(1, 0.0);
    int id = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const size_t f = i + (size_t)n * (j + (size_t)n * k);
                const double h = s.hden ? s.hden[f] : s.Hdens;
                if (!(h > 0.0)) continue;
                m.active[f] = ++id;
                hcell.push_back(h);
            }
    if (id == 0) throw Error("no cell has a positive hydrogen density");
    m.nCells = id;
    // setGrid's starting guess (grid_mod.f90:1583-1645)
    const size_t C1 = (size_t)id + 1, NE = m.elemZ.size(), NS = m.nstages;
    m.Hden = hcell;
    m.Ne = hcell;
    m.Te.assign(C1, s.TeStart);
    m.Te[0] = 0.0;
    m.ionDen.assign(C1 * NE * NS, 0.0);  // (0:nCells, nElem, nstages), cell fastest
    auto X = [&](size_t c, size_t e, size_t i) -> double& { return m.ionDen[c + C1 * (e + NE * i)]; };
    const double H0 = s.H0in;
    for (size_t e = 0; e < NE; ++e) {
        const int z = m.elemZ[e];
        for (size_t c = 1; c < C1; ++c) {
            if (z == 1) {
                X(c, e, 0) = H0;
                X(c, e, 1) = 1.0 - H0;
            } else if (z == 2) {
                X(c, e, 0) = H0;
                X(c, e, 1) = X(c, e, 2) = (1.0 - H0) / 2.0;
            } else {
                const int top = std::min(z + 1, (int)NS);
                X(c, e, 0) = H0;
                for (int i = 1; i < top; ++i) X(c, e, i) = (1.0 - H0) / std::min(z, (int)NS - 1);
            }
        }
    }
}

// the tables of the balance (see mocgpu.h moc_balance_tables)
void build_balance(moc_model& m) {
    const auto& nu = m.nu;
    const int B = (int)nu.size();
    const XSec& xs = m.xs;
    std::map<std::pair<int, int>, int> spec;
    for (size_t i = 0; i < xs.species.size(); ++i) spec[xs.species[i]] = (int)i;
    // shells grouped by species, each group in stack order
    std::vector<int> order(xs.shells.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return spec.at({xs.shells[a].elem, xs.shells[a].ion}) <
               spec.at({xs.shells[b].elem, xs.shells[b].ion});
    });
    const int ns = (int)order.size();
    m.speciesShell.assign(xs.species.size() + 1, 0);
    m.shellNum.clear();
    for (int k : order) {
        ++m.speciesShell[spec.at({xs.shells[k].elem, xs.shells[k].ion}) + 1];
        m.shellNum.push_back(xs.shells[k].shell);
    }
    for (size_t i = 1; i < m.speciesShell.size(); ++i) m.speciesShell[i] += m.speciesShell[i - 1];
    const int secIonP = locate(nu, 7.353);
    const int cRecoilHP = locate(nu, 194.);
    const int cRecoilHeP = locate(nu, 260.);
    const auto rec = recoil_xsec(nu);
    const int NC = 3 * ns + 5;
    m.W.assign((size_t)B * NC, 0.0);
    auto W = [&](int b1, int c) -> double& { return m.W[(size_t)(b1 - 1) * NC + c]; };
    for (int k = 0; k < ns; ++k) {
        const Shell& s = xs.shells[order[k]];
        const int lo = s.lo, hi = std::min(s.hi, B);
        const auto& sig = s.sigma;
        if (s.elem <= 2) {  // H and He (photoionisation, update_mod.f90:3639-3672)
            for (int b = lo; b <= B; ++b) {
                double ph = sig[b - lo];
                if (ph < 1.e-35) ph = 0.0;
                W(b, 3 * k) = ph / (kHcRyd * nu[b - 1]);
                W(b, 3 * k + 1) = ph * (nu[b - 1] - nu[lo - 1]) / nu[b - 1];
            }
        } else {  // getPhotoRates (update_mod.f90:3971-4040)
            const double y1 = m.d.auger[((s.elem * 32 + s.ion) * 11 + s.shell) * 11 + 1];
            const double augerE = kHcRyd * nu[lo - 1] * y1;
            for (int b = lo + 1; b <= std::min(hi, secIonP - 1); ++b) {
                double ph = sig[b - lo];
                if (ph < 1.e-35) ph = 0.0;
                W(b, 3 * k) += ph / (kHcRyd * nu[b - 1]);
                W(b, 3 * k + 1) += ph - augerE * (ph / (kHcRyd * nu[b - 1]));
            }
            const double ph0 = sig[0];
            W(lo, 3 * k) += ph0 / (kHcRyd * nu[lo - 1]);
            W(lo, 3 * k + 1) += ph0 * (kHcRyd * nu[lo - 1] - augerE) / (kHcRyd * nu[lo - 1]);
            double gH = 0.0, hH = 0.0;
            for (int b = std::max(lo + 1, secIonP); b <= hi; ++b) {
                gH += sig[b - lo] / (kHcRyd * nu[b - 1]);
                hH += sig[b - lo];
            }
            W(lo, 3 * k) += gH;
            W(lo, 3 * k + 2) = hH - gH * augerE;
        }
    }
    const int c0 = 3 * ns;
    for (int b = 1; b <= B; ++b) {
        const double v = nu[b - 1];
        W(b, c0) = 1.0;
        W(b, c0 + 1) = 1.0 / v;
        W(b, c0 + 2) = b > cRecoilHP ? rec[b - 1] * (2.66e-5 * v * v - 1.) : 0.0;
        W(b, c0 + 3) = b > cRecoilHeP ? rec[b - 1] * (2.66e-5 * v * v - 1.8) : 0.0;
        W(b, c0 + 4) = 1.0;
    }
    m.outShell.assign(31 * 32, 0);
    m.statW0.assign(31 * 32, 1);
    m.statW1.assign(31 * 32, 1);
    for (int z = 1; z <= 30; ++z)
        for (int nel = 1; nel <= z; ++nel) {
            m.outShell[z * 32 + nel] = outer_shell(z, nel);
            int g0, g1;
            stat_weights(z, nel, &g0, &g1);
            m.statW0[z * 32 + nel] = g0;
            m.statW1[z * 32 + nel] = g1;
        }
    m.chexZ.clear();
    m.chexIon.clear();
    m.chexFit.clear();
    for (const auto& r : kChex) {
        m.chexZ.push_back(r.z);
        m.chexIon.push_back(r.ion);
        m.chexFit.insert(m.chexFit.end(), {r.a, r.b, r.c, r.d});
    }
    const size_t nA = m.d.atom_meta.size() / 7;
    m.meta8.assign(nA * 8, 0);
    for (size_t a = 0; a < nA; ++a)
        for (int c = 0; c < 7; ++c) m.meta8[a * 8 + c] = m.d.atom_meta[a * 7 + c];
    m.qom = m.d.atom_qom;
    m.q2.assign(m.qom.size(), 0.0);
    for (size_t a = 0; a < nA; ++a) {
        const int64_t nl = m.d.atom_meta[a * 7], nt = m.d.atom_meta[a * 7 + 1];
        const int64_t ot = m.d.atom_meta[a * 7 + 2], oq = m.d.atom_meta[a * 7 + 5];
        if (nt < 3) continue;
        std::vector<double> y(nt), y2(nt);
        for (int64_t pr = 0; pr < nl * nl; ++pr) {
            for (int64_t t = 0; t < nt; ++t) y[t] = m.qom[oq + t * nl * nl + pr];
            natural_spline_y2(&m.d.atom_logt[ot], y.data(), (int)nt, y2.data());
            for (int64_t t = 0; t < nt; ++t) m.q2[oq + t * nl * nl + pr] = y2[t];
        }
    }
}

// initGammaCont (ph_mod.f90:96-196): the bin of each table frequency, moved
// up when it lies above the bin midpoint; the first pointer is 1
std::vector<int32_t> edge_pointers(const std::vector<double>& nu, const std::vector<double>& gnu) {
    const int B = (int)nu.size();
    std::vector<int32_t> out(gnu.size());
    for (size_t i = 0; i < gnu.size(); ++i) {
        const double g = (double)(float)gnu[i];
        int e = locate(nu, g);
        if (e >= 1 && e < B && g > (nu[e - 1] + nu[e]) / 2.0) ++e;
        out[i] = e;
    }
    out[0] = 1;
    return out;
}

// the tables of the diffuse emission (see mocgpu.h moc_emission_tables)
void build_emission(moc_model& m) {
    const auto& nu = m.nu;
    const int B = (int)nu.size();
    const AtomicData& d = m.d;
    if (m.xs.clamped)
        throw Error("an ionisation threshold lies below the frequency mesh: the diffuse emission "
                    "needs a mesh that starts below every threshold");
    static const char* kTag[3] = {"H I", "He I", "He II"};
    for (int g = 0; g < 3; ++g) {
        if (d.gam_tk[g] != d.gam_tk[0]) throw Error("gamma tables must share their temperature grid");
        m.edgeP[g] = edge_pointers(nu, d.gam_nu[g]);
        for (size_t i = 0; i + 1 < m.edgeP[g].size(); i += 2)
            if (m.edgeP[g][i] == m.edgeP[g][i + 1])
                throw Error(std::string("frequency grid too coarse for the ") + kTag[g] +
                            " free-bound edges");
        m.logGamma[g].resize(d.gam_val[g].size());
        for (size_t i = 0; i < d.gam_val[g].size(); ++i)
            m.logGamma[g][i] = d.gam_val[g][i] > 0 ? std::log10(d.gam_val[g][i]) : 0.0;
    }
    std::map<std::tuple<int, int, int>, const Shell*> sh;
    for (const auto& s : m.xs.shells) sh[{s.elem, s.ion, s.shell}] = &s;
    auto find = [&](int z, int ion, int s) -> const Shell* {
        auto it = sh.find({z, ion, s});
        return it == sh.end() ? nullptr : it->second;
    };
    m.sHI = find(1, 1, 1);
    if (!m.sHI) throw Error("hydrogen must be switched on");
    m.sHeI = find(2, 1, 1);
    m.sHeII = find(2, 2, 1);
    m.HIPnuP = m.sHI->lo;
    m.HeIPnuP = m.sHeI ? m.sHeI->lo : 0;
    m.HeIIPnuP = m.sHeII ? m.sHeII->lo : 0;
    std::set<int> on;
    for (const auto& s : m.xs.shells) on.insert(s.elem);
    for (const auto& p : m.xs.species) on.insert(p.first);
    int64_t off = 0;
    m.fbXs.clear();
    for (int z = 3; z < 20; ++z) {
        if (!on.count(z)) continue;
        for (int ion = 1; ion <= std::min(z, m.nstages - 1); ++ion) {
            const int nel = z - ion + 1;
            const Shell* s = find(z, ion, outer_shell(z, nel));
            if (!s) continue;  // empty bin range: never emits
            int g0, g1;
            stat_weights(z, nel, &g0, &g1);
            m.fbZ.push_back(z);
            m.fbIon.push_back(ion);
            m.fbLo.push_back(s->lo);
            m.fbHi.push_back(s->hi);
            m.fbOff.push_back(off);
            m.fbGratio.push_back((double)((float)g0 / (float)g1));
            m.fbXs.insert(m.fbXs.end(), s->sigma.begin(), s->sigma.end());
            off += (int64_t)s->sigma.size();
        }
    }
    if (m.fbXs.empty()) m.fbXs.push_back(0.0);
    // HeI2photSub (emission_mod.f90:800-815): the table row reached at each bin
    m.he2Row.assign(B, 0);
    int j = 1;
    for (int i = 0; i < B; ++i) {
        const double y = nu[i] / 1.514;
        if (y > d.hei2phot[(j - 1) * 2]) ++j;
        if (j >= 41) break;
        m.he2Row[i] = j;
    }
    for (int i = 0; i < 4; ++i) m.lymP[i] = locate(nu, (double)(float)d.heiiLyman[i * 2 + 1]);
    m.j2TsP = locate(nu, (double)(float)1.45673);
}

// ------------------------------------------------------------------ records
struct Writer {
    FILE* f;
    void rec(const char* name, int kind, int64_t n, const void* v, size_t esz) {
        char nm[32];
        std::memset(nm, ' ', 32);
        std::memcpy(nm, name, std::min<size_t>(32, std::strlen(name)));
        const int32_t hdr[2] = {kind, 0};
        fwrite(nm, 1, 32, f);
        fwrite(hdr, 4, 2, f);
        fwrite(&n, 8, 1, f);
        if (n) fwrite(v, esz, (size_t)n, f);
    }
    void d(const char* n, const std::vector<double>& v) { rec(n, 1, (int64_t)v.size(), v.data(), 8); }
    void d(const char* n, const double* v, int64_t c) { rec(n, 1, c, v, 8); }
    void i(const char* n, const std::vector<int32_t>& v) { rec(n, 2, (int64_t)v.size(), v.data(), 4); }
    void i(const char* n, const int32_t* v, int64_t c) { rec(n, 2, c, v, 4); }
    void l(const char* n, const std::vector<int64_t>& v) { rec(n, 3, (int64_t)v.size(), v.data(), 8); }
    void d1(const char* n, double v) { rec(n, 1, 1, &v, 8); }
    void i1(const char* n, int32_t v) { rec(n, 2, 1, &v, 4); }
    void l1(const char* n, int64_t v) { rec(n, 3, 1, &v, 8); }
};

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" const char* moc_model_last_error(void) { return g_error.c_str(); }

extern "C" void moc_model_free(moc_model* model) { delete model; }

extern "C" int moc_model_build(const moc_model_spec* spec, moc_model** model) {
    if (!spec || !model) return fail(MOC_ERR_ARG, "moc_model_build: NULL argument");
    *model = nullptr;
    const moc_model_spec& s = *spec;
    if (!s.dataDir || !s.elemZ || !s.abund || s.nElem < 1)
        return fail(MOC_ERR_ARG, "moc_model_build: dataDir, elemZ and abund are required");
    if (s.nstages < 2 || s.nstages > 31)
        return fail(MOC_ERR_ARG, "moc_model_build: nstages must lie in 2..31");
    if (s.nx < 2 || s.Rout <= 0.0) return fail(MOC_ERR_ARG, "moc_model_build: nx >= 2 and Rout > 0");
    if (!(s.nuMin > 0.0) || !(s.nuMax > s.nuMin) || s.nbins < 3)
        return fail(MOC_ERR_ARG, "moc_model_build: 0 < nuMin < nuMax and nbins >= 3");
    if (!(s.Tstellar > 0.0) || !(s.LStar > 0.0) || !(s.nPhotons >= 1.0))
        return fail(MOC_ERR_ARG, "moc_model_build: Tstellar, LStar and nPhotons must be positive");
    if (s.plFraction < 0.0 || s.plFraction > 1.0)
        return fail(MOC_ERR_ARG, "moc_model_build: plFraction must lie in [0, 1]");
    if (s.maxLev < 1 || s.maxLev > 27) return fail(MOC_ERR_ARG, "moc_model_build: maxLev in 1..27");
    if (s.precision != MOC_PREC_FP64 && s.precision != MOC_PREC_MIXED)
        return fail(MOC_ERR_ARG, "moc_model_build: precision must be MOC_PREC_FP64 or MOC_PREC_MIXED");
    std::set<int> on;
    for (int i = 0; i < s.nElem; ++i) {
        if (s.elemZ[i] < 1 || s.elemZ[i] > 30)
            return fail(MOC_ERR_ARG, "moc_model_build: atomic numbers lie in 1..30");
        if (i > 0 && s.elemZ[i] <= s.elemZ[i - 1])
            return fail(MOC_ERR_ARG, "moc_model_build: elemZ must be ascending");
        on.insert(s.elemZ[i]);
    }
    if (!on.count(1)) return fail(MOC_ERR_ARG, "moc_model_build: hydrogen must be switched on");
    try {
        auto m = std::make_unique<moc_model>();
        m->precision = s.precision;
        read_atomic_data(s.dataDir, m->d);
        m->elemZ.assign(s.elemZ, s.elemZ + s.nElem);
        m->abund.assign(s.abund, s.abund + s.nElem);
        m->nstages = s.nstages;
        m->maxLev = s.maxLev;
        m->XHILimit = s.XHILimit;
        m->nu = gas_mesh(s.nbins, s.nuMin, s.nuMax, ion_edges(m->d, on, s.nstages, s.nuMax),
                         s.strictMesh != 0);
        build_source(*m, s);
        m->xs = build_xsec(m->d, m->nu, on, s.nstages);
        for (const auto& p : m->xs.species) {
            m->spE.push_back((int32_t)(std::find(m->elemZ.begin(), m->elemZ.end(), p.first) -
                                       m->elemZ.begin()));
            m->spI.push_back(p.second);
        }
        build_balance(*m);
        build_emission(*m);
        build_grid(*m, s);
        *model = m.release();
        return MOC_OK;
    } catch (const EndOfFile&) {
        return fail(MOC_ERR_ARG, "moc_model_build: a data file ends early");
    } catch (const std::exception& e) {
        return fail(MOC_ERR_ARG, std::string("moc_model_build: ") + e.what());
    }
}

extern "C" int moc_model_get(const moc_model* mp, moc_tables* t, moc_grid_desc* g, moc_atomic* a,
                             moc_balance_tables* b, moc_emission_tables* e) {
    if (!mp) return fail(MOC_ERR_ARG, "moc_model_get: NULL model");
    const moc_model& m = *mp;
    if (t) {
        std::memset(t, 0, sizeof(*t));
        t->nbins = (int32_t)m.nu.size();
        t->precision = m.precision;
        t->nuArray = m.nu.data();
        t->widFlx = m.wid.data();
        t->ionEdge1 = kHEdge;
        t->lgCompton = m.lgCompton;
        t->nAngles = kNAngles;
        t->KNsigmaT = m.KNt.data();
        t->KNsigmaCDF = m.KNcdf.data();
        t->Pcomp = m.Pcomp.data();
        t->nStars = 1;
        t->sourceCDF = m.srcCDF.data();
        t->Lstar = m.Lstar.data();
        t->nPhotons = m.nPhot.data();
        t->starPosition = m.starPos.data();
        t->starIndex = m.starIdx.data();
    }
    if (g) {
        std::memset(g, 0, sizeof(*g));
        g->nx = g->ny = g->nz = m.nx;
        g->xAxis = g->yAxis = g->zAxis = m.axis.data();
        g->active = m.active.data();
        g->nCells = m.nCells;
        g->lgSymmetricXYZ = 1;
        g->R_out = 0.0;
        g->lgPlaneIonization = 0;
    }
    if (a) {
        std::memset(a, 0, sizeof(*a));
        a->nElemUsed = (int32_t)m.elemZ.size();
        a->elemZ = m.elemZ.data();
        a->abund = m.abund.data();
        a->nstages = m.nstages;
        a->nSpecies = (int32_t)m.spE.size();
        a->speciesElem = m.spE.data();
        a->speciesIon = m.spI.data();
        a->sigma = m.xs.sigma.data();
        a->brems = m.xs.brems.data();
    }
    if (b) {
        const AtomicData& d = m.d;
        std::memset(b, 0, sizeof(*b));
        b->maxLev = m.maxLev;
        b->XHILimit = m.XHILimit;
        b->nShells = (int32_t)m.shellNum.size();
        b->shellNum = m.shellNum.data();
        b->speciesShell = m.speciesShell.data();
        b->W = m.W.data();
        b->outShell = m.outShell.data();
        b->statW0 = m.statW0.data();
        b->statW1 = m.statW1.data();
        b->rrb = d.rrb.data();
        b->drb_c = d.drb_c.data();
        b->drb_e = d.drb_e.data();
        b->rrec = d.rrec.data();
        b->rnew = d.rnew.data();
        b->nDiel = d.nDiel;
        b->diel = d.diel.data();
        b->nAldro = d.nAldro;
        b->aldro = d.aldro.data();
        b->cionCF = d.cion_cf.data();
        b->cionTherm = d.cion_th.data();
        b->nAuger = d.nauger.data();
        b->auger = d.auger.data();
        b->nChex = (int32_t)m.chexZ.size();
        b->chexZ = m.chexZ.data();
        b->chexIon = m.chexIon.data();
        b->chexFit = m.chexFit.data();
        b->nAtoms = (int32_t)d.atom_atp.size();
        b->atomIdx = d.atom_idx.data();
        b->atomMeta = m.meta8.data();
        b->atomAtp = d.atom_atp.data();
        b->nLogT = (int64_t)d.atom_logt.size();
        b->nLev = (int64_t)d.atom_g.size();
        b->nA = (int64_t)d.atom_A.size();
        b->nQ = (int64_t)m.qom.size();
        b->atomLogT = d.atom_logt.data();
        b->atomG = d.atom_g.data();
        b->atomE = d.atom_e.data();
        b->atomA = d.atom_A.data();
        b->atomQom = m.qom.data();
        b->atomQ2 = m.q2.data();
        b->atomRecKind = d.atom_rec_kind.data();
        b->atomRec = d.atom_rec.data();
        b->atomRecValid = d.atom_rec_valid.data();
    }
    if (e) {
        const AtomicData& d = m.d;
        std::memset(e, 0, sizeof(*e));
        e->ntk = (int32_t)d.gam_tk[0].size();
        e->tkGamma = d.gam_tk[0].data();
        for (int k = 0; k < 3; ++k) {
            e->nlim[k] = (int32_t)m.edgeP[k].size();
            e->edgeP[k] = m.edgeP[k].data();
            e->logGamma[k] = m.logGamma[k].data();
        }
        e->HIPnuP = m.HIPnuP;
        e->HeIPnuP = m.HeIPnuP;
        e->HeIIPnuP = m.HeIIPnuP;
        e->xsHI = m.sHI->sigma.data();
        e->xsHeI = m.sHeI ? m.sHeI->sigma.data() : nullptr;
        e->xsHeII = m.sHeII ? m.sHeII->sigma.data() : nullptr;
        e->nFb = (int32_t)m.fbZ.size();
        e->fbZ = m.fbZ.data();
        e->fbIon = m.fbIon.data();
        e->fbLo = m.fbLo.data();
        e->fbHi = m.fbHi.data();
        e->fbOff = m.fbOff.data();
        e->nFbXs = (int64_t)m.fbXs.size();
        e->fbXs = m.fbXs.data();
        e->fbGratio = m.fbGratio.data();
        e->he2Row = m.he2Row.data();
        e->hei2phot = d.hei2phot.data();
        e->nHlFiles = d.nHlFiles;
        e->hlMap = d.hl_map.data();
        e->hlDens = d.hl_dens.data();
        e->hlLines = d.hl_lines.data();
        e->hbA = d.hbA.data();
        e->hbB = d.hbB.data();
        e->rbEdge = d.rbEdge.data();
        e->r2aEdge = d.r2aEdge.data();
        e->heiRec = d.heiRec.data();
        e->heiiLyman = d.heiiLyman.data();
        e->j2TsP = m.j2TsP;
        for (int k = 0; k < 4; ++k) e->lymP[k] = m.lymP[k];
    }
    return MOC_OK;
}

extern "C" int moc_model_state(const moc_model* m, int64_t* nCells, const double** Te,
                               const double** Ne, const double** Hden, const double** ionDen) {
    if (!m) return fail(MOC_ERR_ARG, "moc_model_state: NULL model");
    if (nCells) *nCells = m->nCells;
    if (Te) *Te = m->Te.data();
    if (Ne) *Ne = m->Ne.data();
    if (Hden) *Hden = m->Hden.data();
    if (ionDen) *ionDen = m->ionDen.data();
    return MOC_OK;
}

extern "C" int moc_model_load(moc_ctx** ctxp, const moc_model* m) {
    if (!ctxp || !m) return fail(MOC_ERR_ARG, "moc_model_load: NULL argument");
    *ctxp = nullptr;
    moc_tables t;
    moc_grid_desc g;
    moc_atomic a;
    moc_balance_tables b;
    moc_emission_tables e;
    moc_model_get(m, &t, &g, &a, &b, &e);
    moc_ctx* ctx = nullptr;
    int rc = moc_create(&ctx, &t);
    if (rc != MOC_OK) {  // moc_create hands back the failed context with its message
        const std::string msg = std::string("moc_model_load: ") + moc_last_error(ctx);
        if (ctx) moc_destroy(ctx);
        return fail(rc, msg);
    }
    auto step = [&](int r, const char* what) {
        if (r != MOC_OK && rc == MOC_OK) {
            rc = r;
            g_error = std::string("moc_model_load: ") + what + ": " + moc_last_error(ctx);
        }
    };
    const int64_t C1 = (int64_t)m->nCells + 1;
    step(moc_grid_attach(ctx, 1, &g), "moc_grid_attach");
    if (rc == MOC_OK) step(moc_set_atomic(ctx, &a), "moc_set_atomic");
    if (rc == MOC_OK) step(moc_set_balance(ctx, &b), "moc_set_balance");
    if (rc == MOC_OK) step(moc_set_emission(ctx, &e), "moc_set_emission");
    if (rc == MOC_OK)
        step(moc_upload(ctx, MOC_F_IONDEN, 1, m->ionDen.data(), (int64_t)m->ionDen.size(),
                        MOC_DT_F64, MOC_LAYOUT_FORTRAN), "ionDen");
    if (rc == MOC_OK)
        step(moc_upload(ctx, MOC_F_TE, 1, m->Te.data(), C1, MOC_DT_F64, MOC_LAYOUT_FORTRAN), "Te");
    if (rc == MOC_OK)
        step(moc_upload(ctx, MOC_F_NE, 1, m->Ne.data(), C1, MOC_DT_F64, MOC_LAYOUT_FORTRAN), "Ne");
    if (rc == MOC_OK)
        step(moc_upload(ctx, MOC_F_HDEN, 1, m->Hden.data(), C1, MOC_DT_F64, MOC_LAYOUT_FORTRAN),
             "Hden");
    if (rc != MOC_OK) {
        moc_destroy(ctx);
        return rc;
    }
    *ctxp = ctx;
    return MOC_OK;
}

extern "C" int moc_model_write(const moc_model* mp, const char* path) {
    if (!mp || !path) return fail(MOC_ERR_ARG, "moc_model_write: NULL argument");
    const moc_model& m = *mp;
    const AtomicData& d = m.d;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(MOC_ERR_ARG, std::string("moc_model_write: can't open ") + path);
    Writer w{f};
    fwrite("MOCDUMP1", 1, 8, f);
    const int64_t B = (int64_t)m.nu.size();
    // moc_tables
    w.i1("tables.nbins", (int32_t)B);
    w.i1("tables.precision", m.precision);
    w.d("tables.nuArray", m.nu);
    w.d("tables.widFlx", m.wid);
    w.d1("tables.ionEdge1", kHEdge);
    w.i1("tables.lgCompton", m.lgCompton);
    w.i1("tables.nAngles", kNAngles);
    w.d("tables.KNsigmaT", m.KNt);
    w.d("tables.KNsigmaCDF", m.KNcdf);
    w.d("tables.Pcomp", m.Pcomp);
    w.i1("tables.nStars", 1);
    w.d("tables.sourceCDF", m.srcCDF);
    w.d("tables.Lstar", m.Lstar);
    w.d("tables.nPhotons", m.nPhot);
    w.d("tables.starPosition", m.starPos);
    w.i("tables.starIndex", m.starIdx);
    // moc_grid_desc
    w.i1("grid.nx", m.nx);
    w.i1("grid.ny", m.nx);
    w.i1("grid.nz", m.nx);
    w.d("grid.xAxis", m.axis);
    w.d("grid.yAxis", m.axis);
    w.d("grid.zAxis", m.axis);
    w.i("grid.active", m.active);
    w.i1("grid.nCells", m.nCells);
    w.i1("grid.lgSymmetricXYZ", 1);
    w.d1("grid.R_out", 0.0);
    w.i1("grid.lgPlaneIonization", 0);
    // moc_atomic
    w.i1("atomic.nElemUsed", (int32_t)m.elemZ.size());
    w.i("atomic.elemZ", m.elemZ);
    w.d("atomic.abund", m.abund);
    w.i1("atomic.nstages", m.nstages);
    w.i1("atomic.nSpecies", (int32_t)m.spE.size());
    w.i("atomic.speciesElem", m.spE);
    w.i("atomic.speciesIon", m.spI);
    w.d("atomic.sigma", m.xs.sigma);
    w.d("atomic.brems", m.xs.brems);
    // moc_balance_tables
    w.i1("balance.maxLev", m.maxLev);
    w.d1("balance.XHILimit", m.XHILimit);
    w.i1("balance.nShells", (int32_t)m.shellNum.size());
    w.i("balance.shellNum", m.shellNum);
    w.i("balance.speciesShell", m.speciesShell);
    w.d("balance.W", m.W);
    w.i("balance.outShell", m.outShell);
    w.i("balance.statW0", m.statW0);
    w.i("balance.statW1", m.statW1);
    w.d("balance.rrb", d.rrb);
    w.d("balance.drb_c", d.drb_c);
    w.d("balance.drb_e", d.drb_e);
    w.d("balance.rrec", d.rrec);
    w.d("balance.rnew", d.rnew);
    w.i1("balance.nDiel", d.nDiel);
    w.d("balance.diel", d.diel);
    w.i1("balance.nAldro", d.nAldro);
    w.d("balance.aldro", d.aldro);
    w.d("balance.cionCF", d.cion_cf);
    w.d("balance.cionTherm", d.cion_th);
    w.i("balance.nAuger", d.nauger);
    w.d("balance.auger", d.auger);
    w.i1("balance.nChex", (int32_t)m.chexZ.size());
    w.i("balance.chexZ", m.chexZ);
    w.i("balance.chexIon", m.chexIon);
    w.d("balance.chexFit", m.chexFit);
    w.i1("balance.nAtoms", (int32_t)d.atom_atp.size());
    w.i("balance.atomIdx", d.atom_idx);
    w.l("balance.atomMeta", m.meta8);
    w.d("balance.atomAtp", d.atom_atp);
    w.l1("balance.nLogT", (int64_t)d.atom_logt.size());
    w.l1("balance.nLev", (int64_t)d.atom_g.size());
    w.l1("balance.nA", (int64_t)d.atom_A.size());
    w.l1("balance.nQ", (int64_t)m.qom.size());
    w.d("balance.atomLogT", d.atom_logt);
    w.d("balance.atomG", d.atom_g);
    w.d("balance.atomE", d.atom_e);
    w.d("balance.atomA", d.atom_A);
    w.d("balance.atomQom", m.qom);
    w.d("balance.atomQ2", m.q2);
    w.i("balance.atomRecKind", d.atom_rec_kind);
    w.d("balance.atomRec", d.atom_rec);
    w.i("balance.atomRecValid", d.atom_rec_valid);
    // moc_emission_tables
    w.i1("emission.ntk", (int32_t)d.gam_tk[0].size());
    w.d("emission.tkGamma", d.gam_tk[0]);
    const int32_t nlim[3] = {(int32_t)m.edgeP[0].size(), (int32_t)m.edgeP[1].size(),
                             (int32_t)m.edgeP[2].size()};
    w.i("emission.nlim", nlim, 3);
    for (int k = 0; k < 3; ++k) {
        const std::string n = std::to_string(k + 1);
        w.i(("emission.edgeP." + n).c_str(), m.edgeP[k]);
    }
    for (int k = 0; k < 3; ++k) {
        const std::string n = std::to_string(k + 1);
        w.d(("emission.logGamma." + n).c_str(), m.logGamma[k]);
    }
    w.i1("emission.HIPnuP", m.HIPnuP);
    w.i1("emission.HeIPnuP", m.HeIPnuP);
    w.i1("emission.HeIIPnuP", m.HeIIPnuP);
    w.d("emission.xsHI", m.sHI->sigma);
    if (m.sHeI) w.d("emission.xsHeI", m.sHeI->sigma);
    if (m.sHeII) w.d("emission.xsHeII", m.sHeII->sigma);
    w.i1("emission.nFb", (int32_t)m.fbZ.size());
    w.i("emission.fbZ", m.fbZ);
    w.i("emission.fbIon", m.fbIon);
    w.i("emission.fbLo", m.fbLo);
    w.i("emission.fbHi", m.fbHi);
    w.l("emission.fbOff", m.fbOff);
    w.l1("emission.nFbXs", (int64_t)m.fbXs.size());
    w.d("emission.fbXs", m.fbXs);
    w.d("emission.fbGratio", m.fbGratio);
    w.i("emission.he2Row", m.he2Row);
    w.d("emission.hei2phot", d.hei2phot);
    w.i1("emission.nHlFiles", d.nHlFiles);
    w.i("emission.hlMap", d.hl_map);
    w.d("emission.hlDens", d.hl_dens);
    w.d("emission.hlLines", d.hl_lines);
    w.d("emission.hbA", d.hbA);
    w.d("emission.hbB", d.hbB);
    w.d("emission.rbEdge", d.rbEdge);
    w.d("emission.r2aEdge", d.r2aEdge);
    w.d("emission.heiRec", d.heiRec);
    w.d("emission.heiiLyman", d.heiiLyman);
    w.i1("emission.j2TsP", m.j2TsP);
    w.i("emission.lymP", m.lymP, 4);
    // setGrid's starting state
    w.d("state.Te", m.Te);
    w.d("state.Ne", m.Ne);
    w.d("state.Hden", m.Hden);
    w.d("state.ionDen", m.ionDen);
    w.rec("END", 0, 0, nullptr, 1);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    if (!ok) return fail(MOC_ERR_ARG, std::string("moc_model_write: write error on ") + path);
    return MOC_OK;
}

// Per-cell opacity kappa_nu on gfx950 (ionizationDriver, ionization_mod.f90:26-477).
//
// kappa(cell, nu) = free-free(cell, nu) + sum_s n_s(cell) sigma_s(nu) + n_e sigma_KN(nu)
//
// The bound-free part is a contraction over the photoionisable species s =
// (element, stage) of the per-cell ion densities n_s = ionDen * abundance *
// n_H with the per-species cross-sections (the shells of each species summed
// on the host).  A workgroup owns a 64-cell x 64-bin tile; species are staged
// 16 at a time through LDS (densities cell-fastest, cross-sections
// bin-fastest) and each lane accumulates one cell x 16 bins.  Lanes of a wave
// share their bins, so cross-section reads are LDS broadcasts and the output
// (bin-major, cell fastest) is written as 64 consecutive cells per wave.
// The free-free term uses the Hummer (1988) Chebyshev fit of getGauntFF
// (ionization_mod.f90:196-336) at each cell's own T_e.
#include <hip/hip_runtime.h>

#include "moc_internal.h"

namespace {

constexpr double kTe1Ryd = 1.578866e5;  // constants_mod.f90:25

// Hummer (1988) free-free Gaunt-factor coefficients d(8,11), column-major
__constant__ double c_hummer[88] = {
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

__device__ __forceinline__ double hd(int j, int k) { return c_hummer[(k - 1) * 8 + (j - 1)]; }

// Chebyshev coefficients in the photon-energy variable for one temperature
struct GauntT {
    double c[9];
    double con, gamma2;
};

__device__ GauntT gaunt_setup(double z, double log10Te) {
    GauntT g;
    double xlrkt;
    if (log10Te < 2.5) xlrkt = 5.1983649 - 2.5;
    else if (log10Te > 8.) xlrkt = 5.1983649 - 8.;
    else xlrkt = 5.1983649 - log10Te;
    const double txg = 0.66666667 * (2.0 * z + xlrkt);
    g.gamma2 = pow(10., txg * 1.5);
    g.con = 0.72727273 * xlrkt + 0.90909091;
    for (int j = 1; j <= 8; ++j) {
        double b[12];
        b[11] = hd(j, 11);
        b[10] = txg * b[11] + hd(j, 10);
        int ir = 9;
        for (int i = 1; i <= 9; ++i) {
            b[ir] = txg * b[ir + 1] - b[ir + 2] + hd(j, ir);
            --ir;
        }
        g.c[j] = 0.25 * (b[1] - b[3]);
    }
    return g;
}

__device__ double gaunt_eval(const GauntT& g, double xlf) {
    const double txu = 0.72727273 * xlf + g.con;
    const double u = pow(10., (txu - .90909091) / .72727273);
    if (fabs(txu) <= 2.0) {
        double b[9];
        b[8] = g.c[8];
        b[7] = txu * b[8] + g.c[7];
        int ir = 6;
        for (int j = 1; j <= 6; ++j) {
            b[ir] = txu * b[ir + 1] - b[ir + 2] + g.c[ir];
            --ir;
        }
        return b[1] - b[3];
    }
    if (log10(u) < -4.0) {
        if (g.gamma2 < 0.3) return 0.551329 * log(2.24593 / u);
        return 0.551329 * log(0.944931 / u / sqrt(g.gamma2));
    }
    if (txu > 2.0) {
        if (log10(g.gamma2) < -3.) return sqrt(0.9549297 / u);
        if (log10(g.gamma2) < 0.0) {
            const double slope = (sqrt(12 * g.gamma2 / u) - sqrt(0.9549297 / u)) / 3.0;
            return sqrt(12 * g.gamma2 / u) + slope * log10(g.gamma2);
        }
        const double v = sqrt(12. * g.gamma2 / u);
        return v > 1.05 ? 1.05 : v;
    }
    return 0.0;
}

constexpr int TC = 64;   // cells per tile
constexpr int TBIN = 64; // bins per tile
constexpr int KS = 16;   // species per LDS stage
constexpr int BPL = TBIN / 4;  // bins per lane

__global__ void __launch_bounds__(256) k_opacity(MocOpacityArgs a) {
    __shared__ double s_n[KS][TC];
    __shared__ double s_sig[KS][TBIN];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int c0 = blockIdx.x * TC, b0 = blockIdx.y * TBIN;
    const int cell = c0 + tx;
    const int C1 = a.C1, B = a.nbins;
    double acc[BPL];
#pragma unroll
    for (int j = 0; j < BPL; ++j) acc[j] = 0.0;

    for (int s0 = 0; s0 < a.nSpecies; s0 += KS) {
        for (int k = threadIdx.x; k < KS * TC; k += 256) {
            const int sp = k / TC, c = k % TC, s = s0 + sp, cc = c0 + c;
            double v = 0.0;
            if (s < a.nSpecies && cc < C1 && cc > 0) {
                const int e = a.specElem[s];
                const size_t row = (size_t)(a.specIon[s] - 1) * a.nElemUsed + e;
                v = a.ionDen[row * C1 + cc] * a.abund[e] * a.Hden[cc];
            }
            s_n[sp][c] = v;
        }
        for (int k = threadIdx.x; k < KS * TBIN; k += 256) {
            const int sp = k / TBIN, bb = k % TBIN, s = s0 + sp, b = b0 + bb;
            s_sig[sp][bb] = (s < a.nSpecies && b < B) ? a.sigma[(size_t)s * B + b] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int sp = 0; sp < KS; ++sp) {
            const double n = s_n[sp][tx];
#pragma unroll
            for (int j = 0; j < BPL; ++j) acc[j] += n * s_sig[sp][ty * BPL + j];
        }
        __syncthreads();
    }
    if (cell >= C1) return;
    if (cell == 0) {
        for (int j = 0; j < BPL; ++j) {
            const int b = b0 + ty * BPL + j;
            if (b < B) a.opacity[(size_t)b * C1] = 0.0;
        }
        return;
    }
    // free-free prefactor (addOpacity :367-368) with density(n,i), i <= min(n, nstages):
    // the bare nuclei never enter, so only He+ and the heavy ions count
    const double Te = a.Te[cell], Ne = a.Ne[cell], H = a.Hden[cell];
    double dHe2 = 0.0, eSum = 0.0;
    for (int e = 0; e < a.nElemUsed; ++e) {
        const int z = a.elemZ[e];
        if (z == 2 && a.nstages >= 2)
            dHe2 = a.ionDen[((size_t)1 * a.nElemUsed + e) * C1 + cell] * a.abund[e] * H;
        if (z >= 3) {
            const int kmax = min(z - 1, a.nstages - 1);
            for (int k = 1; k <= kmax; ++k)
                eSum += (double)(k * k) *
                        (a.ionDen[((size_t)k * a.nElemUsed + e) * C1 + cell] * a.abund[e] * H);
        }
    }
    const double fac1 = (Ne / 1.e10) * ((dHe2 + eSum) / 1.e10) / sqrt(Te);
    const double fac2 = fac1 * kTe1Ryd / Te;
    const GauntT g = gaunt_setup(0.0, log10(Te));
    const double tb = kTe1Ryd / Te;
    for (int j = 0; j < BPL; ++j) {
        const int b = b0 + ty * BPL + j;
        if (b >= B) break;
        const double nu = a.nuArray[b];
        const double ex = tb + nu;
        // contBoltz is 0 from the first bin whose exponent exceeds 84 (nu ascending)
        const double boltz = ex > 84. ? 0.0 : exp(-ex);
        const double gff = gaunt_eval(g, a.log10nu[b]);
        const double ff = boltz < 0.995 ? a.brems[b] * (1. - boltz) * gff * fac1
                                        : a.brems[b] * nu * gff * fac2;
        double k = acc[j] + ff;
        if (a.lgCompton) k += a.KNsigmaT[b] * Ne;
        a.opacity[(size_t)b * C1 + cell] = k;
    }
}

}  // namespace

int moc_launch_opacity(moc_ctx* ctx, const MocOpacityArgs& a) {
    dim3 grid((unsigned)((a.C1 + TC - 1) / TC), (unsigned)((a.nbins + TBIN - 1) / TBIN));
    hipLaunchKernelGGL(k_opacity, grid, dim3(256), 0, ctx->stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("opacity launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

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

#include "gaunt.h"
#include "moc_internal.h"

namespace {

constexpr double kTe1Ryd = 1.578866e5;  // constants_mod.f90:25

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
    if (moc_unchanged(a.TeOld, a.NeOld, a.Te, a.Ne, cell)) return;
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

// The mixed-precision form's opacity (gas runs): the same tiles in FP32 --
// densities and cross-sections staged as FP32, the bound-free sum, the
// Boltzmann factor and the Gaunt fit in FP32 -- written to the FP32 plane the
// transport reads.  The per-cell prefactors are formed in FP64 and the
// free-free product brems(nu) * prefactor in FP64 before its single rounding
// (it spans 1e-47 .. 1e-3 over an X-ray mesh).  Terms below the FP32 range
// (trace ions) drop out; they are far below the Klein-Nishina floor n_e
// sigma_KN of every bin.
__global__ void __launch_bounds__(256) k_opacity_f32(MocOpacityArgs a) {
    __shared__ float s_n[KS][TC];
    __shared__ float s_sig[KS][TBIN];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int c0 = blockIdx.x * TC, b0 = blockIdx.y * TBIN;
    const int cell = c0 + tx;
    const int C1 = a.C1, B = a.nbins;
    float acc[BPL];
#pragma unroll
    for (int j = 0; j < BPL; ++j) acc[j] = 0.f;
    for (int s0 = 0; s0 < a.nSpecies; s0 += KS) {
        for (int k = threadIdx.x; k < KS * TC; k += 256) {
            const int sp = k / TC, c = k % TC, s = s0 + sp, cc = c0 + c;
            double v = 0.0;
            if (s < a.nSpecies && cc < C1 && cc > 0) {
                const int e = a.specElem[s];
                const size_t row = (size_t)(a.specIon[s] - 1) * a.nElemUsed + e;
                v = a.ionDen[row * C1 + cc] * a.abund[e] * a.Hden[cc];
            }
            s_n[sp][c] = (float)v;
        }
        for (int k = threadIdx.x; k < KS * TBIN; k += 256) {
            const int sp = k / TBIN, bb = k % TBIN, s = s0 + sp, b = b0 + bb;
            s_sig[sp][bb] = (s < a.nSpecies && b < B) ? (float)a.sigma[(size_t)s * B + b] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int sp = 0; sp < KS; ++sp) {
            const float n = s_n[sp][tx];
#pragma unroll
            for (int j = 0; j < BPL; ++j) acc[j] += n * s_sig[sp][ty * BPL + j];
        }
        __syncthreads();
    }
    if (cell >= C1) return;
    if (cell == 0) {
        for (int j = 0; j < BPL; ++j) {
            const int b = b0 + ty * BPL + j;
            if (b < B) a.opacity32[(size_t)b * C1] = 0.f;
        }
        return;
    }
    if (moc_unchanged(a.TeOld, a.NeOld, a.Te, a.Ne, cell)) return;
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
    const GauntF g = gaunt_f32(gaunt_setup(0.0, log10(Te)));
    const float tb = (float)(kTe1Ryd / Te), NeF = (float)Ne;
    for (int j = 0; j < BPL; ++j) {
        const int b = b0 + ty * BPL + j;
        if (b >= B) break;
        const float nu = (float)a.nuArray[b];
        const float ex = tb + nu;
        const float boltz = ex > 84.f ? 0.f : expf(-ex);
        const float gff = gaunt_eval_f32(g, (float)a.log10nu[b]);
        const float ff = boltz < 0.995f ? (float)(a.brems[b] * fac1) * (1.f - boltz) * gff
                                        : (float)(a.brems[b] * fac2) * nu * gff;
        float k = acc[j] + ff;
        if (a.lgCompton) k += (float)a.KNsigmaT[b] * NeF;
        a.opacity32[(size_t)b * C1 + cell] = k;
    }
}

}  // namespace

int moc_launch_opacity(moc_ctx* ctx, const MocOpacityArgs& a) {
    dim3 grid((unsigned)((a.C1 + TC - 1) / TC), (unsigned)((a.nbins + TBIN - 1) / TBIN));
    if (a.opacity32) hipLaunchKernelGGL(k_opacity_f32, grid, dim3(256), 0, ctx->stream, a);
    else hipLaunchKernelGGL(k_opacity, grid, dim3(256), 0, ctx->stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("opacity launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

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
                    const double T = sT[(s - 1) * a.nA + ai - 1];
                    if (!(T > 0. && T < a.tsub[s - 1])) continue;
                    const double c = a.Cabs[((size_t)(s - 1) * a.nA + (ai - 1)) * B + b];
                    v = v + c * moc_planck_h(nu, T) * w * a.weight[ai - 1] * a.abun[s - 1];
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

int moc_launch_dust_temp(moc_ctx* ctx, const MocDustArgs& a, const double* J, const double* W,
                         int ncol, double* R) {
    moc_launch_rates(ctx, J, a.nCells + 1, W, ncol, R);
    if (a.nCells < 1) return check(ctx, "dust absorption integrals");
    hipLaunchKernelGGL(k_dust_temp, dim3((unsigned)((a.nCells + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, ctx->stream, a);
    return check(ctx, "dust temperature launch");
}

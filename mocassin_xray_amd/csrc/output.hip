// Output reductions on gfx950 (outputGas and writeSED, output_mod.f90).
//
//  k_volavg -- the volume averages of outputGas (:642-730): for every ion,
//              sum over the cells that packets reached (lgBlack < 1) of
//              X Te Ne H dV and X Ne H dV.  Blocks of (ion, chunk of cells),
//              the ion's ionDen plane read cell-fastest; each block writes its
//              two partial sums, which the host adds chunk by chunk in order,
//              so the result does not depend on scheduling.
//  k_volavg_sum -- the same in-order chunk sums on the device, for the entry
//              point that keeps its partials in a fixed device buffer.
//  k_sed    -- writeSED (:2947-3064): the escaped energy of every bin summed
//              over the cells (rows 0..nCells), one block per bin.
// All arithmetic is FP64.
#include <hip/hip_runtime.h>

#include "moc_internal.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.;
    for (int k = 0; k < kThreads / 64; ++k) s += sh[k];
    __syncthreads();
    return s;
}

__global__ void __launch_bounds__(kThreads) k_volavg(const double* __restrict__ ionDen,
                                                     const double* __restrict__ Te,
                                                     const double* __restrict__ Ne,
                                                     const double* __restrict__ H,
                                                     const double* __restrict__ vol,
                                                     const int* __restrict__ black, int C1,
                                                     int nChunks, double* __restrict__ part) {
    __shared__ double sh[kThreads / 64];
    const int pair = blockIdx.y, chunk = blockIdx.x;
    const int per = (C1 - 1 + nChunks - 1) / nChunks;
    const int c0 = 1 + chunk * per;
    const int c1 = c0 + per < C1 ? c0 + per : C1;
    const double* X = ionDen + (size_t)pair * C1;
    double sT = 0., sW = 0.;
    for (int c = c0 + threadIdx.x; c < c1; c += kThreads) {
        if (black && black[c] > 0) continue;
        const double f2 = Ne[c] * H[c] * vol[c];           // NeUsed*HdenUsed*dV
        const double f1 = Te[c] * Ne[c] * H[c] * vol[c];   // TeUsed*NeUsed*HdenUsed*dV
        sT += X[c] * f1;
        sW += X[c] * f2;
    }
    sT = block_sum(sT, sh);
    sW = block_sum(sW, sh);
    if (threadIdx.x == 0) {
        part[((size_t)pair * nChunks + chunk) * 2] = sT;
        part[((size_t)pair * nChunks + chunk) * 2 + 1] = sW;
    }
}

__global__ void __launch_bounds__(kThreads) k_sed(const double* __restrict__ escaped, int C1,
                                                  double* __restrict__ sed) {
    __shared__ double sh[kThreads / 64];
    const int b = blockIdx.x;
    const double* row = escaped + (size_t)b * C1;
    double s = 0.;
    for (int c = threadIdx.x; c < C1; c += kThreads) s += row[c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) sed[b] = s;
}

// the partial sums of k_volavg added chunk by chunk in order (the order the
// host uses), one (ion) pair per lane; with no cells the sums are zero
__global__ void __launch_bounds__(kThreads) k_volavg_sum(const double* __restrict__ part, int nPairs,
                                                         int nChunks, int any, double* __restrict__ sums) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= nPairs) return;
    double t = 0., w = 0.;
    if (any)
        for (int c = 0; c < nChunks; ++c) {
            t += part[((size_t)p * nChunks + c) * 2];
            w += part[((size_t)p * nChunks + c) * 2 + 1];
        }
    sums[2 * p] = t;
    sums[2 * p + 1] = w;
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

int moc_launch_volavg(moc_ctx* ctx, const double* ionDen, const double* Te, const double* Ne,
                      const double* H, const double* vol, const int* black, int C1, int nPairs,
                      int nChunks, double* part) {
    if (nPairs < 1 || C1 < 2) return MOC_OK;
    hipLaunchKernelGGL(k_volavg, dim3((unsigned)nChunks, (unsigned)nPairs), dim3(kThreads), 0,
                       ctx->stream, ionDen, Te, Ne, H, vol, black, C1, nChunks, part);
    return check(ctx, "volume averages launch");
}

int moc_launch_sed(moc_ctx* ctx, const double* escaped, int C1, int B, double* sed) {
    if (B < 1) return MOC_OK;
    hipLaunchKernelGGL(k_sed, dim3((unsigned)B), dim3(kThreads), 0, ctx->stream, escaped, C1, sed);
    return check(ctx, "SED launch");
}

int moc_launch_volavg_sum(moc_ctx* ctx, const double* part, int nPairs, int nChunks, int C1,
                          double* sums) {
    if (nPairs < 1) return MOC_OK;
    hipLaunchKernelGGL(k_volavg_sum, dim3((unsigned)((nPairs + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       ctx->stream, part, nPairs, nChunks, C1 >= 2 ? 1 : 0, sums);
    return check(ctx, "volume average sums launch");
}

// Device calibration for the transport's roofline: the rate of scattered
// no-return FP64 atomic additions into global memory, the operation that
// tallies J (one per path segment).  No counterpart in the reference: it
// measures the machine, as a copy kernel measures HBM bandwidth.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "moc_internal.h"

namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// iters additions per lane at hashed indices below mask + 1 (a power of two)
__global__ void __launch_bounds__(256) k_atomic_rate(double* __restrict__ buf, uint32_t mask, int iters) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = 0; i < iters; ++i) atomicAdd(&buf[mix32(id * 2654435761u + (uint32_t)i * 40503u) & mask], 1.0);
}

}  // namespace

extern "C" int moc_calibrate_atomics(uint64_t bytes, double* rate) {
    if (!rate || bytes < (1u << 20) || bytes > ((uint64_t)1 << 32) || (bytes & (bytes - 1)))
        return MOC_ERR_ARG;
    *rate = 0.0;
    double* buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return MOC_ERR_HIP;
    const uint32_t mask = (uint32_t)(bytes / sizeof(double) - 1);
    const int blocks = 1024, iters = 256;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MOC_OK;
    float ms = 0.f, best = 0.f;
    if (hipMemset(buf, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = MOC_ERR_HIP;
    } else {
        for (int rep = 0; rep < 3 && rc == MOC_OK; ++rep) {  // the first launch warms up
            (void)hipEventRecord(e0, 0);
            hipLaunchKernelGGL(k_atomic_rate, dim3(blocks), dim3(256), 0, 0, buf, mask, iters);
            (void)hipEventRecord(e1, 0);
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = MOC_ERR_HIP;
            else if (rep > 0 && (best == 0.f || ms < best))
                best = ms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (rc == MOC_OK && best > 0.f) *rate = (double)blocks * 256.0 * iters / (best * 1e-3);
    return rc;
}

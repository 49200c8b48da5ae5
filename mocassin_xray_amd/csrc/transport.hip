// Energy-packet transport on gfx950: emission, Cartesian traversal with the
// Lucy (1999) path-length estimator of J_nu, Compton scattering, absorption
// and diffuse re-emission, escape tallies.
//
// Semantics follow MOCASSIN's energyPacketDriver (photon_mod.f90:28-414) for a
// single Cartesian grid without sub-grids, plane-parallel illumination, dust
// or fluorescence:
//   energyPacketRun          photon_mod.f90:418-741
//   initPhotonPacket         photon_mod.f90:744-896 (randomUnitVector
//                            vector_mod.f90:303-314, octant folding :875-882)
//   getNu2                   photon_mod.f90:928-972 (bin-1 fold kept)
//   newPhotonPacket          photon_mod.f90:976-1756 (stellar, diffuse)
//   pathSegment              photon_mod.f90:1758-4085
//   comptonScatter           photon_mod.f90:6805-6844
//
// GPU structure: one packet per lane; lanes pull packet ids from a device work
// counter and each loop trip runs either an emission or one cell-crossing
// segment, so the wave stays converged on the segment body while packets of
// very different lifetimes flow through it.  Geometry lives in the L1/L2
// (axes and widths are a few KB); opacity and J are (bin, cell) planes with
// the cell index fastest, which keeps the successive cells of one packet
// (same bin) in one plane.
#include <hip/hip_runtime.h>

#include "moc_internal.h"
#include "philox.h"

namespace {

constexpr double kPiRef = 3.141592654;  // constants_mod.f90:45 / vector_mod.f90:310
constexpr double kRyd2ergQ = 2.1799153e-11;  // photon_mod.f90:1471

enum : int { ST_DONE = 0, ST_EMIT = 1, ST_FLY = 2 };

template <typename T>
struct Pk {
    T x, y, z;     // position [cm]
    T u, v, w;     // direction
    T nu;          // frequency [Ryd]
    T absTau, passProb;
    int ix, iy, iz;  // 0-based cell indices
    int nu0;         // 0-based frequency bin
    int origin;      // active id of the emitting cell
    int stellar;
    int state;
    int emits;       // emission events of this packet
    int segs;        // segments of the current pathSegment call
};

template <typename T>
__device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }

// getNu2 (photon_mod.f90:928-972) on a monotone CDF of n entries: the linear
// scan's "last entry <= r" is an upper-bound bisection; bin 0 folds into bin 1.
template <typename T>
__device__ __forceinline__ int get_nu2(const T* __restrict__ cdf, int n, MocRng& rng,
                                       int* bad) {
    T r = rng.uniform<T>();
    for (int k = 0; k < 16 && (r == T(0) || r == T(1)); ++k) r = rng.uniform<T>();
    if (r == T(0) || r == T(1)) *bad = 1;
    int lo = 0, hi = n;  // first index with cdf[i] > r
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > r) hi = mid; else lo = mid + 1;
    }
    int k = lo;
    if (k < 1) k = 1;
    if (k > n - 1) k = n - 1;
    return k;
}

// locate (interpolation_mod.f90:49-106) returning the 0-based floor bin,
// clamped to [0, n-1]; used for the Compton-shifted frequency.
template <typename T>
__device__ __forceinline__ int locate_floor(const T* __restrict__ xa, int n, T x) {
    if (x > xa[n - 1]) return n - 1;
    if (x < xa[0]) return 0;
    int lo = 0, hi = n;  // first index with xa[i] > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xa[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo - 1;
}

template <typename T>
__device__ __forceinline__ void unit_vector(MocRng& rng, T& u, T& v, T& w) {
    for (int k = 0; k < 64; ++k) {
        const T r1 = rng.uniform<T>();
        w = T(2) * r1 - T(1);
        const T t = sqrt(T(1) - w * w);
        const T r2 = rng.uniform<T>();
        const T ang = T(kPiRef) * (T(2) * r2 - T(1));
        u = t * cos(ang);
        v = t * sin(ang);
        if (u != T(0) && v != T(0) && w != T(0)) return;
    }
}

// distance to the next wall along one axis; walls sit at axis midpoints and
// the outer faces at the first/last axis point (photon_mod.f90:1990-2094).
// Returns false when the packet snaps onto the outer face and is dropped.
template <typename T>
__device__ __forceinline__ bool wall_dist(T& pos, T dir, int& idx, const T* __restrict__ ax,
                                          int n, T& dS) {
    if (dir > T(0)) {
        if (idx < n - 1) {
            const T m = (ax[idx + 1] + ax[idx]) / T(2);
            dS = (m - pos) / dir;
            if (tabs(dS) < T(1.e-5)) { pos = m; ++idx; }
        } else {
            dS = (ax[n - 1] - pos) / dir;
            if (tabs(dS) < T(1.e-5)) { pos = ax[n - 1]; return false; }
        }
    } else if (dir < T(0)) {
        if (idx > 0) {
            const T m = (ax[idx] + ax[idx - 1]) / T(2);
            dS = (m - pos) / dir;
            if (tabs(dS) < T(1.e-5)) { pos = m; --idx; }
        } else {
            dS = (ax[0] - pos) / dir;
            if (tabs(dS) < T(1.e-5)) pos = ax[0];
        }
    } else {
        dS = ax[n - 1];
    }
    return true;
}

// the 2 % index nudge at the start of each segment (photon_mod.f90:1882-1919);
// the ratio uses the signed midpoint as the reference does.
template <typename T>
__device__ __forceinline__ void nudge(T pos, T dir, int& idx, const T* __restrict__ ax, int n) {
    if (idx < n - 1) {
        const T m = (ax[idx] + ax[idx + 1]) / T(2);
        if (dir > T(0) && tabs(pos - m) / m < T(2.e-2)) ++idx;
    }
    if (idx > 0) {
        const T m = (ax[idx] + ax[idx - 1]) / T(2);
        if (dir < T(0) && tabs(pos - m) / m < T(2.e-2)) --idx;
    }
}

// active cell id of (ix, iy, iz); a fully active grid numbered by active_map
// order needs no lookup, which takes one dependent memory access off a segment
template <typename T>
__device__ __forceinline__ int cell_of(const MocTransportArgs<T>& a, int ix, int iy, int iz) {
    if (a.dense) return 1 + iz + a.nz * (iy + a.ny * ix);
    return a.active[(size_t)ix + (size_t)a.nx * ((size_t)iy + (size_t)a.ny * (size_t)iz)];
}

struct Cnt {
    unsigned long long seg = 0, emit = 0, absb = 0, sca = 0, esc = 0, line = 0, lost = 0,
                       trap = 0, err = 0;
    double qphot = 0.0;
};

template <typename T>
__device__ __forceinline__ void tally_escape(const MocTransportArgs<T>& a, const Pk<T>& p,
                                             Cnt& c) {
    atomicAdd(&a.escaped[(size_t)p.nu0 * (size_t)(a.nCells + 1) + (size_t)p.origin],
              (double)a.deltaE);
    ++c.esc;
}

// newPhotonPacket + initPhotonPacket + the escape/line branches of
// energyPacketRun, then the start of pathSegment (tau = -ln(1-R)).
template <typename T>
__device__ void emit(const MocTransportArgs<T>& a, Pk<T>& p, MocRng& rng, Cnt& c) {
    ++c.emit;
    ++p.emits;
    int bad = 0;
    if (p.stellar) {
        p.nu0 = get_nu2(a.sourceCDF, a.nbins, rng, &bad);
        const T nu = a.nuArray[p.nu0];
        if (nu >= T(1)) c.qphot += (double)a.deltaE / (kRyd2ergQ * (double)nu);
    } else {
        const int cell = cell_of(a, p.ix, p.iy, p.iz);
        if (cell <= 0) { ++c.err; p.state = ST_DONE; return; }
        const T r = T(1) - rng.uniform<T>();
        if (r <= a.totalLines[cell]) {  // line photon: leaves the continuum transfer
            ++c.line;
            p.state = ST_DONE;
            return;
        }
        p.nu0 = get_nu2(a.recPDF + (size_t)cell * (size_t)a.nbins, a.nbins, rng, &bad);
    }
    if (bad) ++c.err;
    unit_vector(rng, p.u, p.v, p.w);
    if (a.sym && p.stellar) {
        p.u = tabs(p.u);
        p.v = tabs(p.v);
        p.w = tabs(p.w);
    }
    p.origin = cell_of(a, p.ix, p.iy, p.iz);
    p.nu = a.nuArray[p.nu0];
    if (p.nu < a.ionEdge1) {  // below every edge and no dust: escapes at once
        tally_escape(a, p, c);
        p.state = ST_DONE;
        return;
    }
    p.absTau = T(0);
    p.passProb = -log(T(1) - rng.uniform<T>());
    p.segs = 0;
    p.state = ST_FLY;
}

// Compton scattering (photon_mod.f90:6805-6844): the KN angle is used as the
// lab-frame azimuth with w isotropic, as in the reference.
template <typename T>
__device__ __forceinline__ void compton(const MocTransportArgs<T>& a, Pk<T>& p, MocRng& rng) {
    int bad = 0;
    const int th = get_nu2(a.KNcdf + (size_t)p.nu0 * (size_t)a.nAngles, a.nAngles, rng, &bad);
    const T theta = T(th + 1) * T(kPiRef) / T(180);
    const T w = T(2) * rng.uniform<T>() - T(1);
    const T t = sqrt(T(1) - w * w);
    p.u = t * cos(theta);
    p.v = t * sin(theta);
    p.w = w;
    const T nn = p.nu * a.Pcomp[(size_t)p.nu0 * (size_t)a.nAngles + th];
    p.nu0 = locate_floor(a.nuArray, a.nbins, nn);
    p.nu = nn;
}

// one trip of the pathSegment loop (photon_mod.f90:1880-4038)
template <typename T>
__device__ void segment(const MocTransportArgs<T>& a, Pk<T>& p, MocRng& rng, Cnt& c) {
    if (p.segs >= MOC_SAFE_LIMIT) {
        ++c.lost;
        p.state = ST_DONE;
        return;
    }
    ++p.segs;
    ++c.seg;
    const int nx = a.nx, ny = a.ny, nz = a.nz;

    nudge(p.x, p.u, p.ix, a.ax, nx);
    nudge(p.y, p.v, p.iy, a.ay, ny);
    nudge(p.z, p.w, p.iz, a.az, nz);
    if (p.ix < 0 || p.ix >= nx || p.iy < 0 || p.iy >= ny || p.iz < 0 || p.iz >= nz) {
        ++c.err;
        p.state = ST_DONE;
        return;
    }
    if (a.sym) {  // mirror walls of the octant model
        if (p.x <= a.ax[0]) { if (p.u < T(0)) p.u = -p.u; p.x = a.ax[0]; }
        if (p.y <= a.ay[0]) { if (p.v < T(0)) p.v = -p.v; p.y = a.ay[0]; }
        if (p.z <= a.az[0]) { if (p.w < T(0)) p.w = -p.w; p.z = a.az[0]; }
    }
    // the opacity of the current cell is fetched before the wall distances are
    // worked out; a wall snap below moves the packet to a neighbour (rarely) and
    // the value is fetched again for that cell
    const size_t stride = (size_t)(a.nCells + 1);
    const int cell0 = cell_of(a, p.ix, p.iy, p.iz);
    const T kap0 = a.opacity[(size_t)p.nu0 * stride + (size_t)cell0];
    T dSx = T(0), dSy = T(0), dSz = T(0);
    if (!wall_dist(p.x, p.u, p.ix, a.ax, nx, dSx) || !wall_dist(p.y, p.v, p.iy, a.ay, ny, dSy) ||
        !wall_dist(p.z, p.w, p.iz, a.az, nz, dSz)) {
        ++c.lost;
        p.state = ST_DONE;
        return;
    }
    if (tabs(dSx) < T(1.e-5)) dSx = a.ax[nx - 1];
    if (tabs(dSy) < T(1.e-5)) dSy = a.ay[ny - 1];
    if (tabs(dSz) < T(1.e-5)) dSz = a.az[nz - 1];
    dSx = tabs(dSx);
    dSy = tabs(dSy);
    dSz = tabs(dSz);
    T dS;
    if (dSx <= T(0)) dS = fmin(dSy, dSz);
    else if (dSy <= T(0)) dS = fmin(dSx, dSz);
    else if (dSz <= T(0)) dS = fmin(dSx, dSy);
    else dS = fmin(fmin(dSx, dSy), dSz);
    if (!(dS > T(0))) {
        ++c.err;
        p.state = ST_DONE;
        return;
    }

    const int cell = cell_of(a, p.ix, p.iy, p.iz);
    const T kap = cell == cell0 ? kap0 : a.opacity[(size_t)p.nu0 * stride + (size_t)cell];
    const T tauCell = dS * kap;
    const T dV = a.vx[p.ix] * a.vy[p.iy] * a.vz[p.iz];

    if (p.absTau + tauCell > p.passProb && cell > 0) {
        // the packet interacts inside this cell
        const T dl = (p.passProb - p.absTau) / kap;
        p.x = p.x + dl * p.u;
        p.y = p.y + dl * p.v;
        p.z = p.z + dl * p.w;
        if (a.sym) {
            if (p.x <= a.ax[0]) { if (p.u < T(0)) p.u = -p.u; p.x = a.ax[0]; }
            if (p.y <= a.ay[0]) { if (p.v < T(0)) p.v = -p.v; p.y = a.ay[0]; }
            if (p.z <= a.az[0]) { if (p.w < T(0)) p.w = -p.w; p.z = a.az[0]; }
        }
        atomicAdd(&a.J[(size_t)p.nu0 * stride + (size_t)cell], (double)(dl * a.deltaE / dV));
        if (a.R_out > T(0)) {
            // radius with the 1e10 scaling of photon_mod.f90:2294 (no FP32 overflow)
            const T sx = p.x / T(1.e10), sy = p.y / T(1.e10), sz = p.z / T(1.e10);
            const T rr = sqrt(sx * sx + sy * sy + sz * sz) * T(1.e10);
            if (rr >= a.R_out) {
                tally_escape(a, p, c);
                p.state = ST_DONE;
                return;
            }
        }
        T probSca = T(0);
        if (a.lgCompton) probSca = a.KNsigmaT[p.nu0] * a.Ne[cell] / kap;
        const T r = T(1) - rng.uniform<T>();
        if (r <= probSca) {
            ++c.sca;
            compton(a, p, rng);
            p.origin = cell;
            p.absTau = T(0);
            p.passProb = -log(T(1) - rng.uniform<T>());
            return;
        }
        // absorbed: re-emitted as a diffuse packet from this position
        ++c.absb;
        if (p.emits >= MOC_RECURSION_LIMIT) {
            ++c.trap;
            p.state = ST_DONE;
            return;
        }
        p.stellar = 0;
        p.state = ST_EMIT;
        return;
    }

    // the packet crosses the cell
    if (cell > 0)
        atomicAdd(&a.J[(size_t)p.nu0 * stride + (size_t)cell], (double)(dS * a.deltaE / dV));
    p.absTau = p.absTau + tauCell;
    p.x = p.x + dS * p.u;
    p.y = p.y + dS * p.v;
    p.z = p.z + dS * p.w;
    if (dS == dSx && p.u > T(0)) ++p.ix;
    else if (dS == dSx && p.u < T(0)) --p.ix;
    else if (dS == dSy && p.v > T(0)) ++p.iy;
    else if (dS == dSy && p.v < T(0)) --p.iy;
    else if (dS == dSz && p.w > T(0)) ++p.iz;
    else if (dS == dSz && p.w < T(0)) --p.iz;

    if (!a.sym) {
        bool ret = false;
        if (p.y < a.ay[0] || p.iy < 0) { p.iy = 0; ret = true; }
        if (p.y > a.ay[ny - 1] || p.iy > ny - 1) { p.iy = ny - 1; ret = true; }
        if (p.x <= a.ax[0] || p.ix < 0) { p.ix = 0; ret = true; }
        if (p.x >= a.ax[nx - 1] || p.ix > nx - 1) { p.ix = nx - 1; ret = true; }
        if (p.z <= a.az[0] || p.iz < 0) { p.iz = 0; ret = true; }
        if (p.z >= a.az[nz - 1] || p.iz > nz - 1) { p.iz = nz - 1; ret = true; }
        if (ret) {
            tally_escape(a, p, c);
            p.state = ST_DONE;
            return;
        }
    }
    if ((!a.sym && (p.x <= a.ax[0] || p.y <= a.ay[0] || p.z <= a.az[0])) ||
        p.x >= a.ax[nx - 1] || p.y >= a.ay[ny - 1] || p.z >= a.az[nz - 1] || p.ix > nx - 1 ||
        p.iy > ny - 1 || p.iz > nz - 1) {
        tally_escape(a, p, c);
        p.state = ST_DONE;
        return;
    }
    if (a.sym) {
        if (p.x <= a.ax[0] || p.ix < 0) { if (p.u < T(0)) p.u = -p.u; p.ix = 0; p.x = a.ax[0]; }
        if (p.y <= a.ay[0] || p.iy < 0) { if (p.v < T(0)) p.v = -p.v; p.iy = 0; p.y = a.ay[0]; }
        if (p.z <= a.az[0] || p.iz < 0) { if (p.w < T(0)) p.w = -p.w; p.iz = 0; p.z = a.az[0]; }
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__global__ void __launch_bounds__(256) k_transport(MocTransportArgs<T> a) {
    Cnt c;
    Pk<T> p;
    p.state = ST_DONE;
    MocRng rng;
    bool more = true;
    while (true) {
        if (p.state == ST_DONE) {
            if (more) {
                const unsigned long long id = atomicAdd(&a.counters[MOC_CNT_NEXT], 1ull);
                if ((int64_t)id < a.n) {
                    rng.init(a.seed, id, a.ctr_base);
                    p.x = a.star[0];
                    p.y = a.star[1];
                    p.z = a.star[2];
                    p.ix = a.starIdx[0];
                    p.iy = a.starIdx[1];
                    p.iz = a.starIdx[2];
                    p.stellar = 1;
                    p.emits = 0;
                    p.state = ST_EMIT;
                } else {
                    more = false;
                }
            }
            if (!more) break;
        }
        if (p.state == ST_EMIT) emit(a, p, rng, c);
        if (p.state == ST_FLY) segment(a, p, rng, c);
    }
    // all lanes of the wave have left the loop: reduce the counters once
    unsigned long long* g = a.counters;
    const unsigned long long s0 = wave_sum(c.seg), s1 = wave_sum(c.emit), s2 = wave_sum(c.absb),
                             s3 = wave_sum(c.sca), s4 = wave_sum(c.esc), s5 = wave_sum(c.line),
                             s6 = wave_sum(c.lost), s7 = wave_sum(c.trap), s8 = wave_sum(c.err);
    const double q = wave_sum(c.qphot);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&g[MOC_CNT_SEGMENTS], s0);
        atomicAdd(&g[MOC_CNT_EMISSIONS], s1);
        atomicAdd(&g[MOC_CNT_ABSORBED], s2);
        atomicAdd(&g[MOC_CNT_SCATTERED], s3);
        atomicAdd(&g[MOC_CNT_ESCAPED], s4);
        atomicAdd(&g[MOC_CNT_LINES], s5);
        atomicAdd(&g[MOC_CNT_LOST], s6);
        atomicAdd(&g[MOC_CNT_TRAPPED], s7);
        atomicAdd(&g[MOC_CNT_ERROR], s8);
        atomicAdd(a.Qphot, q);
    }
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) d[k] = (float)s[k];
}

__global__ void k_scale(double* __restrict__ p, int64_t n, double f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) p[k] *= f;
}

template <typename T>
int launch(moc_ctx* ctx, const MocTransportArgs<T>& a, float* ms) {
    if (a.n <= 0) {
        if (ms) *ms = 0.f;
        return MOC_OK;
    }
    // a resident wave count that fills 256 CUs several times over; lanes beyond
    // the packet count exit at their first fetch
    const int64_t lanes_max = 256LL * 32LL * 64LL;
    int64_t lanes = a.n < lanes_max ? a.n : lanes_max;
    const int block = 256;
    const int blocks = (int)((lanes + block - 1) / block);
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    hipLaunchKernelGGL(k_transport<T>, dim3(blocks), dim3(block), 0, ctx->stream, a);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("transport launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    e = hipEventSynchronize(ctx->ev1);
    if (e != hipSuccess) {
        ctx->err = std::string("transport kernel: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    if (ms) (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return MOC_OK;
}

}  // namespace

int moc_launch_transport_f64(moc_ctx* ctx, const MocTransportArgs<double>& a, float* ms) {
    return launch<double>(ctx, a, ms);
}
int moc_launch_transport_f32(moc_ctx* ctx, const MocTransportArgs<float>& a, float* ms) {
    return launch<float>(ctx, a, ms);
}
void moc_convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}
void moc_scale_f64(double* p, int64_t n, double f, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(256), 0, s, p, n, f);
}

// Per-cell ionisation and thermal balance on gfx950 (updateCell,
// update_mod.f90:15-482 and the routines it calls).
//
// Two kernels:
//  k_rates   -- every frequency integral of the balance is linear in J_nu:
//               per-shell photoionisation and heating rates (photoionisation /
//               getPhotoRates, :3584-3832, :3971-4040) and the Compton sums
//               (compton :3834-3897).  They are one contraction
//               R[cell, col] = sum_nu J[nu, cell] W[nu, col] with weight columns
//               built once on the host; tiles of 64 cells x 32 columns stage J
//               and W through LDS, J planes are read cell-fastest (coalesced),
//               and R is written a row per cell (the cell's columns together,
//               as the balance reads them).
//  k_balance -- one wave per cell: the T_e root-find (iterateT :277-482)
//               around the ion-fraction chain (ionBalance :1575-1992) and the
//               heating/cooling integrals (thermBalance :887-1215) with the
//               n-level forbidden-line cooling (equilibrium, emission_mod.f90:
//               2505-3106).  Everything in FP64.  Recombination coefficients
//               (calcAlpha :2298-2413) and collisional ionisation (:3552-3582)
//               are evaluated per iteration from tables resident in HBM/L2;
//               the cell's rate columns and the shell index tables are staged
//               in LDS; the Upsilon(T) splines use second derivatives computed
//               once on the host.
#include <hip/hip_runtime.h>

#include "moc_internal.h"
#include "nlevel.h"

namespace {

constexpr int NEL = 30;
constexpr double kHcRyd = 2.1799153e-11;  // update_mod.f90:64
constexpr double kRyd2erg = 2.1799e-11;   // constants_mod.f90:20
constexpr double kBoltz = 1.380658e-16;   // constants_mod.f90:17
constexpr double kXMax = 0.99999;         // constants_mod.f90:55

// ---------------------------------------------------------------- k_rates
constexpr int RC = 64;  // cells per tile
constexpr int RK = 32;  // bins per LDS stage
constexpr int RN = 32;  // columns per tile

// cells cBase .. cBase + nc - 1 of the J planes (row stride ldJ) into
// R[local cell][col].  Band form (perm != nullptr): W's columns are stored
// sorted by the bins where they are nonzero, column n of W being column
// perm[n] of R, and column tile t is contracted over its bins bLo[t] ..
// bHi[t] only (every column of the tile is zero outside them, so the sums
// are those of the full contraction).
__global__ void __launch_bounds__(256) k_rates(const double* __restrict__ J, int ldJ, int cBase,
                                               int nc, int B, const double* __restrict__ W,
                                               int ncol, double* __restrict__ R,
                                               const int* __restrict__ perm,
                                               const int* __restrict__ bLo,
                                               const int* __restrict__ bHi) {
    __shared__ double sJ[RK][RC];
    __shared__ double sW[RK][RN];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 4 waves x 8 columns
    const int c0 = blockIdx.x * RC, n0 = blockIdx.y * RN;
    const int bFirst = perm ? bLo[blockIdx.y] : 0, bEnd = perm ? bHi[blockIdx.y] + 1 : B;
    double acc[RN / 4];
#pragma unroll
    for (int j = 0; j < RN / 4; ++j) acc[j] = 0.0;
    for (int b0 = bFirst; b0 < bEnd; b0 += RK) {
        for (int k = threadIdx.x; k < RK * RC; k += 256) {
            const int kb = k / RC, kc = k % RC, b = b0 + kb, c = c0 + kc;
            sJ[kb][kc] = (b < B && c < nc) ? J[(size_t)b * ldJ + cBase + c] : 0.0;
        }
        for (int k = threadIdx.x; k < RK * RN; k += 256) {
            const int kb = k / RN, kn = k % RN, b = b0 + kb, n = n0 + kn;
            sW[kb][kn] = (b < B && n < ncol) ? W[(size_t)b * ncol + n] : 0.0;
        }
        __syncthreads();
        for (int kb = 0; kb < RK; ++kb) {
            const double j = sJ[kb][tx];
#pragma unroll
            for (int q = 0; q < RN / 4; ++q) acc[q] += j * sW[kb][ty * (RN / 4) + q];
        }
        __syncthreads();
    }
    // the tile goes out through LDS so that each cell's 32 columns are one
    // contiguous write
    double* sOut = &sJ[0][0];  // [RC][RN]
#pragma unroll
    for (int q = 0; q < RN / 4; ++q) sOut[tx * RN + ty * (RN / 4) + q] = acc[q];
    __syncthreads();
    for (int k = threadIdx.x; k < RC * RN; k += 256) {
        const int c = c0 + k / RN, n = n0 + k % RN;
        if (c < nc && n < ncol) R[(size_t)c * ncol + (perm ? perm[n] : n)] = sOut[k];
    }
}

// ---------------------------------------------------------------- k_balance
// One wave per cell.  The cell's state (ion fractions, recombination and
// collisional-ionisation coefficients, the heavy-element chain work rows and
// the per-ion cooling terms) lives in LDS; per-ion work -- recombination
// coefficients, collisional ionisation, the heavy-element chains of
// ionBalance, the n-level cooling of every ion with data -- is spread over the
// lanes, and every sum is then taken by all lanes in the reference's order, so
// each lane holds the same (wave-uniform) Te, Ne and control state and the
// results equal a sequential evaluation operation for operation.

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// sum over the wave, the same value on every lane (butterfly)
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double t3(const double* a, int i, int z, int n) {
    return a[((size_t)i * 31 + z) * 31 + n];
}

// LDS rows of one cell: ion fractions, recombination and collisional-
// ionisation coefficients, and the work row of the heavy-element chains.
// Each row packs the elements one after the other, element e holding stages
// 0 .. min(Z+1, NS) from off[e].  Loops that spread (element, stage) over the
// lanes run over k = e * S2 + stage (S2 = NS + 2) and skip stages past the end.
// The recombination and collisional-ionisation rows are kept in the fit
// precision S (FP32 in the mixed form, where the fits are FP32 anyway), which
// makes a cell's LDS smaller and lets more cells run per CU.
template <class S>
struct LdsT {
    double* X;
    S *alpha, *cI1;
    double *out, *cool, *diO;
    int* hot;
    int* off;  // [NE + 1]
    int S2;
    __device__ int len(int e) const { return off[e + 1] - off[e]; }
    __device__ double& x(int e, int i) const { return X[off[e] + i]; }
    __device__ S& al(int e, int i) const { return alpha[off[e] + i]; }
    __device__ S& c1(int e, int i) const { return cI1[off[e] + i]; }
};

// LDS of a cell's state rows; fitBytes: the size of one fit value (4 or 8)
__host__ __device__ inline size_t balance_lds_bytes(int nX, int nNlv, size_t fitBytes) {
    const size_t rows = (size_t)2 * nX * sizeof(double) + (size_t)2 * nX * fitBytes;
    return rows + ((size_t)(nNlv > 0 ? nNlv : 1) + MOC_MAXS + 1) * sizeof(double) +
           (MOC_MAXE + 2) * sizeof(int);
}

// the staged rate columns and index tables that follow the state rows
__host__ __device__ inline size_t balance_stage_bytes(const MocBalanceArgs& a) {
    const size_t ints = (size_t)a.NE * 32 + (size_t)a.nSpec + 1 + 2 * (size_t)a.nShells;
    return (size_t)a.ncol * sizeof(double) + ((ints * sizeof(int) + 7) & ~(size_t)7);
}
// staging is used while one workgroup's LDS stays within this
constexpr size_t kMaxBalanceLds = 49152;
// LDS ceiling of one balance workgroup (gfx950: 160 KiB per CU)
constexpr size_t kMaxGrainLds = 65536;

// dielectronic recombination of the ions with more than 11 electrons
// (Aldrovandi & Pequignot above 6e4 K, Nussbaumer & Storey below;
// dielectronic :3204-3305): the last matching row of the data wins, which the
// host finds once per (Z, stage) and temperature side of 2e4 K
template <class R>
__device__ R direc(const MocBalanceArgs& a, int z, int ion, R T) {
    const R tt = T > (R)1000. ? T / (R)10000. : (R)0.1;
    if (T > (R)60000.) {
        const int i = a.aldroIdx[z * 32 + ion];
        if (i < 0) return 0;
        const double* r = a.aldro + 6 * i;
        return (R)r[2] * pow(T, (R)-1.5) * exp(-(R)r[4] / T) * ((R)1. + (R)r[3] * exp(-(R)r[5] / T));
    }
    const int i = a.dielIdx[(z * 32 + ion) * 2 + (T < (R)20000. ? 0 : 1)];
    if (i < 0) return 0;
    const double* r = a.diel + 8 * i;
    return (R)1.e-12 * ((R)r[2] / tt + (R)r[3] + (R)r[4] * tt + (R)r[5] * tt * tt) * pow(tt, (R)-1.5) *
           exp(-(R)r[6] / tt);
}

// recombination coefficients of every ion (calcAlpha :2298-2413), one
// (element, stage) per lane; the fits in precision R, stored in R
template <class R>
__device__ void recomb(const MocBalanceArgs& a, const LdsT<R>& L, double Td, int tid) {
    const int NE = a.NE, NS = a.NS, S2 = L.S2;
    const int eO = a.elemSlot[8];
    const R T = (R)Td;
    for (int k = tid; k < a.nX; k += 64) L.alpha[k] = 0.0;
    for (int i = tid; i <= MOC_MAXS; i += 64)
        L.diO[i] = (eO >= 0 && i >= 1 && i < 7 && i <= NS) ? (double)direc<R>(a, 8, i, T) : 0.0;
    __syncthreads();
    for (int k = tid; k < NE * S2; k += 64) {
        const int e = k / S2, ion = k % S2;
        const int z = a.elemZ[e];
        if (ion < 1 || ion > imin(NS - 1, z)) continue;
        const int nel = z - ion;
        if (nel <= 11) {  // Badnell radiative + dielectronic fits (:2314-2330)
            const R bb = (R)t3(a.rrb, 1, z, nel) + (R)t3(a.rrb, 4, z, nel) * exp(-(R)t3(a.rrb, 5, z, nel) / T);
            const R s0 = sqrt(T / (R)t3(a.rrb, 2, z, nel)), s1 = sqrt(T / (R)t3(a.rrb, 3, z, nel));
            R dr = 0;
            for (int q = 0; q < 9; ++q) {
                const R E = (R)t3(a.drb_e, q, z, nel);
                if (E / T < (R)87.) dr += (R)t3(a.drb_c, q, z, nel) * exp(-E / T);
            }
            L.al(e, ion) = (R)t3(a.rrb, 0, z, nel) /
                               (s0 * pow((R)1. + s0, (R)1. - bb) * pow((R)1. + s1, (R)1. + bb)) +
                           dr * pow(T, (R)-1.5);
            continue;
        }
        // radRecFit + dielectronic, oxygen's dielectronic borrowed when zero (:2334-2362)
        const int n = z - ion + 1;
        R rr;
        if (n <= 3 || n == 11 || (z > 5 && z < 9) || z == 10 || (z == 26 && n > 11)) {
            const R q = sqrt(T / (R)t3(a.rnew, 2, z, n)), p = (R)t3(a.rnew, 1, z, n);
            rr = (R)t3(a.rnew, 0, z, n) /
                 (q * pow(q + (R)1., (R)1. - p) * pow((R)1. + sqrt(T / (R)t3(a.rnew, 3, z, n)), (R)1. + p));
        } else {
            const R q = T * (R)1.e-4;
            if (z == 26 && n <= 13) {
                const double fe0[10] = {4.33e-10, 3.91e-10, 3.49e-10, 3.16e-10, 2.96e-10,
                                        2.59e-10, 2.24e-10, 1.91e-10, 1.68e-10, 1.46e-10};
                const double fe1[10] = {0.531, 0.523, 0.521, 0.534, 0.557, 0.567, 0.579, 0.601, 0.602, 0.597};
                const double fe2[10] = {5.77e-02, 6.15e-02, 6.22e-02, 6.02e-02, 5.79e-02,
                                        5.65e-02, 5.49e-02, 5.10e-02, 5.07e-02, 5.22e-02};
                rr = (R)fe0[n - 4] / pow(q, (R)fe1[n - 4] + (R)fe2[n - 4] * log10(q));
            } else {
                rr = (R)t3(a.rrec, 0, z, n) / pow(q, (R)t3(a.rrec, 1, z, n));
            }
        }
        double di = (double)direc<R>(a, z, ion, T);
        if (di == 0. && ion < 7 && eO >= 0) di = L.diO[ion];
        L.al(e, ion) = (R)((double)rr + (di > 0. ? di : 0.));
    }
    __syncthreads();
}

// Voronov collisional ionisation (:3552-3582), one (element, stage) per lane;
// the reference leaves the loops at the first u > 80, so that entry and every
// later one (in its loop order) stay zero
template <class R>
__device__ void collion(const MocBalanceArgs& a, const LdsT<R>& L, double Te, int tid) {
    const int NE = a.NE, NS = a.NS, S2 = L.S2;
    const R te = (R)Te * (R)8.617385e-05;
    for (int k = tid; k < a.nX; k += 64) L.cI1[k] = 0.0;
    if (tid == 0) *L.hot = NE * S2;
    __syncthreads();
    for (int k = tid; k < NE * S2; k += 64) {
        const int e = k / S2, ion = k % S2;
        const int z = a.elemZ[e];
        if (ion < 1 || ion > imin(NS - 1, z - 1)) continue;
        const int el = z - ion + 1;
        const R u = (R)t3(a.cion_cf, 0, z, el) / te;
        if (u > (R)80.0) {
            atomicMin(L.hot, k);
            continue;
        }
        const R c = (R)t3(a.cion_cf, 2, z, el) * ((R)1.0 + (R)t3(a.cion_cf, 1, z, el) * sqrt(u)) /
                    ((R)t3(a.cion_cf, 3, z, el) + u) * pow(u, (R)t3(a.cion_cf, 4, z, el)) * exp(-u);
        L.c1(e, ion) = c;
    }
    __syncthreads();
    const int hot = *L.hot;
    for (int k = tid; k < NE * S2; k += 64) {
        const int e = k / S2, ion = k % S2;
        if (k >= hot && ion < L.len(e)) L.c1(e, ion) = 0.0;
    }
    __syncthreads();
}

// charge-exchange coefficient [cm^3/s]; T-scaled unless raw (ionBalance keeps the
// heavy elements' coefficients unscaled); the fit in precision R
template <class R>
__device__ double chex(const MocBalanceArgs& a, int z, int ion, double Td, bool scaled) {
    const int k = a.chexIdx[z * 32 + ion];
    if (k < 0) return 0.0;
    const double* c = a.chexTab + 4 * k;
    const double v0 = c[0] * 1.e-9;
    if (!scaled) return v0;
    const R T = (R)Td;
    if (T < (R)6000. || T > (R)5.e4) return 0.0;
    const R t4 = T / (R)10000.;
    const R v = (R)v0 * pow(t4, (R)c[1]) * ((R)1. + (R)c[2] * exp((R)c[3] * t4));
    return v < (R)0. ? 0. : (double)v;
}

__device__ __forceinline__ double dEk(int z, int ion) {
    return (z == 7 && ion == 1) ? 10863. : ((z == 8 && ion == 1) ? 2205. : 0.);
}

// photoionisation rates of the cell from k_rates (per shell), with the
// reverse charge exchange added to the outer shell (:3715-3830), and the
// species/shell index tables the chains walk.  With ST the rate columns and
// the index tables are LDS copies made when the cell starts (the chains read
// them on one lane per element, where every global load is a serial
// round trip); without ST they are read from the [ncol][C1] rate array and
// the tables in HBM.
template <bool ST>
struct Rates {
    const double* R;       // the cell's row of rates (staged: in LDS)
    int ld, cell;
    const int* specOf;     // [NE*32] (slot, stage) -> species or -1
    const int* specShell;  // [S+1]
    const int* shInfo;     // [nShells] Auger count | (1 + reverse charge-exchange slot) << 8
    const int* shAug;      // [nShells] first auger[] entry of the shell
    __device__ double at(int col) const {
        if constexpr (ST) return R[col];
        else return R[(size_t)cell * ld + col];
    }
    __device__ double gamma(int s) const { return at(3 * s); }
    __device__ double heat(int s, double heatef) const { return at(3 * s + 1) + heatef * at(3 * s + 2); }
};

template <class R>
__device__ double rev_chex(const MocBalanceArgs& a, const LdsT<R>& L, double Te, double H, int z, int ion) {
    if (!(dEk(z, ion) > 1.)) return 0.0;
    const int nel = z - ion + 1;
    const double g0g1 = (double)a.statW[z * 32 + nel] / (double)a.statW1[z * 32 + nel];
    return chex<R>(a, z, ion, Te, true) * (1. - L.x(a.elemSlot[1], 1)) * H * 2. *
           (double)exp(-(R)dEk(z, ion) / (R)Te) / g0g1;
}

template <class RT>
__device__ double gamma_out(const RT& r, int spec) {
    double g = 0.0;
    for (int sh = r.specShell[spec]; sh < r.specShell[spec + 1]; ++sh) g += r.gamma(sh);
    return g;
}

// H or He ratio chain (:1715-1790) on the lane that owns element zz
template <class R, class RT>
__device__ void hhe_chain(const MocBalanceArgs& a, const RT& r, const LdsT<R>& L, int zz, double Te,
                          double Ne, double H, double H0, double collIonH) {
    const int NS = a.NS, e = a.elemSlot[zz];
    double ratio[3] = {0., 0., 0.};
    for (int ion = 1; ion <= imin(zz, NS - 1); ++ion) {
        const int spec = r.specOf[e * 32 + ion];
        const double cx = chex<R>(a, zz, ion, Te, true);
        if (zz > 1 || ion > 1) collIonH = 0.;
        const double g = (spec >= 0 ? gamma_out(r, spec) : 0.0) + 1.e-20;
        ratio[ion] = (g + collIonH * Ne) / (Ne * L.al(e, ion) + cx * H * H0);
    }
    double prod[3] = {1., 1., 1.}, den = 1.;
    for (int ion = 1; ion <= imin(zz, NS - 1); ++ion) {
        for (int i = 1; i <= ion; ++i) prod[ion] *= ratio[i];
        den += prod[ion];
    }
    for (int ion = 1; ion <= imin(zz + 1, NS); ++ion) {
        double v = ion == 1 ? 1. / den : prod[ion - 1] / den;
        v = v > kXMax ? kXMax : v;
        v = v < 1.e-20 ? 1.e-20 : v;
        L.x(e, ion) = v;
    }
}

// The heavy-element chains (ionBalance :1826-1929) in three lane-parallel
// steps.  (1) The rates out of each stage: task t of the host's gather lists
// sums, on one lane, the terms the sequential walk over (stage, shell, Auger
// electrons) adds to out(stage) -- collisional ionisation, the shells'
// photoionisation (the N I / O I outer shell with the frozen reverse charge
// exchange) and the Auger terms of lower stages at the ratio X(ion)/X(stage)
// -- in the walk's order, so that each sum is the sequential one.
template <class R, class RT>
__device__ void heavy_out(const MocBalanceArgs& a, const RT& r, const LdsT<R>& L, double Ne,
                          const double revFrozen[2], int tid) {
    constexpr int U = 4;  // terms whose loads are issued together
    for (int t = tid; t < a.nHTask; t += 64) {
        const int e = a.htSlot[t], mx = a.htMx[t];
        const int q1 = a.htBeg[t + 1];
        const double xm = L.x(e, mx), c1Ne = L.c1(e, mx) * Ne;
        double o = 0.;
        for (int q0 = a.htBeg[t]; q0 < q1; q0 += U) {
            int code[U];
            double y[U], g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                code[u] = q0 + u < q1 ? a.htCode[q0 + u] : 3;  // kind 3: past the task's end
                y[u] = q0 + u < q1 ? a.htAug[q0 + u] : 0.;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int sh = code[u] >> 9;
                g[u] = (code[u] & 3) == 1 || (code[u] & 3) == 2 ? r.gamma(sh) : 0.;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kind = code[u] & 3, cx = (code[u] >> 2) & 3, ion = (code[u] >> 4) & 31;
                if (kind == 3) break;
                if (kind == 0) {
                    o += c1Ne;
                    continue;
                }
                double gg = g[u];
                if (cx) gg += revFrozen[cx - 1];
                if (kind == 1) {
                    o += gg;
                } else {
                    const double ratio = xm > 1e-30 ? L.x(e, ion) / xm : 1.;
                    o += gg * y[u] * ratio;
                }
            }
        }
        L.out[L.off[e] + mx] = o;
    }
}

// (2) The chain products of element e on one lane: mat(ion) = mat(ion-1)
// out(ion-1) / in(ion-1) with mat(1) = 1 and in = Ne alpha + chex H0 H (the
// coefficient unscaled); mat replaces out in place and 1 / sum(mat) goes to
// the unused stage-0 entry.
template <class R>
__device__ void heavy_products(const MocBalanceArgs& a, const LdsT<R>& L, int e, double Ne, double H) {
    const int z = a.elemZ[e], top = imin(a.NS, z + 1);
    double* out = L.out + L.off[e];
    const double xH = L.x(a.elemSlot[1], 1);
    double m = 1., sum = 1.;
    for (int ion = 2; ion <= top; ++ion) {
        const double in = Ne * L.al(e, ion - 1) + a.chexU[z * 32 + ion - 1] * H * xH;
        const double mn = m * out[ion - 1] / in;
        out[ion - 1] = m;
        m = mn;
        sum += mn;
    }
    out[top] = m;
    out[0] = 1. / sum;
}

// (3) The new fractions, one (element, stage) per lane: X = mat / sum, on the
// first pass the geometric mean with the old value, clamped to
// [1e-20, 0.99999] below the top stage.
template <class R>
__device__ void heavy_update(const MocBalanceArgs& a, const LdsT<R>& L, int nIterX, int tid) {
    for (int k = tid; k < a.NE * L.S2; k += 64) {
        const int e = k / L.S2, ion = k % L.S2;
        const int z = a.elemZ[e], top = imin(a.NS, z + 1);
        if (z < 3 || ion < 1 || ion > top) continue;
        const double* mat = L.out + L.off[e];
        const double sum = mat[0];
        if (ion == top) {
            L.x(e, top) = mat[top] * sum;
            continue;
        }
        double v;
        if (nIterX < 2 && mat[ion] * sum > 0.) {
            const double old = L.x(e, ion) <= 0. ? 1.e-5 : L.x(e, ion);
            v = pow(10., 0.5 * (log10(old) + log10(mat[ion] * sum)));
        } else {
            v = mat[ion] * sum;
        }
        v = v > kXMax ? kXMax : v;
        v = v < 1.e-20 ? 1.e-20 : v;
        L.x(e, ion) = v;
    }
}

// ionBalance (:1575-1992); returns false when 15 passes did not settle H/He
template <class R, class RT>
__device__ bool ion_balance(const MocBalanceArgs& a, const RT& r, const LdsT<R>& L, double Te,
                            double& Ne, double H, const double revFrozen[2], int tid) {
    const int NS = a.NS, NE = a.NE;
    const int eH = a.elemSlot[1], eHe = a.elemSlot[2];
    int nIterX = 1;
    bool conv = true;
    for (;;) {
        const R Tr = (R)Te;
        R expFact = (R)157893.94 / Tr;
        if (expFact > (R)75.) expFact = (R)75.;
        const double collIonH = (double)((R)2.75E-16 * Tr * sqrt(Tr) * ((R)157893.94 / Tr + (R)2.) * exp(-expFact));
        const double HIOld = L.x(eH, 1);
        const double HeIOld = eHe >= 0 ? L.x(eHe, 1) : 0.0;
        const double HeIIOld = eHe >= 0 ? L.x(eHe, 2) : 0.0;
        const double H0 = L.x(eH, 1);
        __syncthreads();
        if (tid == 0) hhe_chain<R>(a, r, L, 1, Te, Ne, H, H0, collIonH);
        if (tid == 1 && eHe >= 0) hhe_chain<R>(a, r, L, 2, Te, Ne, H, H0, collIonH);
        heavy_out<R>(a, r, L, Ne, revFrozen, tid);
        __syncthreads();
        for (int e = tid; e < NE; e += 64)
            if (a.elemZ[e] >= 3) heavy_products<R>(a, L, e, Ne, H);
        __syncthreads();
        heavy_update<R>(a, L, nIterX, tid);
        __syncthreads();
        // free electrons: the stages of each element over the lanes, the
        // per-lane partial sums then reduced over the wave
        double ne = 0.;
        for (int e = 0; e < NE; ++e) {
            const int top = imin(a.elemZ[e] + 1, NS);
            const double ab = a.abund[e];
            for (int ion = 2 + tid; ion <= top; ion += 64)
                if (L.x(e, ion) >= 1.e-10) ne += (ion - 1) * ab * L.x(e, ion);
        }
        ne = wave_sum(ne);
        ne *= H;
        Ne = ne == 0. ? 1. : ne;
        const double dHI = (L.x(eH, 1) - HIOld) / HIOld;
        const double dHeI = eHe >= 0 ? (L.x(eHe, 1) - HeIOld) / HeIOld : 0.;
        const double dHeII = eHe >= 0 ? (L.x(eHe, 2) - HeIIOld) / HeIIOld : 0.;
        const bool big = fabs(dHI) > 0.01 || fabs(dHeI) > 0.01 || fabs(dHeII) > 0.01;
        if (big && nIterX < 15) { ++nIterX; continue; }
        if (big && nIterX == 15) conv = false;
        return conv;
    }
}


// ------------------------------------------------------------ grains (PE)
// Grain charge, photoelectric heating and gas-grain collisions of gas + dust
// runs (update_mod.f90:484-884), one grain (species, size) per lane inside
// each T-iteration.  Per grain the order of operations is the reference's;
// the gas heating and cooling sums over grains are taken in grain order.
constexpr double kAmu = 1.66053e-24;    // constants_mod.f90:11
constexpr double kMe = 9.109382e-28;    // constants_mod.f90:17
constexpr double kPiB = 3.141592654;     // constants_mod.f90:45
constexpr double kY0 = 0.5, kY1 = 0.2;  // updateCell (Baldwin et al. 1991)
__constant__ double c_aweight[30] = {1.0080, 4.0026, 6.941, 9.0122, 10.811, 12.0111, 14.0067,
                                     15.9994, 18.9984, 20.179, 22.9898, 24.305, 26.9815, 28.086,
                                     30.9738, 32.06, 35.453, 39.948, 39.102, 40.08, 44.956, 47.90,
                                     50.9414, 51.996, 54.9380, 55.847, 58.9332, 58.71, 63.546,
                                     65.37};

// LDS rows of the grain terms: the cell's photon flux J/(hcRyd nu) per bin,
// the ion fractions at the start of the update, and per grain the charge,
// the photoelectric / collision terms of the grains and their gas parts
struct GrainLds {
    double *q, *X0, *Vg, *pe, *coll, *pg, *cg;
};

__host__ __device__ inline size_t grain_lds_bytes(int B, int nX, int nG) {
    return ((size_t)B + (size_t)nX + 5 * (size_t)(nG > 0 ? nG : 1)) * sizeof(double);
}

// locate (interpolation_mod.f90:49-106) on the increasing mesh, 1-based
__device__ int locate_mesh(const double* xa, int n, double x) {
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (xa[m] > x) hi = m; else lo = m + 1;
    }
    return lo;
}

// getGrainEmission (:557-604)
__device__ double g_emission(const MocBalanceArgs& a, const GrainLds& G, int g, int s, int ai,
                             double Vg) {
    const int B = a.B;
    const double Vn = a.gVn[s];
    const double thres = Vn + Vg > Vn ? Vn + Vg : Vn;
    const double* cabs = a.gCabs + (size_t)g * B;
    const double r = a.gRadius[ai];
    double sum = 0.;
    for (int i = locate_mesh(a.gNu, B, thres) + 1; i <= B; ++i) {
        const double nu = a.gNu[i - 1];
        double Yn = kY0 * (1. - Vn / nu);
        if (Yn > kY1) Yn = kY1;
        double f = 1. - Vg / (nu - Vn);
        f = f < 0. ? 0. : (f > 1. ? 1. : f);
        const double Yhat = Yn * f;
        const double photFlux = G.q[i - 1] * (4. * kPiB);
        const double Qa = cabs[i - 1] / (1.e-8 * (r * r));
        sum = sum + Yhat * photFlux * Qa;
    }
    return sum;
}

// getGrainRecombination (:609-682): electrons less the ions of the state at
// the start of the update
template <class LT>
__device__ double g_recombination(const MocBalanceArgs& a, const LT& L, const GrainLds& G,
                                  double T, double Ne, double H, double Vg) {
    const double kT = 6.336e-6 * T, e8 = (1.1045e-15) * T / kPiB;
    double vmean = sqrt(e8 / kMe), eta = -Vg / kT;
    eta = eta <= 0. ? 1. - eta : exp(-eta);
    double rec = 0. + Ne * vmean * 1. * eta;
    for (int e = 0; e < a.NE; ++e) {
        const int z = a.elemZ[e];
        for (int istage = 2; istage <= imin(z + 1, a.NS); ++istage) {
            const double cpDen = G.X0[L.off[e] + istage] * a.abund[e] * H;
            const double mcp = c_aweight[z - 1] * kAmu;
            vmean = sqrt(e8 / mcp);
            double et = (double)(istage - 1) * Vg / kT;
            et = et <= 0. ? 1. - et : exp(-et);
            rec = rec - cpDen * vmean * 1. * et;
        }
    }
    return rec;
}

// setGrainPotential (:484-553): secant on emission = recombination from 0
// and 0.05 Ryd, steps capped at 5 % of the threshold, 100 passes at most;
// the slope starts at 0 for each grain
template <class LT>
__device__ double g_potential(const MocBalanceArgs& a, const LT& L, const GrainLds& G, int g,
                              int s, int ai, double T, double Ne, double H) {
    double Vg = 0., VgOld = 0., EmiOld = 0., RecOld = 0., slope = 0.;
    const double Vn = a.gVn[s];
    for (int n = 1; n <= 100; ++n) {
        if (n == 1) {
            Vg = 0.;
            VgOld = Vg;
            EmiOld = g_emission(a, G, g, s, ai, Vg);
            RecOld = g_recombination(a, L, G, T, Ne, H, Vg);
            Vg = VgOld + 0.05;
        }
        const double thr = Vn + Vg > Vn ? Vn + Vg : Vn;
        const double Emi = g_emission(a, G, g, s, ai, Vg);
        const double Rec = g_recombination(a, L, G, T, Ne, H, Vg);
        double delta = Emi - Rec;
        const double mx = Emi + Rec > 1.e-35 ? Emi + Rec : 1.e-35;
        double d1 = fabs(delta / (0.5 * mx));
        if (d1 < 0.005) return Vg;
        if (Vg != VgOld) {
            const double fac = (Emi - EmiOld) - (Rec - RecOld);
            if (fac != 0.) slope = fac / (Vg - VgOld);
        }
        VgOld = Vg;
        RecOld = Rec;
        EmiOld = Emi;
        d1 = -delta / slope;
        delta = fabs(d1);
        if (0.05 * thr < delta) delta = 0.05 * thr;
        delta = copysign(delta, d1);
        Vg = Vg + delta;
    }
    return Vg;
}

// setPhotoelHeatCool and setDustGasCollHeatCool (:686-884) of one grain:
// its photoelectric cooling and collision heating (into G.pe, G.coll) and
// its shares of the gas heating and cooling (G.pg, G.cg, before the common
// factors)
template <class LT>
__device__ void g_terms(const MocBalanceArgs& a, const LT& L, const GrainLds& G, int cell, int g,
                        int s, int ai, double T, double Ne, double H) {
    const int B = a.B, nS = a.nGS;
    const double Vn = a.gVn[s], V = G.Vg[g];
    const double* cabs = a.gCabs + (size_t)g * B;
    const double th = Vn + V > Vn ? Vn + V : Vn;
    int thP = locate_mesh(a.gNu, B, th);
    if (thP <= 0) thP = 1;
    double pe = 0., pg = 0.;
    for (int i = thP; i <= B; ++i) {
        const double nu = a.gNu[i - 1], dn = nu - Vn;
        double Yn = kY0 * (1. - Vn / nu);
        if (Yn > kY1) Yn = kY1;
        double f = 1. - V / dn;
        f = f < 0. ? 0. : (f > 1. ? 1. : f);
        const double Yhat = Yn * f;
        const double photFlux = G.q[i - 1];
        double m2 = (dn * dn - V * V) / dn;
        if (m2 < 0.) m2 = 0.;
        const double EY = Yn * 0.5 * (dn < m2 ? dn : m2);
        pe = pe + cabs[i - 1] * photFlux * EY * kHcRyd / kPiB;
        pg = pg + cabs[i - 1] * photFlux * (EY - Yhat * V) * a.gAbun[s] * a.gWeight[ai];
    }
    const double kT = 6.336e-6 * T, e8 = (1.1045e-15) * T / kPiB;
    const double r = a.gRadius[ai];
    const double Tg = a.Tdust[((size_t)cell * (a.nGA + 1) + ai + 1) * (nS + 1) + s + 1];
    double coll = 0., cg = 0., IPerg = 0.;
    for (int e = 0; e < a.NE; ++e) {
        const int z = a.elemZ[e];
        for (int istage = 1; istage <= imin(a.NS, z + 1); ++istage) {
            const double Z = (double)(istage - 1);
            const double mcp = c_aweight[z - 1] * kAmu;
            const double vmean = sqrt(e8 / mcp);
            if (istage > 1) IPerg = a.gIp[z * 32 + istage] * kRyd2erg;
            const double S = istage == 1 ? 2. * mcp * a.gMsurf[s] / ((mcp + a.gMsurf[s]) * (mcp + a.gMsurf[s])) : 1.;
            const double psi = Z * V / kT;
            double eta, xi;
            if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
            else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
            const double X = L.x(e, istage);
            coll = coll + X * a.abund[e] * H * kPiB * r * r * 1.e-8 * S * vmean *
                              (2. * kT * kRyd2erg * xi - eta * (Z * V * kRyd2erg - IPerg + 2. * kBoltz * Tg));
            cg = cg + X * a.gWeight[ai] * a.gAbun[s] * a.Ndust[cell] * a.abund[e] * kPiB * r * r * 1.e-8 * S *
                          vmean * (2. * kT * kRyd2erg * xi - eta * 2. * kBoltz * Tg);
        }
    }
    {  // electrons
        const double vmean = sqrt(e8 / kMe), S = 1., Z = -1.;
        const double psi = Z * V / kT;
        double eta, xi;
        if (psi <= 0.) { eta = 1. - psi; xi = 1. - psi / 2.; }
        else { eta = exp(-psi); xi = (1. + psi / 2.) * eta; }
        coll = coll + Ne * kPiB * r * r * 1.e-8 * S * vmean * (2. * kT * kRyd2erg * xi - eta * (Z * V * kRyd2erg));
        cg = cg + Ne * a.gWeight[ai] * a.gAbun[s] * a.Ndust[cell] * kPiB * r * r * 1.e-8 * S * vmean *
                      (2. * kT * kRyd2erg * xi) / H;
    }
    G.pe[g] = pe;
    G.pg[g] = pg;
    G.coll[g] = coll / (4. * kPiB);
    G.cg[g] = cg;
}

// forbidden-line cooling of the ions of list class c (nlevel_team_class),
// P = 64 / W ions at a time, one per team of W lanes; idle teams solve the
// group's first ion again and store nothing
template <int NL, class R>
__device__ void team_cooling(const MocBalanceArgs& a, const LdsT<R>& L, int cnt, const int* list, double T,
                             double Ne, int tid) {
    constexpr int W = NL <= 8 ? 8 : (NL <= 16 ? 16 : 32), P = 64 / W;
    for (int b0 = 0; b0 < cnt; b0 += P) {
        const int bt = b0 + tid / W;
        const int k = list[bt < cnt ? bt : b0];
        const int e = a.nlvSlot[k], ion = a.nlvIon[k];
        const double v = nlevel_team<NL, R>(a, a.nlvAtom[k], T, Ne, 0., false, a.abund[e] * L.x(e, ion), 1.9865e-16);
        if ((tid & (W - 1)) == 0 && bt < cnt) L.cool[k] = v;
    }
}

// thermBalance (:887-1215): heating and cooling [erg/s per n_H].  TEAMS: some
// ions have more than kNlevLane levels (their solves run in teams of lanes).
// The cooling fits and the collision rates of the level solves in precision
// R, every sum in FP64.
template <bool TEAMS, class R, class RT>
__device__ void therm(const MocBalanceArgs& a, const RT& r, const LdsT<R>& L, double T, double Ne,
                      double H, double heatef, const double comp[4], double* heat, double* cool,
                      int tid, double grainHeat = 0., double grainCool = 0., bool grains = false) {
    // forbidden-line cooling: ions of up to kNlevLane levels one per lane with
    // the matrix in registers, larger ones by teams of 8, 16 or 32 lanes
    for (int k = tid; k < a.nNlv; k += 64) {
        const int atom = a.nlvAtom[k];
        if (nlevel_count(a, atom) > kNlevLane) continue;
        const int e = a.nlvSlot[k], ion = a.nlvIon[k];
        L.cool[k] = nlevel_fixed<kNlevLane, R>(a, atom, T, Ne, 0., false, a.abund[e] * L.x(e, ion), 1.9865e-16);
    }
    if constexpr (TEAMS) {
        team_cooling<8, R>(a, L, a.nNlvT[0], a.nlvT[0], T, Ne, tid);
        team_cooling<10, R>(a, L, a.nNlvT[1], a.nlvT[1], T, Ne, tid);
        team_cooling<13, R>(a, L, a.nNlvT[2], a.nlvT[2], T, Ne, tid);
        team_cooling<16, R>(a, L, a.nNlvT[3], a.nlvT[3], T, Ne, tid);
        team_cooling<17, R>(a, L, a.nNlvT[4], a.nlvT[4], T, Ne, tid);
        team_cooling<27, R>(a, L, a.nNlvT[5], a.nlvT[5], T, Ne, tid);
    }
    __syncthreads();
    // The sums run over the lanes (elements, (element, stage) entries, n-level
    // ions, species) and are then reduced over the wave: every lane ends with
    // the same totals.
    const R Tr = (R)T;
    const R sq = (R)kBoltz * Tr / sqrt(Tr);
    double coolFF = 0., coolRec = 0., coolCollH = 0., heatSte = 0., coolNl = 0.;
    for (int e = tid; e < a.NE; e += 64) {
        const int z = a.elemZ[e];
        if (a.NS > z) {  // H-like recombination and free-free (Hummer 1994 fits)
            const double Np = L.x(e, z + 1) * a.abund[e];
            const R lt = log10(Tr / (R)(z * z));
            const R bFF = (R)z * ((R)1.0108464E-11 + (R)9.7930778E-13 * lt - (R)6.6433144E-13 * lt * lt +
                                  (R)2.4793747E-13 * lt * lt * lt - (R)2.3938215E-14 * lt * lt * lt * lt);
            const R bRec = (R)z * ((R)9.4255985E-11 - (R)4.04794384E-12 * lt - (R)1.0055237E-11 * lt * lt +
                                   (R)1.99266862E-12 * lt * lt * lt - (R)1.06681387E-13 * lt * lt * lt * lt);
            coolFF += Np * Ne * (double)(bFF * sq);
            coolRec += Np * Ne * (double)(bRec * sq);
        }
        if (z >= 2 && a.NS > z) {  // He-like (Hummer & Storey 1998 fits)
            const double Np = L.x(e, z) * a.abund[e];
            const R lt = log10((R)4. * Tr / (R)(z * z));
            const R bFF = ((R)z / (R)2.) * ((R)1.070073e-11 - (R)2.5730207e-13 * lt + (R)2.109134e-13 * lt * lt);
            const R bRec = ((R)z / (R)2.) * ((R)1.06926e-10 - (R)2.41756e-11 * lt + (R)1.14542e-12 * lt * lt -
                                             (R)5.06535e-13 * lt * lt * lt + (R)9.16745e-14 * lt * lt * lt * lt);
            coolFF += Np * Ne * (double)(bFF * sq);
            coolRec += Np * Ne * (double)(bRec * sq);
        }
    }
    for (int k = tid; k < a.NE * L.S2; k += 64) {  // collisional-ionisation cooling
        const int e = k / L.S2, ion = k % L.S2;
        const int z = a.elemZ[e];
        if (ion < 1 || ion > imin(z, a.NS - 1)) continue;
        const double c2 = a.cion_th[z * 32 + ion] * L.c1(e, ion) * kBoltz;
        coolCollH += a.abund[e] * L.x(e, ion) * c2 * Ne;
    }
    for (int k = tid; k < a.nNlv; k += 64) coolNl += L.cool[k];
    for (int sp = tid; sp < a.nSpec; sp += 64) {  // photoelectric heating of the gas
        const int e = a.specElem[sp], ion = a.specIon[sp];
        const int z = a.elemZ[e];
        if (ion > imin(z, a.NS - 1)) continue;
        double h = 0.;
        for (int sh = r.specShell[sp]; sh < r.specShell[sp + 1]; ++sh) h += r.heat(sh, z < 3 ? 1.0 : heatef);
        heatSte += h * L.x(e, ion) * a.abund[e];
    }
    coolFF = wave_sum(coolFF);
    coolRec = wave_sum(coolRec);
    coolCollH = wave_sum(coolCollH);
    coolNl = wave_sum(coolNl);
    heatSte = wave_sum(heatSte);
    // H Ly-alpha/beta excitation at the Te4 the reference's He-like loop leaves
    // behind: that of the last element with a He-like stage
    R Te4 = Tr / (R)1.e4;
    for (int e = 0; e < a.NE; ++e) {
        const int z = a.elemZ[e];
        if (z >= 2 && a.NS > z) Te4 = (R)4. * Tr / ((R)(z * z) * (R)1.e4);
    }
    const int eH = a.elemSlot[1];
    const double coolColl =
        (T > 5000. ? (double)((R)2.47e-8 * exp((R)-118338. / Tr) * pow(Te4, (R)-0.228) +
                              (R)1.32e-8 * exp((R)-140252. / Tr) * pow(Te4, (R)-0.460)) *
                         kHcRyd * L.x(eH, 1) * Ne
                   : 0.) +
        coolNl;
    *heat = heatSte;
    *cool = coolFF + coolRec + coolColl + coolCollH;
    if (a.lgCompton) {
        const double cHeat = (6.65e-25 * comp[0] * kRyd2erg / 8.184e-7) / H;
        const double cCool = (6.65e-25 * 6.75e-10 * T * comp[1]) / H;
        *heat += cHeat * Ne + comp[2];
        *cool += cCool * Ne;
    }
    if (grains) {  // thermBalance :1205-1208
        *cool = *cool + grainCool;
        *heat = *heat + grainHeat;
    }
    __syncthreads();  // L.cool is rewritten by the next call
}

// Four waves per SIMD: the per-cell work is chains of dependent FP64 steps,
// so latency hiding comes from other cells' waves.  At the compiler's own
// allocation (256 VGPRs, one wave per SIMD) the nebula32 balance took 20 ms;
// capped at 128 VGPRs with a few hundred bytes of spills it takes 13.6 ms.
// PE: gas + dust runs with grain charge, photoelectric heating and
// gas-grain collisions (the rows of GrainLds follow the other LDS rows).
template <bool TEAMS, bool ST, bool PE, class R>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_balance(MocBalanceArgs a) {
    extern __shared__ double lds[];
    const int cell = a.cell0 + blockIdx.x;
    const int tid = threadIdx.x;
    const int C1 = a.nCells + 1;
    const int NE = a.NE, NS = a.NS, S2 = NS + 2;
    Rates<ST> r{a.R, a.ncol, (int)blockIdx.x, a.specOf, a.specShell, a.shInfo, a.shAug};
    if constexpr (ST) {
        // stage the cell's rate columns and the index tables after the state rows
        double* Rl = lds + balance_lds_bytes(a.nX, a.nNlv, sizeof(R)) / sizeof(double);
        int* iL = reinterpret_cast<int*>(Rl + a.ncol);
        int* sOf = iL;
        int* sSh = sOf + NE * 32;
        int* sIn = sSh + a.nSpec + 1;
        int* sAu = sIn + a.nShells;
        for (int k = tid; k < a.ncol; k += 64) Rl[k] = a.R[(size_t)blockIdx.x * a.ncol + k];
        for (int k = tid; k < NE * 32; k += 64) sOf[k] = a.specOf[k];
        for (int k = tid; k <= a.nSpec; k += 64) sSh[k] = a.specShell[k];
        for (int k = tid; k < a.nShells; k += 64) { sIn[k] = a.shInfo[k]; sAu[k] = a.shAug[k]; }
        __syncthreads();
        r.R = Rl;
        r.specOf = sOf;
        r.specShell = sSh;
        r.shInfo = sIn;
        r.shAug = sAu;
    }
    if (!(r.at(a.colHit) > 0.)) {  // no packet reached the cell (lgHit, :127-171)
        if (tid == 0) {
            a.lgBlack[cell] = 1;
            atomicAdd(&a.counters[0], 1ull);
            a.nIterT[cell] = 0;
        }
        return;
    }
    LdsT<R> L;
    L.S2 = S2;
    L.X = lds;
    L.alpha = reinterpret_cast<R*>(L.X + a.nX);
    L.cI1 = L.alpha + a.nX;
    L.out = reinterpret_cast<double*>(L.cI1 + a.nX);
    L.cool = L.out + a.nX;
    L.diO = L.cool + (a.nNlv > 0 ? a.nNlv : 1);
    L.hot = (int*)(L.diO + MOC_MAXS + 1);
    L.off = L.hot + 1;
    if (tid <= NE) L.off[tid] = a.xoff[tid];
    __syncthreads();
    for (int k = tid; k < NE * S2; k += 64) {
        const int e = k / S2, ion = k % S2;
        if (ion < L.len(e))
            L.x(e, ion) = (ion >= 1 && ion <= NS) ? a.ionDen[((size_t)(ion - 1) * NE + e) * C1 + cell] : 0.0;
    }
    const int nG = PE ? a.nGS * a.nGA : 0;
    GrainLds G{};
    if constexpr (PE) {
        size_t off = balance_lds_bytes(a.nX, a.nNlv, sizeof(R));
        if constexpr (ST) off += balance_stage_bytes(a);
        G.q = lds + off / sizeof(double);
        G.X0 = G.q + a.B;
        G.Vg = G.X0 + a.nX;
        G.pe = G.Vg + nG;
        G.coll = G.pe + nG;
        G.pg = G.coll + nG;
        G.cg = G.pg + nG;
        for (int b = tid; b < a.B; b += 64) G.q[b] = a.J[(size_t)b * C1 + cell] / (kHcRyd * a.gNu[b]);
        __syncthreads();  // X was loaded by other lanes
        for (int k = tid; k < a.nX; k += 64) G.X0[k] = L.X[k];
    }
    __syncthreads();
    double Te = a.Te[cell], Ne = a.Ne[cell];
    const double H = a.Hden[cell];
    const int eH = a.elemSlot[1], eHe = a.elemSlot[2];
    const double XOld = L.x(eH, 1);
    double Thigh = 0., Tlow = 0., rHigh = 0., rLow = 0.;
    int nT = 0;
    bool ib = true, noTe = false;
    for (;;) {
        ++nT;
        recomb<R>(a, L, Te, tid);
        // secondary-ionisation heating efficiency (photoionisation :3605-3627)
        double xe = 0., heatef;
        for (int e = 0; e < NE; ++e) xe += L.x(e, 1) * a.abund[e] * H;
        xe = Ne / xe;
        if (xe > 0.95) heatef = 1.;
        else {
            xe = xe < 1.e-4 ? 1.e-4 : xe;
            heatef = (double)((R)0.9971 * ((R)1. - pow((R)1. - pow((R)xe, (R)0.2663), (R)1.3163)));
        }
        collion<R>(a, L, Te, tid);
        double comp[4] = {0., 0., 0., 0.};
        if (a.lgCompton) {
            comp[0] = r.at(a.colComp + 0);
            comp[1] = r.at(a.colComp + 1);
            comp[2] = (r.at(a.colComp + 2) * L.x(eH, 1) * a.abund[eH] +
                       (eHe >= 0 ? r.at(a.colComp + 3) * L.x(eHe, 1) * a.abund[eHe] * 2. : 0.)) * kRyd2erg;
        }
        double revFrozen[2] = {0., 0.};
        if (a.elemSlot[7] >= 0) revFrozen[0] = rev_chex<R>(a, L, Te, H, 7, 1);
        if (a.elemSlot[8] >= 0) revFrozen[1] = rev_chex<R>(a, L, Te, H, 8, 1);
        ib = ion_balance<R>(a, r, L, Te, Ne, H, revFrozen, tid);
        double heat, cool;
        if constexpr (PE) {  // grain charges and their heating/cooling terms (:318-333)
            for (int g = tid; g < nG; g += 64) {
                const int s = g / a.nGA, ai = g % a.nGA;
                G.Vg[g] = g_potential(a, L, G, g, s, ai, Te, Ne, H);
                g_terms(a, L, G, cell, g, s, ai, Te, Ne, H);
            }
            __syncthreads();
            double pg = 0., cg = 0.;
            for (int g = 0; g < nG; ++g) {
                pg = pg + G.pg[g];
                cg = cg + G.cg[g];
            }
            const double peG = pg * a.Ndust[cell] * kHcRyd / H, collG = cg / (4. * kPiB);
            therm<TEAMS, R>(a, r, L, Te, Ne, H, heatef, comp, &heat, &cool, tid, peG, collG, true);
        } else {
            therm<TEAMS, R>(a, r, L, Te, Ne, H, heatef, comp, &heat, &cool, tid);
        }
        const double res = heat - cool;
        if (fabs(res) >= 0.02 * (fabs(cool) + fabs(heat)) / 2.) {
            if (nT < 20) {
                if (res < 0.) {
                    rLow = res;
                    Tlow = Te;
                    Te = Thigh != 0. ? Tlow - (Thigh - Tlow) * rLow / (rHigh - rLow) : Te / 1.2;
                } else {
                    rHigh = res;
                    Thigh = Te;
                    Te = Tlow != 0. ? Tlow - (Thigh - Tlow) * rLow / (rHigh - rLow) : Te * 1.2;
                }
                if (nT >= 6 && fabs(Thigh - Tlow) <= 0.002 * (Thigh + Tlow)) {
                    if (res < 0.) { Te = Thigh - 100.; Thigh = 0.; }
                    else { Te = Tlow + 100.; Tlow = 0.; }
                }
                if (Te <= 0.) Te = 1.;
                continue;
            }
            Te = Thigh == 0. ? Tlow : (Tlow == 0. ? Thigh : (Thigh + Tlow) * 0.5);
            noTe = true;
        }
        break;
    }
    for (int k = tid; k < NE * S2; k += 64) {
        const int e = k / S2, ion = k % S2;
        if (ion >= 1 && ion <= NS && ion < L.len(e))
            a.ionDen[((size_t)(ion - 1) * NE + e) * C1 + cell] = L.x(e, ion);
    }
    if constexpr (PE) {  // the grain terms of the last T iteration, for getDustT
        for (int g = tid; g < nG; g += 64) {
            a.collD[(size_t)g * C1 + cell] = G.coll[g];
            a.peD[(size_t)g * C1 + cell] = G.pe[g];
        }
    }
    if (tid == 0) {
        if (noTe) atomicAdd(&a.counters[2], 1ull);
        if (!ib) atomicAdd(&a.counters[1], 1ull);
        a.lgBlack[cell] = 0;
        a.Te[cell] = Te;
        a.Ne[cell] = Ne;
        a.lgConverged[cell] = fabs((L.x(eH, 1) - XOld) / XOld) <= a.XHILimit ? 1 : 0;
        a.nIterT[cell] = nT;
    }
}

// the team solves are compiled in only where some ion has more than
// kNlevLane levels
template <bool PE, class R>
void launch_balance_r(moc_ctx* ctx, const MocBalanceArgs& a, bool st, dim3 grid, dim3 block, size_t lds) {
    if (a.maxLev <= kNlevLane) {
        if (st) hipLaunchKernelGGL((k_balance<false, true, PE, R>), grid, block, lds, ctx->stream, a);
        else hipLaunchKernelGGL((k_balance<false, false, PE, R>), grid, block, lds, ctx->stream, a);
    } else {
        if (st) hipLaunchKernelGGL((k_balance<true, true, PE, R>), grid, block, lds, ctx->stream, a);
        else hipLaunchKernelGGL((k_balance<true, false, PE, R>), grid, block, lds, ctx->stream, a);
    }
}

// the team solves are compiled in only where some ion has more than
// kNlevLane levels; the fits in FP32 for the mixed-precision form
template <bool PE>
void launch_balance(moc_ctx* ctx, const MocBalanceArgs& a, bool st, dim3 grid, dim3 block, size_t lds) {
    if (a.fitF32) launch_balance_r<PE, float>(ctx, a, st, grid, block, lds);
    else launch_balance_r<PE, double>(ctx, a, st, grid, block, lds);
}

}  // namespace

void launch_rates(moc_ctx* ctx, const double* J, int ldJ, int c0, int nc, const double* W, int ncol,
                  double* R, const int* perm = nullptr, const int* bLo = nullptr, const int* bHi = nullptr) {
    dim3 g1((unsigned)((nc + RC - 1) / RC), (unsigned)((ncol + RN - 1) / RN));
    hipLaunchKernelGGL(k_rates, g1, dim3(256), 0, ctx->stream, J, ldJ, c0, nc, ctx->nbins, W, ncol, R,
                       perm, bLo, bHi);
}

void moc_launch_rates(moc_ctx* ctx, const double* J, int C1, const double* W, int ncol, double* R) {
    launch_rates(ctx, J, C1, 0, C1, W, ncol, R);
}

int moc_rate_tile_cols() { return RN; }
int moc_rate_tile_bins() { return RK; }

int moc_launch_balance(moc_ctx* ctx, const MocBalanceArgs& a0, const double* J, const double* W,
                       int ncol, double* R, int chunk) {
    MocBalanceArgs a = a0;
    const int C1 = a.nCells + 1;
    const size_t base = balance_lds_bytes(a.nX, a.nNlv, a.fitF32 ? sizeof(float) : sizeof(double));
    const size_t staged = base + balance_stage_bytes(a);
    const bool st = staged <= kMaxBalanceLds;
    const bool pe = a.nGS > 0 && a.nGA > 0;
    const size_t lds = (st ? staged : base) + (pe ? grain_lds_bytes(a.B, a.nX, a.nGS * a.nGA) : 0);
    if (lds > kMaxGrainLds) {
        ctx->err = "balance: a cell needs " + std::to_string(lds) + " bytes of LDS (limit " +
                   std::to_string(kMaxGrainLds) + ")";
        return MOC_ERR_ARG;
    }
    if (chunk < 1) { ctx->err = "balance: no rate buffer"; return MOC_ERR_STATE; }
    a.R = R;
    for (int c0 = 1; c0 <= a.nCells; c0 += chunk) {
        const int n = a.nCells - c0 + 1 < chunk ? a.nCells - c0 + 1 : chunk;
        launch_rates(ctx, J, C1, c0, n, W, ncol, R, a.ratePerm, a.rateLo, a.rateHi);
        a.cell0 = c0;
        a.nLaunch = n;
        const dim3 grid((unsigned)n), block(64);
        if (pe) launch_balance<true>(ctx, a, st, grid, block, lds);
        else launch_balance<false>(ctx, a, st, grid, block, lds);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("balance launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    return MOC_OK;
}

// n-level statistical equilibrium of one ion (equilibrium, emission_mod.f90:
// 2505-3028, with luSlv :3031-3106), shared by the thermal balance (cooling,
// rec = false) and the diffuse emission (line luminosity, rec = true).
#pragma once
#include <hip/hip_runtime.h>

#include "moc_internal.h"

namespace {

// Level count of n-level ion `atom` as solved: its file's levels, cut at maxLev.
__device__ __forceinline__ int nlevel_count(const MocBalanceArgs& a, int atom) {
    const int nf = (int)a.atomMeta[8 * atom];
    return nf < a.maxLev ? nf : a.maxLev;
}

// Doubles of LDS workspace nlevel_wave needs for ions of up to n levels.
__host__ __device__ inline int nlevel_wave_doubles(int n) { return n > 0 ? 2 * n * n + 2 * n : 0; }

// Sum of the line emissivities A_ji dE_ij n_j * scale * unit of ion `atom`
// that exceed 1e-35 (equilibrium, emission_mod.f90:2505-3028; luSlv
// :3031-3106), solved by one whole wave with the level matrix in LDS: for
// the ions whose level count is too large for a per-lane register matrix.
// Every lane of the (single-wave) workgroup calls it with the same
// arguments; the return value is the same on every lane.  ws holds
// nlevel_wave_doubles(n) doubles of LDS.
//   rec      levels 2.. are fed by recombination from the next stage
//            (ionDenUp = X(i+1)/X(i)) at the rates of the ion's data file
//            (:2828-2877; kind 1: power-law radiative plus dielectronic fit per
//            level, kind 2: cubic in log T);
//   lineAcc  with it, every line's emissivity (lower i, upper j: fLineEm(i, j),
//            :2987-3005) times w is added to lineAcc[i * ld + j] (one lane
//            per line).
// Upsilon(T) is the natural cubic spline in log T with second derivatives
// computed on the host, clamped at the table ends.  Each matrix entry sees
// the operations of a sequential evaluation in the same order (row sums and
// the back substitution run on one lane; each elimination step updates the
// rows below the pivot independently), so the result is the sequential one
// bit for bit.
__device__ double nlevel_wave(const MocBalanceArgs& a, int atom, double Te, double Ne,
                              double ionDenUp, bool rec, double scale, double unit, double* ws,
                              int lane, double* lineAcc = nullptr, double w = 0., int ld = 0) {
    const int64_t* m = a.atomMeta + 8 * atom;
    const int nf = (int)m[0], nT = (int)m[1];
    const double* lt = a.atomLogT + m[2];
    const double* g = a.atomG + m[3];
    const double* en = a.atomE + m[3];
    const double* A = a.atomA + m[4];
    const double* q = a.atomQom + m[5];
    const double* q2 = a.atomQ2 + m[5];
    const int iup2p = (int)m[6];
    const double atp = a.atomAtp[atom];
    const int n = nf < a.maxLev ? nf : a.maxLev;
    double* Q = ws;              // [n][n] collision rates qe, later the line emissivities
    double* X = Q + n * n;       // [n][n] level matrix
    double* y = X + n * n;       // [n]
    double* f = y + n;           // [n] elimination factors of one step
    const double sqrTe = sqrt(Te), lT = log10(Te);
    int klo = 0, khi = nT - 1;
    double sa = 0., sb = 0., h = 1.;
    const bool below = lT <= lt[0], above = lT >= lt[nT - 1];
    if (nT > 2 && !below && !above) {
        while (khi - klo > 1) {
            const int k = (khi + klo) >> 1;
            if (lt[k] > lT) khi = k; else klo = k;
        }
        h = lt[khi] - lt[klo];
        sa = (lt[khi] - lT) / h;
        sb = (lT - lt[klo]) / h;
    }
    __syncthreads();  // the workspace may still be read by the previous call
    // collision rates of every pair i < j, one pair per lane
    const int np = n * (n - 1) / 2;
    for (int p = lane; p < np; p += 64) {
        int i = 0, r = p;
        while (r >= n - 1 - i) { r -= n - 1 - i; ++i; }
        const int j = i + 1 + r;
        const size_t pp = (size_t)i * nf + j;
        double ups;
        if (nT == 1) ups = q[pp];
        else if (nT == 2) ups = q[pp] + (q[(size_t)nf * nf + pp] - q[pp]) / (lt[1] - lt[0]) * (lT - lt[0]);
        else if (below) ups = q[pp];
        else if (above) ups = q[(size_t)(nT - 1) * nf * nf + pp];
        else {
            const double ylo = q[(size_t)klo * nf * nf + pp], yhi = q[(size_t)khi * nf * nf + pp];
            const double dlo = q2[(size_t)klo * nf * nf + pp], dhi = q2[(size_t)khi * nf * nf + pp];
            ups = sa * ylo + sb * yhi + ((sa * sa * sa - sa) * dlo + (sb * sb * sb - sb) * dhi) * (h * h) / 6.;
        }
        const double c = 8.63e-6 * ups / sqrTe;
        if (en[i] <= en[j]) {
            Q[i * n + j] = c * exp((en[i] - en[j]) * 1.4388463 / Te) / g[i];
            Q[j * n + i] = c / g[j];
        } else {
            Q[j * n + i] = c * exp((en[j] - en[i]) * 1.4388463 / Te) / g[i];
            Q[i * n + j] = c / g[j];
        }
    }
    __syncthreads();
    // the rows of the matrix, one row per lane: row 0 the normalisation,
    // the diagonal summed in the sequential order
    const int kind = rec ? a.atomRecKind[atom] : 0;
    for (int i = lane; i < n; i += 64) {
        if (i == 0) {
            for (int j = 0; j < n; ++j) X[j] = 1.;
            y[0] = 1.;
            continue;
        }
        double d = 0.;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double v = 0.;
            v += Ne * Q[j * n + i];
            d -= Ne * Q[i * n + j];
            if (j > i) v += A[(size_t)j * nf + i];
            else d -= A[(size_t)i * nf + j];
            X[i * n + j] = v;
        }
        X[i * n + i] = d;
        double yi = 0.;
        if (kind != 0) {
            const int* rv = a.atomRecValid + m[3];
            double al = 0.;
            if (rv[i]) {
                const double* r = a.atomRec + 11 * m[3] + 11 * i;
                const double t4 = Te * 1.e-4;
                if (kind == 1) {
                    const double z = r[1], tz = t4 / (z * z);
                    al = 1.e-13 * r[0] * z * r[2] * pow(tz, r[3]) / (1. + r[4] * pow(tz, r[5])) +
                         1.0e-12 * r[0] * (r[6] / t4 + r[7] + r[8] * t4 + r[9] * t4 * t4) * pow(t4, -1.5) *
                             exp(-r[10] * 1.e4 / Te);
                } else {
                    al = pow(10., r[0] + r[1] * lT + r[2] * lT * lT + r[3] * lT * lT * lT);
                }
            }
            yi = -Ne * ionDenUp * al;
        }
        y[i] = yi;
    }
    __syncthreads();
    // elimination without pivoting, as luSlv: step i updates rows k > i
    for (int i = 0; i < n - 1; ++i) {
        for (int k = i + 1 + lane; k < n; k += 64) f[k] = X[k * n + i] / X[i * n + i];
        __syncthreads();
        const int mm = n - 1 - i;
        for (int t = lane; t < mm * mm; t += 64) {
            const int k = i + 1 + t / mm, j = i + 1 + t % mm;
            X[k * n + j] -= X[i * n + j] * f[k];
        }
        for (int k = i + 1 + lane; k < n; k += 64) y[k] -= y[i] * f[k];
        __syncthreads();
    }
    if (lane == 0) {
        y[n - 1] /= X[(n - 1) * n + n - 1];
        for (int k = n - 2; k >= 0; --k) {
            double v = y[k];
            for (int j = k + 1; j < n; ++j) v -= y[j] * X[k * n + j];
            y[k] = v / X[k * n + k];
        }
    }
    __syncthreads();
    double sumN = 0.;
    for (int i = 0; i < n; ++i) sumN += y[i];
    // line emissivities, one line per lane, into Q (i < j)
    for (int p = lane; p < np; p += 64) {
        int i = 0, r = p;
        while (r >= n - 1 - i) { r -= n - 1 - i; ++i; }
        const int j = i + 1 + r;
        double av = A[(size_t)j * nf + i];
        double em = 0.;
        if (av != 0.) {
            if (j + 1 == iup2p && i == 0 && atp > 0.) av -= atp;
            em = av * (en[j] - en[i]) * (y[j] / sumN) * scale * unit;
            if (lineAcc) lineAcc[i * ld + j] += em * w;
        }
        Q[i * n + j] = em;
    }
    __syncthreads();
    double tot = 0.;
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double em = Q[i * n + j];
            if (em > 1.e-35) tot += em;
        }
    return tot;
}

}  // namespace

namespace {

// The same sum with the level matrix held in registers: NL is a compile-time
// bound on the level count (n <= NL), every loop is unrolled over NL and
// predicated on n, and the collision rates of a pair are recomputed where a
// row needs them instead of being stored.  The arithmetic on the valid
// entries is the generic routine's, operation for operation, including the
// recombination feed when rec is set.
template <int NL>
__device__ double nlevel_fixed(const MocBalanceArgs& a, int atom, double Te, double Ne,
                               double ionDenUp, bool rec, double scale, double unit,
                               double* lineAcc = nullptr, double w = 0., int ld = 0) {
    const int64_t* m = a.atomMeta + 8 * atom;
    const int nf = (int)m[0], nT = (int)m[1];
    const double* lt = a.atomLogT + m[2];
    const double* g = a.atomG + m[3];
    const double* en = a.atomE + m[3];
    const double* A = a.atomA + m[4];
    const double* q = a.atomQom + m[5];
    const double* q2 = a.atomQ2 + m[5];
    const int iup2p = (int)m[6];
    const double atp = a.atomAtp[atom];
    const int n = nf < a.maxLev ? nf : a.maxLev;
    const double sqrTe = sqrt(Te), lT = log10(Te);
    int klo = 0, khi = nT - 1;
    double sa = 0., sb = 0., h = 1.;
    const bool below = lT <= lt[0], above = lT >= lt[nT - 1];
    if (nT > 2 && !below && !above) {
        while (khi - klo > 1) {
            const int k = (khi + klo) >> 1;
            if (lt[k] > lT) khi = k; else klo = k;
        }
        h = lt[khi] - lt[klo];
        sa = (lt[khi] - lT) / h;
        sb = (lT - lt[klo]) / h;
    }
    // effective rate coefficient qe[p][q] of the generic routine
    auto qef = [&](int p, int r) -> double {
        const int i = p < r ? p : r, j = p < r ? r : p;
        const size_t pp = (size_t)i * nf + j;
        double ups;
        if (nT == 1) ups = q[pp];
        else if (nT == 2) ups = q[pp] + (q[(size_t)nf * nf + pp] - q[pp]) / (lt[1] - lt[0]) * (lT - lt[0]);
        else if (below) ups = q[pp];
        else if (above) ups = q[(size_t)(nT - 1) * nf * nf + pp];
        else {
            const double ylo = q[(size_t)klo * nf * nf + pp], yhi = q[(size_t)khi * nf * nf + pp];
            const double dlo = q2[(size_t)klo * nf * nf + pp], dhi = q2[(size_t)khi * nf * nf + pp];
            ups = sa * ylo + sb * yhi + ((sa * sa * sa - sa) * dlo + (sb * sb * sb - sb) * dhi) * (h * h) / 6.;
        }
        const double c = 8.63e-6 * ups / sqrTe;
        if (en[i] <= en[j]) return p == i ? c * exp((en[i] - en[j]) * 1.4388463 / Te) / g[i] : c / g[j];
        return p == i ? c / g[j] : c * exp((en[j] - en[i]) * 1.4388463 / Te) / g[i];
    };
    double x[NL][NL], y[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        y[i] = 0.;
#pragma unroll
        for (int j = 0; j < NL; ++j) x[i][j] = 0.;
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) x[0][j] = 1.;
    y[0] = 1.;
#pragma unroll
    for (int i = 1; i < NL; ++i) {
        if (i >= n) break;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            if (j >= n) break;
            if (j == i) continue;
            x[i][j] += Ne * qef(j, i);
            x[i][i] -= Ne * qef(i, j);
            if (j > i) x[i][j] += A[(size_t)j * nf + i];
            else x[i][i] -= A[(size_t)i * nf + j];
        }
    }
    const int kind = rec ? a.atomRecKind[atom] : 0;
    if (kind != 0) {
        const double* rd = a.atomRec + 11 * m[3];
        const int* rv = a.atomRecValid + m[3];
        const double t4 = Te * 1.e-4;
#pragma unroll
        for (int i = 1; i < NL; ++i) {
            if (i >= n) break;
            double al = 0.;
            if (rv[i]) {
                const double* r = rd + 11 * i;
                if (kind == 1) {
                    const double z = r[1], tz = t4 / (z * z);
                    al = 1.e-13 * r[0] * z * r[2] * pow(tz, r[3]) / (1. + r[4] * pow(tz, r[5])) +
                         1.0e-12 * r[0] * (r[6] / t4 + r[7] + r[8] * t4 + r[9] * t4 * t4) * pow(t4, -1.5) *
                             exp(-r[10] * 1.e4 / Te);
                } else {
                    al = pow(10., r[0] + r[1] * lT + r[2] * lT * lT + r[3] * lT * lT * lT);
                }
            }
            y[i] = -Ne * ionDenUp * al;
        }
    }
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        if (i >= n - 1) break;
#pragma unroll
        for (int k = i + 1; k < NL; ++k) {
            if (k >= n) break;
            const double f = x[k][i] / x[i][i];
#pragma unroll
            for (int j = i + 1; j < NL; ++j) {
                if (j >= n) break;
                x[k][j] -= x[i][j] * f;
            }
            y[k] -= y[i] * f;
        }
    }
#pragma unroll
    for (int k = NL - 1; k >= 0; --k) {
        if (k >= n) continue;
        if (k == n - 1) {
            y[k] /= x[k][k];
            continue;
        }
#pragma unroll
        for (int j = k + 1; j < NL; ++j) {
            if (j >= n) break;
            y[k] -= y[j] * x[k][j];
        }
        y[k] /= x[k][k];
    }
    double sumN = 0.;
#pragma unroll
    for (int i = 0; i < NL; ++i)
        if (i < n) sumN += y[i];
    double tot = 0.;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        if (i >= n - 1) break;
#pragma unroll
        for (int j = i + 1; j < NL; ++j) {
            if (j >= n) break;
            double av = A[(size_t)j * nf + i];
            if (av == 0.) continue;
            if (j + 1 == iup2p && i == 0 && atp > 0.) av -= atp;
            const double em = av * (en[j] - en[i]) * (y[j] / sumN) * scale * unit;
            if (em > 1.e-35) tot += em;
            if (lineAcc) lineAcc[i * ld + j] += em * w;
        }
    }
    return tot;
}

}  // namespace

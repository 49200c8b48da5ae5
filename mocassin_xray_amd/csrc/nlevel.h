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

// Sum of the line emissivities A_ji dE_ij n_j * scale * unit of ion `atom`
// that exceed 1e-35 (equilibrium, emission_mod.f90:2505-3028; luSlv
// :3031-3106), solved by a team of W = 16 (NL = 15) or W = 32 (NL = 27)
// lanes of one wave with the level matrix in registers: team lane c < n owns
// column c of the matrix, the team's last lane owns the right-hand side (as
// one more column, always to the right of the pivot), and the pivot column's
// factors, the back-substitution entries and the emissivities travel between
// lanes by shuffles.  The ion's level count n must not exceed NL.  Every lane
// of the team calls it with the same ion; the return value is the same on
// every lane of the team.  The loops are unrolled over NL and the entries
// past n are carried but never read by the entries below n, whose operations
// are those of a sequential evaluation in the same order (each column built
// from row c of the collision rates, each elimination step x[k][j] -= x[i][j]
// (x[k][i] / x[i][i]), the back substitution row by row on the right-hand
// side's lane, the emissivities summed in (lower, upper) order), so the result
// is the sequential one bit for bit.
//   rec      levels 2.. are fed by recombination from the next stage
//            (ionDenUp = X(i+1)/X(i)) at the rates of the ion's data file
//            (:2828-2877; kind 1: power-law radiative plus dielectronic fit per
//            level, kind 2: cubic in log T);
//   lineAcc  with it, every line's emissivity (lower i, upper j: fLineEm(i, j),
//            :2987-3005) times w is added to lineAcc[i * ld + j] by the lane of
//            column j.
template <int NL>
__device__ __attribute__((noinline)) double nlevel_team(const MocBalanceArgs& a, int atom, double Te, double Ne,
                              double ionDenUp, bool rec, double scale, double unit,
                              double* lineAcc = nullptr, double w = 0., int ld = 0) {
    constexpr int W = NL < 16 ? 16 : 32;
    static_assert(NL < W, "the right-hand side needs a lane of its own");
    const int c = (int)(threadIdx.x & (W - 1));  // column of this lane
    constexpr int RHS = W - 1;
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
    // effective collision rate qe[p][r] (p != r) of the sequential routine
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
        const double cc = 8.63e-6 * ups / sqrTe;
        if (en[i] <= en[j]) return p == i ? cc * exp((en[i] - en[j]) * 1.4388463 / Te) / g[i] : cc / g[j];
        return p == i ? cc / g[j] : cc * exp((en[j] - en[i]) * 1.4388463 / Te) / g[i];
    };
    // column c: row 0 the normalisation, x[i][c] = Ne qe[c][i] (+ A[c][i] above
    // the diagonal), the diagonal minus the sum of row c's rates out of level c
    double x[NL];
    double d = 0.;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        x[k] = 0.;
        if (c < n && k < n && k != c) {
            const double qk = qef(c, k);
            if (k >= 1) {
                x[k] += Ne * qk;
                if (c > k) x[k] += A[(size_t)c * nf + k];
            }
            d -= Ne * qk;
            if (k < c) d -= A[(size_t)c * nf + k];
        }
    }
    x[0] = 1.;
#pragma unroll
    for (int k = 1; k < NL; ++k)
        if (k == c) x[k] = d;
    // the right-hand side: 1 in row 0, the recombination feed of each level
    // (computed on the level's own lane) below
    double feed = 0.;
    const int kind = rec ? a.atomRecKind[atom] : 0;
    if (kind != 0 && c >= 1 && c < n) {
        const int* rv = a.atomRecValid + m[3];
        double al = 0.;
        if (rv[c]) {
            const double* r = a.atomRec + 11 * m[3] + 11 * c;
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
        feed = -Ne * ionDenUp * al;
    }
#pragma unroll
    for (int k = 1; k < NL; ++k) {
        const double v = __shfl(feed, k, W);
        if (c == RHS) x[k] = v;
    }
    // elimination without pivoting, as luSlv: the factors of step i come from
    // column i; columns right of i (and the right-hand side) are updated
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
#pragma unroll
        for (int k = i + 1; k < NL; ++k) {
            const double f = __shfl(x[k] / x[i], i, W);
            if (c > i) x[k] -= x[i] * f;
        }
    }
    // back substitution on the right-hand side's lane (its x[] is y)
#pragma unroll
    for (int k = NL - 1; k >= 0; --k) {
        double v = x[k];
#pragma unroll
        for (int j = k + 1; j < NL; ++j) {
            const double xkj = __shfl(x[k], j, W);
            if (j < n) v -= x[j] * xkj;
        }
        const double xkk = __shfl(x[k], k, W);
        if (c == RHS && k < n) x[k] = v / xkk;
    }
    // populations: the sum on the right-hand side's lane, each level's own on its lane
    double sumN = 0.;
#pragma unroll
    for (int i = 0; i < NL; ++i)
        if (i < n) sumN += x[i];
    sumN = __shfl(sumN, RHS, W);
    double yc = 0.;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const double v = __shfl(x[i], RHS, W);
        if (i == c) yc = v;
    }
    // emissivities of the lines out of level c (upper), into x[lower]
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        double em = 0.;
        if (c < n && i < c) {
            double av = A[(size_t)c * nf + i];
            if (av != 0.) {
                if (c + 1 == iup2p && i == 0 && atp > 0.) av -= atp;
                em = av * (en[c] - en[i]) * (yc / sumN) * scale * unit;
                if (lineAcc) lineAcc[i * ld + c] += em * w;
            }
        }
        x[i] = em;
    }
    double tot = 0.;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i)
#pragma unroll
        for (int j = i + 1; j < NL; ++j) {
            const double em = __shfl(x[i], j, W);
            if (j < n && em > 1.e-35) tot += em;
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

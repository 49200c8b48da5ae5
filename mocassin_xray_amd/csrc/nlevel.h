// n-level statistical equilibrium of one ion (equilibrium, emission_mod.f90:
// 2505-3028, with luSlv :3031-3106), shared by the thermal balance (cooling,
// rec = false) and the diffuse emission (line luminosity, rec = true).
//
// The collision-rate fits (the Upsilon(T) spline, the Boltzmann factor, the
// recombination feed) are evaluated in the fit precision R (double, or float
// in the mixed-precision form); the level matrix, its elimination and the
// populations are always FP64.
#pragma once
#include <hip/hip_runtime.h>

#include "moc_internal.h"

namespace {

// Lane I of this lane's team of W = 8, 16 or 32 consecutive lanes, read by
// every lane of the team without LDS: DPP row_newbcast broadcasts lane I of
// each 16-lane row to its row; teams of 8 pick their half's lane, and teams
// of 32 take the row that holds lane I through v_permlane16_swap (which
// exchanges the odd rows of one operand with the even rows of the other).
// Every lane of the wave must be active, and I a compile-time constant.
template <int I>
__device__ __forceinline__ int row_bcast(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + I, 0xF, 0xF, false);
}

template <int W, int I>
__device__ __forceinline__ int team_bcast_i(int v) {
    static_assert(W == 8 || W == 16 || W == 32, "team width");
    static_assert(I >= 0 && I < W, "team lane");
    if constexpr (W == 16) {
        return row_bcast<I>(v);
    } else if constexpr (W == 8) {
        const int lo = row_bcast<I>(v), hi = row_bcast<8 + I>(v);
        return (threadIdx.x & 8) ? hi : lo;
    } else {
        const int t = row_bcast<I % 16>(v);
        const auto p = __builtin_amdgcn_permlane16_swap(t, t, false, false);
        return I < 16 ? (int)p[0] : (int)p[1];
    }
}

template <int W, int I>
__device__ __forceinline__ double team_bcast_d(double v) {
    const int lo = team_bcast_i<W, I>(__double2loint(v));
    const int hi = team_bcast_i<W, I>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// the same with the lane a constant of the unrolled loops that call it; teams
// of 32 lanes solve ions of 17-27 levels, whose loops are not unrolled, and
// read the lane by a shuffle
template <int W>
__device__ __forceinline__ double team_bcast(double v, int i) {
    if constexpr (W > 16) return __shfl(v, i, W);
#define MOC_TB(I)                                                  \
    case I:                                                         \
        if constexpr (I < W) return team_bcast_d<W, (I < W ? I : 0)>(v); \
        break;
    switch (i) {
        MOC_TB(0) MOC_TB(1) MOC_TB(2) MOC_TB(3) MOC_TB(4) MOC_TB(5) MOC_TB(6) MOC_TB(7)
        MOC_TB(8) MOC_TB(9) MOC_TB(10) MOC_TB(11) MOC_TB(12) MOC_TB(13) MOC_TB(14) MOC_TB(15)
        MOC_TB(16) MOC_TB(17) MOC_TB(18) MOC_TB(19) MOC_TB(20) MOC_TB(21) MOC_TB(22) MOC_TB(23)
        MOC_TB(24) MOC_TB(25) MOC_TB(26) MOC_TB(27) MOC_TB(28) MOC_TB(29) MOC_TB(30) MOC_TB(31)
        default: break;
    }
#undef MOC_TB
    return __shfl(v, i, W);
}

// A table pointer as a global-memory pointer: the solves run in functions of
// their own, where the compiler no longer sees that the tables are in global
// memory and would read them with flat loads.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
    return (gptr<T>)p;
}

// Level count of n-level ion `atom` as solved: its file's levels, cut at maxLev.
__device__ __forceinline__ int nlevel_count(const MocBalanceArgs& a, int atom) {
    const int nf = (int)a.atomMeta[8 * atom];
    return nf < a.maxLev ? nf : a.maxLev;
}

// One ion's data and the interpolation of its Upsilon(T) tables at T_e: the
// natural cubic spline in log T (second derivatives from the host), clamped at
// the table ends; linear in log T with two temperatures; the one value with
// one (:2884-2905).  Every case is written as
//   Upsilon = wa q[lo] + wb q[hi] + ca q2[lo] + cb q2[hi],
// so that a pair costs four loads and no branch.
template <class R>
struct NlevIon {
    int nf, n, iup2p;
    gptr<double> g, en, A, q, q2;
    double atp;
    size_t lo, hi;      // offsets of the bracketing tables
    R wa, wb, ca, cb;   // weights of the bracketing values and second derivatives
    R lT, Te, cs, kT;   // log10 T, T, 8.63e-6 / sqrt(T), 1.4388463 / T

    __device__ NlevIon(const MocBalanceArgs& a, int atom, double Te_) {
        const gptr<int64_t> m = as_global(a.atomMeta + 8 * atom);
        nf = (int)m[0];
        const int nT = (int)m[1];
        const gptr<double> lt = as_global(a.atomLogT + m[2]);
        g = as_global(a.atomG + m[3]);
        en = as_global(a.atomE + m[3]);
        A = as_global(a.atomA + m[4]);
        q = as_global(a.atomQom + m[5]);
        q2 = as_global(a.atomQ2 + m[5]);
        iup2p = (int)m[6];
        atp = as_global(a.atomAtp)[atom];
        n = nf < a.maxLev ? nf : a.maxLev;
        Te = (R)Te_;
        lT = log10(Te);
        cs = (R)8.63e-6 / sqrt(Te);
        kT = (R)1.4388463 / Te;
        int klo = 0, khi = 0;
        wa = 1;
        wb = 0;
        ca = 0;
        cb = 0;
        if (nT == 2) {
            const R d = (R)lt[1] - (R)lt[0];
            khi = 1;
            wa = ((R)lt[1] - lT) / d;
            wb = (lT - (R)lt[0]) / d;
        } else if (nT > 2) {
            if (lT >= (R)lt[nT - 1]) {
                klo = khi = nT - 1;
            } else if (lT > (R)lt[0]) {
                khi = nT - 1;
                while (khi - klo > 1) {
                    const int k = (khi + klo) >> 1;
                    if ((R)lt[k] > lT) khi = k; else klo = k;
                }
                const R h = (R)lt[khi] - (R)lt[klo];
                wa = ((R)lt[khi] - lT) / h;
                wb = (lT - (R)lt[klo]) / h;
                ca = (wa * wa * wa - wa) * (h * h) / (R)6.;
                cb = (wb * wb * wb - wb) * (h * h) / (R)6.;
            }
        }
        lo = (size_t)klo * nf * nf;
        hi = (size_t)khi * nf * nf;
    }

    // collision rates of the pair i < j as equilibrium forms them
    // (:2906-2936): qij from level i to level j and qji back; the Boltzmann
    // factor goes with the upward rate and the lower level's weight
    __device__ void pair(int i, int j, R& qij, R& qji) const {
        const size_t pp = (size_t)i * nf + j;
        const R ups = wa * (R)q[lo + pp] + wb * (R)q[hi + pp] + ca * (R)q2[lo + pp] + cb * (R)q2[hi + pp];
        const R c = ups * cs;
        const R ei = (R)en[i], ej = (R)en[j];
        const bool up = ei <= ej;
        const R bz = c * exp((up ? ei - ej : ej - ei) * kT) / (R)g[i];
        const R dn = c / (R)g[j];
        qij = up ? bz : dn;
        qji = up ? dn : bz;
    }

    // recombination feed coefficient of level lev >= 1 (kind 1: radiative
    // power law plus dielectronic fit, kind 2: cubic in log T; :2828-2877)
    __device__ R rec_alpha(const MocBalanceArgs& a, int atomLevOff, int kind, int lev) const {
        if (!a.atomRecValid[atomLevOff + lev]) return 0;
        const double* r = a.atomRec + 11 * ((size_t)atomLevOff + lev);
        const R t4 = Te * (R)1.e-4;
        if (kind == 1) {
            const R z = (R)r[1], tz = t4 / (z * z);
            return (R)1.e-13 * (R)r[0] * z * (R)r[2] * pow(tz, (R)r[3]) / ((R)1. + (R)r[4] * pow(tz, (R)r[5])) +
                   (R)1.0e-12 * (R)r[0] * ((R)r[6] / t4 + (R)r[7] + (R)r[8] * t4 + (R)r[9] * t4 * t4) *
                       pow(t4, (R)-1.5) * exp(-(R)r[10] * (R)1.e4 / Te);
        }
        return pow((R)10., (R)r[0] + (R)r[1] * lT + (R)r[2] * lT * lT + (R)r[3] * lT * lT * lT);
    }
};

// Sum of the line emissivities A_ji dE_ij n_j * scale * unit of ion `atom`
// that exceed 1e-35, one ion per lane with the level matrix in registers: NL
// is a compile-time bound on the level count (n <= NL), the loops are
// unrolled over NL and predicated on n.  Each collision pair is evaluated
// once and placed where the matrix takes it; every diagonal entry receives
// its terms in the order of the sequential construction (:2944-2965).
//   rec      levels 2.. are fed by recombination from the next stage
//            (ionDenUp = X(i+1)/X(i)) at the rates of the ion's data file;
//   lineAcc  with it, every line's emissivity (lower i, upper j: fLineEm(i, j),
//            :2987-3005) times w is added to lineAcc[i * ld + j].
template <int NL, class R = double>
__device__ double nlevel_fixed(const MocBalanceArgs& a, int atom, double Te, double Ne,
                               double ionDenUp, bool rec, double scale, double unit,
                               double* lineAcc = nullptr, double w = 0., int ld = 0) {
    const NlevIon<R> I(a, atom, Te);
    const int n = I.n, nf = I.nf;
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
    // pairs in (i, j) order: row j takes x[j][i] and its diagonal's (j, i)
    // terms while rows i take theirs when their own i comes round
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
#pragma unroll
        for (int j = i + 1; j < NL; ++j) {
            if (j >= n) continue;
            R qij, qji;
            I.pair(i, j, qij, qji);
            // row i (if not the normalisation row): x[i][j] += Ne qe[j][i] + A[j][i]
            if (i >= 1) {
                x[i][j] += Ne * (double)qji;
                x[i][j] += I.A[(size_t)j * nf + i];
                x[i][i] -= Ne * (double)qij;
            }
            // row j: x[j][i] += Ne qe[i][j]; diagonal -= Ne qe[j][i] + A[j][i]
            x[j][i] += Ne * (double)qij;
            x[j][j] -= Ne * (double)qji;
            x[j][j] -= I.A[(size_t)j * nf + i];
        }
    }
    const int kind = rec ? a.atomRecKind[atom] : 0;
    if (kind != 0) {
        const int off = (int)a.atomMeta[8 * atom + 3];
#pragma unroll
        for (int i = 1; i < NL; ++i) {
            if (i >= n) break;
            y[i] = -Ne * ionDenUp * (double)I.rec_alpha(a, off, kind, i);
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
            double av = I.A[(size_t)j * nf + i];
            if (av == 0.) continue;
            if (j + 1 == I.iup2p && i == 0 && I.atp > 0.) av -= I.atp;
            const double em = av * (I.en[j] - I.en[i]) * (y[j] / sumN) * scale * unit;
            if (em > 1.e-35) tot += em;
            if (lineAcc) lineAcc[i * ld + j] += em * w;
        }
    }
    return tot;
}

// The same sum for ions of up to NL = 8, 16 or 27 levels, solved by a team of
// W = 8, 16 or 32 lanes of one wave with the level matrix in registers: team
// lane k < n owns row k (its entries and its right-hand side).  The pivot
// row's entries travel by DPP broadcasts (one division per lane and step); the back
// substitution is column-oriented (row k's solution broadcast once, every row
// above it updated in parallel); the populations and the line sums are
// gathered on every lane of the team.  Every lane of the team calls it with
// the same ion, and the return value is the same on every lane of the team.
// The ion's level count must not exceed NL.  The operations on each entry are
// those of luSlv; the row sums of the back substitution and the sum of the
// lines are taken in a different (fixed) order than a sequential evaluation.
template <int NL, class R = double>
__device__ __attribute__((noinline)) double nlevel_team(const MocBalanceArgs& a, int atom, double Te,
                                                        double Ne, double ionDenUp, bool rec,
                                                        double scale, double unit,
                                                        double* lineAcc = nullptr, double w = 0.,
                                                        int ld = 0) {
    constexpr int W = NL <= 8 ? 8 : (NL <= 16 ? 16 : 32);
    static_assert(NL <= W, "a team row per level");
    const int k = (int)(threadIdx.x & (W - 1));  // the row of this lane
    const NlevIon<R> I(a, atom, Te);
    const int n = I.n, nf = I.nf;
    const bool mine = k < n;
    // row k: row 0 the normalisation; row k >= 1 x[k][j] = Ne qe[j][k] + A[j][k]
    // (j > k), the diagonal minus the rates out of level k in j order
    double x[NL];
    double d = 0.;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        x[j] = 0.;
        if (mine && j < n && j != k) {
            R qa, qb;  // the pair (min, max): qa from the lower index to the upper
            I.pair(k < j ? k : j, k < j ? j : k, qa, qb);
            const R qkj = k < j ? qa : qb, qjk = k < j ? qb : qa;
            x[j] += Ne * (double)qjk;
            if (j > k) x[j] += I.A[(size_t)j * nf + k];
            d -= Ne * (double)qkj;
            if (j < k) d -= I.A[(size_t)k * nf + j];
        }
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        if (k == 0) x[j] = 1.;
        else if (j == k) x[j] = d;
    }
    double y = k == 0 ? 1. : 0.;
    const int kind = rec ? a.atomRecKind[atom] : 0;
    if (kind != 0 && k >= 1 && mine)
        y = -Ne * ionDenUp * (double)I.rec_alpha(a, (int)a.atomMeta[8 * atom + 3], kind, k);
    // elimination without pivoting: step i subtracts pivot row i from the rows below
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        const double piv = team_bcast<W>(x[i], i);
        const double f = x[i] / piv;
        const bool below = k > i;
#pragma unroll
        for (int j = i + 1; j < NL; ++j) {
            const double xij = team_bcast<W>(x[j], i);
            if (below) x[j] -= xij * f;
        }
        const double yi = team_bcast<W>(y, i);
        if (below) y -= yi * f;
    }
    // back substitution, column-oriented: the last row first
#pragma unroll
    for (int j = NL - 1; j >= 0; --j) {
        if (k == j) y = y / x[j];
        const double yj = team_bcast<W>(y, j);
        if (k < j && j < n) y -= yj * x[j];
    }
    // populations: the sum over the levels on every lane, in level order
    double sumN = 0.;
#pragma unroll
    for (int j = 0; j < NL; ++j) {
        const double yj = team_bcast<W>(y, j);
        if (j < n) sumN += yj;
    }
    // lines out of level k (the upper level), summed on its lane, then over the team
    const double pop = y / sumN;
    double tot = 0.;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        if (mine && i < k) {
            double av = I.A[(size_t)k * nf + i];
            if (av != 0.) {
                if (k + 1 == I.iup2p && i == 0 && I.atp > 0.) av -= I.atp;
                const double em = av * (I.en[k] - I.en[i]) * pop * scale * unit;
                if (em > 1.e-35) tot += em;
                if (lineAcc) lineAcc[i * ld + k] += em * w;
            }
        }
    }
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, W);
    return tot;
}

}  // namespace

namespace {

// Team classes of the n-level solves: ions of up to 5 levels one per lane;
// larger ones by teams of 8 lanes (6-8 levels), 16 lanes (9-10, 11-13, 14-16)
// and 32 lanes (17, 18-27).  A class's solve is unrolled to its level bound,
// so its steps follow the ion's size rather than the team's width.
constexpr int kNlevLane = 5;
__host__ __device__ inline int nlevel_team_class(int n) {
    if (n <= kNlevLane) return -1;
    if (n <= 8) return 0;
    if (n <= 10) return 1;
    if (n <= 13) return 2;
    if (n <= 16) return 3;
    return n <= 17 ? 4 : 5;
}

}  // namespace

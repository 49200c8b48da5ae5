// Free-free Gaunt factor of Hummer (1988): the Chebyshev fit of getGauntFF
// (ionization_mod.f90:196-336), split into a per-temperature set-up and a
// per-frequency evaluation.  Shared by the opacity and the emission kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace {

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
    double lg2;  // log10(gamma2)
};

__device__ GauntT gaunt_setup(double z, double log10Te) {
    GauntT g;
    double xlrkt;
    if (log10Te < 2.5) xlrkt = 5.1983649 - 2.5;
    else if (log10Te > 8.) xlrkt = 5.1983649 - 8.;
    else xlrkt = 5.1983649 - log10Te;
    const double txg = 0.66666667 * (2.0 * z + xlrkt);
    g.gamma2 = pow(10., txg * 1.5);
    g.lg2 = log10(g.gamma2);
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
    if (fabs(txu) <= 2.0) {  // the Chebyshev range needs no u
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
    const double u = exp10((txu - .90909091) / .72727273);
    // txu > 2 puts log10(u) above 1.5, so the small-u branch below never
    // takes these bins: testing txu first skips a log10 and keeps the result
    if (txu > 2.0) {
        if (g.lg2 < -3.) return sqrt(0.9549297 / u);
        if (g.lg2 < 0.0) {
            const double slope = (sqrt(12 * g.gamma2 / u) - sqrt(0.9549297 / u)) / 3.0;
            return sqrt(12 * g.gamma2 / u) + slope * g.lg2;
        }
        const double v = sqrt(12. * g.gamma2 / u);
        return v > 1.05 ? 1.05 : v;
    }
    if (log10(u) < -4.0) {
        if (g.gamma2 < 0.3) return 0.551329 * log(2.24593 / u);
        return 0.551329 * log(0.944931 / u / sqrt(g.gamma2));
    }
    return 0.0;
}

// The same fit in FP32 for the mixed-precision form: coefficients from the
// FP64 set-up, rounded once; the same branches.
struct GauntF {
    float c[9];
    float con, gamma2, lg2;
};

__device__ GauntF gaunt_f32(const GauntT& g) {
    GauntF f;
    for (int j = 0; j < 9; ++j) f.c[j] = (float)g.c[j];
    f.con = (float)g.con;
    f.gamma2 = (float)g.gamma2;
    f.lg2 = (float)g.lg2;
    return f;
}

// gaunt_eval in FP32 (the same branches)
__device__ float gaunt_eval_f32(const GauntF& g, float xlf) {
    const float txu = 0.72727273f * xlf + g.con;
    if (fabsf(txu) <= 2.0f) {
        float b[9];
        b[8] = g.c[8];
        b[7] = txu * b[8] + g.c[7];
        int ir = 6;
        for (int j = 1; j <= 6; ++j) {
            b[ir] = txu * b[ir + 1] - b[ir + 2] + g.c[ir];
            --ir;
        }
        return b[1] - b[3];
    }
    const float u = exp10f((txu - .90909091f) / .72727273f);
    if (txu > 2.0f) {
        if (g.lg2 < -3.f) return sqrtf(0.9549297f / u);
        if (g.lg2 < 0.0f) {
            const float slope = (sqrtf(12.f * g.gamma2 / u) - sqrtf(0.9549297f / u)) / 3.0f;
            return sqrtf(12.f * g.gamma2 / u) + slope * g.lg2;
        }
        const float v = sqrtf(12.f * g.gamma2 / u);
        return v > 1.05f ? 1.05f : v;
    }
    if (log10f(u) < -4.0f) {
        if (g.gamma2 < 0.3f) return 0.551329f * logf(2.24593f / u);
        return 0.551329f * logf(0.944931f / u / sqrtf(g.gamma2));
    }
    return 0.0f;
}

}  // namespace

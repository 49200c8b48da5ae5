/*
 * ORACLE -- test infrastructure only.  Plain-C FP64 restatement of MOCASSIN's
 * output reductions, used by tests/ to check the device path;
 * nothing in the product links it.
 *
 *   oracle_volume_averages  outputGas volume averages   output_mod.f90:642-730
 *                           (Harrington et al. 1982), cells in the
 *                           reference's i, j, k loop order (:281-283),
 *                           dV = getVolume (grid_mod.f90:2926-2985)
 *   oracle_sed              writeSED all-angle column   output_mod.f90:2986-3010
 *
 * Pins (tests/test_output.py): a uniform Te gives <Te> = Te for every ion; the
 * <X> of every element sum to 1; one reached cell gives that cell's values;
 * two cells weight by Ne H dV; cells no packet reached are left out; the SED
 * of a nebula that absorbs nothing carries the whole luminosity.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "moc_oracle_output.h"

double oracle_cell_width(const double* a, int n, int i, int sym, int plane) {
    double d;
    if (i > 0 && i < n - 1) d = fabs(a[i + 1] - a[i - 1]) / 2.;
    else if (i == 0) d = (sym || plane) ? fabs(a[1] - a[0]) / 2. : fabs(a[1] - a[0]);
    else d = plane ? fabs(a[i] - a[i - 1]) / 2. : fabs(a[i] - a[i - 1]);
    return d / 1.e15;
}

int oracle_volume_averages(const oracle_output_in* in, double* TeVol, double* ionDenVol) {
    const int NE = in->NE, NS = in->NS;
    const size_t C1 = (size_t)in->nCells + 1;
    double* T = calloc((size_t)NE * NS, sizeof(double));
    double* W = calloc((size_t)NE * NS, sizeof(double));
    int i, j, k, e, ion;
    if (!T || !W) { free(T); free(W); return -1; }
    for (i = 0; i < in->nx; i++)
        for (j = 0; j < in->ny; j++)
            for (k = 0; k < in->nz; k++) {
                const int c = in->active[(size_t)i + (size_t)in->nx * ((size_t)j + (size_t)in->ny * k)];
                double dV, f1, f2;
                if (c <= 0) continue;
                if (in->lgBlack && in->lgBlack[c] >= 1) continue;
                dV = oracle_cell_width(in->xAxis, in->nx, i, in->sym, in->plane) *
                     oracle_cell_width(in->yAxis, in->ny, j, in->sym, in->plane) *
                     oracle_cell_width(in->zAxis, in->nz, k, in->sym, in->plane);
                f1 = in->Te[c] * in->Ne[c] * in->Hden[c] * dV;
                f2 = in->Ne[c] * in->Hden[c] * dV;
                for (e = 0; e < NE; e++)
                    for (ion = 1; ion <= NS; ion++) {
                        const size_t p = (size_t)(ion - 1) * NE + e;
                        const double X = in->ionDen[p * C1 + c];
                        T[p] += X * f1;
                        W[p] += X * f2;
                    }
            }
    for (e = 0; e < NE; e++) {
        const int top = in->elemZ[e] + 1 < NS ? in->elemZ[e] + 1 : NS;
        double den = 0.;
        for (ion = 1; ion <= top; ion++) den += W[(size_t)(ion - 1) * NE + e];
        for (ion = 1; ion <= NS; ion++) {
            const size_t p = (size_t)(ion - 1) * NE + e;
            if (ion > top) { TeVol[p] = 0.; ionDenVol[p] = 0.; continue; }
            TeVol[p] = W[p] <= 0. ? 0. : T[p] / W[p];
            ionDenVol[p] = den <= 0. ? 0. : W[p] / den;
        }
    }
    free(T);
    free(W);
    return 0;
}

double oracle_sed(const double* escaped, int nCells, int nbins, int sym, double* sed) {
    const size_t C1 = (size_t)nCells + 1;
    double tot = 0.;
    int b;
    size_t c;
    for (b = 0; b < nbins; b++) {
        double s = 0.;
        for (c = 0; c < C1; c++) s += escaped[(size_t)b * C1 + c];
        sed[b] = sym ? s * 8. : s;
        tot += sed[b];
    }
    return tot;
}

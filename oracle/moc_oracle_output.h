/* ORACLE -- test infrastructure only (see moc_oracle_output.c). */
#ifndef MOC_ORACLE_OUTPUT_H
#define MOC_ORACLE_OUTPUT_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_output_in {
    int nx, ny, nz, nCells, sym;
    const double *xAxis, *yAxis, *zAxis; /* [n] cm */
    const int32_t* active;               /* [nx*ny*nz], x fastest */
    int NE, NS;
    const int32_t* elemZ;                /* [NE] */
    const double* Te;                    /* [C+1] */
    const double* Ne;                    /* [C+1] */
    const double* Hden;                  /* [C+1] */
    const double* ionDen;                /* [NS][NE][C+1] */
    const int32_t* lgBlack;              /* [C+1] or NULL */
    int plane;                           /* plane-parallel grid: both end cells half-width */
} oracle_output_in;

/* TeVol, ionDenVol: [NS][NE], element fastest */
int oracle_volume_averages(const oracle_output_in* in, double* TeVol, double* ionDenVol);
/* width of cell i along one axis / 1e15 (getVolume, grid_mod.f90:2926-2985):
 * the first cell is half-width in an octant or plane-parallel grid, the last
 * one in a plane-parallel grid only */
double oracle_cell_width(const double* a, int n, int i, int sym, int plane);
/* sed: [nbins]; escaped: [nbins][C+1]; returns the total */
double oracle_sed(const double* escaped, int nCells, int nbins, int sym, double* sed);

#ifdef __cplusplus
}
#endif
#endif

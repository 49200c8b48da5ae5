/*
 * moc_oracle_emission.h -- ORACLE (test infrastructure only): CPU, FP64
 * restatement of MOCASSIN's per-cell diffuse emission (emission_mod.f90:
 * emissionDriver :36-252 and the routines it calls, setDiffusePDF
 * :1528-1919).  Never linked by the product library.
 */
#ifndef MOC_ORACLE_EMISSION_H
#define MOC_ORACLE_EMISSION_H

#include <stdint.h>

#include "moc_oracle_balance.h"
#include "moc_oracle_output.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Raw tables as the reference reads them; the oracle derives the edge
 * pointers (initGammaCont ph_mod.f90:31-202), the log10 gamma values and the
 * line-bin pointers itself. */
typedef struct oracle_emission_tables {
    const oracle_balance_tables* bal; /* mesh, elements, n-level atoms */
    const double* widFlx;             /* [B] */
    int ntk;
    const double* tk;                 /* [ntk] tkGamma */
    int nlim[3];                      /* H I, He I, He II */
    const double* gnu[3];             /* [nlim] nuGamma */
    const double* gval[3];            /* [ntk][nlim] gamma (linear) */
    /* ground-state cross sections from their threshold pointer to nbins */
    int HIPnuP, HeIPnuP, HeIIPnuP;    /* 1-based; He pointers unused when He is off */
    const double *xsHI, *xsHeI, *xsHeII;
    /* heavy-ion free-bound, indexed z*32+ion: outer-shell bin range [lo, hi)
     * (lo = 0: none), offset of its cross sections in fbXs, g0, g1 */
    const int32_t *fbLo, *fbHi, *fbG0, *fbG1;
    const int64_t* fbOff;
    const double* fbXs;
    const double* hei2phot;           /* [41][2] y, A(y) */
    const int32_t* hlMap;             /* [9][12] file of (Z, T bin) */
    const double* hlDens;             /* [files][14] */
    const double* hlLines;            /* [files][14][16][9] */
    const double *hbA, *hbB;          /* [9][4] */
    const double* rbEdge;             /* [2][3][26] */
    const double* r2aEdge;            /* [3][18] */
    const double* heiRec;             /* [34][3][4] (first coefficient x1e25) */
    const double* heiiLyman;          /* [4][2] ratio, nu */
} oracle_emission_tables;

typedef struct oracle_emission_state {
    int nCells;
    const double* Te;       /* [C+1] */
    const double* Ne;       /* [C+1] */
    const double* ionDen;   /* [NS][NE][C+1] */
    double* recPDF;         /* [C+1][B] out (row 0 untouched) */
    double* totalLines;     /* [C+1] out */
    double* normCont;       /* [C+1] out, may be NULL: HI+HeI+HeII continuum */
    double* normRec;        /* [C+1] out, may be NULL: recombination lines */
    double* normFor;        /* [C+1] out, may be NULL: collisional lines */
    /* grains of gas + dust runs (dustS = 0 without): setDiffusePDF :1685-1752 */
    int dustS, dustA;
    const double* dustAbun;   /* [dustS] */
    const double* dustWeight; /* [dustA] */
    const double* dustTsub;   /* [dustS] */
    const double* dustCabs;   /* [dustS][dustA][B] */
    const double* Tdust;      /* [C+1][dustA+1][dustS+1] */
    const double* Ndust;      /* [C+1] */
    const double* Hden;       /* [C+1] */
    /* quantum heating (qIdx NULL: off): grain g = (species-1) dustA + size-1 with
     * qIdx[g] = q >= 0 emits from qT/qP [C+1][nQ][nTb] */
    int nQ, nTb;
    const int* qIdx;
    const double* qT;
    const double* qP;
} oracle_emission_state;

/* returns 0, or -1 when two gamma edges share a bin (fb_ff stops there) */
/* He I recombination line i of HeIrecLines.dat per n_e n(He+) [erg cm^3/s] */
double oracle_hei_line(const oracle_emission_tables* t, int i, double Te, double Ne);
int oracle_emission(const oracle_emission_tables* t, oracle_emission_state* s, int nthreads);
/* outputGas line luminosities (output_mod.f90:397-432) over the cells of g
 * (reference loop order, lgBlack < 1): lines[(Z*32 + stage)][L][L], entry
 * (i, j) = sum of fLineEm(i+1, j+1) H dV; hbeta = sum of hydroLines(1,4,2) H dV
 * over the cells where that exceeds 1e-35.  s supplies Te, Ne, ionDen; g the
 * grid, Hden and lgBlack. */
int oracle_line_luminosities(const oracle_emission_tables* t, const oracle_emission_state* s,
                             const oracle_output_in* g, int L, double* lines, double* hbeta);

/* the per-bin emission coefficients of one cell [1e-40 erg cm^3/s/Hz per n_H]:
 * out[0..B) HI (fb+ff+2q + heavy fb), [B..2B) HeI, [2B..3B) HeII, and the
 * pieces [3B..) gammaHI, gammaHeI, gammaHeII, gammaHeavies, ffCoeff1,
 * ffCoeff2, twoPhotHI, twoPhotHeI, twoPhotHeII (9 x B) */
int oracle_emission_coeffs(const oracle_emission_tables* t, double Te, double Ne,
                           const double* X /* [NS][NE] */, double* out);

#ifdef __cplusplus
}
#endif
#endif

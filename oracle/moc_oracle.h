/*
 * moc_oracle.h -- CPU restatement (plain C, FP64) of MOCASSIN's Monte Carlo
 * iteration, used ONLY as test infrastructure: the parity tests, the smoke
 * check and the cpu_baseline leg of bench.py call it to check results.  The
 * product (libmocgpu) never links or calls it.
 */
#ifndef MOC_ORACLE_H
#define MOC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block (Salmon et al. SC'11), for known-answer tests. */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* getNu2 (photon_mod.f90:928-972): 1-based bin from a CDF of n entries and a
 * uniform r in (0,1), by the reference's linear scan and bin-1 fold. */
int oracle_getnu2(const double* cdf, int n, double r);

/* locate (interpolation_mod.f90:49-106): 1-based bracket index, 0 or n out
 * of range, exact hits return the hit. */
int oracle_locate(const double* xa, int n, double x);

typedef struct oracle_transport_in {
    int nx, ny, nz, nCells, sym;
    const double *xAxis, *yAxis, *zAxis; /* [n] cm */
    const int32_t* active;               /* [nx*ny*nz], x fastest */
    double R_out;
    int nbins, nAngles, lgCompton;
    double ionEdge1;
    const double* nuArray;    /* [B] */
    const double* KNsigmaT;   /* [B] */
    const double* KNcdf;      /* [B][nAngles] */
    const double* Pcomp;      /* [B][nAngles] */
    const double* sourceCDF;  /* [B] */
    const double* opacity;    /* [B][C+1] cell fastest */
    const double* Ne;         /* [C+1] */
    const double* recPDF;     /* [C+1][B] bin fastest */
    const double* totalLines; /* [C+1] */
    const double* scaOpac;    /* [B][C+1] dust scattering kappa, NULL without grains */
    int lgDust;               /* grains present */
    int dustOnly;             /* no gas: re-emission from recPDF (= dustPDF), no lines */
    double star[3];
    int32_t starIdx[3];       /* 1-based */
    double deltaE;
    int64_t n;
    uint64_t seed;
    uint32_t ctr_base;
    int nthreads;             /* OpenMP threads (<=1: serial) */
    uint64_t id_base;         /* RNG packet id of packet i: id_base + i (a grid of a batch) */
    int plane;                /* plane-parallel illumination (lgPlaneIonization) */
    /* nested sub-grids (nSub > 0; setSubGrids grid_mod.f90:1834, pathSegment
     * photon_mod.f90:1931-1963, 2169-2223, 2884-3948): the grid above is the
     * mother, whose active map holds -g for the cells of sub-grid g
     * (g = 2..nSub+1 is subs[g-2]); packets start in grid startGrid (0: the
     * mother, k: subs[k-1]) at starIdx of that grid */
    int nSub;
    const struct oracle_subgrid* subs;
    int startGrid;
} oracle_transport_in;

/* one nested sub-grid: axes [cm] and active map (x fastest) of global cell ids */
typedef struct oracle_subgrid {
    int nx, ny, nz;
    const double *xAxis, *yAxis, *zAxis;
    const int32_t* active;
} oracle_subgrid;

/* counters[] slots */
enum {
    ORC_SEGMENTS = 0, ORC_EMISSIONS, ORC_ABSORBED, ORC_SCATTERED, ORC_ESCAPED,
    ORC_LINES, ORC_LOST, ORC_TRAPPED, ORC_ERRORS, ORC_N
};

typedef struct oracle_transport_out {
    double* J;        /* [B][C+1], accumulated (not zeroed here) */
    double* escaped;  /* [B][C+1], accumulated */
    double counters[ORC_N];
    double Qphot;
} oracle_transport_out;

/* energyPacketDriver (photon_mod.f90:28-414) for one stellar source. */
int oracle_transport(const oracle_transport_in* in, oracle_transport_out* out);

#ifdef __cplusplus
}
#endif
#endif

#ifndef MOC_ORACLE_OPACITY_H
#define MOC_ORACLE_OPACITY_H
#ifdef __cplusplus
extern "C" {
#endif

/* ionizationDriver / BoltGaunt / getGauntFF / eDenSum / addOpacity
 * (ionization_mod.f90:26-477) for every active cell.  The cross-sections
 * come as the reference's per-shell stack (elementP ranges). */
typedef struct oracle_opacity_in {
    int nbins, nCells, nstages, nElemUsed;
    const double* nuArray;      /* [B] */
    const double* brems;        /* [B] xSecArray(bremsXSecP ...) */
    const double* KNsigmaT;     /* [B] */
    int lgCompton;
    int nShells;
    const int32_t* shZ;         /* [nShells] element Z */
    const int32_t* shIon;       /* [nShells] stage */
    const int32_t* shLo;        /* [nShells] 1-based first bin */
    const int32_t* shHi;        /* [nShells] 1-based last bin */
    const int64_t* shOff;       /* [nShells] offset into xs */
    const double* xs;           /* stacked cross-sections [cm^2] */
    const int32_t* elemZ;       /* [nElemUsed] */
    const double* abund;        /* [nElemUsed] */
    const double* ionDen;       /* [nstages][nElemUsed][C+1] (Fortran ionDen) */
    const double* Ne;           /* [C+1] */
    const double* Te;           /* [C+1] */
    const double* Hden;         /* [C+1] */
} oracle_opacity_in;

void oracle_gaunt_ff(double z, double log10Te, const double* log10nu, int n, double* g);
int oracle_opacity(const oracle_opacity_in* in, double* opacity /* [B][C+1] */);

#ifdef __cplusplus
}
#endif
#endif

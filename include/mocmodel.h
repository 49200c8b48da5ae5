/*
 * mocmodel.h -- native model set-up of libmocgpu: the host work MOCASSIN does
 * once per model before its first Monte Carlo iteration, read from the
 * reference's own data/ directory, so that a Fortran or C host drives the
 * device iteration with no other program in the loop.
 *
 * What one moc_model_build call restates (rwesson/mocassin_xray, source/):
 *   the atomic data readers   phInit ph_mod.f90:523-581 (ph1.dat, ph2.dat),
 *                             initRRDR hydro_mod.f90:334-417 (rrBadnell.dat,
 *                             drBadnell.dat), radRecFit / dielectronic
 *                             update_mod.f90:3094-3305 (radrec.dat,
 *                             dielectronic.dat, aldrovandi.dat),
 *                             makeCollIonData / makeAugerData
 *                             hydro_mod.f90:13-53 (cion.dat, auger.dat),
 *                             equilibrium emission_mod.f90:2505-2877 (the
 *                             n-level files named in fileNames.dat),
 *                             initGammaCont ph_mod.f90:31-202 (gamma*.dat),
 *                             RecLinesEmission emission_mod.f90:856-1153
 *                             (hydroLinesFiles.dat and its tables,
 *                             gaussfitHb.dat, r1bEdge.dat, r2bEdge.dat,
 *                             r2aEdge.dat, HeIrecLines.dat, r2a0100old.dat),
 *                             HeI2photSub emission_mod.f90:764-817
 *                             (HeI2phot.dat);
 *   the frequency mesh        initCartesianGrid grid_mod.f90:160-382 (the
 *                             gas log mesh with the H-series and ion-edge
 *                             triplets, widFlx);
 *   the source                getFlux continuum_mod.f90:358-398, setProbDen
 *                             :415-504 (blackbody, optionally with a
 *                             F_nu ~ 1/nu power law), setCompton/KNsigma/
 *                             Pcomp ph_mod.f90:1883-1963;
 *   the cross-section stack   initXSecArray / makeOpacity ph_mod.f90:204-517,
 *                             861-921 with phFitEl :592-668 and setShells /
 *                             setPointers hydro_mod.f90:208-677;
 *   the balance tables        getOuterShell hydro_mod.f90:89-182, the weight
 *                             columns of getPhotoRates update_mod.f90:3971-4040
 *                             and compton :3834-3897, the charge-exchange fits
 *                             :1624-1690, the Upsilon spline
 *                             interpolation_mod.f90:207-283;
 *   the grid and its start    fillGrid / setMotherGrid grid_mod.f90:651-1295
 *                             (an octant of nx^3 cells with mirror walls),
 *                             setGrid's starting guess :1583-1645.
 * The tables equal those of the package's Python builders (atomic.py,
 * balance.py, emission.py, tables.py, nebula.py) to the last bits of the
 * maths library, integer tables exactly.
 *
 * Every entry point returns MOC_OK (0) or a negative MOC_ERR_* code; the
 * message of the last failure of these calls on the calling thread is
 * returned by moc_model_last_error().
 */
#ifndef MOCMODEL_H
#define MOCMODEL_H

#include "mocgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct moc_model moc_model;

/* One gas model (the keywords of an input deck that the path reads). */
typedef struct moc_model_spec {
    const char* dataDir;       /* the reference's data/ directory */
    int32_t precision;         /* MOC_PREC_* of the context moc_model_load creates */
    int32_t nElem;             /* elements switched on (lgElementOn), H among them */
    const int32_t* elemZ;      /* [nElem] atomic numbers, ascending */
    const double* abund;       /* [nElem] n(X)/n(H) (elemAbun) */
    int32_t nstages;           /* ionisation stages kept (nstages) */
    int32_t maxLev;            /* n-level truncation (27 keeps every level of the files) */
    double XHILimit;           /* convergence limit on X(H0) (0.05) */
    int32_t nbins;             /* points of the gas mesh (nbins) */
    double nuMin, nuMax;       /* mesh limits [Ryd] (nuMin, nuMax) */
    int32_t strictMesh;        /* 1: fail when coinciding edges leave fewer than nbins
                                  distinct points; 0: keep the shorter mesh */
    int32_t lgCompton;         /* Compton scattering of packets (lgCompton) */
    double Tstellar;           /* blackbody temperature [K] (TStellar) */
    double LStar;              /* luminosity [1e36 erg/s] (LStar) */
    double nPhotons;           /* packets per iteration (nPhotons) */
    double plFraction;         /* 0: the blackbody alone; else a F_nu ~ 1/nu power law */
    double plLoKeV, plHiKeV;   /* between these energies carries this fraction of the
                                  luminosity, the blackbody the rest */
    int32_t nx;                /* octant of nx^3 cells, mirror walls at the lower faces */
    double Rout;               /* edge of the octant [cm] (Rout) */
    double Hdens;              /* uniform n_H [cm^-3] (Hdensity), used when hden is NULL */
    const double* hden;        /* [nx*nx*nx] n_H per cell, x fastest, or NULL;
                                  cells at 0 are inactive */
    double TeStart;            /* starting T_e [K] (TeStart, 1e4) */
    double H0in;               /* starting X(H0) (H0Start, 1e-5) */
} moc_model_spec;

/* Reads the data files and builds every table of the model.  The model owns
 * its memory until moc_model_free. */
int moc_model_build(const moc_model_spec* spec, moc_model** model);
void moc_model_free(moc_model* model);
const char* moc_model_last_error(void);

/* The model's ABI structs, their pointers into the model's memory (valid
 * until moc_model_free); NULL arguments are skipped. */
int moc_model_get(const moc_model* model, moc_tables* tables, moc_grid_desc* grid,
                  moc_atomic* atomic, moc_balance_tables* balance,
                  moc_emission_tables* emission);

/* setGrid's starting state: Te, Ne, Hden [nCells+1] and ionDen in the
 * reference's (0:nCells, nElem, nstages) layout, cell fastest. */
int moc_model_state(const moc_model* model, int64_t* nCells, const double** Te,
                    const double** Ne, const double** Hden, const double** ionDen);

/* moc_create with the model's tables, then moc_grid_attach, moc_set_atomic,
 * moc_set_balance, moc_set_emission and the upload of the starting state:
 * the context is ready for moc_iterate.  The model may be freed after. */
int moc_model_load(moc_ctx** ctx, const moc_model* model);

/* Writes the model as a stream of named records (the model-file format of
 * the package's Fortran driver: "MOCDUMP1", then per record a 32-byte name,
 * an int32 kind 1 float64 / 2 int32 / 3 int64, an int32 0, an int64 count
 * and the values; "END" closes it), without run controls. */
int moc_model_write(const moc_model* model, const char* path);

#ifdef __cplusplus
}
#endif

#endif

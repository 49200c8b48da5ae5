/*
 * mocgpu.h -- C ABI of libmocgpu, the MI355X (gfx950) implementation of one
 * MOCASSIN Monte Carlo iteration: packet emission, packet transport with the
 * Lucy path-length J_nu estimator, the per-cell opacity, the per-cell diffuse
 * emission CDF and the per-cell ionization/thermal balance.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * MOC_OK (0) on success and a negative MOC_ERR_* code on failure; the message
 * of the last failure is returned by moc_last_error().  The library owns all
 * device memory; host pointers are read or written only during the call.
 * Calls on one context are blocking at return and not reentrant.
 *
 * Indexing follows MOCASSIN: cells are numbered 1..nCells by the `active`
 * map (row 0 is the inactive dummy cell) and frequency bins 1..nbins in the
 * Fortran arrays; in this ABI the bin index is 0-based (bin b here is bin b+1
 * of the Fortran arrays).
 *
 * Reference interfaces replaced (rwesson/mocassin_xray, source/):
 *   moc_transport   <- energyPacketDriver(iStar,n,grid,plot,gpLoc,cellLoc)
 *                      photon_mod.f90:28 (with energyPacketRun :418,
 *                      newPhotonPacket :976, pathSegment :1758,
 *                      comptonScatter :6805 as device functions)
 *   moc_opacity     <- ionizationDriver(grid,ix,iy,iz) ionization_mod.f90:26,
 *                      called per cell from iterateMC iteration_mod.f90:148
 *   moc_emission    <- emissionDriver(grids,ix,iy,iz,ig) emission_mod.f90:36
 *                      (setDiffusePDF :1528 -> recPDF, totalLines)
 *   moc_update      <- updateCell(grid,xP,yP,zP) update_mod.f90:15
 *   moc_normalize_j <- the J reduction/normalisation of iterateMC
 *                      iteration_mod.f90:944-1031
 *   moc_upload/moc_download <- direct access to grid_type components
 *                      common_mod.f90:308-371
 *   moc_set_dust    <- dustDriver(grid) dust_mod.f90:14 (table part, with
 *                      makeDustXsec's mixture ph_mod.f90:926-1585 and
 *                      dustEmissionInt dust_mod.f90:146-184 as inputs)
 *   moc_dust_opacity <- the dust block of iterateMC iteration_mod.f90:194-246
 *   moc_dust_emission <- setDustPDF emission_mod.f90:1921-1986 (dust only)
 *   moc_dust_temperature <- getDustT update_mod.f90:3346-3453 with the
 *                      dust-only convergence of updateCell :240-267
 */
#ifndef MOCGPU_H
#define MOCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOC_OK 0
#define MOC_ERR_ARG (-1)    /* invalid argument or size */
#define MOC_ERR_HIP (-2)    /* HIP runtime failure */
#define MOC_ERR_STATE (-3)  /* call out of order (e.g. no grid attached) */
#define MOC_ERR_KERNEL (-4) /* a kernel reported an insane packet/cell state */
#define MOC_ERR_NOMEM (-5)  /* device allocation failed */

/* arithmetic of the transport path (chosen at moc_create) */
#define MOC_PREC_FP64 0  /* positions, opacity, CDFs and path sums in FP64 */
#define MOC_PREC_MIXED 1 /* FP32 positions/opacity/CDFs, FP64 J accumulation */

/* element type of a host buffer passed to moc_upload/moc_download */
#define MOC_DT_F32 0
#define MOC_DT_F64 1
#define MOC_DT_I32 2

/* layout of a (cell, bin) host buffer */
#define MOC_LAYOUT_FORTRAN 0 /* (0:nCells, 1:nbins), cell index fastest */
#define MOC_LAYOUT_CELLROW 1 /* (1:nbins, 0:nCells), bin index fastest  */

/* per-cell fields (shape (nCells+1) unless noted) */
#define MOC_F_OPACITY 1    /* kappa_nu [cm^-1], (cell, bin) */
#define MOC_F_JSTE 2       /* J estimator, (cell, bin), FP64 on device */
#define MOC_F_RECPDF 3     /* diffuse emission CDF, (cell, bin) */
#define MOC_F_TOTALLINES 4 /* fraction of diffuse energy in lines */
#define MOC_F_NE 5         /* electron density [cm^-3] */
#define MOC_F_TE 6         /* electron temperature [K] */
#define MOC_F_HDEN 7       /* hydrogen density [cm^-3] */
#define MOC_F_ESCAPED 8    /* escapedPackets(origin cell, bin), angle bin 0 */
#define MOC_F_IONDEN 9     /* ion fractions (cell, elem, stage): ionDen */
#define MOC_F_CONVERGED 10 /* lgConverged (int32) */
#define MOC_F_BLACK 11     /* lgBlack (int32) */
#define MOC_F_EMISSION 12  /* diffuse continuum emissivity (cell, bin) */
#define MOC_F_NDUST 13     /* grain number density [cm^-3] (cell) */
#define MOC_F_TDUST 14     /* Tdust(0:nSpecies, 0:nSizes, 0:nCells) [K], species fastest */
#define MOC_F_DUSTPDF 15   /* dust emission CDF (cell, bin), dust-only runs */
#define MOC_F_ABSOPAC 16   /* dust absorption kappa [cm^-1] (cell, bin) */
#define MOC_F_SCAOPAC 17   /* dust scattering kappa [cm^-1] (cell, bin) */
/* grain terms of the last gas update of a photoelectric gas + dust run,
 * (0:nCells, nSpecies*nSizes), cell fastest, grain (species-1)*nSizes + size-1 */
#define MOC_F_GRAINCOLL 18 /* gas-grain collision heating of the grains (gasDustColl_d) */
#define MOC_F_GRAINPE 19   /* photoelectric cooling of the grains (photoelHeat_d) */

typedef struct moc_ctx moc_ctx;

/* Frequency mesh and source tables (photon_mod.f90 globals, xSec_mod). */
typedef struct moc_tables {
    int32_t nbins;             /* B */
    int32_t precision;         /* MOC_PREC_* */
    const double* nuArray;     /* [B] bin frequencies [Ryd] */
    const double* widFlx;      /* [B] bin widths [Ryd] */
    double ionEdge1;           /* lowest ionisation edge [Ryd] (ionEdge(1)) */
    int32_t lgCompton;         /* Compton scattering of packets on free e- */
    int32_t nAngles;           /* angular bins of the KN tables (180) */
    const double* KNsigmaT;    /* [B] total Klein-Nishina sigma [cm^2] */
    const double* KNsigmaCDF;  /* [B][nAngles] scattering-angle CDF */
    const double* Pcomp;       /* [B][nAngles] nu_out/nu_in */
    int32_t nStars;            /* point sources (>=1) */
    const double* sourceCDF;   /* [nStars][B] inSpectrumProbDen */
    const double* Lstar;       /* [nStars] luminosity [1e36 erg/s] */
    const double* nPhotons;    /* [nStars] packets per iteration (deltaE=L/N) */
    const double* starPosition;/* [nStars][3] [cm] */
    const int32_t* starIndex;  /* [nStars][3] 1-based cell indices (xP,yP,zP) */
} moc_tables;

/* One Cartesian grid (grid_type, common_mod.f90:308-371). */
typedef struct moc_grid_desc {
    int32_t nx, ny, nz;
    const double* xAxis;      /* [nx] cell centres [cm] */
    const double* yAxis;      /* [ny] */
    const double* zAxis;      /* [nz] */
    const int32_t* active;    /* [nx*ny*nz] column-major (x fastest); 1..nCells or 0 */
    int32_t nCells;
    int32_t lgSymmetricXYZ;   /* octant model with mirror walls at the lower faces */
    double R_out;             /* outer radius [cm]; <= 0 disables the test */
    /* plane-parallel illumination (planeIonization, set_input_mod.f90:256-261):
     * stellar packets start on the y = 0 plane at a uniform (x, z), moving
     * along +y (photon_mod.f90:796-859); packets leave through the y faces
     * and are mirrored at the x and z faces (:3249-3606); the end cells of
     * every axis are half-width (:2166-2223); 5000 segments per flight
     * (:1873-1878).  Exclusive with lgSymmetricXYZ; nx, nz >= 3. */
    int32_t lgPlaneIonization;
} moc_grid_desc;

/* Counters of one moc_transport call (photon_mod.f90 Qphot, trapped, ...). */
typedef struct moc_transport_stats {
    double Qphot;            /* ionising photons emitted [1e36/(erg) units as reference] */
    double segments;         /* cell-crossing segments walked */
    double emissions;        /* packet emission events (incl. re-emissions) */
    double absorbed;         /* absorption events */
    double scattered;        /* Compton scattering events */
    double escaped;          /* packets that left the grid (energy tallied) */
    double lines;            /* packets converted to line photons (end of life) */
    double lost;             /* packets dropped by a wall snap or safe limit */
    double trapped;          /* packets that hit the re-emission limit */
    double kernel_ms;        /* device time of the transport kernel(s) [ms] */
} moc_transport_stats;

/* Elements, abundances and the photoionisation cross-section stack
 * (xSec_mod: initXSecArray ph_mod.f90:204-517, setPointers
 * hydro_mod.f90:422-677), summed per species (element, stage). */
typedef struct moc_atomic {
    int32_t nElemUsed;           /* elements switched on (lgElementOn) */
    const int32_t* elemZ;        /* [nElemUsed] atomic numbers */
    const double* abund;         /* [nElemUsed] n(X)/n(H) (elemAbun) */
    int32_t nstages;             /* ionisation stages kept (ionDen extent) */
    int32_t nSpecies;            /* photoionisable species */
    const int32_t* speciesElem;  /* [nSpecies] 0-based index into elemZ */
    const int32_t* speciesIon;   /* [nSpecies] stage, 1 = neutral */
    const double* sigma;         /* [nSpecies][nbins] cross-section [cm^2] */
    const double* brems;         /* [nbins] free-free factor 1.0368e-18/nu^3 */
} moc_atomic;

/* Tables of the ionisation/thermal balance (update_mod.f90 and the data it
 * reads).  Index conventions: Z and stages 1-based; [31][32] tables are
 * indexed z*32+stage (or z*32+nElec).  The frequency integrals of the balance
 * (photoionisation and heating per shell, getPhotoRates; the Compton sums)
 * are given as weight columns W[nbins][ncol], ncol = 3*nShells + 5:
 * per shell (gamma, heat below 7.353 Ryd, heat above -- scaled by the
 * secondary-ionisation efficiency), then sum J, sum J/nu, H recoil, He
 * recoil, and a hit column (1 in every bin). */
typedef struct moc_balance_tables {
    int32_t maxLev;              /* n-level truncation of the cooling ions */
    double XHILimit;             /* convergence limit on X(H0) (XHILimit) */
    int32_t nShells;
    const int32_t* shellNum;     /* [nShells] shell number, grouped by species */
    const int32_t* speciesShell; /* [nSpecies+1] shell range of each species */
    const double* W;             /* [nbins][3*nShells+5] */
    const int32_t* outShell;     /* [31*32] (Z, nElec) */
    const int32_t* statW0;       /* [31*32] g0 (getOuterShell) */
    const int32_t* statW1;       /* [31*32] g1 */
    const double* rrb;           /* [6][31][31] Badnell RR (A,B,T0,T1,C,T2) by (Z, N) */
    const double* drb_c;         /* [9][31][31] Badnell DR c */
    const double* drb_e;         /* [9][31][31] Badnell DR E */
    const double* rrec;          /* [2][31][31] radRecFit power laws */
    const double* rnew;          /* [4][31][31] radRecFit Verner-Ferland fits */
    int32_t nDiel;
    const double* diel;          /* [nDiel][8] elem n a b c d f g */
    int32_t nAldro;
    const double* aldro;         /* [nAldro][6] elem n a b t0 t1 */
    const double* cionCF;        /* [5][31][31] Voronov by (Z, nElec) */
    const double* cionTherm;     /* [31][32] */
    const int32_t* nAuger;       /* [31][32][11] */
    const double* auger;         /* [31][32][11][11] */
    int32_t nChex;
    const int32_t* chexZ;        /* [nChex] */
    const int32_t* chexIon;      /* [nChex] */
    const double* chexFit;       /* [nChex][4] a[1e-9 cm^3/s] b c d */
    int32_t nAtoms;
    const int32_t* atomIdx;      /* [31*32] n-level ion or -1 */
    const int64_t* atomMeta;     /* [nAtoms][8] nLev nT offT offLev offA offQ iup2p 0 */
    const double* atomAtp;       /* [nAtoms] */
    int64_t nLogT, nLev, nA, nQ;
    const double* atomLogT;
    const double* atomG;
    const double* atomE;
    const double* atomA;
    const double* atomQom;       /* Upsilon [nT][nLev][nLev] per ion */
    const double* atomQ2;        /* natural-spline second derivatives in log T */
    /* recombination feed of the upper levels (equilibrium rec=.true.,
     * emission_mod.f90:2828-2877), used by the diffuse emission */
    const int32_t* atomRecKind;  /* [nAtoms] 0 none, 1 (br,z,a_r(4),a_d(5)) fit, 2 log-T cubic */
    const double* atomRec;       /* [nLev][11] per level, same offsets as atomG */
    const int32_t* atomRecValid; /* [nLev] 1 where the level has a rate */
} moc_balance_tables;

typedef struct moc_update_stats {
    double noHit, noIonBal, noTeBal; /* grid%noHit, noIonBal, noTeBal */
    double converged;                /* cells with lgConverged = 1 */
    double kernel_ms;
} moc_update_stats;

/* Tables of the diffuse emission (emissionDriver, emission_mod.f90:36-252):
 * the H I / He I / He II free-bound fits with their edge pointers
 * (initGammaCont, ph_mod.f90:31-202), the ground-state cross-sections above
 * the H / He thresholds, the heavy-ion outer-shell free-bound entries, the
 * He I two-photon shape, the hydrogenic and He I recombination-line tables
 * and the He I 2^3S / He II Lyman line bins.  Built once per mesh and
 * element set (the package's emission.build_emission). */
typedef struct moc_emission_tables {
    int32_t ntk;
    const double* tkGamma;       /* [ntk] */
    int32_t nlim[3];             /* points of the H I, He I, He II fits (even, <= 64) */
    const int32_t* edgeP[3];     /* [nlim] 1-based bin pointers (HINuEdgeP, ...) */
    const double* logGamma[3];   /* [ntk][nlim] log10 gamma, 0 where gamma <= 0 */
    int32_t HIPnuP, HeIPnuP, HeIIPnuP; /* ground-state threshold bins (He: 0 when off) */
    const double* xsHI;          /* [nbins-HIPnuP+1] cm^2 */
    const double* xsHeI;         /* NULL when He is off */
    const double* xsHeII;
    int32_t nFb;                 /* heavy-ion (Z = 3..19) free-bound entries */
    const int32_t* fbZ;          /* [nFb] Z ascending, then ion */
    const int32_t* fbIon;
    const int32_t* fbLo;         /* [nFb] outer-shell bin range [lo, hi), 1-based */
    const int32_t* fbHi;
    const int64_t* fbOff;        /* [nFb] offsets into fbXs */
    int64_t nFbXs;
    const double* fbXs;
    const double* fbGratio;      /* [nFb] g0/g1 (getOuterShell) */
    const int32_t* he2Row;       /* [nbins] He I two-photon table row per bin, 0 = none */
    const double* hei2phot;      /* [41][2] y, A(y) */
    int32_t nHlFiles;
    const int32_t* hlMap;        /* [9][12] table file of (Z, T bin) */
    const double* hlDens;        /* [nHlFiles][14] */
    const double* hlLines;       /* [nHlFiles][14][16][9] relative to Hbeta */
    const double* hbA;           /* [9][4] Hbeta fits */
    const double* hbB;           /* [9][4] */
    const double* rbEdge;        /* [2][3][26] */
    const double* r2aEdge;       /* [3][18] */
    const double* heiRec;        /* [34][3][4] He I line fits (first coefficient x1e25) */
    const double* heiiLyman;     /* [4][2] He II Lyman ratio, nu */
    int32_t j2TsP;               /* bin of He I 2^3S (1.45673 Ryd) */
    int32_t lymP[4];             /* bins of the He II Lyman lines, 0 = below the mesh */
} moc_emission_tables;

/* Grains (dustDriver, dust_mod.f90:14-186): the per-grain cross-sections of
 * every (species, size), the Kirchhoff emission integrals and the mixture.
 * Built once per mesh (the package's dust.build_dust). */
typedef struct moc_dust_tables {
    int32_t nSpecies, nSizes;
    int32_t nTemps;              /* temperatures of emIntegral: 1..nTemps K (3000) */
    int32_t lgGas;               /* 0: dust only (absorbed packets re-emitted from dustPDF);
                                    1: gas + dust (re-emitted as diffuse packets) */
    int32_t lgPhotoelectric;     /* grain charge, photoelectric heating and gas-grain
                                    collisions (update_mod.f90:484-884); gas + dust only,
                                    after moc_set_balance */
    const double* grainAbun;     /* [nSpecies] */
    const double* grainWeight;   /* [nSizes] */
    const double* TdustSublime;  /* [nSpecies] K */
    const double* Cabs;          /* [nSpecies][nSizes][nbins] cm^2 per grain */
    const double* Csca;          /* [nSpecies][nSizes][nbins] */
    const double* emIntegral;    /* [nSpecies][nSizes][nTemps] dustEmIntegral */
    double XHILimit;             /* dust-only convergence limit on the mean grain T */
    /* lgPhotoelectric only (may be NULL otherwise): */
    const double* grainRadius;   /* [nSizes] um */
    const double* grainVn;       /* [nSpecies] potential of a neutral grain [Ryd] */
    const double* MsurfAtom;     /* [nSpecies] mass of a surface atom [amu] */
    const double* ipRyd;         /* [31*32] (Z, stage >= 2): the energy [Ryd] released when the
                                    stage recombines, nu of the first bin of the outer shell of
                                    stage-1 (setDustGasCollHeatCool :796-806) */
    /* quantum heating (qHeat, emission_mod.f90:1993-2504): grains smaller than minaQHeat
     * emit from their temperature distribution once more than minConvQHeat per cent of
     * the cells have converged, from the second iteration on (:1701-1702, :1946-1947);
     * needs grainRadius */
    int32_t lgQHeat;
    int32_t nTbins;              /* temperature bins of a distribution, >= 11 (300) */
    double minaQHeat;            /* um (1e-3) */
    double minConvQHeat;         /* per cent (99) */
    double Tmax;                 /* K, the top bin (700) */
    const int32_t* grainKind;    /* [nSpecies] the first letter of the species' label:
                                    1 'S' (silicate), 2 'C' (carbon), 0 other (no
                                    distribution: the grain then emits nothing) */
    /* lgEquivalentTau (iteration_mod.f90:200-203): the first iteration runs
     * without the grains' opacity, so its escaped SED is the unextinguished
     * one, and later iterations with it; the equivalent optical depth per bin
     * is ln(SED_1 / SED) (output_mod.f90:3020-3035).  moc_iterate knows the
     * iteration; a host that calls the phases one by one says it with
     * moc_set_qheat_state. */
    int32_t lgEquivalentTau;
} moc_dust_tables;

/* Create a context: uploads the mesh and source tables. */
int moc_create(moc_ctx** ctx, const moc_tables* tables);
void moc_destroy(moc_ctx* ctx);
const char* moc_last_error(const moc_ctx* ctx);

/* Attach the (single) Cartesian grid; allocates all per-cell device arrays
 * and zeroes them. */
int moc_grid_attach(moc_ctx* ctx, int ig, const moc_grid_desc* grid);

/* Volume averages of outputGas (output_mod.f90:642-730) over the cells that
 * packets reached (lgBlack < 1): for each element slot e and stage i,
 *   TeVol(e, i)     = sum X Te Ne H dV / sum X Ne H dV
 *   ionDenVol(e, i) = sum X Ne H dV / sum over the element's stages of the same
 * (Harrington et al. 1982), dV the cell volume (getVolume).  Both outputs
 * are [nstages][nElemUsed], element fastest (the reference's TeVol(0, elem,
 * ion) order), 0 above stage min(Z+1, nstages).  One nebula per context:
 * a batch of grids (moc_grid_attach_batch) returns MOC_ERR_STATE, as do the
 * line luminosities and the SED. */
int moc_volume_averages(moc_ctx* ctx, int ig, double* TeVol, double* ionDenVol);

/* moc_volume_averages without a device allocation: the same outputs, the
 * partial sums held in a fixed device buffer of 4 MiB (2^19 doubles) that
 * each GPU keeps from the first moc_create on it until the process ends.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Returns MOC_ERR_ARG when nElemUsed * nstages * (chunks
 * of 4096 cells, at most 256) * 2 + nElemUsed * nstages * 2 exceeds the
 * buffer.  Every call on one GPU shares the buffer: calls that overlap on one
 * GPU (from several threads or contexts) race. */
int moc_volume_averages_noalloc(moc_ctx* ctx, int ig, double* TeVol, double* ionDenVol);

/* Line luminosities of outputGas (output_mod.f90:397-432): over the cells
 * packets reached, the sum of H dV times the emissivity of every line of every
 * n-level ion (forLines, emission_mod.f90:1481-1526: equilibrium with the
 * recombination feed, times abundance, ionDen and 1.9865e9), and Hbeta = the
 * sum of hydroLines(1,4,2) H dV over the cells where that exceeds 1e-35.  The
 * ions are those with atomic data, Z then stage ascending; nIons and nLev (the
 * level count, maxLev) come back first; with lines non-NULL, lines[(k * nLev
 * + i) * nLev + j] is ion k's line from level j+1 down to level i+1 (i < j;
 * other entries 0), count >= nIons * nLev * nLev, and ionZ / ionStage (may be
 * NULL) name the ions.  Lines relative to Hbeta are lines / Hbeta. */
int moc_line_luminosities(moc_ctx* ctx, int ig, int* nIons, int* nLev, int* ionZ, int* ionStage,
                          double* lines, int64_t count, double* Hbeta);

/* writeSED (output_mod.f90:2947-3064), the all-angle column: sed[nbins] =
 * the escaped energy of each bin summed over the cells, times 8 for an
 * octant (lgSymmetricXYZ) model; totalE (may be NULL) = its sum.  nu*F(nu) of
 * SED.out is sed * nuArray / widFlx. */
int moc_sed(moc_ctx* ctx, int ig, double* sed, double* totalE);

/* A batch of independent grids in one context (no counterpart in the
 * reference, which runs one nebula per process; `multiGrids` there is nested
 * refinement, not batching): nGrids Cartesian grids that share the frequency
 * mesh, the element set and every table.  Grid k (0-based) is lit by star k+1
 * of moc_tables, whose position and cell indices are in grid k's own frame.
 * The grids' cells are numbered one grid after the other -- cell c of grid k
 * is cell cellBase[k] + c, cellBase[k] the sum of nCells of grids 0..k-1 --
 * and every per-cell field (MOC_F_*) is over these numbers.  The grids must
 * share lgSymmetricXYZ.  moc_iterate then transports every grid's packets in
 * one launch (packet i of grid k draws from the Philox stream of packet id
 * (k << 32) + i, counter base nIterateMC * 256 + 1) and runs opacity,
 * emission and balance once over all cells; moc_transport(iStar) transports
 * star iStar's packets in grid iStar-1.  Grains are not supported in a batch. */
int moc_grid_attach_batch(moc_ctx* ctx, int nGrids, const moc_grid_desc* grids);

/* Nested sub-grids (multiGrids: setSubGrids grid_mod.f90:1834-2100, fillGrid
 * :907-960, the sub-grid branches of pathSegment photon_mod.f90:1931-1963,
 * 2169-2223, 2884-3948): grids[0] is the mother grid, grids[1..nGrids-1]
 * its sub-grids, each inside the mother's box.  The mother's active map
 * holds -g for the cells that sub-grid g (1-based grid number, 2..nGrids)
 * covers; a sub-grid's own map numbers its cells 1..nCells (0 inactive) and
 * its lgSymmetricXYZ / lgPlaneIonization are ignored (it takes the mother's
 * symmetry; plane-parallel mothers are refused).  Cells are numbered as in a
 * batch (cell c of grid k is cellBase[k] + c); the single source of
 * moc_tables starts in the grid that holds its mother cell.  Packets cross
 * into a sub-grid where they reach a covered mother cell, walk its cells
 * (end cells half-width on every face, no loss at its faces), and go back to
 * the mother cell that holds their position when they leave its box.
 * Opacity, emission, balance, normalisation and the output reductions run
 * over every grid's cells; grains are not supported. */
int moc_grid_attach_nested(moc_ctx* ctx, int nGrids, const moc_grid_desc* grids);

/* Copy a per-cell field between host and device.  `count` is the number of
 * elements of the host buffer; (cell, bin) fields take `layout`. */
int moc_upload(moc_ctx* ctx, int field, int ig, const void* host, int64_t count,
               int dtype, int layout);
int moc_download(moc_ctx* ctx, int field, int ig, void* host, int64_t count,
                 int dtype, int layout);

/* Upload the element set and cross-section stack.  The per-cell ionDen
 * field (MOC_F_IONDEN, (0:nCells, nElemUsed, nstages)) exists once both the
 * grid and the atomic set are known. */
int moc_set_atomic(moc_ctx* ctx, const moc_atomic* atomic);

/* kappa_nu of every cell from ionDen, Ne, Te, Hden (ionizationDriver,
 * ionization_mod.f90:26-477; free-free Gaunt factors at each cell's T_e). */
int moc_opacity(moc_ctx* ctx, int ig);

/* Upload the grain tables (after moc_grid_attach): allocates Ndust, Tdust
 * (50 K, as dustDriver), absOpac, scaOpac.  Dust-only contexts need no
 * atomic, balance or emission tables; their dust emission CDF is held in the
 * recPDF storage (MOC_F_DUSTPDF and MOC_F_RECPDF are the same array). */
int moc_set_dust(moc_ctx* ctx, const moc_dust_tables* tables);

/* The dust block of iterateMC (iteration_mod.f90:194-246): absOpac, scaOpac
 * of every cell from Ndust and the grains below their sublimation
 * temperature, and opacity = gas opacity (gas + dust) + absOpac + scaOpac. */
int moc_dust_opacity(moc_ctx* ctx, int ig);

/* setDustPDF for every active cell (emission_mod.f90:1921-1986; dust only). */
int moc_dust_emission(moc_ctx* ctx, int ig);

/* getDustT for every active cell (update_mod.f90:3346-3453) from the
 * normalised J: Tdust of every (species, size), the size-weighted species
 * means and the abundance-weighted mean; cells no packet reached keep
 * theirs (lgBlack).  In dust-only runs lgConverged comes from the change of
 * the mean grain temperature (:240-267). */
int moc_dust_temperature(moc_ctx* ctx, int ig, moc_update_stats* stats);

/* Upload the balance tables (after moc_set_atomic). */
int moc_set_balance(moc_ctx* ctx, const moc_balance_tables* tables);

/* Upload the emission tables (after moc_set_balance, whose n-level atoms
 * the forbidden lines use). */
int moc_set_emission(moc_ctx* ctx, const moc_emission_tables* tables);

/* emissionDriver for every active cell (emission_mod.f90:36; setDiffusePDF
 * :1528): from Te, Ne, ionDen writes recPDF (the normalised diffuse CDF)
 * and totalLines.  kernel_ms (may be NULL) gets the kernel time. */
int moc_emission(moc_ctx* ctx, int ig, float* kernel_ms);

typedef struct moc_iteration_stats {
    double convPercent;              /* % of active cells with lgConverged */
    double noHitPercent, noIonBalPercent, noTeBalPercent;
    double packets;                  /* energy packets transported, all stars */
    double opacity_ms, emission_ms, transport_ms, update_ms; /* kernel times */
    double segments, emissions, escaped; /* transport counters, all stars */
    double iteration_ms;             /* device time of the iteration: the sum of the four
                                        phases, or the whole graph when it is replayed
                                        (moc_set_graph; the phase times are then 0) */
} moc_iteration_stats;

/* Packets of star iStar per iteration (nPhotons(iStar)); the packet energy
 * is Lstar/nPhotons (photon_mod.f90:86). */
int moc_set_packets(moc_ctx* ctx, int iStar, int64_t nPhotons);

/* One Monte Carlo iteration on the device (iterateMC, iteration_mod.f90:
 * 32-1497, without its file output and MPI reductions): opacity and diffuse
 * emission of every cell -- from iteration 2 on, cells whose Te and Ne moved
 * by less than 0.5% since the last balance keep theirs -- then the packets
 * of every star, the normalisation of J and updateCell of every cell.
 * Needs moc_set_atomic, moc_set_balance and moc_set_emission.  Packet
 * streams are keyed by (seed, nIterateMC, iStar). */
int moc_iterate(moc_ctx* ctx, int ig, int nIterateMC, uint64_t seed, moc_iteration_stats* stats);

/* Replay of moc_iterate as one graph launch.  No counterpart in the
 * reference, whose iterateMC (iteration_mod.f90:32) is host-driven; it is
 * for callers that run many iterations of a large model, where the launch
 * and wait between the phases cost more than the small kernels.  With
 * enable != 0, iterations from the second on capture their device work once
 * per seed (opacity, emission, tally reset, every star's transport, the J
 * normalisation, the balance and the grain temperatures) and replay it; the
 * iteration number reaches the packet streams through device memory, so a
 * replay draws exactly the packets of the eager call.  Any moc_set_* call,
 * a new grid or new packet counts drop the captured graph.  Tuning values
 * and arguments are those of the capture. */
int moc_set_graph(moc_ctx* ctx, int enable);

/* Launch parameters of the transport kernel.  No counterpart in the
 * reference (its photon loop has no launch shape).  None of them changes
 * which packets are drawn: each packet keeps its own RNG stream, so two
 * settings differ only in the order of the atomic additions to J. */
typedef struct moc_tuning {
    int32_t emitBatch;      /* diffuse packets a wave waits for before it emits them
                               together, 1..64 (16) */
    int32_t segsPerTrip;    /* path segments a flying lane steps per trip through the
                               wave's loop, 1..64 (16) */
    int32_t transportWaves; /* resident waves of one transport launch, 64..2^20 (4096:
                               four per SIMD) */
    int32_t graph;          /* moc_get_tuning: 1 when moc_set_graph is on */
    int32_t hotJ;           /* mixed form: 1 sums the J tallies of the cells around the
                               star (times the bins that carry most path there) as FP32
                               partials in LDS, workgroups of 1024 lanes; 0 adds every
                               tally with a global atomic (1) */
} moc_tuning;
int moc_set_tuning(moc_ctx* ctx, const moc_tuning* t);
int moc_get_tuning(moc_ctx* ctx, moc_tuning* t);

/* updateCell for every active cell (update_mod.f90:15): from the normalised
 * J (moc_normalize_j) and Te, Ne, Hden, ionDen, writes Te, Ne, ionDen,
 * lgConverged, lgBlack. */
int moc_update(moc_ctx* ctx, int ig, moc_update_stats* stats);

/* Zero the J estimator and the escape tallies (iteration_mod.f90:626-717). */
int moc_reset_tallies(moc_ctx* ctx, int ig);

/* Transport n packets of source iStar (>=1) through the grid, accumulating
 * J and escape tallies on the device.  The counter-based RNG stream of
 * packet p is Philox4x32-10 with key = seed and counter = (draw, p_lo, p_hi,
 * ctr_base): results do not depend on the launch shape. */
int moc_transport(moc_ctx* ctx, int iStar, int64_t n, uint64_t seed,
                  uint32_t ctr_base, moc_transport_stats* stats);

/* J <- J * 1e-9 (and /8 for symmetric grids), escape tallies /8 likewise
 * (iteration_mod.f90:1002-1031). */
int moc_normalize_j(moc_ctx* ctx, int ig);

/* Device pointer of a field, for callers that keep data in HBM. */
int moc_device_ptr(moc_ctx* ctx, int field, int ig, void** dptr, int64_t* count,
                   int* dtype);

/* Library self-description: "gfx950 mocgpu <version>". */
const char* moc_version(void);

/* Calibration of the transport's bound (no counterpart in the reference: it
 * measures the device): the rate [additions/s] of no-return FP64 atomic adds
 * at scattered addresses of a buffer of `bytes` (a power of two, 1 MiB ..
 * 4 GiB) -- the operation of every J tally.  Allocates and frees the buffer;
 * needs no context. */
int moc_calibrate_atomics(uint64_t bytes, double* rate);

/* The iteration state that switches quantum heating on (lgQHeat): the
 * emission of the next moc_dust_emission / moc_emission uses the grains'
 * temperature distributions when nIterateMC > 1 and convPercent (of the last
 * iteration) > minConvQHeat (emission_mod.f90:1701-1702).  moc_iterate sets
 * it itself; this is for callers that run the phases one by one.  Replaces
 * the module variables nIterateMC and convPercent that qHeat's callers read. */
int moc_set_qheat_state(moc_ctx* ctx, int nIterateMC, double convPercent);

#ifdef __cplusplus
}
#endif

#endif /* MOCGPU_H */

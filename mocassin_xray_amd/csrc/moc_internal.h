// Internal state of a libmocgpu context (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mocgpu.h"

#define MOC_RECURSION_LIMIT 100000   // re-emissions per packet (constants_mod.f90:70)
#define MOC_SAFE_LIMIT 500000        // segments per pathSegment call (photon_mod.f90:1873-1878)
#define MOC_SAFE_LIMIT_PLANE 5000    // the same, plane-parallel illumination

// Device counters of one transport launch (doubles accumulated as u64).
enum {
    MOC_CNT_SEGMENTS = 0,
    MOC_CNT_EMISSIONS,
    MOC_CNT_ABSORBED,
    MOC_CNT_SCATTERED,
    MOC_CNT_ESCAPED,
    MOC_CNT_LINES,
    MOC_CNT_LOST,
    MOC_CNT_TRAPPED,
    MOC_CNT_ERROR,      // insane index / state (Fortran `stop` in the reference)
    MOC_CNT_NEXT,       // work-queue head
    MOC_CNT_N
};

// One grid of a batch as the transport sees it: its axes and volume factors,
// its active map (global cell ids; null when every cell is active and the id
// is base + 1 + iz + nz*(iy + ny*ix)) and its star.
template <typename T>
struct MocGridView {
    const T *ax, *ay, *az, *vx, *vy, *vz;
    const int* active;
    const T* sourceCDF;   // [B] of the grid's star
    int nx, ny, nz, base, sym, plane;
    int sub;              // a nested sub-grid (its mother is view 0)
    int starIdx[3];       // 0-based
    T R_out, deltaE;
    T star[3];
};

// Transport view of a grid in arithmetic type T (see transport.hip).
template <typename T>
struct MocTransportArgs {
    // grid
    int nx, ny, nz, nCells, sym;
    int plane;          // plane-parallel illumination (lgPlaneIonization)
    const T* ax;        // [nx] axes
    const T* ay;
    const T* az;
    const T* vx;        // [nx] cell widths / 1e15 (volume factors)
    const T* vy;
    const T* vz;
    const int* active;  // [nx*ny*nz], x fastest
    int dense;          // all cells active, id = 1 + iz + nz*(iy + ny*ix)
    T R_out;
    // mesh & tables
    int nbins, nAngles, lgCompton;
    T ionEdge1;
    const T* nuArray;     // [B]
    const T* KNsigmaT;    // [B]
    const T* KNcdf;       // [B][nAngles]
    const T* Pcomp;       // [B][nAngles]
    const T* sourceCDF;   // [B] of this source
    // per-cell
    const T* opacity;     // [B][C+1]   (cell fastest)
    const T* Ne;          // [C+1]
    const T* recPDF;      // [C+1][B]   (bin fastest); the dust CDF in dust-only runs
    const T* scaOpac;     // [B][C+1]   dust scattering kappa (lgDust)
    int lgDust;           // grains present: dust scattering/absorption, no early escape
    int dustOnly;         // no gas: re-emission from the dust CDF, no line packets
    const T* totalLines;  // [C+1]
    // tallies
    double* J;            // [B][C+1]
    double* escaped;      // [B][C+1]
    double* Qphot;        // [1]
    unsigned long long* counters;  // [MOC_CNT_N]
    // source
    T star[3];
    int starIdx[3];       // 0-based
    T deltaE;
    // work
    int64_t n;
    uint64_t seed;
    uint32_t ctr_base;
    const uint32_t* ctrDev;  // non-null: the RNG counter base is *ctrDev + ctr_base (graph replays)
    int emitBatch;           // lanes of a wave that wait to emit together
    int segsPerTrip;         // segments per flying lane and trip through the loop
    // batched grids: packet ids gridFirst[k] .. gridFirst[k+1]-1 belong to grid k
    // (RNG packet id (k << 32) + local id); nGrids = 0 for a single grid
    int nGrids;
    const MocGridView<T>* grids;  // [nGrids]
    const int64_t* gridFirst;     // [nGrids + 1]
    // nested sub-grids (nested = 1): view 0 is the mother grid, whose active
    // map holds -g for the cells sub-grid g (view g - 1) covers; the packets
    // start in view nestStart; mir = the mother's lower faces (mirror planes)
    int nested, nestStart;
    T mir[3];
    // J tallies of a hot block of (bin, cell) keys -- a box of cells around
    // the star times the bins that carried the most path there -- summed in
    // LDS per workgroup as FP32 partials, then added to J in FP64 (the mixed
    // form on a single grid; nHotC = 0: every tally is a global atomic)
    int nHotC, nHotB;
    int hotLo[3], hotN[3];   // the box: cell indices lo .. lo + n - 1 per axis (nHotC = product)
    const int* hotBin;       // [B] slot of a bin, or -1
    double* hotS;            // [B] work: J of each bin summed over the hot cells
    const int* hotCellList;  // [nHotC] cell of each slot
    const int* hotBinList;   // [nHotB] bin of each slot, or -1
    float* hotPart;          // [blocks][nHotB * nHotC] the workgroups' partial sums
    int hotBlocks;           // rows of hotPart
};

// Opacity kernel view (opacity.hip)
struct MocOpacityArgs {
    int C1, nbins, nSpecies, nElemUsed, nstages, lgCompton;
    const int* specElem;   // [S] index into elemZ
    const int* specIon;    // [S] stage 1..nstages
    const int* elemZ;      // [NE]
    const double* abund;   // [NE]
    const double* sigma;   // [S][B]
    const double* brems;   // [B]
    const double* nuArray; // [B]
    const double* log10nu; // [B]
    const double* KNsigmaT;// [B]
    const double* ionDen;  // [NS][NE][C1]
    const double* Ne;
    const double* Te;
    const double* Hden;
    double* opacity;       // [B][C1]
    float* opacity32;      // [B][C1]: mixed form (gas runs), the FP32 kernel writes this plane instead
    // cells whose Te and Ne moved by < 0.5% since the last balance keep their
    // opacity (iterateMC, iteration_mod.f90:126-153); NULL: every cell
    const double* TeOld;
    const double* NeOld;
};

// Balance kernel view (balance.hip)
#define MOC_MAXE 30     // elements
#define MOC_MAXS 31     // ionisation stages
#define MOC_MAXLEV 27   // levels of an n-level ion
// n-level ions of more than 5 levels are solved by teams of lanes, in classes
// of at most 8, 10, 13, 16, 17 and 27 levels (nlevel.h)
constexpr int kNlevTeams = 6;

struct MocBalanceArgs {
    int nCells, NE, NS, maxLev, lgCompton, nShells, nSpec;
    int fitF32;            // mixed-precision form: rate and cooling fits in FP32
    double XHILimit;
    const int* elemZ;      // [NE]
    const double* abund;   // [NE]
    const int* elemSlot;   // [31] Z -> slot or -1
    const int* specElem;   // [S] slot
    const int* specIon;    // [S]
    const int* specOf;     // [NE*32] (slot, stage) -> species or -1
    const int* specShell;  // [S+1] shell range of each species
    const int* shellNum;   // [nShells] shell number (1 = K)
    const int* shInfo;     // [nShells] Auger electron count | (1 + N I/O I outer-shell slot) << 8
    const int* shAug;      // [nShells] index of the shell's auger[...][0] entry
    const int* outShell;   // [31*32] (Z, nElec) -> outer shell
    const int* statW;      // [31*32] g0 (Z, nElec)
    const int* statW1;     // [31*32] g1
    const double* rrb;     // [6][31][31]
    const double* drb_c;   // [9][31][31]
    const double* drb_e;   // [9][31][31]
    const double* rrec;    // [2][31][31]
    const double* rnew;    // [4][31][31]
    int nDiel, nAldro;
    const double* diel;    // [nDiel][8]
    const double* aldro;   // [nAldro][6]
    const double* cion_cf; // [5][31][31]
    const double* cion_th; // [31][32]
    const int* nauger;     // [31][32][11]
    const double* auger;   // [31][32][11][11]
    const int* chexIdx;    // [31*32] -> row of chexTab or -1
    const double* chexTab; // [n][4]
    const double* chexU;   // [31*32] unscaled charge-exchange coefficient (0 without a row)
    const int* dielIdx;    // [31*32][2] dielectronic row in force below / at or above 2e4 K, or -1
    const int* aldroIdx;   // [31*32] Aldrovandi & Pequignot row or -1
    // heavy-element chains as gather lists: task t sums the terms
    // htBeg[t] .. htBeg[t+1]-1 into out(htSlot[t], htMx[t]) in the sequential
    // walk's order (term code: kind | rev. charge-exchange slot << 2 | source
    // stage << 4 | shell << 9; htAug: Auger yield of kind-2 terms)
    int nHTask;
    const int* htSlot;
    const int* htMx;
    const int* htBeg;
    const int* htCode;
    const double* htAug;
    const int* atomIdx;    // [31*32]
    const int64_t* atomMeta;  // [nAtoms][8]
    const double* atomAtp;
    const double* atomLogT;
    const double* atomG;
    const double* atomE;
    const double* atomA;
    const double* atomQom;    // [nT][nLev][nLev] per atom
    const double* atomQ2;     // spline second derivatives, same layout
    const int* atomRecKind;   // [nAtoms] recombination feed: 0 none, 1 fit per level, 2 log-T cubic
    const double* atomRec;    // [levels][11] per-level recombination coefficients
    const int* atomRecValid;  // [levels]
    int nNlv;                 // ions with n-level data, in (element slot, stage) order
    const int* nlvSlot;
    const int* nlvIon;
    const int* nlvAtom;
    // ions whose level count (min(file levels, maxLev)) exceeds 5 are solved
    // by teams of lanes with the matrix in registers (nlevel_team): the
    // entries k of the nlv list with [0] 6-8 levels (teams of 8 lanes), [1]
    // 9-16 (teams of 16), [2] 17-27 (teams of 32); the others one ion per lane
    int nNlvT[kNlevTeams];
    const int* nlvT[kNlevTeams];
    int colHit, colComp;      // rate columns of the hit test and the Compton sums
    int ncol;                 // rate columns per cell
    int nX;                   // entries of a packed per-cell stage row (sum over elements
                              // of min(Z+1, NS) + 1)
    int xoff[MOC_MAXE + 1];   // start of each element in a packed row
    // grains with photoelectric heating and gas-grain collisions (nGS = 0: none)
    int nGS, nGA, B;
    const double* gAbun;      // [nGS]
    const double* gWeight;    // [nGA]
    const double* gRadius;    // [nGA] um
    const double* gVn;        // [nGS] Ryd
    const double* gMsurf;     // [nGS] amu
    const double* gCabs;      // [nGS][nGA][B]
    const double* gIp;        // [31*32] Ryd released on recombination of (Z, stage)
    const double* gNu;        // [B]
    const double* J;          // [B][C1] normalised J
    const double* Tdust;      // [C1][nGA+1][nGS+1]
    const double* Ndust;      // [C1]
    double* collD;            // [nGS*nGA][C1] out: gas-grain collision heating of the grains
    double* peD;              // [nGS*nGA][C1] out: photoelectric cooling of the grains
    const double* R;          // [nLaunch][ncol] rates of cells cell0 .. cell0 + nLaunch - 1
    // band form of the rate weights: W's columns sorted by their nonzero bins
    // (column n is rate column ratePerm[n]); column tile t spans bins
    // rateLo[t] .. rateHi[t]
    const int* ratePerm;
    const int* rateLo;
    const int* rateHi;
    int cell0, nLaunch;       // the cells of one launch (a chunk of the grid)
    double* Te;
    double* Ne;
    const double* Hden;
    double* ionDen;
    int* lgConverged;
    int* lgBlack;
    int* nIterT;
    unsigned long long* counters;  // noHit, noIonBal, noTeBal
};

// Emission kernel view (emission.hip)
#define MOC_EM_MAXLIM 64   // points of a gamma fit table
#define MOC_EM_MAXGRAIN 256 // (species, size) pairs of gas + dust emission
#define MOC_EM_MAXFB 1024   // heavy-ion free-bound entries
// the skip test of iterateMC for one cell
__device__ __forceinline__ bool moc_unchanged(const double* TeOld, const double* NeOld,
                                              const double* Te, const double* Ne, int cell) {
    if (!TeOld) return false;
    return fabs(Te[cell] - TeOld[cell]) / TeOld[cell] < 0.005 &&
           fabs(Ne[cell] - NeOld[cell]) / NeOld[cell] < 0.005;
}

struct MocEmissionArgs {
    int nCells, C1, B, NE, NS;
    const double* nu;        // [B]
    const double* widFlx;    // [B]
    const double* log10nu;   // [B]
    const double* abund;     // [NE]
    int ntk;
    const double* tk;        // [ntk]
    int nlim[3];             // H I, He I, He II
    const int* edge[3];      // [nlim] 1-based bin pointers
    const int* gpair;        // [3][B] fit pair (1-based) that covers each bin, 0 for none
    // grains of gas + dust runs (nDS = 0 without): setDiffusePDF :1685-1752
    int nDS, nDA;
    const double* dAbun;     // [nDS]
    const double* dWeight;   // [nDA]
    const double* dTsub;     // [nDS]
    const double* dCabs;     // [nDS][nDA][B]
    const double* Tdust;     // [C+1][nDA+1][nDS+1]
    const double* Ndust;     // [C+1]
    const double* Hden;      // [C+1]
    // quantum heating of the grains (qActive: qIdx, qT, qP as in MocDustArgs)
    int qActive, nQ, nTb;
    const int* qIdx;
    const double* qT;
    const double* qP;
    const double* lg[3];     // [ntk][nlim] log10 gamma
    int HIP, HeIP, HeIIP;    // threshold bins (He: 0 when off)
    const double* xsHI;
    const double* xsHeI;
    const double* xsHeII;
    int hSlot, heSlot;       // element slots of H and He (-1 when off)
    int nFb;                 // heavy-ion free-bound entries (Z, then ion, ascending)
    const int* fbSlot;
    const int* fbIon;
    const int* fbLo;
    const int* fbHi;
    const int64_t* fbOff;
    const double* fbXs;
    const double* fbG;       // g0/g1
    const int* he2Row;       // [B]
    const double* hei2phot;  // [41][2]
    int nHyd;                // hydrogenic line emitters, Z ascending
    const int* hydZ;
    const int* hydSlot;
    int nFor;                // ions with n-level data
    const int* forSlot;
    const int* forIon;
    const int* forAtom;
    const int* forTop;       // min(Z+1, nstages)
    int nForT[kNlevTeams];   // entries k of the for list solved by teams of lanes, per
    const int* forT[kNlevTeams];  // level class (nlevel_team_class)
    const int* hlMap;        // [9][12]
    const double* hlDens;    // [files][14]
    const double* hlLines;   // [files][14][16][9]
    const double* hbA;       // [9][4]
    const double* hbB;
    const double* rbEdge;    // [2][3][26]
    const double* r2aEdge;   // [3][18]
    const double* heiRec;    // [34][3][4]
    const double* heiiLyman; // [4][2]
    int j2TsP;
    int lymP[4];
    const double* Te;        // [C1]
    const double* Ne;
    const double* ionDen;    // [NS][NE][C1]
    double* recPDF;          // [C1][B]
    float* recPDF32;         // [C1][B]: mixed form (gas runs), the continuum in FP32 writes this row instead
    double* totalLines;      // [C1]
    double* lineSum;         // [C1] line emissivity of each cell (lines kernel -> CDF kernel)
    const double* TeOld;     // skip test as for the opacity (iteration_mod.f90:318-349); NULL: none
    const double* NeOld;
    MocBalanceArgs bal;      // n-level atoms
};

struct MocEmission {
    bool set = false;
    MocEmissionArgs a{};          // device pointers (grid-independent part)
    std::vector<void*> owned;
};

// Grain tables of a context (moc_set_dust)
struct MocDust {
    bool set = false;
    int nS = 0, nA = 0, nT = 0, lgGas = 0;
    double XHILimit = 0.;
    std::vector<void*> owned;   // device allocations of this table set
    double *d_abun = nullptr, *d_weight = nullptr, *d_tsub = nullptr;
    // lgEquivalentTau: the first iteration runs without dust opacity (its
    // escaped SED is the unextinguished one); d_tsub0 holds zero
    // sublimation temperatures, which leave no grain in the opacity
    int equivTau = 0;
    bool firstIter = false;
    double* d_tsub0 = nullptr;
    double *d_Cabs = nullptr, *d_Csca = nullptr, *d_emInt = nullptr;
    double* d_widFlx = nullptr; // [B]
    int lgPE = 0;               // grain charge, photoelectric heating, gas-grain collisions
    double *d_radius = nullptr, *d_vn = nullptr, *d_msurf = nullptr, *d_ip = nullptr;
    double* d_W = nullptr;      // [B][nS*nA+1] Cabs/pi and the hit column
    double* d_R = nullptr;      // [C+1][nS*nA+1] absorption integrals, a row per cell
    unsigned long long* d_counters = nullptr;  // [2]
    // quantum heating (qHeat): the grains below minaQHeat, their distributions
    int lgQHeat = 0, nTb = 0, nQ = 0;
    double minaQHeat = 0., minConvQHeat = 0., Tmax = 0.;
    bool qActive = false;       // this iteration's emission uses the distributions
    int* d_kind = nullptr;      // [nS] 1 silicate, 2 carbon, 0 other
    int* d_qIdx = nullptr;      // [nS*nA] slot of a grain among the nQ, or -1
    int* d_qG = nullptr;        // [nQ] grain (s * nA + ai, 0-based) of each slot
};

// Dust kernel view (dust.hip)
struct MocDustArgs {
    int nCells, B, nS, nA, nT, dustOnly;
    const double* abun;     // [nS]
    const double* weight;   // [nA]
    const double* tsub;     // [nS]
    const double* Cabs;     // [nS][nA][B]
    const double* Csca;     // [nS][nA][B]
    const double* emInt;    // [nS][nA][nT]
    const double* nu;       // [B]
    const double* widFlx;   // [B]
    const double* Ndust;    // [C+1]
    double* Tdust;          // [C+1][nA+1][nS+1]
    double* absOpac;        // [B][C+1]
    double* scaOpac;        // [B][C+1]
    const double* gasOpac;  // [B][C+1] or null (dust only)
    double* opacity;        // [B][C+1]
    double* pdf;            // [C+1][B] dust-only emission CDF
    const double* R;        // [C+1][nS*nA+1] absorption integrals, last column the hit test
    const double* collD;    // [nS*nA][C+1] gas-grain collision heating, or null
    const double* peD;      // [nS*nA][C+1] photoelectric cooling, or null
    int* lgConverged;
    int* lgBlack;
    unsigned long long* counters;  // noHit, converged
    double XHILimit;
    // quantum heating: with qActive, grain g with qIdx[g] = q >= 0 emits from
    // its temperature distribution qT/qP [C+1][nQ][nTb] instead of its Tdust
    int qActive, nQ, nTb;
    const int* qIdx;        // [nS*nA]
    const int* qG;          // [nQ]
    const int* kind;        // [nS]
    const double* radius;   // [nA] um
    double Tmax;
    const double* J;        // [B][C+1] normalised J
    double* qT;             // [C+1][nQ][nTb]
    double* qP;             // [C+1][nQ][nTb]
    double* qWork;          // [qBlocks][nTb][nTb]
    int qBlocks;
};

// getFlux('blackbody') (continuum_mod.f90:358-400): the Planck function over
// h in erg/(cm^2 s Hz sr) / h, with its Wien and Rayleigh-Jeans branches
__host__ __device__ inline double moc_planck_h(double nu, double T) {
    const double constant = 0.5250229 / 6.6262e-27;
    const double x = 157893.94 * nu / T;
    if (x > 86.) return constant * nu * nu * nu * exp(-x);
    const double den = exp(x) - 1.;
    if (den <= 0.) return 3.32154e-6 * nu * nu * T / 6.6262e-27;
    return constant * nu * nu * nu / den;
}

struct MocGrid {
    bool attached = false;
    int nx = 0, ny = 0, nz = 0, nCells = 0, sym = 0, plane = 0;
    // a batch of independent grids (nGrids > 1): per-grid sizes, first cell,
    // offsets into the concatenated axes and active maps; grid k is lit by
    // star k
    int nGrids = 1;
    // nested sub-grids (moc_grid_attach_nested): grid 0 the mother, whose
    // active map holds -g for the cells that grid g (1-based) covers; the
    // source starts in grid nestStart (0-based) at cell nestStartIdx (0-based)
    int nested = 0, nestStart = 0;
    int nestStartIdx[3] = {0, 0, 0};
    bool batch() const { return nGrids > 1 && !nested; }
    std::vector<int> gnx, gny, gnz, gbase, gdense;
    std::vector<size_t> goffx, goffy, goffz, goffa;
    std::vector<double> gR_out;
    MocGridView<double>* d_view64 = nullptr;
    MocGridView<float>* d_view32 = nullptr;
    int64_t* d_gridFirst = nullptr;
    bool dense = false;   // every cell active, numbered z fastest, then y, then x
    double R_out = 0.0;
    std::vector<double> ax, ay, az;   // host copies
    std::vector<int> active;
    // device arrays (FP64 master copies)
    double *d_ax = nullptr, *d_ay = nullptr, *d_az = nullptr;
    double *d_vx = nullptr, *d_vy = nullptr, *d_vz = nullptr;
    int* d_active = nullptr;
    double* d_opacity = nullptr;     // [B][C+1]
    double* d_Ne = nullptr;          // [C+1]
    double* d_Te = nullptr;          // [C+1]
    double* d_Hden = nullptr;        // [C+1]
    double* d_TeOld = nullptr;       // [C+1] Te, Ne before the last balance
    double* d_NeOld = nullptr;
    double* d_recPDF = nullptr;      // [C+1][B]
    double* d_totalLines = nullptr;  // [C+1]
    double* d_lineSum = nullptr;     // [C+1] line emissivity, between the two emission kernels
    double* d_vol = nullptr;         // [C+1] cell volumes [1e45 cm^3] (getVolume)
    double* d_J = nullptr;           // [B][C+1]
    double* d_escaped = nullptr;     // [B][C+1]
    // FP32 shadow copies for the mixed path (refreshed when dirty)
    float *f_ax = nullptr, *f_ay = nullptr, *f_az = nullptr;
    float *f_vx = nullptr, *f_vy = nullptr, *f_vz = nullptr;
    float* f_opacity = nullptr;
    float* f_Ne = nullptr;
    float* f_recPDF = nullptr;
    float* f_totalLines = nullptr;
    bool f32_dirty = true;
    // the mixed form's FP32 kernels write the FP32 planes directly: while set,
    // the FP32 plane holds the current opacity / diffuse CDF and the FP64 one
    // is brought up to date before the host reads or writes it
    bool op32_master = false, pdf32_master = false;
    // quantum heating distributions and the solver's per-workgroup work rows
    double *d_qT = nullptr, *d_qP = nullptr, *d_qWork = nullptr;
    int qBlocks = 0;
    // hot block of the J tallies (MocTransportArgs::nHotC): the maps and the
    // per-workgroup partial sums, set up at the first mixed transport
    int nHotC = 0, nHotB = 0, hotBlocks = 0;
    bool hotTried = false;
    int hotLo[3] = {0, 0, 0}, hotN[3] = {0, 0, 0};
    int *d_hotBin = nullptr, *d_hotCellList = nullptr, *d_hotBinList = nullptr;
    float* d_hotPart = nullptr;
    double* d_hotS = nullptr;
    // grains (moc_set_dust)
    double* d_Ndust = nullptr;       // [C+1]
    double* d_Tdust = nullptr;       // [C+1][nSizes+1][nSpecies+1] (species fastest)
    double* d_absOpac = nullptr;     // [B][C+1]
    double* d_scaOpac = nullptr;     // [B][C+1]
    double* d_gasOpac = nullptr;     // [B][C+1] gas kappa of gas + dust runs
    double* d_collD = nullptr;       // [nS*nA][C+1] photoelectric runs: collision heating
    double* d_peD = nullptr;         // [nS*nA][C+1] and photoelectric cooling of the grains
    float* f_scaOpac = nullptr;
    double* d_ionDen = nullptr;      // [NS][NE][C+1]
    int* d_conv = nullptr;           // [C+1] lgConverged
    int* d_black = nullptr;          // [C+1] lgBlack
    int* d_nIterT = nullptr;         // [C+1]
    double* d_rates = nullptr;       // [rates_chunk][ncol]: the rates of one chunk of cells
    int rates_cols = 0;
    int rates_chunk = 0;             // cells per balance launch (bounds the rate buffer)
};

struct MocBalance {
    bool set = false;
    MocBalanceArgs a{};           // device pointers (grid-independent part)
    int ncol = 0;
    double* d_W = nullptr;        // [B][ncol]
    std::vector<void*> owned;     // device allocations of this table set
    unsigned long long* d_counters = nullptr;
    std::vector<int> atomIdx;     // host copy of (Z, stage) -> n-level ion
    std::vector<int> atomLevels;  // level count of each n-level ion as solved (maxLev cut)
};

struct MocAtomic {
    bool set = false;
    int nElemUsed = 0, nstages = 0, nSpecies = 0;
    std::vector<int> elemZ, specElem, specIon;
    std::vector<double> abund;
    int *d_elemZ = nullptr, *d_specElem = nullptr, *d_specIon = nullptr;
    double *d_abund = nullptr, *d_sigma = nullptr, *d_brems = nullptr, *d_log10nu = nullptr;
};

// Runtime tuning parameters (moc_set_tuning)
struct MocTuning {
    int emitBatch = 16;         // transport: lanes that wait to emit together
    int segsPerTrip = 16;       // transport: segments per flying lane and loop trip
    int transportWaves = 4096;  // transport: resident waves (the launch's lanes / 64)
    int hotJ = 1;               // mixed transport: the hot block of J tallies in LDS
};

struct moc_ctx {
    std::string err;
    MocTuning tune;
    // graph replay of moc_iterate (moc_set_graph): the iteration captured once
    // per seed and replayed; the iteration number reaches the transport's RNG
    // counter through d_ctr
    bool graphOn = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphSeed = 0;
    double lastConvPercent = 0.;  // converged cells of the last moc_iterate [%] (quantum heating)
    // planes the captured iteration writes in FP32 (their FP64 copies go stale at each replay)
    bool graphOp32 = false, graphPdf32 = false;
    uint32_t* d_ctr = nullptr;
    uint32_t ctrHost = 0;  // source of the d_ctr upload, alive until the replay is waited for
    int precision = MOC_PREC_FP64;
    int nbins = 0, nAngles = 180, lgCompton = 1, nStars = 0;
    double ionEdge1 = 0.0;
    std::vector<double> nuArray, widFlx, Lstar, nPhotons, starPos;  // nPhotons: current packets per star
    std::vector<int> starIdx;
    // device tables
    double *d_nu = nullptr, *d_KNsigmaT = nullptr, *d_KNcdf = nullptr, *d_Pcomp = nullptr;
    double* d_sourceCDF = nullptr;  // [nStars][B]
    float *f_nu = nullptr, *f_KNsigmaT = nullptr, *f_KNcdf = nullptr, *f_Pcomp = nullptr;
    float* f_sourceCDF = nullptr;
    double* d_Qphot = nullptr;
    unsigned long long* d_counters = nullptr;
    MocGrid grid;
    MocAtomic atomic;
    MocBalance balance;
    MocEmission emission;
    MocDust dust;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double* d_scratch = nullptr;  // the device's fixed buffer (allocation-free entry point), not owned
};

// launchers (transport.hip)
int moc_launch_transport_f64(moc_ctx* ctx, const MocTransportArgs<double>& a, float* ms);
int moc_launch_transport_f32(moc_ctx* ctx, const MocTransportArgs<float>& a, float* ms);
// workgroup size and LDS of the mixed transport with a hot J block of nHot keys
int moc_transport_hot_block();
size_t moc_transport_hot_lds(int nx, int ny, int nz, int nHot);
void moc_convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);
void moc_convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s);
void moc_scale_f64(double* p, int64_t n, double f, hipStream_t s);
void moc_fill_zero_u64(unsigned long long* p, int64_t n, hipStream_t s);
void moc_copy_f64(const double* src, double* dst, int64_t n, hipStream_t s);
int moc_launch_opacity(moc_ctx* ctx, const MocOpacityArgs& a);
int moc_launch_emission(moc_ctx* ctx, const MocEmissionArgs& a);
int moc_launch_dust_opacity(moc_ctx* ctx, const MocDustArgs& a);
int moc_launch_dust_pdf(moc_ctx* ctx, const MocDustArgs& a);
int moc_launch_qheat(moc_ctx* ctx, const MocDustArgs& a);
size_t moc_qheat_lds(int B, int nTb);
int moc_launch_dust_temp(moc_ctx* ctx, const MocDustArgs& a, const double* J, const double* W, int ncol,
                         double* R);
// R[col][cell] = sum_nu J[nu][cell] W[nu][col] (k_rates, balance.hip)
void moc_launch_rates(moc_ctx* ctx, const double* J, int C1, const double* W, int ncol, double* R);
int moc_rate_tile_cols();  // columns per k_rates tile
int moc_rate_tile_bins();  // bins per k_rates stage
int moc_launch_volavg(moc_ctx* ctx, const double* ionDen, const double* Te, const double* Ne,
                      const double* H, const double* vol, const int* black, int C1, int nPairs,
                      int nChunks, double* part);
int moc_launch_sed(moc_ctx* ctx, const double* escaped, int C1, int B, double* sed);
// sums[2p], sums[2p+1] = the volume-average partials of pair p added over the chunks in order
int moc_launch_volavg_sum(moc_ctx* ctx, const double* part, int nPairs, int nChunks, int C1,
                          double* sums);
// Forbidden-line luminosities of ions k0 .. k0+nk-1 (nk <= moc_line_lum_group(L))
// and, with k0 == 0, Hbeta: part[block][nk*L*L + 1].
int moc_line_lum_group(const MocEmissionArgs& a, int L);
int moc_launch_line_lum(moc_ctx* ctx, const MocEmissionArgs& a, const double* vol, const int* black,
                        const double* H, int nChunks, int L, int k0, int nk, double* part);
// the balance over every cell, in chunks of `chunk` cells whose rates fit R ([ncol][chunk])
int moc_launch_balance(moc_ctx* ctx, const MocBalanceArgs& a, const double* J, const double* W,
                       int ncol, double* R, int chunk);

/*
 * moc_oracle_balance.h -- ORACLE (test infrastructure only): CPU, FP64
 * restatement of MOCASSIN's updateCell (update_mod.f90:15-4040).
 */
#ifndef MOC_ORACLE_BALANCE_H
#define MOC_ORACLE_BALANCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_balance_tables {
    int nbins, nstages, nElemUsed, maxLev, lgCompton;
    const int32_t* elemZ;       /* [NE] */
    const double* abund;        /* [NE] */
    const double* nuArray;      /* [B] */
    int nShells;                /* per-shell cross-section stack (elementP ranges) */
    const int32_t *shZ, *shIon, *shNum, *shLo, *shHi;
    const int64_t* shOff;
    const double* xs;
    int secIonP, cRecoilHP, cRecoilHeP; /* 1-based locate() pointers */
    const double* xSecRecoil;   /* [B] */
    const double* rrb;          /* [6][31][31] Badnell RR  (A,B,T0,T1,C,T2)[Z][N] */
    const double* drb_c;        /* [9][31][31] Badnell DR c_i */
    const double* drb_e;        /* [9][31][31] Badnell DR E_i */
    const double* rrec;         /* [2][31][31] */
    const double* rnew;         /* [4][31][31] */
    int nDiel;
    const double* diel;         /* [nDiel][8] elem n a b c d f g */
    int nAldro;
    const double* aldro;        /* [nAldro][6] elem n a b t0 t1 */
    const double* cion_cf;      /* [5][31][31] Voronov (.,Z,nElec) */
    const double* cion_th;      /* [31][32] cionTherm (Z, ion) */
    const int32_t* nauger;      /* [31][32][11] */
    const double* auger;        /* [31][32][11][11] */
    const int32_t* atom_idx;    /* [31][32] n-level ion index or -1 */
    const int64_t* atom_meta;   /* [nAtoms][7] nLev nTemp offT offLev offA offQ iup2p */
    const double* atom_atp;
    const double* atom_logt;
    const double* atom_g;
    const double* atom_e;
    const double* atom_A;
    const float* atom_qom;      /* [nTemp][nLev][nLev] per atom */
    const int32_t* atom_rec_kind;  /* [nAtoms] 0 none, 1 (br,z,a_r,a_d), 2 log-T cubic */
    const double* atom_rec;        /* [levels][11] recombination data per level */
    const int32_t* atom_rec_valid; /* [levels] */
} oracle_balance_tables;

typedef struct oracle_balance_state {
    int nCells;
    const double* J;            /* [B][C+1] normalised Jste */
    double* Te;                 /* [C+1] in/out */
    double* Ne;                 /* [C+1] in/out */
    const double* Hden;         /* [C+1] */
    double* ionDen;             /* [NS][NE][C+1] in/out */
    int32_t* lgConverged;       /* [C+1] out */
    int32_t* lgBlack;           /* [C+1] out */
    int32_t* nIterT;            /* [C+1] out (T iterations), may be NULL */
    double XHILimit;
    double noHit, noIonBal, noTeBal; /* out */
    /* grains of gas + dust runs with photoelectric heating and gas-grain
     * collisions (update_mod.f90:484-884); gS = 0 without */
    int gS, gA;
    const double* gAbun;    /* [gS] grainAbun */
    const double* gWeight;  /* [gA] grainWeight */
    const double* gRadius;  /* [gA] um */
    const double* gVn;      /* [gS] grainVn [Ryd] */
    const double* gMsurf;   /* [gS] MsurfAtom [amu] */
    const double* gCabs;    /* [gS][gA][B] */
    const double* Tdust;    /* [C+1][gA+1][gS+1] */
    const double* Ndust;    /* [C+1] */
    double* collD;          /* [gS*gA][C+1] out: gasDustColl_d of the last T-iteration */
    double* peD;            /* [gS*gA][C+1] out: photoelHeat_d of the last T-iteration */
    double* grainPot;       /* [gS*gA][C+1] out, may be NULL */
} oracle_balance_state;

int oracle_update(const oracle_balance_tables* t, oracle_balance_state* s, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

#ifndef MOC_ORACLE_BALANCE_PIECES
#define MOC_ORACLE_BALANCE_PIECES
#ifdef __cplusplus
extern "C" {
#endif
void oracle_alpha(const oracle_balance_tables* t, double Te, double* out);
void oracle_heavy_chain(const oracle_balance_tables* t, int z, double Te, double Ne, double H, double XH0,
                        const double* alpha, const double* cI1, const double* pI1, int damp, double* X);
void oracle_cooling_parts(const oracle_balance_tables* t, double Te, double Ne, const double* X,
                          double* parts);
void oracle_cell_heat_cool(const oracle_balance_tables* t, const oracle_balance_state* st, int cell,
                           double Te, double* heat, double* cool);
double oracle_nlevel_cooling(const oracle_balance_tables* t, int atom, double Te, double Ne,
                             double scale);
/* equilibrium (emission_mod.f90:2505-3028): sum of the line emissivities of
 * one ion times scale*unit that exceed 1e-35; with rec != 0 the upper levels
 * are fed by recombination from the next stage (ionDenUp = X(i+1)/X(i)). */
double oracle_nlevel_lines(const oracle_balance_tables* t, int atom, double Te, double Ne,
                           double ionDenUp, int rec, double scale, double unit);
/* the same, with every line's emissivity fLineEm(i, j) (lower i, upper j,
 * 1-based levels) also stored in lineEm[(i-1) * ld + (j-1)] */
double oracle_nlevel_line_em(const oracle_balance_tables* t, int atom, double Te, double Ne,
                             double ionDenUp, int rec, double scale, double unit, double* lineEm,
                             int ld);
#ifdef __cplusplus
}
#endif
#endif

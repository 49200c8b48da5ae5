/* ORACLE -- test infrastructure only: grain routines (see moc_oracle_dust.c). */
#ifndef MOC_ORACLE_DUST_H
#define MOC_ORACLE_DUST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_dust_in {
    int nCells, B, nS, nA, nT, dustOnly;
    const double* abun;     /* [nS] grainAbun */
    const double* weight;   /* [nA] grainWeight */
    const double* tsub;     /* [nS] TdustSublime */
    const double* Cabs;     /* [nS][nA][B] */
    const double* Csca;     /* [nS][nA][B] */
    const double* emInt;    /* [nS][nA][nT] dustEmIntegral */
    const double* nu;       /* [B] */
    const double* widFlx;   /* [B] */
    const double* Ndust;    /* [C+1] */
    double XHILimit;
    const double* collD;    /* [nS*nA][C+1] gasDustColl_d, or NULL */
    const double* peD;      /* [nS*nA][C+1] photoelHeat_d, or NULL */
    /* quantum heating (qIdx NULL: off): grain g = (s-1) nA + ai-1 with qIdx[g] = q >= 0
     * emits from its distribution qT/qP [C+1][nQ][nTb] */
    int nQ, nTb;
    const int* qIdx;
    const double* qT;
    const double* qP;
} oracle_dust_in;

double oracle_planck_h(double nu, double T);
void oracle_dust_em_integral(const double* Cabs, const double* nu, const double* widFlx, int B,
                             const double* temps, int nT, double* out);
/* Tdust: [C+1][nA+1][nS+1]; planes [B][C+1]; gasOpac may be NULL */
void oracle_dust_opacity(const oracle_dust_in* in, const double* Tdust, const double* gasOpac,
                         double* opacity, double* absOpac, double* scaOpac);
/* pdf: [C+1][B] */
void oracle_dust_pdf(const oracle_dust_in* in, const double* Tdust, double* pdf);
/* J: [B][C+1] normalised; Tdust updated in place; counters = noHit, converged */
int oracle_dust_temperature(const oracle_dust_in* in, const double* J, double* Tdust,
                            int32_t* lgConverged, int32_t* lgBlack, double counters[2]);

/* quantum heating of one grain in one cell (qHeat, emission_mod.f90:1993-2504) */
typedef struct oracle_qheat_in {
    int B, nTbins;
    double Tmax;            /* K */
    const double* nu;       /* [B] Ryd */
    const double* widFlx;   /* [B] Ryd */
    const double* J;        /* [B] the cell's normalised J */
    const double* Cabs;     /* [B] the grain's absorption cross-section */
    double radius;          /* um */
    int kind;               /* 1 silicate ('S'), 2 carbon ('C'), other: none */
} oracle_qheat_in;
/* temp, pss: [nTbins] the temperature bins and their probabilities (pss of the
 * last bin 0); work: oracle_qheat_work(B, nTbins) doubles; returns 0 for a
 * kind without a distribution (temp and pss 0) */
int oracle_qheat(const oracle_qheat_in* q, double tg, double* temp, double* pss, double* work);
size_t oracle_qheat_work(int B, int nTbins);

#ifdef __cplusplus
}
#endif
#endif

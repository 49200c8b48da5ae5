/*
 * moc_oracle.c -- ORACLE (test infrastructure only): a plain-C, FP64
 * restatement of MOCASSIN's packet transport, written in the shape of the
 * reference's Fortran (1-based indices, one packet after another), for the
 * parity tests and the CPU baseline.  Nothing in the product links it.
 *
 * Reference: rwesson/mocassin_xray source/photon_mod.f90
 *   energyPacketDriver :28-414, energyPacketRun :418-741,
 *   initPhotonPacket :744-896, getNu2 :928-972, newPhotonPacket :976-1756,
 *   pathSegment :1758-4085, comptonScatter :6805-6844;
 *   vector_mod.f90:303-314 (randomUnitVector);
 *   interpolation_mod.f90:49-106 (locate).
 * Scope: one Cartesian grid (whole, octant or plane-parallel slab), grains
 * as scatterers and absorbers, no sub-grids, no 1D mode, no fluorescence,
 * nAngleBins = 0.  The clock-seeded random_number of
 * the reference (photon_mod.f90:97-114) is replaced by Philox4x32-10 streams
 * keyed by (seed, packet, draw, ctr_base), shared in definition (not in code)
 * with the GPU path, so single packet histories can be compared.
 */
#include "moc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_RECURSION_LIMIT 100000 /* constants_mod.f90:70 */
#define ORC_SAFE_LIMIT 500000      /* photon_mod.f90:1873-1878 */
#define ORC_SAFE_LIMIT_PLANE 5000  /* the same lines, plane-parallel illumination */
#define ORC_PI 3.141592654         /* constants_mod.f90:45 */

/* ---- Philox4x32-10, Salmon et al. SC'11 -------------------------------- */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t ka = key[0], kb = key[1];
    int r;
    for (r = 0; r < 10; r++) {
        uint64_t a = (uint64_t)0xD2511F53u * (uint64_t)x0;
        uint64_t b = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
        uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ ka;
        uint32_t y1 = (uint32_t)b;
        uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ kb;
        uint32_t y3 = (uint32_t)a;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        ka += 0x9E3779B9u;
        kb += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

typedef struct {
    uint32_t ctr[4];
    uint32_t key[2];
} orng_t;

static void orng_init(orng_t* g, uint64_t seed, uint64_t packet, uint32_t base) {
    g->key[0] = (uint32_t)(seed & 0xffffffffu);
    g->key[1] = (uint32_t)(seed >> 32);
    g->ctr[0] = 0;
    g->ctr[1] = (uint32_t)(packet & 0xffffffffu);
    g->ctr[2] = (uint32_t)(packet >> 32);
    g->ctr[3] = base;
}

/* random_number(): uniform on [0,1) from 53 bits of one Philox block */
static double orng_u(orng_t* g) {
    uint32_t o[4];
    oracle_philox4x32_10(g->ctr, g->key, o);
    g->ctr[0]++;
    uint64_t m = ((uint64_t)o[0] << 21) ^ (uint64_t)(o[1] >> 11);
    return (double)m / 9007199254740992.0;
}

/* ---- getNu2 and locate --------------------------------------------------- */
int oracle_getnu2(const double* cdf, int n, double r) {
    /* photon_mod.f90:954-965: nuP=1; scan while random >= probDen(i) */
    int nuP = 1, i;
    for (i = 1; i <= n; i++) {
        if (r >= cdf[i - 1]) nuP = i;
        else break;
    }
    if (nuP < n) nuP = nuP + 1;
    return nuP;
}

static int getnu2_draw(const double* cdf, int n, orng_t* g, int* bad) {
    double r = orng_u(g);
    int k;
    for (k = 0; k < 16 && (r == 0.0 || r == 1.0); k++) r = orng_u(g);
    if (r == 0.0 || r == 1.0) *bad = 1;
    return oracle_getnu2(cdf, n, r);
}

int oracle_locate(const double* xa, int n, double x) {
    int jlow, jup, jm, i;
    if (x > xa[n - 1]) return n;
    if (x < xa[0]) return 0;
    jlow = 0;
    jup = n + 1;
    for (i = 0; i < 10000; i++) {
        if (jup - jlow <= 1) break;
        jm = (jup + jlow) / 2;
        if (xa[jm - 1] == x) return jm;
        if ((xa[n - 1] > xa[0]) == (x > xa[jm - 1])) jlow = jm;
        else jup = jm;
    }
    return jlow;
}

/* ---- packet -------------------------------------------------------------- */
typedef struct {
    double x, y, z;    /* position */
    double u, v, w;    /* direction */
    double nu;
    int nuP;           /* 1-based bin */
    int xP, yP, zP;    /* 1-based cell indices */
    int origin;        /* active id of the emitting cell */
    int lgStellar, lgLine;
    int gP;            /* grid: 0 the mother, k sub-grid k (subs[k-1]) */
} opacket_t;

typedef struct {
    double c[ORC_N];
    double qphot;
} ocount_t;

#define ACT(in, i, j, k) \
    ((in)->active[((i)-1) + (size_t)(in)->nx * (((j)-1) + (size_t)(in)->ny * ((k)-1))])
#define PLANE(in, nuP, cell) ((size_t)((nuP)-1) * (size_t)((in)->nCells + 1) + (size_t)(cell))

/* the grid a packet is in: the mother (gP = 0) or a nested sub-grid */
typedef struct {
    int nx, ny, nz;
    const double *xa, *ya, *za;
    const int32_t* act;
} ogrid_t;

static ogrid_t grid_of(const oracle_transport_in* in, int gP) {
    ogrid_t G;
    if (gP == 0) {
        G.nx = in->nx; G.ny = in->ny; G.nz = in->nz;
        G.xa = in->xAxis; G.ya = in->yAxis; G.za = in->zAxis;
        G.act = in->active;
    } else {
        const oracle_subgrid* s = &in->subs[gP - 1];
        G.nx = s->nx; G.ny = s->ny; G.nz = s->nz;
        G.xa = s->xAxis; G.ya = s->yAxis; G.za = s->zAxis;
        G.act = s->active;
    }
    return G;
}

#define GACT(G, i, j, k) \
    ((G).act[((i)-1) + (size_t)(G).nx * (((j)-1) + (size_t)(G).ny * ((k)-1))])

/* locate in a grid's axis, at least 1, then the next cell above the midpoint:
 * the cell of a position entering a sub-grid (photon_mod.f90:1937-1960), or
 * returning to the mother (:3894-3948), where a position on the midpoint
 * goes to the cell ahead of the packet (dir: its direction on the axis).  The
 * reference writes the return's midpoint xAxis(i) + xAxis(i+1)/2, its
 * parenthesis misplaced; and a packet leaving through a sub-grid face sits on
 * a mother wall up to rounding, so the position alone can put it back into a
 * covered cell (path_segment then steps on along the axis it left by). */
static int nest_cell(const double* a, int n, double x, int entering, double dir) {
    int p = oracle_locate(a, n, x);
    if (p < 1) p = 1;
    if (p < n) {
        const double m = (a[p - 1] + a[p]) / 2.;
        if (entering ? x > m : (x > m || (x == m && dir > 0.))) p = p + 1;
    }
    return p;
}

static void add_to(double* p, double v) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    *p += v;
}

static void escape_tally(const oracle_transport_in* in, oracle_transport_out* out,
                         const opacket_t* pk, ocount_t* cnt) {
    add_to(&out->escaped[PLANE(in, pk->nuP, pk->origin)], in->deltaE);
    cnt->c[ORC_ESCAPED] += 1;
}

/* randomUnitVector (vector_mod.f90:303-314), repeated while a component is 0
 * (photon_mod.f90:866-872) */
static void random_unit_vector(orng_t* g, double* u, double* v, double* w) {
    int rep;
    for (rep = 0; rep < 64; rep++) {
        double r1 = orng_u(g), t, r2, ang;
        *w = 2. * r1 - 1.;
        t = sqrt(1. - (*w) * (*w));
        r2 = orng_u(g);
        ang = ORC_PI * (2. * r2 - 1.);
        *u = t * cos(ang);
        *v = t * sin(ang);
        if (*u != 0. && *v != 0. && *w != 0.) return;
    }
}

/* comptonScatter (photon_mod.f90:6805-6844) */
static void compton_scatter(const oracle_transport_in* in, opacket_t* pk, orng_t* g, int* bad) {
    const double* cdf = in->KNcdf + (size_t)(pk->nuP - 1) * in->nAngles;
    int thP = getnu2_draw(cdf, in->nAngles, g, bad);
    double th = (double)thP * ORC_PI / 180.;
    double r = orng_u(g);
    double w = 2. * r - 1.;
    double t = sqrt(1. - w * w);
    double newNu;
    int newNuP;
    pk->u = t * cos(th);
    pk->v = t * sin(th);
    pk->w = w;
    newNu = pk->nu * in->Pcomp[(size_t)(pk->nuP - 1) * in->nAngles + (thP - 1)];
    newNuP = oracle_locate(in->nuArray, in->nbins, newNu);
    if (newNuP < 1) newNuP = 1; /* the reference would index bin 0 here */
    pk->nu = newNu;
    pk->nuP = newNuP;
}

/* pathSegment (photon_mod.f90:1758-4085); returns 1 when the packet was
 * absorbed (re-emission requested, reRun=1), 0 otherwise. */
static int path_segment(const oracle_transport_in* in, oracle_transport_out* out, opacket_t* pk,
                        orng_t* g, ocount_t* cnt) {
    int gP = pk->gP;
    ogrid_t G = grid_of(in, gP);
    int nx = G.nx, ny = G.ny, nz = G.nz;
    const double *xa = G.xa, *ya = G.ya, *za = G.za;
#define SETG(k) do { gP = (k); G = grid_of(in, gP); nx = G.nx; ny = G.ny; nz = G.nz; \
                     xa = G.xa; ya = G.ya; za = G.za; } while (0)
    double rx = pk->x, ry = pk->y, rz = pk->z;
    double vx = pk->u, vy = pk->v, vz = pk->w;
    int xP = pk->xP, yP = pk->yP, zP = pk->zP;
    double absTau = 0., passProb, dSx = 0., dSy = 0., dSz = 0., dS, tauCell, dV, dx, dy, dz;
    int i, cell, bad = 0;

    const int safeLimit = in->plane ? ORC_SAFE_LIMIT_PLANE : ORC_SAFE_LIMIT;

    passProb = -log(1. - orng_u(g));

    for (i = 1; i <= safeLimit; i++) {
        cnt->c[ORC_SEGMENTS] += 1;
        /* 2 % nudge, signed midpoint in the ratio (:1882-1919) */
        if (xP < nx && vx > 0. && fabs(rx - (xa[xP - 1] + xa[xP]) / 2.) / ((xa[xP - 1] + xa[xP]) / 2.) < 2.e-2) xP = xP + 1;
        if (xP > 1 && vx < 0. && fabs(rx - (xa[xP - 1] + xa[xP - 2]) / 2.) / ((xa[xP - 1] + xa[xP - 2]) / 2.) < 2.e-2) xP = xP - 1;
        if (yP < ny && vy > 0. && fabs(ry - (ya[yP - 1] + ya[yP]) / 2.) / ((ya[yP - 1] + ya[yP]) / 2.) < 2.e-2) yP = yP + 1;
        if (yP > 1 && vy < 0. && fabs(ry - (ya[yP - 1] + ya[yP - 2]) / 2.) / ((ya[yP - 1] + ya[yP - 2]) / 2.) < 2.e-2) yP = yP - 1;
        if (zP < nz && vz > 0. && fabs(rz - (za[zP - 1] + za[zP]) / 2.) / ((za[zP - 1] + za[zP]) / 2.) < 2.e-2) zP = zP + 1;
        if (zP > 1 && vz < 0. && fabs(rz - (za[zP - 1] + za[zP - 2]) / 2.) / ((za[zP - 1] + za[zP - 2]) / 2.) < 2.e-2) zP = zP - 1;

        if (xP > nx || xP < 1 || yP > ny || yP < 1 || zP > nz || zP < 1) {
            cnt->c[ORC_ERRORS] += 1; /* "pathSegment: insanity" stop */
            return 0;
        }
        for (;;) {
        /* a mother cell that a sub-grid covers: into the sub-grid (:1931-1963) */
        if (gP == 0 && in->nSub > 0 && GACT(G, xP, yP, zP) < 0) {
            SETG(-GACT(G, xP, yP, zP) - 1);
            xP = nest_cell(xa, nx, rx, 1, vx);
            yP = nest_cell(ya, ny, ry, 1, vy);
            zP = nest_cell(za, nz, rz, 1, vz);
        }
        if (in->sym) { /* the mother's mirror planes (:1966-1979) */
            if (rx <= in->xAxis[0]) { if (vx < 0.) vx = -vx; rx = in->xAxis[0]; }
            if (ry <= in->yAxis[0]) { if (vy < 0.) vy = -vy; ry = in->yAxis[0]; }
            if (rz <= in->zAxis[0]) { if (vz < 0.) vz = -vz; rz = in->zAxis[0]; }
        }
        /* distances from the walls (:1990-2094); a sub-grid's faces drop no packet */
        if (vx > 0.) {
            if (xP < nx) {
                dSx = ((xa[xP] + xa[xP - 1]) / 2. - rx) / vx;
                if (fabs(dSx) < 1.e-5) { rx = (xa[xP] + xa[xP - 1]) / 2.; xP = xP + 1; }
            } else {
                dSx = (xa[nx - 1] - rx) / vx;
                if (fabs(dSx) < 1.e-5) {
                    rx = xa[nx - 1];
                    if (!in->plane && gP == 0) { cnt->c[ORC_LOST] += 1; return 0; }
                }
            }
        } else if (vx < 0.) {
            if (xP > 1) {
                dSx = ((xa[xP - 1] + xa[xP - 2]) / 2. - rx) / vx;
                if (fabs(dSx) < 1.e-5) { rx = (xa[xP - 1] + xa[xP - 2]) / 2.; xP = xP - 1; }
            } else {
                dSx = (xa[0] - rx) / vx;
                if (fabs(dSx) < 1.e-5) rx = xa[0];
            }
        } else {
            dSx = xa[nx - 1];
        }
        if (vy > 0.) {
            if (yP < ny) {
                dSy = ((ya[yP] + ya[yP - 1]) / 2. - ry) / vy;
                if (fabs(dSy) < 1.e-5) { ry = (ya[yP] + ya[yP - 1]) / 2.; yP = yP + 1; }
            } else {
                dSy = (ya[ny - 1] - ry) / vy;
                if (fabs(dSy) < 1.e-5) {
                    ry = ya[ny - 1];
                    if (gP == 0) { cnt->c[ORC_LOST] += 1; return 0; }
                }
            }
        } else if (vy < 0.) {
            if (yP > 1) {
                dSy = ((ya[yP - 1] + ya[yP - 2]) / 2. - ry) / vy;
                if (fabs(dSy) < 1.e-5) { ry = (ya[yP - 1] + ya[yP - 2]) / 2.; yP = yP - 1; }
            } else {
                dSy = (ya[0] - ry) / vy;
                if (fabs(dSy) < 1.e-5) ry = ya[0];
            }
        } else {
            dSy = ya[ny - 1];
        }
        if (vz > 0.) {
            if (zP < nz) {
                dSz = ((za[zP] + za[zP - 1]) / 2. - rz) / vz;
                if (fabs(dSz) < 1.e-5) { rz = (za[zP] + za[zP - 1]) / 2.; zP = zP + 1; }
            } else {
                dSz = (za[nz - 1] - rz) / vz;
                if (fabs(dSz) < 1.e-5) {
                    rz = za[nz - 1];
                    if (!in->plane && gP == 0) { cnt->c[ORC_LOST] += 1; return 0; }
                }
            }
        } else if (vz < 0.) {
            if (zP > 1) {
                dSz = ((za[zP - 1] + za[zP - 2]) / 2. - rz) / vz;
                if (fabs(dSz) < 1.e-5) { rz = (za[zP - 1] + za[zP - 2]) / 2.; zP = zP - 1; }
            } else {
                dSz = (za[0] - rz) / vz;
                if (fabs(dSz) < 1.e-5) rz = za[0];
            }
        } else {
            dSz = za[nz - 1];
        }
        if (xP > nx || xP < 1 || yP > ny || yP < 1 || zP > nz || zP < 1) {
            cnt->c[ORC_ERRORS] += 1;
            return 0;
        }
        /* a wall snap into a sub-grid's cells: again from its walls (:2083) */
        if (gP == 0 && in->nSub > 0 && GACT(G, xP, yP, zP) < 0) continue;
        break;
        }
        /* cells on a wall (:2096-2102), nearest wall (:2104-2137) */
        if (fabs(dSx) < 1.e-5) dSx = xa[nx - 1];
        if (fabs(dSy) < 1.e-5) dSy = ya[ny - 1];
        if (fabs(dSz) < 1.e-5) dSz = za[nz - 1];
        dSx = fabs(dSx);
        dSy = fabs(dSy);
        dSz = fabs(dSz);
        if (dSx <= 0.) dS = fmin(dSy, dSz);
        else if (dSy <= 0.) dS = fmin(dSx, dSz);
        else if (dSz <= 0.) dS = fmin(dSx, dSy);
        else { dS = fmin(dSx, dSy); dS = fmin(dS, dSz); }
        if (!(dS > 0.)) { cnt->c[ORC_ERRORS] += 1; return 0; }

        cell = GACT(G, xP, yP, zP);
        tauCell = dS * in->opacity[PLANE(in, pk->nuP, cell)];

        /* volume element in 1e45 cm^3 (:2166-2223); a plane-parallel grid
         * halves the first and the last cell of every axis */
        {
            const int h0 = in->sym || in->plane || gP > 0, h1 = in->plane || gP > 0;
            if (xP > 1 && xP < nx) dx = fabs(xa[xP] - xa[xP - 2]) / 2.;
            else if (xP == 1) dx = h0 ? fabs(xa[1] - xa[0]) / 2. : fabs(xa[1] - xa[0]);
            else dx = h1 ? fabs(xa[xP - 1] - xa[xP - 2]) / 2. : fabs(xa[xP - 1] - xa[xP - 2]);
            if (yP > 1 && yP < ny) dy = fabs(ya[yP] - ya[yP - 2]) / 2.;
            else if (yP == 1) dy = h0 ? fabs(ya[1] - ya[0]) / 2. : fabs(ya[1] - ya[0]);
            else dy = h1 ? fabs(ya[yP - 1] - ya[yP - 2]) / 2. : fabs(ya[yP - 1] - ya[yP - 2]);
            if (zP > 1 && zP < nz) dz = fabs(za[zP] - za[zP - 2]) / 2.;
            else if (zP == 1) dz = h0 ? fabs(za[1] - za[0]) / 2. : fabs(za[1] - za[0]);
            else dz = h1 ? fabs(za[zP - 1] - za[zP - 2]) / 2. : fabs(za[zP - 1] - za[zP - 2]);
        }
        dx = dx / 1.e15;
        dy = dy / 1.e15;
        dz = dz / 1.e15;
        dV = dx * dy * dz;

        if (absTau + tauCell > passProb && cell > 0) {
            /* the packet interacts in this cell (:2244-2700) */
            double kap = in->opacity[PLANE(in, pk->nuP, cell)];
            double dlLoc = (passProb - absTau) / kap, probSca, r;
            rx = rx + dlLoc * vx;
            ry = ry + dlLoc * vy;
            rz = rz + dlLoc * vz;
            if (in->sym && gP == 0) {
                if (rx <= xa[0]) { if (vx < 0.) vx = -vx; rx = xa[0]; }
                if (ry <= ya[0]) { if (vy < 0.) vy = -vy; ry = ya[0]; }
                if (rz <= za[0]) { if (vz < 0.) vz = -vz; rz = za[0]; }
            }
            add_to(&out->J[PLANE(in, pk->nuP, cell)], dlLoc * in->deltaE / dV);
            if (in->R_out > 0.) {
                double rr = sqrt((rx / 1.e10) * (rx / 1.e10) + (ry / 1.e10) * (ry / 1.e10) +
                                 (rz / 1.e10) * (rz / 1.e10)) * 1.e10;
                if (rr >= in->R_out) { escape_tally(in, out, pk, cnt); return 0; }
            }
            if (in->lgDust) {
                /* grains (:2385-2500): scattered isotropically with probability
                 * scaOpac/opacity, else absorbed; no Compton scattering */
                probSca = in->scaOpac[PLANE(in, pk->nuP, cell)] / kap;
                r = 1. - orng_u(g);
                if (r <= probSca) {
                    cnt->c[ORC_SCATTERED] += 1;
                    random_unit_vector(g, &vx, &vy, &vz);
                    pk->origin = cell;
                    absTau = 0.;
                    passProb = -log(1. - orng_u(g));
                    continue;
                }
                cnt->c[ORC_ABSORBED] += 1;
                pk->x = rx; pk->y = ry; pk->z = rz;
                pk->xP = xP; pk->yP = yP; pk->zP = zP; pk->gP = gP;
                return 1; /* "dustEmi" without gas, "diffuse" with it */
            }
            probSca = in->lgCompton ? in->KNsigmaT[pk->nuP - 1] * in->Ne[cell] / kap : 0.;
            r = 1. - orng_u(g);
            if (r <= probSca) {
                cnt->c[ORC_SCATTERED] += 1;
                pk->xP = xP; pk->yP = yP; pk->zP = zP;
                compton_scatter(in, pk, g, &bad);
                if (bad) cnt->c[ORC_ERRORS] += 1;
                pk->origin = cell;
                vx = pk->u; vy = pk->v; vz = pk->w;
                absTau = 0.;
                passProb = -log(1. - orng_u(g));
                continue;
            }
            cnt->c[ORC_ABSORBED] += 1;
            pk->x = rx; pk->y = ry; pk->z = rz;
            pk->xP = xP; pk->yP = yP; pk->zP = zP; pk->gP = gP;
            return 1; /* packetType = "diffuse", reRun = 1 (:4076-4082) */
        }

        /* the packet crosses the cell (:2706-2874) */
        add_to(&out->J[PLANE(in, pk->nuP, cell)], dS * in->deltaE / dV);
        absTau = absTau + tauCell;
        rx = rx + dS * vx;
        ry = ry + dS * vy;
        rz = rz + dS * vz;
        if (dS == dSx && vx > 0.) xP = xP + 1;
        else if (dS == dSx && vx < 0.) xP = xP - 1;
        else if (dS == dSy && vy > 0.) yP = yP + 1;
        else if (dS == dSy && vy < 0.) yP = yP - 1;
        else if (dS == dSz && vz > 0.) zP = zP + 1;
        else if (dS == dSz && vz < 0.) zP = zP - 1;

        if (gP > 0) { /* in a sub-grid: back to the mother when it leaves the box */
            const double mx = in->xAxis[0], my = in->yAxis[0], mz = in->zAxis[0];
            int ret = 0;
            /* out through a y face of a full model (:2884-3244) */
            if (!in->sym && (ry < ya[0] || yP < 1 || ry > ya[ny - 1] || yP > ny)) ret = 1;
            /* beyond an upper face, or at the mother's lower faces (:3612-3624) */
            if ((!in->sym && (rx <= mx || ry <= my || rz <= mz)) || rx >= xa[nx - 1] ||
                ry >= ya[ny - 1] || rz >= za[nz - 1] || xP > nx || yP > ny || zP > nz)
                ret = 1;
            if (!ret && in->sym) { /* :3866-3885 */
                if (rx <= mx) { if (vx < 0.) vx = -vx; xP = 1; rx = xa[0]; }
                if (ry <= my) { if (vy < 0.) vy = -vy; yP = 1; ry = ya[0]; }
                if (rz <= mz) { if (vz < 0.) vz = -vz; zP = 1; rz = mz; }
            }
            /* out through any face, moving outwards (:3894-3948) */
            if (((rx <= xa[0] || xP < 1) && vx <= 0.) || ((ry <= ya[0] || yP < 1) && vy <= 0.) ||
                ((rz <= za[0] || zP < 1) && vz <= 0.) || ((rx >= xa[nx - 1] || xP > nx) && vx >= 0.) ||
                ((ry >= ya[ny - 1] || yP > ny) && vy >= 0.) || ((rz >= za[nz - 1] || zP > nz) && vz >= 0.))
                ret = 1;
            if (ret) {
                /* the axis left by: the index out of the box, else the face
                 * it sits on moving outwards */
                int axis = 0, t, *idx, n, step;
                if (yP < 1 || yP > ny) axis = 1;
                else if (zP < 1 || zP > nz) axis = 2;
                else if (xP >= 1 && xP <= nx) {
                    if ((ry <= ya[0] && vy <= 0.) || (ry >= ya[ny - 1] && vy >= 0.)) axis = 1;
                    else if ((rz <= za[0] && vz <= 0.) || (rz >= za[nz - 1] && vz >= 0.)) axis = 2;
                }
                SETG(0);
                xP = nest_cell(xa, nx, rx, 0, vx);
                yP = nest_cell(ya, ny, ry, 0, vy);
                zP = nest_cell(za, nz, rz, 0, vz);
                /* a return into a covered cell steps on along that axis */
                idx = axis == 0 ? &xP : (axis == 1 ? &yP : &zP);
                n = axis == 0 ? nx : (axis == 1 ? ny : nz);
                step = (axis == 0 ? vx : (axis == 1 ? vy : vz)) < 0. ? -1 : 1;
                for (t = 0; t < n && GACT(G, xP, yP, zP) < 0; t++) {
                    if (*idx + step < 1 || *idx + step > n) break;
                    *idx += step;
                }
                if (in->sym) {
                    if (rx <= xa[0]) { if (vx < 0.) vx = -vx; xP = 1; rx = xa[0]; }
                    if (ry <= ya[0]) { if (vy < 0.) vy = -vy; yP = 1; ry = ya[0]; }
                    if (rz <= za[0]) { if (vz < 0.) vz = -vz; zP = 1; rz = za[0]; }
                }
            }
            continue;
        }
        if (in->plane) {
            /* plane-parallel slab (:3249-3606): the packet leaves through the
             * lower or upper y face and is mirrored at the x and z faces (the
             * lower z face puts it at yAxis(1), as the reference does) */
            int ret = 0;
            if (ry < ya[0] || yP < 1) { yP = 1; ret = 1; }
            if (ry > ya[ny - 1] || yP > ny) { yP = ny; ret = 1; }
            if (rx <= xa[0] || xP < 1) { xP = 1; rx = xa[0]; vx = -vx; }
            if (rx >= xa[nx - 1] || xP > nx) { xP = nx; rx = xa[nx - 1]; vx = -vx; }
            if (rz <= za[0] || zP < 1) { zP = 1; rz = ya[0]; vz = -vz; }
            if (rz >= za[nz - 1] || zP > nz) { zP = nz; rz = za[nz - 1]; vz = -vz; }
            if (ret) { escape_tally(in, out, pk, cnt); return 0; }
            continue;
        }
        if (!in->sym) { /* :2884-3244 */
            int ret = 0;
            if (ry < ya[0] || yP < 1) { yP = 1; ret = 1; }
            if (ry > ya[ny - 1] || yP > ny) { yP = ny; ret = 1; }
            if (rx <= xa[0] || xP < 1) { xP = 1; ret = 1; }
            if (rx >= xa[nx - 1] || xP > nx) { xP = nx; ret = 1; }
            if (rz <= za[0] || zP < 1) { zP = 1; ret = 1; }
            if (rz >= za[nz - 1] || zP > nz) { zP = nz; ret = 1; }
            if (ret) { escape_tally(in, out, pk, cnt); return 0; }
        }
        /* :3612-3860 */
        if ((!in->sym && (rx <= xa[0] || ry <= ya[0] || rz <= za[0])) || rx >= xa[nx - 1] ||
            ry >= ya[ny - 1] || rz >= za[nz - 1] || xP > nx || yP > ny || zP > nz) {
            escape_tally(in, out, pk, cnt);
            return 0;
        }
        if (in->sym) {
            if (rx <= xa[0] || xP < 1) { if (vx < 0.) vx = -vx; xP = 1; rx = xa[0]; }
            if (ry <= ya[0] || yP < 1) { if (vy < 0.) vy = -vy; yP = 1; ry = ya[0]; }
            if (rz <= za[0] || zP < 1) { if (vz < 0.) vz = -vz; zP = 1; rz = za[0]; }
        }
    }
    cnt->c[ORC_LOST] += 1; /* safe limit exceeded (:4040-4047) */
    return 0;
}

/* cell index of a birth coordinate on one axis (photon_mod.f90:814-819):
 * locate, then the next index from the midpoint on, at least 1 */
static int plane_index(const double* a, int n, double x) {
    int p = oracle_locate(a, n, x);
    if (p >= 1 && p < n && x >= (a[p - 1] + a[p]) / 2.) p = p + 1;
    return p < 1 ? 1 : p;
}

/* a stellar packet of a plane-parallel model (photon_mod.f90:796-859): born
 * on the y = 0 plane at a uniform (x, z) of the slab's face, x drawn first */
static void plane_birth(const oracle_transport_in* in, opacket_t* pk, orng_t* g) {
    const double *xa = in->xAxis, *za = in->zAxis;
    double r = 1. - orng_u(g);
    pk->x = xa[0] + r * (xa[in->nx - 1] - xa[0]);
    pk->xP = plane_index(xa, in->nx, pk->x);
    pk->y = 0.;
    pk->yP = 1;
    r = 1. - orng_u(g);
    pk->z = za[0] + r * (za[in->nz - 1] - za[0]);
    pk->zP = plane_index(za, in->nz, pk->z);
}

/* newPhotonPacket + initPhotonPacket + energyPacketRun; returns reRun */
static int energy_packet_run(const oracle_transport_in* in, oracle_transport_out* out,
                             opacket_t* pk, orng_t* g, ocount_t* cnt) {
    int bad = 0;
    cnt->c[ORC_EMISSIONS] += 1;
    pk->lgLine = 0;
    if (pk->lgStellar) {
        pk->nuP = getnu2_draw(in->sourceCDF, in->nbins, g, &bad);
        if (in->nuArray[pk->nuP - 1] >= 1.)
            cnt->qphot += in->deltaE / (2.1799153e-11 * in->nuArray[pk->nuP - 1]);
        if (in->plane) plane_birth(in, pk, g);
    } else {
        int cell = GACT(grid_of(in, pk->gP), pk->xP, pk->yP, pk->zP);
        double r;
        if (cell <= 0) { cnt->c[ORC_ERRORS] += 1; return 0; }
        if (!in->dustOnly) { /* dustEmi packets (:1673-1741) are never lines */
            r = 1. - orng_u(g);
            if (r <= in->totalLines[cell]) { /* line photon (:1605-1634) */
                cnt->c[ORC_LINES] += 1;
                return 0;
            }
        }
        pk->nuP = getnu2_draw(in->recPDF + (size_t)cell * in->nbins, in->nbins, g, &bad);
    }
    if (bad) cnt->c[ORC_ERRORS] += 1;
    if (in->plane && pk->lgStellar) {
        pk->u = 0.; /* along +y (:851-854) */
        pk->v = 1.;
        pk->w = 0.;
    } else {
        random_unit_vector(g, &pk->u, &pk->v, &pk->w);
    }
    if (in->sym && pk->lgStellar) {
        pk->u = fabs(pk->u);
        pk->v = fabs(pk->v);
        pk->w = fabs(pk->w);
    }
    pk->origin = GACT(grid_of(in, pk->gP), pk->xP, pk->yP, pk->zP);
    pk->nu = in->nuArray[pk->nuP - 1];
    if (!in->lgDust && pk->nu < in->ionEdge1) { /* :634-719 */
        escape_tally(in, out, pk, cnt);
        return 0;
    }
    return path_segment(in, out, pk, g, cnt);
}

int oracle_transport(const oracle_transport_in* in, oracle_transport_out* out) {
    int64_t ip;
    double tot[ORC_N];
    double q = 0.;
    int k;
    memset(tot, 0, sizeof(tot));
#ifdef _OPENMP
    int nt = in->nthreads > 1 ? in->nthreads : 1;
#pragma omp parallel num_threads(nt)
#endif
    {
        ocount_t cnt;
        memset(&cnt, 0, sizeof(cnt));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (ip = 0; ip < in->n; ip++) {
            opacket_t pk;
            orng_t g;
            int e, reRun = 1;
            orng_init(&g, in->seed, in->id_base + (uint64_t)ip, in->ctr_base);
            memset(&pk, 0, sizeof(pk));
            pk.x = in->star[0];
            pk.y = in->star[1];
            pk.z = in->star[2];
            pk.xP = in->starIdx[0];
            pk.yP = in->starIdx[1];
            pk.zP = in->starIdx[2];
            pk.gP = in->nSub > 0 ? in->startGrid : 0;
            pk.lgStellar = 1;
            for (e = 1; e <= ORC_RECURSION_LIMIT; e++) {
                reRun = energy_packet_run(in, out, &pk, &g, &cnt);
                if (!reRun) break;
                pk.lgStellar = 0; /* re-emitted as "diffuse" */
            }
            if (reRun) cnt.c[ORC_TRAPPED] += 1;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (k = 0; k < ORC_N; k++) tot[k] += cnt.c[k];
            q += cnt.qphot;
        }
    }
    for (k = 0; k < ORC_N; k++) out->counters[k] = tot[k];
    out->Qphot = q;
    return tot[ORC_ERRORS] > 0 ? -4 : 0;
}

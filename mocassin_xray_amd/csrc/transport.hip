// Energy-packet transport on gfx950: emission, Cartesian traversal with the
// Lucy (1999) path-length estimator of J_nu, Compton scattering, absorption
// and diffuse re-emission, escape tallies.
//
// Semantics follow MOCASSIN's energyPacketDriver (photon_mod.f90:28-414) for a
// single Cartesian grid without sub-grids, plane-parallel illumination or
// fluorescence, with grains (dust scattering and absorption, dust-emitted
// packets) when the grid has them:
//   energyPacketRun          photon_mod.f90:418-741
//   initPhotonPacket         photon_mod.f90:744-896 (randomUnitVector
//                            vector_mod.f90:303-314, octant folding :875-882)
//   getNu2                   photon_mod.f90:928-972 (bin-1 fold kept)
//   newPhotonPacket          photon_mod.f90:976-1756 (stellar, diffuse)
//   pathSegment              photon_mod.f90:1758-4085
//   comptonScatter           photon_mod.f90:6805-6844
//
// GPU structure: one packet per lane; lanes pull packet ids from a device work
// counter and each loop trip runs either an emission or one cell-crossing
// segment, so the wave stays converged on the segment body while packets of
// very different lifetimes flow through it.  Geometry lives in the L1/L2
// (axes and widths are a few KB); opacity and J are (bin, cell) planes with
// the cell index fastest, which keeps the successive cells of one packet
// (same bin) in one plane.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "moc_internal.h"
#include "philox.h"

namespace {

constexpr double kPiRef = 3.141592654;  // constants_mod.f90:45 / vector_mod.f90:310
constexpr double kRyd2ergQ = 2.1799153e-11;  // photon_mod.f90:1471

enum : int { ST_DONE = 0, ST_EMIT = 1, ST_FLY = 2 };

template <typename T>
struct Pk {
    T x, y, z;     // position [cm]
    T u, v, w;     // direction
    T nu;          // frequency [Ryd]
    T absTau, passProb;
    int ix, iy, iz;  // 0-based cell indices
    int nu0;         // 0-based frequency bin
    int origin;      // active id of the emitting cell
    int stellar;
    int state;
    int emits;       // emission events of this packet
    int segs;        // segments of the current pathSegment call
    int kcell;       // cell whose opacity in bin nu0 is held in kap (-1: none)
    T kap;
    int hb;          // hot-block slot of bin nu0 (-1: none)
    int gi;          // view of the grid the packet is in (nested sub-grids; 0 otherwise)
};

// The J tally of one step: into the workgroup's LDS partial of the hot block
// when both the bin and the cell have a slot there, else a global atomic.
template <typename T>
__device__ __forceinline__ void add_j(const MocTransportArgs<T>& a, float* hot, double* Jc, double dJ,
                                      int hb, int hc) {
    if (hot && hb >= 0 && hc >= 0) atomicAdd(&hot[hb * a.nHotC + hc], (float)dJ);
    else atomicAdd(Jc, dJ);
}

template <typename T>
__device__ __forceinline__ int hot_bin(const MocTransportArgs<T>& a, const float* hot, int nu0) {
    return hot ? a.hotBin[nu0] : -1;
}
// slot of cell (ix, iy, iz) in the hot box of cells around the star, or -1
template <typename T>
__device__ __forceinline__ int hot_cell(const MocTransportArgs<T>& a, const float* hot, int ix, int iy, int iz) {
    if (!hot) return -1;
    const int x = ix - a.hotLo[0], y = iy - a.hotLo[1], z = iz - a.hotLo[2];
    if ((unsigned)x >= (unsigned)a.hotN[0] || (unsigned)y >= (unsigned)a.hotN[1] ||
        (unsigned)z >= (unsigned)a.hotN[2])
        return -1;
    return (x * a.hotN[1] + y) * a.hotN[2] + z;
}

// The grid a packet moves in (MocGridView): for a single grid it is built
// once per lane from the launch's (uniform) arguments, with the axes and
// volume factors staged in LDS -- segment walks read them every step, and LDS
// reads do not queue behind the J atomics in the vector memory counter; in a
// batch of grids each packet loads its grid's view from global memory.
template <typename T>
using View = MocGridView<T>;

template <typename T>
__device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }

// getNu2 (photon_mod.f90:928-972) on a monotone CDF of n entries: the linear
// scan's "last entry <= r" is an upper-bound bisection; bin 0 folds into bin 1.
template <typename T>
__device__ __forceinline__ int get_nu2(const T* __restrict__ cdf, int n, MocRng& rng,
                                       int* bad) {
    T r = rng.uniform<T>();
    for (int k = 0; k < 16 && (r == T(0) || r == T(1)); ++k) r = rng.uniform<T>();
    if (r == T(0) || r == T(1)) *bad = 1;
    int lo = 0, hi = n;  // first index with cdf[i] > r
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > r) hi = mid; else lo = mid + 1;
    }
    int k = lo;
    if (k < 1) k = 1;
    if (k > n - 1) k = n - 1;
    return k;
}

// locate (interpolation_mod.f90:49-106) returning the 0-based floor bin,
// clamped to [0, n-1]; used for the Compton-shifted frequency.
template <typename T>
__device__ __forceinline__ int locate_floor(const T* __restrict__ xa, int n, T x) {
    if (x > xa[n - 1]) return n - 1;
    if (x < xa[0]) return 0;
    int lo = 0, hi = n;  // first index with xa[i] > x
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xa[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo - 1;
}

template <typename T>
__device__ __forceinline__ void unit_vector(MocRng& rng, T& u, T& v, T& w) {
    for (int k = 0; k < 64; ++k) {
        const T r1 = rng.uniform<T>();
        w = T(2) * r1 - T(1);
        const T t = sqrt(T(1) - w * w);
        const T r2 = rng.uniform<T>();
        const T ang = T(kPiRef) * (T(2) * r2 - T(1));
        u = t * cos(ang);
        v = t * sin(ang);
        if (u != T(0) && v != T(0) && w != T(0)) return;
    }
}

// distance to the next wall along one axis; walls sit at axis midpoints and
// the outer faces at the first/last axis point (photon_mod.f90:1990-2094).
// Returns false when the packet snaps onto the upper outer face and is
// dropped, which a plane-parallel slab does not do on its x and z axes
// (keep = true there: the face mirrors the packet after the step).
// One division per axis: the target wall is selected first (the midpoint sum
// is commutative, so both directions form it as ax[idx] + ax[j]).
template <typename T>
__device__ __forceinline__ bool wall_dist(T& pos, T dir, int& idx, const T* ax, int n, T& dS,
                                          bool keep = false) {
    const bool up = dir > T(0);
    if (!up && !(dir < T(0))) {
        dS = ax[n - 1];
        return true;
    }
    const bool inner = up ? idx < n - 1 : idx > 0;
    const int j = up ? idx + 1 : idx - 1;
    const T t = inner ? (ax[idx] + ax[j]) / T(2) : (up ? ax[n - 1] : ax[0]);
    dS = (t - pos) / dir;
    if (tabs(dS) < T(1.e-5)) {
        pos = t;
        if (inner) idx = j;
        else if (up && !keep) return false;
    }
    return true;
}

// the 2 % index nudge at the start of each segment (photon_mod.f90:1882-1919);
// the ratio uses the signed midpoint as the reference does.
// Only the wall ahead can move the index (an upward step cannot be undone by
// the downward test, which needs dir < 0), so one ratio per axis is formed.
template <typename T>
__device__ __forceinline__ void nudge(T pos, T dir, int& idx, const T* ax, int n) {
    const bool up = dir > T(0) && idx < n - 1;
    const bool dn = dir < T(0) && idx > 0;
    if (up || dn) {
        const int j = up ? idx + 1 : idx - 1;
        const T m = (ax[idx] + ax[j]) / T(2);
        if (tabs(pos - m) / m < T(2.e-2)) idx = j;
    }
}

// active cell id of (ix, iy, iz); a fully active grid numbered by active_map
// order needs no lookup, which takes one dependent memory access off a segment
template <typename T>
__device__ __forceinline__ int cell_of(const View<T>& v, int ix, int iy, int iz) {
    if (!v.active) return v.base + 1 + iz + v.nz * (iy + v.ny * ix);
    return v.active[(size_t)ix + (size_t)v.nx * ((size_t)iy + (size_t)v.ny * (size_t)iz)];
}

// per-lane event counts; a lane handles at most a few million events per
// launch, and the wave sums are taken in 64 bits
struct Cnt {
    unsigned seg = 0, emit = 0, absb = 0, sca = 0, esc = 0, line = 0, lost = 0, trap = 0, err = 0;
    double qphot = 0.0;
};

template <typename T>
__device__ __forceinline__ void tally_escape(const MocTransportArgs<T>& a, const View<T>& v,
                                             const Pk<T>& p, Cnt& c) {
    atomicAdd(&a.escaped[(size_t)p.nu0 * (size_t)(a.nCells + 1) + (size_t)p.origin],
              (double)v.deltaE);
    ++c.esc;
}

// cell index of a birth coordinate on one axis (photon_mod.f90:814-819):
// locate, then the next index from the midpoint on
template <typename T>
__device__ __forceinline__ int plane_index(const T* ax, int n, T x) {
    int k = locate_floor(ax, n, x);
    if (k < n - 1 && x >= (ax[k] + ax[k + 1]) / T(2)) ++k;
    return k;
}

// a position's cell on one axis of a sub-grid it enters (photon_mod.f90:
// 1937-1960): locate, at least the first cell, the next one above the midpoint
template <typename T>
__device__ __forceinline__ int nest_index(const T* ax, int n, T x) {
    int k = locate_floor(ax, n, x);
    if (k < n - 1 && x > (ax[k] + ax[k + 1]) / T(2)) ++k;
    return k;
}

// the packet enters the sub-grid of view vi at its position (a mother cell
// that the sub-grid covers, photon_mod.f90:1931-1963)
template <typename T>
__device__ __forceinline__ void nest_enter(const MocTransportArgs<T>& a, View<T>& g, Pk<T>& p, int vi) {
    g = a.grids[vi];
    p.gi = vi;
    p.ix = nest_index(g.ax, g.nx, p.x);
    p.iy = nest_index(g.ay, g.ny, p.y);
    p.iz = nest_index(g.az, g.nz, p.z);
}

// back to the mother grid, in the mother cell that holds the position
// (photon_mod.f90:3894-3948: locate, at least 1, the next cell from the
// midpoint on).  Departures from the reference: it writes that midpoint as
// xAxis(i) + xAxis(i+1)/2, the parenthesis misplaced; and a packet leaving
// through a sub-grid face, which lies on a mother wall, sits on that wall up
// to rounding, so the position alone can put it back into a covered cell,
// where it re-enters the sub-grid and slides along the face.  Here the
// midpoint itself, on it the cell ahead of the packet, and a return that
// lands in a covered cell steps on along the axis the packet left by.
template <typename T>
__device__ __forceinline__ int leave_index(const T* ax, int n, T x, T dir) {
    int k = locate_floor(ax, n, x);
    if (k < n - 1) {
        const T m = (ax[k] + ax[k + 1]) / T(2);
        if (x > m || (x == m && dir > T(0))) ++k;
    }
    return k;
}
template <typename T>
__device__ __forceinline__ void nest_leave(const MocTransportArgs<T>& a, View<T>& g, Pk<T>& p, int axis) {
    g = a.grids[0];
    p.gi = 0;
    p.ix = leave_index(g.ax, g.nx, p.x, p.u);
    p.iy = leave_index(g.ay, g.ny, p.y, p.v);
    p.iz = leave_index(g.az, g.nz, p.z, p.w);
    int* idx = axis == 0 ? &p.ix : (axis == 1 ? &p.iy : &p.iz);
    const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    const int step = (axis == 0 ? p.u : (axis == 1 ? p.v : p.w)) < T(0) ? -1 : 1;
    for (int t = 0; t < n && cell_of(g, p.ix, p.iy, p.iz) < 0; ++t) {
        const int k = *idx + step;
        if (k < 0 || k > n - 1) break;
        *idx = k;
    }
}

// newPhotonPacket + initPhotonPacket + the escape/line branches of
// energyPacketRun, then the start of pathSegment (tau = -ln(1-R)).
template <typename T>
__device__ __forceinline__ void emit(const MocTransportArgs<T>& a, const View<T>& v, Pk<T>& p,
                                     MocRng& rng, Cnt& c, const float* hot) {
    ++c.emit;
    ++p.emits;
    int bad = 0;
    if (p.stellar) {
        p.nu0 = get_nu2(v.sourceCDF, a.nbins, rng, &bad);
        const T nu = a.nuArray[p.nu0];
        if (nu >= T(1)) c.qphot += (double)v.deltaE / (kRyd2ergQ * (double)nu);
        if (v.plane) {
            // plane-parallel illumination (photon_mod.f90:796-859): born on
            // the y = 0 plane at a uniform (x, z) of the slab's face
            const int nx = v.nx, nz = v.nz;
            T r = T(1) - rng.uniform<T>();
            p.x = v.ax[0] + r * (v.ax[nx - 1] - v.ax[0]);
            p.ix = plane_index(v.ax, nx, p.x);
            p.y = T(0);
            p.iy = 0;
            r = T(1) - rng.uniform<T>();
            p.z = v.az[0] + r * (v.az[nz - 1] - v.az[0]);
            p.iz = plane_index(v.az, nz, p.z);
        }
    } else {
        const int cell = cell_of(v, p.ix, p.iy, p.iz);
        if (cell <= 0) { ++c.err; p.state = ST_DONE; return; }
        if (!a.dustOnly) {  // a dustEmi packet (:1673-1741) is never a line
            const T r = T(1) - rng.uniform<T>();
            if (r <= a.totalLines[cell]) {  // line photon: leaves the continuum transfer
                ++c.line;
                p.state = ST_DONE;
                return;
            }
        }
        p.nu0 = get_nu2(a.recPDF + (size_t)cell * (size_t)a.nbins, a.nbins, rng, &bad);
    }
    if (bad) ++c.err;
    if (v.plane && p.stellar) {  // along +y (:851-854)
        p.u = T(0);
        p.v = T(1);
        p.w = T(0);
    } else {
        unit_vector(rng, p.u, p.v, p.w);
    }
    if (v.sym && p.stellar) {
        p.u = tabs(p.u);
        p.v = tabs(p.v);
        p.w = tabs(p.w);
    }
    p.origin = cell_of(v, p.ix, p.iy, p.iz);
    p.nu = a.nuArray[p.nu0];
    if (!a.lgDust && p.nu < a.ionEdge1) {  // below every edge and no dust: escapes at once
        tally_escape(a, v, p, c);
        p.state = ST_DONE;
        return;
    }
    p.absTau = T(0);
    p.passProb = -log(T(1) - rng.uniform<T>());
    p.segs = 0;
    p.state = ST_FLY;
    p.kcell = p.origin;
    p.kap = a.opacity[(size_t)p.nu0 * (size_t)(a.nCells + 1) + (size_t)p.origin];
    p.hb = hot_bin(a, hot, p.nu0);
}

// Compton scattering (photon_mod.f90:6805-6844): the KN angle is used as the
// lab-frame azimuth with w isotropic, as in the reference.
template <typename T>
__device__ __forceinline__ void compton(const MocTransportArgs<T>& a, Pk<T>& p, MocRng& rng) {
    int bad = 0;
    const int th = get_nu2(a.KNcdf + (size_t)p.nu0 * (size_t)a.nAngles, a.nAngles, rng, &bad);
    const T theta = T(th + 1) * T(kPiRef) / T(180);
    const T w = T(2) * rng.uniform<T>() - T(1);
    const T t = sqrt(T(1) - w * w);
    p.u = t * cos(theta);
    p.v = t * sin(theta);
    p.w = w;
    const T nn = p.nu * a.Pcomp[(size_t)p.nu0 * (size_t)a.nAngles + th];
    p.nu0 = locate_floor(a.nuArray, a.nbins, nn);
    p.nu = nn;
}

// the packet interacts inside the cell (photon_mod.f90:2244-2700): it moves
// to the interaction point (mirrored at the octant's walls in the mother
// grid), leaves beyond R_out, scatters off grains or free electrons, or is
// absorbed and re-emitted; the step's J is added in every case
template <typename T>
__device__ __forceinline__ void interact(const MocTransportArgs<T>& a, const View<T>& g, Pk<T>& p,
                                         MocRng& rng, Cnt& c, float* hot, int cell, T kap, size_t row,
                                         double* Jc, T dV, int hc, bool mirror) {
    // the packet interacts inside this cell
    const T dl = (p.passProb - p.absTau) / kap;
    const double dJ = (double)(dl * g.deltaE / dV);
    p.x = p.x + dl * p.u;
    p.y = p.y + dl * p.v;
    p.z = p.z + dl * p.w;
    if (mirror) {
        if (p.x <= g.ax[0]) { if (p.u < T(0)) p.u = -p.u; p.x = g.ax[0]; }
        if (p.y <= g.ay[0]) { if (p.v < T(0)) p.v = -p.v; p.y = g.ay[0]; }
        if (p.z <= g.az[0]) { if (p.w < T(0)) p.w = -p.w; p.z = g.az[0]; }
    }
    if (g.R_out > T(0)) {
        // radius with the 1e10 scaling of photon_mod.f90:2294 (no FP32 overflow)
        const T sx = p.x / T(1.e10), sy = p.y / T(1.e10), sz = p.z / T(1.e10);
        const T rr = sqrt(sx * sx + sy * sy + sz * sz) * T(1.e10);
        if (rr >= g.R_out) {
            add_j(a, hot, Jc, dJ, p.hb, hc);
            tally_escape(a, g, p, c);
            p.state = ST_DONE;
            return;
        }
    }
    if (a.lgDust) {
        // grains (:2385-2500): scattered with probability scaOpac/kappa,
        // isotropically; otherwise absorbed and re-emitted (dustEmi
        // without gas, diffuse with it); no Compton scattering
        const T pSca = a.scaOpac[row + (size_t)cell] / kap;
        const T rd = T(1) - rng.uniform<T>();
        if (rd <= pSca) {
            ++c.sca;
            unit_vector(rng, p.u, p.v, p.w);
            p.origin = cell;
            p.absTau = T(0);
            p.passProb = -log(T(1) - rng.uniform<T>());
            p.kcell = cell;
            p.kap = kap;
            add_j(a, hot, Jc, dJ, p.hb, hc);
            return;
        }
        add_j(a, hot, Jc, dJ, p.hb, hc);
        ++c.absb;
        if (p.emits >= MOC_RECURSION_LIMIT) {
            ++c.trap;
            p.state = ST_DONE;
            return;
        }
        p.stellar = 0;
        p.state = ST_EMIT;
        return;
    }
    T probSca = T(0);
    if (a.lgCompton) probSca = a.KNsigmaT[p.nu0] * a.Ne[cell] / kap;
    const T r = T(1) - rng.uniform<T>();
    if (r <= probSca) {
        ++c.sca;
        compton(a, p, rng);
        p.origin = cell;
        p.absTau = T(0);
        p.passProb = -log(T(1) - rng.uniform<T>());
        p.kcell = cell;
        p.kap = a.opacity[(size_t)p.nu0 * (size_t)(a.nCells + 1) + (size_t)cell];
        add_j(a, hot, Jc, dJ, p.hb, hc);  // the step was taken in the old bin
        p.hb = hot_bin(a, hot, p.nu0);
        return;
    }
    add_j(a, hot, Jc, dJ, p.hb, hc);
    // absorbed: re-emitted as a diffuse packet from this position
    ++c.absb;
    if (p.emits >= MOC_RECURSION_LIMIT) {
        ++c.trap;
        p.state = ST_DONE;
        return;
    }
    p.stellar = 0;
    p.state = ST_EMIT;
    return;
}

// one trip of pathSegment inside a nested sub-grid (gP > 1 in photon_mod.f90:
// 1758-4038): the sub-grid's walls, none of whose faces drops a packet; the
// mother's mirror planes in an octant model (the interaction point is not
// mirrored in a sub-grid, :2254); cell volumes with half-width end cells on
// every face (:2169-2223); the packet goes back to the mother grid when it
// leaves the sub-grid's box (:2884-3244, :3612-3948).  No hot J block.
template <typename T>
__device__ void segment_sub(const MocTransportArgs<T>& a, View<T>& g, Pk<T>& p, MocRng& rng, Cnt& c) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (a.sym) {  // the mother's mirror planes (:1966-1979)
        if (p.x <= a.mir[0]) { if (p.u < T(0)) p.u = -p.u; p.x = a.mir[0]; }
        if (p.y <= a.mir[1]) { if (p.v < T(0)) p.v = -p.v; p.y = a.mir[1]; }
        if (p.z <= a.mir[2]) { if (p.w < T(0)) p.w = -p.w; p.z = a.mir[2]; }
    }
    T dSx = T(0), dSy = T(0), dSz = T(0);
    wall_dist(p.x, p.u, p.ix, g.ax, nx, dSx, true);
    wall_dist(p.y, p.v, p.iy, g.ay, ny, dSy, true);
    wall_dist(p.z, p.w, p.iz, g.az, nz, dSz, true);
    if (tabs(dSx) < T(1.e-5)) dSx = g.ax[nx - 1];
    if (tabs(dSy) < T(1.e-5)) dSy = g.ay[ny - 1];
    if (tabs(dSz) < T(1.e-5)) dSz = g.az[nz - 1];
    dSx = tabs(dSx);
    dSy = tabs(dSy);
    dSz = tabs(dSz);
    T dS;
    if (dSx <= T(0)) dS = fmin(dSy, dSz);
    else if (dSy <= T(0)) dS = fmin(dSx, dSz);
    else if (dSz <= T(0)) dS = fmin(dSx, dSy);
    else dS = fmin(fmin(dSx, dSy), dSz);
    if (!(dS > T(0))) {
        ++c.err;
        p.state = ST_DONE;
        return;
    }
    const size_t row = (size_t)p.nu0 * (size_t)(a.nCells + 1);
    const int cell = cell_of(g, p.ix, p.iy, p.iz);  // >= 0: a sub-grid holds no sub-grids
    const T kap = cell == p.kcell ? p.kap : a.opacity[row + (size_t)cell];
    const T tauCell = dS * kap;
    const T dV = g.vx[p.ix] * g.vy[p.iy] * g.vz[p.iz];
    double* const Jc = a.J + row + (size_t)cell;
    if (p.absTau + tauCell > p.passProb && cell > 0) {
        interact(a, g, p, rng, c, (float*)nullptr, cell, kap, row, Jc, dV, -1, false);
        return;
    }
    // the packet crosses the cell
    const double dJ = (double)(dS * g.deltaE / dV);
    p.absTau = p.absTau + tauCell;
    p.x = p.x + dS * p.u;
    p.y = p.y + dS * p.v;
    p.z = p.z + dS * p.w;
    if (dS == dSx && p.u > T(0)) ++p.ix;
    else if (dS == dSx && p.u < T(0)) --p.ix;
    else if (dS == dSy && p.v > T(0)) ++p.iy;
    else if (dS == dSy && p.v < T(0)) --p.iy;
    else if (dS == dSz && p.w > T(0)) ++p.iz;
    else if (dS == dSz && p.w < T(0)) --p.iz;
    const T* ax = g.ax;
    const T* ay = g.ay;
    const T* az = g.az;
    // out through a y face of a full (non-octant) model (:2884-3244)
    bool out = !a.sym && (p.y < ay[0] || p.iy < 0 || p.y > ay[ny - 1] || p.iy > ny - 1);
    // beyond an upper face, or at the mother's lower faces (:3612-3624)
    out = out || (!a.sym && (p.x <= a.mir[0] || p.y <= a.mir[1] || p.z <= a.mir[2])) ||
          p.x >= ax[nx - 1] || p.y >= ay[ny - 1] || p.z >= az[nz - 1] || p.ix > nx - 1 ||
          p.iy > ny - 1 || p.iz > nz - 1;
    if (!out && a.sym) {  // the mother's mirror planes (:3866-3885)
        if (p.x <= a.mir[0]) { if (p.u < T(0)) p.u = -p.u; p.ix = 0; p.x = ax[0]; }
        if (p.y <= a.mir[1]) { if (p.v < T(0)) p.v = -p.v; p.iy = 0; p.y = ay[0]; }
        if (p.z <= a.mir[2]) { if (p.w < T(0)) p.w = -p.w; p.iz = 0; p.z = a.mir[2]; }
    }
    // out through any face, moving outwards (:3894-3948)
    out = out || ((p.x <= ax[0] || p.ix < 0) && p.u <= T(0)) ||
          ((p.y <= ay[0] || p.iy < 0) && p.v <= T(0)) || ((p.z <= az[0] || p.iz < 0) && p.w <= T(0)) ||
          ((p.x >= ax[nx - 1] || p.ix > nx - 1) && p.u >= T(0)) ||
          ((p.y >= ay[ny - 1] || p.iy > ny - 1) && p.v >= T(0)) ||
          ((p.z >= az[nz - 1] || p.iz > nz - 1) && p.w >= T(0));
    if (out) {
        // the axis the packet left by: the index that stepped out of the box,
        // else the face it sits on moving outwards
        int axis = 0;
        if (p.iy < 0 || p.iy > ny - 1) axis = 1;
        else if (p.iz < 0 || p.iz > nz - 1) axis = 2;
        else if (p.ix >= 0 && p.ix <= nx - 1) {
            if (((p.y <= ay[0] && p.v <= T(0)) || (p.y >= ay[ny - 1] && p.v >= T(0)))) axis = 1;
            else if (((p.z <= az[0] && p.w <= T(0)) || (p.z >= az[nz - 1] && p.w >= T(0)))) axis = 2;
        }
        nest_leave(a, g, p, axis);
        if (g.sym) {  // the mother's own walls after the return (:3866-3885)
            if (p.x <= g.ax[0]) { if (p.u < T(0)) p.u = -p.u; p.ix = 0; p.x = g.ax[0]; }
            if (p.y <= g.ay[0]) { if (p.v < T(0)) p.v = -p.v; p.iy = 0; p.y = g.ay[0]; }
            if (p.z <= g.az[0]) { if (p.w < T(0)) p.w = -p.w; p.iz = 0; p.z = g.az[0]; }
        }
    }
    // the next cell's opacity (none for a mother cell that a sub-grid covers)
    p.kcell = cell_of(g, p.ix, p.iy, p.iz);
    if (p.kcell >= 0) p.kap = a.opacity[row + (size_t)p.kcell];
    if (cell > 0) add_j(a, (float*)nullptr, Jc, dJ, p.hb, -1);
}

// one trip of the pathSegment loop (photon_mod.f90:1880-4038).  The J
// contribution of the step is added last, after the opacity of the next cell
// has been requested: the vector memory counter retires in order, so the next
// step's wait for that opacity does not include the atomic's round trip.
template <typename T, bool NEST>
__device__ __forceinline__ void segment(const MocTransportArgs<T>& a, View<T>& g, Pk<T>& p,
                                        MocRng& rng, Cnt& c, float* hot) {
    if (p.segs >= (g.plane ? MOC_SAFE_LIMIT_PLANE : MOC_SAFE_LIMIT)) {
        ++c.lost;
        p.state = ST_DONE;
        return;
    }
    ++p.segs;
    ++c.seg;
    if constexpr (NEST) {
        if (p.gi > 0) {
            const int nx = g.nx, ny = g.ny, nz = g.nz;
            nudge(p.x, p.u, p.ix, g.ax, nx);
            nudge(p.y, p.v, p.iy, g.ay, ny);
            nudge(p.z, p.w, p.iz, g.az, nz);
            if (p.ix < 0 || p.ix >= nx || p.iy < 0 || p.iy >= ny || p.iz < 0 || p.iz >= nz) {
                ++c.err;
                p.state = ST_DONE;
                return;
            }
            segment_sub(a, g, p, rng, c);
            return;
        }
    }
    const int nx = g.nx, ny = g.ny, nz = g.nz;

    nudge(p.x, p.u, p.ix, g.ax, nx);
    nudge(p.y, p.v, p.iy, g.ay, ny);
    nudge(p.z, p.w, p.iz, g.az, nz);
    if (p.ix < 0 || p.ix >= nx || p.iy < 0 || p.iy >= ny || p.iz < 0 || p.iz >= nz) {
        ++c.err;
        p.state = ST_DONE;
        return;
    }
    if constexpr (NEST) {
        // a mother cell that a sub-grid covers: the step is taken in the sub-grid
        const int cm = cell_of(g, p.ix, p.iy, p.iz);
        if (cm < 0) {
            nest_enter(a, g, p, -cm - 1);
            segment_sub(a, g, p, rng, c);
            return;
        }
    }
    if (g.sym) {  // mirror walls of the octant model
        if (p.x <= g.ax[0]) { if (p.u < T(0)) p.u = -p.u; p.x = g.ax[0]; }
        if (p.y <= g.ay[0]) { if (p.v < T(0)) p.v = -p.v; p.y = g.ay[0]; }
        if (p.z <= g.az[0]) { if (p.w < T(0)) p.w = -p.w; p.z = g.az[0]; }
    }
    // opacity of the current cell: the prefetched value unless a nudge moved
    // the packet; a wall snap below moves it to a neighbour (rarely) and the
    // value is fetched again for that cell
    const size_t row = (size_t)p.nu0 * (size_t)(a.nCells + 1);
    const int cell0 = cell_of(g, p.ix, p.iy, p.iz);
    T kap0 = p.kap;
    if (cell0 != p.kcell) kap0 = a.opacity[row + (size_t)cell0];
    T dSx = T(0), dSy = T(0), dSz = T(0);
    const bool keep = g.plane != 0;
    if (!wall_dist(p.x, p.u, p.ix, g.ax, nx, dSx, keep) || !wall_dist(p.y, p.v, p.iy, g.ay, ny, dSy) ||
        !wall_dist(p.z, p.w, p.iz, g.az, nz, dSz, keep)) {
        ++c.lost;
        p.state = ST_DONE;
        return;
    }
    if (tabs(dSx) < T(1.e-5)) dSx = g.ax[nx - 1];
    if (tabs(dSy) < T(1.e-5)) dSy = g.ay[ny - 1];
    if (tabs(dSz) < T(1.e-5)) dSz = g.az[nz - 1];
    dSx = tabs(dSx);
    dSy = tabs(dSy);
    dSz = tabs(dSz);
    T dS;
    if (dSx <= T(0)) dS = fmin(dSy, dSz);
    else if (dSy <= T(0)) dS = fmin(dSx, dSz);
    else if (dSz <= T(0)) dS = fmin(dSx, dSy);
    else dS = fmin(fmin(dSx, dSy), dSz);
    if (!(dS > T(0))) {
        ++c.err;
        p.state = ST_DONE;
        return;
    }

    const int cell = cell_of(g, p.ix, p.iy, p.iz);
    if constexpr (NEST) {
        if (cell < 0) {  // a wall snap moved the packet into a sub-grid's cells (:2083)
            nest_enter(a, g, p, -cell - 1);
            segment_sub(a, g, p, rng, c);
            return;
        }
    }
    T kap = kap0;
    if (cell != cell0) kap = a.opacity[row + (size_t)cell];
    const int hc = hot_cell(a, hot, p.ix, p.iy, p.iz);  // the hot slot of this cell
    const T tauCell = dS * kap;
    const T dV = g.vx[p.ix] * g.vy[p.iy] * g.vz[p.iz];
    double* const Jc = a.J + row + (size_t)cell;

    if (p.absTau + tauCell > p.passProb && cell > 0) {
        interact(a, g, p, rng, c, hot, cell, kap, row, Jc, dV, hc, g.sym != 0);
        return;
    }

    // the packet crosses the cell
    const double dJ = (double)(dS * g.deltaE / dV);
    p.absTau = p.absTau + tauCell;
    p.x = p.x + dS * p.u;
    p.y = p.y + dS * p.v;
    p.z = p.z + dS * p.w;
    if (dS == dSx && p.u > T(0)) ++p.ix;
    else if (dS == dSx && p.u < T(0)) --p.ix;
    else if (dS == dSy && p.v > T(0)) ++p.iy;
    else if (dS == dSy && p.v < T(0)) --p.iy;
    else if (dS == dSz && p.w > T(0)) ++p.iz;
    else if (dS == dSz && p.w < T(0)) --p.iz;

    bool esc = false;
    if (g.plane) {
        // plane-parallel slab (:3249-3606): out through the y faces, mirrored
        // at the x and z faces (the lower z face puts the packet at yAxis(1),
        // as the reference does)
        if (p.y < g.ay[0] || p.iy < 0) { p.iy = 0; esc = true; }
        if (p.y > g.ay[ny - 1] || p.iy > ny - 1) { p.iy = ny - 1; esc = true; }
        if (p.x <= g.ax[0] || p.ix < 0) { p.ix = 0; p.x = g.ax[0]; p.u = -p.u; }
        if (p.x >= g.ax[nx - 1] || p.ix > nx - 1) { p.ix = nx - 1; p.x = g.ax[nx - 1]; p.u = -p.u; }
        if (p.z <= g.az[0] || p.iz < 0) { p.iz = 0; p.z = g.ay[0]; p.w = -p.w; }
        if (p.z >= g.az[nz - 1] || p.iz > nz - 1) { p.iz = nz - 1; p.z = g.az[nz - 1]; p.w = -p.w; }
    } else if (!g.sym) {
        if (p.y < g.ay[0] || p.iy < 0) { p.iy = 0; esc = true; }
        if (p.y > g.ay[ny - 1] || p.iy > ny - 1) { p.iy = ny - 1; esc = true; }
        if (p.x <= g.ax[0] || p.ix < 0) { p.ix = 0; esc = true; }
        if (p.x >= g.ax[nx - 1] || p.ix > nx - 1) { p.ix = nx - 1; esc = true; }
        if (p.z <= g.az[0] || p.iz < 0) { p.iz = 0; esc = true; }
        if (p.z >= g.az[nz - 1] || p.iz > nz - 1) { p.iz = nz - 1; esc = true; }
    }
    if (!esc && !g.plane &&
        ((!g.sym && (p.x <= g.ax[0] || p.y <= g.ay[0] || p.z <= g.az[0])) ||
         p.x >= g.ax[nx - 1] || p.y >= g.ay[ny - 1] || p.z >= g.az[nz - 1] ||
         p.ix > nx - 1 || p.iy > ny - 1 || p.iz > nz - 1))
        esc = true;
    if (!esc) {
        if (g.sym) {
            if (p.x <= g.ax[0] || p.ix < 0) { if (p.u < T(0)) p.u = -p.u; p.ix = 0; p.x = g.ax[0]; }
            if (p.y <= g.ay[0] || p.iy < 0) { if (p.v < T(0)) p.v = -p.v; p.iy = 0; p.y = g.ay[0]; }
            if (p.z <= g.az[0] || p.iz < 0) { if (p.w < T(0)) p.w = -p.w; p.iz = 0; p.z = g.az[0]; }
        }
        // every index is inside the grid here: request the next cell's opacity
        // (none for a cell that a sub-grid covers)
        p.kcell = cell_of(g, p.ix, p.iy, p.iz);
        if (!NEST || p.kcell >= 0) p.kap = a.opacity[row + (size_t)p.kcell];
    }
    if (cell > 0) add_j(a, hot, Jc, dJ, p.hb, hc);
    if (esc) {
        tally_escape(a, g, p, c);
        p.state = ST_DONE;
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Emissions wait for company: an emission walks a CDF row by bisection, a
// chain of dependent loads, so it runs when at least a.emitBatch lanes of the
// wave need one (or no lane is flying); until then those lanes sit out while
// the others keep stepping their segments, a.segsPerTrip per trip through the
// loop.  Packet ids are taken a wave at a time from the work counter.  Both
// are tuning parameters (moc_set_tuning; the package's autotuner times them
// on the model): on the 32^3 nebula (1e7 packets, 1000 bins) batch
// 4/8/16/24/32/40/48 gave 45/34/21/17/16/17/19 ms, and two segments per trip
// 16 -> 15.5 ms.

extern __shared__ __align__(16) unsigned char transport_lds[];

// BATCH: the packets of a batch of grids (a.nGrids > 0), each in its own grid.
// NEST (with BATCH): one source in a mother grid with nested sub-grids
// (a.nested), the packets moving between the grids' views.
// HOT: workgroups of 1024 lanes (one per CU) sum the tallies of the hot block
// of J keys (a.nHotC x a.nHotB) as FP32 partials in LDS, after the staged
// geometry, and write them to their row of a.hotPart at the end.
template <typename T, bool BATCH, bool HOT, bool NEST>
__global__ void __launch_bounds__(HOT ? 1024 : 256) k_transport(MocTransportArgs<T> a) {
    View<T> g;
    float* hot = nullptr;
    if constexpr (!BATCH) {
        // stage the axes and volume factors: [ax ay az vx vy vz]
        T* const sg = reinterpret_cast<T*>(transport_lds);
        const int nx = a.nx, ny = a.ny, nz = a.nz, nt = nx + ny + nz;
        for (int i = threadIdx.x; i < nt; i += blockDim.x) {
            const bool bx = i < nx, by = !bx && i < nx + ny;
            const int k = bx ? i : (by ? i - nx : i - nx - ny);
            sg[i] = bx ? a.ax[k] : (by ? a.ay[k] : a.az[k]);
            sg[nt + i] = bx ? a.vx[k] : (by ? a.vy[k] : a.vz[k]);
        }
        if constexpr (HOT) {
            hot = reinterpret_cast<float*>(sg + 2 * nt);
            for (int i = threadIdx.x; i < a.nHotB * a.nHotC; i += blockDim.x) hot[i] = 0.f;
        }
        __syncthreads();
        g.ax = sg; g.ay = sg + nx; g.az = sg + nx + ny;
        g.vx = sg + nt; g.vy = sg + nt + nx; g.vz = sg + nt + nx + ny;
        g.active = a.dense ? nullptr : a.active;
        g.sourceCDF = a.sourceCDF;
        g.nx = nx; g.ny = ny; g.nz = nz; g.base = 0; g.sym = a.sym; g.plane = a.plane;
        g.R_out = a.R_out;
        g.deltaE = a.deltaE;
    }
    Cnt c;
    Pk<T> p;
    p.state = ST_DONE;
    p.kcell = -1;
    p.hb = -1;
    p.gi = 0;
    MocRng rng;
    bool more = true;
    const int lane = threadIdx.x & 63;
    const uint32_t ctrBase = a.ctrDev ? *a.ctrDev + a.ctr_base : a.ctr_base;
    const int emitBatch = a.emitBatch, segsPerTrip = a.segsPerTrip;
    const unsigned long long below = (1ull << lane) - 1ull;
    while (true) {
        const unsigned long long want = __ballot(p.state == ST_DONE && more);
        if (want) {
            const int leader = __ffsll((long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&a.counters[MOC_CNT_NEXT], (unsigned long long)__popcll(want));
            base = __shfl(base, leader, 64);
            if (p.state == ST_DONE && more) {
                const unsigned long long id = base + (unsigned long long)__popcll(want & below);
                if ((int64_t)id < a.n) {
                    if constexpr (NEST) {
                        // the source's grid: the mother or the sub-grid around it
                        rng.init(a.seed, id, ctrBase);
                        g = a.grids[a.nestStart];
                        p.gi = a.nestStart;
                        p.x = g.star[0];
                        p.y = g.star[1];
                        p.z = g.star[2];
                        p.ix = g.starIdx[0];
                        p.iy = g.starIdx[1];
                        p.iz = g.starIdx[2];
                    } else if constexpr (BATCH) {
                        // the grid of this packet: last k with gridFirst[k] <= id
                        int lo = 0, hi = a.nGrids;
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (a.gridFirst[mid] <= (int64_t)id) lo = mid; else hi = mid;
                        }
                        const unsigned long long local = id - (unsigned long long)a.gridFirst[lo];
                        rng.init(a.seed, ((unsigned long long)lo << 32) + local, ctrBase);
                        g = a.grids[lo];
                        p.x = g.star[0];
                        p.y = g.star[1];
                        p.z = g.star[2];
                        p.ix = g.starIdx[0];
                        p.iy = g.starIdx[1];
                        p.iz = g.starIdx[2];
                    } else {
                        rng.init(a.seed, id, ctrBase);
                        p.x = a.star[0];
                        p.y = a.star[1];
                        p.z = a.star[2];
                        p.ix = a.starIdx[0];
                        p.iy = a.starIdx[1];
                        p.iz = a.starIdx[2];
                    }
                    p.stellar = 1;
                    p.emits = 0;
                    p.state = ST_EMIT;
                } else {
                    more = false;
                }
            }
        }
        const unsigned long long emitters = __ballot(p.state == ST_EMIT);
        const unsigned long long flyers = __ballot(p.state == ST_FLY);
        if (!emitters && !flyers && !__ballot(more)) break;
        if (emitters && (__popcll(emitters) >= emitBatch || !flyers || !__ballot(more))) {
            if (p.state == ST_EMIT) emit(a, g, p, rng, c, hot);
        }
#pragma unroll 1
        for (int k = 0; k < segsPerTrip; ++k)
            if (p.state == ST_FLY) segment<T, NEST>(a, g, p, rng, c, hot);
    }
    if constexpr (HOT) {  // every wave of the workgroup is done: its partials out
        __syncthreads();
        float* out = a.hotPart + (size_t)blockIdx.x * (size_t)(a.nHotB * a.nHotC);
        for (int i = threadIdx.x; i < a.nHotB * a.nHotC; i += blockDim.x) out[i] = hot[i];
    }
    // all lanes of the wave have left the loop: reduce the counters once
    unsigned long long* cnt = a.counters;
    using u64 = unsigned long long;
    const u64 s0 = wave_sum((u64)c.seg), s1 = wave_sum((u64)c.emit), s2 = wave_sum((u64)c.absb),
              s3 = wave_sum((u64)c.sca), s4 = wave_sum((u64)c.esc), s5 = wave_sum((u64)c.line),
              s6 = wave_sum((u64)c.lost), s7 = wave_sum((u64)c.trap), s8 = wave_sum((u64)c.err);
    const double q = wave_sum(c.qphot);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&cnt[MOC_CNT_SEGMENTS], s0);
        atomicAdd(&cnt[MOC_CNT_EMISSIONS], s1);
        atomicAdd(&cnt[MOC_CNT_ABSORBED], s2);
        atomicAdd(&cnt[MOC_CNT_SCATTERED], s3);
        atomicAdd(&cnt[MOC_CNT_ESCAPED], s4);
        atomicAdd(&cnt[MOC_CNT_LINES], s5);
        atomicAdd(&cnt[MOC_CNT_LOST], s6);
        atomicAdd(&cnt[MOC_CNT_TRAPPED], s7);
        atomicAdd(&cnt[MOC_CNT_ERROR], s8);
        atomicAdd(a.Qphot, q);
    }
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) d[k] = (float)s[k];
}

__global__ void k_f32_to_f64(const float* __restrict__ s, double* __restrict__ d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) d[k] = (double)s[k];
}

__global__ void k_scale(double* __restrict__ p, int64_t n, double f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) p[k] *= f;
}

// zero fill, for code that a caller may capture in a graph (no memset nodes)
__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) p[k] = 0ull;
}

// device-to-device copy, for code that a caller may capture in a graph
__global__ void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += st) dst[k] = src[k];
}

// The hot block's partial sums of every workgroup added to J: one key per
// lane, the rows summed in FP64 in workgroup order (deterministic for given
// partials); each key belongs to one slot, so the read-modify-write of J is
// plain.
__global__ void k_hot_reduce(const float* __restrict__ part, int rows, int nHotC, int nHotB,
                             const int* __restrict__ cellList, const int* __restrict__ binList,
                             double* __restrict__ J, int64_t C1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nHot = nHotC * nHotB;
    if (k >= nHot) return;
    const int bin = binList[k / nHotC], cell = cellList[k % nHotC];
    if (bin < 0 || cell <= 0) return;
    double sum = 0.;
    for (int r = 0; r < rows; ++r) sum += (double)part[(size_t)r * nHot + k];
    if (sum != 0.) J[(size_t)bin * C1 + cell] += sum;
}

// The hot bins of the next launch: the nHotB bins with the most J (path per
// volume) summed over the hot cells, ranked with ties to the lower bin.
// k_hot_binsum: one workgroup per bin, the cells over the lanes (summed in a
// fixed order); k_hot_select: one workgroup ranks the B sums, B <= kHotMaxBins.
constexpr int kHotMaxBins = 8192;
__global__ void __launch_bounds__(256) k_hot_binsum(const double* __restrict__ J, int64_t C1,
                                                    const int* __restrict__ cellList, int nHotC,
                                                    double* __restrict__ S) {
    __shared__ double w[256];
    const int b = blockIdx.x;
    double v = 0.;
    for (int k = threadIdx.x; k < nHotC; k += 256) {
        const int cell = cellList[k];
        if (cell > 0) v += J[(size_t)b * C1 + cell];
    }
    w[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) w[threadIdx.x] += w[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) S[b] = w[0];
}

__global__ void __launch_bounds__(1024) k_hot_select(const double* __restrict__ Sg, int B, int nHotB,
                                                     int* __restrict__ binMap, int* __restrict__ binList) {
    __shared__ double S[kHotMaxBins];
    for (int b = threadIdx.x; b < B; b += blockDim.x) S[b] = Sg[b];
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const double v = S[b];
        int rank = 0;
        for (int q = 0; q < B; ++q) rank += (S[q] > v || (S[q] == v && q < b)) ? 1 : 0;
        const bool in = rank < nHotB;
        binMap[b] = in ? rank : -1;
        if (in) binList[rank] = b;
    }
}

// LDS for the staged geometry of one workgroup (2 x (nx+ny+nz) values)
constexpr size_t kMaxGeoLds = 32768;
// workgroup of the transport with a hot J block: 16 waves, one workgroup per CU
constexpr int kHotBlock = 1024;

// ms == nullptr: enqueue only (no events, no wait), as inside a graph capture
template <typename T>
int launch(moc_ctx* ctx, const MocTransportArgs<T>& a, float* ms) {
    if (a.n <= 0) {
        if (ms) *ms = 0.f;
        return MOC_OK;
    }
    if (a.emitBatch < 1 || a.emitBatch > 64 || a.segsPerTrip < 1 || a.segsPerTrip > 64) {
        ctx->err = "transport: tuning parameters out of range (emit batch 1..64, segments per trip 1..64)";
        return MOC_ERR_ARG;
    }
    constexpr bool kF32 = std::is_same<T, float>::value;
    const bool hot = kF32 && a.nGrids == 0 && a.nHotC > 0 && a.nHotB > 0;
    // the resident wave count (a tuning parameter; the default fills the 256
    // CUs several times over); lanes beyond the packet count exit at their
    // first fetch
    const int64_t lanes_max = 64LL * (int64_t)ctx->tune.transportWaves;
    int64_t lanes = a.n < lanes_max ? a.n : lanes_max;
    const int block = hot ? kHotBlock : 256;
    const int blocks = (int)((lanes + block - 1) / block);
    const size_t geo = a.nGrids > 0 ? 0 : 2 * (size_t)(a.nx + a.ny + a.nz) * sizeof(T);
    if (geo > kMaxGeoLds) {
        ctx->err = "transport: the grid axes (" + std::to_string(a.nx + a.ny + a.nz) +
                   " points) exceed the " + std::to_string(kMaxGeoLds) + " B LDS staging area";
        return MOC_ERR_ARG;
    }
    const size_t lds = hot ? geo + (size_t)a.nHotC * a.nHotB * sizeof(float) : geo;
    if (hot && (blocks > a.hotBlocks || a.nHotB > kHotMaxBins || a.nbins > kHotMaxBins)) {
        ctx->err = "transport: the hot J block's partial buffer does not fit this launch";
        return MOC_ERR_STATE;
    }
    if (ms) (void)hipEventRecord(ctx->ev0, ctx->stream);
    if (a.nGrids > 0 && a.nested)
        hipLaunchKernelGGL((k_transport<T, true, false, true>), dim3(blocks), dim3(block), lds, ctx->stream, a);
    else if (a.nGrids > 0)
        hipLaunchKernelGGL((k_transport<T, true, false, false>), dim3(blocks), dim3(block), lds, ctx->stream, a);
    else if (hot)
        hipLaunchKernelGGL((k_transport<T, false, kF32, false>), dim3(blocks), dim3(block), lds, ctx->stream, a);
    else
        hipLaunchKernelGGL((k_transport<T, false, false, false>), dim3(blocks), dim3(block), lds, ctx->stream, a);
    if (hot) {
        // the partials into J, then the hot bins of the next launch from this J
        const int nHot = a.nHotC * a.nHotB;
        hipLaunchKernelGGL(k_hot_reduce, dim3((unsigned)((nHot + 255) / 256)), dim3(256), 0, ctx->stream,
                           a.hotPart, blocks, a.nHotC, a.nHotB, a.hotCellList, a.hotBinList, a.J,
                           (int64_t)a.nCells + 1);
        hipLaunchKernelGGL(k_hot_binsum, dim3((unsigned)a.nbins), dim3(256), 0, ctx->stream, a.J,
                           (int64_t)a.nCells + 1, a.hotCellList, a.nHotC, a.hotS);
        hipLaunchKernelGGL(k_hot_select, dim3(1), dim3(1024), 0, ctx->stream, a.hotS, a.nbins, a.nHotB,
                           const_cast<int*>(a.hotBin), const_cast<int*>(a.hotBinList));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("transport launch: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    if (!ms) return MOC_OK;
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    e = hipEventSynchronize(ctx->ev1);
    if (e != hipSuccess) {
        ctx->err = std::string("transport kernel: ") + hipGetErrorString(e);
        return MOC_ERR_HIP;
    }
    (void)hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return MOC_OK;
}

}  // namespace

int moc_launch_transport_f64(moc_ctx* ctx, const MocTransportArgs<double>& a, float* ms) {
    return launch<double>(ctx, a, ms);
}
int moc_launch_transport_f32(moc_ctx* ctx, const MocTransportArgs<float>& a, float* ms) {
    return launch<float>(ctx, a, ms);
}
int moc_transport_hot_block() { return kHotBlock; }
size_t moc_transport_hot_lds(int nx, int ny, int nz, int nHot) {
    return 2 * (size_t)(nx + ny + nz) * sizeof(float) + (size_t)nHot * sizeof(float);
}
void moc_convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}
void moc_convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}
void moc_scale_f64(double* p, int64_t n, double f, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(256), 0, s, p, n, f);
}

void moc_fill_zero_u64(unsigned long long* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
}

void moc_copy_f64(const double* src, double* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
}

"""Host-side table builders (FP64) for the Monte Carlo iteration.

These restate MOCASSIN's set-up code for the tables the GPU path consumes:
frequency mesh and bin widths, point-source spectrum CDF, Klein-Nishina
tables, Cartesian axes and the active-cell map.  They run once per model
on the host; the per-iteration work is on the device.
"""
from __future__ import annotations

import numpy as np

# constants_mod.f90 / continuum_mod.f90 (values as the reference defines them)
PI_REF = 3.141592654            # constants_mod.f90:45
H_PLANCK = 6.6262e-27           # constants_mod.f90:14
HCRYD_K = 157893.94             # continuum_mod.f90:14  h c Ryd / k [K]
HCRYD = 2.1799153e-11           # continuum_mod.f90:15  h c Ryd [erg]
RADIO_4P9GHZ = 1.489434e-6      # constants_mod.f90:19 [Ryd]
ME_C2_RYD = 37557.66267         # ph_mod.f90:1933 m_e c^2 [Ryd]
HALF_RE2 = 3.970393838e-26      # ph_mod.f90:1958 0.5 r_e^2 [cm^2]

# H-series edges of the reference gas mesh (grid_mod.f90:169-171)
SERIES_EDGES = np.array([0.0069, 0.0083, 0.01, 0.0123, 0.0156, 0.0204, 0.0278, 0.04,
                         0.0625, 0.11117, 0.11610, 0.12248, 0.13732, 0.24763, 0.24994,
                         0.26630, 0.29189])


def widths_of(nu: np.ndarray) -> np.ndarray:
    """widFlx (grid_mod.f90:378-382): centred differences, one-sided at the ends."""
    w = np.empty_like(nu)
    w[0] = nu[1] - nu[0]
    w[1:-1] = (nu[2:] - nu[:-2]) / 2.0
    w[-1] = nu[-1] - nu[-2]
    return w


def log_mesh(nbins: int, nu_min: float, nu_max: float) -> np.ndarray:
    """nbins log-spaced points from nu_min to nu_max [Ryd] (the build's own mesh)."""
    return np.logspace(np.log10(nu_min), np.log10(nu_max), nbins)


def reference_gas_mesh(nbins: int, nu_min: float, nu_max: float,
                       ion_edges: np.ndarray, strict: bool = True) -> np.ndarray:
    """The gas-only log mesh of initCartesianGrid (grid_mod.f90:251-375).

    Each H-series edge and each ionisation edge below nuMax adds the triplet
    (e - 0.0003, e, e + 0.0003); edges below nuMin pull the log part of the
    mesh down to (edge - 0.0006); the rest of the nbins points are log-spaced
    by repeated multiplication; the result is sorted ascending (sortUp keeps
    distinct values).  Where edges coincide the mesh has fewer than nbins
    distinct points: an error when strict, else the shorter mesh.
    """
    pts = []
    nu_lo, nu_hi = nu_min, nu_max
    if nu_min < RADIO_4P9GHZ:
        pts.append(RADIO_4P9GHZ)
    for e in SERIES_EDGES:
        pts += [e, e - 0.0003, e + 0.0003]
        if e < nu_lo:
            nu_lo = e - 0.0006
        if e > nu_hi:
            nu_hi = e
    for e in sorted(ion_edges):
        if e < nu_hi:
            pts += [e, e - 0.0003, e + 0.0003]
            if e < nu_lo:
                nu_lo = e - 0.0006
    n_count = len(pts) + 1                     # Fortran nuCount after the triplets
    i_count = nbins - n_count + 1
    step = (np.log10(nu_hi) - np.log10(nu_lo)) / (i_count - 1)
    logs = [nu_lo]
    for _ in range(i_count - 1):
        logs.append(10.0 ** (np.log10(logs[-1]) + step))
    mesh = np.unique(np.array(pts + logs, dtype=np.float64))
    if strict and mesh.size != nbins:
        raise ValueError(f"mesh has {mesh.size} distinct points, expected {nbins}")
    return mesh


def blackbody_flux(nu: np.ndarray, T: float) -> np.ndarray:
    """getFlux('blackbody') (continuum_mod.f90:358-392) [erg/cm^2/s/Hz/sr]."""
    const = 0.5250229 / H_PLANCK
    x = HCRYD_K * nu / T
    out = np.empty_like(nu)
    big = x > 86.0
    out[big] = const * nu[big] ** 3 * np.exp(-x[big])
    small = ~big
    den = np.exp(x[small]) - 1.0
    ok = den > 0
    o = np.empty(small.sum())
    o[ok] = const * nu[small][ok] ** 3 / den[ok]
    o[~ok] = 3.32154e-6 * nu[small][~ok] ** 2 * T / H_PLANCK
    out[small] = o
    return out


def powerlaw_flux(nu: np.ndarray, index: float, nu_cut_lo: float = 0.0,
                  nu_cut_hi: float = np.inf) -> np.ndarray:
    """getFlux('powerlaw') (continuum_mod.f90:394-398)."""
    f = nu ** (-index)
    if np.isfinite(nu_cut_hi):
        f = f * np.exp(-nu / nu_cut_hi)
    if nu_cut_lo > 0:
        f = f * np.exp(-nu_cut_lo / nu)
    return f


def source_cdf(flux: np.ndarray, widflx: np.ndarray, start: int = 1) -> np.ndarray:
    """inSpectrumProbDen (setProbDen, continuum_mod.f90:415-504): cumulative
    flux*width from the 1-based bin `start` on (0 below it), normalised,
    entries at the maximum set to exactly 1."""
    s = max(int(start), 1) - 1
    cdf = np.zeros(np.asarray(flux).size)
    cdf[s:] = np.cumsum((flux * widflx)[s:])
    cdf = cdf / cdf[-1]
    cdf[cdf >= cdf.max()] = 1.0
    return cdf


def compton_tables(nu: np.ndarray, n_angles: int = 180):
    """setCompton / Pcomp / KNsigma (ph_mod.f90:1883-1963).

    Returns (KNsigmaT[B], KNcdf[B, n_angles], Pcomp[B, n_angles]).  The
    angular quadrature is the reference's: 1-degree steps weighted by
    2 pi sin(theta) dtheta.
    """
    th_deg = np.arange(1, n_angles + 1, dtype=np.float64)
    th = th_deg * PI_REF / 180.0
    P = 1.0 / (1.0 + (nu[:, None] / ME_C2_RYD) * (1.0 - np.cos(th[None, :])))
    ks = HALF_RE2 * (P - P ** 2 * np.sin(th[None, :]) ** 2 + P ** 3)
    wgt = 2.0 * PI_REF * np.sin(th_deg * PI_REF / 180.0) * PI_REF / 180.0
    contrib = ks * wgt[None, :]
    sig_t = contrib.sum(axis=1)
    cdf = np.cumsum(contrib, axis=1) / sig_t[:, None]
    cdf[:, -1] = 1.0
    return sig_t, cdf, P


def symmetric_axes(n: int, R: float) -> np.ndarray:
    """Octant axis (fillGrid, grid_mod.f90:698-711): (i-1)/(n-1) * R."""
    return np.arange(n, dtype=np.float64) / (n - 1) * R


def full_axes(n: int, R: float) -> np.ndarray:
    """Whole-sphere axis (grid_mod.f90:738-747): 2(i-1)/(n-1) - 1, times R."""
    if n % 2 == 0:
        raise ValueError("the automatic non-symmetric grid requires odd n (grid_mod.f90:731)")
    return (2.0 * np.arange(n, dtype=np.float64) / (n - 1) - 1.0) * R


def active_map(xa, ya, za, hden: np.ndarray, R_in: float = 0.0, R_out: float = 0.0):
    """Active-cell numbering of setMotherGrid (grid_mod.f90:1242-1295).

    hden is indexed [i, j, k].  Cells with r < R_in, r > R_out (if R_out > 0)
    or zero density are inactive; the others are numbered 1..nCells with i
    outermost and k innermost.  Returns (active[nx,ny,nz] int32, nCells,
    hden_per_cell[nCells+1]).
    """
    nx, ny, nz = len(xa), len(ya), len(za)
    X, Y, Z = np.meshgrid(xa, ya, za, indexing="ij")
    r = 1.e10 * np.sqrt((X / 1.e10) ** 2 + (Y / 1.e10) ** 2 + (Z / 1.e10) ** 2)
    on = hden > 0.0
    on &= ~(r < R_in)
    if R_out > 0:
        on &= ~(r > R_out)
    act = np.zeros((nx, ny, nz), dtype=np.int32)
    flat_on = on.reshape(-1)            # C order over (i, j, k): k fastest
    ids = np.cumsum(flat_on).astype(np.int32)
    act.reshape(-1)[flat_on] = ids[flat_on]
    ncells = int(flat_on.sum())
    hcell = np.zeros(ncells + 1)
    hcell[1:] = hden.reshape(-1)[flat_on]
    return act, ncells, hcell


def cell_volumes(xa, ya, za, active, sym: bool, plane: bool = False) -> np.ndarray:
    """dV per active cell [1e45 cm^3] as pathSegment uses it (photon_mod.f90:2169-2223):
    the first cell of an axis is half-width in an octant or plane-parallel
    grid, the last one in a plane-parallel grid only."""
    def w(a):
        n = len(a)
        d = np.empty(n)
        d[1:-1] = np.abs(a[2:] - a[:-2]) / 2.0
        d[0] = np.abs(a[1] - a[0]) / (2.0 if (sym or plane) else 1.0)
        d[-1] = np.abs(a[-1] - a[-2]) / (2.0 if plane else 1.0)
        return d / 1.e15
    V = w(xa)[:, None, None] * w(ya)[None, :, None] * w(za)[None, None, :]
    ncells = int(active.max())
    out = np.zeros(ncells + 1)
    on = active > 0
    out[active[on]] = V[on]
    return out


# ------------------------------------------------ plane-parallel illumination
def _locate(xa: np.ndarray, x: float) -> int:
    """locate (interpolation_mod.f90:49-106): 1-based bracket index, 0 below
    the first point, n at or above the last."""
    xa = np.asarray(xa)
    if x > xa[-1]:
        return xa.size
    if x < xa[0]:
        return 0
    return int(np.searchsorted(xa, x, side="right"))


def _locate_mid(xa: np.ndarray, x: float) -> int:
    """locate, then the next point when x lies above the midpoint
    (grid_mod.f90:1771-1784)."""
    p = _locate(xa, x)
    if 1 <= p < len(xa) and x > (xa[p - 1] + xa[p]) / 2.0:
        p += 1
    return p


def plane_R_out(xa, ya, za) -> float:
    """R_out of a plane-parallel model: the distance of the far corner, so the
    radius test never cuts the slab (grid_mod.f90:852-858)."""
    return float(1.e10 * np.sqrt((xa[-1] / 1.e10) ** 2 + (ya[-1] / 1.e10) ** 2 +
                                 (za[-1] / 1.e10) ** 2))


def plane_area(xa, za) -> float:
    """getPlaneArea (grid_mod.f90:2989-3047): the illuminated face [1e36 cm^2],
    summed over (x, z) cells with half-width end cells."""
    def w(a):
        a = np.asarray(a, dtype=np.float64)
        d = np.empty(a.size)
        d[1:-1] = np.abs(a[2:] - a[:-2]) / 2.0
        d[0] = np.abs(a[1] - a[0]) / 2.0
        d[-1] = np.abs(a[-1] - a[-2]) / 2.0
        return d / 1.e18
    return float(np.sum(w(xa)[:, None] * w(za)[None, :]))


def plane_source(nu: np.ndarray, widflx: np.ndarray, flux_erg: np.ndarray, meanFieldin: float,
                 nu0: float, nu0Add: float, xa, za):
    """The source of a plane-parallel model (planeIonization meanField nu0
    nu0Add).  Returns (Lstar [1e36 erg/s], meanField [erg/cm^2/s], sourceCDF).

    With nu0 > 0, meanFieldin is the photon flux above nu0 [phot/cm^2/s]: the
    spectrum is scaled to it and meanField becomes the energy flux from nu0Add
    on (setMotherGrid, grid_mod.f90:1764-1815); with nu0 = 0 it is the energy
    flux itself.  Lstar = meanField/1e36 times the face (x_n - x_1)(z_n - z_1)
    (:1818-1820).  The CDF starts at locate(nu, nu0Add) when nu0Add > 0
    (setProbDen, continuum_mod.f90:472-482, which runs before setMotherGrid
    and so sees the unclipped spectrum)."""
    nu = np.asarray(nu, dtype=np.float64)
    wid = np.asarray(widflx, dtype=np.float64)
    erg = np.asarray(flux_erg, dtype=np.float64)
    nb = nu.size
    meanField = float(meanFieldin)
    if nu0 > 0.0:
        nu0P = _locate_mid(nu, nu0)
        nu0AddP = _locate_mid(nu, nu0Add) if nu0Add != nu0 else nu0P
        if nu0P >= nb or nu0P < 1:
            raise ValueError("plane_source: nu0 outside the frequency mesh (grid_mod.f90:1786-1790)")
        phot = erg / (nu * HCRYD)
        scale = meanField / float(np.sum(phot[nu0P - 1:] * wid[nu0P - 1:]))
        clip = erg.copy()
        clip[:max(nu0AddP, 1) - 1] = 0.0
        meanField = float(np.sum(clip * wid)) * scale
    Lstar = meanField / 1.e36 * (xa[-1] - xa[0]) * (za[-1] - za[0])
    start = _locate(nu, nu0Add) if nu0Add > 0.0 else 1
    return float(Lstar), meanField, source_cdf(erg, wid, start)

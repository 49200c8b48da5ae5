"""Host-side grain tables (FP64) for the dust path.

Restates the set-up that MOCASSIN's dust runs do once per model:
  grain sizes and weights           makeDustXsec  ph_mod.f90:1033-1081
  optical constants onto the mesh   makeDustXsec  ph_mod.f90:1195-1213 (linearMap)
  per-grain cross-sections          makeDustXsec  ph_mod.f90:1502-1524
  Kirchhoff emission integrals      dustEmissionInt  dust_mod.f90:146-184
  grain number density from MdMg    setMotherGrid  grid_mod.f90:1690-1727
The efficiencies Q_abs, Q_sca follow getQs (ph_mod.f90:1604-1667): Mie
theory for homogeneous spheres (the Bohren & Huffman (1983, appendix A)
recurrences) while |m| x < 1000, with x capped at 100; the Rayleigh-Gans
forms where |m| x >= 1000 and |m - 1| x < 0.001; the reference's geometric
optics otherwise (geometricOptics, :1669-1735).  The device path consumes
whatever tables it is given.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np

from .tables import blackbody_flux, H_PLANCK

DATA = Path(__file__).resolve().parent / "data" / "dust.npz"
C_LIGHT = 2.9979250e10       # constants_mod.f90:12 [cm/s]
FR1RYD = 3.28984e15          # constants_mod.f90:13 [Hz]
AMU = 1.66053e-24            # constants_mod.f90:11 [g]
PI_REF = 3.141592654         # constants_mod.f90:45
N_TEMPS = 3000               # constants_mod.f90:69
# atomic weights (grid_mod.f90:1661-1664)
A_WEIGHT = np.array([1.0080, 4.0026, 6.941, 9.0122, 10.811, 12.0111, 14.0067, 15.9994,
                     18.9984, 20.179, 22.9898, 24.305, 26.9815, 28.086, 30.9738, 32.06, 35.453,
                     39.948, 39.102, 40.08, 44.956, 47.90, 50.9414, 51.996, 54.9380, 55.847,
                     58.9332, 58.71, 63.546, 65.37])


@lru_cache(maxsize=1)
def _data() -> dict:
    d = np.load(DATA)
    return {k: d[k] for k in d.files}


def optical_constants(species: str):
    """(wav [um], n, k, [Tsub, rho, Vn, Msurf]) of a packed n,k file."""
    d = _data()
    return d[f"{species}_wav"], d[f"{species}_n"], d[f"{species}_k"], d[f"{species}_hdr"]


def grain_kind(species: str) -> int:
    """qHeat's class of a species from the first letter of its file label
    (emission_mod.f90:2070-2096): 1 silicate ('S'), 2 carbon ('C'), 0 any
    other (no temperature distribution; e.g. the graphite files' 'g')."""
    name = str(_data()[f"{species}_name"])
    return {"S": 1, "C": 2}.get(name[:1], 0)


def linear_map(y: np.ndarray, x: np.ndarray, x_new: np.ndarray) -> np.ndarray:
    """linearMap (interpolation_mod.f90:111-131): clamp outside [x1, xn]."""
    out = np.empty_like(x_new)
    for i, xv in enumerate(x_new):
        if xv > x[-1]:
            out[i] = y[-1]
            continue
        if xv < x[0]:
            out[i] = y[0]
            continue
        hit = np.nonzero(x == xv)[0]
        ii = int(hit[0]) + 1 if hit.size else int(np.searchsorted(x, xv, side="right"))
        if ii == x.size:
            out[i] = y[-1]
        else:
            out[i] = y[ii - 1] + (y[ii] - y[ii - 1]) * (xv - x[ii - 1]) / (x[ii] - x[ii - 1])
    return out


def refractive_index_on_mesh(species: str, nu: np.ndarray):
    """Re(m), Im(m) on the frequency mesh: wavelengths become Ryd
    (nu = c / (lambda fr1Ryd 1e-4)), reversed to ascending nu, then linearMap."""
    wav, n, k, _ = optical_constants(species)
    e = C_LIGHT / (wav * FR1RYD * 1.e-4)
    e, n, k = e[::-1], n[::-1], k[::-1]
    return linear_map(n, e, nu), linear_map(k, e, nu)


def mie_q(x: np.ndarray, m: np.ndarray):
    """Q_ext, Q_sca of homogeneous spheres, size parameters x (array) and
    complex refractive indices m (same shape): the Bohren & Huffman series
    with D_n(mx) from downward recurrence and psi_n, chi_n upward."""
    with np.errstate(all="ignore"):  # the recurrences of finished (small-x) entries overflow
        return _mie_q(np.asarray(x, dtype=np.float64), np.asarray(m, dtype=np.complex128))


def _mie_q(x: np.ndarray, m: np.ndarray):
    mx = m * x
    nstop = np.floor(x + 4.0 * np.cbrt(x) + 2.0).astype(int)
    nmx = int(max(nstop.max(), np.abs(mx).max())) + 16
    D = np.zeros((nmx + 1,) + x.shape, dtype=np.complex128)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / mx - 1.0 / (D[n] + n / mx)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = psi1 - 1j * chi1
    qext = np.zeros_like(x)
    qsca = np.zeros_like(x)
    for n in range(1, int(nstop.max()) + 1):
        fn = float(n)
        psi = (2.0 * fn - 1.0) * psi1 / x - psi0
        chi = (2.0 * fn - 1.0) * chi1 / x - chi0
        xi = psi - 1j * chi
        da = D[n] / m + fn / x
        db = m * D[n] + fn / x
        an = (da * psi - psi1) / (da * xi - xi1)
        bn = (db * psi - psi1) / (db * xi - xi1)
        live = n <= nstop
        qsca += np.where(live, (2.0 * fn + 1.0) * (np.abs(an) ** 2 + np.abs(bn) ** 2), 0.0)
        qext += np.where(live, (2.0 * fn + 1.0) * np.real(an + bn), 0.0)
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = psi1 - 1j * chi1
    f = 2.0 / (x * x)
    return qext * f, qsca * f


PI_GETQS = 3.14159265        # the pi of getQs's size parameter (ph_mod.f90:1628)


def geometric_optics(x: float, a: float, mre: float, mim: float, lam: float):
    """geometricOptics (ph_mod.f90:1669-1735) as the reference computes it:
    45 incidence angles of 1 degree for Q_abs, then Q_sca from the
    Rayleigh-Gans Q saturated to 2 minus Q_abs.  Kept from the reference:
    1/eps's imaginary part taken as -Re(eps)/|eps|^2, and the refraction angle
    psi of the previous angle in the exit-face terms p21, q21 (psi starts at
    0; the reference reads it uninitialised at the first angle)."""
    absmm1sq = (mre - 1.) ** 2 + mim ** 2
    epsre = mre ** 2 - mim ** 2
    epsim = mre * mim
    epsrem1 = epsre / (epsre ** 2 + epsim ** 2)
    epsimm1 = -(epsre / (epsre ** 2 + epsim ** 2))
    qa = 0.
    psi = 0.
    for i in range(1, 46):
        th = i * PI_REF / 180.
        s2 = np.sin(th) ** 2
        p20 = 0.5 * (s2 - epsre + ((2 * epsim) ** 2 + (epsre - s2)) ** 0.5)
        q20 = 0.5 * (-s2 + epsre + ((2 * epsim) ** 2 + (epsre - s2)) ** 0.5)
        sp2 = np.sin(psi) ** 2
        p21 = 0.5 * (sp2 - epsrem1 + ((2 * epsimm1) ** 2 + (epsrem1 - sp2)) ** 0.5)
        q21 = 0.5 * (-sp2 + epsrem1 + ((2 * epsimm1) ** 2 + (epsrem1 - sp2)) ** 0.5)
        rperp0 = ((np.sqrt(q20) - np.cos(th)) ** 2 + p20) / ((np.sqrt(q20) + np.cos(th)) ** 2 + p20)
        rpara0 = rperp0 * ((np.sqrt(q20) - np.sin(th) * np.tan(th)) ** 2 + p20) / \
            ((np.sqrt(q20) + np.sin(th) * np.tan(th)) ** 2 + p20)
        psi = np.arcsin(np.sin(th) / (q20 + s2) ** 2)
        rperp1 = ((np.sqrt(q21) - np.cos(psi)) ** 2 + p21) / ((np.sqrt(q21) + np.cos(psi)) ** 2 + p21)
        rpara1 = rperp1 * ((np.sqrt(q21) - np.sin(psi) * np.tan(psi)) ** 2 + p21) / \
            ((np.sqrt(q21) + np.sin(psi) * np.tan(psi)) ** 2 + p21)
        at = np.exp(-8. * PI_REF * a * mim * np.cos(psi) / lam)
        qa += np.sin(th) * np.cos(th) * (PI_REF / 180.) * (1 - at) * \
            ((1. - rperp0) / (1. - at * rperp1) + (1. - rpara0) / (1. - at * rpara1))
    qrg = (32. * absmm1sq * x ** 4) / (27. + 16 * x ** 2) + (8. / 3.) * mim * x
    qs = qrg / np.sqrt(1. + 0.25 * qrg ** 2) - qa
    return qa, qs


def get_qs(n_re: np.ndarray, n_im: np.ndarray, a: float, lam_um: np.ndarray,
           scattering: bool = True):
    """Q_abs, Q_sca of grains of radius a [um] over the bins (getQs,
    ph_mod.f90:1604-1667): the regime of each bin from |m| x and |m - 1| x."""
    n_re = np.asarray(n_re, dtype=np.float64)
    n_im = np.asarray(n_im, dtype=np.float64)
    x = 2.0 * PI_GETQS * a / np.asarray(lam_um, dtype=np.float64)
    absm = np.sqrt(n_re ** 2 + n_im ** 2)
    absm1 = np.sqrt((n_re - 1.) ** 2 + n_im ** 2)
    qabs = np.zeros_like(x)
    qsca = np.zeros_like(x)
    mie = absm * x < 1000.
    rg = ~mie & (absm1 * x < 0.001)
    geo = ~mie & ~rg
    if mie.any():
        xm = np.minimum(x[mie], 100.)              # geometrical-optics cap of the series
        qe, qs = mie_q(xm, n_re[mie] + 1j * n_im[mie])
        qabs[mie] = qe - qs
        qsca[mie] = qs
    if rg.any():
        qabs[rg] = (8. / 3.) * n_im[rg] * x[rg]
        qsca[rg] = 32. * absm1[rg] ** 2 * x[rg] ** 4 / (27. + 16. * x[rg] ** 2)
    for i in np.flatnonzero(geo):
        qabs[i], qsca[i] = geometric_optics(x[i], a, n_re[i], n_im[i], lam_um[i])
    if not scattering:
        qsca[:] = 0.
    return qabs, qsca


def size_weights(radius: np.ndarray, weight: np.ndarray) -> np.ndarray:
    """grainWeight (ph_mod.f90:1056-1081): weight x da, normalised; 1 for one size."""
    n = radius.size
    if n == 1:
        return np.array([1.0])
    da = np.empty(n)
    da[0] = radius[1] - radius[0]
    da[1:-1] = (radius[2:] - radius[:-2]) / 2.0
    da[-1] = radius[-1] - radius[-2]
    norm = np.sum(weight * da)
    return weight * da / norm


def mrn_sizes(n: int = 10, a_min: float = 0.005, a_max: float = 0.25, q: float = 3.5):
    """n log-spaced radii [um] with MRN weights a^-q (before the da factor)."""
    a = np.exp(np.linspace(np.log(a_min), np.log(a_max), n))
    return a, a ** (-q)


@dataclass
class DustTables:
    species: list            # packed species names
    grainAbun: np.ndarray    # [nS] number fractions of the species
    radius: np.ndarray       # [nSizes] um
    grainWeight: np.ndarray  # [nSizes]
    Tsub: np.ndarray         # [nS] sublimation temperatures [K]
    rho: np.ndarray          # [nS] bulk densities [g/cm^3]
    Cabs: np.ndarray         # [nS][nSizes][B] cm^2 per grain
    Csca: np.ndarray         # [nS][nSizes][B]
    emIntegral: np.ndarray   # [nS][nSizes][N_TEMPS] dustEmIntegral
    Vn: np.ndarray = None    # [nS] potential of a neutral grain [Ryd] (grainVn)
    Msurf: np.ndarray = None # [nS] surface atom mass [amu] (MsurfAtom)
    kind: np.ndarray = None  # [nS] qHeat class: 1 silicate, 2 carbon, 0 other (grain_kind)

    @property
    def nSpecies(self) -> int:
        return len(self.species)

    @property
    def nSizes(self) -> int:
        return self.radius.size


def emission_integrals(Cabs: np.ndarray, nu: np.ndarray, widflx: np.ndarray,
                       n_temps: int = N_TEMPS) -> np.ndarray:
    """dustEmIntegral(nS, ai, T) = 4 h sum_i Cabs_i B_i(T)/h fr1Ryd widFlx_i for
    T = 1 .. nTemps K (dust_mod.f90:146-184)."""
    nS, nA, B = Cabs.shape
    out = np.zeros((nS, nA, n_temps))
    for t in range(1, n_temps + 1):
        bb = blackbody_flux(nu, float(t)) * FR1RYD * widflx
        out[:, :, t - 1] = np.einsum("sab,b->sa", Cabs, bb)
    return out * H_PLANCK * 4.0


def build_dust(nu: np.ndarray, widflx: np.ndarray, species: list, abundances,
               radius: np.ndarray, weights: np.ndarray, scattering: bool = True) -> DustTables:
    """Grain tables of a species mixture and size distribution on the mesh."""
    radius = np.asarray(radius, dtype=np.float64)
    gw = size_weights(radius, np.asarray(weights, dtype=np.float64))
    nS, nA, B = len(species), radius.size, nu.size
    Cabs = np.zeros((nS, nA, B))
    Csca = np.zeros((nS, nA, B))
    Tsub = np.zeros(nS)
    rho = np.zeros(nS)
    vn = np.zeros(nS)
    ms = np.zeros(nS)
    lam_um = C_LIGHT / (nu * FR1RYD) * 1.e4
    for s, name in enumerate(species):
        n_re, n_im = refractive_index_on_mesh(name, nu)
        hdr = optical_constants(name)[3]
        Tsub[s], rho[s], vn[s], ms[s] = hdr[0], hdr[1], hdr[2], hdr[3]
        for a_i, a in enumerate(radius):
            qa, qs = get_qs(n_re, n_im, a, lam_um, scattering)
            geo = PI_REF * a * a * 1.e-8
            Cabs[s, a_i] = qa * geo
            Csca[s, a_i] = qs * geo
    emi = emission_integrals(Cabs, nu, widflx)
    kind = np.array([grain_kind(s) for s in species], dtype=np.int32)
    return DustTables(list(species), np.asarray(abundances, dtype=np.float64), radius, gw,
                      Tsub, rho, Cabs, Csca, emi, vn, ms, kind)


def ndust_from_mdmg(dt: DustTables, mdmg: float, hden: np.ndarray, elem_z, abund) -> np.ndarray:
    """Grain number density [cm^-3] for a dust-to-gas mass ratio
    (grid_mod.f90:1676-1727): gas mass per H times MdMg times n_H over the
    mean grain mass."""
    gas_cell = float(np.sum(np.asarray(abund) * A_WEIGHT[np.asarray(elem_z) - 1] * AMU))
    den = 0.0
    for s in range(dt.nSpecies):
        for a_i in range(dt.nSizes):
            den += (1.3333 * PI_REF * (dt.radius[a_i] * 1.e-4) ** 3 * dt.rho[s] *
                    dt.grainWeight[a_i] * dt.grainAbun[s])
    return gas_cell * mdmg * np.asarray(hden) / den


def ip_ryd(xsec) -> np.ndarray:
    """[31*32] (Z, stage): nu [Ryd] of the first bin of the outermost shell of
    stage - 1, the energy released when `stage` recombines (the IPerg of
    setDustGasCollHeatCool, update_mod.f90:796-806); 0 where undefined."""
    out = np.zeros(31 * 32)
    best = {}
    for sh in xsec.shells:
        k = (sh.elem, sh.ion)
        if k not in best or sh.shell > best[k].shell:
            best[k] = sh
    for (z, ion), sh in best.items():
        if ion + 1 < 32:
            out[z * 32 + ion + 1] = xsec.nu[sh.lo - 1]
    return out

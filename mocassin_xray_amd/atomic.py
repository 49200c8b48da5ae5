"""Photoionisation cross-section stack (host, FP64).

Restates the table set-up MOCASSIN performs once per model:
  phFitEl      ph_mod.f90:592-668   Verner et al. (1996) analytic fits
  getOuterShell hydro_mod.f90:89-182
  setShells / limitShell hydro_mod.f90:208-332 (elementP bin ranges)
  setPointers  hydro_mod.f90:422-677 (H, He level pointers, kShellLimitP)
  initXSecArray / makeOpacity ph_mod.f90:204-517, 861-921
and returns, besides the per-shell stack (for the per-shell photoionisation
rates of the balance), the dense per-species matrix sigma[species, bin] that
the opacity kernel contracts with the per-cell ion densities.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path

import numpy as np

RYD_TO_EV = 13.6056981          # constants_mod.f90:51
HI_ION_POT = 0.99946            # constants_mod.f90:15
K_SHELL_LIMIT = 7.35e4          # constants_mod.f90:40 [Ryd]
N_TOT = [1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7]
N_INN = [0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3, 3, 3, 3, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5]
LEVEL = [0, 0, 1, 0, 1, 2, 0]
ELEM_LABEL = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P",
              "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn"]
# setPointers calls setShells in this element order (hydro_mod.f90:584-641)
SHELL_ORDER = [6, 8, 7, 10, 11, 12, 13, 14, 15, 16, 17, 26, 18, 19, 20, 21, 22, 23, 24, 25, 9, 3,
               4, 5, 27, 28, 29, 30]


@lru_cache(maxsize=1)
def packed_data() -> dict:
    """Every table of data/atomic.npz (written by data/make_atomic.py)."""
    with np.load(Path(__file__).resolve().parent / "data" / "atomic.npz") as d:
        return {k: d[k] for k in d.files}


def verner():
    d = packed_data()
    return d["ph1"], d["ph2"]


def locate(xa: np.ndarray, x: float) -> int:
    """1-based bracket index (interpolation_mod.f90:49-106)."""
    n = xa.size
    if x > xa[-1]:
        return n
    if x < xa[0]:
        return 0
    hit = np.nonzero(xa == x)[0]
    if hit.size:
        return int(hit[0]) + 1
    return int(np.searchsorted(xa, x, side="right"))


def phfit_el(nz: int, ne: int, shell: int, E: float) -> float:
    """Verner partial cross-section [Mb] of shell `shell` of ion (nz, ne) at E [eV]."""
    ph1, ph2 = verner()
    if nz < 1 or nz > 30 or ne < 1 or ne > nz:
        return 0.0
    nout = N_TOT[ne - 1]
    if nz == ne and nz > 18:
        nout = 7
    if nz == ne + 1 and nz in (20, 21, 22, 25, 26):
        nout = 7
    if shell > nout:
        return 0.0
    p = ph1[:, nz - 1, ne - 1, shell - 1]
    if E < p[0]:
        return 0.0
    nint = N_INN[ne - 1]
    if nz in (15, 17, 19) or (nz > 20 and nz != 26):
        einn = 0.0
    else:
        einn = 1.e30 if ne < 3 else ph1[0, nz - 1, ne - 1, nint - 1]
    if shell < nout and shell > nint and E < einn:
        return 0.0
    if shell <= nint or E >= einn:
        p1 = -p[4]
        y = E / p[1]
        q = -0.5 * p1 - LEVEL[shell - 1] - 5.5
        a = p[2] * ((y - 1.0) ** 2 + p[5] ** 2)
        b = np.sqrt(y / p[3]) + 1.0
        return float(a * y ** q * b ** p1)
    p2 = ph2[:, nz - 1, ne - 1]
    p1 = -p2[3]
    q = -0.5 * p1 - 5.5
    x = E / p2[0] - p2[5]
    z = np.sqrt(x * x + p2[6] * p2[6])
    a = p2[1] * ((x - 1.0) ** 2 + p2[4] ** 2)
    b = np.sqrt(z / p2[2]) + 1.0
    return float(a * z ** q * b ** p1)


def outer_shell(z: int, nelec: int) -> int:
    """Outer shell number of ion (z, nelec) (getOuterShell, hydro_mod.f90:89-182)."""
    ss = [1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7]
    out = ss[nelec - 1]
    if z == nelec and z > 18:
        out = 7
    if z == nelec + 1 and z in (20, 21, 22, 25, 26):
        out = 7
    return out


def ion_edges(elem_on: list, nstages: int, nu_max: float) -> np.ndarray:
    """Ionisation edges [Ryd] of the switched-on ions up to nuMax that the
    gas frequency mesh resolves (initCartesianGrid, grid_mod.f90:170-209);
    each is held in single precision there."""
    ph1, _ = verner()
    out = []
    for z in sorted(set(elem_on)):
        for ion in range(1, min(z, nstages - 1) + 1):
            if z > 2:
                nel = z - ion + 1
                e = ph1[0, z - 1, nel - 1, outer_shell(z, nel) - 1] / RYD_TO_EV
            elif z == 1:
                e = 0.999434
            else:
                e = 1.80804 if ion == 1 else 4.0
            e = float(np.float32(e))
            if e <= nu_max:
                out.append(e)
    return np.sort(np.array(out, dtype=np.float64))


@dataclass
class Shell:
    elem: int      # Z
    ion: int       # ionisation stage (1 = neutral)
    shell: int
    lo: int        # 1-based first bin (elementP(...,1))
    hi: int        # 1-based last bin  (elementP(...,2))
    sigma: np.ndarray  # [hi-lo+1] cm^2


@dataclass
class XSecStack:
    nu: np.ndarray
    species: list            # [(elem, ion)] in opacity order
    sigma: np.ndarray        # [S, B] cm^2: summed shells per species
    shells: list             # [Shell]
    brems: np.ndarray        # [B] 1.0368e-18 / nu^3
    nstages: int
    lo_clamped: int          # shells whose threshold bin fell below the mesh


def build_xsec(nu: np.ndarray, elem_on: list, nstages: int) -> XSecStack:
    """Cross-section stack for the elements switched on (1-based Z list)."""
    B = nu.size
    kshell_p = locate(nu, K_SHELL_LIMIT)
    shells: list = []
    clamped = 0

    def lohi(lo, hi):
        nonlocal clamped
        if lo < 1:
            clamped += 1
            lo = 1
        return lo, min(hi, B)

    on = set(elem_on)
    # H I ground continuum (ph_mod.f90:389-394): from HlevNuP(1), E clamped to threshold
    ph1, _ = verner()
    if 1 in on:
        lo, hi = lohi(locate(nu, HI_ION_POT), B)
        E = np.maximum(nu[lo - 1:hi] * RYD_TO_EV, ph1[0, 0, 0, 0])
        shells.append(Shell(1, 1, 1, lo, hi, np.array([phfit_el(1, 1, 1, e) for e in E]) * 1e-18))
    if 2 in on:
        # He I singlet ground (ph_mod.f90:441-446), no clamp of E
        lo, hi = lohi(locate(nu, 1.80802), B)
        E = nu[lo - 1:hi] * RYD_TO_EV
        shells.append(Shell(2, 1, 1, lo, hi, np.array([phfit_el(2, 2, 1, e) for e in E]) * 1e-18))
        # He II (ph_mod.f90:456-461)
        lo, hi = lohi(locate(nu, 4.0), B)
        E = np.maximum(nu[lo - 1:hi] * RYD_TO_EV, ph1[0, 1, 0, 0])
        shells.append(Shell(2, 2, 1, lo, hi, np.array([phfit_el(2, 1, 1, e) for e in E]) * 1e-18))
    # heavy elements: setShells (bin ranges) then makeOperacity (values)
    eP = {}
    for z in SHELL_ORDER:
        if z not in on:
            continue
        for ion in range(1, z + 1):
            nel = z - ion + 1
            out = outer_shell(z, nel)
            for sh in range(1, out + 1):
                thres = ph1[0, z - 1, nel - 1, sh - 1] / RYD_TO_EV
                if thres <= 0.1:
                    eP[(z, ion, sh)] = [2, 1]
                else:
                    lo = locate(nu, thres)
                    if sh in (1, 2, 3):
                        hi = kshell_p
                    elif sh in (4, 5, 6):
                        hi = eP[(z, ion, 1)][0] - 1
                    else:
                        hi = (eP[(z, ion, 5)][0] if eP[(z, ion, 6)][0] < 3
                              else eP[(z, ion, 6)][0]) - 1
                    eP[(z, ion, sh)] = [lo, hi]
            # the outer shell's lower pointer is re-located with the last threshold
            eP[(z, ion, out)][0] = locate(nu, thres)
    for z in range(3, 31):
        if z not in on:
            continue
        for ion in range(1, min(z, nstages) + 1):
            nel = z - ion + 1
            for sh in range(1, outer_shell(z, nel) + 1):
                lo, hi = eP[(z, ion, sh)]
                if lo > hi:
                    continue
                lo, hi = lohi(lo, hi)
                E = np.maximum(nu[lo - 1:hi] * RYD_TO_EV, ph1[0, z - 1, nel - 1, sh - 1])
                sig = np.array([phfit_el(z, nel, sh, e) for e in E]) * 1e-18
                shells.append(Shell(z, ion, sh, lo, hi, sig))
    species = []
    for z in sorted(on):
        for ion in range(1, min(z, nstages) + 1):
            species.append((z, ion))
    sidx = {s: i for i, s in enumerate(species)}
    sigma = np.zeros((len(species), B))
    for s in shells:
        sigma[sidx[(s.elem, s.ion)], s.lo - 1:s.hi] += s.sigma
    brems = 1.03680e-18 / nu ** 3
    return XSecStack(nu, species, sigma, shells, brems, nstages, clamped)

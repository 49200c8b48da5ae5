"""Host-side tables of the ionisation/thermal balance (FP64).

Builds, from the packed atomic data and the cross-section stack of the
model's mesh, everything libmocgpu's balance kernels read:
  * the weight columns of the frequency integrals (photoionisation rate and
    heating per shell, getPhotoRates update_mod.f90:3971-4040 and the H/He
    branch of photoionisation :3639-3672; the Compton sums of compton
    :3834-3897 with the recoil cross-section of initXSecArray ph_mod.f90:369-385);
  * outer shells and statistical weights (getOuterShell hydro_mod.f90:89-182);
  * recombination, collisional-ionisation, Auger and charge-exchange tables;
  * the n-level ions with the natural-spline second derivatives of their
    Upsilon(log T) tables (spline, interpolation_mod.f90:207-283), which do
    not depend on T and are computed once here.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .atomic import XSecStack, locate, outer_shell, packed_data, verner

HCRYD = 2.1799153e-11  # update_mod.f90:64

# charge-exchange fits (a [1e-9 cm^3/s], b, c, d) by (Z, stage), in the order
# photoionisation / ionBalance list them (update_mod.f90:1624-1690); the last
# two "Ca" rows are written into argon's slots by the reference
CHEX_ROWS = [
    (2, 1, 7.47e-6, 2.06, 9.93, -3.89), (2, 2, 1.e-5, 0., 0., 0.),
    (3, 2, 1.26, 0.96, 3.02, -0.65), (3, 3, 1.e-5, 0., 0., 0.),
    (4, 2, 1.e-5, 0., 0., 0.), (4, 3, 1.e-5, 0., 0., 0.), (4, 4, 5.17, 0.82, -.69, -1.12),
    (5, 2, 2.e-2, 0., 0., 0.), (5, 3, 1.e-5, 0., 0., 0.), (5, 4, 5.27e-1, 0.76, -0.63, -1.17),
    (6, 1, 1.76e-9, 8.33, 4278.78, -6.41), (6, 2, 1.67e-4, 2.79, 304.72, -4.07),
    (6, 3, 3.25, 0.21, 0.19, -3.29), (6, 4, 332.46, -0.11, -0.995, -1.58e-3),
    (7, 1, 1.01e-3, -0.29, -0.92, -8.38), (7, 2, 3.05e-1, 0.60, 2.65, -0.93),
    (7, 3, 4.54, 0.57, -0.65, -0.89), (7, 4, 3.28, 0.52, -0.52, -0.19),
    (8, 1, 1.04, 3.15e-2, -0.61, -9.73), (8, 2, 1.04, 0.27, 2.02, -5.92),
    (8, 3, 3.98, 0.26, 0.56, -2.62), (8, 4, 2.52e-1, 0.63, 2.08, -4.16),
    (9, 2, 1.e-5, 0., 0., 0.), (9, 3, 9.86, 0.29, -0.21, -1.15), (9, 4, 7.15e-1, 1.21, -0.70, -0.85),
    (10, 2, 1.e-5, 0., 0., 0.), (10, 3, 14.73, 4.52e-2, -0.84, -0.31), (10, 4, 6.47, 0.54, 3.59, -5.22),
    (11, 2, 1.e-5, 0., 0., 0.), (11, 3, 1.33, 1.15, 1.20, -0.32), (11, 4, 1.01e-1, 1.34, 10.05, -6.41),
    (12, 2, 8.58e-5, 2.49e-3, 2.93e-2, -4.33), (12, 3, 6.49, 0.53, 2.82, -7.63),
    (12, 4, 6.36, 0.55, 3.86, -5.19), (13, 2, 1.e-5, 0., 0., 0.),
    (13, 3, 7.11e-5, 4.12, 1.72e4, -22.24), (13, 4, 7.52e-1, 0.77, 6.24, -5.67),
    (14, 2, 1.23, 0.24, 3.17, 4.18e-3), (14, 3, 4.900e-1, -8.74e-2, -0.36, -0.79),
    (14, 4, 7.58, 0.37, 1.06, -4.09), (16, 1, 3.82e-7, 11.10, 2.57e4, -8.22),
    (16, 2, 1.e-5, 0., 0., 0.), (16, 3, 2.29, 4.02e-2, 1.59, -6.06), (16, 4, 6.44, 0.13, 2.69, -5.69),
    (18, 2, 1.e-5, 0., 0., 0.), (18, 3, 4.57, 0.27, -0.18, -1.57), (18, 4, 6.37, 2.12, 10.21, -6.22),
    (18, 3, 3.17e-2, 2.12, 12.06, -0.40), (18, 4, 2.68, 0.69, -0.68, -4.47),
    (26, 2, 1.26, 7.72e-2, -0.41, -7.31), (26, 3, 3.42, 0.51, -2.06, -8.99),
]


def stat_weights(z: int, nel: int):
    """(g0, g1) of getOuterShell (hydro_mod.f90:133-179)."""
    gl = [2, 1, 2, 1, 6, 9, 4, 9, 6, 1, 2, 1, 6, 9, 4, 9, 6, 1, 2, 1, 10, 21, 28, 7, 6, 25,
          28, 21, 2, 1]
    glh = {19 + i: v for i, v in enumerate([10, 21, 28, 25, 6, 25, 28, 21, 10, 1, 2, 1])}
    g0 = gl[nel - 1]
    g1 = 1 if nel == 1 else gl[nel - 2]
    if z > 20 and nel >= 19 and z - nel >= 1:
        g0 = glh[nel]
        g1 = gl[nel - 2] if nel == 19 else glh[nel - 1]
    if z == nel:
        if z > 20 and nel >= 21:
            g1 = glh[nel - 1]
        g1 = {21: 15, 22: 28, 25: 7, 26: 30}.get(z, g1)
    if z - nel == 1:
        g0 = {21: 15, 22: 28, 25: 7, 26: 30}.get(z, g0)
    return g0, g1


def natural_spline_y2(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Second derivatives of the natural cubic spline through (x, y[..., k]),
    vectorised over the leading axes of y; the arithmetic follows spline()
    operation by operation."""
    n = x.size
    y2 = np.zeros_like(y)
    u = np.zeros_like(y)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[..., i - 1] + 2.0
        y2[..., i] = (sig - 1.0) / p
        ui = (y[..., i + 1] - y[..., i]) / (x[i + 1] - x[i]) - (y[..., i] - y[..., i - 1]) / (x[i] - x[i - 1])
        u[..., i] = (6.0 * ui / (x[i + 1] - x[i - 1]) - sig * u[..., i - 1]) / p
    y2[..., n - 1] = 0.0
    for k in range(n - 2, -1, -1):
        y2[..., k] = y2[..., k] * y2[..., k + 1] + u[..., k]
    return y2


def recoil_xsec(nu: np.ndarray) -> np.ndarray:
    """xSecRecoil (ph_mod.f90:369-385), the reference's expression as written."""
    thom = 6.65e-25
    xray_p = locate(nu, 20.6)
    out = np.full(nu.size, thom)
    for i in range(xray_p, nu.size):          # 0-based i >= xrayP  <=>  1-based i > xrayP
        dx = nu[i] / 3.7573e4
        out[i] = thom * 3. / 4. * ((1. + dx) / dx ** 3 * (2. * dx * (1. + dx) / (1. + 2. * dx) -
                                                           np.log(1. + 2. * dx)) +
                                   1. / 2. / dx * np.log(1. + 2. * dx) -
                                   (1. + 3. * dx) / (1. + 2. * dx) ** 3)
    return out


@dataclass
class BalanceTables:
    nu: np.ndarray
    xsec: XSecStack
    shells: list             # stack shells in species order
    speciesShell: np.ndarray # [S+1]
    shellNum: np.ndarray     # [nShells]
    W: np.ndarray            # [B, 3*nShells + 5]
    secIonP: int
    cRecoilHP: int
    cRecoilHeP: int
    xSecRecoil: np.ndarray
    outShell: np.ndarray     # [31*32]
    statW0: np.ndarray
    statW1: np.ndarray
    chex: np.ndarray         # rows (Z, stage, a, b, c, d)
    data: dict               # packed atomic arrays
    atomMeta8: np.ndarray
    atomQom: np.ndarray
    atomQ2: np.ndarray
    maxLev: int
    XHILimit: float


def build_balance(xsec: XSecStack, maxLev: int = 27, XHILimit: float = 0.05) -> BalanceTables:
    nu = xsec.nu
    B = nu.size
    d = dict(packed_data())
    # shells ordered by species; each species' shells contiguous
    spec_index = {s: i for i, s in enumerate(xsec.species)}
    order = sorted(range(len(xsec.shells)),
                   key=lambda k: (spec_index[(xsec.shells[k].elem, xsec.shells[k].ion)], k))
    shells = [xsec.shells[k] for k in order]
    starts = np.zeros(len(xsec.species) + 1, dtype=np.int32)
    for s in shells:
        starts[spec_index[(s.elem, s.ion)] + 1] += 1
    starts = np.cumsum(starts).astype(np.int32)
    secIonP = locate(nu, 7.353)
    cRecoilHP = locate(nu, 194.)
    cRecoilHeP = locate(nu, 260.)
    rec = recoil_xsec(nu)
    ns = len(shells)
    W = np.zeros((B, 3 * ns + 5))
    for k, s in enumerate(shells):
        lo, hi = s.lo, min(s.hi, B)
        sig = s.sigma
        if s.elem <= 2:
            for b in range(lo, B + 1):
                ph = sig[b - lo]
                if ph < 1.e-35:
                    ph = 0.0
                W[b - 1, 3 * k] = ph / (HCRYD * nu[b - 1])
                W[b - 1, 3 * k + 1] = ph * (nu[b - 1] - nu[lo - 1]) / nu[b - 1]
        else:
            y1 = d["auger"][s.elem, s.ion, s.shell, 1]
            augerE = HCRYD * nu[lo - 1] * y1
            for b in range(lo + 1, min(hi, secIonP - 1) + 1):
                ph = sig[b - lo]
                if ph < 1.e-35:
                    ph = 0.0
                W[b - 1, 3 * k] += ph / (HCRYD * nu[b - 1])
                W[b - 1, 3 * k + 1] += ph - augerE * (ph / (HCRYD * nu[b - 1]))
            ph0 = sig[0]
            W[lo - 1, 3 * k] += ph0 / (HCRYD * nu[lo - 1])
            W[lo - 1, 3 * k + 1] += ph0 * (HCRYD * nu[lo - 1] - augerE) / (HCRYD * nu[lo - 1])
            gH = hH = 0.0
            for b in range(max(lo + 1, secIonP), hi + 1):
                gH += sig[b - lo] / (HCRYD * nu[b - 1])
                hH += sig[b - lo]
            W[lo - 1, 3 * k] += gH
            W[lo - 1, 3 * k + 2] = hH - gH * augerE
    c0 = 3 * ns
    W[:, c0] = 1.0
    W[:, c0 + 1] = 1.0 / nu
    idx = np.arange(1, B + 1)
    W[:, c0 + 2] = np.where(idx > cRecoilHP, rec * (2.66e-5 * nu * nu - 1.), 0.0)
    W[:, c0 + 3] = np.where(idx > cRecoilHeP, rec * (2.66e-5 * nu * nu - 1.8), 0.0)
    W[:, c0 + 4] = 1.0
    outS = np.zeros(31 * 32, dtype=np.int32)
    g0 = np.ones(31 * 32, dtype=np.int32)
    g1 = np.ones(31 * 32, dtype=np.int32)
    for z in range(1, 31):
        for nel in range(1, z + 1):
            outS[z * 32 + nel] = outer_shell(z, nel)
            g0[z * 32 + nel], g1[z * 32 + nel] = stat_weights(z, nel)
    meta = d["atom_meta"]
    meta8 = np.zeros((meta.shape[0], 8), dtype=np.int64)
    meta8[:, :7] = meta
    qom = d["atom_qom"].astype(np.float64)
    q2 = np.zeros_like(qom)
    for a in range(meta.shape[0]):
        nl, nt, ot, _, _, oq, _ = (int(v) for v in meta[a])
        if nt >= 3:
            blk = qom[oq:oq + nt * nl * nl].reshape(nt, nl * nl).T    # [pairs, nT]
            y2 = natural_spline_y2(d["atom_logt"][ot:ot + nt], blk)
            q2[oq:oq + nt * nl * nl] = y2.T.reshape(-1)
    return BalanceTables(nu=nu, xsec=xsec, shells=shells, speciesShell=starts,
                         shellNum=np.array([s.shell for s in shells], dtype=np.int32), W=W,
                         secIonP=secIonP, cRecoilHP=cRecoilHP, cRecoilHeP=cRecoilHeP,
                         xSecRecoil=rec, outShell=outS, statW0=g0, statW1=g1,
                         chex=np.array(CHEX_ROWS), data=d, atomMeta8=meta8, atomQom=qom,
                         atomQ2=q2, maxLev=maxLev, XHILimit=XHILimit)

"""The text files of outputGas (output_mod.f90:13-1396) from the device
reductions: temperature.out, ionratio.out and lineFlux.out.

The device gives the volume averages (moc_volume_averages, :642-730), the
n-level line luminosities and Hbeta (moc_line_luminosities, :397-432); this
module lays them out as outputGas writes them (:810-1160).  One abundance
component: the "Total" block and component 1 hold the same numbers.

lineFlux.out carries Hbeta and the heavy-ion forbidden lines relative to
Hbeta with their wavelengths (1e8 / (E_upper - E_lower), E in cm^-1, as
equilibrium forms them, emission_mod.f90:2987-3005) and line indices.  The
hydrogenic and He I recombination-line tables of the reference's file are
not written: outputGas computes them from data/e1bx.d, which the reference
distribution does not ship (its hdatx stops there, :2581-2621).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import tables as T

N_ELEMENTS = 30


def _full(vol: np.ndarray, elems: list, nstages: int) -> np.ndarray:
    """(nElemUsed, nstages) -> (30, nstages) with the elements switched off at 0."""
    out = np.zeros((N_ELEMENTS, nstages))
    for e, z in enumerate(elems):
        out[int(z) - 1, :] = vol[e, :nstages]
    return out


def _rec(*vals) -> str:
    parts = []
    for v in vals:
        if isinstance(v, (int, np.integer)):
            parts.append(f"{int(v):12d}")
        elif isinstance(v, (float, np.floating)):
            parts.append(f"{float(v):16.8E}")
        else:
            parts.append(str(v))
    return " " + " ".join(parts) + "\n"


def write_temperature(outdir, TeVol: np.ndarray, elems: list, nstages: int,
                      nComponents: int = 1) -> Path:
    """temperature.out (:1085, :1115-1140): Te(ion) per element and stage."""
    T = _full(TeVol, elems, nstages)
    lines = ["\n", " Mean ionic temperatures Te(ion)\n", "\n",
             " Total, integrated over all nebular components\n"]
    for iAb in range(nComponents + 1):
        if iAb > 0:
            lines += ["\n", " Mean ionic temperatures Te(ion)\n", "\n", _rec("Component : ", iAb), "\n"]
        lines += [" Element      Ion        Te(Element,ion)\n", "\n"]
        for z in range(1, N_ELEMENTS + 1):
            for ion in range(1, nstages + 1):
                lines.append(_rec(z, ion, float(T[z - 1, ion - 1])))
    p = Path(outdir) / "temperature.out"
    p.write_text("".join(lines))
    return p


def write_ionratio(outdir, ionDenVol: np.ndarray, elems: list, nstages: int,
                   nComponents: int = 1) -> Path:
    """ionratio.out (:1099, :1120-1150): <ion> per element and stage."""
    X = _full(ionDenVol, elems, nstages)
    lines = ["\n", " Mean ionic ratio <ion>/<H+>\n", "\n",
             " Total, integrated over all nebular components\n", "\n"]
    for iAb in range(nComponents + 1):
        if iAb > 0:
            lines += ["\n", " Mean ionic ratio <ion>/<H+>\n", "\n", _rec("Component : ", iAb), "\n"]
        lines += [" Element      Ion        <ion>/<H+>I     <ion>/<H+>II\n", "\n"]
        for z in range(1, N_ELEMENTS + 1):
            for ion in range(1, nstages + 1):
                lines.append(_rec(z, ion, float(X[z - 1, ion - 1])))
    p = Path(outdir) / "ionratio.out"
    p.write_text("".join(lines))
    return p


def line_wavelengths(bt, z: int, stage: int, nLev: int) -> np.ndarray:
    """wav(i, j) = 1e8 / (E_j - E_i) [A] of ion (z, stage) from its level
    energies [cm^-1] (0 where E_j <= E_i), levels i < j up to nLev."""
    d = bt.data
    atom = int(d["atom_idx"][z, stage])
    out = np.zeros((nLev, nLev))
    if atom < 0:
        return out
    m = np.asarray(bt.atomMeta8)[atom]
    nf = min(int(m[0]), nLev)
    e = np.asarray(d["atom_e"])[int(m[3]):int(m[3]) + nf]
    for i in range(nf):
        for j in range(i + 1, nf):
            de = e[j] - e[i]
            out[i, j] = 1.e8 / de if de > 0. else 0.
    return out


def first_forbidden_index(nElemOn: int) -> int:
    """The line index of the first forbidden-line slot: the file numbers the
    84 hydrogenic lines (upper 2..15, lower 1..min(8, upper-1)) of every
    element switched on and the 34 He I lines first (:960-995)."""
    per = sum(min(8, up - 1) for up in range(2, 16))
    return per * nElemOn + 34 + 1


def lineflux_rows(lines: np.ndarray, ions: list, Hbeta: float, bt, nForLevels: int = 30,
                  nElemOn: int = 0):
    """The forbidden-line rows of lineFlux.out (:995-1025): (Z, stage, lower,
    upper, wavelength [A], I/I(Hbeta), index) for every line with a positive
    flux; the index continues after the recombination lines and counts every
    (ion, lower, upper) slot of nForLevels^2 levels per ion with data, in the
    reference's order (Z, stage ascending)."""
    nI, nL, _ = lines.shape
    rows = []
    iLine = first_forbidden_index(nElemOn)
    for k in range(nI):
        z, st = ions[k]
        wav = line_wavelengths(bt, z, st, nL)
        rel = lines[k] / Hbeta if Hbeta > 0. else np.zeros_like(lines[k])
        for a in range(nForLevels):
            for b in range(nForLevels):
                if a < nL and b < nL and rel[a, b] > 0.:
                    rows.append((z, st, a + 1, b + 1, float(wav[a, b]), float(rel[a, b]), iLine))
                iLine += 1
    return rows


def write_lineflux(outdir, lines: np.ndarray, ions: list, Hbeta: float, bt, symmetric: bool,
                   nElemOn: int, nComponents: int = 1, planeArea: float | None = None) -> Path:
    """lineFlux.out (:810-1060) with Hbeta [1e36 erg/s] (the sum of
    hydroLines(1,4,2) H dV times 1e-16, and 8 for an octant model, :795-805)
    and the heavy-ion forbidden lines relative to Hbeta.  A plane-parallel
    model (planeArea = getPlaneArea [1e36 cm^2]) adds Hbeta per unit area of
    the lit face [erg/cm^2/s] (:837-840)."""
    hb36 = Hbeta * 1.e-16 * (8. if symmetric else 1.)
    rows = lineflux_rows(lines, ions, Hbeta, bt, nElemOn=nElemOn)
    out = [" Totals integrated over all nebular components: \n"]
    for iAb in range(nComponents + 1):
        out.append("\n")
        if iAb > 0:
            out.append(_rec("Component: ", iAb))
        out += ["\n", " Line Ratios (Hbeta=1.) :               Analytical\n",
                _rec("Hbeta [E36 erg/s]:     ", float(hb36))]
        if planeArea is not None:
            out.append(_rec("Hbeta [erg/cm^2/s]:     ", float(hb36) / planeArea, float(planeArea)))
        out += ["\n",
                " Heavy ions forbidden lines\n", "\n",
                "  Z    Ion    lowlev  highlev   lambda[A] I/I(Hbeta)  index\n"]
        for z, st, lo, hi, wav, rel, idx in rows:
            out.append(f"{z:5d}  {st:5d}  {lo:5d}  {hi:5d}  {wav:12.2f}  {rel:12.2G}  {idx:5d}\n")
    p = Path(outdir) / "lineFlux.out"
    p.write_text("".join(out))
    return p


def write_output_gas(outdir, g, model) -> dict:
    """The three files from a MocGPU context g holding model (after the
    iterations); returns the paths."""
    Path(outdir).mkdir(parents=True, exist_ok=True)
    TeVol, XVol = g.volume_averages()
    lines, ions, hb = g.line_luminosities()
    gr = model.grid
    area = T.plane_area(gr.xAxis, gr.zAxis) if gr.lgPlaneIonization else None
    return {"temperature": write_temperature(outdir, TeVol, model.elems, model.nstages),
            "ionratio": write_ionratio(outdir, XVol, model.elems, model.nstages),
            "lineFlux": write_lineflux(outdir, lines, ions, hb, model.bt,
                                       bool(gr.lgSymmetricXYZ), len(model.elems),
                                       planeArea=area)}


def equivalent_tau(sed_noext, sed):
    """The equivalent optical depth per bin (output_mod.f90:3020-3035):
    ln(SEDnoExt / SED) where both are positive, else -1; SEDnoExt is the
    escaped SED of the first iteration of an lgEquivalentTau run (no dust
    opacity), SED that of the current one."""
    a = np.asarray(sed_noext, dtype=np.float64)
    b = np.asarray(sed, dtype=np.float64)
    ok = (a > 0.0) & (b > 0.0)
    out = np.full(a.shape, -1.0)
    out[ok] = np.log(a[ok] / b[ok])
    return out


def write_equivalent_tau(path, nu, sed_noext, sed) -> None:
    """equivalentTau.out: nu [Ryd], lambda [um], the equivalent tau and SEDnoExt
    per bin, as the reference writes it (output_mod.f90:3023-3035)."""
    c, fr1ryd = 2.9979250e10, 3.28984e15  # constants_mod.f90:12-13
    tau = equivalent_tau(sed_noext, sed)
    with open(path, "w") as f:
        for v, t, s0 in zip(np.asarray(nu, dtype=np.float64), tau, np.asarray(sed_noext)):
            f.write(f" {v:.8E} {c / (v * fr1ryd) * 1.e4:.8E} {t:.8E} {s0:.8E}\n")

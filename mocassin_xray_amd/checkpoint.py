"""Checkpoint files in the reference's layout and the warm start.

writeGrid (grid_mod.f90:2646-2883) leaves a run's state in output/:

  grid0.out        per grid: nGrids; nx ny nz nCells motherP R_out; the x, y
                   and z axes one value a record; then, in (i, j, k) order with
                   k fastest, one record "i j k active lgConverged lgBlack" for
                   every cell whose active value is not 0
  grid1.out        per active cell: Te Ne Hden
  grid2.out        per active cell and element switched on (Z order): the
                   ion fractions of stages 1 .. min(Z+1, nstages)
  dustGrid.out     per active cell: Ndust, then nSizes+1 records of Tdust of
                   species 0..nSpecies; a blank record and the total dust mass
  photoSource.out  nStars; per source: 'contShape' T_eff L_* nPackets x/xmax
                   y/ymax z/zmax spID tStep; a closing comment record
  grid3.out        the run parameters, one group a record

resetGrid (:3051-3690) reads them back with list-directed reads, and
MoCaSSiNwarm (mocassinWarm.f90:8-102) restarts the iteration from that state
(its first iteration recomputes every cell's opacity and emission).

The writer puts one read statement's values in one record and writes reals
with 17 significant digits, so a state written here is read back exactly; the
reader follows list-directed rules (each read starts a new record, continues
over records until its values are read and drops the rest of its last
record; r*value repeats; quoted or bare strings; T/F logicals), so it also
reads files that the reference wrote, whose single-precision values it
widens to FP64.
"""
from __future__ import annotations

import dataclasses
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .mocgpu import CartesianGrid

FILES = ("grid0.out", "grid1.out", "grid2.out", "grid3.out", "photoSource.out", "dustGrid.out")


# ------------------------------------------------------------------ records
def _fmt(v) -> str:
    if isinstance(v, (bool, np.bool_)):
        return "T" if v else "F"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        x = float(v)
        if not np.isfinite(x):
            raise ValueError(f"checkpoint: non-finite value {x}")
        return format(x, ".17g")
    if isinstance(v, str):
        return v
    raise TypeError(type(v))


def _rec(*vals) -> str:
    return " " + " ".join(_fmt(v) for v in vals) + "\n"


def _q(s: str) -> str:
    """A quoted character value (the reference writes '"', name, '"')."""
    if '"' in s:
        raise ValueError(f"checkpoint: string with a double quote: {s!r}")
    return f'"{s}"'


_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^\s,/]+))\s*,?""")


class ListReader:
    """List-directed reads (Fortran READ(u, *)) over the records of a file."""

    def __init__(self, path):
        self.path = Path(path)
        self.lines = self.path.read_text().splitlines()
        self.pos = 0

    def _tokens(self, line: str) -> list:
        out = []
        i = 0
        while i < len(line):
            if line[i].isspace() or line[i] == ",":
                i += 1
                continue
            if line[i] == "/":          # a slash ends the read
                out.append(None)
                break
            m = _TOKEN.match(line, i)
            if not m or m.end() == i:
                raise ValueError(f"{self.path}: cannot read {line[i:]!r}")
            if m.group(1) is not None:
                out.append(("s", m.group(1).replace("''", "'")))
            elif m.group(2) is not None:
                out.append(("s", m.group(2).replace('""', '"')))
            else:
                t = m.group(3)
                rep = re.fullmatch(r"(\d+)\*(.*)", t)
                if rep:
                    out.extend([("v", rep.group(2))] * int(rep.group(1)))
                else:
                    out.append(("v", t))
            i = m.end()
        return out

    def read(self, *kinds: str) -> list:
        """One READ(u, *) of len(kinds) items; kinds: 'i', 'f', 'l', 's'."""
        vals, toks = [], []
        while len(vals) < len(kinds):
            if not toks:
                if self.pos >= len(self.lines):
                    raise EOFError(f"{self.path}: end of file after {len(vals)} of "
                                   f"{len(kinds)} values")
                toks = self._tokens(self.lines[self.pos])
                self.pos += 1
                continue
            t = toks.pop(0)
            if t is None:
                raise ValueError(f"{self.path}: slash before {len(kinds)} values")
            vals.append(self._convert(t, kinds[len(vals)]))
        return vals

    def _convert(self, t, kind):
        tag, s = t
        if kind == "s":
            return s
        if tag == "s":
            raise ValueError(f"{self.path}: a string where a number is read: {s!r}")
        if kind == "i":
            return int(s)
        if kind == "f":
            return float(s.replace("d", "e").replace("D", "e"))
        if kind == "l":
            u = s.upper().lstrip(".")
            if u.startswith("T"):
                return True
            if u.startswith("F"):
                return False
            raise ValueError(f"{self.path}: not a logical: {s!r}")
        raise ValueError(kind)

    def skip(self):
        """An empty READ: moves to the next record."""
        self.pos += 1


# --------------------------------------------------------------- parameters
@dataclass
class RunParams:
    """The records of grid3.out; defaults are readInput's (set_input_mod.f90:
    28-120)."""
    nbins: int = 600
    nstages: int = 17
    lgSymmetricXYZ: bool = False
    convWriteGrid: float = 0.0
    lgAutoPackets: bool = False
    convIncPercent: float = 0.0
    nPhotIncrease: float = 0.0
    maxPhotons: int = 0
    lgTalk: bool = False
    lg1D: bool = False
    nuStepSize: float = 0.075
    nuMax: float = 15.0
    nuMin: float = 1.001e-5
    R_in: float = -1.0
    XHIlimit: float = 0.05
    maxIterateMC: int = 30
    minConvergence: float = 95.0
    lgDebug: bool = False
    lgPlaneIonization: bool = False
    meanFieldin: float = 0.0
    nu0: float = 0.0
    nu0Add: float = 0.0
    abundanceFiles: list = field(default_factory=lambda: ["none"])
    lgOutput: bool = False
    dxSlit: float = 0.0
    dySlit: float = 0.0
    lgDust: bool = False
    lgDustConstant: bool = False
    dustFile: tuple = ("none", "none")
    lgGas: bool = True
    lgRecombination: bool = False
    nSpecies: int = 0
    nSizes: int = 0
    resLinesTransfer: float = 101.0
    lgDustScattering: bool = True
    viewPointTheta: list = field(default_factory=list)
    viewPointPhi: list = field(default_factory=list)
    contCube: tuple = (-1.0, -1.0)
    lgPhotoelectric: bool = True
    lgTraceHeating: bool = False
    Ldiffuse: float = 0.0
    Tdiffuse: float = 0.0
    shapeDiffuse: str = "none"
    nPhotonsDiffuse: int = 0
    emittingGrid: int = 0
    lgMultistars: bool = False
    lgCompton: bool = True
    pwlMin: float = 0.0
    pwlMax: float = 0.0
    lgNahar: bool = False
    lgBadnell: bool = True
    lgSeparateSED: bool = False
    lg2D: bool = False


@dataclass
class PhotoSource:
    """One record of photoSource.out; position in cm."""
    contShape: str
    TStellar: float
    LStar: float          # 1e36 erg/s
    nPhotons: int
    position: tuple       # (x, y, z) cm
    spID: str = "none"
    tStep: float = 0.0
    pwlIndex: float | None = None


@dataclass
class Checkpoint:
    grid: CartesianGrid
    state: dict           # Te, Ne, Hden (C+1,), ionDen (C+1, NE, NS), lgConverged, lgBlack
    params: RunParams
    sources: list
    dust: dict | None = None   # Ndust (C+1,), Tdust (C+1, nSizes+1, nSpecies+1)
    motherP: int = 0


# ------------------------------------------------------------------- writer
def write_grid(outdir, ck: Checkpoint, elems: list, totalDustMass: float = 0.0) -> None:
    """writeGrid (grid_mod.f90:2646-2883) for one grid: grid0-3.out,
    photoSource.out and, with grains, dustGrid.out."""
    out = Path(outdir)
    out.mkdir(parents=True, exist_ok=True)
    g, s, p = ck.grid, ck.state, ck.params
    nx, ny, nz = g.active.shape
    ns = p.nstages
    lines0 = [_rec(1), _rec(nx, ny, nz, int(g.nCells), int(ck.motherP), float(g.R_out))]
    lines0 += [_rec(float(v)) for v in g.xAxis]
    lines0 += [_rec(float(v)) for v in g.yAxis]
    lines0 += [_rec(float(v)) for v in g.zAxis]
    lines1, lines2, linesD = [], [], []
    conv = np.asarray(s.get("lgConverged", np.zeros(g.nCells + 1)), dtype=np.int64)
    black = np.asarray(s.get("lgBlack", np.zeros(g.nCells + 1)), dtype=np.int64)
    order = np.argsort(np.asarray(elems))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                a = int(g.active[i, j, k])
                if a == 0:
                    continue
                c = max(a, 0)
                lines0.append(_rec(i + 1, j + 1, k + 1, a, int(conv[c]), int(black[c])))
                if c == 0:
                    continue
                if p.lgGas:
                    lines1.append(_rec(float(s["Te"][c]), float(s["Ne"][c]), float(s["Hden"][c])))
                    for e in order:
                        z = int(elems[e])
                        lines2.append(_rec(*[float(v) for v in s["ionDen"][c, e, :min(z + 1, ns)]]))
                if p.lgDust:
                    d = ck.dust
                    linesD.append(_rec(float(d["Ndust"][c])))
                    for ai in range(p.nSizes + 1):
                        linesD.append(_rec(*[float(v) for v in d["Tdust"][c, ai, :p.nSpecies + 1]]))
    (out / "grid0.out").write_text("".join(lines0))
    if p.lgGas:
        (out / "grid1.out").write_text("".join(lines1))
        (out / "grid2.out").write_text("".join(lines2))
    if p.lgDust:
        linesD += [" \n", f" Total dust mass [1.e45 g]: {_fmt(float(totalDustMass))}\n"]
        (out / "dustGrid.out").write_text("".join(linesD))
    ps = [_rec(len(ck.sources), "number of photon sources")]
    xm, ym, zm = float(g.xAxis[-1]), float(g.yAxis[-1]), float(g.zAxis[-1])
    for src in ck.sources:
        x, y, z = (float(v) for v in src.position)
        if not src.spID or any(ch.isspace() for ch in src.spID):
            raise ValueError("checkpoint: spID must be one non-blank word")
        ps.append(_rec("'" + src.contShape + "'", float(src.TStellar), float(src.LStar),
                       int(src.nPhotons), x / xm, y / ym, z / zm, src.spID, float(src.tStep)))
        if src.contShape == "powerlaw":
            ps.append(_rec(float(src.pwlIndex or 0.0)))
    ps.append(" (contShape, T_eff[K], L_* [E36 erg/s], nPackets, (x,y,z) position, spID, tstep)\n")
    (out / "photoSource.out").write_text("".join(ps))
    r = [_rec(1), _rec(float(p.convWriteGrid), "convWriteGrid"),
         _rec(p.lgAutoPackets, float(p.convIncPercent), float(p.nPhotIncrease), int(p.maxPhotons),
              "lgAutoPackets"),
         _rec(p.lgSymmetricXYZ, "lgSymmetricXYZ"), _rec(p.lgTalk, "lgTalk"), _rec(p.lg1D, "lg1D"),
         _rec(int(p.nbins), "nbins"), _rec(float(p.nuStepSize), "nuStepSize"),
         _rec(float(p.nuMax), "nuMax"), _rec(float(p.nuMin), "nuMin"), _rec(float(p.R_in), "R_in"),
         _rec(float(p.XHIlimit), "XHIlimit"),
         _rec(int(p.maxIterateMC), float(p.minConvergence), "maxIterateMC"),
         _rec(p.lgDebug, "lgDebug"),
         _rec(p.lgPlaneIonization, float(p.meanFieldin), float(p.nu0), float(p.nu0Add),
              "lgPlaneIonization"),
         _rec(len(p.abundanceFiles), "nAbComponents")]
    r += [_rec(_q(f)) for f in p.abundanceFiles]
    r += [_rec(p.lgOutput, "lgOutput"), _rec(float(p.dxSlit), float(p.dySlit), "dxSlit,dySlit"),
          _rec(p.lgDust, p.lgDustConstant, "lgDust, lgDustConstant"),
          _rec(_q(p.dustFile[0]), "dustFile"), _rec(_q(p.dustFile[1]), "dustFile"),
          _rec(p.lgGas, "lgGas"), _rec(p.lgRecombination, "lgRecombination"),
          _rec(int(p.nSpecies), int(p.nSizes), "nSpecies, nSizes"),
          _rec(float(p.resLinesTransfer), "resLinesTransfer"),
          _rec(p.lgDustScattering, "lgDustScattering"),
          _rec(len(p.viewPointTheta), "nAngleBins")]
    if p.viewPointTheta:
        r += [_rec(*[float(v) for v in p.viewPointTheta]), _rec(*[float(v) for v in p.viewPointPhi])]
    r += [_rec(float(p.contCube[0]), float(p.contCube[1]), "continuumCube"),
          _rec(p.lgPhotoelectric, "lgPhotoelectric"), _rec(p.lgTraceHeating, "lgTraceHeating"),
          _rec(float(p.Ldiffuse), "Ldiffuse"), _rec(float(p.Tdiffuse), "Tdiffuse"),
          _rec(_q(p.shapeDiffuse), "shapeDiffuse"), _rec(int(p.nPhotonsDiffuse), "nPhotonsDiffuse"),
          _rec(int(p.emittingGrid), "emittingGrid"), _rec(int(p.nstages), "nstages"),
          _rec(p.lgMultistars, "lgMultiStars"), _rec(p.lgCompton, "compton?"),
          _rec(float(p.pwlMin), float(p.pwlMax), "pwlMin, pwlMax"), _rec(p.lgNahar, "nahar?"),
          _rec(p.lgBadnell, "badnell?"), _rec(p.lgSeparateSED, "lgSeparateSED"),
          _rec(p.lg2D, "2D geometry?")]
    (out / "grid3.out").write_text("".join(r))


# ------------------------------------------------------------------- reader
def read_params(path) -> RunParams:
    """grid3.out as resetGrid reads it (grid_mod.f90:3122-3199)."""
    R = ListReader(path)
    p = RunParams()
    nG, = R.read("i")
    if nG != 1:
        raise ValueError(f"{path}: {nG} grids (nested sub-grids are not supported)")
    p.convWriteGrid, = R.read("f")
    p.lgAutoPackets, p.convIncPercent, p.nPhotIncrease, p.maxPhotons = R.read("l", "f", "f", "i")
    p.lgSymmetricXYZ, = R.read("l")
    p.lgTalk, = R.read("l")
    p.lg1D, = R.read("l")
    p.nbins, = R.read("i")
    p.nuStepSize, = R.read("f")
    p.nuMax, = R.read("f")
    p.nuMin, = R.read("f")
    p.R_in, = R.read("f")
    p.XHIlimit, = R.read("f")
    p.maxIterateMC, p.minConvergence = R.read("i", "f")
    p.lgDebug, = R.read("l")
    p.lgPlaneIonization, p.meanFieldin, p.nu0, p.nu0Add = R.read("l", "f", "f", "f")
    nAb, = R.read("i")
    p.abundanceFiles = [R.read("s")[0].strip() for _ in range(nAb)]
    p.lgOutput, = R.read("l")
    p.dxSlit, p.dySlit = R.read("f", "f")
    p.lgDust, p.lgDustConstant = R.read("l", "l")
    p.dustFile = (R.read("s")[0].strip(), R.read("s")[0].strip())
    p.lgGas, = R.read("l")
    p.lgRecombination, = R.read("l")
    p.nSpecies, p.nSizes = R.read("i", "i")
    p.resLinesTransfer, = R.read("f")
    p.lgDustScattering, = R.read("l")
    nA, = R.read("i")
    if nA > 0:
        p.viewPointTheta = R.read(*("f",) * nA)
        p.viewPointPhi = R.read(*("f",) * nA)
    p.contCube = tuple(R.read("f", "f"))
    p.lgPhotoelectric, = R.read("l")
    p.lgTraceHeating, = R.read("l")
    p.Ldiffuse, = R.read("f")
    p.Tdiffuse, = R.read("f")
    p.shapeDiffuse = R.read("s")[0].strip()
    p.nPhotonsDiffuse, = R.read("i")
    p.emittingGrid, = R.read("i")
    p.nstages, = R.read("i")
    p.lgMultistars, = R.read("l")
    p.lgCompton, = R.read("l")
    p.pwlMin, p.pwlMax = R.read("f", "f")
    p.lgNahar, = R.read("l")
    p.lgBadnell, = R.read("l")
    p.lgSeparateSED, = R.read("l")
    p.lg2D, = R.read("l")
    return p


def read_sources(path, axes_max) -> list:
    """photoSource.out (grid_mod.f90:3087-3110); positions are stored as
    fractions of the outer axis values and scaled back as setStarPosition
    does (:3706-3710)."""
    R = ListReader(path)
    n, = R.read("i")
    out = []
    for _ in range(n):
        shape, T, L, nP, x, y, z, spID, tStep = R.read("s", "f", "f", "i", "f", "f", "f", "s", "f")
        shape = shape.strip()
        pwl = R.read("f")[0] if shape == "powerlaw" else None
        out.append(PhotoSource(shape, T, L, nP, (x * axes_max[0], y * axes_max[1], z * axes_max[2]),
                               spID.strip(), tStep, pwl))
    return out


def read_grid(outdir, elems: list) -> Checkpoint:
    """resetGrid (grid_mod.f90:3051-3690) for one grid.  elems: the atomic
    numbers switched on (the reference takes them from its abundance file)."""
    d = Path(outdir)
    p = read_params(d / "grid3.out")
    R0 = ListReader(d / "grid0.out")
    R0.read("i")
    nx, ny, nz, nCells, motherP, R_out = R0.read("i", "i", "i", "i", "i", "f")
    ax = np.array([R0.read("f")[0] for _ in range(nx)])
    ay = np.array([R0.read("f")[0] for _ in range(ny)])
    az = np.array([R0.read("f")[0] for _ in range(nz)])
    ns = p.nstages
    C1 = nCells + 1
    order = np.argsort(np.asarray(elems))
    act = np.zeros((nx, ny, nz), dtype=np.int32)
    st = {"Te": np.zeros(C1), "Ne": np.zeros(C1), "Hden": np.zeros(C1),
          "ionDen": np.zeros((C1, len(elems), ns)),
          "lgConverged": np.zeros(C1, dtype=np.int32), "lgBlack": np.zeros(C1, dtype=np.int32)}
    R1 = ListReader(d / "grid1.out") if p.lgGas else None
    R2 = ListReader(d / "grid2.out") if p.lgGas else None
    RD = ListReader(d / "dustGrid.out") if p.lgDust else None
    dust = None
    if p.lgDust:
        dust = {"Ndust": np.zeros(C1), "Tdust": np.zeros((C1, p.nSizes + 1, p.nSpecies + 1))}
    for _ in range(nCells):
        i, j, k, a, conv, black = R0.read("i", "i", "i", "i", "i", "i")
        act[i - 1, j - 1, k - 1] = a
        c = max(a, 0)
        st["lgConverged"][c] = conv
        st["lgBlack"][c] = black
        if c == 0:
            continue
        if p.lgGas:
            st["Te"][c], st["Ne"][c], st["Hden"][c] = R1.read("f", "f", "f")
            for e in order:
                z = int(elems[e])
                n = min(z + 1, ns)
                st["ionDen"][c, e, :n] = R2.read(*("f",) * n)
        if p.lgDust:
            dust["Ndust"][c], = RD.read("f")
            for ai in range(p.nSizes + 1):
                dust["Tdust"][c, ai, :] = RD.read(*("f",) * (p.nSpecies + 1))
    grid = CartesianGrid(ax, ay, az, act, nCells, bool(p.lgSymmetricXYZ), float(R_out),
                         lgPlaneIonization=bool(p.lgPlaneIonization))
    sources = read_sources(d / "photoSource.out", (ax[-1], ay[-1], az[-1]))
    return Checkpoint(grid, st, p, sources, dust, motherP)


# --------------------------------------------------------------- warm start
def checkpoint_of(model, g, params: RunParams | None = None, TStellar=None,
                  contShape: str = "blackbody") -> Checkpoint:
    """The state of a MocGPU context g holding model, as writeGrid sees it
    (TStellar: per source, for photoSource.out)."""
    p = dataclasses.replace(params) if params is not None else RunParams()
    p.nbins = model.tables.nbins
    p.nstages = model.nstages
    p.lgSymmetricXYZ = bool(model.grid.lgSymmetricXYZ)
    p.lgPlaneIonization = bool(model.grid.lgPlaneIonization)
    p.lgCompton = bool(model.tables.lgCompton)
    dust = None
    st = {}
    gas = hasattr(model, "bt")
    p.lgGas = gas
    if gas:
        st = {k: g.download(k) for k in ("Te", "Ne", "Hden", "ionDen")}
    st["lgConverged"] = g.download_flags("lgConverged")
    st["lgBlack"] = g.download_flags("lgBlack")
    if getattr(model, "dust", None) is not None:
        nS, nA = g.dust_shape
        p.lgDust, p.nSpecies, p.nSizes = True, nS, nA
        dust = {"Ndust": g.download("Ndust"), "Tdust": g.download("Tdust")}
    t = model.tables
    pos = np.asarray(t.starPosition).reshape(-1, 3)
    T = np.zeros(pos.shape[0]) if TStellar is None else np.broadcast_to(
        np.asarray(TStellar, dtype=np.float64), (pos.shape[0],))
    sources = [PhotoSource(contShape, float(T[s]), float(np.atleast_1d(t.Lstar)[s]),
                           int(np.atleast_1d(t.nPhotons)[s]), tuple(float(v) for v in pos[s]))
               for s in range(pos.shape[0])]
    return Checkpoint(model.grid, st, p, sources, dust)


def warm_start(model, ck: Checkpoint, g) -> None:
    """Load a checkpoint into a MocGPU context that holds model's tables
    (MoCaSSiNwarm: resetGrid, then the iteration).  The checkpoint's grid must
    be the model's grid; its state replaces the model's starting state."""
    mg = model.grid
    if (ck.grid.active.shape != mg.active.shape or not np.array_equal(ck.grid.active, mg.active)
            or not np.array_equal(ck.grid.xAxis, mg.xAxis)
            or not np.array_equal(ck.grid.yAxis, mg.yAxis)
            or not np.array_equal(ck.grid.zAxis, mg.zAxis)):
        raise ValueError("warm_start: the checkpoint's grid is not the model's grid")
    if ck.params.lgGas:
        for k in ("Te", "Ne", "Hden", "ionDen"):
            g.upload(k, ck.state[k])
    if ck.dust is not None:
        g.upload("Ndust", ck.dust["Ndust"])
        g.upload("Tdust", ck.dust["Tdust"])

"""Python host mirror of the reference's transport interface over libmocgpu.

`MocGPU` owns one libmocgpu context (one Cartesian grid).  Its methods carry
the names of the MOCASSIN routines they replace:

    energyPacketDriver(iStar, n, ...)  photon_mod.f90:28
    reset_tallies()                     iteration_mod.f90:626-717 (Jste=0, escapedPackets=0)
    normalize_j()                       iteration_mod.f90:1002-1031

Per-cell arrays are exchanged in MOCASSIN's shapes: (nCells+1,) for cell
fields and (nCells+1, nbins) for (cell, bin) fields, row 0 being the inactive
dummy cell.  Errors raise MocError with the library's message (the reference
prints and stops).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import _dp, _ip
from ._lib import (CELLBIN_FIELDS, FIELDS, GRAIN_FIELDS, MOC_DT_F64, MOC_DT_I32,
                   MOC_LAYOUT_FORTRAN, MOC_PREC_FP64, MOC_PREC_MIXED, MocAtomic, MocBalanceTables, MocDustTables,
                   MocEmissionTables, MocGridDesc, MocIterationStats, MocTables, MocTransportStats,
                   MocTuning, MocUpdateStats)


class MocError(RuntimeError):
    pass


@dataclass
class SourceTables:
    """Mesh and point-source tables (moc_tables)."""
    nuArray: np.ndarray
    widFlx: np.ndarray
    ionEdge1: float
    KNsigmaT: np.ndarray
    KNcdf: np.ndarray          # [B, nAngles]
    Pcomp: np.ndarray          # [B, nAngles]
    sourceCDF: np.ndarray      # [nStars, B]
    Lstar: np.ndarray          # [nStars] 1e36 erg/s
    nPhotons: np.ndarray       # [nStars]
    starPosition: np.ndarray   # [nStars, 3] cm
    starIndex: np.ndarray      # [nStars, 3] 1-based
    lgCompton: bool = True

    @property
    def nbins(self) -> int:
        return int(self.nuArray.size)


@dataclass
class CartesianGrid:
    """One grid_type (common_mod.f90:308-371) restricted to what the path needs."""
    xAxis: np.ndarray
    yAxis: np.ndarray
    zAxis: np.ndarray
    active: np.ndarray         # [nx, ny, nz] int32, 1..nCells or 0
    nCells: int
    lgSymmetricXYZ: bool
    R_out: float = 0.0
    lgPlaneIonization: bool = False
    extra: dict = field(default_factory=dict)


@dataclass
class BatchGrid:
    """A batch of grids attached together: grid k's cell c is global cell
    cellBase[k] + c."""
    grids: list

    @property
    def nCells(self) -> int:
        return int(sum(g.nCells for g in self.grids))

    @property
    def cellBase(self) -> list:
        out, b = [], 0
        for g in self.grids:
            out.append(b)
            b += g.nCells
        return out


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------------
# The contents of every C struct of the ABI as a dict of field name ->
# numpy array (pointer fields) or Python scalar, in the struct's own order.
# MocGPU builds the ctypes structs from these; fortran_io writes the same
# dicts for the Fortran driver.

def tables_arrays(tables: SourceTables, precision: str) -> dict:
    B = tables.nbins
    kn = _c(tables.KNcdf)
    return {
        "nbins": B, "precision": {"fp64": MOC_PREC_FP64, "mixed": MOC_PREC_MIXED}[precision],
        "nuArray": _c(tables.nuArray), "widFlx": _c(tables.widFlx),
        "ionEdge1": float(tables.ionEdge1), "lgCompton": int(bool(tables.lgCompton)),
        "nAngles": kn.size // B, "KNsigmaT": _c(tables.KNsigmaT), "KNsigmaCDF": kn,
        "Pcomp": _c(tables.Pcomp), "nStars": int(np.atleast_1d(tables.Lstar).size),
        "sourceCDF": _c(tables.sourceCDF), "Lstar": _c(tables.Lstar),
        "nPhotons": _c(tables.nPhotons), "starPosition": _c(tables.starPosition),
        "starIndex": _c(tables.starIndex, np.int32),
    }


def grid_arrays(g: CartesianGrid) -> dict:
    act = np.ascontiguousarray(np.asarray(g.active, dtype=np.int32).reshape(-1, order="F"))
    xa, ya, za = _c(g.xAxis), _c(g.yAxis), _c(g.zAxis)
    return {"nx": xa.size, "ny": ya.size, "nz": za.size, "xAxis": xa, "yAxis": ya, "zAxis": za,
            "active": act, "nCells": int(g.nCells), "lgSymmetricXYZ": int(bool(g.lgSymmetricXYZ)),
            "R_out": float(g.R_out), "lgPlaneIonization": int(bool(g.lgPlaneIonization))}


def atomic_arrays(elem_z, abund, nstages: int, xsec) -> dict:
    z = _c(elem_z, np.int32)
    zi = {int(v): i for i, v in enumerate(z)}
    spec_e = _c([zi[e] for e, _ in xsec.species], np.int32)
    spec_i = _c([i for _, i in xsec.species], np.int32)
    return {"nElemUsed": z.size, "elemZ": z, "abund": _c(abund), "nstages": int(nstages),
            "nSpecies": spec_e.size, "speciesElem": spec_e, "speciesIon": spec_i,
            "sigma": _c(xsec.sigma), "brems": _c(xsec.brems)}


def balance_arrays(bt) -> dict:
    d = bt.data
    diel, aldro, chex = _c(d["diel"]), _c(d["aldro"]), bt.chex
    am = _c(bt.atomMeta8, np.int64)
    ag, aA, aq = _c(d["atom_g"]), _c(d["atom_A"]), _c(bt.atomQom)
    lt = _c(d["atom_logt"])
    return {
        "maxLev": int(bt.maxLev), "XHILimit": float(bt.XHILimit),
        "nShells": int(np.asarray(bt.shellNum).size), "shellNum": _c(bt.shellNum, np.int32),
        "speciesShell": _c(bt.speciesShell, np.int32), "W": _c(bt.W),
        "outShell": _c(bt.outShell, np.int32), "statW0": _c(bt.statW0, np.int32),
        "statW1": _c(bt.statW1, np.int32), "rrb": _c(d["rrb"]), "drb_c": _c(d["drb_c"]),
        "drb_e": _c(d["drb_e"]), "rrec": _c(d["rrec"]), "rnew": _c(d["rnew"]),
        "nDiel": diel.shape[0], "diel": diel, "nAldro": aldro.shape[0], "aldro": aldro,
        "cionCF": _c(d["cion_cf"]), "cionTherm": _c(d["cion_th"]),
        "nAuger": _c(d["nauger"], np.int32), "auger": _c(d["auger"]),
        "nChex": int(chex.shape[0]), "chexZ": _c(chex[:, 0], np.int32),
        "chexIon": _c(chex[:, 1], np.int32), "chexFit": _c(chex[:, 2:6]),
        "nAtoms": am.shape[0], "atomIdx": _c(d["atom_idx"], np.int32), "atomMeta": am,
        "atomAtp": _c(d["atom_atp"]), "nLogT": lt.size, "nLev": ag.size, "nA": aA.size,
        "nQ": aq.size, "atomLogT": lt, "atomG": ag, "atomE": _c(d["atom_e"]), "atomA": aA,
        "atomQom": aq, "atomQ2": _c(bt.atomQ2), "atomRecKind": _c(d["atom_rec_kind"], np.int32),
        "atomRec": _c(d["atom_rec"]), "atomRecValid": _c(d["atom_rec_valid"], np.int32),
    }


def emission_arrays(et) -> dict:
    fbXs = _c(et.fbXs)
    return {
        "ntk": et.tk.size, "tkGamma": _c(et.tk), "nlim": [int(e.size) for e in et.edge],
        "edgeP": [_c(e, np.int32) for e in et.edge], "logGamma": [_c(v) for v in et.loggam],
        "HIPnuP": int(et.HIPnuP), "HeIPnuP": int(et.HeIPnuP), "HeIIPnuP": int(et.HeIIPnuP),
        "xsHI": _c(et.xsHI), "xsHeI": None if et.xsHeI is None else _c(et.xsHeI),
        "xsHeII": None if et.xsHeII is None else _c(et.xsHeII), "nFb": int(et.fbZ.size),
        "fbZ": _c(et.fbZ, np.int32), "fbIon": _c(et.fbIon, np.int32),
        "fbLo": _c(et.fbLo, np.int32), "fbHi": _c(et.fbHi, np.int32),
        "fbOff": _c(et.fbOff, np.int64), "nFbXs": fbXs.size, "fbXs": fbXs,
        "fbGratio": _c(et.fbGratio), "he2Row": _c(et.he2j, np.int32),
        "hei2phot": _c(et.hei2phot), "nHlFiles": int(et.hlDens.shape[0]),
        "hlMap": _c(et.hlMap, np.int32), "hlDens": _c(et.hlDens), "hlLines": _c(et.hlLines),
        "hbA": _c(et.hbA), "hbB": _c(et.hbB), "rbEdge": _c(et.rbEdge), "r2aEdge": _c(et.r2aEdge),
        "heiRec": _c(et.heiRec), "heiiLyman": _c(et.heiiLyman), "j2TsP": int(et.j2TsP),
        "lymP": [int(v) for v in et.lymP],
    }


def dust_arrays(dt, lgGas: bool, XHILimit: float, xsec=None, qheat: dict | None = None,
                equivalent_tau: bool = False) -> dict:
    """xsec (the gas cross-section stack) switches on the grain photoelectric
    and gas-grain collision terms (lgPhotoelectric); qheat (keys minaQHeat,
    minConvQHeat, Tmax, nTbins; the reference's defaults 1e-3 um, 99 %,
    700 K, 300 where left out) switches on quantum heating."""
    from .dust import ip_ryd
    pe = xsec is not None
    q = None if qheat is None else {"minaQHeat": 1.e-3, "minConvQHeat": 99.0, "Tmax": 700.0,
                                    "nTbins": 300, **qheat}
    kind = getattr(dt, "kind", None)
    return {
        "lgQHeat": 1 if q else 0, "nTbins": int(q["nTbins"]) if q else 0,
        "minaQHeat": float(q["minaQHeat"]) if q else 0.0,
        "minConvQHeat": float(q["minConvQHeat"]) if q else 0.0, "Tmax": float(q["Tmax"]) if q else 0.0,
        "grainKind": _c(kind, np.int32) if q else None,
        "lgEquivalentTau": 1 if equivalent_tau else 0,
        "nSpecies": dt.nSpecies, "nSizes": dt.nSizes, "nTemps": int(dt.emIntegral.shape[2]),
        "lgGas": 1 if lgGas else 0, "lgPhotoelectric": 1 if pe else 0,
        "grainAbun": _c(dt.grainAbun), "grainWeight": _c(dt.grainWeight),
        "TdustSublime": _c(dt.Tsub), "Cabs": _c(dt.Cabs), "Csca": _c(dt.Csca),
        "emIntegral": _c(dt.emIntegral), "XHILimit": float(XHILimit),
        "grainRadius": _c(dt.radius) if (pe or q) else None, "grainVn": _c(dt.Vn) if pe else None,
        "MsurfAtom": _c(dt.Msurf) if pe else None, "ipRyd": _c(ip_ryd(xsec)) if pe else None,
    }


def to_struct(cls, d: dict):
    """The ctypes struct `cls` filled from a field dict (arrays stay owned by d)."""
    kw = {}
    for name, ct in cls._fields_:
        v = d[name]
        if isinstance(ct, type) and issubclass(ct, C._Pointer):
            kw[name] = None if v is None else v.ctypes.data_as(ct)
        elif isinstance(ct, type) and issubclass(ct, C.Array):
            el = ct._type_
            if issubclass(el, C._Pointer):
                kw[name] = ct(*[None if a is None else a.ctypes.data_as(el) for a in v])
            else:
                kw[name] = ct(*v)
        else:
            kw[name] = v
    return cls(**kw)


class MocGPU:
    def __init__(self, tables: SourceTables, precision: str = "fp64"):
        self._L = _lib.lib()
        # own copy: set_packets changes nPhotons
        self.tables = dataclasses.replace(tables, nPhotons=np.array(tables.nPhotons, dtype=np.float64))
        self.precision = precision
        self._keep = tables_arrays(self.tables, precision)
        t = to_struct(MocTables, self._keep)
        self._ctx = C.c_void_p()
        rc = self._L.moc_create(C.byref(self._ctx), C.byref(t))
        if rc != 0:
            msg = self._L.moc_last_error(self._ctx).decode()
            self._L.moc_destroy(self._ctx)
            self._ctx = None
            raise MocError(f"moc_create: {msg} (rc={rc})")
        self.grid: CartesianGrid | None = None

    # ------------------------------------------------------------------ utils
    def _check(self, rc: int, what: str):
        if rc != 0:
            raise MocError(f"{what}: {self._L.moc_last_error(self._ctx).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.moc_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nbins(self) -> int:
        return self.tables.nbins

    # ------------------------------------------------------------------- grid
    def grid_attach(self, g: CartesianGrid):
        self._grid_keep = grid_arrays(g)
        d = to_struct(MocGridDesc, self._grid_keep)
        self._check(self._L.moc_grid_attach(self._ctx, 1, C.byref(d)), "moc_grid_attach")
        self.grid = g

    def grid_attach_batch(self, grids: list) -> None:
        """A batch of independent grids (one star per grid, in order); their
        cells are numbered one grid after the other."""
        self._grid_keep = [grid_arrays(g) for g in grids]
        arr = (MocGridDesc * len(grids))(*[to_struct(MocGridDesc, d) for d in self._grid_keep])
        self._check(self._L.moc_grid_attach_batch(self._ctx, len(grids), arr),
                    "moc_grid_attach_batch")
        self.grid = BatchGrid(list(grids))

    def grid_attach_nested(self, grids: list) -> None:
        """A mother grid (grids[0], whose active map holds -g for the cells
        that grid g covers) and its nested sub-grids (moc_grid_attach_nested);
        the cells are numbered one grid after the other."""
        self._grid_keep = [grid_arrays(g) for g in grids]
        arr = (MocGridDesc * len(grids))(*[to_struct(MocGridDesc, d) for d in self._grid_keep])
        self._check(self._L.moc_grid_attach_nested(self._ctx, len(grids), arr),
                    "moc_grid_attach_nested")
        self.grid = BatchGrid(list(grids))

    def _count(self, name: str) -> int:
        c1 = self.grid.nCells + 1
        if name == "ionDen":
            return c1 * self.n_elem * self.nstages
        if name == "Tdust":
            return c1 * (self.dust_shape[0] + 1) * (self.dust_shape[1] + 1)
        if name in GRAIN_FIELDS:
            return c1 * self.dust_shape[0] * self.dust_shape[1]
        return c1 * self.nbins if name in CELLBIN_FIELDS else c1

    # ----------------------------------------------------------------- atomic
    def set_atomic(self, elem_z, abund, nstages: int, xsec) -> None:
        """Element set, abundances and the cross-section stack (an
        atomic.XSecStack built on the same mesh)."""
        self._atomic_keep = atomic_arrays(elem_z, abund, nstages, xsec)
        a = to_struct(MocAtomic, self._atomic_keep)
        self._check(self._L.moc_set_atomic(self._ctx, C.byref(a)), "moc_set_atomic")
        self.n_elem, self.nstages = int(self._atomic_keep["nElemUsed"]), int(nstages)

    def ionizationDriver(self):
        """Opacity of every cell from ionDen, Ne, Te, Hden (ionization_mod.f90:26)."""
        self._check(self._L.moc_opacity(self._ctx, 1), "moc_opacity")

    def set_balance(self, bt) -> None:
        """Balance tables (a balance.BalanceTables built on the same mesh)."""
        self._balance_keep = balance_arrays(bt)
        t = to_struct(MocBalanceTables, self._balance_keep)
        self._check(self._L.moc_set_balance(self._ctx, C.byref(t)), "moc_set_balance")

    def set_emission(self, et) -> None:
        """Emission tables (an emission.EmissionTables built on the same mesh
        and element set); after set_balance."""
        self._emission_keep = emission_arrays(et)
        t = to_struct(MocEmissionTables, self._emission_keep)
        self._check(self._L.moc_set_emission(self._ctx, C.byref(t)), "moc_set_emission")

    def emissionDriver(self) -> float:
        """Diffuse emission of every active cell (emission_mod.f90:36): recPDF
        and totalLines from Te, Ne, ionDen on the device.  Returns kernel ms."""
        ms = C.c_float(0.0)
        self._check(self._L.moc_emission(self._ctx, 1, C.byref(ms)), "moc_emission")
        return float(ms.value)

    def updateCell(self) -> dict:
        """updateCell for every active cell (update_mod.f90:15) from the
        normalised J on the device."""
        st = MocUpdateStats()
        self._check(self._L.moc_update(self._ctx, 1, C.byref(st)), "moc_update")
        return st.as_dict()

    # ----------------------------------------------------------------- output
    def volume_averages(self):
        """outputGas volume averages (output_mod.f90:642-730): (TeVol,
        ionDenVol), each (nElemUsed, nstages), over the cells packets reached."""
        ne, ns = self.n_elem, self.nstages
        T = np.zeros(ne * ns)
        X = np.zeros(ne * ns)
        self._check(self._L.moc_volume_averages(self._ctx, 1, T.ctypes.data_as(_dp),
                                                X.ctypes.data_as(_dp)), "moc_volume_averages")
        return T.reshape(ns, ne).T.copy(), X.reshape(ns, ne).T.copy()

    def volume_averages_noalloc(self):
        """volume_averages through moc_volume_averages_noalloc: the partial
        sums in the library's fixed device buffer, no device allocation."""
        ne, ns = self.n_elem, self.nstages
        T = np.zeros(ne * ns)
        X = np.zeros(ne * ns)
        self._check(self._L.moc_volume_averages_noalloc(self._ctx, 1, T.ctypes.data_as(_dp),
                                                        X.ctypes.data_as(_dp)),
                    "moc_volume_averages_noalloc")
        return T.reshape(ns, ne).T.copy(), X.reshape(ns, ne).T.copy()

    def line_luminosities(self):
        """outputGas line luminosities (output_mod.f90:397-432): (lines, ions,
        Hbeta) with lines[k, i, j] ion k's line from level j+1 to level i+1
        summed over the reached cells times H dV, ions[k] = (Z, stage)."""
        n, L = C.c_int(0), C.c_int(0)
        self._check(self._L.moc_line_luminosities(self._ctx, 1, C.byref(n), C.byref(L), None, None,
                                                  None, C.c_int64(0), None), "moc_line_luminosities")
        nI, nL = n.value, L.value
        lines = np.zeros(max(nI * nL * nL, 1))
        z = np.zeros(max(nI, 1), dtype=np.int32)
        st = np.zeros(max(nI, 1), dtype=np.int32)
        hb = C.c_double(0.0)
        self._check(self._L.moc_line_luminosities(
            self._ctx, 1, C.byref(n), C.byref(L), z.ctypes.data_as(_ip), st.ctypes.data_as(_ip),
            lines.ctypes.data_as(_dp), C.c_int64(lines.size), C.byref(hb)), "moc_line_luminosities")
        return (lines[:nI * nL * nL].reshape(nI, nL, nL), list(zip(z[:nI].tolist(), st[:nI].tolist())),
                hb.value)

    def sed(self):
        """writeSED (output_mod.f90:2947-3064): escaped energy per bin summed
        over the cells (x 8 for an octant model) and its total."""
        out = np.zeros(self.nbins)
        tot = C.c_double(0.0)
        self._check(self._L.moc_sed(self._ctx, 1, out.ctypes.data_as(_dp), C.byref(tot)), "moc_sed")
        return out, tot.value

    # ------------------------------------------------------------------- dust
    def set_dust(self, dt, lgGas: bool, XHILimit: float = 0.05, photoelectric_xsec=None,
                 qheat: dict | None = None, equivalent_tau: bool = False) -> None:
        """Grain tables (a dust.DustTables built on the same mesh); after
        grid_attach.  Tdust starts at 50 K, Ndust at 0.  With gas, passing the
        cross-section stack as photoelectric_xsec switches on the grain
        charge, photoelectric heating and gas-grain collisions (after
        set_balance).  qheat (a dict, see dust_arrays) switches on the quantum
        heating of the grains smaller than its minaQHeat; equivalent_tau
        (lgEquivalentTau) runs the first iteration without the grains'
        opacity, whose escaped SED is then the unextinguished one."""
        self._dust_keep = dust_arrays(dt, lgGas, XHILimit, photoelectric_xsec, qheat,
                                      equivalent_tau)
        t = to_struct(MocDustTables, self._dust_keep)
        self._check(self._L.moc_set_dust(self._ctx, C.byref(t)), "moc_set_dust")
        self.dust_shape = (dt.nSpecies, dt.nSizes)

    def set_qheat_state(self, nIterateMC: int, convPercent: float) -> None:
        """The iteration state that switches quantum heating on for the next
        emission (moc_set_qheat_state); moc_iterate sets it itself."""
        self._check(self._L.moc_set_qheat_state(self._ctx, int(nIterateMC), float(convPercent)),
                    "moc_set_qheat_state")

    def dustOpacity(self):
        """absOpac, scaOpac and the total opacity of every cell (the dust
        block of iterateMC, iteration_mod.f90:194-246)."""
        self._check(self._L.moc_dust_opacity(self._ctx, 1), "moc_dust_opacity")

    def setDustPDF(self):
        """Dust emission CDF of every cell (emission_mod.f90:1921-1986)."""
        self._check(self._L.moc_dust_emission(self._ctx, 1), "moc_dust_emission")

    def getDustT(self) -> dict:
        """Grain temperatures of every cell from the normalised J
        (update_mod.f90:3346-3453)."""
        st = MocUpdateStats()
        self._check(self._L.moc_dust_temperature(self._ctx, 1, C.byref(st)), "moc_dust_temperature")
        return st.as_dict()

    def download_flags(self, name: str) -> np.ndarray:
        """lgConverged / lgBlack as int32 (nCells+1,)."""
        n = self.grid.nCells + 1
        buf = np.zeros(n, dtype=np.int32)
        self._check(self._L.moc_download(self._ctx, FIELDS[name], 1, buf.ctypes.data, n,
                                         MOC_DT_I32, MOC_LAYOUT_FORTRAN), f"moc_download({name})")
        return buf

    def upload(self, name: str, arr: np.ndarray):
        """Upload a field given in MOCASSIN shape ((C+1,) or (C+1, B))."""
        c1 = self.grid.nCells + 1
        a = np.asarray(arr, dtype=np.float64)
        if name == "ionDen":             # (C+1, nElemUsed, nstages) -> Fortran order
            assert a.shape == (c1, self.n_elem, self.nstages), a.shape
            buf = np.ascontiguousarray(a.transpose(2, 1, 0))
        elif name == "Tdust":            # (C+1, nSizes+1, nSpecies+1): species fastest
            nS, nA = self.dust_shape
            assert a.shape == (c1, nA + 1, nS + 1), a.shape
            buf = np.ascontiguousarray(a)
        elif name in CELLBIN_FIELDS:
            assert a.shape == (c1, self.nbins), (name, a.shape)
            buf = np.ascontiguousarray(a.T)      # Fortran order: cell fastest
        else:
            assert a.shape == (c1,), (name, a.shape)
            buf = np.ascontiguousarray(a)
        self._check(self._L.moc_upload(self._ctx, FIELDS[name], 1, buf.ctypes.data, buf.size,
                                       MOC_DT_F64, MOC_LAYOUT_FORTRAN), f"moc_upload({name})")

    def download(self, name: str) -> np.ndarray:
        c1 = self.grid.nCells + 1
        n = self._count(name)
        buf = np.empty(n, dtype=np.float64)
        self._check(self._L.moc_download(self._ctx, FIELDS[name], 1, buf.ctypes.data, n,
                                         MOC_DT_F64, MOC_LAYOUT_FORTRAN), f"moc_download({name})")
        if name in CELLBIN_FIELDS:
            return np.ascontiguousarray(buf.reshape(self.nbins, c1).T)
        if name == "ionDen":
            return np.ascontiguousarray(buf.reshape(self.nstages, self.n_elem, c1).transpose(2, 1, 0))
        if name == "Tdust":
            nS, nA = self.dust_shape
            return buf.reshape(c1, nA + 1, nS + 1)
        if name in GRAIN_FIELDS:         # (nSpecies*nSizes, C+1)
            return buf.reshape(self.dust_shape[0] * self.dust_shape[1], c1)
        return buf

    # -------------------------------------------------------------- transport
    def reset_tallies(self):
        self._check(self._L.moc_reset_tallies(self._ctx, 1), "moc_reset_tallies")

    def energyPacketDriver(self, iStar: int, n: int, seed: int = 20261203,
                           ctr_base: int = 0) -> dict:
        """Transport n packets of star iStar (photon_mod.f90:28); J and escape
        tallies accumulate on the device.  Returns the packet counters."""
        st = MocTransportStats()
        self._check(self._L.moc_transport(self._ctx, int(iStar), int(n), int(seed) & (2**64 - 1),
                                          int(ctr_base) & 0xffffffff, C.byref(st)),
                    "energyPacketDriver")
        return st.as_dict()

    def normalize_j(self):
        self._check(self._L.moc_normalize_j(self._ctx, 1), "moc_normalize_j")

    # -------------------------------------------------------------- iteration
    def set_packets(self, iStar: int, n: int) -> None:
        """nPhotons(iStar) for the following iterations (packet energy L/n)."""
        self._check(self._L.moc_set_packets(self._ctx, int(iStar), int(n)), "moc_set_packets")
        self.tables.nPhotons[int(iStar) - 1] = int(n)

    def iterateMC(self, nIterateMC: int, seed: int = 20261203) -> dict:
        """One Monte Carlo iteration on the device (iteration_mod.f90:32):
        opacity, emission, packets of every star, J normalisation, updateCell."""
        st = MocIterationStats()
        self._check(self._L.moc_iterate(self._ctx, 1, int(nIterateMC), int(seed) & (2**64 - 1),
                                        C.byref(st)), "moc_iterate")
        return st.as_dict()

    def set_graph(self, enable: bool) -> None:
        """Replay iterations from the second on as one captured graph launch
        (moc_set_graph); the packets drawn are those of the eager call."""
        self._check(self._L.moc_set_graph(self._ctx, 1 if enable else 0), "moc_set_graph")

    def get_tuning(self) -> dict:
        """The transport's launch parameters and the graph switch."""
        t = MocTuning()
        self._check(self._L.moc_get_tuning(self._ctx, C.byref(t)), "moc_get_tuning")
        return t.as_dict()

    def set_tuning(self, emitBatch: int | None = None, segsPerTrip: int | None = None,
                   transportWaves: int | None = None, hotJ: int | None = None) -> None:
        """Set some of the transport's launch parameters (moc_set_tuning); the
        others keep their values.  They change no packet's random stream."""
        cur = self.get_tuning()
        t = MocTuning(int(emitBatch if emitBatch is not None else cur["emitBatch"]),
                      int(segsPerTrip if segsPerTrip is not None else cur["segsPerTrip"]),
                      int(transportWaves if transportWaves is not None else cur["transportWaves"]),
                      0, int(hotJ if hotJ is not None else cur["hotJ"]))
        self._check(self._L.moc_set_tuning(self._ctx, C.byref(t)), "moc_set_tuning")

    def MCIterationDriver(self, maxIterateMC: int, minConvergence: float = 100.0,
                          seed: int = 20261203, lgAutoPackets: bool = False,
                          convIncPercent: float = 0.0, nPhotIncrease: float = 2.0,
                          maxPhotons: float = 0.0, callback=None) -> list:
        """The iteration loop of MCIterationDriver (iteration_mod.f90:17-1497):
        iterate until the converged-cell percentage reaches minConvergence or
        maxIterateMC iterations have run; with lgAutoPackets the packets of
        every star grow by nPhotIncrease when convergence improved by no more
        than convIncPercent (:1412-1421).  Returns the per-iteration stats."""
        history = []
        old = 0.0
        n = 1
        while True:
            st = self.iterateMC(n, seed)
            st["nIterateMC"] = n
            history.append(st)
            if callback is not None:
                callback(st)
            tot = st["convPercent"]
            n_tot = sum(self.tables.nPhotons) + self.tables.nPhotons[0]
            if (lgAutoPackets and n > 1 and tot < 95.0 and n_tot < maxPhotons and old > 0.0
                    and (tot - old) / old <= convIncPercent):
                for s, np_ in enumerate(list(self.tables.nPhotons)):
                    self.set_packets(s + 1, int(np_ * nPhotIncrease))
            old = tot
            if tot >= minConvergence or n >= maxIterateMC:
                return history
            n += 1


def version() -> str:
    return _lib.lib().moc_version().decode()

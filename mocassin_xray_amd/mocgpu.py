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
from ._lib import (CELLBIN_FIELDS, FIELDS, MOC_DT_F64, MOC_DT_I32, MOC_LAYOUT_FORTRAN,
                   MOC_PREC_FP64, MOC_PREC_MIXED, MocAtomic, MocBalanceTables, MocEmissionTables,
                   MocGridDesc, MocIterationStats, MocTables, MocTransportStats, MocUpdateStats, _dp, _ip, dptr,
                   iptr)


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
    extra: dict = field(default_factory=dict)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class MocGPU:
    def __init__(self, tables: SourceTables, precision: str = "fp64"):
        self._L = _lib.lib()
        # own copy: set_packets changes nPhotons
        self.tables = dataclasses.replace(tables, nPhotons=np.array(tables.nPhotons, dtype=np.float64))
        self.precision = precision
        prec = {"fp64": MOC_PREC_FP64, "mixed": MOC_PREC_MIXED}[precision]
        B = tables.nbins
        self._keep = [_c(tables.nuArray), _c(tables.widFlx), _c(tables.KNsigmaT),
                      _c(tables.KNcdf), _c(tables.Pcomp), _c(tables.sourceCDF),
                      _c(tables.Lstar), _c(tables.nPhotons), _c(tables.starPosition),
                      _c(tables.starIndex, np.int32)]
        k = self._keep
        nst = int(np.atleast_1d(tables.Lstar).size)
        assert k[5].size == nst * B and k[3].size == k[4].size and k[3].size % B == 0
        t = MocTables(nbins=B, precision=prec, nuArray=dptr(k[0]), widFlx=dptr(k[1]),
                      ionEdge1=float(tables.ionEdge1), lgCompton=int(bool(tables.lgCompton)),
                      nAngles=k[3].size // B, KNsigmaT=dptr(k[2]), KNsigmaCDF=dptr(k[3]),
                      Pcomp=dptr(k[4]), nStars=nst, sourceCDF=dptr(k[5]), Lstar=dptr(k[6]),
                      nPhotons=dptr(k[7]), starPosition=dptr(k[8]), starIndex=iptr(k[9]))
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
        act = np.asfortranarray(g.active, dtype=np.int32).reshape(-1, order="F")
        act = np.ascontiguousarray(act)
        xa, ya, za = _c(g.xAxis), _c(g.yAxis), _c(g.zAxis)
        d = MocGridDesc(nx=xa.size, ny=ya.size, nz=za.size, xAxis=dptr(xa), yAxis=dptr(ya),
                        zAxis=dptr(za), active=iptr(act), nCells=int(g.nCells),
                        lgSymmetricXYZ=int(bool(g.lgSymmetricXYZ)), R_out=float(g.R_out))
        self._check(self._L.moc_grid_attach(self._ctx, 1, C.byref(d)), "moc_grid_attach")
        self.grid = g

    def _count(self, name: str) -> int:
        c1 = self.grid.nCells + 1
        if name == "ionDen":
            return c1 * self.n_elem * self.nstages
        return c1 * self.nbins if name in CELLBIN_FIELDS else c1

    # ----------------------------------------------------------------- atomic
    def set_atomic(self, elem_z, abund, nstages: int, xsec) -> None:
        """Element set, abundances and the cross-section stack (an
        atomic.XSecStack built on the same mesh)."""
        z = _c(elem_z, np.int32)
        zi = {int(v): i for i, v in enumerate(z)}
        spec_e = _c([zi[e] for e, _ in xsec.species], np.int32)
        spec_i = _c([i for _, i in xsec.species], np.int32)
        self._atomic_keep = [z, _c(abund), spec_e, spec_i, _c(xsec.sigma), _c(xsec.brems)]
        k = self._atomic_keep
        a = MocAtomic(nElemUsed=z.size, elemZ=iptr(k[0]), abund=dptr(k[1]), nstages=int(nstages),
                      nSpecies=spec_e.size, speciesElem=iptr(k[2]), speciesIon=iptr(k[3]),
                      sigma=dptr(k[4]), brems=dptr(k[5]))
        self._check(self._L.moc_set_atomic(self._ctx, C.byref(a)), "moc_set_atomic")
        self.n_elem, self.nstages = int(z.size), int(nstages)

    def ionizationDriver(self):
        """Opacity of every cell from ionDen, Ne, Te, Hden (ionization_mod.f90:26)."""
        self._check(self._L.moc_opacity(self._ctx, 1), "moc_opacity")

    def set_balance(self, bt) -> None:
        """Balance tables (a balance.BalanceTables built on the same mesh)."""
        d = bt.data
        lp = C.POINTER(C.c_int64)
        k = {
            "sn": _c(bt.shellNum, np.int32), "ss": _c(bt.speciesShell, np.int32), "W": _c(bt.W),
            "os": _c(bt.outShell, np.int32), "g0": _c(bt.statW0, np.int32),
            "g1": _c(bt.statW1, np.int32), "rrb": _c(d["rrb"]), "dc": _c(d["drb_c"]),
            "de": _c(d["drb_e"]), "rrec": _c(d["rrec"]), "rnew": _c(d["rnew"]),
            "diel": _c(d["diel"]), "aldro": _c(d["aldro"]), "cf": _c(d["cion_cf"]),
            "th": _c(d["cion_th"]), "na": _c(d["nauger"], np.int32), "au": _c(d["auger"]),
            "cz": _c(bt.chex[:, 0], np.int32), "ci": _c(bt.chex[:, 1], np.int32),
            "cfit": _c(bt.chex[:, 2:6]), "ai": _c(d["atom_idx"], np.int32),
            "am": _c(bt.atomMeta8, np.int64), "atp": _c(d["atom_atp"]), "lt": _c(d["atom_logt"]),
            "ag": _c(d["atom_g"]), "ae": _c(d["atom_e"]), "aA": _c(d["atom_A"]),
            "aq": _c(bt.atomQom), "aq2": _c(bt.atomQ2), "rk": _c(d["atom_rec_kind"], np.int32),
            "rt": _c(d["atom_rec"]), "rv": _c(d["atom_rec_valid"], np.int32),
        }
        self._balance_keep = k
        t = MocBalanceTables(
            maxLev=int(bt.maxLev), XHILimit=float(bt.XHILimit), nShells=k["sn"].size,
            shellNum=iptr(k["sn"]), speciesShell=iptr(k["ss"]), W=dptr(k["W"]),
            outShell=iptr(k["os"]), statW0=iptr(k["g0"]), statW1=iptr(k["g1"]),
            rrb=dptr(k["rrb"]), drb_c=dptr(k["dc"]), drb_e=dptr(k["de"]), rrec=dptr(k["rrec"]),
            rnew=dptr(k["rnew"]), nDiel=k["diel"].shape[0], diel=dptr(k["diel"]),
            nAldro=k["aldro"].shape[0], aldro=dptr(k["aldro"]), cionCF=dptr(k["cf"]),
            cionTherm=dptr(k["th"]), nAuger=iptr(k["na"]), auger=dptr(k["au"]),
            nChex=k["cz"].size, chexZ=iptr(k["cz"]), chexIon=iptr(k["ci"]),
            chexFit=dptr(k["cfit"]), nAtoms=k["am"].shape[0], atomIdx=iptr(k["ai"]),
            atomMeta=k["am"].ctypes.data_as(lp), atomAtp=dptr(k["atp"]), nLogT=k["lt"].size,
            nLev=k["ag"].size, nA=k["aA"].size, nQ=k["aq"].size, atomLogT=dptr(k["lt"]),
            atomG=dptr(k["ag"]), atomE=dptr(k["ae"]), atomA=dptr(k["aA"]),
            atomQom=dptr(k["aq"]), atomQ2=dptr(k["aq2"]), atomRecKind=iptr(k["rk"]),
            atomRec=dptr(k["rt"]), atomRecValid=iptr(k["rv"]))
        self._check(self._L.moc_set_balance(self._ctx, C.byref(t)), "moc_set_balance")

    def set_emission(self, et) -> None:
        """Emission tables (an emission.EmissionTables built on the same mesh
        and element set); after set_balance."""
        lp = C.POINTER(C.c_int64)
        k = {"tk": _c(et.tk), "edge": [_c(e, np.int32) for e in et.edge],
             "lg": [_c(v) for v in et.loggam], "xsHI": _c(et.xsHI),
             "xsHeI": None if et.xsHeI is None else _c(et.xsHeI),
             "xsHeII": None if et.xsHeII is None else _c(et.xsHeII),
             "fbZ": _c(et.fbZ, np.int32), "fbIon": _c(et.fbIon, np.int32),
             "fbLo": _c(et.fbLo, np.int32), "fbHi": _c(et.fbHi, np.int32),
             "fbOff": _c(et.fbOff, np.int64), "fbXs": _c(et.fbXs), "fbG": _c(et.fbGratio),
             "he2j": _c(et.he2j, np.int32), "he2": _c(et.hei2phot), "map": _c(et.hlMap, np.int32),
             "dens": _c(et.hlDens), "lines": _c(et.hlLines), "hbA": _c(et.hbA), "hbB": _c(et.hbB),
             "rb": _c(et.rbEdge), "r2a": _c(et.r2aEdge), "heiRec": _c(et.heiRec),
             "lym": _c(et.heiiLyman)}
        self._emission_keep = k
        opt = lambda a: dptr(a) if a is not None else None
        t = MocEmissionTables(
            ntk=k["tk"].size, tkGamma=dptr(k["tk"]), nlim=(C.c_int32 * 3)(*[e.size for e in k["edge"]]),
            edgeP=(_ip * 3)(*[iptr(e) for e in k["edge"]]), logGamma=(_dp * 3)(*[dptr(v) for v in k["lg"]]),
            HIPnuP=int(et.HIPnuP), HeIPnuP=int(et.HeIPnuP), HeIIPnuP=int(et.HeIIPnuP),
            xsHI=dptr(k["xsHI"]), xsHeI=opt(k["xsHeI"]), xsHeII=opt(k["xsHeII"]),
            nFb=k["fbZ"].size, fbZ=iptr(k["fbZ"]), fbIon=iptr(k["fbIon"]), fbLo=iptr(k["fbLo"]),
            fbHi=iptr(k["fbHi"]), fbOff=k["fbOff"].ctypes.data_as(lp), nFbXs=k["fbXs"].size,
            fbXs=dptr(k["fbXs"]), fbGratio=dptr(k["fbG"]), he2Row=iptr(k["he2j"]),
            hei2phot=dptr(k["he2"]), nHlFiles=k["dens"].shape[0], hlMap=iptr(k["map"]),
            hlDens=dptr(k["dens"]), hlLines=dptr(k["lines"]), hbA=dptr(k["hbA"]),
            hbB=dptr(k["hbB"]), rbEdge=dptr(k["rb"]), r2aEdge=dptr(k["r2a"]),
            heiRec=dptr(k["heiRec"]), heiiLyman=dptr(k["lym"]), j2TsP=int(et.j2TsP),
            lymP=(C.c_int32 * 4)(*[int(v) for v in et.lymP]))
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

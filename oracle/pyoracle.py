"""ctypes front end of the ORACLE (liboracle.so) -- test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
where it checks the device path; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SubGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("xAxis", _dp), ("yAxis", _dp),
                ("zAxis", _dp), ("active", _ip)]


class TransportIn(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nCells", C.c_int), ("sym", C.c_int),
        ("xAxis", _dp), ("yAxis", _dp), ("zAxis", _dp), ("active", _ip), ("R_out", C.c_double),
        ("nbins", C.c_int), ("nAngles", C.c_int), ("lgCompton", C.c_int), ("ionEdge1", C.c_double),
        ("nuArray", _dp), ("KNsigmaT", _dp), ("KNcdf", _dp), ("Pcomp", _dp), ("sourceCDF", _dp),
        ("opacity", _dp), ("Ne", _dp), ("recPDF", _dp), ("totalLines", _dp),
        ("scaOpac", _dp), ("lgDust", C.c_int), ("dustOnly", C.c_int),
        ("star", C.c_double * 3), ("starIdx", C.c_int32 * 3), ("deltaE", C.c_double),
        ("n", C.c_int64), ("seed", C.c_uint64), ("ctr_base", C.c_uint32), ("nthreads", C.c_int),
        ("id_base", C.c_uint64), ("plane", C.c_int),
        ("nSub", C.c_int), ("subs", C.POINTER(SubGrid)), ("startGrid", C.c_int),
    ]


class TransportOut(C.Structure):
    _fields_ = [("J", _dp), ("escaped", _dp), ("counters", C.c_double * 9), ("Qphot", C.c_double)]


COUNTER_NAMES = ["segments", "emissions", "absorbed", "scattered", "escaped", "lines", "lost",
                 "trapped", "errors"]

_L = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32)]
        L.oracle_getnu2.argtypes = [_dp, C.c_int, C.c_double]
        L.oracle_getnu2.restype = C.c_int
        L.oracle_locate.argtypes = [_dp, C.c_int, C.c_double]
        L.oracle_locate.restype = C.c_int
        L.oracle_transport.argtypes = [C.POINTER(TransportIn), C.POINTER(TransportOut)]
        L.oracle_transport.restype = C.c_int
        _L = L
    return _L


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return list(o)


def getnu2(cdf, r) -> int:
    a = np.ascontiguousarray(cdf, dtype=np.float64)
    return lib().oracle_getnu2(a.ctypes.data_as(_dp), a.size, float(r))


def locate(xa, x) -> int:
    a = np.ascontiguousarray(xa, dtype=np.float64)
    return lib().oracle_locate(a.ctypes.data_as(_dp), a.size, float(x))


def nest_start(grids, star, idx):
    """The grid (0: the mother) and 1-based cell of a source in nested grids:
    a mother cell that a sub-grid covers puts it in that sub-grid, at the cell
    that locate and the midpoint rule give (setStarPosition)."""
    m = grids[0]
    cm = int(np.asarray(m.active)[idx[0] - 1, idx[1] - 1, idx[2] - 1])
    if cm >= 0:
        return 0, [int(v) for v in idx]
    g = grids[-cm - 1]
    out = []
    for a, x in ((g.xAxis, star[0]), (g.yAxis, star[1]), (g.zAxis, star[2])):
        a = np.asarray(a, dtype=np.float64)
        p = max(locate(a, float(x)), 1)
        if p < a.size and x > (a[p - 1] + a[p]) / 2.0:
            p += 1
        out.append(p)
    return -cm - 1, out


def transport(model_tables, grid, fields, n, seed, ctr_base=0, nthreads=1, istar=0, id_base=0,
              nested=None):
    """Run the oracle transport.  Returns (J[C+1,B], escaped[C+1,B], counters, Qphot).
    With grains, fields holds scaOpac (C+1, B) and dustOnly (bool).  Packet i
    draws from the RNG stream of packet id id_base + i.  nested: a list of
    grids, the mother first (whose active map holds -g for the cells of grid
    g), the cells numbered one grid after the other; grid is then ignored."""
    t = model_tables
    B = t.nbins
    if nested is not None:
        grid = nested[0]
    C1 = (sum(g.nCells for g in nested) if nested is not None else grid.nCells) + 1
    keep = {}

    def d(name, a):
        keep[name] = np.ascontiguousarray(a, dtype=np.float64)
        return keep[name].ctypes.data_as(_dp)

    act = np.ascontiguousarray(np.asarray(grid.active).reshape(-1, order="F"), dtype=np.int32)
    keep["act"] = act
    src = np.asarray(t.sourceCDF).reshape(-1, B)[istar]
    deltaE = float(np.atleast_1d(t.Lstar)[istar] / np.atleast_1d(t.nPhotons)[istar])
    inp = TransportIn(
        nx=len(grid.xAxis), ny=len(grid.yAxis), nz=len(grid.zAxis), nCells=C1 - 1,
        sym=int(bool(grid.lgSymmetricXYZ)), xAxis=d("x", grid.xAxis), yAxis=d("y", grid.yAxis),
        zAxis=d("z", grid.zAxis), active=act.ctypes.data_as(_ip), R_out=float(grid.R_out),
        nbins=B, nAngles=np.asarray(t.KNcdf).shape[1], lgCompton=int(bool(t.lgCompton)),
        ionEdge1=float(t.ionEdge1), nuArray=d("nu", t.nuArray), KNsigmaT=d("kn", t.KNsigmaT),
        KNcdf=d("kncdf", t.KNcdf), Pcomp=d("pc", t.Pcomp), sourceCDF=d("src", src),
        opacity=d("op", np.asarray(fields["opacity"]).T), Ne=d("ne", fields["Ne"]),
        recPDF=d("rec", fields["recPDF"]), totalLines=d("tl", fields["totalLines"]),
        scaOpac=d("sca", np.asarray(fields["scaOpac"]).T) if "scaOpac" in fields else None,
        lgDust=int("scaOpac" in fields), dustOnly=int(bool(fields.get("dustOnly", False))),
        deltaE=deltaE, n=int(n), seed=int(seed), ctr_base=int(ctr_base), nthreads=int(nthreads),
        id_base=int(id_base), plane=int(bool(getattr(grid, "lgPlaneIonization", False))))
    pos = np.asarray(t.starPosition).reshape(-1, 3)[istar]
    idx = np.asarray(t.starIndex).reshape(-1, 3)[istar]
    if nested is not None:
        subs = (SubGrid * (len(nested) - 1))()
        base = nested[0].nCells
        for k, g in enumerate(nested[1:]):
            a = np.asarray(g.active, dtype=np.int32).reshape(-1, order="F")
            ga = np.ascontiguousarray(np.where(a > 0, a + base, 0), dtype=np.int32)
            keep[f"sub{k}"] = ga
            subs[k] = SubGrid(nx=len(g.xAxis), ny=len(g.yAxis), nz=len(g.zAxis),
                              xAxis=d(f"sx{k}", g.xAxis), yAxis=d(f"sy{k}", g.yAxis),
                              zAxis=d(f"sz{k}", g.zAxis), active=ga.ctypes.data_as(_ip))
            base += g.nCells
        keep["subs"] = subs
        inp.nSub = len(nested) - 1
        inp.subs = C.cast(subs, C.POINTER(SubGrid))
        inp.startGrid, idx = nest_start(nested, pos, idx)
    for k in range(3):
        inp.star[k] = float(pos[k])
        inp.starIdx[k] = int(idx[k])
    J = np.zeros(B * C1)
    E = np.zeros(B * C1)
    out = TransportOut(J=J.ctypes.data_as(_dp), escaped=E.ctypes.data_as(_dp))
    rc = lib().oracle_transport(C.byref(inp), C.byref(out))
    cnt = {n_: out.counters[i] for i, n_ in enumerate(COUNTER_NAMES)}
    cnt["rc"] = rc
    return J.reshape(B, C1).T.copy(), E.reshape(B, C1).T.copy(), cnt, out.Qphot


class OpacityIn(C.Structure):
    _fields_ = [
        ("nbins", C.c_int), ("nCells", C.c_int), ("nstages", C.c_int), ("nElemUsed", C.c_int),
        ("nuArray", _dp), ("brems", _dp), ("KNsigmaT", _dp), ("lgCompton", C.c_int),
        ("nShells", C.c_int), ("shZ", _ip), ("shIon", _ip), ("shLo", _ip), ("shHi", _ip),
        ("shOff", C.POINTER(C.c_int64)), ("xs", _dp), ("elemZ", _ip), ("abund", _dp),
        ("ionDen", _dp), ("Ne", _dp), ("Te", _dp), ("Hden", _dp),
    ]


def gaunt_ff(z, log10Te, log10nu):
    L = lib()
    L.oracle_gaunt_ff.argtypes = [C.c_double, C.c_double, _dp, C.c_int, _dp]
    x = np.ascontiguousarray(log10nu, dtype=np.float64)
    g = np.zeros_like(x)
    L.oracle_gaunt_ff(float(z), float(log10Te), x.ctypes.data_as(_dp), x.size,
                      g.ctypes.data_as(_dp))
    return g


def opacity(xsec, kn_sigma, lgCompton, elem_z, abund, nstages, ionDen, Ne, Te, Hden):
    """Oracle opacity (C+1, B) from the per-shell stack of an atomic.XSecStack.
    ionDen has shape (C+1, nElemUsed, nstages)."""
    L = lib()
    L.oracle_opacity.argtypes = [C.POINTER(OpacityIn), _dp]
    L.oracle_opacity.restype = C.c_int
    B = xsec.nu.size
    C1 = ionDen.shape[0]
    sh = xsec.shells
    keep = {
        "nu": np.ascontiguousarray(xsec.nu, dtype=np.float64),
        "br": np.ascontiguousarray(xsec.brems, dtype=np.float64),
        "kn": np.ascontiguousarray(kn_sigma, dtype=np.float64),
        "z": np.array([s.elem for s in sh], dtype=np.int32),
        "ion": np.array([s.ion for s in sh], dtype=np.int32),
        "lo": np.array([s.lo for s in sh], dtype=np.int32),
        "hi": np.array([s.hi for s in sh], dtype=np.int32),
        "off": np.cumsum([0] + [s.sigma.size for s in sh[:-1]]).astype(np.int64),
        "xs": np.ascontiguousarray(np.concatenate([s.sigma for s in sh]) if sh else np.zeros(1)),
        "ez": np.ascontiguousarray(elem_z, dtype=np.int32),
        "ab": np.ascontiguousarray(abund, dtype=np.float64),
        "id": np.ascontiguousarray(np.asarray(ionDen, dtype=np.float64).transpose(2, 1, 0)),
        "ne": np.ascontiguousarray(Ne, dtype=np.float64),
        "te": np.ascontiguousarray(Te, dtype=np.float64),
        "h": np.ascontiguousarray(Hden, dtype=np.float64),
    }
    ip = lambda k: keep[k].ctypes.data_as(_ip)
    dp = lambda k: keep[k].ctypes.data_as(_dp)
    inp = OpacityIn(nbins=B, nCells=C1 - 1, nstages=int(nstages), nElemUsed=keep["ez"].size,
                    nuArray=dp("nu"), brems=dp("br"), KNsigmaT=dp("kn"),
                    lgCompton=int(bool(lgCompton)), nShells=len(sh), shZ=ip("z"), shIon=ip("ion"),
                    shLo=ip("lo"), shHi=ip("hi"),
                    shOff=keep["off"].ctypes.data_as(C.POINTER(C.c_int64)), xs=dp("xs"),
                    elemZ=ip("ez"), abund=dp("ab"), ionDen=dp("id"), Ne=dp("ne"), Te=dp("te"),
                    Hden=dp("h"))
    out = np.zeros(B * C1)
    L.oracle_opacity(C.byref(inp), out.ctypes.data_as(_dp))
    return out.reshape(B, C1).T.copy()


class BalanceTablesC(C.Structure):
    _fields_ = [
        ("nbins", C.c_int), ("nstages", C.c_int), ("nElemUsed", C.c_int), ("maxLev", C.c_int),
        ("lgCompton", C.c_int), ("elemZ", _ip), ("abund", _dp), ("nuArray", _dp),
        ("nShells", C.c_int), ("shZ", _ip), ("shIon", _ip), ("shNum", _ip), ("shLo", _ip),
        ("shHi", _ip), ("shOff", C.POINTER(C.c_int64)), ("xs", _dp), ("secIonP", C.c_int),
        ("cRecoilHP", C.c_int), ("cRecoilHeP", C.c_int), ("xSecRecoil", _dp), ("rrb", _dp),
        ("drb_c", _dp), ("drb_e", _dp), ("rrec", _dp), ("rnew", _dp), ("nDiel", C.c_int),
        ("diel", _dp), ("nAldro", C.c_int), ("aldro", _dp), ("cion_cf", _dp), ("cion_th", _dp),
        ("nauger", _ip), ("auger", _dp), ("atom_idx", _ip), ("atom_meta", C.POINTER(C.c_int64)),
        ("atom_atp", _dp), ("atom_logt", _dp), ("atom_g", _dp), ("atom_e", _dp), ("atom_A", _dp),
        ("atom_qom", C.POINTER(C.c_float)), ("atom_rec_kind", _ip), ("atom_rec", _dp),
        ("atom_rec_valid", _ip),
    ]


class BalanceStateC(C.Structure):
    _fields_ = [
        ("nCells", C.c_int), ("J", _dp), ("Te", _dp), ("Ne", _dp), ("Hden", _dp), ("ionDen", _dp),
        ("lgConverged", _ip), ("lgBlack", _ip), ("nIterT", _ip), ("XHILimit", C.c_double),
        ("noHit", C.c_double), ("noIonBal", C.c_double), ("noTeBal", C.c_double),
        ("gS", C.c_int), ("gA", C.c_int), ("gAbun", _dp), ("gWeight", _dp), ("gRadius", _dp),
        ("gVn", _dp), ("gMsurf", _dp), ("gCabs", _dp), ("Tdust", _dp), ("Ndust", _dp),
        ("collD", _dp), ("peD", _dp), ("grainPot", _dp),
    ]


def balance_tables(bt, elem_z, abund, nstages, lgCompton):
    """ctypes view of the oracle's balance tables; returns (struct, keep-alive dict)."""
    d = bt.data
    sh = bt.xsec.shells
    k = {
        "ez": np.ascontiguousarray(elem_z, dtype=np.int32),
        "ab": np.ascontiguousarray(abund, dtype=np.float64),
        "nu": np.ascontiguousarray(bt.nu, dtype=np.float64),
        "z": np.array([s.elem for s in sh], dtype=np.int32),
        "ion": np.array([s.ion for s in sh], dtype=np.int32),
        "num": np.array([s.shell for s in sh], dtype=np.int32),
        "lo": np.array([s.lo for s in sh], dtype=np.int32),
        "hi": np.array([s.hi for s in sh], dtype=np.int32),
        "off": np.cumsum([0] + [s.sigma.size for s in sh[:-1]]).astype(np.int64),
        "xs": np.ascontiguousarray(np.concatenate([s.sigma for s in sh])),
        "rec": np.ascontiguousarray(bt.xSecRecoil, dtype=np.float64),
        "am": np.ascontiguousarray(d["atom_meta"], dtype=np.int64),
        "aq": np.ascontiguousarray(d["atom_qom"], dtype=np.float32),
    }
    for name in ("rrb", "drb_c", "drb_e", "rrec", "rnew", "diel", "aldro", "cion_cf", "cion_th",
                 "auger", "atom_atp", "atom_logt", "atom_g", "atom_e", "atom_A"):
        k[name] = np.ascontiguousarray(d[name], dtype=np.float64)
    for name in ("nauger", "atom_idx", "atom_rec_kind", "atom_rec_valid"):
        k[name] = np.ascontiguousarray(d[name], dtype=np.int32)
    k["atom_rec"] = np.ascontiguousarray(d["atom_rec"], dtype=np.float64)
    dp = lambda n: k[n].ctypes.data_as(_dp)
    ip = lambda n: k[n].ctypes.data_as(_ip)
    t = BalanceTablesC(
        nbins=bt.nu.size, nstages=int(nstages), nElemUsed=k["ez"].size, maxLev=int(bt.maxLev),
        lgCompton=int(bool(lgCompton)), elemZ=ip("ez"), abund=dp("ab"), nuArray=dp("nu"),
        nShells=len(sh), shZ=ip("z"), shIon=ip("ion"), shNum=ip("num"), shLo=ip("lo"),
        shHi=ip("hi"), shOff=k["off"].ctypes.data_as(C.POINTER(C.c_int64)), xs=dp("xs"),
        secIonP=bt.secIonP, cRecoilHP=bt.cRecoilHP, cRecoilHeP=bt.cRecoilHeP,
        xSecRecoil=dp("rec"), rrb=dp("rrb"), drb_c=dp("drb_c"), drb_e=dp("drb_e"),
        rrec=dp("rrec"), rnew=dp("rnew"), nDiel=k["diel"].shape[0], diel=dp("diel"),
        nAldro=k["aldro"].shape[0], aldro=dp("aldro"), cion_cf=dp("cion_cf"),
        cion_th=dp("cion_th"), nauger=ip("nauger"), auger=dp("auger"), atom_idx=ip("atom_idx"),
        atom_meta=k["am"].ctypes.data_as(C.POINTER(C.c_int64)), atom_atp=dp("atom_atp"),
        atom_logt=dp("atom_logt"), atom_g=dp("atom_g"), atom_e=dp("atom_e"), atom_A=dp("atom_A"),
        atom_qom=k["aq"].ctypes.data_as(C.POINTER(C.c_float)), atom_rec_kind=ip("atom_rec_kind"),
        atom_rec=dp("atom_rec"), atom_rec_valid=ip("atom_rec_valid"))
    return t, k


def alpha(bt, elem_z, nstages, Te):
    """alphaTot(Z, stage) [31, 32] of the oracle's calcAlpha at Te."""
    L = lib()
    L.oracle_alpha.argtypes = [C.POINTER(BalanceTablesC), C.c_double, _dp]
    t, k = balance_tables(bt, elem_z, np.ones(len(elem_z)), nstages, True)
    out = np.zeros(31 * 32)
    L.oracle_alpha(C.byref(t), float(Te), out.ctypes.data_as(_dp))
    return out.reshape(31, 32)


def nlevel_cooling(bt, atom, Te, Ne, scale=1.0):
    L = lib()
    L.oracle_nlevel_cooling.argtypes = [C.POINTER(BalanceTablesC), C.c_int, C.c_double,
                                        C.c_double, C.c_double]
    L.oracle_nlevel_cooling.restype = C.c_double
    t, k = balance_tables(bt, [1], np.ones(1), 2, True)
    return L.oracle_nlevel_cooling(C.byref(t), int(atom), float(Te), float(Ne), float(scale))


def update(bt, elem_z, abund, nstages, lgCompton, J, Te, Ne, Hden, ionDen, nthreads=1,
           grains=None):
    """Oracle updateCell over all cells.  J is (C+1, B) normalised; ionDen is
    (C+1, nElemUsed, nstages).  Returns a dict with the new Te, Ne, ionDen,
    flags, iteration counts and counters.  grains = (DustTables, Tdust
    (C+1, nA+1, nS+1), Ndust (C+1,)) adds the grain charge, photoelectric and
    gas-grain collision terms; the result then also holds collD, peD and
    grainPot, each (nS*nA, C+1)."""
    L = lib()
    L.oracle_update.argtypes = [C.POINTER(BalanceTablesC), C.POINTER(BalanceStateC), C.c_int]
    L.oracle_update.restype = C.c_int
    t, k = balance_tables(bt, elem_z, abund, nstages, lgCompton)
    C1 = J.shape[0]
    s = {
        "J": np.ascontiguousarray(np.asarray(J, dtype=np.float64).T),
        "Te": np.array(Te, dtype=np.float64), "Ne": np.array(Ne, dtype=np.float64),
        "H": np.ascontiguousarray(Hden, dtype=np.float64),
        "X": np.ascontiguousarray(np.asarray(ionDen, dtype=np.float64).transpose(2, 1, 0)),
        "conv": np.zeros(C1, dtype=np.int32), "black": np.zeros(C1, dtype=np.int32),
        "nit": np.zeros(C1, dtype=np.int32),
    }
    st = BalanceStateC(nCells=C1 - 1, J=s["J"].ctypes.data_as(_dp), Te=s["Te"].ctypes.data_as(_dp),
                       Ne=s["Ne"].ctypes.data_as(_dp), Hden=s["H"].ctypes.data_as(_dp),
                       ionDen=s["X"].ctypes.data_as(_dp), lgConverged=s["conv"].ctypes.data_as(_ip),
                       lgBlack=s["black"].ctypes.data_as(_ip), nIterT=s["nit"].ctypes.data_as(_ip),
                       XHILimit=float(bt.XHILimit))
    if grains is not None:
        dt, Tdust, Ndust = grains
        nG = dt.nSpecies * dt.nSizes
        for name, v in (("gAbun", dt.grainAbun), ("gWeight", dt.grainWeight), ("gRadius", dt.radius),
                        ("gVn", dt.Vn), ("gMsurf", dt.Msurf), ("gCabs", dt.Cabs), ("Tdust", Tdust),
                        ("Ndust", Ndust)):
            s[name] = np.ascontiguousarray(v, dtype=np.float64)
            setattr(st, name, s[name].ctypes.data_as(_dp))
        for name in ("collD", "peD", "grainPot"):
            s[name] = np.zeros((nG, C1))
            setattr(st, name, s[name].ctypes.data_as(_dp))
        st.gS, st.gA = dt.nSpecies, dt.nSizes
    L.oracle_update(C.byref(t), C.byref(st), int(nthreads))
    out = {"Te": s["Te"], "Ne": s["Ne"],
           "ionDen": s["X"].transpose(2, 1, 0).copy(), "lgConverged": s["conv"],
           "lgBlack": s["black"], "nIterT": s["nit"], "noHit": st.noHit,
           "noIonBal": st.noIonBal, "noTeBal": st.noTeBal}
    if grains is not None:
        out.update(collD=s["collD"], peD=s["peD"], grainPot=s["grainPot"])
    return out


def heavy_chain(bt, z, Te, Ne, H, XH0, alpha, cI1, pI1, X, damp=False, nstages=31):
    """The oracle's heavy-element chain of ionBalance for element z alone:
    alpha, cI1 [32] and pI1 [11][32] (shell, stage) rates, X [32] the
    stage fractions in; returns the new X [32]."""
    L = lib()
    L.oracle_heavy_chain.argtypes = [C.POINTER(BalanceTablesC), C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, _dp]
    t, k = balance_tables(bt, [1, z], np.ones(2), nstages, True)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    c = np.ascontiguousarray(cI1, dtype=np.float64)
    p = np.ascontiguousarray(pI1, dtype=np.float64)
    x = np.array(X, dtype=np.float64)
    L.oracle_heavy_chain(C.byref(t), int(z), float(Te), float(Ne), float(H), float(XH0),
                         a.ctypes.data_as(_dp), c.ctypes.data_as(_dp), p.ctypes.data_as(_dp),
                         int(bool(damp)), x.ctypes.data_as(_dp))
    return x


def cell_heat_cool(bt, elem_z, abund, nstages, lgCompton, J, Ne, Hden, ionDen, cell, Te):
    """One T iteration of the oracle's updateCell for `cell` at Te from the
    given state (J (C+1, B), ionDen (C+1, nElemUsed, nstages)): (heat, cool)."""
    L = lib()
    L.oracle_cell_heat_cool.argtypes = [C.POINTER(BalanceTablesC), C.POINTER(BalanceStateC), C.c_int,
                                        C.c_double, _dp, _dp]
    t, k = balance_tables(bt, elem_z, abund, nstages, lgCompton)
    C1 = J.shape[0]
    s = {"J": np.ascontiguousarray(np.asarray(J, dtype=np.float64).T),
         "Ne": np.array(Ne, dtype=np.float64), "H": np.ascontiguousarray(Hden, dtype=np.float64),
         "X": np.ascontiguousarray(np.asarray(ionDen, dtype=np.float64).transpose(2, 1, 0))}
    st = BalanceStateC(nCells=C1 - 1, J=s["J"].ctypes.data_as(_dp), Ne=s["Ne"].ctypes.data_as(_dp),
                       Hden=s["H"].ctypes.data_as(_dp), ionDen=s["X"].ctypes.data_as(_dp))
    h, c = C.c_double(0.), C.c_double(0.)
    L.oracle_cell_heat_cool(C.byref(t), C.byref(st), int(cell), float(Te), C.byref(h), C.byref(c))
    return h.value, c.value


def cooling_parts(bt, elem_z, abund, nstages, Te, Ne, X):
    """The oracle's thermBalance cooling terms at Te, Ne for ion fractions X
    [31, 32] (Z, stage), no radiation field: (free-free, recombination,
    Ly-alpha/beta + forbidden lines, collisional ionisation) [erg/s per n_H]."""
    L = lib()
    L.oracle_cooling_parts.argtypes = [C.POINTER(BalanceTablesC), C.c_double, C.c_double, _dp, _dp]
    t, k = balance_tables(bt, elem_z, abund, nstages, False)
    x = np.ascontiguousarray(X, dtype=np.float64).reshape(31 * 32)
    out = np.zeros(4)
    L.oracle_cooling_parts(C.byref(t), float(Te), float(Ne), x.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


class EmissionTablesC(C.Structure):
    _fields_ = [
        ("bal", C.POINTER(BalanceTablesC)), ("widFlx", _dp), ("ntk", C.c_int), ("tk", _dp),
        ("nlim", C.c_int * 3), ("gnu", _dp * 3), ("gval", _dp * 3), ("HIPnuP", C.c_int),
        ("HeIPnuP", C.c_int), ("HeIIPnuP", C.c_int), ("xsHI", _dp), ("xsHeI", _dp),
        ("xsHeII", _dp), ("fbLo", _ip), ("fbHi", _ip), ("fbG0", _ip), ("fbG1", _ip),
        ("fbOff", C.POINTER(C.c_int64)), ("fbXs", _dp), ("hei2phot", _dp), ("hlMap", _ip),
        ("hlDens", _dp), ("hlLines", _dp), ("hbA", _dp), ("hbB", _dp), ("rbEdge", _dp),
        ("r2aEdge", _dp), ("heiRec", _dp), ("heiiLyman", _dp),
    ]


class EmissionStateC(C.Structure):
    _fields_ = [
        ("nCells", C.c_int), ("Te", _dp), ("Ne", _dp), ("ionDen", _dp), ("recPDF", _dp),
        ("totalLines", _dp), ("normCont", _dp), ("normRec", _dp), ("normFor", _dp),
        ("dustS", C.c_int), ("dustA", C.c_int), ("dustAbun", _dp), ("dustWeight", _dp),
        ("dustTsub", _dp), ("dustCabs", _dp), ("Tdust", _dp), ("Ndust", _dp), ("Hden", _dp),
        ("nQ", C.c_int), ("nTb", C.c_int), ("qIdx", _ip), ("qT", _dp), ("qP", _dp),
    ]


def emission_tables(et, bt, elem_z, abund, nstages):
    """ctypes view of the oracle's emission tables built from the RAW data of
    `et` (the oracle derives pointers and logarithms itself)."""
    t, k = balance_tables(bt, elem_z, abund, nstages, True)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    fbLo, fbHi, g0, g1 = (np.zeros(31 * 32, dtype=np.int32) for _ in range(4))
    fbOff = np.zeros(31 * 32, dtype=np.int64)
    for n in range(et.fbZ.size):
        idx = int(et.fbZ[n]) * 32 + int(et.fbIon[n])
        fbLo[idx], fbHi[idx] = et.fbLo[n], et.fbHi[n]
        g0[idx], g1[idx], fbOff[idx] = et.fbG0[n], et.fbG1[n], et.fbOff[n]
    g1[g1 == 0] = 1
    k.update(e_wid=f64(et.widFlx), e_tk=f64(et.tk), e_gnu=[f64(a) for a in et.gnu],
             e_gval=[f64(a) for a in et.gval], e_xsHI=f64(et.xsHI),
             e_xsHeI=None if et.xsHeI is None else f64(et.xsHeI),
             e_xsHeII=None if et.xsHeII is None else f64(et.xsHeII),
             e_fbLo=fbLo, e_fbHi=fbHi, e_g0=g0, e_g1=g1, e_off=fbOff, e_fbXs=f64(et.fbXs),
             e_he2=f64(et.hei2phot), e_map=i32(et.hlMap), e_dens=f64(et.hlDens),
             e_lines=f64(et.hlLines), e_hbA=f64(et.hbA), e_hbB=f64(et.hbB),
             e_rb=f64(et.rbEdge), e_r2a=f64(et.r2aEdge), e_heiRec=f64(et.heiRec),
             e_lym=f64(et.heiiLyman), e_bal=t)
    dp = lambda a: a.ctypes.data_as(_dp) if a is not None else None
    ip = lambda a: a.ctypes.data_as(_ip)
    e = EmissionTablesC(
        bal=C.pointer(t), widFlx=dp(k["e_wid"]), ntk=et.tk.size, tk=dp(k["e_tk"]),
        nlim=(C.c_int * 3)(*[a.size for a in et.gnu]), gnu=(_dp * 3)(*[dp(a) for a in k["e_gnu"]]),
        gval=(_dp * 3)(*[dp(a) for a in k["e_gval"]]), HIPnuP=et.HIPnuP, HeIPnuP=et.HeIPnuP,
        HeIIPnuP=et.HeIIPnuP, xsHI=dp(k["e_xsHI"]), xsHeI=dp(k["e_xsHeI"]),
        xsHeII=dp(k["e_xsHeII"]), fbLo=ip(fbLo), fbHi=ip(fbHi), fbG0=ip(g0), fbG1=ip(g1),
        fbOff=fbOff.ctypes.data_as(C.POINTER(C.c_int64)), fbXs=dp(k["e_fbXs"]),
        hei2phot=dp(k["e_he2"]), hlMap=ip(k["e_map"]), hlDens=dp(k["e_dens"]),
        hlLines=dp(k["e_lines"]), hbA=dp(k["e_hbA"]), hbB=dp(k["e_hbB"]), rbEdge=dp(k["e_rb"]),
        r2aEdge=dp(k["e_r2a"]), heiRec=dp(k["e_heiRec"]), heiiLyman=dp(k["e_lym"]))
    return e, k


def emission(et, bt, elem_z, abund, nstages, Te, Ne, ionDen, nthreads=1, dust=None):
    """Oracle emissionDriver over all cells.  ionDen is (C+1, nElemUsed,
    nstages).  Returns recPDF (C+1, B), totalLines and the per-cell sums.
    dust = (DustTables, Tdust (C+1, nA+1, nS+1), Ndust, Hden[, (qIdx, qT, qP)])
    adds the grains of a gas + dust run (the last item: quantum heating)."""
    L = lib()
    L.oracle_emission.argtypes = [C.POINTER(EmissionTablesC), C.POINTER(EmissionStateC), C.c_int]
    L.oracle_emission.restype = C.c_int
    t, k = emission_tables(et, bt, elem_z, abund, nstages)
    C1 = len(Te)
    B = et.nu.size
    s = {"Te": np.ascontiguousarray(Te, dtype=np.float64),
         "Ne": np.ascontiguousarray(Ne, dtype=np.float64),
         "X": np.ascontiguousarray(np.asarray(ionDen, dtype=np.float64).transpose(2, 1, 0)),
         "pdf": np.zeros((C1, B)), "tl": np.zeros(C1), "nc": np.zeros(C1), "nr": np.zeros(C1),
         "nf": np.zeros(C1)}
    dp = lambda n: s[n].ctypes.data_as(_dp)
    st = EmissionStateC(nCells=C1 - 1, Te=dp("Te"), Ne=dp("Ne"), ionDen=dp("X"), recPDF=dp("pdf"),
                        totalLines=dp("tl"), normCont=dp("nc"), normRec=dp("nr"), normFor=dp("nf"))
    if dust is not None:
        dt, Td, Nd, Hd = dust[:4]
        if len(dust) > 4 and dust[4] is not None:
            qi, qT, qP = dust[4]
            s["qi"] = np.ascontiguousarray(qi, dtype=np.int32)
            s["qT"] = np.ascontiguousarray(qT, dtype=np.float64)
            s["qP"] = np.ascontiguousarray(qP, dtype=np.float64)
            st.nQ, st.nTb = int(s["qT"].shape[1]), int(s["qT"].shape[2])
            st.qIdx = s["qi"].ctypes.data_as(_ip)
            st.qT, st.qP = dp("qT"), dp("qP")
        for name, a in (("da", dt.grainAbun), ("dw", dt.grainWeight), ("dt", dt.Tsub),
                        ("dc", dt.Cabs), ("td", Td), ("nd", Nd), ("hd", Hd)):
            s[name] = np.ascontiguousarray(a, dtype=np.float64)
        st.dustS, st.dustA = dt.nSpecies, dt.nSizes
        st.dustAbun, st.dustWeight, st.dustTsub = dp("da"), dp("dw"), dp("dt")
        st.dustCabs, st.Tdust, st.Ndust, st.Hden = dp("dc"), dp("td"), dp("nd"), dp("hd")
    if L.oracle_emission(C.byref(t), C.byref(st), int(nthreads)) != 0:
        raise ValueError("frequency grid too coarse for the free-bound edges")
    return {"recPDF": s["pdf"], "totalLines": s["tl"], "normCont": s["nc"], "normRec": s["nr"],
            "normFor": s["nf"]}


EMISSION_PARTS = ["eHI", "eHeI", "eHeII", "gammaHI", "gammaHeI", "gammaHeII", "gammaHeavies",
                  "ffCoeff1", "ffCoeff2", "twoPhotHI", "twoPhotHeI", "twoPhotHeII"]


def emission_coeffs(et, bt, elem_z, abund, nstages, Te, Ne, X):
    """Per-bin emission coefficients of one cell (X is (nElemUsed, nstages))."""
    L = lib()
    L.oracle_emission_coeffs.argtypes = [C.POINTER(EmissionTablesC), C.c_double, C.c_double, _dp,
                                         _dp]
    L.oracle_emission_coeffs.restype = C.c_int
    t, k = emission_tables(et, bt, elem_z, abund, nstages)
    x = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T)
    B = et.nu.size
    out = np.zeros(12 * B)
    if L.oracle_emission_coeffs(C.byref(t), float(Te), float(Ne), x.ctypes.data_as(_dp),
                                out.ctypes.data_as(_dp)) != 0:
        raise ValueError("frequency grid too coarse for the free-bound edges")
    return {n: out[i * B:(i + 1) * B] for i, n in enumerate(EMISSION_PARTS)}


def hei_line(et, bt, elem_z, abund, nstages, i, Te, Ne):
    """He I recombination line i (order of HeIrecLines.dat) per n_e n(He+)
    [erg cm^3/s] of the oracle's RecLinesEmission."""
    L = lib()
    L.oracle_hei_line.argtypes = [C.POINTER(EmissionTablesC), C.c_int, C.c_double, C.c_double]
    L.oracle_hei_line.restype = C.c_double
    t, k = emission_tables(et, bt, elem_z, abund, nstages)
    return L.oracle_hei_line(C.byref(t), int(i), float(Te), float(Ne))


def nlevel_lines(bt, atom, Te, Ne, ionDenUp, rec, scale=1.0, unit=1.9865e9):
    L = lib()
    L.oracle_nlevel_lines.argtypes = [C.POINTER(BalanceTablesC), C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_double]
    L.oracle_nlevel_lines.restype = C.c_double
    t, k = balance_tables(bt, [1], np.ones(1), 2, True)
    return L.oracle_nlevel_lines(C.byref(t), int(atom), float(Te), float(Ne), float(ionDenUp),
                                 int(rec), float(scale), float(unit))


# ----------------------------------------------------------------------- dust
class DustIn(C.Structure):
    _fields_ = [
        ("nCells", C.c_int), ("B", C.c_int), ("nS", C.c_int), ("nA", C.c_int), ("nT", C.c_int),
        ("dustOnly", C.c_int), ("abun", _dp), ("weight", _dp), ("tsub", _dp), ("Cabs", _dp),
        ("Csca", _dp), ("emInt", _dp), ("nu", _dp), ("widFlx", _dp), ("Ndust", _dp),
        ("XHILimit", C.c_double), ("collD", _dp), ("peD", _dp),
        ("nQ", C.c_int), ("nTb", C.c_int), ("qIdx", _ip), ("qT", _dp), ("qP", _dp),
    ]


def _dust_in(dt, nu, widflx, Ndust, dustOnly, XHILimit=0.05, collD=None, peD=None):
    keep = {}

    def d(name, a):
        keep[name] = np.ascontiguousarray(a, dtype=np.float64)
        return keep[name].ctypes.data_as(_dp)

    inp = DustIn(nCells=int(np.asarray(Ndust).size - 1), B=int(np.asarray(nu).size),
                 nS=dt.nSpecies, nA=dt.nSizes, nT=int(dt.emIntegral.shape[2]),
                 dustOnly=int(bool(dustOnly)), abun=d("a", dt.grainAbun), weight=d("w", dt.grainWeight),
                 tsub=d("t", dt.Tsub), Cabs=d("ca", dt.Cabs), Csca=d("cs", dt.Csca),
                 emInt=d("e", dt.emIntegral), nu=d("nu", nu), widFlx=d("wf", widflx),
                 Ndust=d("nd", Ndust), XHILimit=float(XHILimit))
    if collD is not None:
        inp.collD = d("cd", collD)
        inp.peD = d("pd", peD)
    return inp, keep


def planck_h(nu, T):
    f = lib().oracle_planck_h
    f.restype = C.c_double
    f.argtypes = [C.c_double, C.c_double]
    return np.array([f(float(v), float(T)) for v in np.atleast_1d(nu)])


def dust_em_integral(Cabs_row, nu, widflx, temps):
    c = np.ascontiguousarray(Cabs_row, dtype=np.float64)
    n_ = np.ascontiguousarray(nu, dtype=np.float64)
    w = np.ascontiguousarray(widflx, dtype=np.float64)
    t = np.ascontiguousarray(temps, dtype=np.float64)
    out = np.zeros(t.size)
    lib().oracle_dust_em_integral(c.ctypes.data_as(_dp), n_.ctypes.data_as(_dp),
                                  w.ctypes.data_as(_dp), C.c_int(n_.size), t.ctypes.data_as(_dp),
                                  C.c_int(t.size), out.ctypes.data_as(_dp))
    return out


def dust_opacity(dt, nu, widflx, Ndust, Tdust, gasOpac=None):
    """(opacity, absOpac, scaOpac) as (C+1, B); Tdust (C+1, nA+1, nS+1)."""
    inp, keep = _dust_in(dt, nu, widflx, Ndust, gasOpac is None)
    C1, B = np.asarray(Ndust).size, np.asarray(nu).size
    td = np.ascontiguousarray(Tdust, dtype=np.float64)
    gas = None if gasOpac is None else np.ascontiguousarray(np.asarray(gasOpac).T, dtype=np.float64)
    op, ab, sc = np.zeros(B * C1), np.zeros(B * C1), np.zeros(B * C1)
    if gas is not None:
        op[:] = gas.reshape(-1)
    lib().oracle_dust_opacity(C.byref(inp), td.ctypes.data_as(_dp),
                              None if gas is None else gas.ctypes.data_as(_dp),
                              op.ctypes.data_as(_dp), ab.ctypes.data_as(_dp), sc.ctypes.data_as(_dp))
    f = lambda a: a.reshape(B, C1).T.copy()
    return f(op), f(ab), f(sc)


def _q(inp, keep, qheat):
    """qheat = (qIdx [nS*nA], qT, qP [C+1][nQ][nTb]) into an oracle input."""
    if qheat is None:
        return
    qi, qT, qP = qheat
    keep["qi"] = np.ascontiguousarray(qi, dtype=np.int32)
    keep["qT"] = np.ascontiguousarray(qT, dtype=np.float64)
    keep["qP"] = np.ascontiguousarray(qP, dtype=np.float64)
    inp.nQ, inp.nTb = int(keep["qT"].shape[1]), int(keep["qT"].shape[2])
    inp.qIdx = keep["qi"].ctypes.data_as(_ip)
    inp.qT, inp.qP = keep["qT"].ctypes.data_as(_dp), keep["qP"].ctypes.data_as(_dp)


def dust_pdf(dt, nu, widflx, Ndust, Tdust, qheat=None):
    """setDustPDF rows (C+1, B); qheat = (qIdx, qT, qP) for the grains with
    temperature distributions."""
    inp, keep = _dust_in(dt, nu, widflx, Ndust, True)
    _q(inp, keep, qheat)
    C1, B = np.asarray(Ndust).size, np.asarray(nu).size
    td = np.ascontiguousarray(Tdust, dtype=np.float64)
    pdf = np.zeros(C1 * B)
    lib().oracle_dust_pdf(C.byref(inp), td.ctypes.data_as(_dp), pdf.ctypes.data_as(_dp))
    return pdf.reshape(C1, B)


def dust_temperature(dt, nu, widflx, Ndust, Tdust, J, dustOnly=True, XHILimit=0.05, collD=None,
                     peD=None):
    """getDustT: returns (Tdust, lgConverged, lgBlack, noHit, converged); J (C+1, B).
    collD, peD (nS*nA, C+1): the grains' collision heating and photoelectric
    cooling of a photoelectric gas + dust update."""
    inp, keep = _dust_in(dt, nu, widflx, Ndust, dustOnly, XHILimit, collD, peD)
    C1 = np.asarray(Ndust).size
    td = np.ascontiguousarray(Tdust, dtype=np.float64).copy()
    j = np.ascontiguousarray(np.asarray(J).T, dtype=np.float64)
    conv = np.zeros(C1, dtype=np.int32)
    black = np.zeros(C1, dtype=np.int32)
    cnt = (C.c_double * 2)()
    lib().oracle_dust_temperature(C.byref(inp), j.ctypes.data_as(_dp), td.ctypes.data_as(_dp),
                                  conv.ctypes.data_as(_ip), black.ctypes.data_as(_ip), cnt)
    return td, conv, black, cnt[0], cnt[1]


class QHeatIn(C.Structure):
    _fields_ = [("B", C.c_int), ("nTbins", C.c_int), ("Tmax", C.c_double), ("nu", _dp), ("widFlx", _dp),
                ("J", _dp), ("Cabs", _dp), ("radius", C.c_double), ("kind", C.c_int)]


def qheat(nu, widflx, J, Cabs, radius, kind, tg, nTbins=300, Tmax=700.0):
    """qHeat of one grain in one cell: (temp, pss) over nTbins bins; J (B,) the
    cell's normalised J, Cabs (B,) the grain's cross-section, kind 1 silicate
    or 2 carbon."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (nu, widflx, J, Cabs)]
    B = keep[0].size
    q = QHeatIn(B=B, nTbins=int(nTbins), Tmax=float(Tmax), nu=keep[0].ctypes.data_as(_dp),
                widFlx=keep[1].ctypes.data_as(_dp), J=keep[2].ctypes.data_as(_dp),
                Cabs=keep[3].ctypes.data_as(_dp), radius=float(radius), kind=int(kind))
    L = lib()
    L.oracle_qheat_work.restype = C.c_size_t
    work = np.zeros(int(L.oracle_qheat_work(C.c_int(B), C.c_int(int(nTbins)))))
    temp, pss = np.zeros(nTbins), np.zeros(nTbins)
    L.oracle_qheat(C.byref(q), C.c_double(float(tg)), temp.ctypes.data_as(_dp), pss.ctypes.data_as(_dp),
                   work.ctypes.data_as(_dp))
    return temp, pss


# ---- output reductions (outputGas volume averages, writeSED) ----------------
class OutputIn(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nCells", C.c_int), ("sym", C.c_int),
        ("xAxis", _dp), ("yAxis", _dp), ("zAxis", _dp), ("active", _ip), ("NE", C.c_int),
        ("NS", C.c_int), ("elemZ", _ip), ("Te", _dp), ("Ne", _dp), ("Hden", _dp), ("ionDen", _dp),
        ("lgBlack", _ip), ("plane", C.c_int),
    ]


def output_in(grid, elem_z, Te, Ne, Hden, ionDen, lgBlack=None):
    """(OutputIn, arrays it points into); ionDen (C+1, nElemUsed, nstages)."""
    keep = {}

    def d(name, a, dt=np.float64, ptr=_dp):
        keep[name] = np.ascontiguousarray(a, dtype=dt)
        return keep[name].ctypes.data_as(ptr)

    X = np.asarray(ionDen, dtype=np.float64)
    C1, NE, NS = X.shape
    inp = OutputIn(nx=len(grid.xAxis), ny=len(grid.yAxis), nz=len(grid.zAxis), nCells=C1 - 1,
                   sym=int(bool(grid.lgSymmetricXYZ)), xAxis=d("x", grid.xAxis),
                   yAxis=d("y", grid.yAxis), zAxis=d("z", grid.zAxis),
                   active=d("a", np.asarray(grid.active).reshape(-1, order="F"), np.int32, _ip),
                   NE=NE, NS=NS, elemZ=d("ez", elem_z, np.int32, _ip), Te=d("te", Te),
                   Ne=d("ne", Ne), Hden=d("h", Hden), ionDen=d("X", X.transpose(2, 1, 0)),
                   lgBlack=None if lgBlack is None else d("bl", lgBlack, np.int32, _ip),
                   plane=int(bool(getattr(grid, "lgPlaneIonization", False))))
    return inp, keep


def volume_averages(grid, elem_z, Te, Ne, Hden, ionDen, lgBlack=None):
    """outputGas volume averages: (TeVol, ionDenVol), each (nElemUsed, nstages);
    ionDen (C+1, nElemUsed, nstages)."""
    inp, keep = output_in(grid, elem_z, Te, Ne, Hden, ionDen, lgBlack)
    NE, NS = inp.NE, inp.NS
    T = np.zeros(NS * NE)
    W = np.zeros(NS * NE)
    f = lib().oracle_volume_averages
    f.restype = C.c_int
    rc = f(C.byref(inp), T.ctypes.data_as(_dp), W.ctypes.data_as(_dp))
    assert rc == 0
    return T.reshape(NS, NE).T.copy(), W.reshape(NS, NE).T.copy()


def sed(escaped, sym):
    """writeSED all-angle column: (sed[B], total); escaped (C+1, B)."""
    E = np.ascontiguousarray(np.asarray(escaped, dtype=np.float64).T)
    B, C1 = E.shape
    out = np.zeros(B)
    f = lib().oracle_sed
    f.restype = C.c_double
    f.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp]
    tot = f(E.ctypes.data_as(_dp), C1 - 1, B, int(bool(sym)), out.ctypes.data_as(_dp))
    return out, tot


def line_luminosities(et, bt, elem_z, abund, nstages, grid, Te, Ne, Hden, ionDen, L,
                      lgBlack=None):
    """outputGas line luminosities: (lines, Hbeta) with lines[Z, stage] the
    (L, L) array whose [i, j] is the line from level j+1 to level i+1 summed
    over the reached cells times Hden dV (zero for ions without atomic data)."""
    t, k = emission_tables(et, bt, elem_z, abund, nstages)
    inp, keep = output_in(grid, elem_z, Te, Ne, Hden, ionDen, lgBlack)
    st = EmissionStateC(nCells=inp.nCells, Te=inp.Te, Ne=inp.Ne, ionDen=inp.ionDen)
    out = np.zeros((31 * 32, L, L))
    hb = C.c_double(0.0)
    f = lib().oracle_line_luminosities
    f.argtypes = [C.POINTER(EmissionTablesC), C.POINTER(EmissionStateC), C.POINTER(OutputIn),
                  C.c_int, _dp, C.POINTER(C.c_double)]
    f.restype = C.c_int
    assert f(C.byref(t), C.byref(st), C.byref(inp), int(L), out.ctypes.data_as(_dp), C.byref(hb)) == 0
    return out.reshape(31, 32, L, L), hb.value

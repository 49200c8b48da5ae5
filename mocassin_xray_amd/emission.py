"""Tables of the per-cell diffuse emission (host, FP64).

Restates the one-off set-up behind MOCASSIN's emissionDriver
(emission_mod.f90:36-252):
  initGammaCont        ph_mod.f90:31-202     H I / He I / He II free-bound fits,
                                             their edge pointers in the mesh
  fb_ff                emission_mod.f90:336-618  ground cross-sections above the
                                             H / He edges, outer-shell bin ranges
                                             and g0/g1 of the heavy ions
  HeI2photSub          emission_mod.f90:764-817  the bin -> table-row walk of
                                             the He I two-photon shape
  RecLinesEmission     emission_mod.f90:856-1153 hydrogenic line tables
  setDiffusePDF        emission_mod.f90:1528-1919 He I 2^3S and He II Lyman bins
The data files are the reference's, packed by data/make_atomic.py.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .atomic import XSecStack, locate, outer_shell, packed_data
from .balance import stat_weights


@dataclass
class EmissionTables:
    nu: np.ndarray
    widFlx: np.ndarray
    tk: np.ndarray                 # [ntk] tkGamma
    gnu: list                      # 3 x [nlim] table frequencies (raw)
    gval: list                     # 3 x [ntk, nlim] gamma (raw, linear)
    edge: list                     # 3 x [nlim] int32 1-based bin pointers
    loggam: list                   # 3 x [ntk, nlim] log10 gamma (0 where gamma <= 0)
    HIPnuP: int
    HeIPnuP: int
    HeIIPnuP: int
    xsHI: np.ndarray
    xsHeI: np.ndarray | None
    xsHeII: np.ndarray | None
    fbZ: np.ndarray                # [nFb] heavy-ion free-bound entries, Z ascending
    fbIon: np.ndarray
    fbLo: np.ndarray               # outer-shell bin range [lo, hi), 1-based
    fbHi: np.ndarray
    fbOff: np.ndarray              # offset of each entry's cross-sections in fbXs
    fbXs: np.ndarray
    fbG0: np.ndarray
    fbG1: np.ndarray
    fbGratio: np.ndarray           # real(g0)/real(g1) in single precision, as FP64
    he2j: np.ndarray               # [B] He I two-photon table row per bin, 0 = none
    hei2phot: np.ndarray           # [41, 2]
    hlMap: np.ndarray              # [9, 12]
    hlDens: np.ndarray             # [files, 14]
    hlLines: np.ndarray            # [files, 14, 16, 9]
    hbA: np.ndarray                # [9, 4]
    hbB: np.ndarray
    rbEdge: np.ndarray             # [2, 3, 26]
    r2aEdge: np.ndarray            # [3, 18]
    heiRec: np.ndarray             # [34, 3, 4]
    heiiLyman: np.ndarray          # [4, 2]
    j2TsP: int
    lymP: np.ndarray               # [4] int32, 0 = below the mesh


def edge_pointers(nu: np.ndarray, gnu: np.ndarray) -> np.ndarray:
    """initGammaCont (ph_mod.f90:96-196): bin of each table frequency, moved
    up when it lies above the bin midpoint; the first pointer is 1."""
    B = nu.size
    out = np.zeros(gnu.size, dtype=np.int32)
    for i, g in enumerate(gnu):
        g = float(np.float32(g))
        e = locate(nu, g)
        if 1 <= e < B and g > (nu[e - 1] + nu[e]) / 2.0:
            e += 1
        out[i] = e
    out[0] = 1
    return out


def hei2phot_rows(nu: np.ndarray, yd: np.ndarray) -> np.ndarray:
    """HeI2photSub (emission_mod.f90:800-815): the table row j reached at
    each bin by the walk that steps j at most once per bin; 0 once j hits 41."""
    out = np.zeros(nu.size, dtype=np.int32)
    j = 1
    for i in range(nu.size):
        y = nu[i] / 1.514
        if y > yd[j - 1, 0]:
            j += 1
        if j >= 41:
            break
        out[i] = j
    return out


def build_emission(xsec: XSecStack, widFlx: np.ndarray) -> EmissionTables:
    nu = np.asarray(xsec.nu, dtype=np.float64)
    B = nu.size
    ns = xsec.nstages
    d = packed_data()
    if xsec.lo_clamped:
        raise ValueError("an ionisation threshold lies below the frequency mesh: the diffuse "
                         "emission needs a mesh that starts below every threshold")
    tk = d["gam_tk_HI"].astype(np.float64)
    gnu, gval, edge, loggam = [], [], [], []
    for tag in ("HI", "HeI", "HeII"):
        if not np.array_equal(d[f"gam_tk_{tag}"], tk):
            raise ValueError("gamma tables must share their temperature grid")
        f, v = d[f"gam_nu_{tag}"].astype(np.float64), d[f"gam_val_{tag}"].astype(np.float64)
        e = edge_pointers(nu, f)
        pairs = e.reshape(-1, 2)
        if np.any(pairs[:, 0] == pairs[:, 1]):
            raise ValueError(f"frequency grid too coarse for the {tag} free-bound edges")
        lg = np.zeros_like(v)
        pos = v > 0
        lg[pos] = np.log10(v[pos])
        gnu.append(f), gval.append(v), edge.append(e), loggam.append(lg)
    sh = {(s.elem, s.ion, s.shell): s for s in xsec.shells}
    if (1, 1, 1) not in sh:
        raise ValueError("hydrogen must be switched on")
    sHI = sh[(1, 1, 1)]
    sHeI, sHeII = sh.get((2, 1, 1)), sh.get((2, 2, 1))
    on = {s.elem for s in xsec.shells} | {z for z, _ in xsec.species}
    fb = {k: [] for k in ("z", "ion", "lo", "hi", "off", "g0", "g1")}
    xs_parts, off = [], 0
    for z in range(3, 20):
        if z not in on:
            continue
        for ion in range(1, min(z, ns - 1) + 1):
            nel = z - ion + 1
            s = sh.get((z, ion, outer_shell(z, nel)))
            if s is None:
                continue                      # empty bin range: never emits
            g0, g1 = stat_weights(z, nel)
            for k, v in (("z", z), ("ion", ion), ("lo", s.lo), ("hi", s.hi), ("off", off),
                         ("g0", g0), ("g1", g1)):
                fb[k].append(v)
            xs_parts.append(s.sigma)
            off += s.sigma.size
    g0 = np.array(fb["g0"], dtype=np.int32)
    g1 = np.array(fb["g1"], dtype=np.int32)
    ratio = (g0.astype(np.float32) / g1.astype(np.float32)).astype(np.float64)
    he2 = d["hei2phot"].astype(np.float64)
    lym = d["heiiLyman"].astype(np.float64)
    lymP = np.array([locate(nu, float(np.float32(v))) for v in lym[:, 1]], dtype=np.int32)
    return EmissionTables(
        nu=nu, widFlx=np.asarray(widFlx, dtype=np.float64), tk=tk, gnu=gnu, gval=gval, edge=edge,
        loggam=loggam, HIPnuP=sHI.lo, HeIPnuP=sHeI.lo if sHeI else 0,
        HeIIPnuP=sHeII.lo if sHeII else 0, xsHI=sHI.sigma,
        xsHeI=sHeI.sigma if sHeI else None, xsHeII=sHeII.sigma if sHeII else None,
        fbZ=np.array(fb["z"], dtype=np.int32), fbIon=np.array(fb["ion"], dtype=np.int32),
        fbLo=np.array(fb["lo"], dtype=np.int32), fbHi=np.array(fb["hi"], dtype=np.int32),
        fbOff=np.array(fb["off"], dtype=np.int64),
        fbXs=np.concatenate(xs_parts) if xs_parts else np.zeros(1),
        fbG0=g0, fbG1=g1, fbGratio=ratio, he2j=hei2phot_rows(nu, he2), hei2phot=he2,
        hlMap=d["hl_map"].astype(np.int32), hlDens=d["hl_dens"].astype(np.float64),
        hlLines=d["hl_lines"].astype(np.float64), hbA=d["hbA"].astype(np.float64),
        hbB=d["hbB"].astype(np.float64), rbEdge=d["rbEdge"].astype(np.float64),
        r2aEdge=d["r2aEdge"].astype(np.float64), heiRec=d["heiRec"].astype(np.float64),
        heiiLyman=lym, j2TsP=locate(nu, float(np.float32(1.45673))), lymP=lymP)

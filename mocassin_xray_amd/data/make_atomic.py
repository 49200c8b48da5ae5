"""Pack the atomic-data tables the Monte Carlo iteration needs into atomic.npz.

Reads MOCASSIN's text data files (default: /root/reference/data) and writes
their numbers, unchanged, as arrays.  Each reader follows the read sequence
of the routine that consumes the file:
  ph1, ph2          Verner et al. (1996) shell fits           phInit       ph_mod.f90:523-581
  rrb_*             Badnell radiative recombination fits     initRRDR     hydro_mod.f90:334-417
  drb_c, drb_e      Badnell dielectronic recombination fits  initRRDR
  rrec, rnew        Verner & Ferland / Pequignot RR fits     radRecFit    update_mod.f90:3094-3202
  diel              Nussbaumer & Storey DR fits (rows)       dielectronic update_mod.f90:3204-3305
  aldro             Aldrovandi & Pequignot DR fits (rows)    dielectronic
  cion_cf, cion_th  Voronov (1997) collisional ionisation    makeCollIonData hydro_mod.f90:13-33
  nauger, auger     Auger electron yields                    makeAugerData hydro_mod.f90:35-53
  atom_*            n-level ions (levels, A, Upsilon tables) equilibrium emission_mod.f90:2505-2754
                    for every file named in fileNames.dat     makeElements hydro_mod.f90:679-768
Run once where the data directory exists; the .npz travels with the package.
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

N_TOT = [1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7]
_SPLIT = re.compile(r"[\s,]+")


def toks(line: str):
    return [t for t in _SPLIT.split(line.strip().replace("D", "E").replace("d", "e")) if t]


_FEXP = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))([+-]\d+)$")


def fnum(t: str) -> float:
    """A Fortran list-directed real: also accepts exponents without a letter (2.066-1)."""
    m = _FEXP.match(t)
    if m:
        return float(m.group(1) + "e" + m.group(2))
    return float(t)


def records(path: Path):
    return path.read_text(errors="replace").split("\n")


def numbers(path: Path):
    for line in records(path):
        for tok in toks(line):
            yield float(tok)


def read_ph(data: Path):
    it1 = numbers(data / "ph1.dat")
    ph1 = np.zeros((6, 30, 30, 7))
    for j in range(1, 31):
        for k in range(1, j + 1):
            nt = N_TOT[k - 1]
            if j == k and k > 18:
                nt = 7
            if j == k + 1 and j in (20, 21, 22, 25, 26):
                nt = 7
            for l in range(1, nt + 1):
                for i in range(6):
                    ph1[i, j - 1, k - 1, l - 1] = next(it1)
    it2 = numbers(data / "ph2.dat")
    ph2 = np.zeros((7, 30, 30))
    for k in range(1, 31):
        for j in range(1, k + 1):
            if k not in (15, 17, 19) and (k <= 20 or k == 26):
                for i in range(7):
                    ph2[i, k - 1, j - 1] = next(it2)
    return ph1, ph2


def read_badnell(data: Path):
    rr = np.zeros((6, 31, 31))            # A, B, T0, T1, C, T2 indexed [., Z, N]
    for line in records(data / "rrBadnell.dat")[2:]:
        t = toks(line)
        if len(t) < 6:
            continue
        z, n, m = int(t[0]), int(t[1]), int(t[2])
        vals = [fnum(x) for x in t[4:10]] + [0.0] * (10 - len(t))
        if z <= 30 and m == 1 and n <= 30:
            rr[:, z, n] = vals[:6]
    dc = np.zeros((9, 31, 31))
    de = np.zeros((9, 31, 31))
    lines = records(data / "drBadnell.dat")
    # coefficient block then energy block, each after a header line
    blocks, cur = [], None
    for line in lines:
        t = toks(line)
        if not t:
            continue
        if not re.match(r"^-?\d", t[0]):
            cur = []
            blocks.append(cur)
            continue
        cur.append(t)
    for arr, block in zip((dc, de), (blocks[0] + blocks[1], blocks[2])):
        for t in block:
            z, n, m = int(t[0]), int(t[1]), int(t[2])
            v = [fnum(x) for x in t[4:13]]
            v += [0.0] * (9 - len(v))
            if z <= 30 and m == 1 and n <= 30:
                arr[:, z, n] = v[:9]
    return rr, dc, de


def read_radrec(data: Path):
    it = (toks(l) for l in records(data / "radrec.dat"))
    rrec = np.zeros((2, 31, 31))
    rnew = np.zeros((4, 31, 31))

    def nxt(k):
        while True:
            t = next(it)
            if t:
                return [fnum(x) for x in t[:k]]
    for ion in range(4, 31):
        if ion != 11:
            for elem in range(ion, 31):
                if elem != 26 or ion >= 14:
                    rrec[:, elem, ion] = nxt(2)
    for ion in range(1, 4):
        for elem in range(ion, 31):
            rnew[:, elem, ion] = nxt(4)
    for elem in range(11, 31):
        rnew[:, elem, 11] = nxt(4)
    for ion in range(4, 11):
        for elem in range(max(6, ion), 9):
            rnew[:, elem, ion] = nxt(4)
        rnew[:, 10, ion] = nxt(4)
    for ion in range(12, 27):
        rnew[:, 26, ion] = nxt(4)
    return rrec, rnew


def read_rows(path: Path, ncol: int):
    rows = []
    for line in records(path):
        t = toks(line)
        if len(t) >= ncol:
            rows.append([fnum(x) for x in t[:ncol]])
    return np.array(rows)


def read_cion(data: Path):
    lines = [toks(l) for l in records(data / "cion.dat")]
    lines = [t for t in lines]
    cf = np.zeros((5, 31, 31))
    k = 0
    for elem in range(1, 31):
        for ion in range(elem, 0, -1):
            while not lines[k]:
                k += 1
            cf[:, elem, ion] = [fnum(x) for x in lines[k][:5]]
            k += 1
    k += 2                                   # two comment records
    th = np.zeros((31, 32))
    for elem in range(3, 31):
        vals = [fnum(x) for x in lines[k][:elem]]
        th[elem, 1:elem + 1] = vals
        k += 1
    return cf, th


def read_auger(data: Path):
    lines = [toks(l) for l in records(data / "auger.dat")]
    lines = [t for t in lines if t]
    nauger = np.zeros((31, 32, 11), dtype=np.int32)
    auger = np.zeros((31, 32, 11, 11))
    k = 0
    for _ in range(1696):
        elem, ion, shell, nelec = (int(x) for x in lines[k][:4])
        nauger[elem, ion, shell] = nelec
        auger[elem, ion, shell, 1:11] = [fnum(x) for x in lines[k + 1][:10]]
        k += 2
    return nauger, auger


def read_atom(path: Path):
    """One n-level ion file in the sequence equilibrium() reads it."""
    rec = records(path)
    k = 0

    def line():
        nonlocal k
        r = rec[k]
        k += 1
        return r
    num_lines = int(toks(line())[0])
    for _ in range(num_lines):
        line()
    t = toks(line())
    nlev, ntemp = int(t[0]), int(t[1])
    for _ in range(nlev):
        line()
    logt = np.array([np.log10(fnum(toks(line())[0])) for _ in range(ntemp)])
    line()                                    # iRats
    qom = np.zeros((ntemp, nlev + 1, nlev + 1))
    ilow1 = iup1 = 0
    kt = 0
    while True:
        t = toks(line())
        il, iu, qx = int(t[0]), int(t[1]), fnum(t[2])
        if qx == 0.0:
            break
        if il == 0:
            il = ilow1
            kt += 1
        else:
            ilow1 = il
            kt = 1
        if iu == 0:
            iu = iup1
        else:
            iup1 = iu
        qom[kt - 1, il, iu] = qx
    A = np.zeros((nlev + 1, nlev + 1))
    atp, iup2p, jj = 0.0, -1, 0
    for _ in range(nlev * (nlev - 1) // 2):
        t = toks(line())
        i, j, ax = int(t[0]), int(t[1]), fnum(t[2])
        if i < 0 and j < 0:
            atp = ax
            iup2p = abs(j)
            i, j = abs(i), abs(j)
            if A[j, i] > 0.0:
                jj += 1
            A[j, i] += atp
        else:
            A[j, i] = ax
    for _ in range(jj):
        t = toks(line())
        A[int(t[1]), int(t[0])] = fnum(t[2])
    g = np.zeros(nlev + 1)
    e = np.zeros(nlev + 1)
    for _ in range(nlev):
        t = toks(line())
        i = int(t[0])
        g[i] = float(t[1])
        e[i] = fnum(t[2])
        if i > 1 and e[i] == e[i - 1]:
            e[i] = e[i] + 0.1
    return nlev, ntemp, logt, qom, A, g, e, atp, iup2p, k


class ListReader:
    """Fortran list-directed input over a file's records: each read starts on
    a new record, spans as many records as it needs (blank ones skipped) and
    drops what is left of its last record."""

    def __init__(self, path: Path):
        self.rec = records(path)
        self.k = 0

    def eof(self) -> bool:
        j = self.k
        while j < len(self.rec) and not toks(self.rec[j].split("!")[0]):
            j += 1
        return j >= len(self.rec)

    def read(self, n: int):
        out = []
        while len(out) < n:
            if self.k >= len(self.rec):
                raise EOFError
            t = toks(self.rec[self.k].split("!")[0])
            self.k += 1
            out.extend(t)
        return out[:n]

    def line(self) -> str:
        r = self.rec[self.k]
        self.k += 1
        return r


REC_SPECIAL = {"ci.dat", "ciii.dat", "ni.dat", "neiv.dat", "oii.dat", "oiii.dat", "nii.dat",
               "nev.dat", "niv.dat"}


def read_rec_data(path: Path, nlev: int, lr: "ListReader"):
    """Recombination data after the level block (equilibrium rec=.true.,
    emission_mod.f90:2828-2877): returns (kind, table[nlev+1, 11], valid[nlev+1])."""
    tab = np.zeros((nlev + 1, 11))
    valid = np.zeros(nlev + 1, dtype=np.int32)
    special = path.name in REC_SPECIAL
    kind = 1 if special else 2
    any_read = False
    for j in range(2, nlev + 1):
        if lr.eof():
            break
        if special:
            v = [fnum(x) for x in lr.read(11)]
            any_read = True
            if v[1] > 0.0:
                tab[j] = v
                valid[j] = 1
        else:
            save = lr.k
            t = lr.read(1)
            jj = int(float(t[0]))
            if jj == -9999:
                break
            lr.k = save
            try:
                v = lr.read(5)
            except EOFError:        # a truncated last record ends the table
                break
            jj = int(float(v[0]))
            if jj > nlev:
                break
            any_read = True
            tab[jj, :4] = [fnum(x) for x in v[1:5]]
            valid[jj] = 1
    if any_read:
        valid[1] = 1
    return (kind if any_read else 0), tab, valid


def read_atoms(data: Path):
    names = [l.strip() for l in records(data / "fileNames.dat")]
    idx = np.full((31, 32), -1, dtype=np.int32)
    meta, logt, g, e, A, qom = [], [], [], [], [], []
    rkind, rtab, rvalid = [], [], []
    k = 0
    for elem in range(3, 31):
        for ion in range(1, min(elem + 1, 31) + 1):
            name = names[k]
            k += 1
            p = data / Path(name).name
            if not name or not p.exists():
                continue
            nlev, ntemp, lt, q, a, gg, ee, atp, iup2p, kend = read_atom(p)
            lr = ListReader(p)
            lr.k = kend
            kind, tab, valid = read_rec_data(p, nlev, lr)
            rkind.append(kind)
            rtab.extend(tab[1:].reshape(-1))
            rvalid.extend(valid[1:])
            idx[elem, ion] = len(meta)
            meta.append([nlev, ntemp, len(logt), len(g), len(A), len(qom), iup2p])
            logt.extend(lt)
            g.extend(gg[1:])
            e.extend(ee[1:])
            A.extend(a[1:, 1:].reshape(-1))
            qom.extend(q[:, 1:, 1:].reshape(-1))
            meta[-1].append(atp)
    meta_i = np.array([m[:7] for m in meta], dtype=np.int64)
    atp = np.array([m[7] for m in meta])
    return (idx, meta_i, atp, np.array(logt), np.array(g), np.array(e), np.array(A),
            np.array(qom, dtype=np.float32), np.array(rkind, dtype=np.int32), np.array(rtab),
            np.array(rvalid, dtype=np.int32))


def read_emission_tables(data: Path) -> dict:
    """Tables of the diffuse emission (emissionDriver and its callers)."""
    out = {}
    # H, He I, He II free-bound fits (initGammaCont ph_mod.f90:31-202)
    for tag, fname in (("HI", "gammaHI.dat"), ("HeI", "gammaHeI.dat"), ("HeII", "gammaHeII.dat")):
        lr = ListReader(data / fname)
        ntk, nlim = (int(float(x)) for x in lr.read(2))
        tk = np.array([fnum(x) for x in lr.read(ntk)])
        nu = np.zeros(nlim)
        val = np.zeros((ntk, nlim))
        for i in range(nlim):
            v = [fnum(x) for x in lr.read(ntk + 1)]
            nu[i] = v[0]
            val[:, i] = v[1:]
        out[f"gam_tk_{tag}"] = tk
        out[f"gam_nu_{tag}"] = nu
        out[f"gam_val_{tag}"] = val
    lr = ListReader(data / "HeI2phot.dat")
    he2 = np.array([[fnum(x) for x in lr.read(2)] for _ in range(41)])
    out["hei2phot"] = he2
    # hydrogenic recombination-line tables (RecLinesEmission emission_mod.f90:856-1153)
    names = [l.strip().strip("'\"").strip() for l in records(data / "hydroLinesFiles.dat")
             if l.strip()][:9 * 12]
    uniq = sorted(set(names))
    fmap = np.array([uniq.index(n) for n in names], dtype=np.int32).reshape(9, 12)
    dens = np.zeros((len(uniq), 14))
    lines = np.zeros((len(uniq), 14, 16, 9))
    for f, name in enumerate(uniq):
        lr = ListReader(data / Path(name).name)
        for iden in range(14):
            if lr.eof():
                break
            dens[f, iden] = fnum(lr.read(1)[0])
            for iup in range(15, 1, -1):
                v = lr.read(min(8, iup - 1))
                lines[f, iden, iup, 1:1 + len(v)] = [fnum(x) for x in v]
    out["hl_map"] = fmap
    out["hl_dens"] = dens
    out["hl_lines"] = lines
    lr = ListReader(data / "gaussfitHb.dat")
    hb = np.array([[fnum(x) for x in lr.read(8)] for _ in range(9)])
    out["hbA"] = hb[:, :4]
    out["hbB"] = hb[:, 4:8]
    rb = np.zeros((2, 3, 26))
    for z, fname in ((0, "r1bEdge.dat"), (1, "r2bEdge.dat")):
        lr = ListReader(data / fname)
        for i in range(26):
            rb[z, :, i] = [fnum(x) for x in lr.read(3)]
    out["rbEdge"] = rb
    lr = ListReader(data / "r2aEdge.dat")
    out["r2aEdge"] = np.array([[fnum(x) for x in lr.read(3)] for _ in range(18)]).T.copy()
    lr = ListReader(data / "HeIrecLines.dat")
    he = np.zeros((34, 3, 4))
    for iden in range(3):
        for il in range(34):
            v = [fnum(x) for x in lr.read(4)]
            he[il, iden] = v
            he[il, iden, 0] *= 1.e25
    out["heiRec"] = he
    lr = ListReader(data / "r2a0100old.dat")
    out["heiiLyman"] = np.array([[fnum(x) for x in lr.read(2)] for _ in range(4)])
    return out


def main(data_dir: str = "/root/reference/data"):
    data = Path(data_dir)
    ph1, ph2 = read_ph(data)
    rr, dc, de = read_badnell(data)
    rrec, rnew = read_radrec(data)
    diel = read_rows(data / "dielectronic.dat", 8)
    aldro = read_rows(data / "aldrovandi.dat", 6)
    cf, th = read_cion(data)
    nauger, auger = read_auger(data)
    aidx, ameta, aatp, alogt, ag, ae, aA, aq, rk, rt, rv = read_atoms(data)
    em = read_emission_tables(data)
    out = Path(__file__).resolve().parent / "atomic.npz"
    np.savez_compressed(out, ph1=ph1, ph2=ph2, rrb=rr, drb_c=dc, drb_e=de, rrec=rrec, rnew=rnew,
                        diel=diel, aldro=aldro, cion_cf=cf, cion_th=th, nauger=nauger,
                        auger=auger, atom_idx=aidx, atom_meta=ameta, atom_atp=aatp,
                        atom_logt=alogt, atom_g=ag, atom_e=ae, atom_A=aA, atom_qom=aq,
                        atom_rec_kind=rk, atom_rec=rt, atom_rec_valid=rv, **em)
    print(out, out.stat().st_size)


if __name__ == "__main__":
    main(*sys.argv[1:])

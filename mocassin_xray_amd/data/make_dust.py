"""Pack the grain optical constants the dust tables need into dust.npz.

Reads MOCASSIN's n,k files (default directory: /root/reference/dustData) and
writes their numbers, unchanged, as arrays.  The read sequence is that of
makeDustXsec for an 'nk' file (ph_mod.f90:1108-1213):
  line 1   file type 'nk'
  line 2   label, sublimation temperature [K], bulk density [g/cm^3],
           work function grainVn [Ryd], surface atom mass MsurfAtom
  then     wavelength [um], Re(m), Im(m) per line, wavelength ascending
For species <label> the archive holds <label>_wav, <label>_n, <label>_k,
<label>_hdr = [Tsub, rho, Vn, Msurf] and <label>_name, the file's own label
(grainLabel; its first letter is the species class qHeat goes by, 'S'
silicate and 'C' carbon, emission_mod.f90:2070-2096).  Also packed: nuDustRyd.dat, the dust
frequency points that merge into the mesh of gas+dust runs
(grid_mod.f90:341-362), as nu_dust_ryd.
Run once where the data directory exists; the .npz travels with the package.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

FILES = ["sil-dlee.nk", "grph1-dl.nk", "grph2-dl.nk", "amC-hann.nk", "SiC-peg.nk", "glSil.nk",
         "crSil.nk", "OssOdef.nk", "OssOrich.nk"]


def read_nk(path: Path):
    lines = [ln for ln in path.read_text().splitlines() if ln.strip()]
    if lines[0].split()[0] != "nk":
        raise ValueError(f"{path}: not an nk file")
    hdr = lines[1].split()
    label = hdr[0]
    vals = [float(v) for v in hdr[1:5]]
    rows = []
    for ln in lines[2:]:
        t = ln.split()
        if len(t) < 3:
            break
        rows.append([float(t[0]), float(t[1]), float(t[2])])
    a = np.array(rows)
    return label, np.array(vals), a[:, 0], a[:, 1], a[:, 2]


def main(src: str = "/root/reference/dustData", out: str | None = None):
    src = Path(src)
    dst = Path(out) if out else Path(__file__).with_name("dust.npz")
    arrays = {}
    labels = []
    for f in FILES:
        label, hdr, wav, n, k = read_nk(src / f)
        key = f.rsplit(".", 1)[0]
        labels.append(key)
        arrays[f"{key}_hdr"] = hdr
        arrays[f"{key}_wav"] = wav
        arrays[f"{key}_n"] = n
        arrays[f"{key}_k"] = k
        arrays[f"{key}_name"] = np.array(label)
    nu = [float(ln.split()[0]) for ln in (src / "nuDustRyd.dat").read_text().splitlines() if ln.strip()]
    arrays["nu_dust_ryd"] = np.array(nu)
    arrays["species"] = np.array(labels)
    np.savez_compressed(dst, **arrays)
    print(f"wrote {dst} ({len(labels)} species, {len(nu)} dust mesh points)")


if __name__ == "__main__":
    main(*sys.argv[1:])

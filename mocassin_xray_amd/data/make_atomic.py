"""Pack the atomic-data tables the Monte Carlo iteration needs into atomic.npz.

Reads MOCASSIN's text data files (default: /root/reference/data) and writes
their numbers, unchanged, as arrays:
  ph1[6, 30, 30, 7]  Verner et al. (1996) shell fits   (read as in phInit, ph_mod.f90:523-581)
  ph2[7, 30, 30]     Verner et al. (1996) outer-shell fits
Run once where the data directory exists; the .npz travels with the package.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

N_TOT = [1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7]


def numbers(path: Path):
    for line in path.read_text().split("\n"):
        for tok in line.replace("D", "E").split():
            yield float(tok)


def read_ph(data: Path):
    it1 = numbers(data / "ph1.dat")
    ph1 = np.zeros((6, 30, 30, 7))
    for j in range(1, 31):               # Z
        for k in range(1, j + 1):        # number of electrons
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
    for k in range(1, 31):               # Z
        for j in range(1, k + 1):        # number of electrons
            if k not in (15, 17, 19) and (k <= 20 or k == 26):
                for i in range(7):
                    ph2[i, k - 1, j - 1] = next(it2)
    return ph1, ph2


def main(data_dir: str = "/root/reference/data"):
    data = Path(data_dir)
    ph1, ph2 = read_ph(data)
    out = Path(__file__).resolve().parent / "atomic.npz"
    np.savez_compressed(out, ph1=ph1, ph2=ph2)
    print(out, out.stat().st_size)


if __name__ == "__main__":
    main(*sys.argv[1:])

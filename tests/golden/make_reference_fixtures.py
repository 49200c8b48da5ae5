"""Writes reference_data.json: small fixtures cut from the reference's data
files (rwesson/mocassin_xray data/), for the CPU tests of the oracle and of
the packed atomic tables.  Run in a tree that holds the reference:

    python tests/golden/make_reference_fixtures.py /path/to/reference

Fixtures (data only):
  hummer94_betaFF, hummer94_betaRec  data/betaFF.dat, data/betaRec.dat: the
      free-free and recombination energy-loss coefficients of hydrogen,
      (log10 T, beta) per row (Hummer 1994, MNRAS 268, 109, Table 1), that the
      polynomial fits of thermBalance (update_mod.f90:933-951) approximate;
  dielectronic_rows   data/dielectronic.dat, every row (elem n a b c d f g);
  aldrovandi_rows     data/aldrovandi.dat, the first 20 and last 5 rows;
  heirec_lines        data/HeIrecLines.dat, the rows of 4471, 5876, 6678 A in
      each of the three density blocks: (row index, wavelength, c0, c1, c2).
"""
import json
import sys
from pathlib import Path


def rows(path, n=None):
    out = []
    for line in Path(path).read_text().splitlines():
        t = line.replace(",", " ").split()
        if not t:
            continue
        out.append([float(v) for v in (t if n is None else t[:n])])
    return out


def main(ref: Path) -> None:
    d = ref / "data"
    he = rows(d / "HeIrecLines.dat")
    blocks = [he[k * 34:(k + 1) * 34] for k in range(3)]
    sel = []
    for k, blk in enumerate(blocks):
        for i, r in enumerate(blk):
            if int(r[3]) in (4473, 5877, 6680) and not (i == 0 and int(r[3]) == 4473):
                sel.append([k, i, int(r[3])] + r[:3])
    aldro = rows(d / "aldrovandi.dat")
    out = {
        "hummer94_betaFF": rows(d / "betaFF.dat"),
        "hummer94_betaRec": rows(d / "betaRec.dat"),
        "dielectronic_rows": rows(d / "dielectronic.dat"),
        "aldrovandi_rows": aldro[:20] + aldro[-5:],
        "aldrovandi_tail_start": len(aldro) - 5,
        "heirec_lines": sel,
    }
    dst = Path(__file__).resolve().parent / "reference_data.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(dst)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))

"""The outputGas text files (output_mod.f90:810-1160) from the device
reductions: the element/stage rows of temperature.out and ionratio.out, and
lineFlux.out's Hbeta scaling, forbidden-line rows, wavelengths and indices."""
import numpy as np
import pytest

from mocassin_xray_amd import checkpoint as K
from mocassin_xray_amd import nebula, outputs


def model():
    return nebula.gas_model("out4", 4, 1.e18, 1000., nebula.NEBULA_ELEMS,
                            10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                            10 ** 1.5, 1000, maxLev=27)


def test_temperature_and_ionratio_rows(tmp_path):
    elems, ns = [1, 2, 8], 7
    T = np.arange(1.0, 1.0 + 3 * ns).reshape(3, ns) * 1000.
    X = np.linspace(0.01, 0.99, 3 * ns).reshape(3, ns)
    for fn, vol, name in ((outputs.write_temperature, T, "temperature.out"),
                          (outputs.write_ionratio, X, "ionratio.out")):
        p = fn(tmp_path, vol, elems, ns)
        assert p.name == name
        rows = [l.split() for l in p.read_text().splitlines()]
        data = [r for r in rows if len(r) == 3 and r[0].isdigit() and r[1].isdigit()]
        assert len(data) == 2 * 30 * ns          # the total and component 1
        got = {(int(a), int(b)): float(c) for a, b, c in data[:30 * ns]}
        for e, z in enumerate(elems):
            for ion in range(1, ns + 1):
                assert got[(z, ion)] == pytest.approx(vol[e, ion - 1], rel=1e-8)
        assert got[(6, 3)] == 0.0                # carbon is off


def test_lineflux_wavelengths_ratios_and_indices(tmp_path):
    m = model()
    L = 27
    ions = [(8, 2), (8, 3)]
    lines = np.zeros((2, L, L))
    lines[1, 2, 3] = 3.0                         # [O III] 1D2 -> 3P2
    lines[1, 1, 3] = 1.0                         # [O III] 1D2 -> 3P1
    lines[0, 0, 1] = 0.5                         # [O II] 2D5/2 -> 4S3/2
    hb = 2.0
    p = outputs.write_lineflux(tmp_path, lines, ions, hb, m.bt, symmetric=True, nElemOn=7)
    text = p.read_text().splitlines()
    hbl = [l for l in text if "Hbeta [E36" in l][0]
    assert float(hbl.split()[-1]) == pytest.approx(hb * 1e-16 * 8.0, rel=1e-7)
    rows = outputs.lineflux_rows(lines, ions, hb, m.bt, nElemOn=7)
    assert [(r[0], r[1], r[2], r[3]) for r in rows] == [(8, 2, 1, 2), (8, 3, 2, 4), (8, 3, 3, 4)]
    wav = {(r[1], r[2], r[3]): r[4] for r in rows}
    # vacuum wavelengths of [O III] 4960/5008 and [O II] 3729 from the level energies
    assert wav[(3, 3, 4)] == pytest.approx(5008.2, abs=1.0)
    assert wav[(3, 2, 4)] == pytest.approx(4960.3, abs=1.0)
    assert wav[(2, 1, 2)] == pytest.approx(3729.9, abs=1.5)
    assert [r[5] for r in rows] == [0.25, 0.5, 1.5]
    # index: after 7 x 84 hydrogenic and 34 He I lines, 900 slots per ion,
    # row-major over (lower, upper)
    f = 7 * 84 + 34 + 1
    assert outputs.first_forbidden_index(7) == f
    assert [r[6] for r in rows] == [f + 1, f + 900 + 1 * 30 + 3, f + 900 + 2 * 30 + 3]
    body = [l for l in text if l.startswith("    8")]
    assert len(body) == 2 * 3 and body[0].split()[:4] == ["8", "2", "1", "2"]


@pytest.mark.gpu
def test_output_gas_files_from_the_device(tmp_path):
    from mocassin_xray_amd import MocGPU
    m = model()
    print(f"case output files: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.iterateMC(1, 3)
        paths = outputs.write_output_gas(tmp_path, g, m)
        TeVol, XVol = g.volume_averages()
        lines, ions, hb = g.line_luminosities()
    assert all(p.exists() for p in paths.values())
    rows = outputs.lineflux_rows(lines, ions, hb, m.bt)
    o3 = [r for r in rows if (r[0], r[1], r[2], r[3]) == (8, 3, 3, 4)]
    assert len(o3) == 1 and o3[0][5] > 0.0
    R = K.ListReader(paths["temperature"])
    for _ in range(6):
        R.skip()
    z, ion, t = R.read("i", "i", "f")
    assert (z, ion) == (1, 1) and t == pytest.approx(TeVol[m.elems.index(1), 0], rel=1e-7)

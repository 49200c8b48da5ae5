"""Checkpoint files in the reference's layout (writeGrid / resetGrid,
grid_mod.f90:2646-2883, 3051-3690) and the warm start (mocassinWarm.f90):
list-directed record rules, an exact round trip of every field and
parameter, a file in the reference's own single-precision list-directed
style, and on the GPU a warm start that continues the run it was written
from."""
import dataclasses

import numpy as np
import pytest

from mocassin_xray_amd import checkpoint as K
from mocassin_xray_amd.mocgpu import CartesianGrid


def test_list_directed_records(tmp_path):
    f = tmp_path / "r.txt"
    f.write_text(" 3 4 5 9 9\n"                 # extra values on the record are dropped
                 " 1.5d0, 2*7.25\n"              # d exponent, comma, repeat count
                 "   'black body' \"x y\" bare\n"
                 " T .FALSE. t\n"
                 " 1 2\n 3\n"                    # a read continues over records
                 "\n 8\n")                       # an empty record is passed over
    R = K.ListReader(f)
    assert R.read("i", "i", "i") == [3, 4, 5]
    assert R.read("f", "f", "f") == [1.5, 7.25, 7.25]
    assert R.read("s", "s", "s") == ["black body", "x y", "bare"]
    assert R.read("l", "l", "l") == [True, False, True]
    assert R.read("i", "i", "i") == [1, 2, 3]
    assert R.read("i") == [8]
    with pytest.raises(EOFError):
        R.read("i")


def synthetic(seed=3, dust=True):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 4, 3, 5
    act = np.zeros((nx, ny, nz), dtype=np.int32)
    on = rng.random((nx, ny, nz)) < 0.8
    n = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if on[i, j, k]:
                    n += 1
                    act[i, j, k] = n
    grid = CartesianGrid(np.sort(rng.random(nx)) * 1e18, np.sort(rng.random(ny)) * 1e18,
                         np.sort(rng.random(nz)) * 1e18, act, n, True, 3.3e18)
    elems = [8, 1, 2, 6]                       # not in Z order: the files are
    ns = 7
    C1 = n + 1
    X = rng.random((C1, len(elems), ns))
    for e, z in enumerate(elems):
        X[:, e, min(z + 1, ns):] = 0.0
    X[0] = 0.0
    st = {"Te": np.r_[0.0, rng.uniform(5e3, 2e4, n)], "Ne": np.r_[0.0, rng.uniform(1, 1e4, n)],
          "Hden": np.r_[0.0, rng.uniform(1, 1e4, n)], "ionDen": X,
          "lgConverged": np.r_[0, rng.integers(0, 2, n)].astype(np.int32),
          "lgBlack": np.r_[0, rng.integers(0, 2, n)].astype(np.int32)}
    p = K.RunParams(nbins=321, nstages=ns, lgSymmetricXYZ=True, convWriteGrid=12.5,
                    lgAutoPackets=True, convIncPercent=0.3, nPhotIncrease=2.0, maxPhotons=10**8,
                    nuMax=100.0, R_in=1.25e17, maxIterateMC=12, minConvergence=97.5,
                    abundanceFiles=["abunA.in", "abun B.in"], dxSlit=1e17, dySlit=2e17,
                    lgDust=dust, dustFile=("grains.dat", "sizes.dat"), nSpecies=2, nSizes=3,
                    viewPointTheta=[0.0, 45.0], viewPointPhi=[0.0, 90.0], contCube=(1.0, 2.0),
                    Ldiffuse=3.5, Tdiffuse=1e5, shapeDiffuse="blackbody", nPhotonsDiffuse=1000,
                    lgNahar=True, lgBadnell=False, pwlMin=0.1, pwlMax=100.0)
    srcs = [K.PhotoSource("blackbody", 1e5, 31.62, 10**7, (0.0, 0.0, 0.0)),
            K.PhotoSource("powerlaw", 0.0, 2.5, 5 * 10**6, (grid.xAxis[1], grid.yAxis[2], 0.0),
                          "pl1", 10.0, -1.5)]
    dd = None
    if dust:
        dd = {"Ndust": np.r_[0.0, rng.uniform(1e-12, 1e-10, n)],
              "Tdust": np.concatenate([np.zeros((1, 4, 3)), rng.uniform(20, 300, (n, 4, 3))])}
    return K.Checkpoint(grid, st, p, srcs, dd, 0), elems


def test_round_trip_is_exact(tmp_path):
    ck, elems = synthetic()
    K.write_grid(tmp_path, ck, elems, totalDustMass=1.5e-3)
    assert {f.name for f in tmp_path.iterdir()} == set(K.FILES)
    back = K.read_grid(tmp_path, elems)
    g, h = ck.grid, back.grid
    assert np.array_equal(g.active, h.active) and g.nCells == h.nCells
    for a in ("xAxis", "yAxis", "zAxis"):
        assert np.array_equal(getattr(g, a), getattr(h, a))
    assert h.R_out == g.R_out and h.lgSymmetricXYZ
    for k, v in ck.state.items():
        assert np.array_equal(v, back.state[k]), k
    for k, v in ck.dust.items():
        assert np.array_equal(v, back.dust[k]), k
    assert back.params == ck.params
    for a, b in zip(ck.sources, back.sources):
        assert (a.contShape, a.TStellar, a.LStar, a.nPhotons, a.spID, a.tStep, a.pwlIndex) == \
               (b.contShape, b.TStellar, b.LStar, b.nPhotons, b.spID, b.tStep, b.pwlIndex)
        np.testing.assert_allclose(b.position, a.position, rtol=1e-15)


def test_grid2_records_follow_the_atomic_number_order(tmp_path):
    ck, elems = synthetic(dust=False)
    K.write_grid(tmp_path, ck, elems)
    first_cell = (tmp_path / "grid2.out").read_text().splitlines()[:4]
    # H (2 stages), He (3), C (7 = min(Z+1, nstages)), O (7)
    assert [len(r.split()) for r in first_cell] == [2, 3, 7, 7]
    c = int(ck.grid.active[ck.grid.active > 0].min())
    assert float(first_cell[0].split()[0]) == ck.state["ionDen"][c, elems.index(1), 0]


def test_reads_the_reference_list_directed_style(tmp_path):
    """Single-precision values as the reference's list-directed writes lay
    them out, with an ion-fraction record wrapped over two lines."""
    (tmp_path / "grid0.out").write_text(
        "           1\n           2           1           1           2           0"
        "   1.00000002E+18\n   0.00000000    \n   5.00000000E+17\n   0.00000000    \n"
        "   0.00000000    \n           1           1           1           1           1"
        "           0\n           2           1           1           2           0"
        "           1\n")
    (tmp_path / "grid1.out").write_text("   10000.0000       1000.00000       1000.00000    \n"
                                        "   8123.45020       950.000000       1000.00000    \n")
    (tmp_path / "grid2.out").write_text(
        "   1.00000005E-03  0.999000013    \n   1.00000005E-03  0.499500006      0.499500006    \n"
        "   2.00000009E-03  0.998000026    \n   2.00000009E-03  0.499000013    \n"
        "  0.499000013    \n")
    (tmp_path / "photoSource.out").write_text(
        "           1  number of photon sources\n 'blackbody'   100000.000       31.6200008"
        "         100000   0.00000000       0.00000000       0.00000000     none"
        "   0.00000000    \n (contShape, T_eff[K], L_* [E36 erg/s], nPackets, (x,y,z) position, "
        "spID, tstep)\n")
    K.write_grid(tmp_path / "p", K.Checkpoint(
        CartesianGrid(np.zeros(1), np.zeros(1), np.zeros(1), np.zeros((1, 1, 1), np.int32), 0,
                      False), {}, K.RunParams(nstages=3, lgGas=True), [], None), [1, 2])
    (tmp_path / "grid3.out").write_text((tmp_path / "p" / "grid3.out").read_text()
                                        .replace(" T lgGas", "           T  lgGas"))
    ck = K.read_grid(tmp_path, [1, 2])
    assert ck.grid.nCells == 2 and ck.grid.active[:, 0, 0].tolist() == [1, 2]
    assert ck.grid.xAxis.tolist() == [0.0, 5.00000000e17]
    assert ck.state["Te"][1:].tolist() == [10000.0, 8123.4502]
    assert ck.state["ionDen"][2, 1].tolist() == [2.00000009e-03, 0.499000013, 0.499000013]
    assert ck.state["lgConverged"][1] == 1 and ck.state["lgBlack"][2] == 1
    assert ck.sources[0].contShape == "blackbody" and ck.sources[0].nPhotons == 100000


def test_writer_refuses_what_the_reader_could_not_split(tmp_path):
    ck, elems = synthetic(dust=False)
    ck.sources[0].spID = "two words"
    with pytest.raises(ValueError):
        K.write_grid(tmp_path, ck, elems)


@pytest.mark.gpu
def test_warm_start_continues_the_run(tmp_path):
    """Three iterations, the checkpoint written and read back, a fresh context
    warm-started from it: the state it holds equals the run's exactly, and the
    next iteration (MoCaSSiNwarm restarts at nIterateMC = 1, every cell
    recomputed) draws the same packets and lands on the same state as the
    run continued with the same iteration number."""
    from mocassin_xray_amd import MocGPU, nebula
    m = nebula.gas_model("warm6", 6, 1.e18, 1000., nebula.NEBULA_ELEMS,
                         10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                         10 ** 1.5, 20000, maxLev=27)
    print(f"case warm-start nebula6: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as ga, MocGPU(m.tables, "fp64") as gb:
        m.load(ga)
        for it in (1, 2, 3):
            ga.iterateMC(it, 61)
        ck = K.checkpoint_of(m, ga, TStellar=1e5)
        K.write_grid(tmp_path, ck, m.elems)
        back = K.read_grid(tmp_path, m.elems)
        for k in ("Te", "Ne", "Hden", "ionDen", "lgConverged", "lgBlack"):
            assert np.array_equal(back.state[k], ck.state[k]), k
        m.load(gb)
        K.warm_start(m, back, gb)
        for k in ("Te", "Ne", "Hden", "ionDen"):
            assert np.array_equal(gb.download(k), ga.download(k)), k
        sa = ga.iterateMC(1, 62)
        sb = gb.iterateMC(1, 62)
        for k in ("segments", "emissions", "escaped"):
            assert sa[k] == sb[k], k
        for k in ("opacity", "recPDF", "totalLines"):   # computed before the packets
            assert np.array_equal(ga.download(k), gb.download(k)), k
        for k in ("Te", "Ne", "ionDen", "Jste"):
            np.testing.assert_allclose(gb.download(k), ga.download(k), rtol=1e-9, atol=1e-30,
                                       err_msg=k)

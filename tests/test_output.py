"""Output reductions: the volume averages of outputGas (Harrington et al.
1982) and the SED of writeSED.  The oracle is pinned on the CPU by what the
definitions fix; the GPU reductions are compared with it."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T


def small_state(seed=3):
    m = nebula.gas_model("neb6", 6, 1.e18, 1000., nebula.NEBULA_ELEMS,
                         10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 200, 100.0, 1.e5,
                         10 ** 1.5, 1000)
    rng = np.random.default_rng(seed)
    C1 = m.grid.nCells + 1
    X = rng.uniform(0.0, 1.0, size=(C1, 7, 7))
    for e, z in enumerate(m.elems):
        X[:, e, min(z + 1, 7):] = 0.0
        X[:, e] /= X[:, e].sum(axis=1, keepdims=True)
    Te = rng.uniform(5e3, 2e4, C1)
    Ne = rng.uniform(500., 1500., C1)
    H = np.full(C1, 1000.)
    return m, Te, Ne, H, X


def test_volume_averages_pinned(oracle):
    m, Te, Ne, H, X = small_state()
    C1 = m.grid.nCells + 1
    Tv, Xv = oracle.volume_averages(m.grid, m.elems, np.full(C1, 1.2e4), Ne, H, X)
    for e, z in enumerate(m.elems):
        top = min(z + 1, 7)
        np.testing.assert_allclose(Tv[e, :top], 1.2e4, rtol=1e-13)   # uniform Te
        assert np.all(Tv[e, top:] == 0.0) and np.all(Xv[e, top:] == 0.0)
        assert Xv[e, :top].sum() == pytest.approx(1.0, rel=1e-13)
    # one reached cell: its own values
    black = np.ones(C1, dtype=np.int32)
    black[5] = 0
    Tv, Xv = oracle.volume_averages(m.grid, m.elems, Te, Ne, H, X, lgBlack=black)
    for e, z in enumerate(m.elems):
        top = min(z + 1, 7)
        on = X[5, e, :top] > 0
        np.testing.assert_allclose(Tv[e, :top][on], Te[5], rtol=1e-13)
        np.testing.assert_allclose(Xv[e, :top], X[5, e, :top] / X[5, e, :top].sum(), rtol=1e-13)
    # two reached cells weight by Ne H dV
    black[7] = 0
    dV = T.cell_volumes(m.grid.xAxis, m.grid.yAxis, m.grid.zAxis, m.grid.active, True)
    w5, w7 = Ne[5] * H[5] * dV[5], Ne[7] * H[7] * dV[7]
    Tv, Xv = oracle.volume_averages(m.grid, m.elems, Te, Ne, H, X, lgBlack=black)
    x5, x7 = X[5, 0, 1], X[7, 0, 1]
    assert Tv[0, 1] == pytest.approx((x5 * w5 * Te[5] + x7 * w7 * Te[7]) / (x5 * w5 + x7 * w7), rel=1e-12)


def test_sed_sums_cells(oracle):
    rng = np.random.default_rng(9)
    E = rng.uniform(0, 1, size=(9, 5))
    s, tot = oracle.sed(E, True)
    np.testing.assert_allclose(s, 8 * E.sum(axis=0), rtol=1e-14)
    assert tot == pytest.approx(8 * E.sum(), rel=1e-14)


@pytest.mark.gpu
def test_gpu_volume_averages_and_sed_match_oracle(oracle):
    from mocassin_xray_amd import MocGPU
    m, Te, Ne, H, X = small_state()
    print("case output neb6", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        for k, v in (("Te", Te), ("Ne", Ne), ("Hden", H), ("ionDen", X)):
            g.upload(k, v)
        Tg, Xg = g.volume_averages()                   # no balance yet: every cell counts
        To, Xo = oracle.volume_averages(m.grid, m.elems, Te, Ne, H, X)
        np.testing.assert_allclose(Tg, To, rtol=1e-12, atol=0)
        np.testing.assert_allclose(Xg, Xo, rtol=1e-12, atol=1e-300)
        st = g.iterateMC(1, 5)
        black = g.download_flags("lgBlack")
        E = g.download("escaped")
        s, tot = g.sed()
        state = {k: g.download(k) for k in ("Te", "Ne", "Hden", "ionDen")}
        Tg, Xg = g.volume_averages()
    so, toto = oracle.sed(E, True)
    np.testing.assert_allclose(s, so, rtol=1e-12, atol=1e-300)
    assert tot == pytest.approx(toto, rel=1e-12)
    To, Xo = oracle.volume_averages(m.grid, m.elems, state["Te"], state["Ne"], state["Hden"],
                                    state["ionDen"], lgBlack=black)
    np.testing.assert_allclose(Tg, To, rtol=1e-12, atol=0)
    np.testing.assert_allclose(Xg, Xo, rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_gpu_sed_of_a_transparent_nebula_carries_the_luminosity():
    """Gas so thin that no packet interacts: every packet escapes, and the SED
    (the octant's tallies x 8) sums to the source luminosity."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_model("thin6", 6, 1.e18, 1.e-12, nebula.NEBULA_ELEMS,
                         10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 200, 100.0, 1.e5,
                         10 ** 1.5, 4000)
    print("case transparent SED", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.ionizationDriver()
        g.reset_tallies()
        ts = g.energyPacketDriver(1, 4000, seed=3, ctr_base=257)
        g.normalize_j()
        s, tot = g.sed()
    assert ts["escaped"] == 4000 and ts["absorbed"] == 0
    assert tot == pytest.approx(float(m.tables.Lstar[0]), rel=1e-12)

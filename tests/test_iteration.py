"""The Monte Carlo iteration (iterateMC): its starting state on the CPU, and on
the GPU every stage of one moc_iterate call against the oracle -- opacity and
emission of the starting state, the packets' J, the balance on that J -- the
skip of unchanged cells in the next iteration, and a converged pure-hydrogen
sphere against the Stromgren volume."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T


def test_initial_state_matches_setgrid_guess():
    hd = np.r_[0.0, np.full(5, 300.0)]
    st = nebula.initial_state([1, 2, 6, 16], 7, hd)
    X = st["ionDen"]
    np.testing.assert_allclose(X[1:].sum(axis=2), 1.0, rtol=1e-15)
    assert X[1, 0, 0] == 1e-5 and X[1, 1, 1] == X[1, 1, 2] == (1 - 1e-5) / 2
    assert np.all(X[1, 2, 1:7] == (1 - 1e-5) / 6)            # carbon: 6 ionised stages
    assert np.all(X[1, 3, 1:7] == (1 - 1e-5) / 6) and X[1, 3, 0] == 1e-5
    assert np.all(st["Ne"] == hd) and np.all(st["Te"][1:] == 1e4)


def small_nebula(n_packets, maxLev=27):
    return nebula.gas_model("nebula6", 6, 1.e18, 1000., nebula.NEBULA_ELEMS,
                            10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                            10 ** 1.5, n_packets, maxLev=maxLev)


@pytest.mark.gpu
@pytest.mark.parametrize("maxLev", [5, 27])
def test_one_iteration_matches_oracle_chain(oracle, maxLev):
    """maxLev 5: every n-level ion is solved one per lane in registers;
    maxLev 27 (the data files' level counts, up to 17 here): the ions of more
    than 8 levels are solved by a whole wave with the matrix in LDS, in the
    balance (cooling) and in the emission (lines)."""
    m = small_nebula(20000, maxLev)
    s0 = m.state
    seed = 4242
    print(f"case iteration-nebula6 maxLev {maxLev}: cells={m.grid.nCells} bins={m.tables.nbins}",
          flush=True)
    from mocassin_xray_amd import MocGPU
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines", "Jste", "Te", "Ne",
                                          "ionDen")}
    assert st["packets"] == 20000
    E, Z, NS = m.elems, m.abund, m.nstages
    # 1. opacity and emission of the starting state
    op = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"],
                        s0["Te"], s0["Hden"])
    np.testing.assert_allclose(got["opacity"][1:], op[1:], rtol=1e-11, atol=1e-300)
    em = oracle.emission(m.et, m.bt, E, Z, NS, s0["Te"], s0["Ne"], s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["recPDF"][1:], em["recPDF"][1:], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(got["totalLines"][1:], em["totalLines"][1:], rtol=1e-11)
    # 2. the packets (iteration 1, star 1 stream) and the normalisation of J
    fields = {"opacity": got["opacity"], "Ne": s0["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"]}
    Jo, Eo, co, qo = oracle.transport(m.tables, m.grid, fields, 20000, seed, ctr_base=257,
                                      nthreads=8)
    Jo = Jo * 1.e-9 / 8.0
    l1 = np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 2e-3, l1
    # 3. the balance on the J the device used
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["Ne"][1:], up["Ne"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    assert st["convPercent"] == pytest.approx(100.0 * up["lgConverged"][1:].mean())


@pytest.mark.gpu
def test_second_iteration_keeps_unchanged_cells(oracle):
    """Cells whose Te and Ne moved by < 0.5% since the last balance keep the
    opacity and diffuse spectrum of the previous iteration."""
    m = small_nebula(5000)
    from mocassin_xray_amd import MocGPU
    print(f"case iteration-skip: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.iterateMC(1, 7)
        op1, pdf1 = g.download("opacity"), g.download("recPDF")
        Te1, Ne1, X1 = g.download("Te"), g.download("Ne"), g.download("ionDen")
        keep = np.zeros(Te1.size, dtype=bool)
        keep[1::3] = True                       # these cells go back to their old Te, Ne
        Te2, Ne2 = Te1.copy(), Ne1.copy()
        Te2[keep] = m.state["Te"][keep] * (1 + 1e-3)
        Ne2[keep] = m.state["Ne"][keep]
        g.upload("Te", Te2)
        g.upload("Ne", Ne2)
        g.iterateMC(2, 7)
        op2, pdf2 = g.download("opacity"), g.download("recPDF")
    fresh = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, m.elems, m.abund, m.nstages, X1, Ne2,
                           Te2, m.state["Hden"])
    em = oracle.emission(m.et, m.bt, m.elems, m.abund, m.nstages, Te2, Ne2, X1, nthreads=8)
    keep[0] = False
    chg = ~keep
    chg[0] = False
    np.testing.assert_array_equal(op2[keep], op1[keep])
    np.testing.assert_array_equal(pdf2[keep], pdf1[keep])
    np.testing.assert_allclose(op2[chg], fresh[chg], rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(pdf2[chg], em["recPDF"][chg], rtol=1e-11, atol=1e-14)


@pytest.mark.gpu
def test_hydrogen_sphere_converges_to_stromgren_radius():
    """Pure-H sphere (16^3 octant, n_H = 100, 40 kK star of 1e37 erg/s): the
    radius of the ionised volume sum X(H+) dV over the cells packets reached
    (cells never hit keep their starting values, update_mod.f90:152-171)
    matches R_S^3 = 3 Q(H0) / (4 pi n_H^2 alpha_B(T)) (case B, Osterbrock &
    Ferland 2006) to 10%: re-emitting each absorbed packet with the local
    emissivity spectrum (which has no collisional cooling lines for pure H)
    returns more energy to the Lyman continuum than case B, ~6% in radius."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_hsphere(n=16, nbins=400, n_packets=300_000)
    print("case stromgren-hsphere16", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        hist = g.MCIterationDriver(maxIterateMC=8, minConvergence=101.0, seed=99)
        X = g.download("ionDen")
        Te = g.download("Te")
        black = g.download_flags("lgBlack")
    nu = m.tables.nuArray
    pdf = np.diff(np.r_[0.0, m.tables.sourceCDF[0]])
    ion = nu >= 1.0
    Q = m.tables.Lstar[0] * 1e36 * np.sum(pdf[ion] / (2.1799153e-11 * nu[ion]))
    dV = T.cell_volumes(m.grid.xAxis, m.grid.yAxis, m.grid.zAxis, m.grid.active, True) * 1.e45
    hit = black == 0
    hit[0] = False
    w = X[:, 0, 1] * dV
    Vion = 8.0 * np.sum(w[hit])
    Tm = np.sum(Te[hit] * w[hit]) / np.sum(w[hit])
    alphaB = 2.59e-13 * (Tm / 1e4) ** -0.8     # case B, 5e3-2e4 K
    Vs = Q / (100.0 ** 2 * alphaB)
    print(f"Q={Q:.4e} Tm={Tm:.0f} Vion={Vion:.4e} Vs={Vs:.4e}", flush=True)
    assert 8000.0 < Tm < 20000.0          # no metal coolants: a hot pure-H nebula
    assert (Vion / Vs) ** (1 / 3) == pytest.approx(1.0, abs=0.10)
    assert hist[-1]["convPercent"] > hist[0]["convPercent"]


@pytest.mark.gpu
def test_balance_in_cell_chunks_matches_one_launch(monkeypatch):
    """A rate buffer smaller than the grid runs the rates and the balance in
    chunks of cells (here 100 of 512, the last one short): on the same J the
    state is bit for bit that of one launch."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_model("nebula8", 8, 1.e18, 1000., nebula.NEBULA_ELEMS,
                         10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                         10 ** 1.5, 5000)
    monkeypatch.delenv("MOCGPU_RATES_BYTES", raising=False)
    print("case balance chunks: J of one iteration", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.iterateMC(1, 99)
        J = g.download("Jste")
    out = []
    for budget in (None, 100 * 8 * m.bt.W.shape[1]):
        if budget is not None:
            monkeypatch.setenv("MOCGPU_RATES_BYTES", str(budget))
        print(f"case balance chunks: budget={budget}", flush=True)
        with MocGPU(m.tables, "fp64") as g:
            m.load(g)
            g.upload("Jste", J)
            st = g.updateCell()
            out.append(({k: g.download(k) for k in ("Te", "Ne", "ionDen")}, st))
    for k in ("Te", "Ne", "ionDen"):
        np.testing.assert_array_equal(out[1][0][k], out[0][0][k])
    assert out[1][1]["converged"] == out[0][1]["converged"]

"""Batched grids: independent nebulae of uneven size attached together, their
cells numbered grid after grid, every grid's packets in one transport launch
and every cell in one balance launch.  On the GPU each grid of a batch is
checked against the oracle run on that grid alone (its own star, its slice of
the state, its packets' RNG streams)."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula


def small_batch(n_grids=6):
    return nebula.batch_nebulae(n_grids=n_grids, nbins=200, edge=(4, 10), packets=(2000, 5000))


def test_batch_builder_concatenates_members():
    m = small_batch()
    sizes = [g.nCells for g in m.grids]
    assert m.cellBase == list(np.cumsum([0] + sizes[:-1]))
    C1 = sum(sizes) + 1
    assert m.state["ionDen"].shape == (C1, 7, 7) and m.state["Hden"][0] == 0.0
    assert m.tables.sourceCDF.shape == (6, m.tables.nbins)
    assert m.n_packets == int(m.tables.nPhotons.sum())
    for k in range(6):
        mk = m.member(k)
        b = m.cellBase[k]
        np.testing.assert_array_equal(mk.state["Hden"][1:], m.state["Hden"][b + 1: b + 1 + sizes[k]])
        np.testing.assert_array_equal(mk.tables.sourceCDF[0], m.tables.sourceCDF[k])
        # a uniform member: every cell at its own n_H
        assert np.unique(mk.state["Hden"][1:]).size == 1
    # uneven sizes and densities across the batch
    assert len(set(sizes)) > 1 and np.unique(m.state["Hden"][1:]).size == 6


@pytest.mark.gpu
def test_gpu_batch_iteration_matches_oracle_per_grid(oracle):
    from mocassin_xray_amd import MocGPU
    m = small_batch()
    s0 = m.state
    seed = 606
    print(f"case batch6: cells={sum(g.nCells for g in m.grids)} packets={m.n_packets}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines", "Jste", "Te", "Ne",
                                          "ionDen", "lgConverged")}
    assert st["packets"] == m.n_packets
    E, Z, NS = m.elems, m.abund, m.nstages
    conv = 0
    for k in range(len(m.grids)):
        mk = m.member(k)
        b, n = m.cellBase[k], m.grids[k].nCells
        rows = np.r_[0, np.arange(b + 1, b + n + 1)]
        part = {key: v[rows] for key, v in got.items()}
        sk = mk.state
        op = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, sk["ionDen"], sk["Ne"],
                            sk["Te"], sk["Hden"])
        np.testing.assert_allclose(part["opacity"][1:], op[1:], rtol=1e-11, atol=1e-300)
        em = oracle.emission(m.et, m.bt, E, Z, NS, sk["Te"], sk["Ne"], sk["ionDen"], nthreads=8)
        np.testing.assert_allclose(part["recPDF"][1:], em["recPDF"][1:], rtol=1e-11, atol=1e-14)
        fields = {"opacity": part["opacity"], "Ne": sk["Ne"], "recPDF": part["recPDF"],
                  "totalLines": part["totalLines"]}
        Jo, Eo, co, qo = oracle.transport(mk.tables, mk.grid, fields, mk.n_packets, seed,
                                          ctr_base=257, nthreads=8, id_base=k << 32)
        assert co["errors"] == 0
        Jo = Jo * 1.e-9 / 8.0
        l1 = np.abs(part["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum()
        assert l1 < 2e-3, (k, l1)
        up = oracle.update(m.bt, E, Z, NS, True, part["Jste"], sk["Te"], sk["Ne"], sk["Hden"],
                           sk["ionDen"], nthreads=8)
        np.testing.assert_allclose(part["Te"][1:], up["Te"][1:], rtol=1e-9)
        np.testing.assert_allclose(part["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
        conv += int(up["lgConverged"][1:].sum())
    assert st["convPercent"] == pytest.approx(100.0 * conv / sum(g.nCells for g in m.grids))


@pytest.mark.gpu
def test_gpu_batch_single_star_transport(oracle):
    """moc_transport(iStar) in a batch sends star iStar's packets through grid
    iStar-1 only."""
    from mocassin_xray_amd import MocGPU
    m = small_batch(3)
    print("case batch3 single star", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.ionizationDriver()
        g.emissionDriver()
        g.reset_tallies()
        ts = g.energyPacketDriver(2, 3000, seed=77, ctr_base=9)
        J = g.download("Jste")
        op, rec, tl = g.download("opacity"), g.download("recPDF"), g.download("totalLines")
        for reduce in (g.volume_averages, g.sed, g.line_luminosities):  # one nebula per output
            with pytest.raises(RuntimeError, match="batch of grids"):
                reduce()
    b, n = m.cellBase[1], m.grids[1].nCells
    outside = np.r_[np.arange(1, b + 1), np.arange(b + n + 1, J.shape[0])]
    assert np.all(J[outside] == 0.0) and J[b + 1: b + n + 1].sum() > 0
    mk = m.member(1)
    rows = np.r_[0, np.arange(b + 1, b + n + 1)]
    fields = {"opacity": op[rows], "Ne": mk.state["Ne"], "recPDF": rec[rows], "totalLines": tl[rows]}
    Jo, Eo, co, qo = oracle.transport(mk.tables, mk.grid, fields, 3000, 77, ctr_base=9, nthreads=8,
                                      id_base=1 << 32)
    assert ts["segments"] == pytest.approx(co["segments"], rel=2e-3)
    l1 = np.abs(J[rows][1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 2e-3, l1

"""Graph replay of the iteration (moc_set_graph) and the transport's launch
parameters (moc_set_tuning): a replayed iteration draws the packets of the
eager call and lands on its state, replay after replay, through a re-seed and
a change of packet counts; other launch parameters leave every packet's path
unchanged and J equal up to the order of its atomic additions."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula


def model(n_packets):
    return nebula.gas_model("nebula6", 6, 1.e18, 1000., nebula.NEBULA_ELEMS,
                            10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                            10 ** 1.5, n_packets, maxLev=27)


FIELDS = ("Te", "Ne", "ionDen", "Jste", "opacity", "recPDF")
DUST_FIELDS = ("Tdust", "Jste", "opacity", "recPDF")


def state(g, fields=FIELDS):
    return {k: g.download(k) for k in fields}


def same_dust_state(a, b, what):
    for k in DUST_FIELDS:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-7, atol=1e-30, err_msg=f"{what}: {k}")


def same_state(a, b, what, mixed=False):
    """mixed: the mixed form sums the tallies of its hot (cell, bin) block as
    FP32 partials per workgroup, in an order that depends on which workgroup
    ran which packet; a key's sum then moves between two runs by up to n eps
    for n FP32 additions per key and workgroup (n reaches 1e4 on this 6^3 grid,
    whose every cell is in the box: up to ~1e-3 relative, against a Monte
    Carlo noise of a few per cent per key), and the balance passes a part of
    that on to the ion fractions and T_e."""
    rT, rX, rJ = (1e-5, 1e-4, 2e-3) if mixed else (1e-7, 1e-6, 1e-7)
    for k in ("Te", "Ne"):
        np.testing.assert_allclose(a[k][1:], b[k][1:], rtol=rT, err_msg=f"{what}: {k}")
    np.testing.assert_allclose(a["ionDen"][1:], b["ionDen"][1:], rtol=rX, atol=1e-20,
                               err_msg=f"{what}: ionDen")
    np.testing.assert_allclose(a["Jste"][1:], b["Jste"][1:], rtol=rJ, atol=1e-30,
                               err_msg=f"{what}: Jste")
    for k in ("opacity", "recPDF"):
        np.testing.assert_allclose(a[k][1:], b[k][1:], rtol=rT if mixed else 1e-7, atol=1e-30,
                                   err_msg=f"{what}: {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replays_match_eager_iterations(prec):
    from mocassin_xray_amd import MocGPU
    m = model(20000)
    print(f"case graph-replay {prec}: cells={m.grid.nCells} bins={m.tables.nbins}", flush=True)
    with MocGPU(m.tables, prec) as ge, MocGPU(m.tables, prec) as gg:
        m.load(ge)
        m.load(gg)
        gg.set_graph(True)
        assert gg.get_tuning()["graph"] == 1
        seed = 515
        for it in range(1, 7):
            if it == 4:  # new packet counts: the captured graph is dropped
                ge.set_packets(1, 15000)
                gg.set_packets(1, 15000)
            if it == 5:  # a new seed: captured again
                seed = 777
            se = ge.iterateMC(it, seed)
            sg = gg.iterateMC(it, seed)
            print(f"  iteration {it}: eager {se['iteration_ms']:.2f} ms, "
                  f"graph {sg['iteration_ms']:.2f} ms", flush=True)
            assert sg["packets"] == se["packets"]
            assert sg["segments"] == se["segments"], it
            assert sg["emissions"] == se["emissions"], it
            assert sg["escaped"] == se["escaped"], it
            # a cell's convergence flag is a threshold on its change since the
            # last iteration: the order of J's atomic additions can flip a
            # cell that sits on it, in two eager runs as well
            cell = 100.0 / m.grid.nCells
            assert abs(sg["convPercent"] - se["convPercent"]) <= 2 * cell + 1e-9, it
            assert abs(sg["noTeBalPercent"] - se["noTeBalPercent"]) <= 2 * cell + 1e-9, it
            if it > 1:
                assert sg["transport_ms"] == 0.0 and sg["iteration_ms"] > 0.0
            same_state(state(ge), state(gg), f"{prec} iteration {it}", mixed=prec == "mixed")


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["gas+grains", "grains"])
def test_graph_replays_match_eager_with_dust(kind):
    """Grains with photoelectric heating inside the balance, and a pure dust
    shell (kappa, setDustPDF and the grain temperatures in the graph)."""
    from mocassin_xray_amd import MocGPU
    if kind == "grains":
        m = nebula.dust_shell(n=8, nbins=150, n_packets=20000)
        fields = DUST_FIELDS
    else:
        m = nebula.gas_dust_nebula(n=6, nbins=300, n_packets=20000)
        fields = FIELDS + ("Tdust",)
    print(f"case graph-replay {kind}: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as ge, MocGPU(m.tables, "fp64") as gg:
        m.load(ge)
        m.load(gg)
        gg.set_graph(True)
        for it in range(1, 5):
            se = ge.iterateMC(it, 31)
            sg = gg.iterateMC(it, 31)
            assert sg["segments"] == se["segments"], it
            cell = 100.0 / m.grid.nCells
            assert abs(sg["convPercent"] - se["convPercent"]) <= 2 * cell + 1e-9, it
            a, b = state(ge, fields), state(gg, fields)
            same_dust_state(a, b, f"{kind} iteration {it}")
            if kind != "grains":
                same_state(a, b, f"{kind} iteration {it}")


@pytest.mark.gpu
def test_launch_parameters_change_no_packet():
    from mocassin_xray_amd import MocGPU
    m = model(30000)
    print(f"case tuning: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as ga, MocGPU(m.tables, "fp64") as gb:
        m.load(ga)
        m.load(gb)
        ga.ionizationDriver()
        gb.ionizationDriver()
        ga.emissionDriver()
        gb.emissionDriver()
        gb.set_tuning(emitBatch=8, segsPerTrip=5, transportWaves=256)
        t = gb.get_tuning()
        assert (t["emitBatch"], t["segsPerTrip"], t["transportWaves"]) == (8, 5, 256)
        sa = ga.energyPacketDriver(1, 30000, 99, 3)
        sb = gb.energyPacketDriver(1, 30000, 99, 3)
        for k in ("segments", "emissions", "absorbed", "scattered", "escaped", "lines"):
            assert sa[k] == sb[k], k
        np.testing.assert_allclose(gb.download("Jste"), ga.download("Jste"), rtol=1e-11,
                                   atol=1e-300)
        with pytest.raises(RuntimeError):
            gb.set_tuning(emitBatch=65)


@pytest.mark.gpu
def test_hot_j_block_changes_no_packet():
    """The mixed form's LDS sums of the hot J block (the cells around the star
    times the bins chosen from the previous launch's J) against global atomics
    for every tally: the same packets take the same paths, and J agrees to
    the FP32 rounding of the partial sums -- in the hot keys too, which the
    second launch fills (its hot bins come from the first)."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_nebula(n=12, nbins=600, n_packets=200_000, maxLev=5)
    print(f"case hot J block: cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "mixed") as ga, MocGPU(m.tables, "mixed") as gb:
        m.load(ga)
        m.load(gb)
        gb.set_tuning(hotJ=0)
        assert ga.get_tuning()["hotJ"] == 1 and gb.get_tuning()["hotJ"] == 0
        for g in (ga, gb):
            g.ionizationDriver()
            g.emissionDriver()
        for launch in (1, 2):
            ga.reset_tallies()
            gb.reset_tallies()
            sa = ga.energyPacketDriver(1, 200_000, 99, launch)
            sb = gb.energyPacketDriver(1, 200_000, 99, launch)
            for k in ("segments", "emissions", "absorbed", "scattered", "escaped", "lines"):
                assert sa[k] == sb[k], (launch, k)
            Ja, Jb = ga.download("Jste"), gb.download("Jste")
            # each key: FP32 partials of at most a few thousand additions
            np.testing.assert_allclose(Ja, Jb, rtol=2e-4, atol=1e-12 * Jb.max())
            print(f"launch {launch}: J sum {Ja.sum():.6e} vs {Jb.sum():.6e}", flush=True)
        assert abs(Ja.sum() - Jb.sum()) <= 1e-5 * Jb.sum()
        with pytest.raises(RuntimeError):
            ga.set_tuning(hotJ=2)


@pytest.mark.gpu
def test_autotuner_leaves_a_valid_choice(tmp_path):
    from mocassin_xray_amd import MocGPU
    from mocassin_xray_amd.autotune import tune_transport
    m = model(20000)
    print("case autotune", flush=True)
    with MocGPU(m.tables, "mixed") as g:
        m.load(g)
        g.ionizationDriver()
        g.emissionDriver()
        out = tune_transport(g, packets=20000, repeats=1, store=tmp_path / "tune.json")
        t = g.get_tuning()
        assert {k: t[k] for k in out["best"]} == out["best"]
        assert out["ms"] <= out["default_ms"]
        assert (tmp_path / "tune.json").exists()
        assert not g.download("Jste").any()  # the tallies are left zeroed
        st = g.iterateMC(1, 5)
        assert st["packets"] == 20000

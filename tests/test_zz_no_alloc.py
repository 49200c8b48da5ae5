"""The allocation-free entry point moc_volume_averages_noalloc against
moc_volume_averages, through the C ABI: two calls in a row on one GPU, the
second on a smaller grid, each bit for bit equal to the allocating entry
point on the same state (the partial sums are added in the same order)."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula


def state(n, seed):
    m = nebula.gas_model(f"neb{n}", n, 1.e18, 1000., nebula.NEBULA_ELEMS,
                         10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 200, 100.0, 1.e5,
                         10 ** 1.5, 1000, maxLev=5)
    rng = np.random.default_rng(seed)
    C1 = m.grid.nCells + 1
    X = rng.uniform(0.0, 1.0, size=(C1, 7, 7))
    for e, z in enumerate(m.elems):
        X[:, e, min(z + 1, 7):] = 0.0
        X[:, e] /= X[:, e].sum(axis=1, keepdims=True)
    return m, {"Te": rng.uniform(5e3, 2e4, C1), "Ne": rng.uniform(500., 1500., C1),
               "Hden": np.full(C1, 1000.), "ionDen": X}


@pytest.mark.gpu
def test_noalloc_volume_averages_match_the_allocating_entry_point():
    from mocassin_xray_amd import MocGPU
    for n, seed in ((40, 3), (6, 4)):     # 8000 cells (2 chunks), then 27 cells (1 chunk)
        m, st = state(n, seed)
        print(f"case volume averages noalloc: {m.grid.nCells} cells", flush=True)
        with MocGPU(m.tables, "fp64") as g:
            m.load(g)
            for k, v in st.items():
                g.upload(k, v)
            T1, X1 = g.volume_averages_noalloc()
            T0, X0 = g.volume_averages()
        np.testing.assert_array_equal(T1, T0)
        np.testing.assert_array_equal(X1, X0)
        assert np.all(T1[0, :2] > 0)

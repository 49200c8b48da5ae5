"""The X-ray nebula (30 elements H-Zn, all 31 stages, inner shells and Auger
yields, a power-law + blackbody source, a clumpy grid): its inputs on the
CPU, and on the GPU one moc_iterate call against the oracle at every stage."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T


def test_clumpy_density_and_source():
    hd = nebula.clumpy_density(20, fill=0.1)
    assert np.sum(hd == 1000.0) == 800 and np.sum(hd == 10.0) == 7200
    nu = T.log_mesh(3000, 1e-3, 1000.0)
    w = T.widths_of(nu)
    f = nebula.xray_source(nu, pl_fraction=0.3)
    np.testing.assert_allclose(np.sum(f * w), 1.0, rtol=1e-12)
    ryd_kev = 13.6057e-3
    band = (nu >= 0.01 / ryd_kev) & (nu <= 10.0 / ryd_kev)
    pl = np.where(band, 1.0 / nu, 0.0)
    pl_part = 0.3 * pl / np.sum(pl * w)
    np.testing.assert_allclose(np.sum(pl_part * w), 0.3, rtol=1e-12)
    # nu F_nu of the power law is flat across its band
    nfn = (nu * pl_part)[band]
    np.testing.assert_allclose(nfn, nfn[0], rtol=1e-12)


def test_xray_mesh_resolves_ion_edges():
    """The ionisation edge of every hydrogen-like ion up to Fe XXVI (9.28 keV,
    682 Ryd) falls on a mesh triplet; the mesh ends at 10 keV."""
    m = nebula.xray_nebula(n=4, nbins=2000, n_packets=1000)
    nu = m.tables.nuArray
    assert nu[-1] == pytest.approx(735.0, rel=1e-6) and nu.size > 1900
    hlike = [sh for sh in m.xsec.shells if sh.ion == sh.elem and sh.shell == 1]
    assert len(hlike) == 30
    for sh in hlike:
        if sh.elem > 26:
            continue                   # above nuMax
        e = nu[sh.lo - 1]                # the threshold bin is one of the triplet (e +- 3e-4)
        gap = min(e - nu[sh.lo - 2], nu[sh.lo] - e)
        assert gap <= 0.00031, (sh.elem, nu[sh.lo - 2: sh.lo + 1])
    fe = [sh for sh in hlike if sh.elem == 26][0]
    assert nu[fe.lo - 1] == pytest.approx(682.0, rel=0.01)


@pytest.mark.gpu
def test_xray_iteration_matches_oracle_chain(oracle):
    m = nebula.xray_nebula(n=6, nbins=2000, n_packets=20000)
    s0 = m.state
    seed = 515
    print(f"case xray6: cells={m.grid.nCells} bins={m.tables.nbins} elements=30 stages=31",
          flush=True)
    from mocassin_xray_amd import MocGPU
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines", "Jste", "Te", "Ne",
                                          "ionDen")}
    E, Z, NS = m.elems, m.abund, m.nstages
    op = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"],
                        s0["Te"], s0["Hden"])
    np.testing.assert_allclose(got["opacity"][1:], op[1:], rtol=1e-11, atol=1e-300)
    em = oracle.emission(m.et, m.bt, E, Z, NS, s0["Te"], s0["Ne"], s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["recPDF"][1:], em["recPDF"][1:], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(got["totalLines"][1:], em["totalLines"][1:], rtol=1e-11)
    fields = {"opacity": got["opacity"], "Ne": s0["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"]}
    Jo, Eo, co, qo = oracle.transport(m.tables, m.grid, fields, 20000, seed, ctr_base=257,
                                      nthreads=8)
    assert co["errors"] == 0
    Jo = Jo * 1.e-9 / 8.0
    l1 = np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 2e-3, l1
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["Ne"][1:], up["Ne"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    assert st["convPercent"] == pytest.approx(100.0 * up["lgConverged"][1:].mean())
    # X-rays keep the gas hot and ionise iron high
    assert np.all(got["Te"][1:] > 5e3)

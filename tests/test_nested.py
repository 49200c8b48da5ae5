"""Nested sub-grids (multiGrids): a mother grid whose cells around the star
are replaced by a finer sub-grid (setSubGrids grid_mod.f90:1834, fillGrid
:907-960, the sub-grid branches of pathSegment photon_mod.f90:1931-1963,
2169-2223, 2884-3948).

Pins of the oracle's restatement (CPU), on a sub-grid whose box ends on the
mother's cell walls, so that it replaces the covered mother cells exactly:
  * transparent gas: no packet interacts, so a packet's path -- and the J
    dV it leaves -- does not depend on how the volume is cut into cells: the
    total sum J dV of the nested run equals that of the mother alone, and
    the sub-grid's sum equals the covered mother cells' (less the pieces
    that the 2 % index nudge attributes across a wall); a wrong sub-grid
    volume, a packet dropped at a sub-grid face or a wrong re-entry into the
    mother fails one of them;
  * every packet escapes, none is lost, with and without the octant's
    mirror planes.
On the GPU: the FP64 transport on nested grids against the oracle with the
same packet streams, and a whole gas iteration on nested grids against the
oracle chain.
"""
import numpy as np
import pytest

from mocassin_xray_amd import MocGPU, nebula
from mocassin_xray_amd import tables as T

KEYS = ["segments", "emissions", "absorbed", "scattered", "escaped", "lines", "lost", "trapped"]


def volumes(grids):
    """getVolume of every cell of nested grids in the global numbering."""
    out = [np.zeros(1)]
    for k, g in enumerate(grids):
        v = T.cell_volumes(g.xAxis, g.yAxis, g.zAxis, np.maximum(g.active, 0), g.lgSymmetricXYZ,
                           plane=k > 0)
        out.append(v[1:])
    return np.concatenate(out)


def transparent_pair(sym):
    """A nested model and the mother alone (the covered cells back), both
    transparent.  The whole cube (sym False) sits at positive coordinates:
    the 2 % index nudge compares the distance to a wall with the wall's
    signed coordinate (photon_mod.f90:1882-1919), so at negative coordinates
    it always fires and moves path across walls, which would blur the
    comparison of the covered cells with the sub-grid."""
    n, R = 9, 3.e18
    if sym:
        ax = T.symmetric_axes(n, R)
        sax = np.linspace(0.0, (ax[2] + ax[3]) / 2.0, 13)
        star, idx = [0.0, 0.0, 0.0], [1, 1, 1]
    else:
        ax = T.full_axes(n, R) + 3.5e18
        c = n // 2
        sax = np.linspace((ax[c - 3] + ax[c - 2]) / 2.0, (ax[c + 2] + ax[c + 3]) / 2.0, 13)
        star, idx = [ax[c]] * 3, [c + 1] * 3
    nested, _ = nebula.nested_grids((ax, ax, ax), 100., [((sax, sax, sax), 100.)], sym)
    plain, _ = nebula.nested_grids((ax, ax, ax), 100., [], sym)
    nu = T.log_mesh(60, 1.0, 25.0)
    tabs = nebula.source_tables(nu, 4.e4, 10.0, 4000, sym, n, n, n)
    tabs.starPosition = np.array([star])
    tabs.starIndex = np.array([idx], dtype=np.int32)

    def fields(c1):
        return {"opacity": np.zeros((c1, nu.size)), "Ne": np.zeros(c1),
                "recPDF": np.ones((c1, nu.size)), "totalLines": np.zeros(c1)}
    return tabs, nested, plain, fields(sum(g.nCells for g in nested) + 1), fields(plain[0].nCells + 1)


@pytest.mark.parametrize("sym,tol", [(True, 2e-3), (False, 1e-2)])
def test_transparent_paths_do_not_depend_on_the_sub_grid(oracle, sym, tol):
    """tol bounds the path that the 2 % nudge attributes across the sub-grid's
    faces (its band is 2 % of the wall's coordinate)."""
    tabs, nested, plain, fn, fp = transparent_pair(sym)
    n = 4000
    Jn, En, cn, _ = oracle.transport(tabs, None, fn, n, 11, nested=nested)
    Jp, Ep, cp, _ = oracle.transport(tabs, plain[0], fp, n, 11)
    for c in (cn, cp):
        assert c["errors"] == 0 and c["lost"] == 0 and c["escaped"] == n, c
        assert c["absorbed"] == 0 and c["scattered"] == 0
    Vn, Vp = volumes(nested), volumes(plain)
    tot_n = (Jn * Vn[:, None]).sum()
    tot_p = (Jp * Vp[:, None]).sum()
    assert abs(tot_n - tot_p) <= 1e-10 * tot_p, (tot_n, tot_p)
    # the sub-grid's cells against the mother cells it covers
    nm = nested[0].nCells
    sub_n = (Jn[nm + 1:] * Vn[nm + 1:, None]).sum()
    ids = plain[0].active[nested[0].active < 0]
    sub_p = (Jp[ids] * Vp[ids, None]).sum()
    assert abs(sub_n - sub_p) <= tol * sub_p, (sub_n, sub_p)
    # the escaped energy per bin is the same
    assert np.allclose(En.sum(axis=0), Ep.sum(axis=0), rtol=1e-12, atol=0)


def test_sub_grid_cell_marking():
    """fillGrid's rule: a mother centre strictly inside the box, or on its
    lower face at 0 in an octant model, is covered."""
    ax = T.symmetric_axes(5, 4.0)            # 0 1 2 3 4
    sub = np.linspace(0.0, 2.5, 6)
    grids, h = nebula.nested_grids((ax, ax, ax), 1.0, [((sub, sub, sub), 1.0)], True)
    cov = grids[0].active < 0
    assert cov.sum() == 27 and cov[:3, :3, :3].all()
    assert (grids[0].active[cov] == -2).all()
    assert grids[0].nCells == 125 - 27 and grids[1].nCells == 216 and h.size == 98 + 216 + 1
    grids2, _ = nebula.nested_grids((ax, ax, ax), 1.0, [((sub, sub, sub), 1.0)], False)
    assert (grids2[0].active < 0).sum() == 8     # the centres at 0 stay in the mother


# ---------------------------------------------------------------- GPU parity
def run_gpu_nested(m, n, seed, prec="fp64", fields=None):
    f = m.fields if fields is None else fields
    with MocGPU(m.tables, prec) as g:
        g.grid_attach_nested(m.grids)
        for k in ("opacity", "Ne", "recPDF", "totalLines"):
            g.upload(k, f[k])
        st = g.energyPacketDriver(1, n, seed=seed)
        return g.download("Jste"), g.download("escaped"), st


@pytest.mark.gpu
@pytest.mark.parametrize("sym", [True, False])
def test_gpu_nested_transport_matches_oracle(oracle, sym):
    m = nebula.frozen_nested(n=9, sub_n=13, sub_cells=3, nbins=200, n_packets=20000, sym=sym)
    n, seed = 20000, 20261204
    print(f"case nested sym={sym}: n={n} cells={[g.nCells for g in m.grids]}", flush=True)
    J, E, st = run_gpu_nested(m, n, seed)
    Jo, Eo, co, qo = oracle.transport(m.tables, None, m.fields, n, seed, nthreads=8,
                                      nested=m.grids)
    assert co["errors"] == 0 and st["lost"] == co["lost"]
    for k in KEYS:
        assert abs(st[k] - co[k]) <= 2e-3 * max(co[k], 1.0) + 1, (k, st[k], co[k])
    l1 = np.abs(J[1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 1e-3, l1
    nm = m.grids[0].nCells
    l1s = np.abs(J[nm + 1:] - Jo[nm + 1:]).sum() / Jo[nm + 1:].sum()
    assert l1s < 1e-3, l1s          # the sub-grid's own cells
    assert abs(st["Qphot"] - qo) <= 1e-9 * qo


@pytest.mark.gpu
def test_gpu_nested_mixed_form_agrees_statistically():
    """The FP32 transport on nested grids: the same counters as the FP64 one
    within Monte Carlo noise, and the sub-grid's J within a few per cent."""
    m = nebula.frozen_nested(n=9, sub_n=13, sub_cells=3, nbins=200, n_packets=100000)
    n = 100000
    J64, _, s64 = run_gpu_nested(m, n, 3)
    J32, _, s32 = run_gpu_nested(m, n, 3, prec="mixed")
    for k in ("segments", "escaped", "absorbed"):
        assert abs(s32[k] - s64[k]) <= 0.02 * s64[k] + 30, (k, s32[k], s64[k])
    nm = m.grids[0].nCells
    a, b = J32[nm + 1:].sum(), J64[nm + 1:].sum()
    assert abs(a - b) <= 0.03 * b, (a, b)
    assert s32["lost"] == 0


@pytest.mark.gpu
def test_gpu_nested_gas_iteration_matches_oracle_chain(oracle):
    """One Monte Carlo iteration of a gas model on nested grids (opacity,
    emission, transport, normalisation and balance over every grid's cells)
    against the oracle: J from the device's opacity and CDFs with the same
    packet streams, then the balance on the device's J."""
    from mocassin_xray_amd import atomic, balance, emission
    n, sub_n, R = 7, 9, 1.e18
    ax = T.symmetric_axes(n, R)
    hi = (ax[1] + ax[2]) / 2.0
    sax = np.linspace(0.0, hi, sub_n)
    grids, hcell = nebula.nested_grids((ax, ax, ax), 1000., [((sax, sax, sax), 1000.)], True)
    elems, abund = nebula.NEBULA_ELEMS, 10.0 ** np.array(nebula.NEBULA_LOG_ABUND)
    nu = T.reference_gas_mesh(300, 1.0, 100.0, atomic.ion_edges(elems, 7, 100.0))
    npk = 30000
    tabs = nebula.source_tables(nu, 1.e5, 10 ** 1.5, npk, True, n, n, n)
    xs = atomic.build_xsec(nu, elems, 7)
    bt = balance.build_balance(xs, maxLev=5)
    et = emission.build_emission(xs, tabs.widFlx)
    st = nebula.initial_state(elems, 7, hcell)
    print(f"case nested gas: cells {[g.nCells for g in grids]}, {npk} packets", flush=True)
    with MocGPU(tabs, "fp64") as g:
        g.grid_attach_nested(grids)
        g.set_atomic(elems, abund, 7, xs)
        g.set_balance(bt)
        g.set_emission(et)
        for k in ("ionDen", "Te", "Ne", "Hden"):
            g.upload(k, st[k])
        g.iterateMC(1, 99)
        got = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines", "Jste", "Te", "Ne",
                                          "ionDen")}
        TeVol, _ = g.volume_averages()
    fields = {"opacity": got["opacity"], "Ne": st["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"]}
    Jo = oracle.transport(tabs, None, fields, npk, 99, ctr_base=257, nthreads=8,
                          nested=grids)[0] * 1.e-9 / 8.0
    l1 = np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 2e-3, l1
    up = oracle.update(bt, elems, abund, 7, True, got["Jste"], st["Te"], st["Ne"], st["Hden"],
                       st["ionDen"])
    r = np.max(np.abs(got["Te"][1:] - up["Te"][1:]) / up["Te"][1:])
    assert r < 1e-9, r
    nm = grids[0].nCells
    assert (got["Te"][nm + 1:] > 5000).all() and np.isfinite(TeVol).all()

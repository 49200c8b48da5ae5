"""Plane-parallel illumination (planeIonization): packets born on the y = 0
face of a slab at a uniform (x, z), moving along +y (photon_mod.f90:796-859),
leaving through the y faces and mirrored at the x and z faces (:3249-3606),
half-width end cells (:2166-2223, getVolume and getPlaneArea), 5000 segments
per flight (:1873-1878), and the field's luminosity (grid_mod.f90:1764-1820).

CPU: the host tables against closed forms, and the oracle against what the
geometry fixes -- a transparent slab (every packet crosses the full depth,
born in x/z columns in proportion to their half-width-ended widths), Beer-
Lambert escape through an absorbing slab, and a pure-scattering slab whose
mirrored walls make the field independent of x and z and let every packet
out through a y face.  GPU: libmocgpu against the oracle on the same packet
streams, and one whole iteration of a gas slab.
"""
import numpy as np
import pytest

from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T
from mocassin_xray_amd.mocgpu import CartesianGrid, SourceTables


def _axes(rng, n, lo, hi):
    a = np.sort(rng.uniform(lo, hi, n))
    a[0], a[-1] = lo, hi
    return a


# ------------------------------------------------------------ host tables
def test_plane_area_is_the_face():
    """getPlaneArea sums half-width-ended cells: the face (x_n-x_1)(z_n-z_1)
    exactly, on any axes."""
    rng = np.random.default_rng(3)
    xa, za = _axes(rng, 7, 1.e17, 3.e18), _axes(rng, 5, 0.0, 2.e18)
    assert T.plane_area(xa, za) == pytest.approx((xa[-1] - xa[0]) * (za[-1] - za[0]) / 1.e36,
                                                 rel=1e-13)


def test_plane_cell_volumes_fill_the_box():
    """Half-width end cells on every axis: the volumes add up to the box; the
    octant rule (full-width last cell) overshoots by the half cells."""
    rng = np.random.default_rng(5)
    xa, ya, za = _axes(rng, 6, 0.0, 2.e18), _axes(rng, 9, 0.0, 5.e18), _axes(rng, 4, 0.0, 1.e18)
    act = np.arange(1, 6 * 9 * 4 + 1, dtype=np.int32).reshape(6, 9, 4)
    box = (xa[-1] - xa[0]) * (ya[-1] - ya[0]) * (za[-1] - za[0]) / 1.e45
    v = T.cell_volumes(xa, ya, za, act, False, plane=True)
    assert v[1:].sum() == pytest.approx(box, rel=1e-13)
    vs = T.cell_volumes(xa, ya, za, act, True)
    assert vs[1:].sum() > box * (1 + 1e-3)
    # a corner cell of the slab is an eighth of an interior-sized cell's spans
    w = lambda a, i: abs(a[1] - a[0]) / 2 if i == 0 else abs(a[i] - a[i - 1]) / 2
    assert v[act[5, 8, 3]] == pytest.approx(w(xa, 5) * w(ya, 8) * w(za, 3) / 1.e45, rel=1e-13)


def test_plane_luminosity_energy_flux_and_face():
    """nu0 = 0: meanField is the energy flux itself; Lstar = flux x face."""
    nu = T.log_mesh(50, 1.0, 20.0)
    xa, za = np.linspace(0, 3.e18, 4), np.linspace(0, 2.e18, 5)
    L, mf, cdf = T.plane_source(nu, T.widths_of(nu), T.blackbody_flux(nu, 5.e4), 2.5, 0.0, 0.0,
                                xa, za)
    assert mf == 2.5
    assert L == pytest.approx(2.5 * 6.e36 / 1.e36, rel=1e-14)
    np.testing.assert_allclose(cdf, T.source_cdf(T.blackbody_flux(nu, 5.e4), T.widths_of(nu)))


def test_plane_luminosity_photon_flux_scaling():
    """nu0 > 0: meanFieldin counts photons from nu0P on; the energy flux is
    photons x h nu.  Two lit bins k-1 and k (photon fluxes p1, p2): with nu0
    below their midpoint both count, above it only bin k does (the midpoint
    rule of grid_mod.f90:1771-1775); nu0Add at bin k drops bin k-1's energy."""
    nu = T.log_mesh(40, 1.0, 10.0)
    wid = T.widths_of(nu)
    k = 20
    p1, p2 = 3.0, 5.0                        # photons per bin (flux x width)
    erg = np.zeros(nu.size)
    erg[k - 1] = p1 * T.HCRYD * nu[k - 1] / wid[k - 1]
    erg[k] = p2 * T.HCRYD * nu[k] / wid[k]
    e1, e2 = p1 * T.HCRYD * nu[k - 1], p2 * T.HCRYD * nu[k]
    xa = za = np.linspace(0.0, 1.e18, 3)
    phi = 1.e10
    mid = (nu[k - 1] + nu[k]) / 2
    below, above = mid * (1 - 1e-9), mid * (1 + 1e-9)
    _, mf, _ = T.plane_source(nu, wid, erg, phi, below, below, xa, za)
    assert mf == pytest.approx(phi * (e1 + e2) / (p1 + p2), rel=1e-12)
    _, mf, _ = T.plane_source(nu, wid, erg, phi, above, above, xa, za)
    assert mf == pytest.approx(phi * e2 / p2, rel=1e-12)
    L, mf, cdf = T.plane_source(nu, wid, erg, phi, below, nu[k], xa, za)
    assert mf == pytest.approx(phi * e2 / (p1 + p2), rel=1e-12)
    assert L == pytest.approx(mf / 1.e36 * 1.e36, rel=1e-12)
    # the CDF starts at locate(nu, nu0Add) = k + 1 (1-based): bin k-1 is gone
    assert np.all(cdf[:k] == 0.0) and cdf[k] == 1.0
    with pytest.raises(ValueError):
        T.plane_source(nu, wid, erg, phi, 100.0, 100.0, xa, za)


def test_source_cdf_start_bin():
    nu = T.log_mesh(30, 1.0, 5.0)
    f = T.blackbody_flux(nu, 3.e4)
    c = T.source_cdf(f, T.widths_of(nu), start=11)
    assert np.all(c[:10] == 0.0) and c[10] > 0.0 and c[-1] == 1.0
    fw = f * T.widths_of(nu)
    np.testing.assert_allclose(np.diff(c[10:]), fw[11:] / fw[10:].sum(), rtol=1e-9, atol=1e-15)


def test_lineflux_hbeta_per_unit_face(tmp_path):
    """lineFlux.out of a slab: Hbeta [1e36 erg/s] over getPlaneArea [1e36 cm^2]."""
    from mocassin_xray_amd import outputs
    xa, za = np.linspace(0, 2.e18, 4), np.linspace(0, 3.e18, 3)
    area = T.plane_area(xa, za)
    p = outputs.write_lineflux(tmp_path, np.zeros((0, 5, 5)), [], 4.0e16, None, False, 2,
                               planeArea=area)
    rows = [ln for ln in p.read_text().splitlines() if "erg/cm^2/s" in ln]
    assert len(rows) == 2                      # the total and component 1
    v = [float(t) for t in rows[0].split()[-2:]]
    assert v[0] == pytest.approx(4.0 / 6.0, rel=1e-8) and v[1] == pytest.approx(6.0, rel=1e-8)


def test_checkpoint_carries_the_plane_flag(tmp_path):
    from mocassin_xray_amd import checkpoint as K
    grid, hcell = nebula.slab_grid((3, 4, 3), 1.e18, 2.e18, 50.0)
    C1 = grid.nCells + 1
    st = {"Te": np.r_[0.0, np.full(C1 - 1, 1.e4)], "Ne": hcell.copy(), "Hden": hcell.copy(),
          "ionDen": np.zeros((C1, 1, 2)), "lgConverged": np.zeros(C1, np.int32),
          "lgBlack": np.zeros(C1, np.int32)}
    st["ionDen"][1:, 0, :] = 0.5
    p = K.RunParams(nbins=50, nstages=2, lgPlaneIonization=True, meanFieldin=1.e10, nu0=1.0)
    ck = K.Checkpoint(grid, st, p, [K.PhotoSource("blackbody", 4.e4, 1.2, 1000, (0., 0., 0.))],
                      None, 0)
    K.write_grid(tmp_path, ck, [1])
    back = K.read_grid(tmp_path, [1])
    assert back.grid.lgPlaneIonization and not back.grid.lgSymmetricXYZ
    assert back.params.meanFieldin == 1.e10 and back.params.nu0 == 1.0
    assert back.grid.R_out == grid.R_out


# ------------------------------------------------------------- oracle pins
def _slab(shape, side, depth, nbins=40, n=20000, lgCompton=False, edge=0.5):
    grid, hcell = nebula.slab_grid(shape, side, depth, 100.0)
    nu = T.log_mesh(nbins, 1.0, 10.0)
    tabs = nebula.slab_source_tables(nu, 4.e4, 1.e10, 1.0, 0.0, grid, n, lgCompton,
                                     ion_edge1=edge)
    return grid, hcell, nu, tabs


def _fields(grid, nu, kap, sca=None, lines=1.0):
    C1 = grid.nCells + 1
    op = np.full((C1, nu.size), kap)
    op[0] = 0.0
    cdf = np.tile(np.linspace(0.0, 1.0, nu.size), (C1, 1))
    f = {"opacity": op, "Ne": np.zeros(C1), "recPDF": cdf, "totalLines": np.full(C1, lines)}
    if sca is not None:
        f["scaOpac"] = np.where(op > 0, sca, 0.0)
        f["dustOnly"] = True
    return f


def _columns(grid, J, axis):
    """Path length [cm] per column along `axis` summed over the others."""
    dV = T.cell_volumes(grid.xAxis, grid.yAxis, grid.zAxis, grid.active, False, plane=True)
    pl = J.sum(axis=1) * dV                       # deltaE x path [cm] per cell
    cube = np.zeros(grid.active.shape)
    on = grid.active > 0
    cube[on] = pl[grid.active[on]]
    other = tuple(a for a in range(3) if a != axis)
    return cube.sum(axis=other)


def test_transparent_slab_every_packet_crosses(oracle):
    shape, side, depth, n = (5, 7, 5), 2.e18, 6.e18, 40000
    grid, _, nu, tabs = _slab(shape, side, depth, n=n)
    f = _fields(grid, nu, 0.0)
    f["opacity"][1:] = 1.e-40                     # no interaction on any path
    J, E, c, _ = oracle.transport(tabs, grid, f, n, 11, nthreads=8)
    dE = tabs.Lstar[0] / n
    assert c["errors"] == 0 and c["lost"] == 0 and c["escaped"] == n and c["absorbed"] == 0
    assert c["emissions"] == n
    assert E.sum() == pytest.approx(n * dE, rel=1e-12)
    # every packet walks the full depth along y
    cols = _columns(grid, J, 1)
    assert cols.sum() == pytest.approx(n * dE * depth, rel=1e-10)
    # births over the x (and z) columns: widths with half-width ends, 1/8 1/4 1/4 1/4 1/8
    for ax in (0, 2):
        frac = _columns(grid, J, ax) / (dE * depth) / n
        want = np.array([1, 2, 2, 2, 1]) / 8.0
        sig = np.sqrt(want * (1 - want) / n)
        assert np.all(np.abs(frac - want) < 5 * sig), (ax, frac)


def test_absorbing_slab_beer_lambert(oracle):
    shape, side, depth, n = (3, 10, 3), 1.e18, 4.e18, 60000
    grid, _, nu, tabs = _slab(shape, side, depth, n=n)
    tau = 1.3
    f = _fields(grid, nu, tau / depth, lines=1.0)  # absorbed packets end as lines
    J, E, c, _ = oracle.transport(tabs, grid, f, n, 23, nthreads=8)
    assert c["errors"] == 0 and c["lost"] == 0
    assert c["absorbed"] + c["escaped"] == n and c["lines"] == c["absorbed"]
    p = np.exp(-tau)
    assert abs(c["escaped"] / n - p) < 5 * np.sqrt(p * (1 - p) / n)
    # mean path through the slab: (1 - e^-tau)/kappa per packet
    path = _columns(grid, J, 1).sum() / (tabs.Lstar[0] / n)
    want = n * (1 - p) * depth / tau
    assert abs(path - want) < 5 * depth * np.sqrt(n) * 0.5 / tau


def test_scattering_slab_mirrors(oracle):
    """Pure isotropic scattering (grains with albedo 1): every packet leaves
    through a y face and the mirrored x/z walls make the mean intensity per
    volume the same in every x (and z) column, end columns included."""
    shape, side, depth, n = (5, 6, 5), 1.e18, 2.e18, 40000
    grid, _, nu, tabs = _slab(shape, side, depth, n=n)
    kap = 2.0 / depth
    f = _fields(grid, nu, kap, sca=kap)
    J, E, c, _ = oracle.transport(tabs, grid, f, n, 5, nthreads=8)
    dE = tabs.Lstar[0] / n
    assert c["errors"] == 0 and c["lost"] == 0 and c["absorbed"] == 0
    assert c["escaped"] == n and c["scattered"] > n
    assert E.sum() == pytest.approx(n * dE, rel=1e-12)
    on = grid.active > 0
    Jc = np.zeros(grid.active.shape)
    Jc[on] = J.sum(axis=1)[grid.active[on]]
    for ax in (0, 2):
        prof = Jc.mean(axis=tuple(a for a in range(3) if a != ax))
        # per-volume tallies: the half-width end columns are not halved
        assert np.max(np.abs(prof / prof.mean() - 1)) < 0.03, (ax, prof / prof.mean())


def test_thick_scattering_slab_hits_the_plane_safe_limit(oracle):
    """An optically very thick pure scatterer: packets that random-walk into
    the slab are dropped after 5000 segments (the plane-parallel limit; the
    500000 of other grids would drop ~10x fewer).  Fewer than 5000 segments
    per packet on average, and a percent-level loss."""
    shape, side, depth, n = (3, 4, 3), 1.e18, 1.e18, 6000
    grid, _, nu, tabs = _slab(shape, side, depth, n=n)
    kap = 1.e4 / depth
    f = _fields(grid, nu, kap, sca=kap)
    _, _, c, _ = oracle.transport(tabs, grid, f, n, 9, nthreads=8)
    assert c["errors"] == 0 and c["escaped"] + c["lost"] == n
    assert 0.004 * n < c["lost"] < 0.05 * n, c["lost"]
    assert c["segments"] <= 5000 * c["lost"] + 5000 * c["escaped"]


# ------------------------------------------------------------------- GPU
def _compare(m, n, seed, oracle, tol_cnt=2e-3, tol_j=1e-3):
    from mocassin_xray_amd import MocGPU
    print(f"case {m.name}: n={n} seed={seed} cells={m.grid.nCells} bins={m.tables.nbins}",
          flush=True)
    with MocGPU(m.tables, "fp64") as g:
        g.grid_attach(m.grid)
        for k in ("opacity", "Ne", "recPDF", "totalLines"):
            g.upload(k, m.fields[k])
        st = g.energyPacketDriver(1, n, seed=seed)
        J, E = g.download("Jste"), g.download("escaped")
    Jo, Eo, co, qo = oracle.transport(m.tables, m.grid, m.fields, n, seed, nthreads=8)
    assert co["errors"] == 0 and st["lost"] == co["lost"] == 0
    for k in ("segments", "emissions", "absorbed", "scattered", "escaped", "lines"):
        assert abs(st[k] - co[k]) <= tol_cnt * max(co[k], 1.0) + 1, (k, st[k], co[k])
    l1 = np.abs(J[1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < tol_j, l1
    assert np.abs(E[1:] - Eo[1:]).sum() / Eo[1:].sum() < 5 * tol_j
    return st, co


@pytest.mark.gpu
def test_gpu_slab_matches_oracle(oracle):
    m = nebula.frozen_slab(shape=(5, 16, 5), nbins=200, n_packets=20000)
    _compare(m, 20000, 2026, oracle)


@pytest.mark.gpu
def test_gpu_thomson_thick_slab_mirrors_match_oracle(oracle):
    """Electron scattering dominates (tau_T ~ 3 across the slab): packets
    bounce off the x/z mirrors many times."""
    grid, hcell = nebula.slab_grid((4, 8, 3), 1.e16, 4.e16, 1.e8)
    nu = T.log_mesh(120, 1.0, 25.0)
    tabs = nebula.slab_source_tables(nu, 4.e4, 1.e14, 1.0, 0.0, grid, 8000, True)
    C1 = grid.nCells + 1
    ne = np.r_[0.0, np.full(C1 - 1, 1.e8)]
    op = ne[:, None] * tabs.KNsigmaT[None, :] * 1.25
    cdf = np.tile(np.linspace(0.0, 1.0, nu.size), (C1, 1))
    cdf[0] = 0.0
    f = {"opacity": op, "Ne": ne, "recPDF": cdf, "totalLines": np.r_[0.0, np.full(C1 - 1, 0.4)]}
    m = nebula.FrozenModel("thomson-slab", tabs, grid, f, 8000)
    st, co = _compare(m, 8000, 99, oracle)
    assert co["scattered"] > 8000


@pytest.mark.gpu
def test_gpu_gas_slab_iteration_matches_oracle_chain(oracle):
    m = nebula.gas_slab(n_packets=20000)
    s0 = m.state
    seed = 808
    print(f"case gas-slab: cells={m.grid.nCells} bins={m.tables.nbins}", flush=True)
    from mocassin_xray_amd import MocGPU
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines", "Jste", "Te", "Ne",
                                          "ionDen")}
        TeV, XV = g.volume_averages()
    E, Z, NS = m.elems, m.abund, m.nstages
    op = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"],
                        s0["Te"], s0["Hden"])
    np.testing.assert_allclose(got["opacity"][1:], op[1:], rtol=1e-11, atol=1e-300)
    fields = {"opacity": got["opacity"], "Ne": s0["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"]}
    Jo, _, co, _ = oracle.transport(m.tables, m.grid, fields, 20000, seed, ctr_base=257,
                                    nthreads=8)
    assert co["errors"] == 0
    Jo = Jo * 1.e-9                                # no octant factor on a slab
    l1 = np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 2e-3, l1
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    # the ionised layer sits at the lit face
    Xh = got["ionDen"][1:, 0, 1].reshape(m.grid.active.shape)
    assert Xh[:, 0, :].mean() > 0.9
    # volume averages with the slab's half-width end cells
    to, Xo = oracle.volume_averages(m.grid, E, got["Te"], got["Ne"], s0["Hden"], got["ionDen"])
    np.testing.assert_allclose(TeV, to, rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(XV, Xo, rtol=1e-10, atol=1e-300)


@pytest.mark.gpu
def test_gpu_plane_grid_rejects_symmetry_and_thin_faces():
    from mocassin_xray_amd import MocError, MocGPU
    grid, _ = nebula.slab_grid((3, 4, 3), 1.e18, 1.e18, 10.0)
    tabs = nebula.slab_source_tables(T.log_mesh(30, 1.0, 5.0), 3.e4, 1.e9, 1.0, 0.0, grid, 10)
    with MocGPU(tabs, "fp64") as g:
        bad = CartesianGrid(grid.xAxis, grid.yAxis, grid.zAxis, grid.active, grid.nCells, True,
                            grid.R_out, lgPlaneIonization=True)
        with pytest.raises(MocError):
            g.grid_attach(bad)
        g2, _ = nebula.slab_grid((2, 4, 3), 1.e18, 1.e18, 10.0)
        with pytest.raises(MocError):
            g.grid_attach(g2)
        g.grid_attach(grid)

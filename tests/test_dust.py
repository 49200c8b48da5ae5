"""Grains: the host tables, the oracle's dust routines pinned by closed forms
and conservation laws, and on the GPU the dust opacity, emission CDF, grain
temperatures, dust transport and dust-only iterations against the oracle.

Pins (CPU):
  Planck function      Wien and Rayleigh-Jeans limits of getFlux, and the
                       Stefan-Boltzmann integral over a fine mesh
  emission integrals   a grey grain radiates 4 Cabs sigma T^4 (dust_mod.f90:
                       146-184 in closed form); host table = oracle sum
  getDustT             a grain in a blackbody field of T0 (J = 4 pi B dnu,
                       Kirchhoff) comes out at T0
  transport            without gas every packet ends escaping: the escaped
                       energy is n deltaE exactly, with and without absorption
  sublimation          grains at or above TdustSublime drop out of kappa and
                       of the emission CDF
  Mie                  the Bohren & Huffman test sphere and the Rayleigh limit
"""
import numpy as np
import pytest

from mocassin_xray_amd import dust as D
from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T

SIGMA_SB = 5.67040e-5
H = 6.6262e-27
FR1RYD = 3.28984e15


@pytest.fixture(scope="module")
def shell():
    return nebula.dust_shell(n=8, nbins=120, n_packets=4000)


def tdust_of(m, T=None, rng=None):
    C1 = m.grid.nCells + 1
    nS, nA = m.dust.nSpecies, m.dust.nSizes
    Td = np.full((C1, nA + 1, nS + 1), 50.0)
    if rng is not None:
        Td[1:, 1:, 1:] = rng.uniform(20.0, 900.0, size=(C1 - 1, nA, nS))
    elif T is not None:
        Td[:] = T
    return Td


# ---------------------------------------------------------------- CPU pins
def test_mie_reference_sphere_and_rayleigh_limit():
    x = 2 * np.pi * 0.525 / 0.6328
    qe, qs = D.mie_q(np.array([x]), np.array([1.55 + 0j]))
    assert abs(qe[0] - 3.10543) < 2e-5 and abs(qs[0] - 3.10543) < 2e-5
    m = np.array([1.7 + 0.03j])
    x = np.array([1e-3])
    qe, qs = D.mie_q(x, m)
    a = (m ** 2 - 1) / (m ** 2 + 2)
    np.testing.assert_allclose(qe - qs, 4 * x * a.imag, rtol=1e-5)
    np.testing.assert_allclose(qs, 8 / 3 * x ** 4 * abs(a) ** 2, rtol=1e-5)


def test_planck_limits_and_stefan_boltzmann(oracle):
    T0 = 300.0
    nu = np.array([1e-7, 0.5])                 # x = 5e-5 and 263
    b = oracle.planck_h(nu, T0)
    rj = 3.32154e-6 * nu[0] ** 2 * T0 / H
    assert abs(b[0] / rj - 1) < 2e-3           # the reference's RJ constant is 0.1 % off
    wien = 0.5250229 / H * nu[1] ** 3 * np.exp(-157893.94 * nu[1] / T0)
    assert b[1] == pytest.approx(wien, rel=1e-13)
    mesh = np.geomspace(1e-7, 1.0, 20000)
    w = T.widths_of(mesh)
    integral = np.sum(oracle.planck_h(mesh, T0) * H * FR1RYD * w)
    assert integral == pytest.approx(SIGMA_SB * T0 ** 4 / np.pi, rel=2e-3)


def test_grey_grain_emission_integral(oracle):
    mesh = np.geomspace(1e-7, 1.0, 6000)
    w = T.widths_of(mesh)
    sigma = 3.0e-10
    C = np.full((1, 1, mesh.size), sigma)
    emi = D.emission_integrals(C, mesh, w, n_temps=400)
    temps = np.array([20.0, 100.0, 400.0])
    want = 4 * sigma * SIGMA_SB * temps ** 4 / np.pi
    np.testing.assert_allclose(emi[0, 0, temps.astype(int) - 1], want, rtol=3e-3)
    got = oracle.dust_em_integral(C[0, 0], mesh, w, temps)
    np.testing.assert_allclose(emi[0, 0, temps.astype(int) - 1], got, rtol=1e-12)


def test_grain_temperature_of_a_blackbody_field(oracle, shell):
    m = shell
    nu, w = m.tables.nuArray, m.tables.widFlx
    C1 = m.grid.nCells + 1
    T0 = np.array([0.0] + [37.25 + 11.0 * (c % 70) for c in range(1, C1)])
    J = np.zeros((C1, nu.size))
    for c in range(1, C1):
        J[c] = 4 * np.pi * oracle.planck_h(nu, T0[c]) * H * FR1RYD * w
    Td, conv, black, nohit, nconv = oracle.dust_temperature(m.dust, nu, w, m.Ndust, tdust_of(m), J)
    assert nohit == 0 and black[1:].sum() == 0
    np.testing.assert_allclose(Td[1:, 1, 1], T0[1:], atol=0.02)
    np.testing.assert_allclose(Td[1:, 0, 1], Td[1:, 1, 1] * m.dust.grainWeight[0], rtol=1e-15)
    assert conv[1:].sum() == nconv
    J[3] = 0.0                                   # a cell no packet reached keeps its grains
    Td2, conv2, black2, nohit2, _ = oracle.dust_temperature(m.dust, nu, w, m.Ndust,
                                                            tdust_of(m, 77.0), J)
    assert nohit2 == 1 and black2[3] == 1 and np.all(Td2[3] == 77.0)


def test_sublimated_grains_leave_kappa_and_emission(oracle):
    rng = np.random.default_rng(20261203 + 3)
    nu = np.geomspace(1e-4, 5.0, 40)
    w = T.widths_of(nu)
    dt = D.build_dust(nu, w, ["sil-dlee", "grph1-dl"], [0.6, 0.4], [0.01, 0.05, 0.2],
                      [1.0, 0.5, 0.1])
    C1 = 6
    nd = np.r_[0.0, rng.uniform(1e-9, 1e-7, C1 - 1)]
    Td = np.full((C1, 4, 3), 100.0)
    op, ab, sc = oracle.dust_opacity(dt, nu, w, nd, Td)
    want_s = np.zeros((C1, nu.size))
    want_a = np.zeros((C1, nu.size))
    for s in range(2):
        for a in range(3):
            f = dt.grainAbun[s] * dt.grainWeight[a]
            want_s[1:] += f * nd[1:, None] * dt.Csca[s, a][None]
            want_a[1:] += f * nd[1:, None] * dt.Cabs[s, a][None]
    np.testing.assert_allclose(sc, want_s, rtol=1e-13)
    np.testing.assert_allclose(ab, want_a, rtol=1e-13)
    np.testing.assert_allclose(op, want_s + want_a, rtol=1e-13)
    Td[2, :, 1] = dt.Tsub[0]                     # species 1 sublimated in cell 2
    op2, ab2, sc2 = oracle.dust_opacity(dt, nu, w, nd, Td)
    part = sum(dt.grainAbun[1] * dt.grainWeight[a] * nd[2] * dt.Csca[1, a] for a in range(3))
    np.testing.assert_allclose(sc2[2], part, rtol=1e-13)
    np.testing.assert_allclose(sc2[3], sc[3], rtol=0)
    pdf = oracle.dust_pdf(dt, nu, w, nd, Td)
    assert np.all(pdf[1:, -1] == 1.0) and np.all(np.diff(pdf[1:], axis=1) >= 0)
    inc = np.zeros(nu.size)
    for a in range(3):
        inc += dt.Cabs[1, a] * oracle.planck_h(nu, 100.0) * w * dt.grainWeight[a] * dt.grainAbun[1]
    np.testing.assert_allclose(pdf[2, :-1], (np.cumsum(inc) / inc.sum())[:-1], rtol=1e-12)


def test_dust_only_transport_conserves_energy(oracle, shell):
    m = shell
    C1, B = m.grid.nCells + 1, m.tables.nbins
    for absorb in (False, True):
        Td = tdust_of(m, 40.0)
        dt = m.dust
        if not absorb:       # pure scatterers: every packet escapes with its first frequency
            dt = D.DustTables(dt.species, dt.grainAbun, dt.radius, dt.grainWeight, dt.Tsub, dt.rho,
                              np.zeros_like(dt.Cabs), dt.Csca * 5.0, dt.emIntegral)
        op, ab, sc = oracle.dust_opacity(dt, m.tables.nuArray, m.tables.widFlx, m.Ndust * 3.0, Td)
        pdf = oracle.dust_pdf(m.dust, m.tables.nuArray, m.tables.widFlx, m.Ndust, Td)
        fields = {"opacity": op, "Ne": np.zeros(C1), "recPDF": pdf, "totalLines": np.zeros(C1),
                  "scaOpac": sc, "dustOnly": True}
        n = 3000
        J, E, cnt, _ = oracle.transport(m.tables, m.grid, fields, n, 11, nthreads=4)
        dE = m.tables.Lstar[0] / m.tables.nPhotons[0]
        assert cnt["errors"] == 0 and cnt["lines"] == 0 and cnt["lost"] == 0
        assert cnt["escaped"] == n
        assert E.sum() == pytest.approx(n * dE, rel=1e-12)
        assert cnt["scattered"] > n
        if absorb:
            assert cnt["absorbed"] > 0 and cnt["emissions"] == n + cnt["absorbed"]
        else:
            assert cnt["absorbed"] == 0
            # the escaped spectrum is the stellar one: no bin below the source's support
            assert E[:, m.tables.sourceCDF[0] == 0].sum() == 0.0


# ------------------------------------------- grain charge and photoelectric terms
HCRYD = 2.1799153e-11
PI_REF = 3.141592654
ME = 9.109382e-28
AMU = 1.66053e-24


@pytest.fixture(scope="module")
def pe_model():
    return nebula.gas_dust_model("pe4", 4, 1.e18, 300., nebula.NEBULA_ELEMS,
                                 10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 1.001e-5, 25.0,
                                 5.e4, 10 ** 1.5, 1000, ["sil-dlee", "grph1-dl"], [0.5, 0.5],
                                 [0.01, 0.05, 0.2], [1.0, 0.3, 0.05], 0.01, photoelectric=True)


def dilute_field(oracle, m, mask=None):
    """4 pi W B_nu(50 kK) dnu per cell, W = 1e-15 x (1 .. 5); bins outside mask zero."""
    nu, w = m.tables.nuArray, m.tables.widFlx
    C1 = m.grid.nCells + 1
    bb = 4 * np.pi * oracle.planck_h(nu, 5.e4) * H * FR1RYD * w
    if mask is not None:
        bb = np.where(mask, bb, 0.0)
    J = np.zeros((C1, nu.size))
    for c in range(1, C1):
        J[c] = 1.e-15 * (1 + c % 5) * bb
    return J


def pe_update(oracle, m, J):
    s0 = m.state
    Td = np.full((m.grid.nCells + 1, m.dust.nSizes + 1, m.dust.nSpecies + 1), 50.0)
    up = oracle.update(m.bt, m.elems, m.abund, m.nstages, True, J, s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8, grains=(m.dust, Td, m.Ndust))
    return up, Td


def charge_rates(m, J, cell, g, V, Te, Ne):
    """Photoemission and net recombination currents of grain g at potential V
    (Baldwin et al. 1991 yields; electrons less the ions of the starting
    state), evaluated from the definitions."""
    dt, nu = m.dust, m.tables.nuArray
    s, a = divmod(g, dt.nSizes)
    Vn, r = dt.Vn[s], dt.radius[a]
    on = nu > Vn + max(V, 0.0)
    Yn = np.minimum(0.5 * (1 - Vn / nu), 0.2)
    Yhat = Yn * np.clip(1 - V / (nu - Vn), 0.0, 1.0)
    q = J[cell] / (HCRYD * nu) * 4 * PI_REF
    emi = np.sum(np.where(on, Yhat * q * dt.Cabs[s, a] / (1e-8 * r * r), 0.0))
    kT, e8 = 6.336e-6 * Te, 1.1045e-15 * Te / PI_REF
    eta = lambda x: 1 - x if x <= 0 else np.exp(-x)
    rec = Ne * np.sqrt(e8 / ME) * eta(-V / kT)
    X0, H_ = m.state["ionDen"][cell], m.state["Hden"][cell]
    for e, z in enumerate(m.elems):
        for st in range(2, min(z + 1, m.nstages) + 1):
            rec -= (X0[e, st - 1] * m.abund[e] * H_ * np.sqrt(e8 / (D.A_WEIGHT[z - 1] * AMU)) *
                    eta((st - 1) * V / kT))
    return emi, rec


def test_grain_potential_balances_charging_currents(oracle, pe_model):
    m = pe_model
    J = dilute_field(oracle, m)
    up, _ = pe_update(oracle, m, J)
    nG = m.dust.nSpecies * m.dust.nSizes
    checked = 0
    for c in range(1, m.grid.nCells + 1):
        if up["nIterT"][c] >= 20:        # Te moved after the last grain solve
            continue
        for g in range(nG):
            V = up["grainPot"][g, c]
            emi, rec = charge_rates(m, J, c, g, V, up["Te"][c], up["Ne"][c])
            assert abs(emi - rec) < 0.005 * 0.5 * max(emi + rec, 1e-35), (c, g, V, emi, rec)
            assert V > 0.0                 # a UV-lit grain charges positive
            checked += 1
    assert checked >= nG * m.grid.nCells // 2
    # hot gas heats 50 K grains; photoelectrons cool them
    assert np.all(up["collD"][:, 1:] > 0) and np.all(up["peD"][:, 1:] > 0)


def test_grain_without_photoemission(oracle, pe_model):
    """No photon above any work function: no photoelectric cooling of any
    grain (the sum starts at the bin below the threshold, which holds no J
    here).  Without ionising photons Te never balances, so the potential is
    not checked against the currents (Te moves after the last grain solve)."""
    m = pe_model
    nu = m.tables.nuArray
    J = dilute_field(oracle, m, nu < 0.5 * m.dust.Vn.min())
    up, _ = pe_update(oracle, m, J)
    assert np.all(up["peD"][:, 1:] == 0.0)
    assert np.all(np.isfinite(up["grainPot"][:, 1:]))
    assert np.all(up["collD"][:, 1:] != 0.0)


def test_monochromatic_photoelectric_cooling(oracle, pe_model):
    """One bin at nu0: peD = Cabs (J0 / h nu0) <E> / pi with the mean
    photoelectron energy <E> = Y (nu0 - Vn)/2, cut where the grain potential
    reflects the slower half (min(dn, (dn^2 - V^2)/dn))."""
    m = pe_model
    nu = m.tables.nuArray
    i0 = int(np.searchsorted(nu, 1.5))
    J = dilute_field(oracle, m, np.arange(nu.size) == i0)
    up, _ = pe_update(oracle, m, J)
    dt = m.dust
    for c in range(1, m.grid.nCells + 1):
        for g in range(dt.nSpecies * dt.nSizes):
            s, a = divmod(g, dt.nSizes)
            V, Vn, n0 = up["grainPot"][g, c], dt.Vn[s], nu[i0]
            assert n0 > Vn + max(V, 0)
            dn = n0 - Vn
            EY = min(0.5 * (1 - Vn / n0), 0.2) * 0.5 * min(dn, max((dn * dn - V * V) / dn, 0.0))
            want = dt.Cabs[s, a, i0] * J[c, i0] / (HCRYD * n0) * EY * HCRYD / PI_REF
            assert up["peD"][g, c] == pytest.approx(want, rel=1e-12)
            if up["nIterT"][c] < 20:
                emi, rec = charge_rates(m, J, c, g, V, up["Te"][c], up["Ne"][c])
                assert abs(emi - rec) < 0.005 * 0.5 * max(emi + rec, 1e-35)


def test_recombination_energies_are_the_ionisation_edges(pe_model):
    """The energy a recombining ion releases to a grain is its ionisation
    potential: the mesh point at the H, He and He+ edges (1, 1.807, 4 Ryd)."""
    ip = D.ip_ryd(pe_model.xsec)
    assert ip[1 * 32 + 2] == pytest.approx(1.0, rel=2e-2)
    assert ip[2 * 32 + 2] == pytest.approx(1.807, rel=2e-2)
    assert ip[2 * 32 + 3] == pytest.approx(4.0, rel=2e-2)
    assert ip[1 * 32 + 1] == 0.0 and ip[8 * 32 + 2] == pytest.approx(1.0, rel=5e-2)


def test_dust_temperature_takes_grain_heating(oracle, pe_model):
    """getDustT absorbs sum Cabs J/pi + collD - peD: equal terms cancel, net
    collisional heating warms the grains."""
    m = pe_model
    nu, w = m.tables.nuArray, m.tables.widFlx
    J = dilute_field(oracle, m)
    Td0 = tdust_of(m)
    nG, C1 = m.dust.nSpecies * m.dust.nSizes, m.grid.nCells + 1
    base = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, J, dustOnly=False)[0]
    x = np.full((nG, C1), 3.e-6)
    same = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, J, dustOnly=False, collD=x,
                                   peD=x)[0]
    np.testing.assert_allclose(same[1:], base[1:], rtol=1e-10)
    warm = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, J, dustOnly=False, collD=x,
                                   peD=np.zeros_like(x))[0]
    assert np.all(warm[1:, 1:, 1:] > base[1:, 1:, 1:])


# ---------------------------------------------------------------- GPU parity
@pytest.mark.gpu
def test_gpu_dust_opacity_pdf_temperature_match_oracle(oracle):
    from mocassin_xray_amd import MocGPU
    m = nebula.dust_shell(n=10, nbins=150, n_packets=20000)
    rng = np.random.default_rng(20261203 + 3)
    Td = tdust_of(m, rng=rng)
    Td[5, 1, 1] = 2000.0                          # above the 1400 K sublimation point
    nu, w = m.tables.nuArray, m.tables.widFlx
    print(f"case dust_shell10: cells={m.grid.nCells} bins={nu.size}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.upload("Tdust", Td)
        g.dustOpacity()
        op, ab, sc = (g.download(k) for k in ("opacity", "absOpac", "scaOpac"))
        g.setDustPDF()
        pdf = g.download("dustPDF")
        C1 = m.grid.nCells + 1
        J = np.zeros((C1, nu.size))
        for c in range(1, C1):
            J[c] = 4 * np.pi * oracle.planck_h(nu, 30.0 + 7.0 * (c % 40)) * H * FR1RYD * w
        J[7] = 0.0
        g.upload("Jste", J)
        st = g.getDustT()
        Tg = g.download("Tdust")
        conv = g.download_flags("lgConverged")
        black = g.download_flags("lgBlack")
    op_o, ab_o, sc_o = oracle.dust_opacity(m.dust, nu, w, m.Ndust, Td)
    np.testing.assert_allclose(op[1:], op_o[1:], rtol=1e-13, atol=0)
    np.testing.assert_allclose(ab[1:], ab_o[1:], rtol=1e-13, atol=0)
    np.testing.assert_allclose(sc[1:], sc_o[1:], rtol=1e-13, atol=0)
    pdf_o = oracle.dust_pdf(m.dust, nu, w, m.Ndust, Td)
    np.testing.assert_allclose(pdf[1:], pdf_o[1:], rtol=1e-11, atol=1e-14)
    To, conv_o, black_o, nohit_o, nconv_o = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td, J)
    np.testing.assert_allclose(Tg[1:], To[1:], rtol=1e-10)
    assert st["noHit"] == nohit_o == 1 and np.all(black == black_o)
    assert np.all(conv[1:] == conv_o[1:]) and st["converged"] == nconv_o


@pytest.mark.gpu
def test_gpu_dust_transport_matches_oracle(oracle):
    from mocassin_xray_amd import MocGPU
    m = nebula.dust_shell(n=10, nbins=150, n_packets=20000, tau=10.0)
    C1 = m.grid.nCells + 1
    nu, w = m.tables.nuArray, m.tables.widFlx
    Td = tdust_of(m, rng=np.random.default_rng(5))
    n, seed = 20000, 20261203
    print(f"case dust_shell10 tau10: n={n} cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        g.upload("Tdust", Td)
        g.dustOpacity()
        g.setDustPDF()
        st = g.energyPacketDriver(1, n, seed=seed)
        J, E = g.download("Jste"), g.download("escaped")
    op, ab, sc = oracle.dust_opacity(m.dust, nu, w, m.Ndust, Td)
    pdf = oracle.dust_pdf(m.dust, nu, w, m.Ndust, Td)
    fields = {"opacity": op, "Ne": np.zeros(C1), "recPDF": pdf, "totalLines": np.zeros(C1),
              "scaOpac": sc, "dustOnly": True}
    Jo, Eo, co, _ = oracle.transport(m.tables, m.grid, fields, n, seed, nthreads=8)
    assert co["errors"] == 0 and co["absorbed"] > n and co["scattered"] > n
    for k in ("segments", "emissions", "absorbed", "scattered", "escaped", "lines", "lost"):
        assert abs(st[k] - co[k]) <= 2e-3 * max(co[k], 1.0) + 1, (k, st[k], co[k])
    l1 = np.abs(J[1:] - Jo[1:]).sum() / Jo[1:].sum()
    assert l1 < 1e-3, l1
    dE = m.tables.Lstar[0] / m.tables.nPhotons[0]
    assert E.sum() == pytest.approx(st["escaped"] * dE, rel=1e-12)


@pytest.mark.gpu
def test_gpu_dust_only_iterations_match_oracle_chain(oracle):
    from mocassin_xray_amd import MocGPU
    m = nebula.dust_shell(n=8, nbins=120, n_packets=30000)
    nu, w = m.tables.nuArray, m.tables.widFlx
    C1 = m.grid.nCells + 1
    print(f"case dust_shell8: two iterations, cells={m.grid.nCells}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        Td0 = g.download("Tdust")
        st1 = g.iterateMC(1, 99)
        J1, Td1 = g.download("Jste"), g.download("Tdust")
        st2 = g.iterateMC(2, 99)
        Td2 = g.download("Tdust")
    # iteration 1 on the oracle from the same start
    op, ab, sc = oracle.dust_opacity(m.dust, nu, w, m.Ndust, Td0)
    pdf = oracle.dust_pdf(m.dust, nu, w, m.Ndust, Td0)
    fields = {"opacity": op, "Ne": np.zeros(C1), "recPDF": pdf, "totalLines": np.zeros(C1),
              "scaOpac": sc, "dustOnly": True}
    Jo = oracle.transport(m.tables, m.grid, fields, 30000, 99, ctr_base=257, nthreads=8)[0]
    Jo = Jo * 1.e-9 / 8.0
    assert np.abs(J1[1:] - Jo[1:]).sum() / Jo[1:].sum() < 2e-3
    To = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, J1)[0]
    np.testing.assert_allclose(Td1[1:], To[1:], rtol=1e-10)
    # the grains warm from the 50 K start and the second iteration moves them less
    d1 = np.abs(Td1[1:, 0, 0] - 50.0).mean()
    d2 = np.abs(Td2[1:, 0, 0] - Td1[1:, 0, 0]).mean()
    assert d2 < d1
    assert st2["convPercent"] >= st1["convPercent"]


def small_dusty(n_packets):
    return nebula.gas_dust_model("dusty6", 6, 1.e18, 300., nebula.NEBULA_ELEMS,
                                 10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 1.001e-5, 25.0,
                                 5.e4, 10 ** 1.5, n_packets, ["sil-dlee", "grph1-dl"], [0.5, 0.5],
                                 [0.01, 0.05, 0.2], [1.0, 0.3, 0.05], 0.01)


@pytest.mark.gpu
def test_gpu_gas_dust_iteration_matches_oracle_chain(oracle):
    """Gas + grains (noPhotoelectric): kappa = gas + dust, the dust emission in
    the diffuse CDF, dust scattering/absorption of packets, then the gas balance
    and the grain temperatures on the device's J."""
    from mocassin_xray_amd import MocGPU
    m = small_dusty(20000)
    s0 = m.state
    nu, w = m.tables.nuArray, m.tables.widFlx
    seed = 31
    print(f"case gas+dust dusty6: cells={m.grid.nCells} bins={nu.size}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        Td0 = g.download("Tdust")
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "absOpac", "scaOpac", "recPDF", "totalLines",
                                          "Jste", "Te", "Ne", "ionDen", "Tdust")}
    E, Z, NS = m.elems, m.abund, m.nstages
    gas = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"],
                         s0["Te"], s0["Hden"])
    op_o, ab_o, sc_o = oracle.dust_opacity(m.dust, nu, w, m.Ndust, Td0, gasOpac=gas)
    np.testing.assert_allclose(got["opacity"][1:], op_o[1:], rtol=1e-11, atol=1e-300)
    np.testing.assert_allclose(got["scaOpac"][1:], sc_o[1:], rtol=1e-13, atol=0)
    assert np.all(got["absOpac"][1:] > 0)
    em = oracle.emission(m.et, m.bt, E, Z, NS, s0["Te"], s0["Ne"], s0["ionDen"], nthreads=8,
                         dust=(m.dust, Td0, m.Ndust, s0["Hden"]))
    np.testing.assert_allclose(got["recPDF"][1:], em["recPDF"][1:], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(got["totalLines"][1:], em["totalLines"][1:], rtol=1e-11)
    fields = {"opacity": got["opacity"], "Ne": s0["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"], "scaOpac": got["scaOpac"], "dustOnly": False}
    Jo, Eo, co, qo = oracle.transport(m.tables, m.grid, fields, 20000, seed, ctr_base=257,
                                      nthreads=8)
    assert co["errors"] == 0 and co["scattered"] > 0 and co["absorbed"] > 0
    Jo = Jo * 1.e-9 / 8.0
    assert np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum() < 2e-3
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8)
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    To = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, got["Jste"], dustOnly=False)[0]
    np.testing.assert_allclose(got["Tdust"][1:], To[1:], rtol=1e-10)
    assert st["convPercent"] == pytest.approx(100.0 * up["lgConverged"][1:].mean())


@pytest.mark.gpu
def test_gpu_photoelectric_gas_dust_iteration_matches_oracle_chain(oracle):
    """Gas + grains with grain charging, photoelectric heating and gas-grain
    collisions: the balance's grain terms (charge solve per grain inside every
    T iteration) and the grain temperatures that take them, on the device's J."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_dust_model("dustype6", 6, 1.e18, 300., nebula.NEBULA_ELEMS,
                              10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 1.001e-5, 25.0,
                              5.e4, 10 ** 1.5, 20000, ["sil-dlee", "grph1-dl"], [0.5, 0.5],
                              [0.01, 0.05, 0.2], [1.0, 0.3, 0.05], 0.01, photoelectric=True)
    s0 = m.state
    nu, w = m.tables.nuArray, m.tables.widFlx
    print(f"case photoelectric gas+dust: cells={m.grid.nCells} bins={nu.size}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        Td0 = g.download("Tdust")
        st = g.iterateMC(1, 37)
        got = {k: g.download(k) for k in ("Jste", "Te", "Ne", "ionDen", "Tdust", "grainColl",
                                          "grainPE")}
    E, Z, NS = m.elems, m.abund, m.nstages
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8, grains=(m.dust, Td0, m.Ndust))
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["Ne"][1:], up["Ne"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    np.testing.assert_allclose(got["grainColl"][:, 1:], up["collD"][:, 1:], rtol=1e-9)
    np.testing.assert_allclose(got["grainPE"][:, 1:], up["peD"][:, 1:], rtol=1e-9)
    assert np.all(got["grainPE"][:, 1:] > 0)
    To = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, got["Jste"], dustOnly=False,
                                 collD=up["collD"], peD=up["peD"])[0]
    np.testing.assert_allclose(got["Tdust"][1:], To[1:], rtol=1e-10)
    assert st["convPercent"] == pytest.approx(100.0 * up["lgConverged"][1:].mean())


@pytest.mark.gpu
def test_gpu_dusty_lognormal_nebula_matches_oracle_chain(oracle):
    """The dusty-nebula workload at 6^3: seeded log-normal n_H (mean 300,
    sigma_ln 1), silicate and graphite at 10 MRN sizes each (20 grains, one
    photoelectric grain per lane), grain charging and photoelectric heating
    on, every n-level ion at its file's level count; one moc_iterate call
    against the oracle at every stage."""
    from mocassin_xray_amd import MocGPU
    m = nebula.gas_dust_nebula(n=6, nbins=300, n_packets=20000)
    assert m.dust.nSizes == 10 and m.dust.nSpecies == 2 and m.bt.maxLev == 27
    s0 = m.state
    hd = s0["Hden"][1:]
    assert hd.min() < 100.0 and hd.max() > 1000.0        # the densities vary cell to cell
    nu, w = m.tables.nuArray, m.tables.widFlx
    seed = 43
    print(f"case dustneb6 lognormal 20 grains: cells={m.grid.nCells} bins={nu.size}", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        Td0 = g.download("Tdust")
        st = g.iterateMC(1, seed)
        got = {k: g.download(k) for k in ("opacity", "scaOpac", "recPDF", "totalLines", "Jste",
                                          "Te", "Ne", "ionDen", "Tdust", "grainColl", "grainPE")}
    E, Z, NS = m.elems, m.abund, m.nstages
    gas = oracle.opacity(m.xsec, m.tables.KNsigmaT, True, E, Z, NS, s0["ionDen"], s0["Ne"],
                         s0["Te"], s0["Hden"])
    op_o, ab_o, sc_o = oracle.dust_opacity(m.dust, nu, w, m.Ndust, Td0, gasOpac=gas)
    np.testing.assert_allclose(got["opacity"][1:], op_o[1:], rtol=1e-11, atol=1e-300)
    em = oracle.emission(m.et, m.bt, E, Z, NS, s0["Te"], s0["Ne"], s0["ionDen"], nthreads=8,
                         dust=(m.dust, Td0, m.Ndust, s0["Hden"]))
    np.testing.assert_allclose(got["recPDF"][1:], em["recPDF"][1:], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(got["totalLines"][1:], em["totalLines"][1:], rtol=1e-11)
    fields = {"opacity": got["opacity"], "Ne": s0["Ne"], "recPDF": got["recPDF"],
              "totalLines": got["totalLines"], "scaOpac": got["scaOpac"], "dustOnly": False}
    Jo, Eo, co, qo = oracle.transport(m.tables, m.grid, fields, 20000, seed, ctr_base=257,
                                      nthreads=8)
    assert co["errors"] == 0 and co["scattered"] > 0
    Jo = Jo * 1.e-9 / 8.0
    assert np.abs(got["Jste"][1:] - Jo[1:]).sum() / Jo[1:].sum() < 2e-3
    up = oracle.update(m.bt, E, Z, NS, True, got["Jste"], s0["Te"], s0["Ne"], s0["Hden"],
                       s0["ionDen"], nthreads=8, grains=(m.dust, Td0, m.Ndust))
    np.testing.assert_allclose(got["Te"][1:], up["Te"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["Ne"][1:], up["Ne"][1:], rtol=1e-9)
    np.testing.assert_allclose(got["ionDen"][1:], up["ionDen"][1:], rtol=1e-8, atol=1e-25)
    np.testing.assert_allclose(got["grainColl"][:, 1:], up["collD"][:, 1:], rtol=1e-9)
    np.testing.assert_allclose(got["grainPE"][:, 1:], up["peD"][:, 1:], rtol=1e-9)
    To = oracle.dust_temperature(m.dust, nu, w, m.Ndust, Td0, got["Jste"], dustOnly=False,
                                 collD=up["collD"], peD=up["peD"])[0]
    np.testing.assert_allclose(got["Tdust"][1:], To[1:], rtol=1e-10)
    assert st["convPercent"] == pytest.approx(100.0 * up["lgConverged"][1:].mean())


def test_getqs_rayleigh_gans_branch_is_the_mie_limit():
    """Where |m| x >= 1000 and |m - 1| x < 0.001 getQs takes the Rayleigh-Gans
    forms Q_abs = 8/3 Im(m) x and Q_sca = 32 |m-1|^2 x^4 / (27 + 16 x^2)
    (ph_mod.f90:1640-1645): the weak-contrast limit that the Mie series itself
    reaches there (2 x |m - 1| << 1)."""
    from mocassin_xray_amd import dust as D
    x = 1000.0
    m = 1.0 + 1e-7 + 1e-7j
    lam = 1.0
    a = x * lam / (2.0 * D.PI_GETQS)
    qa, qs = D.get_qs(np.array([m.real]), np.array([m.imag]), a, np.array([lam]))
    assert qa[0] == pytest.approx(8.0 / 3.0 * 1e-7 * x, rel=1e-9)
    qe_m, qs_m = D.mie_q(np.array([x]), np.array([m]))
    assert qa[0] == pytest.approx(qe_m[0] - qs_m[0], rel=0.02)
    assert qs[0] == pytest.approx(qs_m[0], rel=0.05)


def test_getqs_mie_branch_caps_the_size_parameter():
    """In the Mie regime the series runs at min(x, 100)."""
    from mocassin_xray_amd import dust as D
    m = 1.7 + 0.05j
    lam = np.array([1.0, 1.0])
    a150 = 150.0 / (2.0 * D.PI_GETQS)
    qa, qs = D.get_qs(np.array([m.real] * 2), np.array([m.imag] * 2), a150, lam)
    qe, qss = D.mie_q(np.array([100.0]), np.array([m]))
    assert qa[0] == pytest.approx(qe[0] - qss[0], rel=1e-12)
    assert qs[0] == pytest.approx(qss[0], rel=1e-12)


def test_geometric_optics_saturates_extinction():
    """Large, strongly absorbing grains (|m| x >= 1000): Q_sca is the
    saturated Rayleigh-Gans Q minus Q_abs, so Q_abs + Q_sca -> 2 (the
    extinction paradox), and Q_abs lies in (0, 1)."""
    from mocassin_xray_amd import dust as D
    lam = 0.01
    for mre, mim in ((1.5, 0.3), (2.0, 1.0), (1.2, 0.02)):
        a = 3000.0 * lam / (2.0 * D.PI_GETQS)
        qa, qs = D.get_qs(np.array([mre]), np.array([mim]), a, np.array([lam]))
        assert 0.0 < qa[0] < 1.0
        assert qa[0] + qs[0] == pytest.approx(2.0, abs=1e-6)


def test_xray_grain_tables_switch_regimes():
    """0.25 um silicate grains on an X-ray mesh: the short wavelengths leave
    the Mie regime; every cross-section stays finite and positive."""
    from mocassin_xray_amd import dust as D
    from mocassin_xray_amd import tables as T
    nu = np.geomspace(0.01, 700.0, 300)
    wid = np.gradient(nu)
    n_re, n_im = D.refractive_index_on_mesh("sil-dlee", nu)
    lam = D.C_LIGHT / (nu * D.FR1RYD) * 1.e4
    x = 2.0 * D.PI_GETQS * 0.25 / lam
    assert np.any(np.sqrt(n_re ** 2 + n_im ** 2) * x >= 1000.0)
    dt = D.build_dust(nu, wid, ["sil-dlee"], [1.0], np.array([0.01, 0.25]), np.array([1.0, 1.0]))
    assert np.all(np.isfinite(dt.Cabs)) and np.all(dt.Cabs > 0.0)
    assert np.all(np.isfinite(dt.Csca)) and np.all(dt.Csca >= 0.0)


# ---------------------------------------------------------------- equivalent tau
def test_equivalent_tau_formula(tmp_path):
    """ln(SEDnoExt / SED) where both are positive, -1 elsewhere
    (output_mod.f90:3026-3031); the file's wavelength column is c / (nu fr1Ryd)
    in um: 1 Ryd is 0.0911 um."""
    from mocassin_xray_amd import outputs
    a = np.array([2.0, 1.0, 0.0, 3.0])
    b = np.array([1.0, 1.0, 1.0, 0.0])
    t = outputs.equivalent_tau(a, b)
    assert np.allclose(t, [np.log(2.0), 0.0, -1.0, -1.0])
    outputs.write_equivalent_tau(tmp_path / "equivalentTau.out", np.array([1.0, 2.0, 3.0, 4.0]), a, b)
    rows = np.loadtxt(tmp_path / "equivalentTau.out")
    assert rows.shape == (4, 4)
    assert abs(rows[0, 1] - 0.0911267) < 1e-6 and np.allclose(rows[:, 2], t)


@pytest.mark.gpu
def test_gpu_equivalent_tau_first_iteration_without_dust_opacity():
    """lgEquivalentTau (iteration_mod.f90:200-203): the first iteration of a
    dust shell sees no grain opacity -- every packet escapes, the SED is the
    source's -- and the second sees the grains, so the equivalent tau of the
    bins the grains absorb is positive."""
    from mocassin_xray_amd import MocGPU, outputs
    m = nebula.dust_shell(n=8, nbins=120, n_packets=20000)
    print("case equivalent tau: dust shell 8^3, 120 bins, 20000 packets", flush=True)
    with MocGPU(m.tables, "fp64") as g:
        g.grid_attach(m.grid)
        g.set_dust(m.dust, lgGas=False, XHILimit=m.XHILimit, equivalent_tau=True)
        g.upload("Ndust", m.Ndust)
        s1 = g.iterateMC(1, 5)
        sed1, tot1 = g.sed()
        s2 = g.iterateMC(2, 5)
        sed2, tot2 = g.sed()
    assert s1["escaped"] == s1["emissions"] and s1["escaped"] > 0
    L = float(np.atleast_1d(m.tables.Lstar)[0])
    assert abs(tot1 - L) <= 1e-9 * L               # every packet's energy escapes unabsorbed
    assert abs(tot2 - tot1) <= 1e-3 * tot1         # grains re-emit what they absorb
    assert s2["escaped"] < s2["emissions"]         # grains absorb and re-emit
    tau = outputs.equivalent_tau(sed1, sed2)
    # the source's strongest bins lose energy to the grains (which re-emit it
    # at lower frequencies, where tau turns negative)
    top = np.argsort(sed1)[-10:]
    assert (tau[top] > 0).all(), tau[top]

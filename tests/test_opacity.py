"""Opacity: pins of the cross-section stack and the oracle (CPU), and GPU
parity of the opacity kernel against the oracle."""
import numpy as np
import pytest

from mocassin_xray_amd import atomic
from mocassin_xray_amd import tables as T

ELEMS7 = [1, 2, 6, 7, 8, 10, 16]


def test_verner_threshold_cross_sections():
    # published threshold values (Verner et al. 1996): H I 6.30e-18, He I 7.4e-18,
    # He II 1.58e-18 (= H I / Z^2), O I 2p ~ 2.9e-18 ... within the fits' accuracy
    assert atomic.phfit_el(1, 1, 1, 13.6057) * 1e-18 == pytest.approx(6.30e-18, rel=0.01)
    assert atomic.phfit_el(2, 2, 1, 24.59) * 1e-18 == pytest.approx(7.40e-18, rel=0.01)
    assert atomic.phfit_el(2, 1, 1, 54.42) * 1e-18 == pytest.approx(6.30e-18 / 4, rel=0.01)
    # hydrogenic scaling sigma ~ 1/Z^2 at threshold for H-like ions
    for z in (6, 8, 10):
        thr = atomic.verner()[0][0, z - 1, 0, 0]
        s = atomic.phfit_el(z, 1, 1, thr * 1.0001) * 1e-18
        assert s == pytest.approx(6.30e-18 / z ** 2, rel=0.05)
    # below threshold: zero; far above: ~E^-3 (hydrogenic)
    assert atomic.phfit_el(1, 1, 1, 13.0) == 0.0
    r = atomic.phfit_el(1, 1, 1, 1000.) / atomic.phfit_el(1, 1, 1, 500.)
    assert 2 ** -3.6 < r < 2 ** -2.9


def test_xsec_stack_species_and_ranges():
    nu = T.log_mesh(300, 1.0, 100.0)
    xs = atomic.build_xsec(nu, ELEMS7, 7)
    assert len(xs.species) == 1 + 2 + 6 + 7 + 7 + 7 + 7
    assert xs.sigma.shape == (len(xs.species), 300)
    # every shell lies in the mesh and its values sum into its species row
    tot = np.zeros_like(xs.sigma)
    idx = {s: i for i, s in enumerate(xs.species)}
    for s in xs.shells:
        assert 1 <= s.lo <= s.hi <= 300
        tot[idx[(s.elem, s.ion)], s.lo - 1:s.hi] += s.sigma
    assert np.array_equal(tot, xs.sigma)
    # H I row: zero below 1 Ryd bin, equal to the Verner fit above
    h = xs.sigma[idx[(1, 1)]]
    E = np.maximum(nu * atomic.RYD_TO_EV, atomic.verner()[0][0, 0, 0, 0])
    assert h[5] == pytest.approx(atomic.phfit_el(1, 1, 1, E[5]) * 1e-18, rel=1e-14)


def random_cells(C1, ne, ns, elem_z, seed=1):
    rng = np.random.default_rng(seed)
    ion = rng.random((C1, ne, ns)) ** 3
    for e, z in enumerate(elem_z):
        ion[:, e, min(z + 1, ns):] = 0.0        # stages beyond the bare nucleus
    ion /= ion.sum(axis=2, keepdims=True)
    ion[0] = 0.0
    Te = rng.uniform(4.e3, 3.e4, C1)
    H = rng.uniform(10., 1.e4, C1)
    Ne = H * rng.uniform(0.1, 1.2, C1)
    return ion, Ne, Te, H


def test_oracle_neutral_hydrogen_is_nH_sigma(oracle):
    nu = T.log_mesh(120, 1.0, 30.0)
    xs = atomic.build_xsec(nu, [1], 2)
    C1 = 5
    ion = np.zeros((C1, 1, 2))
    ion[1:, 0, 0] = 1.0
    H = np.array([0., 10., 100., 1000., 1.e4])
    kn, _, _ = T.compton_tables(nu)
    k = oracle.opacity(xs, kn, True, [1], [1.0], 2, ion, np.zeros(C1), np.full(C1, 1.e4), H)
    want = H[:, None] * xs.sigma[0][None, :]
    assert np.allclose(k, want, rtol=1e-14, atol=0)


def test_oracle_free_free_and_electron_scattering_terms(oracle):
    """Fully ionised He-free gas: kappa = n_e sigma_KN + ff(He+ only, none here)
    -- the reference's density array stops at stage min(Z, nstages), so the
    bare protons do not enter the free-free term."""
    nu = T.log_mesh(80, 1.0, 30.0)
    xs = atomic.build_xsec(nu, [1], 2)
    C1 = 3
    ion = np.zeros((C1, 1, 2))
    ion[1:, 0, 1] = 1.0
    H = np.array([0., 100., 1000.])
    kn, _, _ = T.compton_tables(nu)
    k = oracle.opacity(xs, kn, True, [1], [1.0], 2, ion, H.copy(), np.full(C1, 1.e4), H)
    assert np.allclose(k[1:], H[1:, None] * kn[None, :], rtol=1e-14)
    # with He+ present the free-free term is positive and scales as n_e n_He+ / sqrt(T)
    xs2 = atomic.build_xsec(nu, [1, 2], 3)
    ion2 = np.zeros((C1, 2, 3))
    ion2[1:, 0, 1] = 1.0
    ion2[1:, 1, 1] = 1.0
    a = oracle.opacity(xs2, kn, False, [1, 2], [1.0, 0.1], 3, ion2, H.copy(), np.full(C1, 1.e4), H)
    below = nu < 3.9                                  # under the He II edge: pure free-free
    ff = a[1:, below]
    assert np.all(ff > 0) and np.all(xs2.sigma[2][below] == 0)
    assert np.allclose(ff[1] / ff[0], 100.0, rtol=1e-12)


def test_gaunt_factor_chebyshev_domain_meets_its_asymptote(oracle):
    """Inside the fit's domain (|txu| <= 2) the Hummer series is smooth and
    of order unity; at its low-frequency edge (u = h nu / kT -> 1e-4) it meets
    the classical limit 0.5513 ln(2.2459/u) (small gamma^2) that the routine
    switches to below the edge."""
    for lt in (3.5, 4.0, 4.5, 5.0):
        xlrkt = 5.1983649 - lt
        # txu = 0.72727273 (log10 nu + xlrkt) + 0.90909091: domain |txu| <= 2
        lo = (-2.0 - 0.90909091) / 0.72727273 - xlrkt
        hi = (2.0 - 0.90909091) / 0.72727273 - xlrkt
        lnu = np.linspace(lo + 1e-6, hi - 1e-6, 400)
        g = oracle.gaunt_ff(0.0, lt, lnu)
        assert np.all((g > 0.8) & (g < 6.0))
        assert np.max(np.abs(np.diff(g))) < 0.02
        gamma2 = 10 ** (0.66666667 * xlrkt * 1.5)
        u_edge = 10 ** (lo + xlrkt)
        edge = g[0]
        if gamma2 < 0.3:
            assert edge == pytest.approx(0.551329 * np.log(2.24593 / u_edge), rel=0.05)
        else:
            assert edge == pytest.approx(0.551329 * np.log(0.944931 / u_edge / np.sqrt(gamma2)),
                                         rel=0.05)


@pytest.mark.gpu
def test_gpu_opacity_matches_oracle(oracle):
    from mocassin_xray_amd import CartesianGrid, MocGPU, nebula
    n = 10
    ax = T.symmetric_axes(n, 1.e18)
    act, nc, hcell = T.active_map(ax, ax, ax, np.full((n, n, n), 100.))
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, nc, True, 0.0)
    nu = T.log_mesh(333, 1.0, 100.0)
    tabs = nebula.source_tables(nu, 1.e5, 10., 1000, True, n, n, n)
    ns = 7
    xs = atomic.build_xsec(nu, ELEMS7, ns)
    ab = np.array([1.0, 0.1, 10 ** -3.6, 10 ** -4.1, 10 ** -3.3, 10 ** -4.1, 10 ** -5.0])
    C1 = nc + 1
    ion, Ne, Te, H = random_cells(C1, len(ELEMS7), ns, ELEMS7)
    print(f"case opacity: cells={nc} bins=333 species={len(xs.species)}", flush=True)
    with MocGPU(tabs, "fp64") as g:
        g.grid_attach(grid)
        g.set_atomic(ELEMS7, ab, ns, xs)
        g.upload("ionDen", ion)
        g.upload("Ne", Ne)
        g.upload("Te", Te)
        g.upload("Hden", H)
        g.ionizationDriver()
        k = g.download("opacity")
        assert np.array_equal(g.download("ionDen"), ion)
    ko = oracle.opacity(xs, tabs.KNsigmaT, True, ELEMS7, ab, ns, ion, Ne, Te, H)
    assert np.all(k[0] == 0)
    np.testing.assert_allclose(k[1:], ko[1:], rtol=1e-11, atol=0)

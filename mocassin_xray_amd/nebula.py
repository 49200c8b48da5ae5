"""Synthetic model set-ups of the benchmark workloads.

Each builder returns the grid, the mesh/source tables and the per-cell fields
the transport consumes.  `frozen_*` builders fill the per-cell fields from a
closed-form ionisation structure (a neutral-hydrogen profile with a
hydrogenic cross-section and a recombination-like diffuse spectrum) so the
transport can be exercised and timed on its own at the real shapes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import tables as T
from .mocgpu import CartesianGrid, SourceTables

A0 = 6.30e-18          # H0 photoionisation cross-section at threshold [cm^2] (constants_mod.f90:10)
H_EDGE = 0.999434      # ionEdge of H I [Ryd] (grid_mod.f90:194)


@dataclass
class FrozenModel:
    name: str
    tables: SourceTables
    grid: CartesianGrid
    fields: dict           # name -> array in MOCASSIN shape
    n_packets: int


def _star_at_origin(sym: bool, nx, ny, nz):
    if sym:
        return np.zeros((1, 3)), np.array([[1, 1, 1]], dtype=np.int32)
    return np.zeros((1, 3)), np.array([[(nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2]],
                                      dtype=np.int32)


def source_tables(nu, T_star, L36, n_packets, sym, nx, ny, nz, lgCompton=True,
                  ion_edge1=H_EDGE, flux=None) -> SourceTables:
    wid = T.widths_of(nu)
    f = T.blackbody_flux(nu, T_star) if flux is None else flux
    cdf = T.source_cdf(f, wid)
    kn_t, kn_cdf, pc = T.compton_tables(nu)
    pos, idx = _star_at_origin(sym, nx, ny, nz)
    return SourceTables(nuArray=nu, widFlx=wid, ionEdge1=ion_edge1, KNsigmaT=kn_t, KNcdf=kn_cdf,
                        Pcomp=pc, sourceCDF=cdf[None, :], Lstar=np.array([L36]),
                        nPhotons=np.array([float(n_packets)]), starPosition=pos,
                        starIndex=idx, lgCompton=lgCompton)


def frozen_fields(grid: CartesianGrid, nu: np.ndarray, hden_cell: np.ndarray, R_s: float,
                  x_in: float = 1.e-3, T_e: float = 1.e4, lines: float = 0.3,
                  kn_sigma: np.ndarray | None = None, rng_seed: int = 7) -> dict:
    """Per-cell opacity / Ne / recPDF / totalLines of a Stromgren-like sphere
    (or, on a plane-parallel grid, a slab ionised to the depth R_s along y).

    Neutral fraction x_in inside R_s and 1 outside (smoothed over a shell);
    kappa = n_H0 * A0 (nu/nu0)^-3 above the H edge (+ n_e sigma_KN);
    diffuse emission ~ exp(-(nu - nu0) Te1Ryd/T_e) above the edge plus a flat
    sub-edge continuum; totalLines = `lines` everywhere.
    """
    C1 = grid.nCells + 1
    B = nu.size
    xa, ya, za = grid.xAxis, grid.yAxis, grid.zAxis
    X, Y, Z = np.meshgrid(xa, ya, za, indexing="ij")
    r = Y if grid.lgPlaneIonization else np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    on = grid.active > 0
    rcell = np.zeros(C1)
    rcell[grid.active[on]] = r[on]
    width = 0.05 * R_s
    xh0 = x_in + (1.0 - x_in) / (1.0 + np.exp(-(rcell - R_s) / width))
    xh0[0] = 0.0
    nH0 = xh0 * hden_cell
    ne = (1.0 - xh0) * hden_cell
    ne[0] = 0.0
    sig = np.where(nu >= H_EDGE, A0 * (nu / H_EDGE) ** -3.0, 0.0)
    kap = nH0[:, None] * sig[None, :]
    if kn_sigma is not None:
        kap = kap + ne[:, None] * kn_sigma[None, :]
    kap[0, :] = 0.0
    wid = T.widths_of(nu)
    Te1Ryd = 1.578866e5
    em = np.where(nu >= H_EDGE, np.exp(-(nu - H_EDGE) * Te1Ryd / T_e), 0.02)
    cdf = np.cumsum(em * wid)
    cdf /= cdf[-1]
    rec = np.tile(cdf, (C1, 1))
    rec[0, :] = 0.0
    tl = np.full(C1, lines)
    tl[0] = 0.0
    return {"opacity": kap, "Ne": ne, "recPDF": rec, "totalLines": tl}


def nested_grids(mother: tuple, hden_mother, subs: list, sym: bool, R_out: float = 0.0):
    """A mother grid and nested sub-grids (multiGrids).

    mother = (xa, ya, za); hden_mother [nx, ny, nz] or a scalar; subs =
    [((xa, ya, za), hden)] with each sub-grid inside the mother's box.  A
    mother cell is covered by sub-grid g (g = 2, 3, ...) when its centre lies
    inside that sub-grid's box -- strictly, except that in an octant model a
    centre on a lower face at 0 counts (fillGrid, grid_mod.f90:907-960); the
    mother's active map holds -g there, and its other cells with n_H > 0 are
    numbered as setMotherGrid does.  Returns (grids, hcell): the
    CartesianGrids, mother first, and n_H per cell of all grids in their
    global numbering (row 0 the dummy cell)."""
    xa, ya, za = (np.asarray(a, dtype=np.float64) for a in mother)
    X, Y, Z = np.meshgrid(xa, ya, za, indexing="ij")
    cover = np.zeros(X.shape, dtype=np.int32)
    for jg, ((sx, sy, sz), _) in enumerate(subs, start=2):
        def inside(V, a):
            lo = (V > a[0]) | ((V >= a[0]) & (a[0] == 0.0)) if sym else (V > a[0])
            return lo & (V < a[-1])
        cover[inside(X, sx) & inside(Y, sy) & inside(Z, sz)] = -jg
    hd = np.broadcast_to(np.asarray(hden_mother, dtype=np.float64), X.shape)
    act, nc, hm = T.active_map(xa, ya, za, np.where(cover < 0, 0.0, hd))
    act = np.where(cover < 0, cover, act).astype(np.int32)
    grids = [CartesianGrid(xa, ya, za, act, nc, sym, R_out)]
    hcell = [hm]
    for (sx, sy, sz), hs in subs:
        sx, sy, sz = (np.asarray(a, dtype=np.float64) for a in (sx, sy, sz))
        hsub = np.broadcast_to(np.asarray(hs, dtype=np.float64), (sx.size, sy.size, sz.size))
        a, n, h = T.active_map(sx, sy, sz, np.ascontiguousarray(hsub))
        grids.append(CartesianGrid(sx, sy, sz, a, n, sym, R_out))
        hcell.append(h[1:])
    return grids, np.concatenate(hcell)


@dataclass
class NestedFrozenModel:
    name: str
    tables: SourceTables
    grids: list            # the mother first
    fields: dict           # over every grid's cells, global numbering
    n_packets: int


def frozen_nested(n: int = 9, sub_n: int = 9, sub_cells: int = 4, nbins: int = 200,
                  n_packets: int = 20000, R: float = 3.e18, hden: float = 100.,
                  T_star: float = 4.e4, L36: float = 10.0, R_s: float = 1.6e18,
                  sym: bool = True, lgCompton: bool = True) -> NestedFrozenModel:
    """A frozen pure-H model on an octant (or, sym=False, a whole cube of odd
    n) of n^3 cells with a sub-grid of sub_n^3 points around the star: its box
    ends on the mother's wall after sub_cells cells from the centre (the star
    at the origin sits in it), so the sub-grid exactly replaces the mother
    cells it covers; the Stromgren-like fields of frozen_fields per cell."""
    if sym:
        ax = T.symmetric_axes(n, R)
        hi = (ax[sub_cells - 1] + ax[sub_cells]) / 2.0
        sax = np.linspace(0.0, hi, sub_n)
    else:
        ax = T.full_axes(n, R)
        c = n // 2
        hi = (ax[c + sub_cells - 1] + ax[c + sub_cells]) / 2.0
        sax = np.linspace(-hi, hi, sub_n)
    grids, hcell = nested_grids((ax, ax, ax), hden, [((sax, sax, sax), hden)], sym)
    nu = T.log_mesh(nbins, 1.0, 25.0)
    tabs = source_tables(nu, T_star, L36, n_packets, sym, n, n, n, lgCompton)
    parts = []
    for g, h in zip(grids, np.split(hcell[1:], np.cumsum([g.nCells for g in grids])[:-1])):
        parts.append(frozen_fields(g, nu, np.concatenate([[0.0], h]), R_s, kn_sigma=tabs.KNsigmaT))
    fields = {k: np.concatenate([parts[0][k][:1]] + [p[k][1:] for p in parts]) for k in parts[0]}
    return NestedFrozenModel("nested%d+%d" % (n, sub_n), tabs, grids, fields, n_packets)


def frozen_hsphere(n: int = 16, nbins: int = 400, n_packets: int = 1_000_000,
                   mesh: str = "reference", R: float = 3.e18, hden: float = 100.,
                   T_star: float = 4.e4, L36: float = 10.0, R_s: float = 2.4e18,
                   lgCompton: bool = True) -> FrozenModel:
    """Pure-hydrogen homogeneous sphere, one octant (16^3, edge 3e18 cm,
    n_H = 100, 40 kK blackbody of 1e37 erg/s, 400 bins over 1-25 Ryd)."""
    ax = T.symmetric_axes(n, R)
    hd = np.full((n, n, n), hden)
    act, ncells, hcell = T.active_map(ax, ax, ax, hd)
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    if mesh == "reference":
        nu = T.reference_gas_mesh(nbins, 1.0, 25.0, np.array([H_EDGE]))
    else:
        nu = T.log_mesh(nbins, 1.0, 25.0)
    tabs = source_tables(nu, T_star, L36, n_packets, True, n, n, n, lgCompton)
    f = frozen_fields(grid, nu, hcell, R_s, kn_sigma=tabs.KNsigmaT if lgCompton else None)
    f["Hden"] = hcell
    return FrozenModel("hsphere%d" % n, tabs, grid, f, n_packets)


def frozen_nebula(n: int = 32, nbins: int = 1000, n_packets: int = 10_000_000,
                  R: float = 1.e18, hden: float = 1000., T_star: float = 1.e5,
                  L36: float = 10 ** 1.5, R_s: float = 0.8e18) -> FrozenModel:
    """Multi-element nebula shape: 32^3 octant, n_H = 1000, 100 kK blackbody of
    10^37.5 erg/s, 1000 log bins over 1-100 Ryd."""
    ax = T.symmetric_axes(n, R)
    hd = np.full((n, n, n), hden)
    act, ncells, hcell = T.active_map(ax, ax, ax, hd)
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    nu = T.log_mesh(nbins, 1.0, 100.0)
    tabs = source_tables(nu, T_star, L36, n_packets, True, n, n, n, True)
    f = frozen_fields(grid, nu, hcell, R_s, kn_sigma=tabs.KNsigmaT)
    f["Hden"] = hcell
    return FrozenModel("nebula%d" % n, tabs, grid, f, n_packets)


# ------------------------------------------------ plane-parallel (slab) models
def slab_grid(shape: tuple, side: float, depth: float, hden: float) -> tuple:
    """A plane-parallel grid: x and z over [0, side], y (the depth, along
    which the field enters at y = 0) over [0, depth], every cell at hden and
    active (no radius cut, grid_mod.f90:1258-1266); R_out at the far corner.
    Returns (grid, hden per cell)."""
    nx, ny, nz = shape
    xa = np.linspace(0.0, side, nx)
    ya = np.linspace(0.0, depth, ny)
    za = np.linspace(0.0, side, nz)
    act, ncells, hcell = T.active_map(xa, ya, za, np.full((nx, ny, nz), hden))
    grid = CartesianGrid(xa, ya, za, act, ncells, False, T.plane_R_out(xa, ya, za),
                         lgPlaneIonization=True)
    return grid, hcell


def slab_source_tables(nu, T_star, meanField, nu0, nu0Add, grid, n_packets, lgCompton=True,
                       ion_edge1=H_EDGE, flux=None) -> SourceTables:
    """Mesh and source tables of a slab lit by a blackbody field of photon
    flux meanField [phot/cm^2/s] above nu0 (planeIonization meanField nu0
    nu0Add); the source position is unused (packets are born on the face)."""
    wid = T.widths_of(nu)
    f = T.blackbody_flux(nu, T_star) if flux is None else flux
    L36, _, cdf = T.plane_source(nu, wid, f, meanField, nu0, nu0Add, grid.xAxis, grid.zAxis)
    kn_t, kn_cdf, pc = T.compton_tables(nu)
    return SourceTables(nuArray=nu, widFlx=wid, ionEdge1=ion_edge1, KNsigmaT=kn_t, KNcdf=kn_cdf,
                        Pcomp=pc, sourceCDF=cdf[None, :], Lstar=np.array([L36]),
                        nPhotons=np.array([float(n_packets)]), starPosition=np.zeros((1, 3)),
                        starIndex=np.array([[1, 1, 1]], dtype=np.int32), lgCompton=lgCompton)


def frozen_slab(shape: tuple = (5, 16, 5), nbins: int = 200, n_packets: int = 100_000,
                side: float = 2.e18, depth: float = 6.e18, hden: float = 100.,
                T_star: float = 4.e4, phi: float = 1.e10, y_s: float = 3.8e18,
                lgCompton: bool = True) -> FrozenModel:
    """Pure-hydrogen slab: 5 x 16 x 5 cells, 2e18 cm square face, 6e18 cm
    deep, n_H = 100, 40 kK blackbody field of 1e10 ionising photons/cm^2/s,
    ionised to ~3.8e18 cm (the Stromgren depth phi / (n_H^2 alpha_B))."""
    grid, hcell = slab_grid(shape, side, depth, hden)
    nu = T.log_mesh(nbins, 1.0, 25.0)
    tabs = slab_source_tables(nu, T_star, phi, 1.0, 0.0, grid, n_packets, lgCompton)
    f = frozen_fields(grid, nu, hcell, y_s, kn_sigma=tabs.KNsigmaT if lgCompton else None)
    f["Hden"] = hcell
    return FrozenModel("slab%dx%dx%d" % shape, tabs, grid, f, n_packets)


def gas_slab(shape: tuple = (5, 16, 5), nbins: int = 400, n_packets: int = 1_000_000,
             side: float = 2.e18, depth: float = 6.e18, hden: float = 100.,
             elems: list = (1, 2, 8), log_abund: list = (0.0, -1.0, -3.3), nstages: int = 4,
             T_star: float = 4.e4, phi: float = 1.e10, maxLev: int = 27,
             lgCompton: bool = True) -> GasModel:
    """Gas slab lit on its y = 0 face: 5 x 16 x 5 cells, 2e18 cm square face,
    6e18 cm deep, n_H = 100, H He O, 40 kK blackbody field of 1e10 ionising
    photons/cm^2/s above 1 Ryd, 400 points over 1-25 Ryd."""
    from . import atomic, balance, emission
    grid, hcell = slab_grid(shape, side, depth, hden)
    nu = T.reference_gas_mesh(nbins, 1.0, 25.0, atomic.ion_edges(list(elems), nstages, 25.0))
    tabs = slab_source_tables(nu, T_star, phi, 1.0, 0.0, grid, n_packets, lgCompton)
    xs = atomic.build_xsec(nu, list(elems), nstages)
    bt = balance.build_balance(xs, maxLev=maxLev)
    et = emission.build_emission(xs, tabs.widFlx)
    st = initial_state(list(elems), nstages, hcell)
    return GasModel("slab%dx%dx%d" % tuple(shape), tabs, grid, list(elems),
                    10.0 ** np.asarray(log_abund, dtype=np.float64), nstages, xs, bt, et, st,
                    n_packets)


# ------------------------------------------------------------- full gas models
@dataclass
class GasModel:
    """A gas model ready for the Monte Carlo iteration: grid, mesh/source
    tables, atomic tables (cross-section stack, balance, emission) and the
    starting per-cell state."""
    name: str
    tables: SourceTables
    grid: CartesianGrid
    elems: list
    abund: np.ndarray
    nstages: int
    xsec: object
    bt: object
    et: object
    state: dict            # Te, Ne, Hden (C+1,), ionDen (C+1, NE, NS)
    n_packets: int

    def load(self, g) -> None:
        """Upload everything into a MocGPU context."""
        g.grid_attach(self.grid)
        g.set_atomic(self.elems, self.abund, self.nstages, self.xsec)
        g.set_balance(self.bt)
        g.set_emission(self.et)
        for k in ("ionDen", "Te", "Ne", "Hden"):
            g.upload(k, self.state[k])


def initial_state(elems: list, nstages: int, hden_cells: np.ndarray, Te: float = 1.e4,
                  H0in: float = 1.e-5) -> dict:
    """Starting guess of setGrid (grid_mod.f90:1583-1645): X(H0) = H0in, He
    split evenly between He+ and He++, heavy elements neutral at H0in and the
    rest spread evenly over the ionised stages, Te = TeStart, Ne = n_H."""
    C1 = hden_cells.size
    X = np.zeros((C1, len(elems), nstages))
    for e, z in enumerate(elems):
        if z == 1:
            X[:, e, 0], X[:, e, 1] = H0in, 1.0 - H0in
        elif z == 2:
            X[:, e, 0] = H0in
            X[:, e, 1] = X[:, e, 2] = (1.0 - H0in) / 2.0
        else:
            top = min(z + 1, nstages)
            X[:, e, 0] = H0in
            X[:, e, 1:top] = (1.0 - H0in) / min(z, nstages - 1)
    X[0] = 0.0
    Te_a = np.full(C1, Te)
    Te_a[0] = 0.0
    return {"Te": Te_a, "Ne": hden_cells.copy(), "Hden": hden_cells.copy(), "ionDen": X}


def gas_model(name: str, n: int, R: float, hden: float, elems: list, abund: np.ndarray,
              nstages: int, nbins: int, nu_max: float, T_star: float, L36: float,
              n_packets: int, maxLev: int = 27, lgCompton: bool = True) -> GasModel:
    """maxLev cuts the n-level ions at that many levels; the default, 27, is
    the largest level count of the packed atomic data, so every ion keeps the
    levels of its file as equilibrium reads them (emission_mod.f90:2505-2754)."""
    from . import atomic, balance, emission
    ax = T.symmetric_axes(n, R)
    act, ncells, hcell = T.active_map(ax, ax, ax, np.full((n, n, n), hden))
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    nu = T.reference_gas_mesh(nbins, 1.0, nu_max, atomic.ion_edges(elems, nstages, nu_max))
    tabs = source_tables(nu, T_star, L36, n_packets, True, n, n, n, lgCompton)
    xs = atomic.build_xsec(nu, elems, nstages)
    bt = balance.build_balance(xs, maxLev=maxLev)
    et = emission.build_emission(xs, tabs.widFlx)
    st = initial_state(elems, nstages, hcell)
    return GasModel(name, tabs, grid, list(elems), np.asarray(abund, dtype=np.float64), nstages,
                    xs, bt, et, st, n_packets)


def gas_hsphere(n: int = 16, nbins: int = 400, n_packets: int = 1_000_000,
                lgCompton: bool = True) -> GasModel:
    """Pure-hydrogen homogeneous sphere, one octant: 16^3, edge 3e18 cm,
    n_H = 100, 40 kK blackbody of 1e37 erg/s, 400 points over 1-25 Ryd."""
    return gas_model("hsphere%d" % n, n, 3.e18, 100., [1], np.ones(1), 2, nbins, 25.0, 4.e4,
                     10.0, n_packets, lgCompton=lgCompton)


NEBULA_ELEMS = [1, 2, 6, 7, 8, 10, 16]
NEBULA_LOG_ABUND = [0.0, -1.0, -3.6, -4.1, -3.3, -4.1, -5.0]


def gas_nebula(n: int = 32, nbins: int = 1000, n_packets: int = 10_000_000,
               nstages: int = 7, maxLev: int = 5) -> GasModel:
    """Multi-element nebula, one octant: 32^3, edge 1e18 cm, n_H = 1000,
    H He C N O Ne S at log abundances 0 -1 -3.6 -4.1 -3.3 -4.1 -5, 100 kK
    blackbody of 10^37.5 erg/s, 1000 points over 1-100 Ryd, forbidden-line
    cooling with the ions cut at 5 levels (the workload's own definition;
    maxLev=27 keeps every level of the data files, as the reference does)."""
    return gas_model("nebula%d" % n, n, 1.e18, 1000., NEBULA_ELEMS,
                     10.0 ** np.array(NEBULA_LOG_ABUND), nstages, nbins, 100.0, 1.e5,
                     10 ** 1.5, n_packets, maxLev=maxLev)


# ------------------------------------------------------------------ dust models
@dataclass
class DustModel:
    """A dust-only model (nebComposition noGas) ready for the Monte Carlo
    iteration: grid, mesh/source tables, grain tables and Ndust."""
    name: str
    tables: SourceTables
    grid: CartesianGrid
    dust: object           # dust.DustTables
    Ndust: np.ndarray      # (C+1,)
    n_packets: int
    XHILimit: float = 0.05

    def load(self, g) -> None:
        """Upload everything into a MocGPU context."""
        g.grid_attach(self.grid)
        g.set_dust(self.dust, lgGas=False, XHILimit=self.XHILimit)
        g.upload("Ndust", self.Ndust)


def shell_axes(n: int, R: float, a1: float) -> np.ndarray:
    """Octant axis 0, a1, ... R with geometric spacing after the first cell."""
    return np.concatenate([[0.0], np.geomspace(a1, R, n - 1)])


def dust_model(name: str, n: int, R: float, R_in: float, ndust: float, species: list,
               abundances, radius, weights, nbins: int, nu_min: float, nu_max: float,
               T_star: float, L36: float, n_packets: int) -> DustModel:
    from . import dust as D
    ax = shell_axes(n, R, R / 2.0 ** (n - 1) * 2.0)
    act, ncells, nd = T.active_map(ax, ax, ax, np.full((n, n, n), ndust), R_in=R_in, R_out=R)
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    nu = T.log_mesh(nbins, nu_min, nu_max)
    tabs = source_tables(nu, T_star, L36, n_packets, True, n, n, n, lgCompton=False,
                         ion_edge1=H_EDGE)
    dt = D.build_dust(nu, tabs.widFlx, species, abundances, radius, weights)
    return DustModel(name, tabs, grid, dt, nd, n_packets)


def dust_shell(n: int = 16, nbins: int = 300, n_packets: int = 1_000_000,
               tau: float = 1.0) -> DustModel:
    """Spherical silicate shell around a cool star, one octant (the layout of
    the reference's dust benchmarks p0tau1/10/100): 16^3 with geometric
    axes out to 2.18e17 cm, 0.16 um grains of sil-dlee with n_d = 1.117e-8
    tau cm^-3 beyond 1.5e14 cm, 2500 K blackbody of 3.826e37 erg/s, 300
    log points over 1e-5 - 10 Ryd."""
    return dust_model("dustshell%d" % n, n, 2.18e17, 1.5e14, 1.117e-8 * tau, ["sil-dlee"], [1.0],
                      [0.16], [1.0], nbins, 1.e-5, 10.0, 2500.0, 38.26, n_packets)


@dataclass
class GasDustModel(GasModel):
    """A gas model with grains (MdMg dust-to-gas mass ratio).  photoelectric
    switches on grain charging, photoelectric heating and gas-grain
    collisions; without it the run is the reference's noPhotoelectric one."""
    dust: object = None     # dust.DustTables
    Ndust: np.ndarray = None
    photoelectric: bool = False

    def load(self, g) -> None:
        super().load(g)
        g.set_dust(self.dust, lgGas=True,
                   photoelectric_xsec=self.xsec if self.photoelectric else None)
        g.upload("Ndust", self.Ndust)


def gas_dust_model(name: str, n: int, R: float, hden, elems: list, abund: np.ndarray,
                   nstages: int, nbins: int, nu_min: float, nu_max: float, T_star: float,
                   L36: float, n_packets: int, species: list, grain_abun, radius, weights,
                   mdmg: float, maxLev: int = 27, photoelectric: bool = False) -> GasDustModel:
    """hden is a number or an (n, n, n) array of n_H per cell."""
    from . import atomic, balance, emission
    from . import dust as D
    ax = T.symmetric_axes(n, R)
    hd = np.broadcast_to(np.asarray(hden, dtype=np.float64), (n, n, n))
    act, ncells, hcell = T.active_map(ax, ax, ax, np.ascontiguousarray(hd))
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    nu = T.reference_gas_mesh(nbins, nu_min, nu_max, atomic.ion_edges(elems, nstages, nu_max))
    tabs = source_tables(nu, T_star, L36, n_packets, True, n, n, n, True)
    xs = atomic.build_xsec(nu, elems, nstages)
    bt = balance.build_balance(xs, maxLev=maxLev)
    et = emission.build_emission(xs, tabs.widFlx)
    st = initial_state(elems, nstages, hcell)
    dt = D.build_dust(nu, tabs.widFlx, species, grain_abun, radius, weights)
    nd = D.ndust_from_mdmg(dt, mdmg, hcell, elems, abund)
    nd[0] = 0.0
    return GasDustModel(name, tabs, grid, list(elems), np.asarray(abund, dtype=np.float64),
                        nstages, xs, bt, et, st, n_packets, dust=dt, Ndust=nd,
                        photoelectric=photoelectric)


def lognormal_density(n: int, mean: float = 300.0, sigma_ln: float = 1.0,
                      seed: int = 20261203 + 3) -> np.ndarray:
    """n_H per cell drawn log-normal with the given mean (mu = ln mean - sigma^2/2)."""
    rng = np.random.default_rng(seed)
    return rng.lognormal(np.log(mean) - 0.5 * sigma_ln ** 2, sigma_ln, size=(n, n, n))


def gas_dust_nebula(n: int = 64, nbins: int = 700, n_packets: int = 10_000_000,
                    photoelectric: bool = True) -> GasDustModel:
    """Dusty nebula, one octant: n^3 with log-normal n_H (mean 300, sigma_ln 1),
    edge 1e18 cm, the seven elements of the gas nebula, dust-to-gas mass ratio
    0.01 of sil-dlee and grph1-dl (0.5 each) in 10 MRN sizes 0.005-0.25 um,
    50 kK blackbody of 10^37.5 erg/s, a gas mesh over 1e-5 - 25 Ryd.  Grain
    charging and photoelectric heating are on, as in the reference's default
    (lgPhotoelectric, common_mod.f90:424)."""
    a, w = _mrn()
    return gas_dust_model("dustneb%d" % n, n, 1.e18, lognormal_density(n), NEBULA_ELEMS,
                          10.0 ** np.array(NEBULA_LOG_ABUND), 7, nbins, 1.001e-5, 25.0, 5.e4,
                          10 ** 1.5, n_packets, ["sil-dlee", "grph1-dl"], [0.5, 0.5], a, w, 0.01,
                          photoelectric=photoelectric)


# ------------------------------------------------------------- batched grids
@dataclass
class BatchModel:
    """Independent gas nebulae in one context: one grid and one star each,
    a shared mesh, element set and tables; cells numbered grid after grid."""
    name: str
    tables: SourceTables       # nStars = number of grids, star k lights grid k
    grids: list                # CartesianGrid per member
    members: list              # SourceTables of each member alone (its star)
    elems: list
    abund: np.ndarray
    nstages: int
    xsec: object
    bt: object
    et: object
    state: dict                # concatenated: row 0, then each grid's cells
    n_packets: int             # total per iteration

    @property
    def cellBase(self) -> list:
        out, b = [], 0
        for g in self.grids:
            out.append(b)
            b += g.nCells
        return out

    def load(self, g) -> None:
        g.grid_attach_batch(self.grids)
        g.set_atomic(self.elems, self.abund, self.nstages, self.xsec)
        g.set_balance(self.bt)
        g.set_emission(self.et)
        for k in ("ionDen", "Te", "Ne", "Hden"):
            g.upload(k, self.state[k])

    def member(self, k: int) -> GasModel:
        """Grid k on its own, with its star and the slice of the state."""
        b, n = self.cellBase[k], self.grids[k].nCells
        rows = np.r_[0, np.arange(b + 1, b + n + 1)]
        st = {key: np.ascontiguousarray(v[rows]) for key, v in self.state.items()}
        return GasModel(f"{self.name}[{k}]", self.members[k], self.grids[k], self.elems,
                        self.abund, self.nstages, self.xsec, self.bt, self.et, st,
                        int(self.members[k].nPhotons[0]))


def batch_nebulae(n_grids: int = 256, nbins: int = 400, seed: int = 20261203 + 4,
                  edge=(8, 48), hden=(10.0, 1.e4), T_star=(3.e4, 1.5e5), packets=(1.e5, 1.e6),
                  L36: float = 10.0, nstages: int = 7, maxLev: int = 27) -> BatchModel:
    """n_grids seeded nebulae of uneven size, one octant each: n ~ U{edge}
    cells per axis, uniform n_H log-uniform over `hden`, a blackbody of T
    uniform over T_star and L36 x 1e36 erg/s, packets log-uniform over
    `packets`, the seven elements of the gas nebula; each grid's edge is 1.5
    Stromgren radii of 1.8e47 ionising photons/s at its n_H (alpha_B
    2.6e-13).  A shared mesh of nbins points over 1-100 Ryd."""
    from . import atomic, balance, emission
    rng = np.random.default_rng(seed)
    elems = NEBULA_ELEMS
    abund = 10.0 ** np.array(NEBULA_LOG_ABUND)
    nu = T.reference_gas_mesh(nbins, 1.0, 100.0, atomic.ion_edges(elems, nstages, 100.0))
    xs = atomic.build_xsec(nu, elems, nstages)
    bt = balance.build_balance(xs, maxLev=maxLev)
    grids, members, hcells = [], [], []
    for k in range(n_grids):
        n = int(rng.integers(edge[0], edge[1] + 1))
        nh = float(np.exp(rng.uniform(np.log(hden[0]), np.log(hden[1]))))
        ts = float(rng.uniform(T_star[0], T_star[1]))
        npk = int(np.exp(rng.uniform(np.log(packets[0]), np.log(packets[1]))))
        R = 1.5 * (3.0 * 1.8e47 / (4.0 * np.pi * nh * nh * 2.6e-13)) ** (1.0 / 3.0)
        ax = T.symmetric_axes(n, R)
        act, nc, hc = T.active_map(ax, ax, ax, np.full((n, n, n), nh))
        grids.append(CartesianGrid(ax, ax.copy(), ax.copy(), act, nc, True, 0.0))
        members.append(source_tables(nu, ts, L36, npk, True, n, n, n, True))
        hcells.append(hc)
    t0 = members[0]
    tabs = SourceTables(nuArray=nu, widFlx=t0.widFlx, ionEdge1=t0.ionEdge1, KNsigmaT=t0.KNsigmaT,
                        KNcdf=t0.KNcdf, Pcomp=t0.Pcomp,
                        sourceCDF=np.vstack([m.sourceCDF for m in members]),
                        Lstar=np.concatenate([m.Lstar for m in members]),
                        nPhotons=np.concatenate([m.nPhotons for m in members]),
                        starPosition=np.vstack([np.reshape(m.starPosition, (1, 3)) for m in members]),
                        starIndex=np.vstack([np.reshape(m.starIndex, (1, 3)) for m in members]),
                        lgCompton=True)
    et = emission.build_emission(xs, tabs.widFlx)
    hall = np.concatenate([[0.0]] + [h[1:] for h in hcells])
    st = initial_state(elems, nstages, hall)
    return BatchModel("batch%d" % n_grids, tabs, grids, members, list(elems), abund, nstages,
                      xs, bt, et, st, int(tabs.nPhotons.sum()))


# ------------------------------------------------------------- X-ray nebula
# Solar abundances H-Zn, log eps (H = 12), Anders & Grevesse (1989, Geochim.
# Cosmochim. Acta 53, 197), the photospheric values with meteoritic ones for
# Li, Be, B where the photosphere is depleted.
SOLAR_LOG_EPS = [12.00, 10.99, 3.31, 1.42, 2.88, 8.56, 8.05, 8.93, 4.48, 8.09, 6.31, 7.58, 6.48,
                 7.55, 5.57, 7.27, 5.27, 6.56, 5.13, 6.34, 3.09, 4.93, 4.02, 5.68, 5.53, 7.51,
                 4.91, 6.25, 4.29, 4.67]


def clumpy_density(n: int, fill: float = 0.1, n_clump: float = 1000.0, n_inter: float = 10.0,
                   seed: int = 20261203 + 5) -> np.ndarray:
    """n^3 cells: a fraction `fill` of them (drawn without replacement) at
    n_clump, the rest at n_inter."""
    rng = np.random.default_rng(seed)
    hd = np.full(n ** 3, n_inter)
    hd[rng.choice(n ** 3, size=int(round(fill * n ** 3)), replace=False)] = n_clump
    return hd.reshape(n, n, n)


def xray_source(nu: np.ndarray, T_bb: float = 1.e5, pl_lo: float = 0.01, pl_hi: float = 10.0,
                pl_fraction: float = 0.5) -> np.ndarray:
    """F_nu of a power law F_nu ~ nu^-1 between pl_lo and pl_hi keV plus a
    blackbody of T_bb, the power law carrying pl_fraction of the luminosity
    (a tabulated source spectrum, contShape file)."""
    ryd_kev = 13.6057e-3
    wid = T.widths_of(nu)
    pl = np.where((nu >= pl_lo / ryd_kev) & (nu <= pl_hi / ryd_kev), 1.0 / nu, 0.0)
    bb = T.blackbody_flux(nu, T_bb)
    return (pl_fraction * pl / np.sum(pl * wid) + (1.0 - pl_fraction) * bb / np.sum(bb * wid))


def xray_nebula(n: int = 128, nbins: int = 2000, n_packets: int = 1_000_000_000,
                R: float = 3.e18, L36: float = 10 ** 1.5, maxLev: int = 27) -> GasModel:
    """Clumpy X-ray-illuminated nebula, one octant: n^3 cells, 10 % of them at
    n_H = 1000 and the rest at 10, edge R; all 30 elements H-Zn at solar
    abundances with every ionisation stage (31), inner-shell photoionisation
    and Auger yields; a source of F_nu ~ nu^-1 over 0.01-10 keV plus a 1e5 K
    blackbody in equal luminosity; the gas mesh of nbins points over
    1.001e-5 - 735 Ryd with every ion edge resolved (edges that coincide
    leave a few points fewer)."""
    from . import atomic, balance, emission
    elems = list(range(1, 31))
    abund = 10.0 ** (np.array(SOLAR_LOG_EPS) - 12.0)
    nstages = 31
    nu_max = 735.0
    ax = T.symmetric_axes(n, R)
    act, ncells, hcell = T.active_map(ax, ax, ax, np.ascontiguousarray(clumpy_density(n)))
    grid = CartesianGrid(ax, ax.copy(), ax.copy(), act, ncells, True, 0.0)
    nu = T.reference_gas_mesh(nbins, 1.001e-5, nu_max, atomic.ion_edges(elems, nstages, nu_max),
                              strict=False)
    tabs = source_tables(nu, 1.e5, L36, n_packets, True, n, n, n, True, flux=xray_source(nu))
    xs = atomic.build_xsec(nu, elems, nstages)
    bt = balance.build_balance(xs, maxLev=maxLev)
    et = emission.build_emission(xs, tabs.widFlx)
    st = initial_state(elems, nstages, hcell)
    return GasModel("xray%d" % n, tabs, grid, elems, abund, nstages, xs, bt, et, st, n_packets)


def _mrn():
    from . import dust as D
    return D.mrn_sizes(10, 0.005, 0.25, 3.5)

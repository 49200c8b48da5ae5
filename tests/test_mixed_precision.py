"""The mixed-precision form (FP32 transport, FP32 rate and cooling fits, FP64
ion chain, level solves and root-find) against the all-FP64 form, gated by
the Monte Carlo noise band: the FP64 form run again with another packet
stream from the same start."""
import numpy as np
import pytest

from mocassin_xray_amd import nebula
from mocassin_xray_amd import tables as T


def run(m, prec, seed, iters=3):
    from mocassin_xray_amd import MocGPU
    with MocGPU(m.tables, prec) as g:
        m.load(g)
        for it in range(1, iters + 1):
            g.iterateMC(it, seed)
        return {k: g.download(k) for k in ("Te", "ionDen")}


def volavg(m, s):
    dV = T.cell_volumes(m.grid.xAxis, m.grid.yAxis, m.grid.zAxis, m.grid.active, True)[1:]
    w = dV * m.state["Hden"][1:]
    Te = np.sum(w * s["Te"][1:]) / w.sum()
    X = np.einsum("c,ces->es", w, s["ionDen"][1:]) / w.sum()
    return Te, X


def errors(m, a, b):
    Ta, Xa = volavg(m, a)
    Tb, Xb = volavg(m, b)
    big = Xb > 1e-3
    return abs(Ta - Tb) / Tb, float(np.max(np.abs(Xa[big] - Xb[big]) / Xb[big]))


@pytest.mark.gpu
def test_mixed_precision_within_the_monte_carlo_band():
    """After three iterations of a 12^3 multi-element nebula, the volume-
    averaged T_e and ion fractions (those above 1e-3) of the mixed form differ
    from the FP64 form's by at most 1 % or twice the difference between two
    FP64 runs with different packet streams, whichever is larger."""
    m = nebula.gas_nebula(n=12, nbins=600, n_packets=400_000, maxLev=27)
    print(f"case mixed-vs-fp64 nebula12: cells={m.grid.nCells}", flush=True)
    a = run(m, "fp64", 11)
    b = run(m, "fp64", 12)
    c = run(m, "mixed", 11)
    bandT, bandX = errors(m, b, a)
    errT, errX = errors(m, c, a)
    print(f"band Te {bandT:.2e} X {bandX:.2e}; mixed Te {errT:.2e} X {errX:.2e}", flush=True)
    assert errT <= max(0.01, 2.0 * bandT)
    assert errX <= max(0.01, 2.0 * bandX)


def fp32_kernels_vs_fp64(m, cdf_tol, kappa_tol):
    """Opacity and diffuse CDF of the same state from the FP32 kernels of the
    mixed form and the FP64 kernels; returns the largest differences."""
    from mocassin_xray_amd import MocGPU
    out = {}
    for prec in ("fp64", "mixed"):
        with MocGPU(m.tables, prec) as g:
            m.load(g)
            g.ionizationDriver()
            g.emissionDriver()
            out[prec] = {k: g.download(k) for k in ("opacity", "recPDF", "totalLines")}
    a, b = out["fp64"], out["mixed"]
    ka, kb = a["opacity"][1:], b["opacity"][1:]
    # relative to the cell's kappa in that bin; bins below 1e-30 of the cell's
    # largest kappa carry nothing the transport can see
    sig = ka > 1e-30 * ka.max(axis=1, keepdims=True)
    kap = float(np.max(np.abs(kb[sig] - ka[sig]) / ka[sig]))
    cdf = float(np.max(np.abs(b["recPDF"][1:] - a["recPDF"][1:])))
    tl = float(np.max(np.abs(b["totalLines"][1:] - a["totalLines"][1:]) /
                      np.maximum(a["totalLines"][1:], 1e-300)))
    print(f"FP32 kernels: kappa rel {kap:.2e}, CDF abs {cdf:.2e}, totalLines rel {tl:.2e}", flush=True)
    assert kap < kappa_tol
    assert cdf < cdf_tol
    assert tl < 1e-5
    assert np.all(b["recPDF"][1:, -1] == 1.0)
    assert np.all(np.diff(b["recPDF"][1:], axis=1) >= 0.0)


@pytest.mark.gpu
def test_fp32_opacity_and_continuum_match_fp64_nebula():
    """The mixed form's FP32 opacity and continuum kernels against the FP64
    ones on the starting state of a multi-element nebula: kappa within 1e-5
    relative in every bin that matters, the diffuse CDF within 2e-6 absolute
    (FP32 bin energies, FP64 scan), the line fraction within 1e-5; the CDF
    stays monotone with its last entry 1."""
    m = nebula.gas_nebula(n=8, nbins=600, n_packets=10_000, maxLev=27)
    print(f"case fp32 kernels nebula8: cells={m.grid.nCells}", flush=True)
    fp32_kernels_vs_fp64(m, cdf_tol=2e-6, kappa_tol=1e-5)


@pytest.mark.gpu
def test_fp32_opacity_and_continuum_match_fp64_xray():
    """The same on the X-ray nebula's physics (30 elements, 31 stages, a mesh
    from 1e-5 to 735 Ryd), where the free-free and heavy-ion terms span the
    widest range."""
    m = nebula.xray_nebula(n=6, nbins=2000, n_packets=10_000)
    print(f"case fp32 kernels xray6: cells={m.grid.nCells}", flush=True)
    fp32_kernels_vs_fp64(m, cdf_tol=2e-6, kappa_tol=1e-5)

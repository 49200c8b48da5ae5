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

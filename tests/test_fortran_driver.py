"""The Fortran host side: model files written for mocgpu_driver read back
exactly, and on the GPU a run of MCIterationDriver from Fortran (through the
ISO_C_BINDING module) against the same iterations driven from Python."""
import subprocess

import numpy as np
import pytest

from mocassin_xray_amd import build, fortran_io, nebula

DRIVER = build.FORTRAN / "mocgpu_driver"


def small_model(n_packets=20000):
    return nebula.gas_model("nebula6", 6, 1.e18, 1000., nebula.NEBULA_ELEMS,
                            10.0 ** np.array(nebula.NEBULA_LOG_ABUND), 7, 300, 100.0, 1.e5,
                            10 ** 1.5, n_packets)


def test_model_file_round_trip(tmp_path):
    m = small_model()
    p = tmp_path / "model.bin"
    fortran_io.write_model(p, m, "mixed", maxIterateMC=4, seed=77)
    d = fortran_io.read_dump(p)
    assert d["tables.nbins"][0] == m.tables.nbins and d["tables.precision"][0] == 1
    np.testing.assert_array_equal(d["tables.nuArray"], m.tables.nuArray)
    np.testing.assert_array_equal(d["grid.active"], m.grid.active.reshape(-1, order="F"))
    np.testing.assert_array_equal(d["balance.atomMeta"], m.bt.atomMeta8.reshape(-1))
    assert d["balance.nLev"].dtype == np.int64 and d["emission.fbOff"].dtype == np.int64
    np.testing.assert_array_equal(d["emission.edgeP.2"], m.et.edge[1])
    X = d["state.ionDen"].reshape(m.nstages, len(m.elems), -1)
    np.testing.assert_array_equal(X.transpose(2, 1, 0), m.state["ionDen"])
    assert d["run.maxIterateMC"][0] == 4 and d["run.seed"][0] == 77


@pytest.mark.skipif(not DRIVER.exists(), reason="flang driver not built")
def test_driver_reports_a_missing_model_file(tmp_path):
    r = subprocess.run([str(DRIVER), str(tmp_path / "none.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=60)
    assert "can't open model file" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_matches_python_driver(tmp_path):
    if not DRIVER.exists():
        raise RuntimeError(f"{DRIVER} is missing: run __graft_entry__.build() with flang present")
    from mocassin_xray_amd import MocGPU
    m = small_model()
    model, out = tmp_path / "model.bin", tmp_path / "out.bin"
    fortran_io.write_model(model, m, "fp64", maxIterateMC=3, minConvergence=101.0, seed=4242)
    print("case fortran-driver nebula6 3 iterations", flush=True)
    r = subprocess.run([str(DRIVER), str(model), str(out)], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "iterations" in r.stdout
    got = fortran_io.read_dump(out)
    with MocGPU(m.tables, "fp64") as g:
        m.load(g)
        hist = g.MCIterationDriver(maxIterateMC=3, minConvergence=101.0, seed=4242)
        Te, X = g.download("Te"), g.download("ionDen")
    assert int(got["nIterateMC"][0]) == len(hist) == 3
    C1 = m.grid.nCells + 1
    rel = np.abs(got["Te"][1:] - Te[1:]) / Te[1:]
    # J is summed with device atomics in either run: the last bits differ
    assert np.mean(rel < 1e-6) > 0.95 and np.median(rel) < 1e-9
    Xf = got["ionDen"].reshape(m.nstages, len(m.elems), C1).transpose(2, 1, 0)
    assert np.median(np.abs(Xf[1:] - X[1:])) < 1e-9
    assert got["totPercent"][0] == pytest.approx(hist[-1]["convPercent"], abs=2.0)

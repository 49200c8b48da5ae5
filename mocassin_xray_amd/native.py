"""The library's own model set-up (include/mocmodel.h) from Python.

moc_model_build reads the reference's data/ directory and builds the mesh,
the source, the cross-section stack, the balance and emission tables and the
starting state in C++, with no Python in the loop; this module only fills
its spec and hands back what it built, for tests and for hosts that want the
library's tables.  `NativeModel.write` stores the model in the model-file
format of the Fortran driver (fortran_io.read_dump reads it back).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import MOC_PREC_FP64, MOC_PREC_MIXED, MocModelSpec


class NativeModelError(RuntimeError):
    pass


class NativeModel:
    """One gas model built by moc_model_build: an octant of n^3 cells of edge
    R [cm] at n_H = hden (a scalar, or an [n, n, n] array indexed [i, j, k]
    whose zero cells are inactive), the elements `elems` at abundances
    `abund`, `nstages` stages, the gas mesh of nbins points over
    nu_min-nu_max Ryd, a blackbody of T_star [K] and L36 [1e36 erg/s]
    (optionally with a F_nu ~ 1/nu power law between pl_keV = (lo, hi)
    carrying pl_fraction of the luminosity), n_packets packets."""

    def __init__(self, data_dir: str, elems, abund, nstages: int, nbins: int, nu_min: float,
                 nu_max: float, T_star: float, L36: float, n_packets: float, n: int, R: float,
                 hden=100.0, maxLev: int = 27, XHILimit: float = 0.05, precision: str = "fp64",
                 lgCompton: bool = True, strict: bool = True, pl_fraction: float = 0.0,
                 pl_keV=(0.01, 10.0), TeStart: float = 1.e4, H0in: float = 1.e-5):
        self._L = _lib.lib()
        self._keep = []
        z = np.ascontiguousarray(elems, dtype=np.int32)
        a = np.ascontiguousarray(abund, dtype=np.float64)
        spec = MocModelSpec()
        spec.dataDir = str(data_dir).encode()
        spec.precision = {"fp64": MOC_PREC_FP64, "mixed": MOC_PREC_MIXED}[precision]
        spec.nElem = z.size
        spec.elemZ = z.ctypes.data_as(C.POINTER(C.c_int32))
        spec.abund = a.ctypes.data_as(C.POINTER(C.c_double))
        spec.nstages, spec.maxLev, spec.XHILimit = int(nstages), int(maxLev), float(XHILimit)
        spec.nbins, spec.nuMin, spec.nuMax = int(nbins), float(nu_min), float(nu_max)
        spec.strictMesh, spec.lgCompton = int(bool(strict)), int(bool(lgCompton))
        spec.Tstellar, spec.LStar, spec.nPhotons = float(T_star), float(L36), float(n_packets)
        spec.plFraction = float(pl_fraction)
        spec.plLoKeV, spec.plHiKeV = float(pl_keV[0]), float(pl_keV[1])
        spec.nx, spec.Rout = int(n), float(R)
        if np.ndim(hden) == 0:
            spec.Hdens = float(hden)
        else:
            h = np.asarray(hden, dtype=np.float64)
            assert h.shape == (n, n, n), h.shape
            hf = np.ascontiguousarray(h.reshape(-1, order="F"))     # x fastest
            self._keep.append(hf)
            spec.hden = hf.ctypes.data_as(C.POINTER(C.c_double))
        spec.TeStart, spec.H0in = float(TeStart), float(H0in)
        self._keep += [z, a]
        self._m = C.c_void_p()
        rc = self._L.moc_model_build(C.byref(spec), C.byref(self._m))
        if rc != 0:
            raise NativeModelError(self._L.moc_model_last_error().decode())

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise NativeModelError(f"{what}: {self._L.moc_model_last_error().decode()}")

    def write(self, path) -> None:
        """The model as named records (fortran_io.read_dump reads them)."""
        self._check(self._L.moc_model_write(self._m, str(path).encode()), "moc_model_write")

    def state(self) -> dict:
        """setGrid's starting state: Te, Ne, Hden (C+1,) and ionDen (C+1, NE, NS)."""
        n = C.c_int64()
        p = [C.POINTER(C.c_double)() for _ in range(4)]
        self._check(self._L.moc_model_state(self._m, C.byref(n), *[C.byref(q) for q in p]),
                    "moc_model_state")
        c1 = n.value + 1
        t = _lib.MocAtomic()
        self._check(self._L.moc_model_get(self._m, None, None, C.byref(t), None, None),
                    "moc_model_get")
        ne, ns = t.nElemUsed, t.nstages
        out = {k: np.ctypeslib.as_array(q, shape=(c1,)).copy()
               for k, q in zip(("Te", "Ne", "Hden"), p[:3])}
        X = np.ctypeslib.as_array(p[3], shape=(ns * ne * c1,)).copy()
        out["ionDen"] = X.reshape(ns, ne, c1).transpose(2, 1, 0).copy()
        return out

    def close(self) -> None:
        if self._m:
            self._L.moc_model_free(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

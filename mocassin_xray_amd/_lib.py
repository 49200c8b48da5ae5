"""ctypes binding of libmocgpu (include/mocgpu.h, include/mocmodel.h).

The library is loaded from the package directory; a missing or unloadable
library raises at once -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmocgpu.so"

MOC_OK = 0
MOC_PREC_FP64 = 0
MOC_PREC_MIXED = 1
MOC_DT_F32, MOC_DT_F64, MOC_DT_I32 = 0, 1, 2
MOC_LAYOUT_FORTRAN, MOC_LAYOUT_CELLROW = 0, 1

FIELDS = {
    "opacity": 1, "Jste": 2, "recPDF": 3, "totalLines": 4, "Ne": 5, "Te": 6,
    "Hden": 7, "escaped": 8, "ionDen": 9, "lgConverged": 10, "lgBlack": 11,
    "emission": 12, "Ndust": 13, "Tdust": 14, "dustPDF": 15, "absOpac": 16, "scaOpac": 17,
    "grainColl": 18, "grainPE": 19,
}
GRAIN_FIELDS = {"grainColl", "grainPE"}
CELLBIN_FIELDS = {"opacity", "Jste", "recPDF", "escaped", "emission", "dustPDF", "absOpac",
                  "scaOpac"}

# every symbol include/mocgpu.h declares
EXPORTS = [
    "moc_create", "moc_destroy", "moc_last_error", "moc_grid_attach", "moc_upload",
    "moc_download", "moc_reset_tallies", "moc_transport", "moc_normalize_j",
    "moc_device_ptr", "moc_version", "moc_set_atomic", "moc_opacity", "moc_set_balance",
    "moc_update", "moc_set_emission", "moc_emission", "moc_set_packets", "moc_iterate",
    "moc_set_dust", "moc_dust_opacity", "moc_dust_emission", "moc_dust_temperature",
    "moc_grid_attach_batch", "moc_volume_averages", "moc_sed", "moc_line_luminosities",
    "moc_volume_averages_noalloc", "moc_set_graph", "moc_set_tuning", "moc_get_tuning",
    "moc_calibrate_atomics", "moc_set_qheat_state", "moc_grid_attach_nested",
]
# every symbol include/mocmodel.h declares
MODEL_EXPORTS = [
    "moc_model_build", "moc_model_free", "moc_model_last_error", "moc_model_get",
    "moc_model_state", "moc_model_load", "moc_model_write",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MocModelSpec(C.Structure):
    _fields_ = [
        ("dataDir", C.c_char_p), ("precision", C.c_int32), ("nElem", C.c_int32),
        ("elemZ", C.POINTER(C.c_int32)), ("abund", C.POINTER(C.c_double)),
        ("nstages", C.c_int32), ("maxLev", C.c_int32), ("XHILimit", C.c_double),
        ("nbins", C.c_int32), ("nuMin", C.c_double), ("nuMax", C.c_double),
        ("strictMesh", C.c_int32), ("lgCompton", C.c_int32), ("Tstellar", C.c_double),
        ("LStar", C.c_double), ("nPhotons", C.c_double), ("plFraction", C.c_double),
        ("plLoKeV", C.c_double), ("plHiKeV", C.c_double), ("nx", C.c_int32),
        ("Rout", C.c_double), ("Hdens", C.c_double), ("hden", C.POINTER(C.c_double)),
        ("TeStart", C.c_double), ("H0in", C.c_double),
    ]


class MocTables(C.Structure):
    _fields_ = [
        ("nbins", C.c_int32), ("precision", C.c_int32),
        ("nuArray", _dp), ("widFlx", _dp), ("ionEdge1", C.c_double),
        ("lgCompton", C.c_int32), ("nAngles", C.c_int32),
        ("KNsigmaT", _dp), ("KNsigmaCDF", _dp), ("Pcomp", _dp),
        ("nStars", C.c_int32), ("sourceCDF", _dp), ("Lstar", _dp), ("nPhotons", _dp),
        ("starPosition", _dp), ("starIndex", _ip),
    ]


class MocGridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xAxis", _dp), ("yAxis", _dp), ("zAxis", _dp), ("active", _ip),
        ("nCells", C.c_int32), ("lgSymmetricXYZ", C.c_int32), ("R_out", C.c_double),
        ("lgPlaneIonization", C.c_int32),
    ]


class MocAtomic(C.Structure):
    _fields_ = [
        ("nElemUsed", C.c_int32), ("elemZ", _ip), ("abund", _dp), ("nstages", C.c_int32),
        ("nSpecies", C.c_int32), ("speciesElem", _ip), ("speciesIon", _ip), ("sigma", _dp),
        ("brems", _dp),
    ]


_lp = C.POINTER(C.c_int64)


class MocBalanceTables(C.Structure):
    _fields_ = [
        ("maxLev", C.c_int32), ("XHILimit", C.c_double), ("nShells", C.c_int32),
        ("shellNum", _ip), ("speciesShell", _ip), ("W", _dp), ("outShell", _ip),
        ("statW0", _ip), ("statW1", _ip), ("rrb", _dp), ("drb_c", _dp), ("drb_e", _dp),
        ("rrec", _dp), ("rnew", _dp), ("nDiel", C.c_int32), ("diel", _dp),
        ("nAldro", C.c_int32), ("aldro", _dp), ("cionCF", _dp), ("cionTherm", _dp),
        ("nAuger", _ip), ("auger", _dp), ("nChex", C.c_int32), ("chexZ", _ip),
        ("chexIon", _ip), ("chexFit", _dp), ("nAtoms", C.c_int32), ("atomIdx", _ip),
        ("atomMeta", _lp), ("atomAtp", _dp), ("nLogT", C.c_int64), ("nLev", C.c_int64),
        ("nA", C.c_int64), ("nQ", C.c_int64), ("atomLogT", _dp), ("atomG", _dp),
        ("atomE", _dp), ("atomA", _dp), ("atomQom", _dp), ("atomQ2", _dp),
        ("atomRecKind", _ip), ("atomRec", _dp), ("atomRecValid", _ip),
    ]


class MocEmissionTables(C.Structure):
    _fields_ = [
        ("ntk", C.c_int32), ("tkGamma", _dp), ("nlim", C.c_int32 * 3), ("edgeP", _ip * 3),
        ("logGamma", _dp * 3), ("HIPnuP", C.c_int32), ("HeIPnuP", C.c_int32),
        ("HeIIPnuP", C.c_int32), ("xsHI", _dp), ("xsHeI", _dp), ("xsHeII", _dp),
        ("nFb", C.c_int32), ("fbZ", _ip), ("fbIon", _ip), ("fbLo", _ip), ("fbHi", _ip),
        ("fbOff", _lp), ("nFbXs", C.c_int64), ("fbXs", _dp), ("fbGratio", _dp),
        ("he2Row", _ip), ("hei2phot", _dp), ("nHlFiles", C.c_int32), ("hlMap", _ip),
        ("hlDens", _dp), ("hlLines", _dp), ("hbA", _dp), ("hbB", _dp), ("rbEdge", _dp),
        ("r2aEdge", _dp), ("heiRec", _dp), ("heiiLyman", _dp), ("j2TsP", C.c_int32),
        ("lymP", C.c_int32 * 4),
    ]


class MocDustTables(C.Structure):
    _fields_ = [
        ("nSpecies", C.c_int32), ("nSizes", C.c_int32), ("nTemps", C.c_int32),
        ("lgGas", C.c_int32), ("lgPhotoelectric", C.c_int32), ("grainAbun", _dp),
        ("grainWeight", _dp), ("TdustSublime", _dp), ("Cabs", _dp), ("Csca", _dp),
        ("emIntegral", _dp), ("XHILimit", C.c_double), ("grainRadius", _dp), ("grainVn", _dp),
        ("MsurfAtom", _dp), ("ipRyd", _dp), ("lgQHeat", C.c_int32), ("nTbins", C.c_int32),
        ("minaQHeat", C.c_double), ("minConvQHeat", C.c_double), ("Tmax", C.c_double),
        ("grainKind", _ip), ("lgEquivalentTau", C.c_int32),
    ]


class MocUpdateStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("noHit", "noIonBal", "noTeBal", "converged",
                                           "kernel_ms")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class MocIterationStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "convPercent", "noHitPercent", "noIonBalPercent", "noTeBalPercent", "packets",
        "opacity_ms", "emission_ms", "transport_ms", "update_ms", "segments", "emissions",
        "escaped", "iteration_ms")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class MocTuning(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("emitBatch", "segsPerTrip", "transportWaves", "graph", "hotJ")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class MocTransportStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "Qphot", "segments", "emissions", "absorbed", "scattered", "escaped", "lines",
        "lost", "trapped", "kernel_ms")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_LIB = None


def lib() -> C.CDLL:
    """Load libmocgpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.moc_create.argtypes = [C.POINTER(vp), C.POINTER(MocTables)]
    L.moc_create.restype = C.c_int
    L.moc_destroy.argtypes = [vp]
    L.moc_destroy.restype = None
    L.moc_last_error.argtypes = [vp]
    L.moc_last_error.restype = C.c_char_p
    L.moc_grid_attach.argtypes = [vp, C.c_int, C.POINTER(MocGridDesc)]
    L.moc_grid_attach.restype = C.c_int
    L.moc_upload.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int64, C.c_int, C.c_int]
    L.moc_upload.restype = C.c_int
    L.moc_download.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int64, C.c_int, C.c_int]
    L.moc_download.restype = C.c_int
    L.moc_reset_tallies.argtypes = [vp, C.c_int]
    L.moc_reset_tallies.restype = C.c_int
    L.moc_transport.argtypes = [vp, C.c_int, C.c_int64, C.c_uint64, C.c_uint32,
                                C.POINTER(MocTransportStats)]
    L.moc_transport.restype = C.c_int
    L.moc_normalize_j.argtypes = [vp, C.c_int]
    L.moc_normalize_j.restype = C.c_int
    L.moc_device_ptr.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int)]
    L.moc_device_ptr.restype = C.c_int
    L.moc_set_atomic.argtypes = [vp, C.POINTER(MocAtomic)]
    L.moc_set_atomic.restype = C.c_int
    L.moc_opacity.argtypes = [vp, C.c_int]
    L.moc_opacity.restype = C.c_int
    L.moc_set_balance.argtypes = [vp, C.POINTER(MocBalanceTables)]
    L.moc_set_balance.restype = C.c_int
    L.moc_update.argtypes = [vp, C.c_int, C.POINTER(MocUpdateStats)]
    L.moc_update.restype = C.c_int
    L.moc_set_emission.argtypes = [vp, C.POINTER(MocEmissionTables)]
    L.moc_set_emission.restype = C.c_int
    L.moc_emission.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
    L.moc_emission.restype = C.c_int
    L.moc_set_packets.argtypes = [vp, C.c_int, C.c_int64]
    L.moc_set_packets.restype = C.c_int
    L.moc_iterate.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.POINTER(MocIterationStats)]
    L.moc_iterate.restype = C.c_int
    L.moc_set_graph.argtypes = [vp, C.c_int]
    L.moc_set_graph.restype = C.c_int
    L.moc_set_tuning.argtypes = [vp, C.POINTER(MocTuning)]
    L.moc_set_tuning.restype = C.c_int
    L.moc_get_tuning.argtypes = [vp, C.POINTER(MocTuning)]
    L.moc_get_tuning.restype = C.c_int
    L.moc_version.argtypes = []
    L.moc_version.restype = C.c_char_p
    L.moc_calibrate_atomics.argtypes = [C.c_uint64, _dp]
    L.moc_calibrate_atomics.restype = C.c_int
    L.moc_set_qheat_state.argtypes = [vp, C.c_int, C.c_double]
    L.moc_set_qheat_state.restype = C.c_int
    L.moc_grid_attach_nested.argtypes = [vp, C.c_int, C.POINTER(MocGridDesc)]
    L.moc_grid_attach_nested.restype = C.c_int
    L.moc_model_build.argtypes = [C.POINTER(MocModelSpec), C.POINTER(vp)]
    L.moc_model_build.restype = C.c_int
    L.moc_model_free.argtypes = [vp]
    L.moc_model_free.restype = None
    L.moc_model_last_error.argtypes = []
    L.moc_model_last_error.restype = C.c_char_p
    L.moc_model_get.argtypes = [vp, C.POINTER(MocTables), C.POINTER(MocGridDesc),
                                C.POINTER(MocAtomic), C.POINTER(MocBalanceTables),
                                C.POINTER(MocEmissionTables)]
    L.moc_model_get.restype = C.c_int
    L.moc_model_state.argtypes = [vp, C.POINTER(C.c_int64)] + [C.POINTER(_dp)] * 4
    L.moc_model_state.restype = C.c_int
    L.moc_model_load.argtypes = [C.POINTER(vp), vp]
    L.moc_model_load.restype = C.c_int
    L.moc_model_write.argtypes = [vp, C.c_char_p]
    L.moc_model_write.restype = C.c_int
    _LIB = L
    return L


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)

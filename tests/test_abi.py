"""The C ABI: libmocgpu.so loads without a GPU and exports every function
include/mocgpu.h and include/mocmodel.h declare (no compute calls here)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(moc_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("moc_create", "moc_transport", "moc_opacity", "moc_update", "moc_set_balance"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from mocassin_xray_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) | set(_lib.MODEL_EXPORTS) == set(declared())


def test_version_string_needs_no_device():
    from mocassin_xray_amd import version
    assert version().startswith("gfx950 mocgpu")


def test_create_without_gpu_fails_cleanly():
    """No CPU fallback: without a device moc_create reports an error."""
    import numpy as np
    import pytest
    from mocassin_xray_amd import MocError, MocGPU, nebula
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    m = nebula.frozen_hsphere(n=4, nbins=40, n_packets=10, mesh="log")
    with pytest.raises(MocError):
        MocGPU(m.tables, "fp64")

"""Model files of the Fortran driver (fortran/mocgpu_driver.f90).

A model file holds every C struct of the ABI as named records -- the same
field dicts MocGPU builds its ctypes structs from -- plus the starting state
in the reference's array layout and the iteration controls.  Format: the 8
bytes "MOCDUMP1", then per record a 32-byte space-padded name, an int32 kind
(1 float64, 2 int32, 3 int64), an int32 0, an int64 count and the values;
a record named "END" closes the file.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

from .mocgpu import atomic_arrays, balance_arrays, emission_arrays, grid_arrays, tables_arrays

_KIND = {np.dtype(np.float64): 1, np.dtype(np.int32): 2, np.dtype(np.int64): 3}
_DTYPE = {1: np.float64, 2: np.int32, 3: np.int64}


def _record(f, name: str, a: np.ndarray):
    a = np.ascontiguousarray(a)
    f.write(name.ljust(32).encode()[:32])
    f.write(struct.pack("<iiq", _KIND[a.dtype], 0, a.size))
    f.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def _value(name: str, v):
    """Scalars keep their C type: the int64 fields of the structs are int64."""
    if isinstance(v, np.ndarray):
        return v
    if isinstance(v, float):
        return np.array([v], dtype=np.float64)
    if name in ("balance.nLogT", "balance.nLev", "balance.nA", "balance.nQ", "emission.nFbXs"):
        return np.array([v], dtype=np.int64)
    return np.array([v], dtype=np.int32)


def write_model(path, model, precision: str = "fp64", maxIterateMC: int = 5,
                minConvergence: float = 100.0, seed: int = 20261203, lgAutoPackets: bool = False,
                convIncPercent: float = 0.0, nPhotIncrease: float = 2.0,
                maxPhotons: float = 0.0) -> None:
    """Write a nebula.GasModel and the run controls for mocgpu_driver."""
    groups = {
        "tables": tables_arrays(model.tables, precision),
        "grid": grid_arrays(model.grid),
        "atomic": atomic_arrays(model.elems, model.abund, model.nstages, model.xsec),
        "balance": balance_arrays(model.bt),
        "emission": emission_arrays(model.et),
    }
    with open(path, "wb") as f:
        f.write(b"MOCDUMP1")
        for g, fields in groups.items():
            for k, v in fields.items():
                name = f"{g}.{k}"
                if v is None:
                    continue
                if k in ("edgeP", "logGamma"):
                    for i, a in enumerate(v):
                        _record(f, f"{name}.{i + 1}", a)
                elif k in ("nlim", "lymP"):
                    _record(f, name, np.array(v, dtype=np.int32))
                else:
                    _record(f, name, _value(name, v))
        st = model.state
        for k in ("Te", "Ne", "Hden"):
            _record(f, f"state.{k}", np.asarray(st[k], dtype=np.float64))
        # (0:nCells, nElem, nstages) column-major: cell index fastest
        _record(f, "state.ionDen", np.asarray(st["ionDen"], dtype=np.float64).transpose(2, 1, 0))
        for k, v in (("maxIterateMC", int(maxIterateMC)), ("lgAutoPackets", int(lgAutoPackets))):
            _record(f, f"run.{k}", np.array([v], dtype=np.int32))
        for k, v in (("minConvergence", minConvergence), ("convIncPercent", convIncPercent),
                     ("nPhotIncrease", nPhotIncrease), ("maxPhotons", maxPhotons)):
            _record(f, f"run.{k}", np.array([float(v)]))
        _record(f, "run.seed", np.array([seed], dtype=np.int64))
        f.write("END".ljust(32).encode() + struct.pack("<iiq", 0, 0, 0))


def read_dump(path) -> dict:
    """Records of a model or output file as numpy arrays."""
    data = Path(path).read_bytes()
    if data[:8] != b"MOCDUMP1":
        raise ValueError(f"{path}: not a model file")
    out, p = {}, 8
    while True:
        name = data[p:p + 32].decode().strip()
        kind, _, n = struct.unpack("<iiq", data[p + 32:p + 48])
        p += 48
        if name == "END":
            return out
        dt = np.dtype(_DTYPE[kind]).newbyteorder("<")
        out[name] = np.frombuffer(data, dtype=dt, count=n, offset=p).astype(_DTYPE[kind])
        p += n * dt.itemsize

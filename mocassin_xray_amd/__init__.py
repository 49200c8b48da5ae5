"""mocassin_xray_amd -- MI355X (gfx950) Monte Carlo iteration of MOCASSIN.

The compute path is libmocgpu.so (hand-written HIP kernels behind a C ABI,
include/mocgpu.h); this package holds its sources (csrc/), its build recipe,
the Python host mirror of the reference's transport interface (mocgpu.py),
the host-side table builders (tables.py) and the synthetic model set-ups
(nebula.py).
"""
from .mocgpu import CartesianGrid, MocError, MocGPU, SourceTables, version  # noqa: F401

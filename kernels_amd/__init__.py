"""MI355X (gfx950) implementation of the PRK Cxx11 stencil path.

The compute lives in ``libprkstencil.so`` (hand-written HIP kernels behind
the C ABI declared in ``include/prk_stencil.h``).  This package loads that
library with ctypes and exposes it to Python callers (tests, the benchmark)
with torch tensors as device buffers.  There is no fallback: if the library
is missing every call raises.
"""
from __future__ import annotations

from ._lib import (  # noqa: F401
    BF16, F32, F64, GRID, INIT_ANALYTIC, INIT_SEEDED, STAR, PrkError, Tuning, check, lib, lib_path,
)
from .ops import (  # noqa: F401
    DTYPES, add, alloc_grid, alloc_slab, autotune, default_pitch, init, l1norm, nominal_flops, run, slab_bytes_per_iter, slab_halos, stencil,
    stencil_bytes_per_iter, stencil_ex, sweep, sweep_slab, torch_dtype, OPERATORS, transpose,
    transpose_bytes_per_iter, transpose_check, transpose_init, transpose_run,
)

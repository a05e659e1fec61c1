"""Python face of the C ABI, with torch tensors as device buffers.

Grids are 2-D tensors of shape (n, pitch) on a HIP device; pitch >= n is the
row stride in elements.  BF16 storage uses torch.bfloat16 tensors (the
library reads the raw bits).
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import BF16, F32, F64, GRID, STAR, check, lib

DTYPES = {"f64": F64, "f32": F32, "bf16": BF16}
_TORCH = {F64: torch.float64, F32: torch.float32, BF16: torch.bfloat16}
_ESIZE = {F64: 8, F32: 4, BF16: 2}

# The reference's operator table (Cxx11/stencil.cc:142-159), by name.
OPERATORS = {f"{p}{r}": (STAR if p == "star" else GRID, r) for p in ("star", "grid") for r in range(1, 6)}


def torch_dtype(dtype: int) -> torch.dtype:
    return _TORCH[dtype]


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _stream(stream) -> int | None:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _check_grids(dtype: int, *grids, square: bool = True, rows: int | None = None) -> tuple[int, int]:
    """Checks the buffers of one call before the kernels see them; returns
    (rows, pitch).  Every grid must be a 2-D tensor of the storage type's
    torch dtype with unit column stride; all share shape and row stride
    (pitch >= columns); each one's storage must hold rows * pitch elements
    from its first element (the kernels may read a whole padded row); and
    they must live on a HIP device.  None entries (an absent out) are
    skipped.  Raises ValueError (TypeError for a non-tensor)."""
    if dtype not in _TORCH:
        raise ValueError(f"unknown dtype {dtype}")
    want, es = _TORCH[dtype], _ESIZE[dtype]
    ts = [t for t in grids if t is not None]
    if not ts:
        raise ValueError("no grid given")
    for t in ts:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"grid must be a torch tensor, got {type(t).__name__}")
        if t.dtype != want:
            raise ValueError(f"grid of dtype {t.dtype} passed for storage type {dtype} ({want})")
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError(f"grid must be 2-D with unit column stride, got shape {tuple(t.shape)} "
                             f"strides {t.stride()}")
    ref = ts[0]
    for t in ts[1:]:
        if t.shape != ref.shape or t.stride(0) != ref.stride(0):
            raise ValueError(f"grids differ in shape or row stride: {tuple(ref.shape)}/{ref.stride(0)} vs "
                             f"{tuple(t.shape)}/{t.stride(0)}")
    nrows, ncols = ref.shape
    if square and nrows != ncols:
        raise ValueError(f"grid must be square, got {tuple(ref.shape)}")
    if rows is not None and nrows != rows:
        raise ValueError(f"grid has {nrows} rows, expected {rows}")
    pitch = ref.stride(0) if nrows > 1 else max(ref.stride(0), ncols)
    if pitch < ncols:
        raise ValueError(f"row stride {pitch} < {ncols} columns")
    need = nrows * pitch * es
    for t in ts:
        avail = t.untyped_storage().nbytes() - t.storage_offset() * es
        if avail < need:
            raise ValueError(f"grid storage holds {avail} bytes from its first element; the kernels "
                             f"address {need} ({nrows} rows of pitch {pitch})")
    for t in ts:
        if not t.is_cuda:
            raise ValueError("grid must be a HIP device tensor (there is no CPU path)")
        if t.device != ref.device:
            raise ValueError("grids are on different devices")
    return nrows, pitch


def _check_dense(n: int, *grids) -> None:
    """The drop-in operators' buffers: n x n FP64 with pitch n (the
    reference's prk::vector layout)."""
    for t in grids:
        if isinstance(t, torch.Tensor) and (tuple(t.shape) != (n, n) or not t.is_contiguous()):
            raise ValueError(f"drop-in operators take a contiguous {n} x {n} FP64 grid, got shape "
                             f"{tuple(t.shape)} strides {t.stride()}")
    _check_grids(F64, *grids)


def default_pitch(n: int, dtype: int) -> int:
    """Row stride (elements) of alloc_grid and alloc_slab.  Rows are padded to
    128 bytes, so that every row starts on a cache line (the vector sweeps
    need 16; on MI355X rows on a line boundary ran 11-17 % faster where n is
    not a multiple of 128 bytes), and rows of 64 KiB or more to 1 KiB, so
    that each wave's 1 KiB row segment starts on a 1 KiB boundary (at
    n = 30000 that ran 1-2.6 % faster in FP64 and FP32 for at most 1.6 %
    more memory)."""
    es = _ESIZE[dtype]
    per = (1024 if n * es >= 65536 else 128) // es
    return (n + per - 1) // per * per


def alloc_grid(n: int, dtype: int = F64, device: str = "cuda", pitch: int | None = None,
               fill: float | None = None) -> torch.Tensor:
    """n x n grid stored with row stride `pitch` (default: default_pitch,
    rows rounded up to 128 bytes, or to 1 KiB for rows of 64 KiB or more).
    Returns the (n, n) view; its storage is (n, pitch)."""
    if pitch is None:
        pitch = default_pitch(n, dtype)
    base = torch.empty((n, pitch), dtype=_TORCH[dtype], device=device)
    if fill is not None:
        base.fill_(fill)
    return base[:, :n]


def stencil(name_or_pattern, n: int, t: int, inp: torch.Tensor, out: torch.Tensor, radius: int | None = None):
    """Drop-in operator: out += S(in), FP64, pitch n, null stream."""
    _check_dense(n, inp, out)
    if isinstance(name_or_pattern, str):
        pattern, radius = OPERATORS[name_or_pattern]
        check(getattr(lib(), f"prk_{name_or_pattern}")(n, t, inp.data_ptr(), out.data_ptr()))
    else:
        check(lib().prk_stencil(int(name_or_pattern), int(radius), n, t, inp.data_ptr(), out.data_ptr()))


def add(n: int, t: int, inp: torch.Tensor) -> None:
    _check_dense(n, inp)
    check(lib().prk_add(n, t, inp.data_ptr()))


def sweep(pattern: int, radius: int, dtype: int, inp: torch.Tensor, in_next: torch.Tensor,
          out: torch.Tensor, stream=None) -> None:
    n, pitch = _check_grids(dtype, inp, in_next, out)
    check(lib().prk_sweep(pattern, radius, dtype, n, pitch, _ptr(inp), _ptr(in_next), _ptr(out),
                          _stream(stream)))


def slab_halos(n: int, radius: int, row0: int, rows: int) -> tuple[int, int]:
    """Halo rows above and below a slab owning global rows [row0, row0+rows)."""
    return min(radius, row0), min(radius, n - row0 - rows)


def alloc_slab(n: int, radius: int, row0: int, rows: int, dtype: int = F64, device: str = "cuda",
               pitch: int | None = None) -> torch.Tensor:
    """Buffer of one slab: its rows plus its halo rows, n columns, row stride
    `pitch` (default that of alloc_grid)."""
    top, bot = slab_halos(n, radius, row0, rows)
    if pitch is None:
        pitch = default_pitch(n, dtype)
    return torch.empty((top + rows + bot, pitch), dtype=_TORCH[dtype], device=device)[:, :n]


def sweep_slab(pattern: int, radius: int, dtype: int, n: int, row0: int, rows: int, inp: torch.Tensor,
               in_next: torch.Tensor, out: torch.Tensor, stream=None) -> None:
    """Fused sweep of the slab owning global rows [row0, row0+rows) of an
    n x n grid (prk_sweep_slab): buffers hold those rows and their halo rows."""
    top, bot = slab_halos(n, radius, row0, rows)
    for t in (inp, in_next, out):
        if t.dim() != 2 or t.shape[0] != top + rows + bot or t.shape[1] != n:
            raise ValueError(f"slab buffer of shape {tuple(t.shape)}, expected ({top + rows + bot}, {n})")
    _, pitch = _check_grids(dtype, inp, in_next, out, square=False)
    check(lib().prk_sweep_slab(pattern, radius, dtype, n, row0, rows, pitch, _ptr(inp), _ptr(in_next),
                               _ptr(out), _stream(stream)))


def autotune(pattern: int, radius: int, dtype: int, buf0: torch.Tensor, buf1: torch.Tensor, out: torch.Tensor,
             reps: int = 2, stream=None):
    """Tune the sweep of this shape on the caller's three grids (contents
    overwritten; no device allocation) and record the result for later
    sweeps.  Returns the chosen Tuning."""
    from ._lib import Tuning
    t = Tuning()
    n, pitch = _check_grids(dtype, buf0, buf1, out)
    check(lib().prk_autotune_buffers(pattern, radius, dtype, n, pitch, _ptr(buf0), _ptr(buf1), _ptr(out),
                                     reps, _stream(stream), ctypes.byref(t)))
    return t


def stencil_ex(pattern: int, radius: int, dtype: int, inp: torch.Tensor, out: torch.Tensor, stream=None) -> None:
    n, pitch = _check_grids(dtype, inp, out)
    check(lib().prk_stencil_ex(pattern, radius, dtype, n, pitch, _ptr(inp), _ptr(out), _stream(stream)))


def init(kind: int, seed: int, dtype: int, inp: torch.Tensor, out: torch.Tensor | None, stream=None) -> None:
    n, pitch = _check_grids(dtype, inp, out)
    check(lib().prk_init(kind, ctypes.c_uint64(seed), dtype, n, pitch, _ptr(inp), _ptr(out),
                         _stream(stream)))


def l1norm(dtype: int, radius: int, out: torch.Tensor, stream=None) -> float:
    n, pitch = _check_grids(dtype, out)
    v = ctypes.c_double(0.0)
    check(lib().prk_l1norm(dtype, n, pitch, radius, _ptr(out), ctypes.byref(v), _stream(stream)))
    return v.value


def run(pattern: int, radius: int, dtype: int, iterations: int, buf0: torch.Tensor, buf1: torch.Tensor,
        out: torch.Tensor, stream=None) -> tuple[int, float]:
    """iterations+1 fused sweeps; returns (index of the buffer holding the
    final in, average seconds per timed sweep)."""
    n, pitch = _check_grids(dtype, buf0, buf1, out)
    fb = ctypes.c_int(0)
    avg = ctypes.c_double(0.0)
    check(lib().prk_run(pattern, radius, dtype, n, pitch, iterations, _ptr(buf0), _ptr(buf1),
                        _ptr(out), _stream(stream), ctypes.byref(fb), ctypes.byref(avg)))
    return fb.value, avg.value


def stencil_bytes_per_iter(n: int, radius: int, dtype: int) -> int:
    """Algorithmic HBM bytes of one fused sweep: read in and write in_next
    on all n^2 points, read and write out on the (n-2R)^2 interior."""
    s = _ESIZE[dtype]
    m = n - 2 * radius
    return 2 * s * n * n + 2 * s * m * m


def slab_bytes_per_iter(n: int, radius: int, dtype: int, row0: int, rows: int) -> int:
    """Algorithmic HBM bytes of one slab sweep: read in on the slab's rows and
    halo rows, write in_next on its own rows, read and write out on its own
    rows inside the interior."""
    s = _ESIZE[dtype]
    top, bot = slab_halos(n, radius, row0, rows)
    out_rows = max(0, min(row0 + rows, n - radius) - max(row0, radius))
    return s * n * (top + rows + bot) + s * n * rows + 2 * s * (n - 2 * radius) * out_rows


def nominal_flops(pattern: int, radius: int, n: int) -> int:
    """The reference's flop count per sweep (Cxx11/stencil.cc:221-222)."""
    size = 4 * radius + 1 if pattern == STAR else (2 * radius + 1) ** 2
    return (2 * size + 1) * (n - 2 * radius) ** 2


# ---------------------------------------------------------------------------
# Matrix transpose (B += A^T, A += 1), order x order FP64 row-major.
# ---------------------------------------------------------------------------
def _check_matrices(*mats) -> int:
    """Transpose operands: contiguous order x order FP64 device tensors."""
    order, _ = _check_grids(F64, *mats)
    for m in mats:
        if not m.is_contiguous():
            raise ValueError(f"transpose operands must be contiguous, got strides {m.stride()}")
    return order


def transpose(A: torch.Tensor, B: torch.Tensor, stream=None) -> None:
    order = _check_matrices(A, B)
    check(lib().prk_transpose_ex(order, _ptr(A), _ptr(B), _stream(stream)))


def transpose_init(A: torch.Tensor, B: torch.Tensor, stream=None) -> None:
    check(lib().prk_transpose_init(_check_matrices(A, B), _ptr(A), _ptr(B), _stream(stream)))


def transpose_run(iterations: int, A: torch.Tensor, B: torch.Tensor, stream=None) -> float:
    """iterations + 1 transposes; returns the average seconds of the last
    `iterations`."""
    avg = ctypes.c_double(0.0)
    check(lib().prk_transpose_run(_check_matrices(A, B), iterations, _ptr(A), _ptr(B), _stream(stream),
                                  ctypes.byref(avg)))
    return avg.value


def transpose_check(iterations: int, B: torch.Tensor, stream=None) -> float:
    """The reference's validation sum of absolute differences (0 = exact)."""
    v = ctypes.c_double(0.0)
    check(lib().prk_transpose_check(_check_matrices(B), iterations, _ptr(B), ctypes.byref(v), _stream(stream)))
    return v.value


def transpose_bytes_per_iter(order: int) -> int:
    """Read and write A and B once: the reference's 4 * 8 * order^2."""
    return 4 * 8 * order * order


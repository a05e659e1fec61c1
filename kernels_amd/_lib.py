"""ctypes binding of libprkstencil.so (C ABI: include/prk_stencil.h)."""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

STAR, GRID = 0, 1
F64, F32, BF16 = 0, 1, 2
INIT_ANALYTIC, INIT_SEEDED = 0, 1

_HERE = Path(__file__).resolve().parent
_LIB = None

c_int, c_ll, c_double, c_void_p, c_u64 = (ctypes.c_int, ctypes.c_longlong, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_uint64)
c_int_p, c_double_p = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "prk_last_error": (ctypes.c_char_p, []),
    "prk_abi_version": (c_int, []),
    **{f"prk_{p}{r}": (c_int, [c_int, c_int, c_void_p, c_void_p])
       for p in ("star", "grid") for r in range(1, 6)},
    "prk_stencil": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "prk_add": (c_int, [c_int, c_int, c_void_p]),
    "prk_sweep": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_void_p]),
    "prk_sweep_slab": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p,
                               c_void_p]),
    "prk_stencil_ex": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p]),
    "prk_add_ex": (c_int, [c_int, c_int, c_ll, c_void_p, c_void_p]),
    "prk_init": (c_int, [c_int, c_u64, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p]),
    "prk_l1norm": (c_int, [c_int, c_int, c_ll, c_int, c_void_p, c_double_p, c_void_p]),
    "prk_l1norm_noalloc": (c_int, [c_int, c_int, c_ll, c_int, c_void_p, c_void_p, c_void_p]),
    "prk_run": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_int, c_void_p, c_void_p, c_void_p,
                        c_void_p, c_int_p, c_double_p]),
    "prk_set_rows_per_wave": (c_int, [c_int]),
    "prk_set_sweep_variant": (c_int, [c_int]),
    "prk_set_sweep_order": (c_int, [c_int, c_int]),
    "prk_triad": (c_int, [c_ll, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "prk_copy": (c_int, [c_ll, c_void_p, c_void_p, c_void_p]),
    "prk_copy_shape": (c_int, [c_ll, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "prk_crash_backtrace": (c_int, [c_int]),
    "prk_stream_mix": (c_int, [c_ll, c_void_p, c_void_p, c_void_p, c_void_p]),
    "prk_transpose": (c_int, [c_int, c_int, c_void_p, c_void_p]),
    "prk_transpose_ex": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "prk_transpose_init": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "prk_transpose_run": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_double_p]),
    "prk_transpose_check": (c_int, [c_int, c_int, c_void_p, c_double_p, c_void_p]),
    "prk_set_transpose_variant": (c_int, [c_int, c_int]),
}


class Tuning(ctypes.Structure):
    _fields_ = [("variant", c_int), ("rows_per_wave", c_int), ("wgs_per_cu", c_int), ("order", c_int),
                ("gbs", c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


SIGNATURES.update({
    "prk_autotune": (c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(Tuning)]),
    "prk_autotune_buffers": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                     ctypes.POINTER(Tuning)]),
    "prk_tuning_get": (c_int, [c_int, c_int, c_int, c_int, ctypes.POINTER(Tuning)]),
    "prk_tuning_set": (c_int, [c_int, c_int, c_int, c_int, ctypes.POINTER(Tuning)]),
    "prk_tuning_clear": (c_int, []),
    "prk_tuning_save": (c_int, [ctypes.c_char_p]),
    "prk_tuning_load": (c_int, [ctypes.c_char_p]),
})


class Solve(ctypes.Structure):
    _fields_ = [("pattern", c_int), ("radius", c_int), ("dtype", c_int), ("n", c_int), ("iterations", c_int),
                ("init", c_int), ("seed", c_u64)]


_SOLVE_P = ctypes.POINTER(Solve)
SIGNATURES.update({
    "prk_batch_create": (c_int, [_SOLVE_P, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "prk_batch_run": (c_int, [c_void_p, c_void_p]),
    "prk_batch_wait": (c_int, [c_void_p, c_double_p]),
    "prk_batch_output": (c_int, [c_void_p, c_int, c_int, c_void_p, ctypes.c_size_t]),
    "prk_batch_info": (c_int, [c_void_p, ctypes.POINTER(c_ll), ctypes.POINTER(c_ll), c_int_p]),
    "prk_batch_errors": (c_int, [c_void_p, c_void_p, c_double_p]),
    "prk_batch_destroy": (None, [c_void_p]),
    "prk_solve_one_by_one": (c_int, [_SOLVE_P, c_int, c_int, c_double_p, c_void_p]),
    "prk_set_batch_round_tiles": (c_int, [c_ll]),
    "prk_batch_autotune": (c_int, [_SOLVE_P, c_int, c_int, c_int, c_int_p]),
    "prk_batch_tuning_get": (c_int, [c_int, c_int, c_int, c_int, c_int_p]),
    "prk_batch_tuning_set": (c_int, [c_int, c_int, c_int, c_int, c_int]),
    "prk_batch_tuning_clear": (c_int, []),
})


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", c_ll), ("max_pending", c_int), ("max_active", c_int), ("round_tiles", c_ll),
                ("shadow_every", c_int)]


class ServerResult(ctypes.Structure):
    _fields_ = [("id", c_u64), ("norm", c_double), ("t_submit", c_double), ("t_admit", c_double),
                ("t_done", c_double), ("status", c_int), ("err_valid", c_int), ("err_max_abs", c_double),
                ("err_mean_abs", c_double), ("err_normwise", c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(f, c_ll) for f in ("submitted", "served", "rejected", "windows", "launches", "pending", "active",
                                    "pool_used", "budget")] + [("build_seconds", c_double), ("wait_seconds", c_double),
                                                               ("graph_updates", c_ll), ("shadow_solves", c_ll)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


SIGNATURES.update({
    "prk_server_create": (c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(c_void_p)]),
    "prk_server_submit": (c_int, [c_void_p, _SOLVE_P, c_int, ctypes.POINTER(c_u64)]),
    "prk_server_poll": (c_int, [c_void_p, ctypes.POINTER(ServerResult), c_int, c_int_p]),
    "prk_server_drain": (c_int, [c_void_p, c_double]),
    "prk_server_fetch": (c_int, [c_void_p, c_u64, c_int, c_void_p, ctypes.c_size_t]),
    "prk_server_release": (c_int, [c_void_p, c_u64]),
    "prk_server_stats_get": (c_int, [c_void_p, ctypes.POINTER(ServerStats)]),
    "prk_server_clock": (c_double, []),
    "prk_server_destroy": (None, [c_void_p]),
})


class PrkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"prk error {code}: {msg}")
        self.code = code


def lib_path() -> Path:
    return Path(os.environ.get("PRK_STENCIL_LIB", _HERE / "libprkstencil.so"))


def lib() -> ctypes.CDLL:
    """Load libprkstencil.so (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not p.exists():
            raise FileNotFoundError(
                f"{p} not found: build it with `python -m kernels_amd.build` (no CPU fallback exists)")
        h = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(rc: int) -> None:
    if rc != 0:
        raise PrkError(rc, lib().prk_last_error().decode(errors="replace"))

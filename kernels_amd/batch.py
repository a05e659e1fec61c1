"""Batched solves: every grid resident in one device pool, round kernels per
(storage type, pattern, radius) class replayed from one hipGraph (C ABI:
prk_batch_* in include/prk_stencil.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import Solve, check, lib

_NP = {0: np.float64, 1: np.float32, 2: np.uint16}


def make_solves(specs) -> ctypes.Array:
    """specs: iterable of dicts/tuples (pattern, radius, dtype, n, iterations, init, seed)."""
    specs = list(specs)
    arr = (Solve * len(specs))()
    for i, s in enumerate(specs):
        if isinstance(s, dict):
            s = (s["pattern"], s["radius"], s["dtype"], s["n"], s["iterations"], s.get("init", 1), s.get("seed", 0))
        arr[i] = Solve(*s)
    return arr


PERSISTENT = 1  # PRK_BATCH_PERSISTENT: disabled in the library


class Batch:
    def __init__(self, specs, rows_per_tile: int = 0, flags: int = 0):
        self.solves = make_solves(specs)
        self.count = len(self.solves)
        self.h = None
        h = ctypes.c_void_p()
        check(lib().prk_batch_create(self.solves, self.count, rows_per_tile, flags, ctypes.byref(h)))
        self.h = h

    def run(self, stream=None) -> None:
        check(lib().prk_batch_run(self.h, stream))

    def wait(self) -> np.ndarray:
        norms = np.zeros(self.count)
        check(lib().prk_batch_wait(self.h, norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return norms

    def output(self, idx: int, which: int = 0) -> np.ndarray:
        s = self.solves[idx]
        a = np.empty((s.n, s.n), _NP[s.dtype])
        check(lib().prk_batch_output(self.h, idx, which, a.ctypes.data, a.nbytes))
        return a

    def info(self) -> dict:
        b, t, q = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        check(lib().prk_batch_info(self.h, ctypes.byref(b), ctypes.byref(t), ctypes.byref(q)))
        return {"device_bytes": b.value, "tiles": t.value, "launches": q.value}

    def errors(self, ref: "Batch") -> np.ndarray:
        """Error of every result against `ref`, the same solves stored in
        FP64 (both run and waited for): rows (max abs, mean abs, normwise
        L2) per solve, zeros for solves stored in FP64 here
        (prk_batch_errors)."""
        e = np.zeros((self.count, 3))
        check(lib().prk_batch_errors(self.h, ref.h, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return e

    def close(self) -> None:
        if self.h:
            lib().prk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def one_by_one(specs, dtype_override: int = -1, stream=None) -> np.ndarray:
    solves = make_solves(specs)
    norms = np.zeros(len(solves))
    check(lib().prk_solve_one_by_one(solves, len(solves), dtype_override,
                                     norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), stream))
    return norms


def autotune(specs, sample: int = 16, reps: int = 2) -> int:
    """Tune tile heights per (pattern, radius, dtype, size class) on sample
    batches drawn from specs; later Batch(..., rows_per_tile=0) uses them.
    Returns the number of keys tuned."""
    arr = make_solves(specs)
    k = ctypes.c_int()
    check(lib().prk_batch_autotune(arr, len(arr), sample, reps, ctypes.byref(k)))
    return k.value


def tuning_get(pattern: int, radius: int, dtype: int, n: int) -> int:
    r = ctypes.c_int()
    check(lib().prk_batch_tuning_get(pattern, radius, dtype, n, ctypes.byref(r)))
    return r.value


def tuning_set(pattern: int, radius: int, dtype: int, n: int, rows: int) -> None:
    check(lib().prk_batch_tuning_set(pattern, radius, dtype, n, rows))


def tuning_clear() -> None:
    check(lib().prk_batch_tuning_clear())


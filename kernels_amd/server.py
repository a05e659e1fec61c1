"""Continuous-batching solve server (C ABI: prk_server_* in include/prk_stencil.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import ServerConfig, ServerResult, ServerStats, Solve, check, lib

_NP = {0: np.float64, 1: np.float32, 2: np.uint16}


class Server:
    def __init__(self, budget_bytes: int, max_pending: int = 1024, max_active: int = 4096,
                 round_tiles: int = 16384, shadow_every: int = 0):
        if os.environ.get("PRK_CRASH_BACKTRACE"):
            # a host fault (in the worker thread too) prints its native frames
            check(lib().prk_crash_backtrace(1))
        cfg = ServerConfig(budget_bytes, max_pending, max_active, round_tiles, shadow_every)
        self.h = None
        h = ctypes.c_void_p()
        check(lib().prk_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.solves = {}

    def submit(self, spec, keep: bool = False) -> int:
        """spec: (pattern, radius, dtype, n, iterations, init, seed).  Returns the
        request id; raises PrkError (code -5 queue full, -4 too large)."""
        s = Solve(*spec)
        rid = ctypes.c_uint64()
        check(lib().prk_server_submit(self.h, ctypes.byref(s), int(keep), ctypes.byref(rid)))
        self.solves[rid.value] = tuple(spec)
        return rid.value

    def try_submit(self, spec, keep: bool = False):
        """Like submit, but returns (id, 0) or (None, error code)."""
        s = Solve(*spec)
        rid = ctypes.c_uint64()
        rc = lib().prk_server_submit(self.h, ctypes.byref(s), int(keep), ctypes.byref(rid))
        if rc == 0:
            self.solves[rid.value] = tuple(spec)
            return rid.value, 0
        return None, rc

    def poll(self, max_results: int = 4096) -> list[dict]:
        buf = (ServerResult * max_results)()
        n = ctypes.c_int()
        check(lib().prk_server_poll(self.h, buf, max_results, ctypes.byref(n)))
        out = []
        for r in buf[: n.value]:
            d = {"id": r.id, "norm": r.norm, "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done}
            if r.err_valid:  # shadowed: error against the same solve in FP64
                d["error_vs_f64"] = {"max_abs": r.err_max_abs, "mean_abs": r.err_mean_abs,
                                     "normwise": r.err_normwise}
            out.append(d)
        return out

    def drain(self, timeout_s: float = 60.0) -> None:
        check(lib().prk_server_drain(self.h, timeout_s))

    def fetch(self, rid: int, which: int = 0) -> np.ndarray:
        spec = self.solves[rid]
        n, dtype = spec[3], spec[2]
        a = np.empty((n, n), _NP[dtype])
        check(lib().prk_server_fetch(self.h, rid, which, a.ctypes.data, a.nbytes))
        return a

    def release(self, rid: int) -> None:
        check(lib().prk_server_release(self.h, rid))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().prk_server_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    @staticmethod
    def clock() -> float:
        return lib().prk_server_clock()

    def close(self) -> None:
        if self.h:
            lib().prk_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

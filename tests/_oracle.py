"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: only tests, the smoke check and the benchmark's
CPU-baseline leg load it.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_SO = REPO / "oracle" / "liboracle.so"
REF_SO = REPO / "oracle" / "_ref" / "libprkref.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64

F64, F32, BF16 = 0, 1, 2
NP_LP = {F32: np.float32, BF16: np.uint16}


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, path: Path = ORACLE_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(str(path))
        L.prko_weight.restype = ctypes.c_double
        L.prko_weight.argtypes = [ctypes.c_int] * 4
        L.prko_taps.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp]
        L.prko_seeded_value.restype = ctypes.c_double
        L.prko_seeded_value.argtypes = [_u64, _u64]
        L.prko_init.argtypes = [ctypes.c_int, _u64, ctypes.c_int, _dp, _dp]
        L.prko_stencil.argtypes = [ctypes.c_int] * 4 + [_dp, _dp]
        L.prko_add.argtypes = [ctypes.c_int, _dp]
        L.prko_l1norm.restype = ctypes.c_double
        L.prko_l1norm.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.prko_run.argtypes = [ctypes.c_int] * 6 + [_u64, _dp, _dp]
        L.prko_band.argtypes = [ctypes.c_int] * 5 + [_u64, ctypes.c_int, ctypes.c_int, _dp]
        L.prko_run_lp.argtypes = [ctypes.c_int] * 6 + [_u64, _vp, _vp]
        L.prko_band_lp.argtypes = [ctypes.c_int] * 6 + [_u64, ctypes.c_int, ctypes.c_int, _vp]
        L.prko_sweep_lp.argtypes = [ctypes.c_int] * 4 + [_vp, _vp]
        L.prko_l1norm_lp.restype = ctypes.c_double
        L.prko_l1norm_lp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        L.prko_f32_to_bf16.restype = ctypes.c_uint16
        L.prko_f32_to_bf16.argtypes = [ctypes.c_float]
        L.prko_bf16_to_f32.restype = ctypes.c_float
        L.prko_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.prko_omp_run.argtypes = [ctypes.c_int] * 6 + [_dp, _dp, _ip]
        L.prko_transpose.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.prko_transpose_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.prko_transpose_abserr.restype = ctypes.c_double
        L.prko_transpose_abserr.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        self.L = L

    def taps(self, pattern: int, radius: int):
        m = (2 * radius + 1) ** 2
        di = (ctypes.c_int * m)()
        dj = (ctypes.c_int * m)()
        w = (ctypes.c_double * m)()
        k = self.L.prko_taps(pattern, radius, di, dj, w)
        assert k > 0
        return [(di[i], dj[i], w[i]) for i in range(k)]

    def weight(self, pattern, radius, di, dj) -> float:
        return self.L.prko_weight(pattern, radius, di, dj)

    def seeded_value(self, seed: int, idx: int) -> float:
        return self.L.prko_seeded_value(seed, idx)

    def init(self, kind: int, seed: int, n: int):
        a = np.empty((n, n)); o = np.empty((n, n))
        assert self.L.prko_init(kind, seed, n, _d(a), _d(o)) == 0
        return a, o

    def stencil(self, pattern, radius, n, t, inp, out) -> None:
        assert self.L.prko_stencil(pattern, radius, n, t, _d(inp), _d(out)) == 0

    def add(self, n, inp) -> None:
        assert self.L.prko_add(n, _d(inp)) == 0

    def l1norm(self, n, radius, out) -> float:
        return self.L.prko_l1norm(n, radius, _d(out))

    def run(self, pattern, radius, n, t, iterations, init, seed=0):
        out = np.empty((n, n)); fin = np.empty((n, n))
        assert self.L.prko_run(pattern, radius, n, t, iterations, init, seed, _d(out), _d(fin)) == 0
        return out, fin

    def band(self, pattern, radius, n, iterations, init, seed, a, b):
        out = np.empty((b - a, n))
        assert self.L.prko_band(pattern, radius, n, iterations, init, seed, a, b, _d(out)) == 0
        return out

    def run_lp(self, dtype, pattern, radius, n, iterations, init, seed=0):
        out = np.empty((n, n), NP_LP[dtype]); fin = np.empty((n, n), NP_LP[dtype])
        assert self.L.prko_run_lp(dtype, pattern, radius, n, iterations, init, seed,
                                  out.ctypes.data, fin.ctypes.data) == 0
        return out, fin

    def band_lp(self, dtype, pattern, radius, n, iterations, init, seed, a, b):
        out = np.empty((b - a, n), NP_LP[dtype])
        assert self.L.prko_band_lp(dtype, pattern, radius, n, iterations, init, seed, a, b,
                                   out.ctypes.data) == 0
        return out

    def sweep_lp(self, dtype, pattern, radius, inp, out) -> None:
        """One reduced-precision sweep out += S(inp) in place (n x n arrays of
        the storage dtype)."""
        n = inp.shape[0]
        assert inp.dtype == NP_LP[dtype] and out.dtype == NP_LP[dtype]
        assert inp.shape == (n, n) and out.shape == (n, n)
        assert inp.flags.c_contiguous and out.flags.c_contiguous
        assert self.L.prko_sweep_lp(dtype, pattern, radius, n, inp.ctypes.data, out.ctypes.data) == 0

    def l1norm_lp(self, dtype, n, radius, out) -> float:
        return self.L.prko_l1norm_lp(dtype, n, radius, np.ascontiguousarray(out).ctypes.data)

    def f32_to_bf16(self, x: float) -> int:
        return self.L.prko_f32_to_bf16(x)

    def bf16_to_f32(self, h: int) -> float:
        return self.L.prko_bf16_to_f32(h)

    def transpose(self, order, tile, A, B) -> None:
        assert self.L.prko_transpose(order, tile, _d(A), _d(B)) == 0

    def transpose_run(self, order, tile, iterations):
        A = np.empty((order, order)); B = np.empty((order, order))
        assert self.L.prko_transpose_run(order, tile, iterations, _d(A), _d(B)) == 0
        return A, B

    def transpose_abserr(self, order, iterations, B) -> float:
        return self.L.prko_transpose_abserr(order, iterations, _d(np.ascontiguousarray(B)))

    def omp_run(self, pattern, radius, n, t, iterations, threads=0):
        avg = ctypes.c_double(); norm = ctypes.c_double(); team = ctypes.c_int()
        assert self.L.prko_omp_run(pattern, radius, n, t, iterations, threads, ctypes.byref(avg),
                                   ctypes.byref(norm), ctypes.byref(team)) == 0
        return avg.value, norm.value, team.value


def bf16_bits_to_f64(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def ref_lib():
    """The reference's generated serial operators behind a C harness, if built."""
    if not REF_SO.exists():
        return None
    L = ctypes.CDLL(str(REF_SO))
    L.prkref_stencil.argtypes = [ctypes.c_int] * 4 + [_dp, _dp]
    return L

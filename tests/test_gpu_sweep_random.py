"""Seeded random sweep configurations against the oracle, bit for bit.

Each case draws a pattern, a radius (1..8), a storage type, n (2R+1..600,
odd or even), a row pitch (n itself, rows padded to 16 bytes, the helpers'
default, or 1 KiB), a pipeline variant (0..8, or the default), a row-block
height, a tile order, a cap on workgroups per CU, the init and 1..3
iterations.  Tiling and layout never change results, so every case must
give the oracle's out and final in exactly: FP64 against the reference's
arithmetic, FP32/BF16 against the oracle's arithmetic of those forms.
"""
import numpy as np
import pytest
import torch

import kernels_amd as K
from tests._oracle import F64

pytestmark = pytest.mark.gpu

ES = {0: 8, 1: 4, 2: 2}


def _cases(count: int = 64, seed: int = 20261203):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(count):
        pattern = int(rng.integers(0, 2))
        radius = int(rng.integers(1, 9))
        dtype = int(rng.integers(0, 3))
        n = int(rng.integers(2 * radius + 1, 601))
        es = ES[dtype]
        pitch = [n, (n * es + 15) // 16 * 16 // es, K.default_pitch(n, dtype),
                 (n * es + 1023) // 1024 * 1024 // es][int(rng.integers(0, 4))]
        variant = int(rng.integers(-1, 9))
        rows = int(rng.choice([0, 1, 3, 4, 6, 8, 13, 32, 64]))
        order = int(rng.integers(-1, 8))
        wgs = int(rng.integers(-1, 5))
        iters = int(rng.integers(1, 4))
        init = int(rng.integers(0, 2))
        cases.append((pattern, radius, dtype, n, pitch, variant, rows, order, wgs, iters, init))
    return cases


def _bits(a: np.ndarray) -> np.ndarray:
    return a.view({8: np.uint64, 4: np.uint32, 2: np.uint16}[a.dtype.itemsize])


def _host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    if t.dtype == torch.bfloat16:
        return t.contiguous().view(torch.int16).cpu().numpy().view(np.uint16)
    return t.contiguous().cpu().numpy()


@pytest.mark.parametrize("case", _cases(), ids=lambda c: "p{}r{}d{}n{}pitch{}v{}rows{}o{}w{}k{}i{}".format(*c))
def test_random_configuration_bitwise(gpu, oracle, case):
    pattern, radius, dtype, n, pitch, variant, rows, order, wgs, iters, init = case
    print(f"random case {case}", flush=True)
    L = K.lib()
    seed = 1000 + n
    try:
        K.check(L.prk_set_sweep_variant(variant))
        K.check(L.prk_set_rows_per_wave(rows))
        K.check(L.prk_set_sweep_order(order, wgs))
        b0, b1, out = (K.alloc_grid(n, dtype, pitch=pitch) for _ in range(3))
        K.init(init, seed, dtype, b0, out)
        fb, _ = K.run(pattern, radius, dtype, iters, b0, b1, out)
        if dtype == F64:
            ref_out, ref_in = oracle.run(pattern, radius, n, 32, iters, init, seed)
        else:
            ref_out, ref_in = oracle.run_lp(dtype, pattern, radius, n, iters, init, seed)
        assert np.array_equal(_bits(_host(out)), _bits(ref_out)), case
        assert np.array_equal(_bits(_host((b0, b1)[fb])), _bits(ref_in)), case
    finally:
        K.check(L.prk_set_sweep_variant(-1))
        K.check(L.prk_set_rows_per_wave(0))
        K.check(L.prk_set_sweep_order(-1, -1))

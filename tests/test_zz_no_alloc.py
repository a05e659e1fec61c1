"""prk_l1norm_noalloc (the norm with its partials in a static device array)
against prk_l1norm (the same kernels on the library's workspace), through the
C ABI: bitwise equal for every storage type, on two calls in a row where the
second has a smaller grid (so partials left by the first call are still in
the array), on the default stream and on a stream of its own."""
import ctypes

import numpy as np
import pytest
import torch

import kernels_amd as K

pytestmark = pytest.mark.gpu


def norm_reference(dtype, radius, out, stream):
    return K.l1norm(dtype, radius, out, stream)


def norm_noalloc(dtype, radius, out, stream):
    d = torch.full((1,), -1.0, dtype=torch.float64, device=out.device)
    n = out.shape[0]
    K.check(K.lib().prk_l1norm_noalloc(dtype, n, out.stride(0), radius, out.data_ptr(), d.data_ptr(),
                                       stream.cuda_stream))
    stream.synchronize()
    return float(d.cpu().item())


def prepared(dtype, n, radius, seed):
    b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
    K.init(K.INIT_SEEDED, seed, dtype, b0, out)
    K.run(K.GRID, radius, dtype, 2, b0, b1, out)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("dtype", [K.F64, K.F32, K.BF16])
@pytest.mark.parametrize("own_stream", [False, True])
def test_noalloc_norm_equals_workspace_norm(gpu, dtype, own_stream):
    stream = torch.cuda.Stream() if own_stream else torch.cuda.current_stream()
    # 2100 uses all 1024 partials; 150 then 37 use fewer, with the first
    # call's partials still in the array
    for n, radius, seed in ((2100, 2, 3), (150, 3, 4), (37, 1, 5)):
        print(f"noalloc dtype={dtype} n={n} R={radius} own_stream={own_stream}", flush=True)
        out = prepared(dtype, n, radius, seed)
        ref = norm_reference(dtype, radius, out, stream)
        got = norm_noalloc(dtype, radius, out, stream)
        assert np.float64(got).view(np.uint64) == np.float64(ref).view(np.uint64), (n, got, ref)
        assert got > 0


def test_noalloc_rejects_bad_arguments(gpu):
    lib = K.lib()
    d = torch.zeros(1, dtype=torch.float64, device=gpu)
    out = K.alloc_grid(16, K.F64)
    assert lib.prk_l1norm_noalloc(K.F64, 16, out.stride(0), 8, out.data_ptr(), d.data_ptr(), None) == -1
    assert lib.prk_l1norm_noalloc(K.F64, 16, out.stride(0), 2, out.data_ptr(), None, None) == -1
    assert lib.prk_l1norm_noalloc(9, 16, out.stride(0), 2, out.data_ptr(), d.data_ptr(), None) == -1
    assert b"bad norm arguments" in lib.prk_last_error() or b"dtype" in lib.prk_last_error()

"""The Python wrappers check every buffer before the library sees it (no
GPU needed: the checks run in the order type, dtype, layout, agreement
between buffers, storage size, device, so each refusal shows on CPU
tensors, and a well-formed CPU grid is refused for living on the host)."""
import pytest
import torch

import kernels_amd as K
from kernels_amd.ops import _check_dense, _check_grids


def grid(n, dtype=K.F64, pitch=None):
    pitch = pitch or n
    return torch.zeros((n, pitch), dtype=K.torch_dtype(dtype))[:, :n]


def test_wrong_torch_dtype_is_refused():
    with pytest.raises(ValueError, match="storage type"):
        _check_grids(K.F64, grid(8, K.F32))
    with pytest.raises(ValueError, match="storage type"):
        _check_grids(K.BF16, grid(8, K.F64))


def test_unknown_dtype_and_non_tensor():
    with pytest.raises(ValueError, match="unknown dtype"):
        _check_grids(7, grid(8))
    with pytest.raises(TypeError):
        _check_grids(K.F64, [[0.0]])


def test_layout_is_checked():
    with pytest.raises(ValueError, match="unit column stride"):
        _check_grids(K.F64, torch.zeros((8, 8), dtype=torch.float64).t())
    with pytest.raises(ValueError, match="unit column stride"):
        _check_grids(K.F64, torch.zeros(64, dtype=torch.float64))
    with pytest.raises(ValueError, match="square"):
        _check_grids(K.F64, torch.zeros((8, 9), dtype=torch.float64))


def test_buffers_must_agree():
    with pytest.raises(ValueError, match="shape or row stride"):
        _check_grids(K.F64, grid(8), grid(9))
    with pytest.raises(ValueError, match="shape or row stride"):
        _check_grids(K.F64, grid(8, pitch=8), grid(8, pitch=10))


def test_storage_must_cover_every_row():
    # the kernels read whole padded rows, the last one included: a view of
    # pitch 10 whose storage ends after the last row's 8th element is refused
    short = torch.zeros(7 * 10 + 8, dtype=torch.float64).as_strided((8, 8), (10, 1))
    with pytest.raises(ValueError, match="storage holds"):
        _check_grids(K.F64, short)
    shifted = torch.zeros(8 * 10, dtype=torch.float64)[2:].as_strided((8, 8), (10, 1))
    with pytest.raises(ValueError, match="storage holds"):
        _check_grids(K.F64, shifted)


def test_host_tensors_are_refused_last():
    with pytest.raises(ValueError, match="HIP device"):
        _check_grids(K.F64, grid(8, pitch=10), grid(8, pitch=10), None)


def test_dropin_layout():
    # the reference's operators take pitch n: a padded grid is refused
    with pytest.raises(ValueError, match="contiguous 8 x 8"):
        _check_dense(8, grid(8, pitch=10))
    with pytest.raises(ValueError, match="contiguous 8 x 8"):
        K.stencil("star2", 8, 32, grid(9), grid(9))
    with pytest.raises(ValueError, match="storage type"):
        _check_dense(8, grid(8, K.F32))


def test_wrappers_validate_before_calling_the_library():
    # K.sweep with an FP32 out and FP64 in/in_next never reaches prk_sweep
    a, b = grid(16), grid(16)
    with pytest.raises(ValueError, match="storage type"):
        K.sweep(K.STAR, 2, K.F64, a, b, grid(16, K.F32))
    with pytest.raises(ValueError, match="shape or row stride"):
        K.run(K.STAR, 2, K.F64, 3, a, grid(16, pitch=18), b)
    with pytest.raises(ValueError, match="HIP device"):
        K.init(K.INIT_ANALYTIC, 0, K.F64, a, b)


def test_default_pitch():
    # rows padded to 128 bytes, rows of 64 KiB or more to 1 KiB
    for dtype, es in ((K.F64, 8), (K.F32, 4), (K.BF16, 2)):
        for n in (1, 7, 64, 100, 1000, 1025, 8191, 8192, 16384, 30000, 46340):
            p = K.default_pitch(n, dtype)
            align = 1024 if n * es >= 65536 else 128
            assert p >= n and (p * es) % align == 0 and (p - n) * es < align, (dtype, n, p)
    assert K.default_pitch(30000, K.F64) == 30080 and K.default_pitch(30000, K.BF16) == 30016
    assert K.default_pitch(16384, K.F32) == 16384

"""The matrix transpose (B += A^T, A += 1) on the GPU through the C ABI,
against the oracle's restatement of Cxx11/transpose.cc and the reference's
own validation sum.  Every value is an exact integer (or a sum of two
doubles done once), so parity is bitwise."""
import numpy as np
import pytest
import torch

import kernels_amd as K

pytestmark = pytest.mark.gpu
ORDERS = [1, 2, 3, 63, 64, 65, 127, 128, 129, 500, 1001]


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("order", ORDERS)
def test_transpose_driver_bitwise(gpu, oracle, order):
    iters = 3
    print(f"transpose order={order}", flush=True)
    A = torch.empty((order, order), dtype=torch.float64, device=gpu)
    B = torch.empty_like(A)
    K.transpose_init(A, B)
    K.transpose_run(iters, A, B)
    rA, rB = oracle.transpose_run(order, 32, iters)
    assert np.array_equal(host(A), rA) and np.array_equal(host(B), rB)
    assert K.transpose_check(iters, B) == 0.0


@pytest.mark.parametrize("order", [5, 64, 130, 777])
def test_transpose_one_iteration_random(gpu, oracle, order):
    # non-integer values: one rounding per element on both sides
    rng = np.random.default_rng(order)
    a = rng.uniform(-1, 1, (order, order)); b = rng.uniform(-1, 1, (order, order))
    A = torch.from_numpy(a.copy()).to(gpu); B = torch.from_numpy(b.copy()).to(gpu)
    # the drop-in form: null stream, tiling hint ignored
    K.check(K.lib().prk_transpose(order, 32, A.data_ptr(), B.data_ptr()))
    oracle.transpose(order, 32, a, b)
    assert np.array_equal(host(A).view(np.uint64), a.view(np.uint64))
    assert np.array_equal(host(B).view(np.uint64), b.view(np.uint64))


def test_transpose_check_sees_errors(gpu):
    order, iters = 1000, 5
    A = torch.empty((order, order), dtype=torch.float64, device=gpu)
    B = torch.empty_like(A)
    K.transpose_init(A, B)
    K.transpose_run(iters, A, B)
    assert K.transpose_check(iters, B) == 0.0
    B[3, 5] += 0.5
    B[999, 0] -= 2.0
    assert K.transpose_check(iters, B) == 2.5


@pytest.mark.parametrize("order", [16384, 46339, 46340])
def test_transpose_large_validates(gpu, order):
    # row offsets beyond 2^31 bytes; the largest order the reference allows
    # (even: 16-byte path; odd: 8-byte path)
    iters = 1 if order > 20000 else 4
    print(f"transpose large order={order}", flush=True)
    A = torch.empty((order, order), dtype=torch.float64, device=gpu)
    B = torch.empty_like(A)
    K.transpose_init(A, B)
    avg = K.transpose_run(iters, A, B)
    assert avg > 0
    assert K.transpose_check(iters, B) == 0.0
    # A = iota + K + 1 on a sample of rows
    rows = [0, order // 2, order - 1]
    got = A[rows].cpu().numpy()
    expect = np.array(rows, dtype=np.float64)[:, None] * order + np.arange(order, dtype=np.float64) + iters + 1
    assert np.array_equal(got, expect)
    del A, B
    torch.cuda.empty_cache()


def test_transpose_rejects_bad_arguments(gpu):
    A = torch.zeros((8, 8), dtype=torch.float64, device=gpu)
    L = K.lib()
    assert L.prk_transpose(0, 32, A.data_ptr(), A.data_ptr() + 8) == -1
    assert L.prk_transpose(8, 32, A.data_ptr(), A.data_ptr()) == -1      # A and B must differ
    assert L.prk_transpose(46341, 32, A.data_ptr(), A.data_ptr() + 8) == -1


def test_transpose_tile_shapes_do_not_change_results(gpu, oracle):
    L = K.lib()
    try:
        for order in (300, 301):
            rA, rB = oracle.transpose_run(order, 32, 2)
            for shape in range(9):
                for mode in range(3):
                    K.check(L.prk_set_transpose_variant(shape, mode))
                    A = torch.empty((order, order), dtype=torch.float64, device=gpu)
                    B = torch.empty_like(A)
                    K.transpose_init(A, B)
                    K.transpose_run(2, A, B)
                    assert np.array_equal(host(A), rA) and np.array_equal(host(B), rB), (order, shape, mode)
    finally:
        K.check(L.prk_set_transpose_variant(-1, -1))
    assert L.prk_set_transpose_variant(9, 0) == -1

"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle (a dropped
tap, a wrong sign, row/column offsets swapped, a different summation order)
fails at least one of them:
  * the term lists parsed from the reference's generated header
    (tests/golden/taps.json) -- order and values;
  * outputs of the reference's own generated operators compiled from the
    reference tree (tests/golden/outputs.npz, and oracle/_ref when built);
  * identities the mathematics fixes: sum W = 0, W(-d) = -W(d), W = W^T,
    first moments 1, closed forms S(linear) and S(i^2);
  * the driver's analytic validation norm = 2(K+1) (Cxx11/stencil.cc:209-214)
    for every operator, as the reference CI runs them
    (ci/build-run-prk.sh:340-345: 10 iterations, n=200, tile 20);
  * SplitMix64's published output sequence for the seeded init.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests._oracle import BF16, F32, REPO, bf16_bits_to_f64, ref_lib

GOLDEN = Path(__file__).resolve().parent / "golden"
OPS = [(p, r) for p in (0, 1) for r in range(1, 6)]
# radii 6..8 extend the reference's table with the generator's formulas; the
# checks that need no reference data cover them too
OPS_ALL = [(p, r) for p in (0, 1) for r in range(1, 9)]
NAME = {0: "star", 1: "grid"}


@pytest.mark.parametrize("pattern,radius", OPS)
def test_taps_match_reference_header(oracle, pattern, radius):
    golden = json.loads((GOLDEN / "taps.json").read_text())["operators"][f"{NAME[pattern]}{radius}"]
    mine = oracle.taps(pattern, radius)
    assert len(mine) == len(golden)
    for (di, dj, w), (gdi, gdj, gw) in zip(mine, golden):
        assert (di, dj) == (gdi, gdj)
        assert w == gw  # the literal round-trips to the same double


@pytest.mark.parametrize("pattern,radius", OPS_ALL)
def test_tap_order_is_generator_order(oracle, pattern, radius):
    # column offset outer, row offset inner, every offset inside the square
    taps = oracle.taps(pattern, radius)
    keys = [(dj, di) for di, dj, _ in taps]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    assert all(max(abs(di), abs(dj)) <= radius for di, dj, _ in taps)


@pytest.mark.parametrize("pattern,radius", OPS_ALL)
def test_weight_identities(oracle, pattern, radius):
    taps = oracle.taps(pattern, radius)
    nnz = 4 * radius if pattern == 0 else 4 * radius * radius + 2 * radius
    assert len(taps) == nnz
    W = {(di, dj): w for di, dj, w in taps}
    assert abs(sum(W.values())) < 1e-15
    for (di, dj), w in W.items():
        assert W[(-di, -dj)] == -w          # antisymmetric
        assert W[(dj, di)] == w             # symmetric under transpose
    assert abs(sum(w * di for (di, dj), w in W.items()) - 1.0) < 1e-14
    assert abs(sum(w * dj for (di, dj), w in W.items()) - 1.0) < 1e-14
    # sum |w| (harmonic numbers): star H_R / R, grid 3/4 * H_R / R ... closed form
    H = sum(1.0 / k for k in range(1, radius + 1))
    expect = 2 * H / radius if pattern == 0 else 1.5 * H / radius
    assert abs(sum(abs(w) for w in W.values()) - expect) < 1e-14


@pytest.mark.parametrize("pattern,radius", OPS)
def test_oracle_matches_reference_outputs(oracle, pattern, radius):
    g = np.load(GOLDEN / "outputs.npz")
    x, o = g["in"].copy(), g["out0"].copy()
    n = x.shape[0]
    oracle.stencil(pattern, radius, n, 5, x, o)
    expect = g[f"{NAME[pattern]}{radius}"]
    assert np.array_equal(o.view(np.uint64), expect.view(np.uint64))


@pytest.mark.parametrize("pattern,radius", OPS)
@pytest.mark.parametrize("n,t", [(11, 3), (40, 32), (65, 7)])
def test_oracle_matches_reference_build(oracle, pattern, radius, n, t):
    ref = ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    if 2 * radius + 1 > n:
        pytest.skip("radius too large")
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(1000 * n + 10 * radius + pattern)
    x = rng.uniform(-1, 1, (n, n)) * 1e3
    o0 = rng.uniform(-1, 1, (n, n))
    a, b = o0.copy(), o0.copy()
    oracle.stencil(pattern, radius, n, t, x, a)
    assert ref.prkref_stencil(pattern, radius, n, t, x.ctypes.data_as(dp), b.ctypes.data_as(dp)) == 0
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("pattern,radius", OPS_ALL)
def test_closed_forms(oracle, pattern, radius):
    n = 48
    i, j = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    inner = (slice(radius, n - radius),) * 2
    # S(cx*i + cy*j + c) = cx + cy
    x = 3.0 * i - 5.0 * j + 7.0
    o = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, 16, x, o)
    np.testing.assert_allclose(o[inner], -2.0, rtol=1e-13)
    assert np.all(o[:radius] == 0) and np.all(o[:, :radius] == 0)  # frame untouched
    assert np.all(o[n - radius:] == 0) and np.all(o[:, n - radius:] == 0)
    # S(i^2) = 2i  (sum w*b^2 = 0 by antisymmetry)
    o = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, 16, i * i, o)
    np.testing.assert_allclose(o[inner], 2.0 * i[inner], rtol=1e-13, atol=1e-12)
    # S(j^2) = 2j ; S(i*j) = i + j
    o = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, 16, j * j, o)
    np.testing.assert_allclose(o[inner], 2.0 * j[inner], rtol=1e-13, atol=1e-12)
    o = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, 16, i * j, o)
    np.testing.assert_allclose(o[inner], (i + j)[inner], rtol=1e-13, atol=1e-12)


def test_tile_size_does_not_change_results(oracle):
    n = 37
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (n, n))
    outs = []
    for t in (1, 5, 32, n):
        o = np.zeros((n, n))
        oracle.stencil(1, 3, n, t, x, o)
        outs.append(o)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("pattern,radius", OPS_ALL)
def test_driver_validates_like_reference_ci(oracle, pattern, radius):
    # ci/build-run-prk.sh:340-345: <prog> 10 200 20 {star,grid} {1..5}
    out, fin = oracle.run(pattern, radius, 200, 20, 10, 0)
    norm = oracle.l1norm(200, radius, out)
    assert abs(norm - 2.0 * 11) <= 1e-8
    i, j = np.meshgrid(np.arange(200.0), np.arange(200.0), indexing="ij")
    assert np.array_equal(fin, i + j + 11.0)  # in refreshed K+1 times


def test_config1_norm(oracle):
    # 10 iterations, n=1000, tile 32, star 2 (Cxx11/stencil.cc:209-216)
    out, _ = oracle.run(0, 2, 1000, 32, 10, 0)
    assert oracle.l1norm(1000, 2, out) == 22.0
    assert np.all(out[2:-2, 2:-2] == 22.0)  # dyadic weights: exact at every point


@pytest.mark.parametrize("pattern,radius", [(0, 2), (1, 3), (0, 5)])
def test_band_equals_full_run(oracle, pattern, radius):
    n, K = 60, 4
    out, _ = oracle.run(pattern, radius, n, 8, K, 1, 99)
    for a, b in [(0, 5), (radius, radius + 3), (27, 33), (n - 6, n)]:
        band = oracle.band(pattern, radius, n, K, 1, 99, a, b)
        assert np.array_equal(band, out[a:b])


def test_seeded_value_is_splitmix64(oracle):
    # Published SplitMix64 outputs for seed 1234567 (Vigna's reference
    # implementation, splitmix64.c): the k-th output feeds element k-1.
    known = [6457827717110365317, 3203168211198807973, 9817491932198370423,
             4593380528125082431, 16408922859458223821]
    for k, z in enumerate(known):
        u = (z >> 11) * 2.0 ** -53
        assert oracle.seeded_value(1234567, k) == 2.0 * u - 1.0


def test_seeded_init_range(oracle):
    a, o = oracle.init(1, 42, 64)
    assert np.all(a >= -1.0) and np.all(a < 1.0) and np.all(o == 0.0)
    assert abs(a.mean()) < 0.05 and 0.3 < a.std() < 0.7


def test_bf16_rounding(oracle):
    f = oracle.f32_to_bf16
    assert f(1.0) == 0x3F80
    assert f(-2.0) == 0xC000
    assert f(1.0 + 2.0 ** -8) == 0x3F80          # tie -> even
    assert f(1.0 + 3 * 2.0 ** -8) == 0x3F82      # tie -> even (up)
    assert f(1.0 + 2.0 ** -8 + 2.0 ** -20) == 0x3F81
    assert f(float("inf")) == 0x7F80
    assert (f(float("nan")) & 0x7FC0) == 0x7FC0
    assert oracle.bf16_to_f32(0x4049) == np.float32(3.140625)


def test_lp_forms_close_to_f64(oracle):
    n, K = 64, 6
    ref, _ = oracle.run(1, 2, n, 32, K, 1, 5)
    inner = (slice(2, n - 2),) * 2
    scale = np.abs(ref[inner]).max()
    f32, _ = oracle.run_lp(F32, 1, 2, n, K, 1, 5)
    assert np.abs(f32.astype(np.float64)[inner] - ref[inner]).max() < 1e-5 * scale
    b16, _ = oracle.run_lp(BF16, 1, 2, n, K, 1, 5)
    assert np.abs(bf16_bits_to_f64(b16)[inner] - ref[inner]).max() < 5e-2 * scale
    # analytic init, star 2: dyadic weights and integer data -> exact in FP32
    f32a, _ = oracle.run_lp(F32, 0, 2, n, K, 0)
    assert np.all(f32a[inner] == 2.0 * (K + 1))
    # BF16, star 1, analytic: in = i + j + k stays below 2^8 (exact in an
    # 8-bit significand), the weights are +-1/2 and every partial sum is a
    # small integer, so every interior point is 2(K+1) and in ends at
    # i + j + K + 1 exactly
    m = 40
    b16a, b16in = oracle.run_lp(BF16, 0, 1, m, K, 0)
    assert np.all(bf16_bits_to_f64(b16a)[1:-1, 1:-1] == 2.0 * (K + 1))
    i = np.arange(m, dtype=np.float64)
    assert np.array_equal(bf16_bits_to_f64(b16in), i[:, None] + i[None, :] + K + 1)


@pytest.mark.parametrize("pattern,radius", OPS_ALL)
def test_lp_operator_on_exact_inputs(oracle, pattern, radius):
    """One FP32 sweep against the FP64 operator on integer inputs.

    Every sum and difference the FP32 form builds from integers in
    [-1000, 1000] stays below 2^24 in magnitude, so it is exact: the only
    roundings are the weights' to float and the multiply / fma steps (2R for
    a grid, one per tap pair for a star).  So at every point the result is
    within (steps + 1) * 2^-24 * sum_k |w_k x_k| of the FP64 operator.  A
    dropped, doubled or misplaced tap, or a wrong sign, moves a point by
    |w_k x_k|, at least 2.6e-4 * |x_k|: far beyond the bound at almost every
    point.  The BF16 form on BF16-exact integers is the FP32 result rounded
    to BF16, point for point."""
    n = 2 * radius + 1 + 24
    rng = np.random.default_rng(700 + 10 * pattern + radius)
    x = rng.integers(-1000, 1001, (n, n)).astype(np.float64)
    ref = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, n, x, ref)
    out = np.zeros((n, n), np.float32)
    oracle.sweep_lp(F32, pattern, radius, x.astype(np.float32), out)
    inner = (slice(radius, n - radius),) * 2
    A = np.zeros((n, n))
    for di, dj, w in oracle.taps(pattern, radius):
        A[inner] += abs(w) * np.abs(np.roll(x, (-di, -dj), axis=(0, 1)))[inner]
    steps = 2 * radius  # a grid's rings, or a star's tap pairs
    err = np.abs(out.astype(np.float64) - ref)
    assert np.all(err[inner] <= (steps + 1) * 2.0 ** -24 * A[inner])
    assert np.count_nonzero(A[inner] > 0) == (n - 2 * radius) ** 2
    frame = np.ones((n, n), bool)
    frame[inner] = False
    assert np.all(out[frame] == 0.0)
    # BF16: integers in [-128, 128] are exact with an 8-bit significand
    xb = rng.integers(-128, 129, (n, n)).astype(np.float32)
    f = np.zeros((n, n), np.float32)
    oracle.sweep_lp(F32, pattern, radius, xb, f)
    bits = np.array([oracle.f32_to_bf16(float(v)) for v in xb.ravel()], np.uint16).reshape(n, n)
    ob = np.zeros((n, n), np.uint16)
    oracle.sweep_lp(BF16, pattern, radius, bits, ob)
    expect = np.array([oracle.f32_to_bf16(float(v)) for v in f.ravel()], np.uint16).reshape(n, n)
    assert np.array_equal(ob, expect)


def test_lp_band_equals_full(oracle):
    n, K = 50, 3
    full, _ = oracle.run_lp(F32, 1, 3, n, K, 1, 8)
    band = oracle.band_lp(F32, 1, 3, n, K, 1, 8, 20, 26)
    assert np.array_equal(full[20:26], band)


def test_reference_binaries_validate():
    exe = REPO / "oracle" / "_ref" / "stencil"
    if not exe.exists():
        pytest.skip("oracle/_ref not built")
    for args in (["10", "1000", "32", "star", "2"], ["10", "200", "20", "grid", "5"]):
        r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and "Solution validates" in r.stdout


def test_openmp_baseline_validates(oracle):
    avg, norm, team = oracle.omp_run(0, 2, 400, 32, 3, 2)
    assert norm == 8.0 and team == 2 and avg > 0


# ---------------------------------------------------------------------------
# Matrix transpose (Cxx11/transpose.cc)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order,tile", [(1, 32), (7, 3), (33, 5), (64, 32), (65, 32), (100, 0), (100, 100)])
def test_transpose_oracle_is_b_plus_a_transposed(oracle, order, tile):
    # random non-symmetric A: swapped loop indices or a missed refresh fail
    rng = np.random.default_rng(order * 100 + tile)
    A = rng.integers(-1000, 1000, (order, order)).astype(np.float64)
    B = rng.integers(-1000, 1000, (order, order)).astype(np.float64)
    A0, B0 = A.copy(), B.copy()
    oracle.transpose(order, tile, A, B)
    assert np.array_equal(B, B0 + A0.T)
    assert np.array_equal(A, A0 + 1.0)


def test_transpose_driver_validates(oracle):
    order, K = 100, 10
    A, B = oracle.transpose_run(order, 32, K)
    assert oracle.transpose_abserr(order, K, B) == 0.0
    # closed form: B[j][i] = (i*order + j)(K+1) + K(K+1)/2, A = iota + K + 1
    i = np.arange(order, dtype=np.float64)
    assert np.array_equal(B, (i[None, :] * order + i[:, None]) * (K + 1) + K * (K + 1) / 2)
    assert np.array_equal(A, np.arange(order * order, dtype=np.float64).reshape(order, order) + K + 1)
    # the check sees a planted error of 0.5
    B[3, 5] += 0.5
    assert oracle.transpose_abserr(order, K, B) == 0.5


@pytest.mark.parametrize("exe", ["transpose", "transpose-openmp"])
def test_reference_transpose_binaries_validate(exe):
    p = REPO / "oracle" / "_ref" / exe
    if not p.exists():
        pytest.skip("oracle/_ref not built (no reference tree here)")
    for args in (["10", "1000", "32"], ["4", "257", "0"]):
        r = subprocess.run([str(p), *args], capture_output=True, text=True, timeout=120,
                           env={"OMP_NUM_THREADS": "4", "PATH": "/usr/bin:/bin"})
        assert r.returncode == 0 and "Solution validates" in r.stdout, r.stdout

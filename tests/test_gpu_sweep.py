"""GPU parity of the stencil path against the CPU oracle, through the C ABI.

FP64 is bitwise: the kernels sum the taps in the generator's order with no
contraction, exactly as the oracle (pinned to the reference's own generated
operators in test_oracle.py).  FP32/BF16 storage forms are bitwise against
the oracle's restatement of the same FP32 arithmetic (paired sums for the
stars, ring sums for the grids), and their error against the FP64 oracle is
bounded separately.
"""
import numpy as np
import pytest
import torch

import kernels_amd as K
from tests._oracle import BF16, F32, bf16_bits_to_f64

pytestmark = pytest.mark.gpu
OPS = [(p, r) for p in (0, 1) for r in range(1, 6)]
NAME = {0: "star", 1: "grid"}


def bits(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32 if a.dtype == np.float32 else np.uint16)


def to_dev(a: np.ndarray, gpu) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(gpu)


def to_host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    if t.dtype == torch.bfloat16:
        return t.contiguous().view(torch.int16).cpu().numpy().view(np.uint16)
    return t.contiguous().cpu().numpy()


@pytest.mark.parametrize("pattern,radius", OPS)
@pytest.mark.parametrize("n", [11, 64, 129, 333, 700])
def test_dropin_operator_bitwise(gpu, oracle, pattern, radius, n):
    if 2 * radius + 1 > n:
        pytest.skip("radius too large")
    print(f"dropin {NAME[pattern]}{radius} n={n}", flush=True)
    rng = np.random.default_rng(n * 100 + radius * 10 + pattern)
    x = rng.uniform(-1, 1, (n, n)) * 10.0
    o = rng.uniform(-1, 1, (n, n))
    xd, od = to_dev(x, gpu), to_dev(o, gpu)
    K.stencil(f"{NAME[pattern]}{radius}", n, 32, xd, od)
    oracle.stencil(pattern, radius, n, 32, x, o)
    got = to_host(od)
    assert np.array_equal(bits(got), bits(o)), np.abs(got - o).max()
    assert np.array_equal(to_host(xd), x)  # in is read-only


@pytest.mark.parametrize("n", [5, 64, 257, 1000])
def test_dropin_add_bitwise(gpu, oracle, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1e3, 1e3, (n, n))
    xd = to_dev(x, gpu)
    K.add(n, 32, xd)
    oracle.add(n, x)
    assert np.array_equal(bits(to_host(xd)), bits(x))


@pytest.mark.parametrize("pattern,radius", OPS)
@pytest.mark.parametrize("n,pitch", [(64, 64), (100, 104), (129, 130), (515, 520), (37, 37)])
def test_fused_sweep_equals_stencil_plus_add(gpu, oracle, pattern, radius, n, pitch):
    if 2 * radius + 1 > n:
        pytest.skip("radius too large")
    print(f"fused {NAME[pattern]}{radius} n={n} pitch={pitch}", flush=True)
    rng = np.random.default_rng(7 * n + radius)
    x = rng.uniform(-1, 1, (n, n))
    o = rng.uniform(-1, 1, (n, n))
    xin = K.alloc_grid(n, K.F64, pitch=pitch, fill=float("nan"))
    xnx = K.alloc_grid(n, K.F64, pitch=pitch, fill=-7.0)
    od = K.alloc_grid(n, K.F64, pitch=pitch, fill=float("nan"))
    xin.copy_(to_dev(x, gpu))
    od.copy_(to_dev(o, gpu))
    K.sweep(pattern, radius, K.F64, xin, xnx, od)
    oracle.stencil(pattern, radius, n, 32, x, o)
    assert np.array_equal(bits(to_host(od)), bits(o))
    oracle.add(n, x)
    assert np.array_equal(bits(to_host(xnx)), bits(x))
    # padding columns n..pitch-1 are never written
    if pitch > n:
        pad = xnx.as_strided((n, pitch - n), (pitch, 1), xnx.storage_offset() + n)
        assert torch.all(pad == -7.0)


@pytest.mark.parametrize("pattern,radius", OPS)
@pytest.mark.parametrize("init,seed", [(0, 0), (1, 42)])
def test_run_bitwise(gpu, oracle, pattern, radius, init, seed):
    n, iters = 150, 5
    print(f"run {NAME[pattern]}{radius} init={init}", flush=True)
    b0 = K.alloc_grid(n, K.F64)
    b1 = K.alloc_grid(n, K.F64)
    out = K.alloc_grid(n, K.F64)
    K.init(init, seed, K.F64, b0, out)
    fb, avg = K.run(pattern, radius, K.F64, iters, b0, b1, out)
    ref_out, ref_in = oracle.run(pattern, radius, n, 32, iters, init, seed)
    assert np.array_equal(bits(to_host(out)), bits(ref_out))
    assert np.array_equal(bits(to_host((b0, b1)[fb])), bits(ref_in))
    assert fb == (iters + 1) % 2 and avg > 0
    norm = K.l1norm(K.F64, radius, out)
    ref_norm = oracle.l1norm(n, radius, ref_out)
    assert abs(norm - ref_norm) <= 1e-14 * abs(ref_norm)
    if init == 0:
        assert abs(norm - 2.0 * (iters + 1)) <= 1e-8


def test_init_matches_oracle(gpu, oracle):
    n = 77
    for kind, seed in ((0, 0), (1, 5), (1, 2**63 + 11)):
        a = K.alloc_grid(n, K.F64)
        o = K.alloc_grid(n, K.F64, fill=3.0)
        K.init(kind, seed, K.F64, a, o)
        ra, ro = oracle.init(kind, seed, n)
        assert np.array_equal(bits(to_host(a)), bits(ra))
        assert np.all(to_host(o) == 0.0)


def test_config1_on_gpu(gpu):
    # 10 iterations, n=1000, tile 32, star 2: the reference's CPU-runnable case
    n = 1000
    b0, b1, out = K.alloc_grid(n), K.alloc_grid(n), K.alloc_grid(n)
    K.init(K.INIT_ANALYTIC, 0, K.F64, b0, out)
    K.run(K.STAR, 2, K.F64, 10, b0, b1, out)
    assert K.l1norm(K.F64, 2, out) == 22.0
    assert torch.all(out[2:-2, 2:-2] == 22.0)


@pytest.mark.parametrize("dtype", [F32, BF16])
@pytest.mark.parametrize("pattern,radius", OPS)
def test_storage_forms_bitwise(gpu, oracle, dtype, pattern, radius):
    n, iters = 97, 4
    print(f"lp dtype={dtype} {NAME[pattern]}{radius}", flush=True)
    b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
    K.init(K.INIT_SEEDED, 17, dtype, b0, out)
    fb, _ = K.run(pattern, radius, dtype, iters, b0, b1, out)
    ref_out, ref_in = oracle.run_lp(dtype, pattern, radius, n, iters, 1, 17)
    assert np.array_equal(to_host(out).view(ref_out.dtype), ref_out)
    assert np.array_equal(to_host((b0, b1)[fb]).view(ref_in.dtype), ref_in)
    norm = K.l1norm(dtype, radius, out)
    assert abs(norm - oracle.l1norm_lp(dtype, n, radius, ref_out)) <= 1e-13 * abs(norm)


# normwise error against FP64 of this case: FP32 2.23e-7, BF16 0.0145
# (from the oracle's FP32/BF16 arithmetic, which the kernels match bit for
# bit); the bounds are 1.5x those
@pytest.mark.parametrize("dtype,bound", [(F32, 3.4e-7), (BF16, 0.022)])
def test_storage_form_error_vs_f64(gpu, oracle, dtype, bound):
    n, iters, pattern, radius = 128, 10, 1, 2
    b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
    K.init(K.INIT_SEEDED, 3, dtype, b0, out)
    K.run(pattern, radius, dtype, iters, b0, b1, out)
    ref, _ = oracle.run(pattern, radius, n, 32, iters, 1, 3)
    got = to_host(out)
    got = bf16_bits_to_f64(got) if dtype == BF16 else got.astype(np.float64)
    inner = (slice(radius, n - radius),) * 2
    err = np.abs(got[inner] - ref[inner])
    normwise = np.linalg.norm(err) / np.linalg.norm(ref[inner])
    print(f"dtype={dtype} normwise={normwise:.3e} max={err.max():.3e}")
    assert normwise < bound


def test_triad_and_copy(gpu):
    n = 1 << 20
    A = torch.full((n + 1,), 1.0, dtype=torch.float64, device=gpu)
    B = torch.full((n + 1,), 2.0, dtype=torch.float64, device=gpu)
    C = torch.full((n + 1,), 2.0, dtype=torch.float64, device=gpu)
    for _ in range(3):
        K.check(K.lib().prk_triad(n + 1, 3.0, A.data_ptr(), B.data_ptr(), C.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.all(A == 1.0 + 3 * 8.0)
    D = torch.empty_like(A)
    K.check(K.lib().prk_copy(n * 8, D.data_ptr(), A.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.equal(D[:n], A[:n])
    # the sweep's stream mix: in_next = in + 1, out += in (odd length: tail element)
    x = torch.arange(n + 1, dtype=torch.float64, device=gpu)
    nx = torch.zeros_like(x)
    out = torch.full_like(x, 5.0)
    for _ in range(2):
        K.check(K.lib().prk_stream_mix(n + 1, x.data_ptr(), nx.data_ptr(), out.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.equal(nx, x + 1) and torch.equal(out, 5.0 + 2 * x)


def test_autotune_on_caller_grids(gpu, oracle):
    # prk_autotune_buffers tunes on the caller's grids (overwriting them) and
    # records the choice; a run on the same grids afterwards is unchanged
    n, iters, pattern, radius = 257, 2, 0, 4
    L = K.lib()
    try:
        b0, b1, out = (K.alloc_grid(n, K.F64) for _ in range(3))
        t = K.autotune(pattern, radius, K.F64, b0, b1, out, reps=1)
        got = K.Tuning()
        K.check(L.prk_tuning_get(pattern, radius, K.F64, n, got))
        assert (got.variant, got.rows_per_wave, got.wgs_per_cu, got.order) == \
            (t.variant, t.rows_per_wave, t.wgs_per_cu, t.order) and t.gbs > 0
        K.init(K.INIT_SEEDED, 4, K.F64, b0, out)
        K.run(pattern, radius, K.F64, iters, b0, b1, out)
        ref, _ = oracle.run(pattern, radius, n, 32, iters, 1, 4)
        assert np.array_equal(bits(to_host(out)), bits(ref))
    finally:
        K.check(L.prk_tuning_clear())


def test_autotune_and_tiling_do_not_change_results(gpu, oracle):
    n, iters, pattern, radius = 300, 3, 1, 3
    L = K.lib()
    best = K.Tuning()
    K.check(L.prk_autotune(pattern, radius, K.F64, n, 1, best))
    assert best.gbs > 0 and 1 <= best.rows_per_wave <= 32
    ref, _ = oracle.run(pattern, radius, n, 32, iters, 1, 9)
    try:
        for v, rows, order, wgs in [(-1, 0, -1, -1), (0, 1, 0, 0), (3, 5, 2, 1), (5, 64, 1, 4), (2, 300, 1, 2),
                                    (4, 8, 3, 3), (4, 8, 4, 3), (5, 4, 5, 2), (4, 3, 6, 3), (1, 7, 7, 4),
                                    (8, 8, 1, 3), (8, 5, 3, 2), (8, 300, 0, 4)]:
            K.check(L.prk_set_sweep_variant(v)); K.check(L.prk_set_rows_per_wave(rows))
            K.check(L.prk_set_sweep_order(order, wgs))
            b0, b1, out = (K.alloc_grid(n, K.F64) for _ in range(3))
            K.init(K.INIT_SEEDED, 9, K.F64, b0, out)
            K.run(pattern, radius, K.F64, iters, b0, b1, out)
            assert np.array_equal(bits(to_host(out)), bits(ref)), (v, rows, order, wgs)
    finally:
        K.check(L.prk_set_sweep_variant(-1)); K.check(L.prk_set_rows_per_wave(0))
        K.check(L.prk_set_sweep_order(-1, -1)); K.check(L.prk_tuning_clear())


@pytest.mark.parametrize("dtype", [F32, BF16])
@pytest.mark.parametrize("radius", [1, 2, 3, 4, 5, 7])
def test_grid_ring_sums_every_tiling(gpu, oracle, dtype, radius):
    # The FP32/BF16 grid ring sums in every pipeline variant of the LDS-ring
    # sweep (1-4 rows in flight, cached or non-temporal stores) and, at
    # R <= 3, the register window, over several row heights, tile orders and
    # occupancies: bitwise the oracle's every time (BF16 grid R=5 and R=7
    # read their windows whole, the others narrow).
    n, iters, pattern = 301, 3, 1
    L = K.lib()
    ref, _ = oracle.run_lp(dtype, pattern, radius, n, iters, 1, 21)
    tilings = [(0, 1, 0, 0), (1, 5, 2, 1), (2, 64, 1, 4), (3, 7, 3, 3), (4, 8, 1, 2), (5, 12, 1, 4), (5, 300, 0, 1),
               (8, 8, 1, 3), (8, 6, 3, 2)]
    if radius <= 3:
        tilings += [(6, 6, 1, 4), (7, 13, 1, 2)]
    try:
        for v, rows, order, wgs in tilings:
            K.check(L.prk_set_sweep_variant(v)); K.check(L.prk_set_rows_per_wave(rows))
            K.check(L.prk_set_sweep_order(order, wgs))
            b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
            K.init(K.INIT_SEEDED, 21, dtype, b0, out)
            K.run(pattern, radius, dtype, iters, b0, b1, out)
            assert np.array_equal(bits(to_host(out)), bits(ref)), (v, rows, order, wgs)
    finally:
        K.check(L.prk_set_sweep_variant(-1)); K.check(L.prk_set_rows_per_wave(0))
        K.check(L.prk_set_sweep_order(-1, -1)); K.check(L.prk_tuning_clear())


EXT = [(p, r) for p in (0, 1) for r in (6, 7, 8)]


@pytest.mark.parametrize("pattern,radius", EXT)
@pytest.mark.parametrize("n", [90, 91])
def test_extended_radius_bitwise(gpu, oracle, pattern, radius, n):
    # radii 6..8: the generator's formulas beyond the reference's table,
    # FP64 and both storage forms bitwise against the oracle
    iters = 3
    print(f"ext {NAME[pattern]}{radius} n={n}", flush=True)
    b0, b1, out = (K.alloc_grid(n, K.F64) for _ in range(3))
    K.init(K.INIT_SEEDED, 5, K.F64, b0, out)
    fb, _ = K.run(pattern, radius, K.F64, iters, b0, b1, out)
    ref_out, ref_in = oracle.run(pattern, radius, n, 32, iters, 1, 5)
    assert np.array_equal(bits(to_host(out)), bits(ref_out))
    assert np.array_equal(bits(to_host((b0, b1)[fb])), bits(ref_in))
    # the drop-in switch with pitch n
    x = to_dev(ref_in, gpu)
    o = torch.zeros((n, n), dtype=torch.float64, device=gpu)
    K.stencil(pattern, n, 32, x, o, radius=radius)
    oo = np.zeros((n, n))
    oracle.stencil(pattern, radius, n, 32, ref_in.copy(), oo)
    assert np.array_equal(bits(to_host(o)), bits(oo))
    for dtype in (F32, BF16):
        b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
        K.init(K.INIT_SEEDED, 5, dtype, b0, out)
        K.run(pattern, radius, dtype, iters, b0, b1, out)
        lp_out, _ = oracle.run_lp(dtype, pattern, radius, n, iters, 1, 5)
        assert np.array_equal(to_host(out).view(lp_out.dtype), lp_out), dtype


# Star radii that sweep from a register ring (FP64 R >= 6, FP32 R >= 5,
# BF16 R >= 6): every row-block height from one row up, each pipeline depth,
# aligned and unaligned pitch, bitwise against the oracle.
RING = [(K.F64, 6), (K.F64, 8), (F32, 4), (F32, 5), (BF16, 7)]


@pytest.mark.parametrize("dtype,radius", RING)
def test_star_register_ring_tilings(gpu, oracle, dtype, radius):
    n, iters, seed = 203, 3, 21
    L = K.lib()
    if dtype == K.F64:
        ref, _ = oracle.run(0, radius, n, 32, iters, 1, seed)
    else:
        ref, _ = oracle.run_lp(dtype, 0, radius, n, iters, 1, seed)
    try:
        for v, rows, pitch in [(4, 1, None), (4, 2, None), (1, 5, n), (4, 13, None), (0, 40, None),
                               (2, 203, n), (5, 7, None), (3, 3, n)]:
            print(f"ring dtype={dtype} star{radius} variant={v} rows={rows} pitch={pitch}", flush=True)
            K.check(L.prk_set_sweep_variant(v)); K.check(L.prk_set_rows_per_wave(rows))
            b0, b1, out = (K.alloc_grid(n, dtype, pitch=pitch) for _ in range(3))
            K.init(K.INIT_SEEDED, seed, dtype, b0, out)
            K.run(0, radius, dtype, iters, b0, b1, out)
            got = to_host(out)
            assert np.array_equal(bits(got), bits(ref)) if dtype == K.F64 else \
                np.array_equal(got.view(ref.dtype), ref), (v, rows, pitch)
    finally:
        K.check(L.prk_set_sweep_variant(-1)); K.check(L.prk_set_rows_per_wave(0))


UNFUSED = [(0, 2), (1, 3), (0, 5), (0, 7), (1, 6)]


@pytest.mark.parametrize("dtype", [K.F64, F32, BF16])
@pytest.mark.parametrize("pattern,radius", UNFUSED)
def test_unfused_forms_equal_fused_sweep(gpu, dtype, pattern, radius):
    # prk_stencil_ex (out += S(in)) then prk_add_ex (in += 1), in any dtype,
    # give bit for bit what one fused sweep gives (the fused sweep is pinned
    # to the oracle above); pitch n covers the element-access path.
    L = K.lib()
    for n, pitch in [(131, None), (131, 131)]:
        print(f"unfused dtype={dtype} {NAME[pattern]}{radius} n={n} pitch={pitch}", flush=True)
        b0, b1, out = (K.alloc_grid(n, dtype, pitch=pitch) for _ in range(3))
        K.init(K.INIT_SEEDED, 33, dtype, b0, out)
        xs = K.alloc_grid(n, dtype, pitch=pitch)
        os_ = K.alloc_grid(n, dtype, pitch=pitch)
        xs.copy_(b0)
        os_.copy_(out)
        K.sweep(pattern, radius, dtype, b0, b1, out)
        stream = torch.cuda.current_stream()
        K.stencil_ex(pattern, radius, dtype, xs, os_, stream)
        K.check(L.prk_add_ex(dtype, n, xs.stride(0), xs.data_ptr(), stream.cuda_stream))
        torch.cuda.synchronize()
        assert np.array_equal(bits(to_host(os_)), bits(to_host(out)))
        assert np.array_equal(bits(to_host(xs)), bits(to_host(b1)))


# The register-window sweep (variants 6 and 7: the FP32/BF16 forms at
# R <= 3, each row widened once into a ring of windows held in registers,
# neighbour columns by DPP lane shifts): bitwise against the oracle for
# star and grid R=1..3, one strip (lanes past n, both halo lanes at the
# grid's edge) and several strips (interior strips whose neighbours come
# from the halo vectors of the adjacent strips), row blocks from one row to
# the whole grid, both pipeline depths.
RW = [(p, r) for p in (0, 1) for r in (1, 2, 3)]


@pytest.mark.parametrize("dtype", [F32, BF16])
@pytest.mark.parametrize("pattern,radius", RW)
def test_register_window_tilings(gpu, oracle, dtype, pattern, radius):
    L = K.lib()
    iters, seed = 3, 29
    try:
        for n in (40, 1104):  # multiples of the vector width: the whole-vector path
            ref, ref_in = oracle.run_lp(dtype, pattern, radius, n, iters, 1, seed)
            for v, rows in [(6, 1), (7, 2), (6, 7), (7, 16), (6, 64), (7, n)]:
                print(f"regwin dtype={dtype} {NAME[pattern]}{radius} n={n} variant={v} rows={rows}", flush=True)
                K.check(L.prk_set_sweep_variant(v)); K.check(L.prk_set_rows_per_wave(rows))
                b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
                K.init(K.INIT_SEEDED, seed, dtype, b0, out)
                fb, _ = K.run(pattern, radius, dtype, iters, b0, b1, out)
                assert np.array_equal(to_host(out).view(ref.dtype), ref), (n, v, rows)
                assert np.array_equal(to_host((b0, b1)[fb]).view(ref_in.dtype), ref_in), (n, v, rows)
    finally:
        K.check(L.prk_set_sweep_variant(-1)); K.check(L.prk_set_rows_per_wave(0))


@pytest.mark.parametrize("dtype", [K.F64, F32, BF16])
def test_frame_columns_keep_their_bits(gpu, oracle, dtype):
    # out's frame (the first and last R rows and columns) is never written:
    # a sentinel there survives every sweep variant, and the interior still
    # matches the oracle (the frame of out feeds no interior point)
    L = K.lib()
    n, iters, seed = 616, 2, 8  # a multiple of every vector width: the whole-vector path
    cases = [(0, 2, v) for v in (1, 4, 5, 8)] + [(1, 3, 5), (1, 3, 8), (0, 7, 4)]
    if dtype != K.F64:
        cases += [(1, 2, 6), (0, 3, 7)]
    try:
        for pattern, radius, v in cases:
            print(f"frame dtype={dtype} {NAME[pattern]}{radius} variant={v}", flush=True)
            K.check(L.prk_set_sweep_variant(v))
            b0, b1, out = (K.alloc_grid(n, dtype) for _ in range(3))
            K.init(K.INIT_SEEDED, seed, dtype, b0, out)
            frame = torch.ones((n, n), dtype=torch.bool, device=gpu)
            frame[radius:n - radius, radius:n - radius] = False
            out.view(torch.int16 if dtype == BF16 else torch.int32 if dtype == F32 else torch.int64)[frame] = -7
            before = to_host(out).copy()
            K.run(pattern, radius, dtype, iters, b0, b1, out)
            got = to_host(out)
            fr = frame.cpu().numpy()
            assert np.array_equal(bits(got)[fr], bits(before)[fr]), (pattern, radius, v)
            if dtype == K.F64:
                ref, _ = oracle.run(pattern, radius, n, 32, iters, 1, seed)
            else:
                ref, _ = oracle.run_lp(dtype, pattern, radius, n, iters, 1, seed)
            assert np.array_equal(bits(got)[~fr], bits(ref.view(got.dtype))[~fr]), (pattern, radius, v)
    finally:
        K.check(L.prk_set_sweep_variant(-1))


# Edge waves park the vectors that straddle a frame boundary (the one holding
# column R and the one holding columns n-R-1, n-R) in LDS and write their
# interior elements after the row loop when the tile has at most 64 output
# rows and the launch has LDS room for the stash; taller tiles and the
# register-window sweep store them at once.  Cases: one boundary per edge
# wave, both boundaries in one wave (n <= 64 V), a second edge strip that
# holds frame columns only (n = 130 in FP64), register-ring and LDS-ring
# stars, grids; tilings of 63, 64 and 65 rows, occupancy caps that leave the
# stash in the padding or add it, padded pitch.  The frame keeps a sentinel
# and the interior matches the oracle bit for bit.
STASH = [(K.F64, 0, 3, 616), (K.F64, 1, 5, 130), (K.F64, 0, 7, 90), (F32, 1, 3, 616), (F32, 0, 6, 40),
         (BF16, 0, 5, 616), (BF16, 1, 3, 104)]


@pytest.mark.parametrize("dtype,pattern,radius,n", STASH)
def test_edge_stash_paths(gpu, oracle, dtype, pattern, radius, n):
    L = K.lib()
    iters, seed = 2, 13
    V = 16 // np.dtype(np.float64 if dtype == K.F64 else np.float32 if dtype == F32 else np.uint16).itemsize
    if dtype == K.F64:
        ref, _ = oracle.run(pattern, radius, n, 32, iters, 1, seed)
    else:
        ref, _ = oracle.run_lp(dtype, pattern, radius, n, iters, 1, seed)
    tilings = [(4, 64, 2, None), (5, 64, 0, None), (4, 63, 4, None), (4, 65, 3, None), (1, 100, 2, None),
               (5, 8, 3, n + 4 * V), (2, 1, 0, None)]
    if dtype != K.F64 and radius <= 3:
        tilings.append((7, 64, 2, None))
    try:
        for v, rows, wgs, pitch in tilings:
            print(f"stash dtype={dtype} {NAME[pattern]}{radius} n={n} variant={v} rows={rows} wgs={wgs} "
                  f"pitch={pitch}", flush=True)
            K.check(L.prk_set_sweep_variant(v)); K.check(L.prk_set_rows_per_wave(rows))
            K.check(L.prk_set_sweep_order(1, wgs))
            b0, b1, out = (K.alloc_grid(n, dtype, pitch=pitch) for _ in range(3))
            K.init(K.INIT_SEEDED, seed, dtype, b0, out)
            frame = torch.ones((n, n), dtype=torch.bool, device=gpu)
            frame[radius:n - radius, radius:n - radius] = False
            out.view(torch.int16 if dtype == BF16 else torch.int32 if dtype == F32 else torch.int64)[frame] = -7
            before = to_host(out).copy()
            K.run(pattern, radius, dtype, iters, b0, b1, out)
            got = to_host(out)
            fr = frame.cpu().numpy()
            assert np.array_equal(bits(got)[fr], bits(before)[fr]), (v, rows, wgs, pitch)
            assert np.array_equal(bits(got)[~fr], bits(ref.view(got.dtype))[~fr]), (v, rows, wgs, pitch)
    finally:
        K.check(L.prk_set_sweep_variant(-1)); K.check(L.prk_set_rows_per_wave(0))
        K.check(L.prk_set_sweep_order(-1, -1))

"""Batched solves against the oracle, alone vs in a mix, and across graph
replays.  Every test runs in the round form (one launch per round of the
schedule).  The persistent tile-queue form hung on MI355X (a wave waiting on
another wave's completion counter never saw it advance); it is disabled in
the library and its cases stay here, skipped."""
import numpy as np
import pytest

from kernels_amd.batch import Batch, one_by_one
from tests._oracle import BF16, F32, F64

pytestmark = pytest.mark.gpu

PERSISTENT_SKIP = pytest.mark.skip(reason="persistent tile-queue form hung on MI355X; disabled (PRK_BATCH_PERSISTENT)")
FORMS = [pytest.param(0, id="rounds"), pytest.param(1, id="persistent", marks=PERSISTENT_SKIP)]


def ref_solve(oracle, s):
    pattern, radius, dtype, n, iters, init, seed = s
    if dtype == F64:
        out, fin = oracle.run(pattern, radius, n, 32, iters, init, seed)
        norm = oracle.l1norm(n, radius, out)
    else:
        out, fin = oracle.run_lp(dtype, pattern, radius, n, iters, init, seed)
        norm = oracle.l1norm_lp(dtype, n, radius, out)
    return out, fin, norm


MIX = [
    # pattern, radius, dtype, n, iterations, init, seed
    (0, 2, F64, 130, 3, 1, 11),
    (1, 3, F32, 77, 5, 1, 12),
    (1, 4, BF16, 300, 2, 1, 13),
    (0, 1, F32, 1100, 1, 0, 0),
    (1, 2, F64, 9, 0, 1, 14),
    (0, 4, F64, 257, 4, 0, 0),
    (1, 1, BF16, 64, 6, 0, 0),
    (0, 3, F32, 515, 2, 1, 15),
    (1, 4, F32, 33, 3, 1, 16),
    (0, 2, BF16, 96, 1, 1, 17),
]


@pytest.mark.parametrize("flags", FORMS)
def test_batch_matches_oracle(gpu, oracle, flags):
    b = Batch(MIX, flags=flags)
    b.run()
    norms = b.wait()
    for i, s in enumerate(MIX):
        print(f"solve {i}: {s}", flush=True)
        out, fin, norm = ref_solve(oracle, s)
        got = b.output(i, 0)
        assert np.array_equal(got.view(np.uint8), out.view(np.uint8)), f"out differs for {s}"
        assert np.array_equal(b.output(i, 1).view(np.uint8), fin.view(np.uint8)), f"in differs for {s}"
        assert abs(norms[i] - norm) <= 1e-13 * abs(norm) + 1e-300, (s, norms[i], norm)
    b.close()


def test_batch_errors_against_f64(gpu, oracle):
    # prk_batch_errors over a mix of storage types: FP64 solves get zeros,
    # FP32 and BF16 solves the errors of their oracle results against the
    # FP64 oracle (the max exactly, the sums to 1e-12); two calls agree bit
    # for bit; a reference batch that is not the same solves in FP64 is
    # refused.
    b = Batch(MIX)
    b.run()
    b.wait()
    ref = Batch([(s[0], s[1], F64) + tuple(s[3:]) for s in MIX])
    ref.run()
    ref.wait()
    errs = b.errors(ref)
    assert np.array_equal(errs.view(np.uint64), b.errors(ref).view(np.uint64))
    for i, s in enumerate(MIX):
        pattern, radius, dtype, n, iters, init, seed = s
        print(f"solve {i}: {s} errors {errs[i]}", flush=True)
        if dtype == F64:
            assert np.all(errs[i] == 0.0)
            continue
        x = ref_solve(oracle, s)[0]
        x = (x.astype(np.uint32) << 16).view(np.float32) if dtype == BF16 else x
        y, _ = oracle.run(pattern, radius, n, 32, iters, init, seed)
        inner = (slice(radius, n - radius),) * 2
        d = np.abs(x.astype(np.float64)[inner] - y[inner])
        normwise = np.linalg.norm(d) / np.linalg.norm(y[inner])
        assert errs[i, 0] == d.max(), (s, errs[i, 0], d.max())
        assert abs(errs[i, 1] - d.mean()) <= 1e-12 * d.mean(), (s, errs[i, 1], d.mean())
        assert abs(errs[i, 2] - normwise) <= 1e-12 * normwise, (s, errs[i, 2], normwise)
    with pytest.raises(RuntimeError):
        b.errors(b)  # not stored in FP64
    ref.close()
    b.close()


@pytest.mark.parametrize("flags", FORMS)
def test_alone_equals_in_mix_and_replays(gpu, flags):
    mix = MIX * 3
    b = Batch(mix, rows_per_tile=12, flags=flags)
    runs = []
    for _ in range(3):          # graph replays, with eager work in between
        b.run()
        runs.append(b.wait())
        _ = one_by_one(MIX[:2])
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])
    outs = [b.output(i) for i in range(len(mix))]
    for i, s in enumerate(MIX):
        alone = Batch([s], flags=flags)
        alone.run()
        na = alone.wait()
        assert na[0] == runs[0][i] == runs[0][i + len(MIX)]
        assert np.array_equal(alone.output(0).view(np.uint8), outs[i].view(np.uint8))
        alone.close()
    b.close()


def test_one_by_one_baseline(gpu, oracle):
    specs = [s for s in MIX if s[2] == F64]
    norms = one_by_one(specs)
    for s, nm in zip(specs, norms):
        ref = ref_solve(oracle, s)[2]
        assert abs(nm - ref) <= 1e-13 * abs(ref)
    # forced FP64 storage of the reduced-precision solves
    n64 = one_by_one(MIX, dtype_override=F64)
    for s, nm in zip(MIX, n64):
        s64 = (s[0], s[1], F64) + s[3:]
        ref = ref_solve(oracle, s64)[2]
        assert abs(nm - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("flags", FORMS)
def test_batch_many_small(gpu, oracle, flags):
    rng = np.random.default_rng(5)
    specs = []
    for i in range(300):
        n = int(np.exp(rng.uniform(np.log(20), np.log(400))))
        r = int(rng.integers(1, 5))
        n = max(n, 2 * r + 1)
        specs.append((int(rng.integers(0, 2)), r, int(rng.integers(0, 3)), n, int(rng.integers(0, 8)), 1, i))
    b = Batch(specs, flags=flags)
    b.run()
    norms = b.wait()
    for i in range(0, 300, 7):
        _, _, ref = ref_solve(oracle, specs[i])
        assert abs(norms[i] - ref) <= 1e-13 * abs(ref), specs[i]
        out, _, _ = ref_solve(oracle, specs[i])
        assert np.array_equal(b.output(i).view(np.uint8), out.view(np.uint8)), specs[i]
    info = b.info()
    assert info["launches"] >= 1 and info["tiles"] > 300
    b.close()


def test_batch_large_radii(gpu, oracle):
    # radii 5..8 (the reference's R=5 and the generator's 6..8) in every
    # storage type, batched together with small ones, against the oracle
    specs = [
        (0, 5, F64, 40, 3, 1, 61), (1, 5, F32, 23, 2, 1, 62), (1, 6, BF16, 70, 2, 1, 63),
        (0, 7, F64, 100, 1, 1, 64), (1, 8, F64, 17, 2, 1, 65), (0, 8, F32, 300, 2, 0, 0),
        (1, 7, F32, 131, 3, 1, 66), (0, 6, BF16, 45, 4, 1, 67), (1, 1, F64, 50, 2, 1, 68),
    ]
    b = Batch(specs)
    b.run()
    norms = b.wait()
    for i, s in enumerate(specs):
        print(f"solve {i}: {s}", flush=True)
        out, fin, norm = ref_solve(oracle, s)
        assert np.array_equal(b.output(i, 0).view(np.uint8), out.view(np.uint8)), s
        assert np.array_equal(b.output(i, 1).view(np.uint8), fin.view(np.uint8)), s
        assert abs(norms[i] - norm) <= 1e-13 * abs(norm) + 1e-300, (s, norms[i], norm)
    b.close()


def test_batch_autotune_keeps_results(gpu):
    from kernels_amd.batch import autotune, tuning_clear, tuning_get
    specs = [((i % 2), 1 + i % 4, 1, 60 + 97 * i, 3 + i % 5, 1, 300 + i) for i in range(24)]
    tuning_clear()
    base = Batch(specs)
    base.run(); ref = base.wait(); outs = [base.output(i) for i in (0, 7, 23)]
    base.close()
    try:
        keys = autotune(specs, sample=4, reps=1)
        assert keys >= 1
        heights = {tuning_get(p, r, d, n) for p, r, d, n, *_ in specs}
        assert heights <= {16, 32, 64, 128} | {n for _, _, _, n, *_ in specs}
        tuned = Batch(specs)
        tuned.run(); got = tuned.wait()
        assert np.array_equal(got, ref)
        for i, o in zip((0, 7, 23), outs):
            assert np.array_equal(tuned.output(i).view(np.uint8), o.view(np.uint8))
        tuned.close()
    finally:
        tuning_clear()


def fp32_normwise_bound(iters: int) -> float:
    """Normwise error of an FP32-storage solve against FP64 after iters+1
    sweeps grows with the sweeps (the FP32 rounding of in_0 + k enters each
    one): over the config-4 sample it measured 1.1e-7 at K=4, 6.2e-7 to
    7.2e-7 at K=28..33 and 1.16e-6 at K=50 (oracle arithmetic, which the
    kernels match bit for bit).  The bound is 1.6x to 1.9x those."""
    return 3.5e-8 * (iters + 2)


def test_config4_mix(gpu, oracle):
    # The 4096-grid mix of the batched benchmark (seed 7: n log-uniform in
    # [64, 2048], star/grid, R 1..4, K 1..50, FP32 storage, seeded U(-1,1)):
    # every batched norm equals the one-grid-per-launch FP32 norm bit for
    # bit; a spread sample (every 400th solve, the largest grid and the
    # largest grid R=4 at K=50) equals the oracle bit for bit on three row
    # bands (the whole grid up to n=600), in FP32 and in FP64; every FP32
    # result is within the stated bound of its FP64 result by the device's
    # per-solve errors (prk_batch_errors), whose figures equal the host's on
    # the sample.
    import bench
    from concurrent.futures import ThreadPoolExecutor

    specs = bench.config4_solves()
    big = max(range(len(specs)), key=lambda i: specs[i][3])
    big_g4 = max((i for i, s in enumerate(specs) if s[0] == 1 and s[1] == 4), key=lambda i: specs[i][3] * specs[i][4])
    sample = list(range(0, len(specs), 400)) + [big, big_g4]
    assert specs[big][3] >= 2000 and specs[big_g4][1] == 4 and specs[big_g4][4] == 50
    print("config 4: batch of 4096", flush=True)
    b = Batch(specs)
    b.run()
    norms = b.wait()
    outs = {i: (b.output(i, 0), b.output(i, 1)) for i in sample}
    print("config 4: the same 4096 solves in FP64", flush=True)
    b64 = Batch([(s[0], s[1], F64) + tuple(s[3:]) for s in specs])
    b64.run()
    b64.wait()
    outs64 = [(b64.output(i, 0), b64.output(i, 1)) for i in sample]
    errs = b.errors(b64)
    b64.close()
    b.close()
    bounds = np.array([fp32_normwise_bound(s[4]) for s in specs])
    print(f"config 4: FP32 normwise error against FP64 over 4096 solves: max {errs[:, 2].max():.3e}, "
          f"median {np.median(errs[:, 2]):.3e}, max of error / bound {np.max(errs[:, 2] / bounds):.3f}", flush=True)
    assert errs[:, 2].max() > 0 and np.all(errs[:, 2] <= bounds), int(np.sum(errs[:, 2] > bounds))
    print("config 4: one grid per launch", flush=True)
    n32 = one_by_one(specs)
    assert np.array_equal(norms, n32), int(np.sum(norms != n32))

    def spans(s):
        n, r = s[3], s[1]
        if n <= 600:
            return [(0, n)]
        return [(0, r + 6), (n // 2 - 4, n // 2 + 4), (n - r - 6, n)]

    def check(j):
        i = sample[j]
        p, r, d, n, k, init, seed = specs[i]
        msgs = []
        for a, bb in spans(specs[i]):
            lp = oracle.band_lp(F32, p, r, n, k, init, seed, a, bb)
            if not np.array_equal(outs[i][0][a:bb].view(np.uint32), lp.view(np.uint32)):
                msgs.append(f"FP32 band {a}:{bb} of {specs[i]}")
            f64 = oracle.band(p, r, n, k, init, seed, a, bb)
            if not np.array_equal(outs64[j][0][a:bb].view(np.uint64), f64.view(np.uint64)):
                msgs.append(f"FP64 band {a}:{bb} of {specs[i]}")
        return msgs

    with ThreadPoolExecutor(8) as ex:
        bad = [m for ms in ex.map(check, range(len(sample))) for m in ms]
    assert not bad, bad
    for j, i in enumerate(sample):
        r, n, k = specs[i][1], specs[i][3], specs[i][4]
        inner = (slice(r, n - r),) * 2
        y = outs64[j][0][inner]
        dd = np.abs(outs[i][0].astype(np.float64)[inner] - y)
        normwise = float(np.linalg.norm(dd) / np.linalg.norm(y))
        print(f"config 4 solve {i} {specs[i]}: FP32 normwise {normwise:.3e} max abs {dd.max():.3e}", flush=True)
        assert normwise <= fp32_normwise_bound(k), (specs[i], normwise)
        # the device's figures for the same solve: the max exactly, the sums
        # (summed in another order) to 1e-12
        assert errs[i, 0] == dd.max(), (specs[i], errs[i, 0], dd.max())
        assert abs(errs[i, 1] - dd.mean()) <= 1e-12 * dd.mean(), (specs[i], errs[i, 1], dd.mean())
        assert abs(errs[i, 2] - normwise) <= 1e-12 * normwise, (specs[i], errs[i, 2], normwise)


def test_config4_mix_bf16(gpu, oracle):
    # The same 4096 solves stored in BF16: every batched norm equals the
    # one-grid-per-launch BF16 norm bit for bit, two solves (the largest
    # grid and the largest grid R=4 at K=50) equal the oracle's BF16
    # arithmetic on row bands, and every result is within 0.11 normwise of
    # the same solve in FP64 (BF16 stores in_0 + k to within 2^-3 near 51:
    # 0.073-0.075 at n=16384, K=50; the bound of the precision sweep).
    import bench

    specs = [(s[0], s[1], BF16) + tuple(s[3:]) for s in bench.config4_solves()]
    big = max(range(len(specs)), key=lambda i: specs[i][3])
    big_g4 = max((i for i, s in enumerate(specs) if s[0] == 1 and s[1] == 4), key=lambda i: specs[i][3] * specs[i][4])
    print("config 4 in BF16: batch of 4096", flush=True)
    b = Batch(specs)
    b.run()
    norms = b.wait()
    outs = {i: b.output(i, 0) for i in (big, big_g4)}
    b64 = Batch([(s[0], s[1], F64) + tuple(s[3:]) for s in specs])
    b64.run()
    b64.wait()
    errs = b.errors(b64)
    b64.close()
    b.close()
    print(f"config 4 in BF16: normwise error against FP64: max {errs[:, 2].max():.3e}, "
          f"median {np.median(errs[:, 2]):.3e}", flush=True)
    assert errs[:, 2].max() > 0 and np.all(errs[:, 2] <= 0.11), int(np.sum(errs[:, 2] > 0.11))
    print("config 4 in BF16: one grid per launch", flush=True)
    n16 = one_by_one(specs)
    assert np.array_equal(norms, n16), int(np.sum(norms != n16))
    for i in (big, big_g4):
        p, r, d, n, k, init, seed = specs[i]
        for a, bb in ((0, r + 6), (n // 2 - 4, n // 2 + 4), (n - r - 6, n)):
            ref = oracle.band_lp(BF16, p, r, n, k, init, seed, a, bb)
            assert np.array_equal(outs[i][a:bb], ref), (specs[i], a, bb)


@pytest.mark.parametrize("rows", [0, 8, 17])
def test_random_mix_bitwise(gpu, oracle, rows):
    # A seeded random mix of 48 solves (pattern, radius 1..8, storage type,
    # n 9..300, 0..3 iterations, either init), one batch per tile height:
    # every out, final in and norm equals the oracle's (the norms to 1e-13).
    rng = np.random.default_rng(777 + rows)
    mix = []
    for _ in range(48):
        r = int(rng.integers(1, 9))
        mix.append((int(rng.integers(0, 2)), r, int(rng.integers(0, 3)), int(rng.integers(max(9, 2 * r + 1), 301)),
                    int(rng.integers(0, 4)), int(rng.integers(0, 2)), int(rng.integers(0, 1000))))
    b = Batch(mix, rows_per_tile=rows)
    b.run()
    norms = b.wait()
    for i, s in enumerate(mix):
        out, fin, norm = ref_solve(oracle, s)
        assert np.array_equal(b.output(i, 0).view(np.uint8), out.view(np.uint8)), s
        assert np.array_equal(b.output(i, 1).view(np.uint8), fin.view(np.uint8)), s
        assert abs(norms[i] - norm) <= 1e-13 * abs(norm) + 1e-300, (s, norms[i], norm)
    b.close()


"""The continuous-batching server: results against the oracle, the same
alone and in a mix (bitwise), admission under a tight HBM budget, queue-full
rejection, and requests arriving while windows run."""
import threading
import time

import numpy as np
import pytest

from kernels_amd.server import Server
from tests._oracle import BF16, F32, F64
from tests.test_gpu_batch import ref_solve

pytestmark = pytest.mark.gpu

MIX = [
    (0, 2, F64, 130, 3, 1, 21),
    (1, 3, F32, 77, 9, 1, 22),
    (1, 4, BF16, 300, 2, 1, 23),
    (0, 1, F32, 700, 1, 0, 0),
    (1, 2, F64, 9, 0, 1, 24),
    (0, 4, F64, 257, 12, 0, 0),
    (1, 1, BF16, 64, 6, 0, 0),
    (0, 3, F32, 515, 2, 1, 25),
    (1, 4, F32, 33, 17, 1, 26),
    (0, 2, BF16, 96, 1, 1, 27),
]


def run_all(srv, specs, keep=True):
    ids = [srv.submit(s, keep=keep) for s in specs]
    srv.drain(120)
    res = {r["id"]: r for r in srv.poll()}
    assert set(res) == set(ids)
    return ids, res


def test_server_matches_oracle(gpu, oracle):
    srv = Server(budget_bytes=1 << 30)
    ids, res = run_all(srv, MIX)
    for rid, s in zip(ids, MIX):
        print(f"request {rid}: {s}", flush=True)
        out, fin, norm = ref_solve(oracle, s)
        assert np.array_equal(srv.fetch(rid).view(np.uint8), out.view(np.uint8)), s
        assert np.array_equal(srv.fetch(rid, 1).view(np.uint8), fin.view(np.uint8)), s
        assert abs(res[rid]["norm"] - norm) <= 1e-13 * abs(norm), s
        r = res[rid]
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
        srv.release(rid)
    st = srv.stats()
    assert st["served"] == len(MIX) and st["pool_used"] == 0 and st["windows"] >= 1
    srv.close()


def test_alone_equals_in_mix(gpu):
    mix = MIX * 2
    srv = Server(budget_bytes=1 << 30)
    ids, res = run_all(srv, mix)
    mixed = {i: (srv.fetch(rid), res[rid]["norm"]) for i, rid in enumerate(ids)}
    srv.close()
    for i, s in enumerate(MIX):
        alone = Server(budget_bytes=1 << 30)
        aid, ares = run_all(alone, [s])
        out = alone.fetch(aid[0])
        for j in (i, i + len(MIX)):
            assert np.array_equal(out.view(np.uint8), mixed[j][0].view(np.uint8)), s
            assert ares[aid[0]]["norm"] == mixed[j][1], s
        alone.close()


def test_tight_budget_and_rejections(gpu, oracle):
    # room for about two 300x300 FP64 requests at a time
    srv = Server(budget_bytes=5 << 20, max_pending=64)
    specs = [(i % 2, 1 + i % 4, F64, 300, 3 + i % 5, 1, 100 + i) for i in range(12)]
    ids, res = run_all(srv, specs, keep=False)
    for rid, s in zip(ids, specs):
        _, _, norm = ref_solve(oracle, s)
        assert abs(res[rid]["norm"] - norm) <= 1e-13 * abs(norm)
    rid, rc = srv.try_submit((0, 2, F64, 4000, 1, 1, 0))   # can never fit 5 MiB
    assert rid is None and rc == -4
    srv.close()
    small = Server(budget_bytes=1 << 20, max_pending=2)
    codes = [small.try_submit((0, 1, F32, 200, 200, 1, k))[1] for k in range(8)]
    assert -5 in codes and 0 in codes
    small.drain(120)
    assert small.stats()["rejected"] == codes.count(-5)
    small.close()


def test_arrivals_while_running(gpu, oracle):
    srv = Server(budget_bytes=1 << 30)
    specs = [(i % 2, 1 + i % 4, (F32, F64, BF16)[i % 3], 64 + 37 * i, 5 + (7 * i) % 40, 1, 200 + i)
             for i in range(40)]
    ids = []

    def client():
        for s in specs:
            ids.append(srv.submit(s))
            time.sleep(0.002)

    th = threading.Thread(target=client)
    th.start()
    th.join()
    srv.drain(120)
    res = {r["id"]: r for r in srv.poll()}
    assert len(res) == len(specs)
    for rid, s in list(zip(ids, specs))[::5]:
        _, _, norm = ref_solve(oracle, s)
        assert abs(res[rid]["norm"] - norm) <= 1e-13 * abs(norm)
    assert srv.stats()["windows"] >= 2
    srv.close()


def test_graph_exec_updates_keep_results(gpu, oracle):
    # Windows with the same classes and rounds take their new kernel
    # arguments in the slot's existing graph exec instead of a new one; the
    # finished norms of each window come back in one copy.  Every result is
    # still the oracle's.
    srv = Server(budget_bytes=1 << 30)
    specs = [(0, 2, F32, 200 + 16 * i, 60, 1, 300 + i) for i in range(6)] + [(1, 3, F64, 150, 45, 1, 310)]
    ids, res = run_all(srv, specs)
    for rid, s in zip(ids, specs):
        print(f"request {rid}: {s}", flush=True)
        out, _, norm = ref_solve(oracle, s)
        assert np.array_equal(srv.fetch(rid).view(np.uint8), out.view(np.uint8)), s
        assert abs(res[rid]["norm"] - norm) <= 1e-13 * abs(norm), s
    st = srv.stats()
    print(st, flush=True)
    assert st["windows"] >= 6 and st["graph_updates"] >= 1, st
    srv.close()


def test_shadow_solves_report_error_vs_f64(gpu, oracle):
    # With shadow_every=1 every FP32/BF16 request also runs in FP64 inside
    # the server; its result carries the error of out against that solve.
    # Both solves are bitwise the oracle's (FP64 and the FP32/BF16
    # restatement), so the error must be the oracle outputs' error: the max
    # exactly, the sums up to their order.
    from tests._oracle import bf16_bits_to_f64
    specs = [(0, 2, F32, 130, 5, 1, 41), (1, 3, BF16, 96, 4, 1, 42), (1, 1, F32, 300, 7, 1, 43),
             (0, 4, BF16, 257, 3, 1, 44), (1, 2, F64, 120, 3, 1, 45)]
    srv = Server(budget_bytes=1 << 30, shadow_every=1)
    ids, res = run_all(srv, specs)
    assert srv.stats()["shadow_solves"] == 4
    for rid, s in zip(ids, specs):
        print(f"request {rid}: {s}", flush=True)
        r = res[rid]
        pattern, radius, dtype, n, iters, init, seed = s
        if dtype == F64:
            assert "error_vs_f64" not in r
            continue
        x, _ = oracle.run_lp(dtype, pattern, radius, n, iters, init, seed)
        x = bf16_bits_to_f64(x) if dtype == BF16 else x.astype(np.float64)
        y, _ = oracle.run(pattern, radius, n, 32, iters, init, seed)
        inner = (slice(radius, n - radius),) * 2
        d = np.abs(x[inner] - y[inner])
        got = r["error_vs_f64"]
        assert got["max_abs"] == d.max(), (s, got)
        assert abs(got["mean_abs"] - d.mean()) <= 1e-12 * d.mean(), (s, got)
        nw = float(np.sqrt((d * d).sum() / (y[inner] ** 2).sum()))
        assert abs(got["normwise"] - nw) <= 1e-12 * nw, (s, got, nw)
        assert r["t_admit"] <= r["t_done"]
    srv.close()

"""Test infrastructure: bench.py's open-loop serving client (Poisson
arrivals of the served request stream, submit, drain, poll) driving the
Python Server wrapper, run on the CPU against the server built with the HIP
stand-in (PRK_STENCIL_LIB names that library).  Checks that every accepted
request is answered once, with the stand-in's analytic norm 2(K+1).

usage: serve_client.py <rate per s> <seconds>"""
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402
from kernels_amd.server import Server  # noqa: E402


def main() -> int:
    rate, seconds = float(sys.argv[1]), float(sys.argv[2])
    arrivals, specs = bench.config5_requests(rate, seconds)
    srv = Server(budget_bytes=200 << 30, max_pending=256)
    t0 = Server.clock() + 0.05
    ids, rejected = {}, 0
    for off, spec in zip(arrivals, specs):
        while Server.clock() < t0 + off:
            time.sleep(0.0005)
        rid, rc = srv.try_submit(spec)
        if rid is None:
            rejected += 1
        else:
            ids[rid] = spec
    srv.drain(600)
    res = srv.poll(len(ids) + 16)
    st = srv.stats()
    srv.close()
    seen = [r["id"] for r in res]
    wrong = [r for r in res if r["norm"] != 2.0 * (ids[r["id"]][4] + 1)]
    print(f"offered {len(specs)} accepted {len(ids)} rejected {rejected} answered {len(res)} "
          f"wrong {len(wrong)} windows {st['windows']}", flush=True)
    ok = sorted(seen) == sorted(ids) and not wrong and st["pool_used"] == 0
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())

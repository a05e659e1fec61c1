"""The prk-stencil driver on the GPU: the reference's command line and report
(Cxx11/stencil-hip.cc:97-149,246-264), the fused and unfused structures, the
storage types, and the unsupported-radius path."""
import os
import re
import subprocess
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.gpu
EXE = REPO / "kernels_amd" / "prk-stencil"


def run(args, **env):
    if not EXE.exists():
        pytest.fail("prk-stencil not built (python -m kernels_amd.build)")
    print("prk-stencil", *args, env, flush=True)
    r = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=120,
                       env={**os.environ, **env})
    return r.returncode, r.stdout


@pytest.mark.parametrize("args,env", [
    ((10, 1000), {}),                                        # the reference's CPU case, default tile/star/R=2
    ((10, 1000, 32, "star", 2), {"PRK_MODE": "unfused"}),    # stencil + add per iteration, device sync
    ((10, 1000, 32, "grid", 3), {}),
    ((10, 1000, 0, "star", 5), {}),                          # tile <= 0 -> n
    ((10, 1000, 32, "star", 2), {"PRK_DTYPE": "f32"}),
    ((4, 777, 32, "grid", 1), {"PRK_DTYPE": "f32"}),         # ragged n
    ((5, 300, 32, "grid", 8), {}),                           # radii 6..8 extend the reference's table
    ((5, 300, 32, "star", 7), {"PRK_MODE": "unfused"}),
])
def test_cli_validates(gpu, args, env):
    rc, out = run(args, **env)
    assert rc == 0, out
    assert "Solution validates" in out
    m = re.search(r"Rate \(MFlops/s\): ([0-9.eE+-]+) Avg time \(s\): ([0-9.eE+-]+)", out)
    assert m and float(m.group(1)) > 0 and float(m.group(2)) > 0, out
    assert f"Number of iterations = {args[0]}" in out and f"Grid size            = {args[1]}" in out


def test_cli_seeded_norm_matches_oracle(gpu, oracle):
    n, k = 300, 6
    rc, out = run((k, n, 32, "grid", 2), PRK_INIT="seeded", PRK_SEED="42")
    assert rc == 0, out
    norm = float(re.search(r"L1 norm = ([0-9.eE+-]+)", out).group(1))
    ref_out, _ = oracle.run(1, 2, n, 32, k, 1, 42)
    ref = oracle.l1norm(n, 2, ref_out)
    assert abs(norm - ref) <= 1e-5 * abs(ref), (norm, ref)


def test_cli_unsupported_radius_fails(gpu):
    # beyond the table (the reference's stops at 5, this one at 8) the
    # reference selects its no-op operator and then fails validation; the
    # library refuses the operator and the driver exits with 1
    rc, out = run((5, 100, 32, "star", 9))
    assert rc == 1 and "ERROR" in out, out


TEXE = REPO / "kernels_amd" / "prk-transpose"


@pytest.mark.parametrize("args", [(10, 1000), (5, 1001, 0), (3, 64, 1), (2, 4097)])
def test_transpose_cli_validates(gpu, args):
    if not TEXE.exists():
        pytest.fail("prk-transpose not built (python -m kernels_amd.build)")
    print("prk-transpose", *args, flush=True)
    r = subprocess.run([str(TEXE), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "Solution validates" in r.stdout
    m = re.search(r"Rate \(MB/s\): ([0-9.eE+-]+) Avg time \(s\): ([0-9.eE+-]+)", r.stdout)
    assert m and float(m.group(1)) > 0, r.stdout


# The reference CI's stencil invocations, literally.  The Cxx11 drivers
# (ci/build-run-prk.sh:334,340-345 stencil-vector; :372,375-380
# stencil-openmp) run "10 1000" and "10 200 20 {star,grid} {1..5}".  The C11
# section (:151,154-158) runs "10 200 {star,grid} {1..5}", which passes the
# pattern where the tile goes: the driver parses tile = atoi("star") = 0 ->
# n, pattern "1".."5" -> star, radius 2 -- exactly as the reference's own
# argument handling does.  Every run must validate.
CI_CASES = ([("10", "1000")] + [("10", "200", "20", s, r) for s in ("star", "grid") for r in "12345"]
            + [("10", "200", s, r) for s in ("star", "grid") for r in "12345"])


@pytest.mark.parametrize("args", CI_CASES, ids=["-".join(a) for a in CI_CASES])
def test_reference_ci_invocations(gpu, args):
    rc, out = run(args)
    assert rc == 0 and "Solution validates" in out, out
    if len(args) == 5:
        kind = "star" if args[3] == "star" else "grid"
        assert "Tile size            = 20" in out and f"Type of stencil      = {kind}" in out
        assert f"Radius of stencil    = {args[4]}" in out
    if len(args) == 4:
        assert "Tile size            = 200" in out and "Type of stencil      = star" in out
        assert "Radius of stencil    = 2" in out

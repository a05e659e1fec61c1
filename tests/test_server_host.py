"""The solve server's host logic (kernels_amd/csrc/server.hip) on the CPU:
built host-only against a stand-in for the HIP runtime (tests/host/mock_hip.cc)
with AddressSanitizer and with ThreadSanitizer, and driven from a client
thread at serving rates (tests/host/server_driver.cc).  The stand-in replays
each captured window and checks every work entry against its grid table, so
admission, pool placement, window building and retirement are checked
without a GPU."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
SOURCES = [ROOT / "kernels_amd/csrc/server.hip", ROOT / "tests/host/mock_hip.cc", ROOT / "tests/host/server_driver.cc"]

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="needs hipcc")


def build(tmp: Path, sanitizer: str) -> Path:
    objs = []
    for src in SOURCES:
        obj = tmp / (src.name + ".o")
        subprocess.run([HIPCC, "--offload-host-only", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
                        f"-I{ROOT / 'include'}", f"-I{ROOT / 'kernels_amd/csrc'}",
                        "-Xarch_host", f"-fsanitize={sanitizer}", "-x", "hip", "-c", str(src), "-o", str(obj)],
                       check=True, capture_output=True, text=True)
        objs.append(str(obj))
    exe = tmp / f"server_driver_{sanitizer}"
    subprocess.run([CLANG, f"-fsanitize={sanitizer}", "-o", str(exe), *objs, "-lpthread"], check=True,
                   capture_output=True, text=True)
    return exe


def build_shared(tmp: Path) -> Path:
    """The server and the HIP stand-in as one shared library (AddressSanitizer)."""
    objs = []
    for src in SOURCES[:2]:
        obj = tmp / (src.name + ".pic.o")
        subprocess.run([HIPCC, "--offload-host-only", "-O1", "-g", "-std=c++17", "-fPIC", "-fno-omit-frame-pointer",
                        f"-I{ROOT / 'include'}", f"-I{ROOT / 'kernels_amd/csrc'}",
                        "-Xarch_host", "-fsanitize=address", "-x", "hip", "-c", str(src), "-o", str(obj)],
                       check=True, capture_output=True, text=True)
        objs.append(str(obj))
    so = tmp / "libprkserver_standin.so"
    subprocess.run([CLANG, "-shared", "-fsanitize=address", "-o", str(so), *objs, "-lpthread"], check=True,
                   capture_output=True, text=True)
    return so


def test_serving_client_loop(tmp_path):
    # bench.py's serving client (the served request stream at 500 req/s, a
    # 200 GiB budget) through the Python wrapper, against the stand-in
    so = build_shared(tmp_path)
    rt = sorted(Path(CLANG).resolve().parents[1].glob("lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt:
        pytest.skip("no AddressSanitizer runtime")
    env = dict(os.environ, PRK_STENCIL_LIB=str(so), LD_PRELOAD=str(rt[-1]),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run([sys.executable, str(ROOT / "tests/host/serve_client.py"), "500", "3"], capture_output=True,
                       text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr


# gbs > 0: the stand-in's device takes time (bytes / gbs) and aborts when the
# host frees or reads memory that work still in flight uses; at these rates
# the served request stream saturates it, so the queue fills, requests wait
# for pool memory and the window tables grow while windows run.
@pytest.mark.parametrize("sanitizer,args,gbs", [("address", ["500", "2", "200", "0"], 0),
                                                ("address", ["100", "1", "200", "1"], 0),
                                                ("thread", ["200", "1", "200", "0"], 0),
                                                ("address", ["2000", "3", "200", "0"], 6000),
                                                ("thread", ["500", "2", "200", "0"], 6000),
                                                ("undefined", ["8000", "2", "200", "0"], 6000)])
def test_server_host_logic(tmp_path, sanitizer, args, gbs):
    exe = build(tmp_path, sanitizer)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    if gbs:
        env["PRK_MOCK_GBS"] = str(gbs)
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "failures 0" in r.stdout
    # windows whose topology repeats take the new arguments in their exec
    updates = int(r.stdout.split("graph updates ")[-1].split(",")[0])
    assert updates > 0
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert "runtime error" not in r.stderr

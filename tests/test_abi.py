"""The C ABI without a GPU: the library loads, exports every function
include/*.h declares, and rejects bad arguments with the documented codes
before touching the device."""
import ctypes
import re
import subprocess
import sys
from pathlib import Path

import pytest

import kernels_amd
from kernels_amd._lib import PrkError, check

REPO = Path(__file__).resolve().parent.parent
HEADERS = sorted((REPO / "include").glob("*.h"))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(prk_\w+)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"#define.*", "", text)
        names.update(DECL.findall(text))
    return sorted(names)


def test_headers_exist():
    assert HEADERS, "no include/*.h"
    assert len(declared_functions()) >= 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    h = kernels_amd.lib()
    assert hasattr(h, name), f"{name} declared but not exported"


def test_symbols_exported_by_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", str(kernels_amd.lib_path())],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (prk_\w+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing


def test_abi_version():
    assert kernels_amd.lib().prk_abi_version() == 1


def _rc(fn, *args):
    return fn(*args)


def test_bad_arguments_rejected_before_device():
    L = kernels_amd.lib()
    # radius outside the table (1..8): the reference's `nothing`
    assert L.prk_stencil(0, 9, 100, 32, 16, 16) == -2
    assert b"radius 9" in L.prk_last_error()
    assert L.prk_stencil(0, 0, 100, 32, 16, 16) == -2
    assert L.prk_stencil(2, 2, 100, 32, 16, 16) == -1       # unknown pattern
    assert L.prk_stencil(0, 2, 4, 32, 16, 16) == -1         # 2R+1 > n
    assert L.prk_stencil(0, 2, 100, 0, 16, 16) == -1        # tile < 1
    assert L.prk_stencil(0, 2, 100, 32, None, 16) == -1     # null grid
    assert L.prk_star2(0, 32, 16, 16) == -1                 # n < 1
    assert L.prk_sweep(0, 2, 7, 100, 100, 16, 32, 48, None) == -1    # dtype
    assert L.prk_sweep(0, 2, 0, 100, 99, 16, 32, 48, None) == -1     # pitch < n
    assert L.prk_sweep(0, 2, 0, 100, 100, 16, 16, 48, None) == -1    # in_next aliases in
    assert L.prk_sweep(0, 2, 0, 100, 100, 16, None, 48, None) == -1  # fused needs in_next
    # slabs: rows [row0, row0+rows) must lie inside the n rows
    assert L.prk_sweep_slab(0, 2, 0, 100, -1, 10, 100, 16, 32, 48, None) == -1
    assert L.prk_sweep_slab(0, 2, 0, 100, 0, 0, 100, 16, 32, 48, None) == -1
    assert L.prk_sweep_slab(0, 2, 0, 100, 95, 6, 100, 16, 32, 48, None) == -1
    assert b"not inside" in L.prk_last_error()
    assert L.prk_sweep_slab(0, 2, 0, 100, 2**31 - 5, 10, 100, 16, 32, 48, None) == -1  # no overflow
    assert L.prk_sweep_slab(0, 9, 0, 100, 0, 10, 100, 16, 32, 48, None) == -2
    assert L.prk_sweep_slab(0, 2, 0, 100, 0, 10, 100, 16, None, 48, None) == -1
    assert L.prk_run(0, 2, 0, 100, 100, 0, 16, 32, 48, None, None, None) == -1  # iterations < 1
    assert L.prk_stream_mix(0, 16, 32, 48, None) == -1       # len < 1
    assert L.prk_stream_mix(10, 16, 16, 48, None) == -1      # in_next aliases in
    assert L.prk_stream_mix(10, 16, 40, 48, None) == -1      # misaligned
    # tuning on caller grids: three separate grids, pitch >= n, reps >= 1
    assert L.prk_autotune_buffers(0, 2, 0, 100, 100, 16, 16, 48, 2, None, None) == -1
    assert L.prk_autotune_buffers(0, 2, 0, 100, 99, 16, 32, 48, 2, None, None) == -1
    assert L.prk_autotune_buffers(0, 2, 0, 100, 100, 16, 32, 48, 0, None, None) == -1
    assert L.prk_autotune_buffers(0, 2, 5, 100, 100, 16, 32, 48, 2, None, None) == -1
    assert L.prk_init(5, 0, 0, 10, 10, 16, 32, None) == -1
    assert L.prk_set_rows_per_wave(-1) == -1
    with pytest.raises(PrkError) as e:
        check(L.prk_stencil(1, 9, 100, 32, 16, 16))
    assert e.value.code == -2


def test_cli_argument_errors():
    exe = REPO / "kernels_amd" / "prk-stencil"
    if not exe.exists():
        pytest.skip("prk-stencil not built")
    cases = [
        ([], "Usage: <# iterations> <array dimension> [<tile_size> <star/grid> <radius>]"),
        (["0", "100"], "ERROR: iterations must be >= 1"),
        (["10", "0"], "ERROR: grid dimension must be positive"),
        (["10", "46341"], "ERROR: grid dimension too large - overflow risk"),
        (["10", "100", "32", "star", "0"], "ERROR: Stencil radius negative or too large"),
        (["10", "4", "32", "star", "2"], "ERROR: Stencil radius negative or too large"),
    ]
    for args, msg in cases:
        r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (args, r.stdout)
        assert msg in r.stdout, (args, r.stdout)


def test_transpose_cli_argument_errors():
    exe = REPO / "kernels_amd" / "prk-transpose"
    if not exe.exists():
        pytest.skip("prk-transpose not built")
    cases = [
        ([], "Usage: <# iterations> <matrix order> [variant (0/1/2)]"),
        (["0", "100"], "ERROR: iterations must be >= 1"),
        (["10", "0"], "ERROR: Matrix Order must be greater than 0"),
        (["10", "46341"], "ERROR: matrix dimension too large - overflow risk"),
        (["10", "100", "3"], "Please select a valid variant (0: naive 1: coalesced, 2: no bank conflicts)"),
    ]
    for args, msg in cases:
        r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, (args, r.stdout)
        assert msg in r.stdout, (args, r.stdout)


def test_tuning_table_roundtrip(tmp_path):
    L = kernels_amd.lib()
    T = kernels_amd.Tuning
    check(L.prk_tuning_clear())
    d = T()
    check(L.prk_tuning_get(0, 2, 0, 30000, d))       # built-in default
    assert 0 <= d.variant < 6 and d.rows_per_wave >= 1 and d.order == 1
    t = T(variant=4, rows_per_wave=12, wgs_per_cu=3, order=1, gbs=5000.0)
    check(L.prk_tuning_set(0, 2, 0, 30000, t))
    g = T()
    check(L.prk_tuning_get(0, 2, 0, 25000, g))       # nearest tuned size within 1.5x
    assert (g.variant, g.rows_per_wave, g.wgs_per_cu) == (4, 12, 3)
    check(L.prk_tuning_get(0, 2, 0, 10000, g))       # too far: default
    assert (g.variant, g.rows_per_wave) == (d.variant, min(d.rows_per_wave, 10000))
    path = str(tmp_path / "tune.txt").encode()
    check(L.prk_tuning_save(path))
    check(L.prk_tuning_clear())
    assert L.prk_tuning_load(path) == 1
    check(L.prk_tuning_get(0, 2, 0, 30000, g))
    assert (g.variant, g.rows_per_wave, g.wgs_per_cu, g.gbs) == (4, 12, 3, 5000.0)
    assert L.prk_tuning_set(0, 2, 0, 100, T(variant=9, rows_per_wave=8)) == -1
    check(L.prk_tuning_clear())


def test_batch_persistent_form_is_refused():
    import ctypes
    from kernels_amd._lib import Solve
    arr = (Solve * 1)(Solve(0, 2, 1, 64, 1, 1, 0))
    h = ctypes.c_void_p()
    L = kernels_amd.lib()
    assert L.prk_batch_create(arr, 1, 0, 1, ctypes.byref(h)) == -1
    assert b"disabled" in L.prk_last_error()
    assert L.prk_batch_create(arr, 1, 0, 4, ctypes.byref(h)) == -1
    bad = (Solve * 1)(Solve(0, 9, 1, 64, 1, 1, 0))
    assert L.prk_batch_create(bad, 1, 0, 0, ctypes.byref(h)) == -2   # radius 9: no stencil


def test_server_config_rejected_before_device():
    import ctypes
    from kernels_amd._lib import ServerConfig
    L = kernels_amd.lib()
    h = ctypes.c_void_p()
    for cfg in (ServerConfig(0, 1, 1, 1), ServerConfig(1 << 30, 0, 1, 1), ServerConfig(1 << 30, 1, 0, 1),
                ServerConfig(1 << 30, 1, 1, 0)):
        assert L.prk_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert L.prk_server_clock() > 0


def test_batch_tuning_table_defaults():
    L = kernels_amd.lib()
    import ctypes
    r = ctypes.c_int()
    check(L.prk_batch_tuning_clear())
    check(L.prk_batch_tuning_get(0, 2, 1, 1000, ctypes.byref(r)))
    assert r.value == 32
    check(L.prk_batch_tuning_get(1, 4, 0, 20, ctypes.byref(r)))
    assert r.value == 20                       # never taller than the grid
    assert L.prk_batch_tuning_get(0, 9, 1, 1000, ctypes.byref(r)) == -2   # radius 9: no stencil
    assert L.prk_batch_autotune(None, 0, 4, 1, None) == -1



def test_crash_backtrace_on_host_fault():
    # In a child process (host only, no device): with the handlers installed
    # a host fault prints the faulting thread's native frames, then the
    # signal reaches the previous handler (here the default action), so the
    # process still dies of SIGSEGV.  Uninstalling puts the old handler back.
    code = (
        "import ctypes, signal, sys\n"
        "import kernels_amd\n"
        "L = kernels_amd.lib()\n"
        "kernels_amd.check(L.prk_crash_backtrace(1))\n"
        "kernels_amd.check(L.prk_crash_backtrace(0))\n"
        "assert signal.getsignal(signal.SIGSEGV) == signal.SIG_DFL\n"
        "kernels_amd.check(L.prk_crash_backtrace(1))\n"
        "sys.stderr.write('about to fault\\n'); sys.stderr.flush()\n"
        "ctypes.string_at(0)\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd=str(Path(__file__).resolve().parents[1]))
    assert r.returncode == -11, (r.returncode, r.stderr[-2000:])
    assert "about to fault" in r.stderr
    assert "prk: fatal signal 11, native frames of the faulting thread:" in r.stderr
    frames = r.stderr.split("native frames of the faulting thread:")[1].strip().splitlines()
    assert len(frames) >= 3 and any("libprkstencil" in f for f in frames), frames[:8]

"""Build recipe for the gfx950 stencil library, its command-line driver and
the CPU oracle.

    python -m kernels_amd.build          # everything
    python -m kernels_amd.build --lib    # libprkstencil.so only

Outputs (in-tree, git-ignored, they travel with the source tree):
    kernels_amd/libprkstencil.so   the HIP kernels behind the C ABI of
                                   include/prk_stencil.h
    kernels_amd/prk-stencil        the driver with the reference's command line
    oracle/liboracle.so            the CPU oracle (test infrastructure)
    oracle/_ref/*                  reference binaries, only where the reference
                                   tree is present
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "kernels_amd"
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
ORACLE = ROOT / "oracle"
OBJ = PKG / "_build"
REFERENCE = Path("/root/reference")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep hipcc from emitting packed FP32 arithmetic.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

LIB_SOURCES = ["prk_api.hip", "sweep_f64.hip", "sweep_f32.hip", "sweep_bf16.hip",
               "batch.hip", "server.hip", "autotune.hip", "transpose.hip"]
# command-line drivers: executable -> source
CLIS = {"prk-stencil": "prk_cli.cpp", "prk-transpose": "prk_transpose_cli.cpp"}


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def _compile(src: Path, obj: Path) -> Path:
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in LIB_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, OBJ / (s.stem + ".o")), srcs))
    lib = PKG / "libprkstencil.so"
    tmp = PKG / "libprkstencil.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


def build_cli() -> None:
    for name, src in CLIS.items():
        if (CSRC / src).exists():
            _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", str(PKG / name), str(CSRC / src),
                  f"-L{PKG}", "-lprkstencil", f"-Wl,-rpath,$ORIGIN"])


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-C", str(ORACLE), "-j4", "all"])
    if (REFERENCE / "Cxx11" / "stencil.cc").exists():
        _run([make, "-C", str(ORACLE), "-j4", "ref"])


def build_all(jobs: int = 8) -> None:
    build_lib(jobs)
    build_cli()
    build_oracle()


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if a.lib:
        build_lib(a.j)
    else:
        build_all(a.j)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

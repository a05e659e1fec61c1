"""Writes the golden fixtures of the stencil path from the reference tree.

Run in a container where the reference tree is mounted (it is not needed to
run the tests):

    python tests/golden/make_golden.py [/root/reference]

Outputs (data only, committed beside this script):
  taps.json      for each generated serial operator star1..5 / grid1..5 of
                 Cxx11/stencil_seq.hpp, its terms in source order as
                 [row offset, column offset, weight literal], read from the
                 header's text.
  outputs.npz    reference outputs: for each operator, out += S(in) at n=17
                 on seeded inputs, computed by the reference's own generated
                 operator compiled from the reference tree
                 (oracle/_ref/libprkref.so, built by `make -C oracle ref`).
"""
from __future__ import annotations

import ctypes
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent

TERM = re.compile(r"\+in\[\(i([+-]\d+)?\)\*n\+\(j([+-]\d+)?\)\] \* (-?[0-9.e+-]+)")
FUNC = re.compile(r"void (star|grid)(\d)\(const int n, const int t, prk::vector<double> & in, "
                  r"prk::vector<double> & out\) \{(.*?)\n\}", re.S)

N_OUT = 17
SEED = 20240917


def parse_taps(header_text: str) -> dict[str, list[list]]:
    taps = {}
    for m in FUNC.finditer(header_text):
        name = f"{m.group(1)}{m.group(2)}"
        body = m.group(3)
        terms = [[int(a or 0), int(b or 0), float(w)] for a, b, w in TERM.findall(body)]
        taps[name] = terms
    return taps


def reference_outputs(ref_lib: Path) -> dict[str, np.ndarray]:
    lib = ctypes.CDLL(str(ref_lib))
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(SEED)
    arrays = {}
    x = rng.uniform(-1.0, 1.0, (N_OUT, N_OUT))
    o0 = rng.uniform(-1.0, 1.0, (N_OUT, N_OUT))
    arrays["in"] = x
    arrays["out0"] = o0
    for pat, pname in ((0, "star"), (1, "grid")):
        for r in range(1, 6):
            o = o0.copy()
            rc = lib.prkref_stencil(pat, r, N_OUT, 5, x.ctypes.data_as(dp), o.ctypes.data_as(dp))
            assert rc == 0
            arrays[f"{pname}{r}"] = o
    return arrays


def main(argv: list[str]) -> int:
    ref = Path(argv[1]) if len(argv) > 1 else Path("/root/reference")
    header = (ref / "Cxx11" / "stencil_seq.hpp").read_text()
    taps = parse_taps(header)
    assert len(taps) == 10, sorted(taps)
    (HERE / "taps.json").write_text(json.dumps(
        {"source": "Cxx11/stencil_seq.hpp", "order": "source order", "fields": ["di", "dj", "w"],
         "operators": taps}, indent=1) + "\n")
    ref_lib = REPO / "oracle" / "_ref" / "libprkref.so"
    if not ref_lib.exists():
        print(f"{ref_lib} missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    np.savez_compressed(HERE / "outputs.npz", **reference_outputs(ref_lib))
    print("wrote", HERE / "taps.json", HERE / "outputs.npz")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv))

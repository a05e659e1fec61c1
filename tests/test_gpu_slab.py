"""One GPU's share of a row-strip decomposition (prk_sweep_slab).

The grid is cut into horizontal slabs; each slab sweeps its own rows and
then takes its R halo rows of the new `in` from its neighbours' own rows, the
exchange of MPI1/Stencil/stencil.c:338-370 done here by device copies on one
GPU.  After every sweep each slab's rows must equal, bit for bit, the same
rows of the whole grid swept by prk_sweep (which the oracle pins elsewhere):
the slab kernel does each point's arithmetic exactly as the whole-grid
sweep, so any difference is a slab-boundary mistake (a halo row not
refreshed, an output row written twice or skipped, a refresh of a halo row).
"""
import numpy as np
import pytest
import torch

import kernels_amd as K
from tests._oracle import BF16, F32, F64

pytestmark = pytest.mark.gpu


_INT = {torch.float64: torch.int64, torch.float32: torch.int32, torch.bfloat16: torch.int16}


def same_bits(a: torch.Tensor, b: torch.Tensor) -> bool:
    """Bitwise equality, compared on the device."""
    return a.shape == b.shape and torch.equal(a.view(_INT[a.dtype]), b.view(_INT[b.dtype]))


class Slabs:
    """Slabs owning rows [bounds[s], bounds[s+1]) of an n x n grid."""

    def __init__(self, n, radius, dtype, bounds, in0, out0, pitch=None):
        self.n, self.radius, self.dtype, self.bounds = n, radius, dtype, bounds
        self.parts = []
        for s in range(len(bounds) - 1):
            row0, rows = bounds[s], bounds[s + 1] - bounds[s]
            top, bot = K.slab_halos(n, radius, row0, rows)
            bufs = [K.alloc_slab(n, radius, row0, rows, dtype, pitch=pitch) for _ in range(3)]
            lo = row0 - top
            bufs[0].copy_(in0[lo: row0 + rows + bot])
            bufs[1].fill_(float("nan"))  # halo rows must come from the exchange
            bufs[2].copy_(out0[lo: row0 + rows + bot])
            self.parts.append({"row0": row0, "rows": rows, "top": top, "bot": bot, "in": bufs[0],
                               "nx": bufs[1], "out": bufs[2]})

    def local(self, p, g):
        return g - (p["row0"] - p["top"])

    def sweep(self, pattern):
        for p in self.parts:
            K.sweep_slab(pattern, self.radius, self.dtype, self.n, p["row0"], p["rows"], p["in"], p["nx"],
                         p["out"])
        # halo exchange of the new in: rows above from the slab above, rows
        # below from the slab below (slabs are at least R rows tall)
        for i, p in enumerate(self.parts):
            if p["top"]:
                q = self.parts[i - 1]
                g0 = p["row0"] - p["top"]
                p["nx"][: p["top"]].copy_(q["nx"][self.local(q, g0): self.local(q, p["row0"])])
            if p["bot"]:
                q = self.parts[i + 1]
                g1 = p["row0"] + p["rows"]
                p["nx"][p["top"] + p["rows"]:].copy_(q["nx"][self.local(q, g1): self.local(q, g1 + p["bot"])])
        for p in self.parts:
            p["in"], p["nx"] = p["nx"], p["in"]

    def own(self, p, key):
        return p[key][p["top"]: p["top"] + p["rows"]]


def check_against_whole(pattern, radius, dtype, n, bounds, sweeps, init, seed, pitch=None):
    g_in, g_nx, g_out = (K.alloc_grid(n, dtype, pitch=pitch) for _ in range(3))
    K.init(init, seed, dtype, g_in, g_out)
    slabs = Slabs(n, radius, dtype, bounds, g_in, g_out, pitch=pitch)
    for k in range(sweeps):
        K.sweep(pattern, radius, dtype, g_in, g_nx, g_out)
        g_in, g_nx = g_nx, g_in
        slabs.sweep(pattern)
        torch.cuda.synchronize()
        for p in slabs.parts:
            a, b = p["row0"], p["row0"] + p["rows"]
            assert same_bits(slabs.own(p, "out"), g_out[a:b]), ("out", k, a, b)
            assert same_bits(slabs.own(p, "in"), g_in[a:b]), ("in", k, a, b)
            # halo rows of the new in hold the neighbours' rows
            lo = a - p["top"]
            assert same_bits(p["in"], g_in[lo: b + p["bot"]]), ("halo", k, a, b)


def uneven_bounds(n, parts, radius, seed):
    rng = np.random.default_rng(seed)
    while True:
        cuts = np.sort(rng.choice(np.arange(1, n), parts - 1, replace=False))
        b = [0] + [int(c) for c in cuts] + [n]
        if min(np.diff(b)) >= radius:
            return b


CASES = [
    (0, 1, F64), (0, 2, F64), (1, 2, F64), (0, 5, F64), (1, 3, F64), (0, 8, F64),
    (0, 2, F32), (1, 4, F32), (0, 2, BF16), (1, 1, BF16),
    # star radii swept from a register ring
    (0, 6, F64), (0, 5, F32), (0, 7, BF16),
]


@pytest.mark.parametrize("pattern,radius,dtype", CASES)
def test_uneven_slabs_equal_whole_grid(gpu, pattern, radius, dtype):
    n = 517
    bounds = uneven_bounds(n, 5, radius, seed=radius * 10 + pattern)
    print(f"slabs {bounds} pattern={pattern} R={radius} dtype={dtype}", flush=True)
    check_against_whole(pattern, radius, dtype, n, bounds, sweeps=4, init=1, seed=31)


@pytest.mark.parametrize("radius", [1, 2, 4, 6])
def test_thin_slabs(gpu, radius):
    # slabs of exactly R rows (the thinnest the exchange allows), slabs that
    # lie inside the top and bottom frame, and one whole-grid slab
    n = 64 + 2 * radius
    bounds = list(range(0, 8 * radius, radius)) + [n - 3 * radius, n - radius, n]
    print(f"thin slabs {bounds} R={radius}", flush=True)
    check_against_whole(1, radius, F64, n, bounds, sweeps=3, init=1, seed=7)
    check_against_whole(0, radius, F64, n, [0, n], sweeps=2, init=1, seed=8)


@pytest.mark.parametrize("dtype", [F64, F32, BF16])
def test_slabs_with_unaligned_rows(gpu, dtype):
    # pitch = n with n odd: rows are not 16-byte aligned, so the sweep takes
    # its element-access path, in the slabs as in the whole grid
    n, radius = 301, 3
    bounds = uneven_bounds(n, 4, radius, seed=dtype + 40)
    print(f"unaligned slabs {bounds} dtype={dtype}", flush=True)
    check_against_whole(1, radius, dtype, n, bounds, sweeps=3, init=1, seed=12, pitch=n)


def test_shard_shape_n30000(gpu):
    # n = 30000 on eight slabs of 3750 rows: the share of each of eight GPUs
    n, radius = 30000, 2
    bounds = [3750 * s for s in range(9)]
    print("n=30000 star2 8 slabs", flush=True)
    check_against_whole(0, radius, F64, n, bounds, sweeps=3, init=0, seed=0)
    torch.cuda.empty_cache()

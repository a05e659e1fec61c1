"""Slab geometry on the host (no GPU): the halo rows a slab of a row-strip
split holds, and the algorithmic bytes of the slab sweeps adding up to the
whole grid's plus the halo rows each slab reads."""
import pytest

import kernels_amd as K


@pytest.mark.parametrize("n,radius,parts", [(30000, 2, 8), (517, 3, 5), (64, 8, 3), (11, 5, 1), (1000, 1, 7)])
def test_slab_bytes_add_up(n, radius, parts):
    bounds = [n * k // parts for k in range(parts + 1)]
    total, halo_rows = 0, 0
    for a, b in zip(bounds[:-1], bounds[1:]):
        top, bot = K.slab_halos(n, radius, a, b - a)
        assert (top, bot) == (min(radius, a), min(radius, n - b))
        halo_rows += top + bot
        total += K.slab_bytes_per_iter(n, radius, K.F64, a, b - a)
    assert total == K.stencil_bytes_per_iter(n, radius, K.F64) + 8 * n * halo_rows


def test_slab_inside_the_frame_has_no_outputs():
    # a slab of the first R rows reads and refreshes them but writes no out row
    n, radius = 100, 3
    assert K.slab_bytes_per_iter(n, radius, K.F32, 0, radius) == 4 * n * (2 * radius) + 4 * n * radius

"""GPU check of the raster stage's exact quotient path.

The depth and barycentric divisions (rasterizer.go:149-153) divide by
per-triangle divisors; the HIP path computes them from the correctly rounded
reciprocal of the divisor with two residual corrections. This compares that
sequence with the device's IEEE float32 division bit for bit over the ranges
the raster stage admits, including quotients next to rounding midpoints,
extreme divisor mantissas and integer numerators over integer divisors.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 0x5EED])
def test_exact_division_matches_ieee(ctx, seed):
    n = 1 << 26
    bad = ctypes.c_uint64(0)
    first = (ctypes.c_float * 4)()
    from gorender_amd import _lib
    _lib.check(ctx.lib.gr_check_exact_division(ctx.handle, n, seed, ctypes.byref(bad), first))
    assert bad.value == 0, (f"{bad.value} of {n} quotients differ; first: a={first[0]!r} "
                            f"d={first[1]!r} ieee={first[2]!r} got={first[3]!r}")

"""Mixed precision against FP64 (itself bit-identical to the oracle): the
stated bounds are 1e-4 relative on contact forces and 1 % on the restitution
coefficient and on the packing fraction of a settled bed."""
import math
import sys

import pytest

from gpudem_amd import GDEM_FP64, GDEM_MIXED, accuracy

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_restitution_mixed_within_one_percent(ctx):
    for cor in (0.3, 0.5, 0.9):
        e64 = accuracy.restitution(ctx, GDEM_FP64, cor=cor)
        e32 = accuracy.restitution(ctx, GDEM_MIXED, cor=cor)
        closed = cor ** (math.pi / 3.1415)
        say(f"cor {cor}: fp64 {e64:.6f} mixed {e32:.6f} closed form {closed:.6f}")
        assert e64 == pytest.approx(closed, rel=3e-3)
        assert abs(e32 - e64) / e64 <= 0.01


def test_packing_fraction_mixed_within_one_percent(ctx):
    p64, v64 = accuracy.packing_fraction(ctx, GDEM_FP64)
    p32, v32 = accuracy.packing_fraction(ctx, GDEM_MIXED)
    say(f"packing fraction fp64 {p64:.4f} mixed {p32:.4f}, max speed {max(v64, v32):.3g} m/s")
    assert 0.45 < p64 < 0.7  # a random packing of spheres
    assert max(v64, v32) < 0.05  # settled
    assert abs(p32 - p64) / p64 <= 0.01

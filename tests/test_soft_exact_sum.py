"""The soft pass's exact tap count: for a shadow intensity si in [0, 1] on a
2^-16 grid, the PCF loop's in-order FP32 sum of tap values (1 lit, si
shadowed; AccurateSoftShadow.frag:101-119) equals lit + (taps - lit) * si
computed in FP32, for every pattern of 9 taps and sampled patterns of 16.
Off the grid (0.3) the two can differ, which is why the kernel keeps the
ordered sum there."""
import itertools

import numpy as np
import pytest


def _ordered(pattern, si):
    acc = np.float32(0.0)
    for lit in pattern:
        acc = np.float32(acc + (np.float32(1.0) if lit else si))
    return acc


def _counted(pattern, si):
    lit = int(sum(pattern))
    return np.float32(np.float32(lit) + np.float32(np.float32(len(pattern) - lit) * si))


def _on_grid(si):
    g = np.float32(si) * np.float32(65536.0)
    return 0.0 <= si <= 1.0 and g == np.floor(g)


@pytest.mark.parametrize("si", [0.0, 0.25, 0.5, 0.75, 1.0, 0.5 + 2.0 ** -13, 1.0 - 2.0 ** -16, 2.0 ** -16])
def test_counted_sum_equals_ordered_sum_on_grid(si):
    si = np.float32(si)
    assert _on_grid(si)
    for pattern in itertools.product((0, 1), repeat=9):
        assert _ordered(pattern, si) == _counted(pattern, si)
    rng = np.random.default_rng(1)
    for _ in range(4000):
        pattern = rng.integers(0, 2, 16)
        assert _ordered(pattern, si) == _counted(pattern, si)


def test_off_grid_sums_can_differ():
    si = np.float32(0.3)
    assert not _on_grid(si)
    diffs = sum(_ordered(p, si) != _counted(p, si) for p in itertools.product((0, 1), repeat=9))
    assert diffs > 0

"""Finite-grid PA oracle (restatement of pa_dyadic_bitmask.py:55-97): the
union count against brute force over every pixel and disk, row spans against
the defining inequality, closed forms, and a planted off-by-one."""
import math

import numpy as np
import pytest


def _disks(rng, k, n, extent=10.0):
    xy = rng.uniform(-extent, extent, size=(k, 2))
    r = rng.uniform(0.05, 4.0, size=k)
    lower = np.array([-extent, -extent])
    return xy, r, lower, 2 * extent / n


@pytest.mark.parametrize("n", [8, 16, 32, 64])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_union_count_equals_brute_force(oracle, n, seed):
    rng = np.random.default_rng(seed)
    xy, r, lower, sp = _disks(rng, 1 + 5 * seed, n)
    assert oracle.grid_count(xy, r, lower, sp, n)[0] == oracle.grid_brute(xy, r, lower, sp, n)


def test_row_span_is_exactly_the_inequality(oracle):
    rng = np.random.default_rng(7)
    n = 64
    lower = np.array([-8.0, -8.0])
    sp = 16.0 / n
    for _ in range(200):
        cx, cy = rng.uniform(-9, 9, size=2)
        r = rng.uniform(0, 5)
        y = int(rng.integers(0, n))
        lo, hi = oracle.row_span(cx, cy, r, lower, sp, y, n)
        dy = lower[1] + (y + .5) * sp - cy
        rem = r * r - dy * dy
        inside = [x for x in range(n) if rem >= 0 and (lower[0] + (x + .5) * sp - cx) ** 2 <= rem]
        assert (lo, hi) == ((inside[0], inside[-1] + 1) if inside else (lo, lo)), (cx, cy, r, y)
        if not inside:
            assert hi <= lo


def test_closed_forms(oracle):
    n = 1024
    # one centred disk: area within the pixel-boundary band of its perimeter
    c, _ = oracle.grid_count(np.zeros((1, 2)), np.array([3.0]), np.array([-4.0, -4.0]), 8.0 / n, n)
    sp = 8.0 / n
    assert abs(c * sp * sp - math.pi * 9.0) <= 2 * math.pi * 3.0 * sp
    # disjoint disks add, identical disks do not
    xy = np.array([[-5.0, 0.0], [5.0, 0.0]])
    r = np.array([2.0, 2.0])
    lower = np.array([-8.0, -8.0])
    sp = 16.0 / 256
    both = oracle.grid_count(xy, r, lower, sp, 256)[0]
    one = oracle.grid_count(xy[:1], r[:1], lower, sp, 256)[0]
    two = oracle.grid_count(xy[1:], r[1:], lower, sp, 256)[0]
    assert both == one + two
    assert oracle.grid_count(np.repeat(xy[:1], 3, axis=0), np.repeat(r[:1], 3), lower, sp, 256)[0] == one
    # a disk entirely outside the grid covers nothing
    assert oracle.grid_count(np.array([[50.0, 50.0]]), np.array([2.0]), lower, sp, 256)[0] == 0


def test_planted_off_by_one_is_caught(oracle):
    """Counting [lo, hi] instead of [lo, hi) must disagree with brute force."""
    rng = np.random.default_rng(3)
    n = 32
    xy, r, lower, sp = _disks(rng, 6, n)
    brute = oracle.grid_brute(xy, r, lower, sp, n)
    planted = 0
    for y in range(n):
        cover = np.zeros(n + 1, bool)
        for a in range(len(r)):
            lo, hi = oracle.row_span(xy[a, 0], xy[a, 1], r[a], lower, sp, y, n)
            if hi > lo:
                cover[lo:hi + 1] = True
        planted += int(cover[:n].sum())
    assert planted != brute


def test_projection_frame_and_directions(oracle):
    ds = oracle.adaptive_directions(64, 190003)
    assert ds.shape == (64, 3)
    np.testing.assert_allclose(np.linalg.norm(ds, axis=1), 1.0, atol=1e-12)
    c = np.random.default_rng(1).normal(size=(40, 3))
    r = np.full(40, 1.5)
    for d in ds[:8]:
        xy, lower, sp, u, v, dd = oracle.project_grid(c, r, d, 64)
        assert abs(u @ dd) < 1e-12 and abs(v @ dd) < 1e-12 and abs(u @ v) < 1e-12
        assert np.all(np.abs(xy) + r[:, None] <= -lower[0])
        assert sp == pytest.approx(-2 * lower[0] / 64)

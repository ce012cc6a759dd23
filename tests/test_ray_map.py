"""The wave-coherent ray map (host code of the library, no GPU needed): a
bijection of the net, and each wave's 64 rays lie in one elementary cell."""
import ctypes as ct

import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd.engine import scramble


def rays_of(sob, power):
    cols = np.zeros(32, np.uint32)
    L.call("ehss_ray_map", ct.byref(sob), power, L.ptr(cols))
    t = np.arange(1 << power, dtype=np.uint64)
    if power < 6:
        return t ^ (t >> np.uint64(1))
    w, lane = t >> np.uint64(6), t & np.uint64(63)
    y = w | (lane << np.uint64(power - 6))
    g = np.zeros_like(t)
    for k in range(power):
        g ^= np.where((y >> np.uint64(k)) & np.uint64(1), np.uint64(cols[k]), np.uint64(0))
    return g


@pytest.mark.parametrize("power", [3, 6, 8, 11, 14, 16, 20])
@pytest.mark.parametrize("seed", [0, 1, 190003])
def test_ray_map_is_a_bijection_with_coherent_waves(oracle, power, seed):
    sob = scramble(seed)
    g = rays_of(sob, power)
    assert np.array_equal(np.sort(g), np.arange(1 << power, dtype=np.uint64))
    if power < 8:
        return
    sv, shift = oracle.sobol_scramble(seed)
    q = np.tile(shift.astype(np.uint64), (len(g), 1))
    for j in range(30):
        sel = ((g >> np.uint64(j)) & np.uint64(1)).astype(bool)
        q[sel] ^= sv[:, j].astype(np.uint64)
    u = q.astype(np.float64) / 2 ** 30
    waves = u.reshape(-1, 64, 4)
    widths = np.ptp(waves, axis=1).max(axis=0)
    # sum over dimensions of the fixed bit counts is power - 6: the cell volume
    assert np.prod(np.maximum(widths, 1e-12)) <= 2.0 ** -(power - 6)
    assert widths.max() <= 2.0 ** -((power - 6) // 4 - 1)


def _points(oracle, seed, g):
    sv, shift = oracle.sobol_scramble(seed)
    q = np.tile(shift.astype(np.uint64), (len(g), 1))
    for j in range(30):
        sel = ((g >> np.uint64(j)) & np.uint64(1)).astype(bool)
        q[sel] ^= sv[:, j].astype(np.uint64)
    return q


@pytest.mark.parametrize("power,sub", [(3, 0), (7, 0), (11, 0), (14, 0), (11, 2), (14, 4), (9, 5), (16, 2)])
def test_cross_replicate_schedule(oracle, power, sub):
    """Every replicate's (index -> Gray index) is a bijection, its slot is the one
    its own map gives, and all points with one cell index -- across replicates
    and sub-indices -- lie in one cell of the shared partition.  sub = 0: the
    finest partition (one point of each replicate per cell, the schedule of many
    replicates of small nets); sub = b > 0: a partition of power - b bits, 2^b
    points of each replicate per cell (the schedule of 2^(6 - b) replicates)."""
    seeds = [3 + 7 * k for k in range(20)]
    sobs = (L.Sobol * len(seeds))(*[scramble(s) for s in seeds])
    out = np.zeros((len(seeds), 68), np.uint32)
    L.call("ehss_cross_map", sobs, len(seeds), power, power - sub if sub else 0, L.ptr(out))
    n = 1 << power
    c = np.arange(n, dtype=np.uint64)
    cells = []
    for r, seed in enumerate(seeds):
        inv2, rows1, shift = out[r, :32], out[r, 32:64], np.uint64(out[r, 64])
        if sub:
            assert int(shift) < (1 << (power - sub))  # the digital shift touches cell bits only
        u = c ^ shift
        g = np.zeros(n, np.uint64)
        for k in range(power):
            g ^= np.where((u >> np.uint64(k)) & np.uint64(1), np.uint64(inv2[k]), np.uint64(0))
        assert np.array_equal(np.sort(g), c)
        # slot: u1 = T1 g, t = rotate(u1); the replicate's own map sends t back to g
        u1 = np.zeros(n, np.uint64)
        for k in range(power):
            par = np.array([bin(int(x)).count("1") & 1 for x in (g & np.uint64(rows1[k]))], np.uint64)
            u1 |= par << np.uint64(k)
        if power < 6:
            t = g.copy()
            for sh in (1, 2, 4, 8, 16):
                t ^= t >> np.uint64(sh)
        else:
            lo = np.uint64((1 << (power - 6)) - 1)
            t = ((u1 & lo) << np.uint64(6)) | (u1 >> np.uint64(power - 6))
        own = rays_of(scramble(seed), power)
        assert np.array_equal(own[t.astype(np.int64)], g)
        cells.append(_points(oracle, seed, g))
    # index = cell | sub-index << (power - sub): the points of one cell (every
    # replicate, every sub-index) share the partition's top bits, so their spread
    # in the four coordinates has volume at most 2^-(power - sub)
    q = np.stack(cells).astype(np.float64) / 2 ** 30  # (reps, n, 4)
    q = q.reshape(len(seeds), 1 << sub, n >> sub, 4).transpose(0, 1, 3, 2).reshape(-1, 4, n >> sub)
    spread = np.ptp(q, axis=0)  # (4, cells)
    assert np.prod(np.maximum(spread, 1e-12), axis=0).max() <= 2.0 ** -(power - sub) * (1 + 1e-9)


@pytest.mark.parametrize("power", [5, 11, 14])
def test_cross_maps_of_many_replicates_equal_one_at_a_time(power):
    """The host builds the maps of large launches on several threads; the result
    equals building each replicate's map alone."""
    sob_list = [scramble(11 + 13 * k) for k in range(40)] * 13  # 520 replicates
    sobs = (L.Sobol * len(sob_list))(*sob_list)
    many = np.zeros((len(sob_list), 68), np.uint32)
    L.call("ehss_cross_map", sobs, len(sob_list), power, 0, L.ptr(many))
    for r in range(40):
        one = np.zeros((1, 68), np.uint32)
        L.call("ehss_cross_map", ct.byref(sobs[r]), 1, power, 0, L.ptr(one))
        for k in range(13):
            np.testing.assert_array_equal(many[r + 40 * k], one[0])

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

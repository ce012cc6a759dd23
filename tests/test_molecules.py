"""Synthetic benchmark molecules: geometry and composition as specified."""
import math

import numpy as np
import pytest

from ehss_amd import molecules as M


def test_c60_geometry():
    xyz, el = M.c60()
    assert len(xyz) == 60 and set(el) == {"C"}
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=2)
    np.fill_diagonal(d, np.inf)
    assert d.min() == pytest.approx(1.40, abs=1e-9)
    assert np.all(np.sum(np.isclose(d, 1.40, atol=1e-9), axis=1) == 3)  # three bonds per atom
    np.testing.assert_allclose(np.linalg.norm(xyz, axis=1), 3.4692, atol=2e-4)
    np.testing.assert_allclose(xyz.mean(axis=0), 0, atol=1e-12)
    np.testing.assert_allclose(M.radii_for(el), 3.10)


@pytest.mark.parametrize("n,seed,axes", [(650, 1, (1, 1, 1)), (2000, 2, (1, 1, 1.4))])
def test_globule(n, seed, axes):
    xyz, el = M.globule(n, seed, axes)
    assert len(xyz) == n
    a = (3 * n / (4 * math.pi * M.DENSITY * np.prod(axes))) ** (1 / 3)
    assert np.all(np.sum((xyz / (a * np.asarray(axes))) ** 2, axis=1) <= 1 + 1e-12)
    from scipy.spatial import cKDTree
    pairs = cKDTree(xyz).query_pairs(1.0 - 1e-9)
    assert not pairs
    counts = {e: el.count(e) for e in ("H", "C", "O", "N", "S")}
    total = sum(f for _, f in M.MIX)
    for e, f in M.MIX:
        assert abs(counts[e] - f / total * n) < 1
    again, _ = M.globule(n, seed, axes)
    np.testing.assert_array_equal(xyz, again)


def test_conformer_jitter():
    xyz, _ = M.globule(300, 9)
    j = M.conformer(xyz, 4, 0.3)
    assert np.std(j - xyz) == pytest.approx(0.3, rel=0.1)

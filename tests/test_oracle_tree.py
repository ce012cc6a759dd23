"""Pins the oracle's grouping and flattened trees: two independent builders must
agree (static grouping == PA-tree leaf order, tests/test_ehss_static.py:10-21),
plus the topology algebra and box containment."""
import numpy as np
import pytest

from tests import scenes

KINDS = ["random", "ties", "coincident", "single", "translated"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("leaf", [1, 4, 9])
def test_static_groups_equal_pa_leaf_order(oracle, kind, leaf):
    c, r = scenes.fixture_static(kind)
    groups, ng = oracle.static_groups(c, r, leaf)
    pt = oracle.pa_tree(c, r, leaf)
    np.testing.assert_array_equal(groups, oracle.pa_groups(pt, len(r)))
    assert ng == int((pt["count"] > 0).sum())
    counts = np.bincount(groups)
    if kind not in ("coincident",):
        assert counts.max() <= max(leaf, 1) or kind == "ties"


@pytest.mark.parametrize("ng", [1, 2, 3, 7, 16, 33, 256, 1000])
def test_topology_algebra(oracle, ng):
    t = oracle.topology(ng)
    nn = 2 * ng - 1
    assert len(t["end"]) == nn and t["end"][0] == nn
    leaves = np.flatnonzero(t["leaf"] >= 0)
    np.testing.assert_array_equal(t["leaf"][leaves], np.arange(ng))  # leaves in preorder = group order
    for node in range(nn):
        if t["left"][node] >= 0:
            l, rr = t["left"][node], t["right"][node]
            assert l == node + 1 and t["end"][l] == rr and t["end"][rr] == t["end"][node]
        else:
            assert t["end"][node] == node + 1


@pytest.mark.parametrize("name", ["f29", "random47", "coincident47", "translated47", "c60"])
@pytest.mark.parametrize("leaf", [1, 4])
def test_boxes_nest_and_contain_spheres(oracle, name, leaf):
    c, r = scenes.SCENES[name]()
    t = oracle.static_tree(c, r, leaf)
    pad = 1e-10 * max(1., r.max())
    lo, hi, end, leafg, off, ids = t["lo"], t["hi"], t["end"], t["leaf_group"], t["offsets"], t["ids"]
    for node in range(len(end)):
        g = leafg[node]
        if g >= 0:
            atoms = ids[off[g]:off[g + 1]]
            np.testing.assert_array_equal(lo[node], (c[atoms] - r[atoms, None]).min(axis=0) - pad)
            np.testing.assert_array_equal(hi[node], (c[atoms] + r[atoms, None]).max(axis=0) + pad)
        else:
            sub = range(node + 1, end[node])
            assert np.all(lo[node] <= lo[list(sub)].min(axis=0)) and np.all(hi[node] >= hi[list(sub)].max(axis=0))
    np.testing.assert_array_equal(np.sort(ids), np.arange(len(r)))


def test_pa_tree_layout(oracle):
    c, r = scenes.globule(650, 1)
    pt = oracle.pa_tree(c, r, 4)
    n = len(pt["end"])
    assert pt["end"][0] == n
    leaves = pt["count"] > 0
    assert pt["count"][leaves].sum() == len(r)
    np.testing.assert_array_equal(np.sort(pt["ids"]), np.arange(len(r)))
    # preorder: first[] of leaves increases and tiles the id array
    f = pt["first"][leaves]
    assert np.all(np.diff(f) > 0) and f[0] == 0

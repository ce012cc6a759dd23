"""CPU: the tree shape the shaped build relies on (ehss_tree_counts).

Without two atoms sharing a key triple, the median-split tree's group count,
node counts and depth are functions of (n, leaf_size).  Checked against the
oracle's trees (ehss_static.py / pa_spatial.py restated) for random structures,
ties on single axes and sizes with several halvings; a structure with coincident
atoms shows why the precondition is needed (zero-span leaves come early)."""
import numpy as np
import pytest

from ehss_amd import _lib as L


def _counts(n, leaf):
    out = np.zeros(4, np.int64)
    L.call("ehss_tree_counts", int(n), int(leaf), L.ptr(out))
    return [int(v) for v in out]


def _depth(pa):
    """Deepest PA node: preorder with skip pointers (a node's subtree ends at end)."""
    stack, deepest = [], 0
    for i, e in enumerate(pa["end"]):
        while stack and stack[-1] <= i:
            stack.pop()
        deepest = max(deepest, len(stack))
        stack.append(int(e))
    return deepest


def _structure(rng, n, ties):
    c = rng.normal(size=(n, 3)) * 2 * max(1.0, n ** (1 / 3))
    r = rng.uniform(1.0, 2.0, n)
    if ties:
        c, r = np.round(c, 1), np.round(r, 1)
    keys = ((c - r[:, None]) + (c + r[:, None])) / 2
    keep = np.sort(np.unique(keys, axis=0, return_index=True)[1])
    return c[keep], r[keep]


@pytest.mark.parametrize("leaf", [1, 2, 4, 9, 32])
def test_counts_equal_the_oracle_trees(oracle, leaf):
    rng = np.random.default_rng(leaf)
    sizes = list(range(1, 70)) + [127, 128, 129, 1000, 4095, 4096, 4097, 8193, 20000]
    for n in sizes:
        c, r = _structure(rng, n, ties=n % 2 == 0)
        tree, pa = oracle.static_tree(c, r, leaf), oracle.pa_tree(c, r, leaf)
        ng = int(tree["groups"].max()) + 1
        assert _counts(len(r), leaf) == [ng, len(tree["end"]), len(pa["end"]), _depth(pa)], (len(r), leaf)


def test_zero_span_segments_break_the_shape(oracle):
    """Six coincident atoms with leaf size 1 stop splitting at once (zero span):
    one group instead of six -- such structures take the general build."""
    c, r = np.zeros((6, 3)), np.full(6, 1.5)
    tree = oracle.static_tree(c, r, 1)
    assert int(tree["groups"].max()) + 1 == 1
    assert _counts(6, 1)[0] == 6


def test_counts_argument_checks():
    out = np.zeros(4, np.int64)
    assert L.fn("ehss_tree_counts")(0, 4, L.ptr(out)) == L.EHSS_ERR_INVALID
    assert L.fn("ehss_tree_counts")(10, 0, L.ptr(out)) == L.EHSS_ERR_INVALID

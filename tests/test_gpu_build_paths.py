"""GPU: the shaped tree build (build_shaped.hip: counts known from (n, leaf_size),
no synchronisation) against the level-by-level build (forced with
EHSS_BUILD_PATH=general) and the oracle, bit for bit, at sizes around the
one-workgroup limit (4096 atoms) and with several top levels; the path choice
(duplicated key triples, caller groups, leaf sizes above 32 take the general
build); and ordering: rebuilds enqueued back to back, a plan on its own stream
right after a rebuild (the allocation-free solve on a caller stream right after
a rebuild is tested with the other allocation-free tests)."""
import contextlib
import os

import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd.engine import GPUPlan, GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

TREE_KEYS = ("lo", "hi", "end", "leaf_group", "offsets", "ids", "groups")
PA_KEYS = ("lo", "hi", "end", "first", "count", "ids")


@contextlib.contextmanager
def general_build():
    os.environ["EHSS_BUILD_PATH"] = "general"
    try:
        yield
    finally:
        del os.environ["EHSS_BUILD_PATH"]


def _has_duplicate_keys(c, r):
    keys = ((c - r[:, None]) + (c + r[:, None])) / 2 + 0.0
    return len(np.unique(keys, axis=0)) < len(r)


def _structures():
    rng = np.random.default_rng(5)
    out = []
    for n in (1, 2, 3, 5, 33, 129, 650, 4095, 4096, 4097, 8193, 12001, 20000):
        out.append((f"random{n}", rng.normal(size=(n, 3)) * 1.5 * n ** (1 / 3), rng.uniform(1.0, 2.0, n)))
    # coordinates on a 0.1 A grid with three radii: many ties on single axes
    c = np.round(rng.normal(size=(9000, 3)) * 12, 1)
    r = rng.choice([1.2, 1.5, 1.7], 9000)
    keep = np.unique(((c - r[:, None]) + (c + r[:, None])) / 2, axis=0, return_index=True)[1]
    out.append(("grid9000", c[np.sort(keep)], r[np.sort(keep)]))
    out.append(("ties47",) + scenes.fixture_static("ties"))
    out.append(("globule650",) + scenes.globule(650, 1))
    return out


def _same(a, b, keys, tag):
    for k in keys:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{tag}/{k}")


def _counts(n, leaf):
    out = np.zeros(4, np.int64)
    L.call("ehss_tree_counts", int(n), int(leaf), L.ptr(out))
    return out


@pytest.mark.parametrize("leaf", [1, 4, 9, 32])
def test_shaped_build_equals_general_build_and_oracle(oracle, leaf):
    for name, c, r in _structures():
        say("build paths", name, "leaf", leaf)
        shaped = not _has_duplicate_keys(c, r)
        with GPUScene(c, r, leaf_size=leaf) as fast:
            assert fast.info["build_path"] == (L.BUILD_SHAPED if shaped else L.BUILD_GENERAL), name
            tf, pf = fast.export_tree(), fast.export_pa_tree()
            with general_build(), GPUScene(c, r, leaf_size=leaf) as gen:
                assert gen.info["build_path"] == L.BUILD_GENERAL
                for k in ("groups", "nodes", "pa_nodes", "max_depth"):
                    assert fast.info[k] == gen.info[k], (name, k)
                _same(tf, gen.export_tree(), TREE_KEYS, f"{name}/{leaf}/general")
                _same(pf, gen.export_pa_tree(), PA_KEYS, f"{name}/{leaf}/general/pa")
            if shaped:
                assert list(_counts(len(r), leaf)) == [fast.info[k] for k in ("groups", "nodes", "pa_nodes",
                                                                               "max_depth")]
        _same(tf, oracle.static_tree(c, r, leaf), TREE_KEYS, f"{name}/{leaf}/oracle")
        _same(pf, oracle.pa_tree(c, r, leaf), PA_KEYS, f"{name}/{leaf}/oracle/pa")


def test_path_choice():
    c, r = scenes.fixture_static("coincident")  # every atom at the origin: zero-span segments
    with GPUScene(c, r) as sc:
        assert sc.info["build_path"] == L.BUILD_GENERAL
    c, r = scenes.globule(300, 4)
    with GPUScene(c, r, leaf_size=33) as sc:
        assert sc.info["build_path"] == L.BUILD_GENERAL
    with GPUScene(c, r, groups=np.arange(len(r)) // 5) as sc:
        assert sc.info["build_path"] == L.BUILD_GENERAL
    with GPUScene(c, r) as sc:
        assert sc.info["build_path"] == L.BUILD_SHAPED
        sc.rebuild(*scenes.fixture_static("coincident"))
        assert sc.info["build_path"] == L.BUILD_GENERAL


def test_back_to_back_rebuilds_then_solve_equal_a_fresh_scene():
    """Rebuilds return before their kernels run; growing, shrinking and
    switching build paths in a row must still end in the last structure."""
    chain = [scenes.globule(200, 1), scenes.globule(9000, 2), scenes.fixture_static("coincident"),
             scenes.globule(700, 3), scenes.globule(5000, 4), scenes.globule(650, 1)]
    with GPUScene(*chain[0]) as sc:
        for c, r in chain[1:]:
            sc.rebuild(c, r)
        with GPUScene(*chain[-1]) as fresh:
            _same(sc.export_tree(), fresh.export_tree(), TREE_KEYS, "chain")
            a = sc.solve(scramble(3), 14, 64, histogram=True)
            b = fresh.solve(scramble(3), 14, 64, histogram=True)
            for k in ("omega", "pa", "escaped", "unresolved", "max_bounces"):
                assert a[k] == b[k], k
            np.testing.assert_array_equal(a["histogram"], b["histogram"])


def test_plan_right_after_rebuild_follows_the_rebuild():
    """A plan launches on its own stream: it must wait for its scenes' rebuilds."""
    structs = [scenes.globule(6000, 7), scenes.globule(2500, 8), scenes.globule(400, 9)]
    alone = []
    for c, r in structs:
        with GPUScene(c, r) as s:
            alone.append(s.solve(scramble(11), 13, 64))
    scs = [GPUScene(*scenes.globule(100, 20 + k)) for k in range(3)]
    try:
        for sc, (c, r) in zip(scs, structs):
            sc.rebuild(c, r)
        with GPUPlan(scs, [scramble(11)] * 3, [13] * 3, [64] * 3) as plan:
            got = plan.run()
        for k in range(3):
            assert got[k]["omega"] == alone[k]["omega"] and got[k]["escaped"] == alone[k]["escaped"], k
    finally:
        for sc in scs:
            sc.close()


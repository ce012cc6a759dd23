"""Pins the oracle's trace, PA and readout: against the reference's own compiled
C++ kernel (oracle/_ref, bitwise), closed forms and invariants re-expressed
from the reference's tests, and brute force."""
import numpy as np
import pytest

from tests import scenes


def _src(O, c, r, kind, power, seed):
    return O.external_source(c, r, power, seed) if kind == "external" else O.reverse_pa_source(c, r, power, seed)


def _same(a, b, keys=("positions", "outgoing", "escaped", "unresolved", "bounces", "collider_ids")):
    for k in keys:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("name,kind,power,seed", [("f29", "external", 10, 881), ("f29", "reverse", 10, 881),
                                                  ("f37", "external", 9, 321), ("coincident47", "external", 7, 5),
                                                  ("translated47", "external", 8, 572)])
@pytest.mark.parametrize("cap", [1, 4, 64])
def test_oracle_equals_reference_cpp_kernel(oracle, name, kind, power, seed, cap):
    """tests/test_ehss_cpu.py:163-175 re-expressed: the reference's src/native/ehss_cpu.cpp
    (compiled from its own source into oracle/_ref) is bitwise equal to the oracle."""
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("reference C++ kernel not built (oracle/_ref needs the reference checkout)")
    c, r = scenes.SCENES[name]()
    tree = oracle.static_tree(c, r, 4)
    src = _src(oracle, c, r, kind, power, seed)
    a = oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"], src["initial_bounces"], cap=cap)
    for threads in (1, 4):
        b = oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"], src["initial_bounces"],
                         cap=cap, lib_handle=ref, threads=threads)
        _same(a, b)
        assert a["box_tests"] == b["box_tests"] and a["sphere_tests"] == b["sphere_tests"]


@pytest.mark.parametrize("radius", [.01, 1., 100.])
def test_single_sphere_analytic(oracle, radius):
    """tests/test_ehss_reference.py:8-17: Omega = PA = pi r^2."""
    c, r = np.array([[2., -1., 4.]]), np.array([radius])
    tree = oracle.static_tree(c, r, 4)
    for src in (oracle.external_source(c, r, 14, 98), oracle.reverse_pa_source(c, r, 14, 98)):
        res = oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"], src["initial_bounces"],
                           cap=4)
        ro = oracle.readout(src, res, 4)
        assert res["escaped"].all() and not res["unresolved"].any() and np.all(res["bounces"] == 1)
        assert ro["omega"] == pytest.approx(np.pi * radius * radius, rel=1e-6)
        assert ro["pa"] == pytest.approx(np.pi * radius * radius, rel=1e-7)
        assert np.allclose(np.linalg.norm(res["outgoing"], axis=1), 1., atol=1e-14)


def test_closed_form_reflection(oracle):
    """tests/test_ehss_reference.py:27-34."""
    c, r = np.zeros((1, 3)), np.ones(1)
    tree = oracle.static_tree(c, r, 4)
    b = np.array([0., .2, .8, .999])
    o = np.column_stack((np.full(4, -3.), b, np.zeros(4)))
    d = np.tile([1., 0., 0.], (4, 1))
    res = oracle.trace(tree, c, r, o, d, cap=64)
    src = dict(incoming=d, weights=np.ones(4))
    ro = oracle.readout(src, res, 64)
    assert np.allclose(res["outgoing"][:, 0], 2 * b * b - 1., atol=2e-14)
    assert np.allclose(ro["contributions"], 2 * (1 - b * b), atol=2e-14)


def test_distance_order_small_gap_and_trapped_pair(oracle):
    """tests/test_ehss_reference.py:20-24, 37-44, 69-73."""
    c, r = np.array([[5., 0., 0.], [0., 0., 0.]]), np.ones(2)
    t = oracle.trace_flat(c, r, [[-3., 0., 0.]], [[1., 0., 0.]], cap=1)
    assert t["collider_ids"][0, 0] == 1 and t["positions"][0, 0] == -1.
    keep = [np.array(v, float) for v in ([-3., 1.01, 0.], [1., 0., 0.], [0., 0., 0.])]
    assert np.isinf(oracle.lib().oracle_entry(*(oracle._ptr(a) for a in keep), 1.))
    gap = 1e-10
    c, r = np.array([[0., 0., 0.], [2. + gap, 0., 0.]]), np.ones(2)
    t = oracle.trace_flat(c, r, [[1., 0., 0.]], [[1., 0., 0.]], last_atom=[0], initial_bounces=[1], cap=2)
    assert t["collider_ids"][0, 1] == 1
    c, r = np.array([[-2., 0., 0.], [2., 0., 0.]]), np.ones(2)
    tree = oracle.static_tree(c, r, 4, groups=np.arange(2))
    res = oracle.trace(tree, c, r, [[0., 0., 0.]], [[1., 0., 0.]], cap=7)
    assert res["unresolved"][0] and not res["escaped"][0]
    assert res["collider_ids"][0].tolist() == [1, 0, 1, 0, 1, 0, 1]


def test_tangency_grid(oracle):
    """tests/test_ehss_cpu.py:140-148: y in {0, .9r, r, nextafter(r)} x r in {1e-3, 1, 100}."""
    center = np.zeros(3)
    d = np.array([1., 0., 0.])
    L = oracle.lib()
    for radius in (.001, 1., 100.):
        for y in (0., .9 * radius, radius, np.nextafter(radius, np.inf)):
            for x in (-3 * radius, -.5 * radius, 2 * radius):
                o = np.array([x, y, 0.])
                t = L.oracle_entry(oracle._ptr(o), oracle._ptr(d), oracle._ptr(center), radius)
                # closed form for a ray along +x: hit iff x < 0 and y <= r; entry -x - sqrt(r^2 - y^2),
                # starts inside beyond the 64 eps r band -> no hit, within it -> 0
                if x >= 0 or y > radius:
                    assert np.isinf(t)
                    continue
                near = -x - np.sqrt(radius * radius - y * y)
                if near < -64 * np.finfo(float).eps * radius:
                    assert np.isinf(t)
                else:
                    assert t == pytest.approx(max(0., near), rel=1e-13, abs=1e-13 * radius)


def test_cap_monotone_and_bound(oracle):
    """tests/test_ehss_reference.py:47-56."""
    c, r = scenes.fixture_cross()
    tree = oracle.static_tree(c, r, 4)
    src = oracle.reverse_pa_source(c, r, 14, 114)
    small = oracle.readout(src, oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"],
                                             src["initial_bounces"], cap=1), 1)
    full_t = oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"], src["initial_bounces"],
                          cap=64)
    full = oracle.readout(src, full_t, 64)
    assert np.count_nonzero(full_t["bounces"] >= 2) > 100
    assert full["omega"] >= small["omega"] - 1e-12
    assert full["omega"] <= small["omega"] + small["tail_bound"] + 1e-12


@pytest.mark.parametrize("name", ["f37", "f29", "coincident47", "ties47"])
def test_flat_equals_tree_and_independent_reference(oracle, name):
    """tests/test_ehss_native_ablation.py:10-28: traversal independence (bitwise)
    and the independent all-atom twin within 1e-10."""
    c, r = scenes.SCENES[name]()
    tree = oracle.static_tree(c, r, 4)
    src = oracle.external_source(c, r, 10, 321)
    a = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=64)
    f = oracle.trace_flat(c, r, src["origins"], src["directions"], cap=64)
    _same(a, f)
    ind = oracle.trace_allatom(c, r, src["origins"], src["directions"], cap=64)
    np.testing.assert_allclose(ind["outgoing"], a["outgoing"], rtol=0, atol=1e-10)


def test_thread_count_invariance(oracle):
    """tests/test_ehss_cpu.py:36-40: 1/2/4 threads bit-identical."""
    c, r = scenes.fixture_29()
    tree = oracle.static_tree(c, r, 4)
    src = oracle.external_source(c, r, 11, 44)
    base = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=64, threads=1)
    for th in (2, 4):
        _same(oracle.trace(tree, c, r, src["origins"], src["directions"], cap=64, threads=th), base)


@pytest.mark.parametrize("scale", [.01, 1., 100.])
def test_rigid_motion_covariance(oracle, scale):
    """tests/test_ehss_publication_invariance.py:14-30."""
    from scipy.spatial.transform import Rotation
    c, r = scenes.fixture_pose()
    src = oracle.external_source(c, r, 12, 170100001)
    tree = oracle.static_tree(c, r, 4)
    orig = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=256)
    rot = Rotation.from_rotvec([.31, -.44, .28]).as_matrix()
    shift = scale * np.array([9., -4., 7.])
    c2, r2 = scale * c @ rot.T + shift, scale * r
    o2, d2 = scale * src["origins"] @ rot.T + shift, src["directions"] @ rot.T
    moved = oracle.trace(oracle.static_tree(c2, r2, 4), c2, r2, o2, d2, cap=256)
    for k in ("escaped", "bounces", "collider_ids"):
        np.testing.assert_array_equal(moved[k], orig[k])
    np.testing.assert_allclose(moved["outgoing"], orig["outgoing"] @ rot.T, rtol=1e-8, atol=1e-8)


def test_permutation_invariance(oracle):
    """tests/test_ehss_publication_invariance.py:33-43."""
    c, r = scenes.fixture_pose()
    src = oracle.external_source(c, r, 12, 170100002)
    orig = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=256)
    order = np.array([3, 1, 0, 2])
    c2, r2 = c[order], r[order]
    res = oracle.trace(oracle.static_tree(c2, r2, 4), c2, r2, src["origins"], src["directions"], cap=256)
    np.testing.assert_array_equal(res["outgoing"], orig["outgoing"])
    mapped = res["collider_ids"].copy()
    valid = mapped >= 0
    mapped[valid] = order[mapped[valid]]
    np.testing.assert_array_equal(mapped, orig["collider_ids"])


@pytest.mark.parametrize("name", ["f29", "f23", "c60", "coincident47"])
def test_pa_tree_query_equals_brute_force(oracle, name):
    c, r = scenes.SCENES[name]()
    tree = oracle.pa_tree(c, r, 4)
    src = oracle.external_source(c, r, 11, 3)
    hits, boxes, tests = oracle.pa_query(tree, c, r, src["origins"], src["directions"])
    np.testing.assert_array_equal(hits, oracle.pa_brute(c, r, src["origins"], src["directions"]))
    assert boxes > 0 and tests > 0


def test_pa_bvh_close_to_ehss_pa(oracle):
    """tests/test_ccs_cli.py:73-85: BVH PA on the EHSS source matches the EHSS-side PA."""
    c, r = scenes.fixture_23()
    src = oracle.external_source(c, r, 10, 190003)
    hits, _, _ = oracle.pa_query(oracle.pa_tree(c, r, 4), c, r, src["origins"], src["directions"])
    res = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=1)
    ro = oracle.readout(src, res, 1)
    assert res["unresolved"].any()
    assert float(src["weights"][hits].sum()) == pytest.approx(ro["pa"])


def test_readout_sum_and_histogram(oracle):
    import math
    c, r = scenes.c60()
    src = oracle.external_source(c, r, 12, 0)
    res = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=64)
    ro = oracle.readout(src, res, 64)
    assert ro["omega"] == pytest.approx(math.fsum(ro["contributions"]), rel=1e-14)
    assert ro["histogram"].sum() == len(src["weights"])
    assert np.all((ro["contributions"] >= 0) & (ro["contributions"] <= 2 * src["weights"]))


def test_sobol_single_sphere_pi_r2(oracle):
    """tests/test_ehss_native_ablation.py:79-91: Sobol single-sphere Omega within 1e-6."""
    c, r = np.zeros((1, 3)), np.array([3.1])
    src = oracle.external_source(c, r, 12, 91)
    res = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=64)
    assert oracle.readout(src, res, 64)["omega"] == pytest.approx(np.pi * 3.1 ** 2, rel=1e-6)


def test_tie_rule_lowest_id_wins_against_traversal_order(oracle):
    """ehss_cpu.cpp:67: on equal distances the lower atom id wins.  The mirrored
    pair below has bit-identical entry distances and the median split puts atom 1
    (y = -1) in the first group, so a first-found rule would return atom 1."""
    c = np.array([[5., 1., 0.], [5., -1., 0.]])
    r = np.array([1.5, 1.5])
    tree = oracle.static_tree(c, r, 1)
    assert tree["groups"][1] < tree["groups"][0]
    L = oracle.lib()
    o, d = np.array([-3., 0., 0.]), np.array([1., 0., 0.])
    t0 = L.oracle_entry(oracle._ptr(o), oracle._ptr(d), oracle._ptr(c[0]), 1.5)
    t1 = L.oracle_entry(oracle._ptr(o), oracle._ptr(d), oracle._ptr(c[1]), 1.5)
    assert t0 == t1 and np.isfinite(t0)
    res = oracle.trace(tree, c, r, [o], [d], cap=1)
    assert res["collider_ids"][0, 0] == 0

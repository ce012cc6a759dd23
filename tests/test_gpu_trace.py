"""GPU parity of the trace kernels through the C-ABI: bit-exact against the
oracle on identical rays (the reference's own fixtures, analytic geometries and
the synthetic molecules), for the plain FP64 kernel and the FP32-filter kernel."""
import numpy as np
import pytest

from ehss_amd import _lib as L
from tests import scenes
from tests.gpu_helpers import assert_same_paths, compat_trace, say

pytestmark = pytest.mark.gpu


def test_device_fp64_division_and_sqrt_are_ieee():
    rng = np.random.default_rng(2024)
    n = 1 << 20
    a = np.abs(rng.normal(size=n)) * 10.0 ** rng.uniform(-300, 300, n)
    b = rng.normal(size=n) * 10.0 ** rng.uniform(-300, 300, n)
    special = np.array([1.0, 2.0, 0.0, -0.0, np.inf, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                        np.nextafter(1, 2), np.nextafter(1, 0), 3.0, 9.0, 0.1, 1e-310])
    a[:len(special)] = special
    b[:len(special)] = special[::-1] + 1.0
    q, r = np.empty(n), np.empty(n)
    say("selftest fp64 div/sqrt", n)
    L.call("ehss_selftest_fp64", n, L.ptr(a), L.ptr(b), L.ptr(q), L.ptr(r))
    assert np.array_equal(q.view(np.int64), (a / b).view(np.int64))
    assert np.array_equal(r.view(np.int64), np.sqrt(a).view(np.int64))


CASES = [
    ("f29", "external", 10, 881), ("f29", "reverse", 9, 781), ("f37", "external", 10, 321),
    ("random47", "external", 8, 572), ("ties47", "external", 8, 572), ("coincident47", "external", 8, 572),
    ("single", "external", 8, 572), ("translated47", "external", 8, 572), ("pose4", "external", 12, 170100001),
    ("cross6", "reverse", 12, 114), ("f23", "external", 10, 190003), ("c60", "external", 12, 0),
]


def _source(O, c, r, kind, power, seed):
    if kind == "external":
        return O.external_source(c, r, power, seed)
    return O.reverse_pa_source(c, r, power, seed)


@pytest.mark.parametrize("precision", [0, 1])
def test_trace_matches_oracle_bitwise(oracle, precision):
    O = oracle
    for name, kind, power, seed in CASES:
        c, r = scenes.SCENES[name]()
        tree = O.static_tree(c, r, 4)
        src = _source(O, c, r, kind, power, seed)
        for cap in (1, 4, 64):
            label = f"{name}/{kind}/2^{power}/cap{cap}/prec{precision}"
            say(label)
            exp = O.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"], src["initial_bounces"],
                          cap=cap)
            got = compat_trace(tree, c, r, src, cap, precision=precision)
            assert_same_paths(got, exp, label, counters=(precision == 0))


@pytest.mark.parametrize("leaf", [1, 4, 9])
def test_trace_leaf_sizes_and_globule(oracle, leaf):
    O = oracle
    c, r = scenes.globule(650, 1)
    tree = O.static_tree(c, r, leaf)
    src = O.external_source(c, r, 12, 7)
    exp = O.trace(tree, c, r, src["origins"], src["directions"], cap=64, threads=8)
    for precision in (0, 1):
        label = f"globule650/leaf{leaf}/prec{precision}"
        say(label)
        got = compat_trace(tree, c, r, src, 64, precision=precision)
        assert_same_paths(got, exp, label, counters=(precision == 0))


def test_reference_void_abi_reports_threads_and_counters(oracle):
    O = oracle
    c, r = scenes.fixture_29()
    tree = O.static_tree(c, r, 4)
    src = O.external_source(c, r, 10, 881)
    exp = O.trace(tree, c, r, src["origins"], src["directions"], cap=64)
    say("void ehss_trace abi")
    got = compat_trace(tree, c, r, src, 64, void_abi=True, threads=4)
    assert got["observed_threads"] == 4
    assert_same_paths(got, exp, "void-abi")


def test_trapped_pair_and_empty_input(oracle):
    """tests/test_ehss_cpu.py:151-160 and test_ehss_reference.py:37-44."""
    O = oracle
    c, r = np.array([[-2., 0., 0.], [2., 0., 0.]]), np.ones(2)
    tree = O.static_tree(c, r, 4, groups=np.arange(2))
    for n in (0, 8):
        src = dict(origins=np.zeros((n, 3)), directions=np.tile([1., 0., 0.], (n, 1)),
                   last_atom=np.full(n, -1), initial_bounces=np.zeros(n, dtype=np.int64))
        for precision in (0, 1):
            say("trapped pair", n, precision)
            got = compat_trace(tree, c, r, src, 7, precision=precision)
            assert got["unresolved"].all() and not got["escaped"].any()
            if n:
                assert got["collider_ids"][0].tolist() == [1, 0, 1, 0, 1, 0, 1]


def test_reflection_closed_form(oracle):
    """tests/test_ehss_reference.py:27-34: v_x = 2 b^2 - 1 for impact parameter b."""
    O = oracle
    c, r = np.zeros((1, 3)), np.ones(1)
    tree = O.static_tree(c, r, 4)
    b = np.array([0., .2, .8, .999])
    src = dict(origins=np.column_stack((np.full(4, -3.), b, np.zeros(4))), directions=np.tile([1., 0., 0.], (4, 1)),
               last_atom=np.full(4, -1), initial_bounces=np.zeros(4, dtype=np.int64))
    for precision in (0, 1):
        got = compat_trace(tree, c, r, src, 64, precision=precision)
        assert np.allclose(got["outgoing"][:, 0], 2 * b * b - 1., atol=2e-14)
        assert np.all(got["bounces"] == 1)


def test_invalid_arguments_are_rejected(oracle):
    O = oracle
    c, r = scenes.fixture_29()
    tree = O.static_tree(c, r, 4)
    src = O.external_source(c, r, 4, 1)
    with pytest.raises(L.EHSSError, match="cap"):
        compat_trace(tree, c, r, src, 0)
    bad = dict(src, last_atom=np.full(16, -2))
    with pytest.raises(L.EHSSError, match="last atom"):
        compat_trace(tree, c, r, bad, 8)

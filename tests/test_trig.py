"""CPU pins of the source-angle trig (ehss_amd/csrc/ehss_trig.h, host build).

* Correct rounding is checked against mpmath at 160 bits (an independent
  implementation) on structured and random angles, Sobol and generic.
* The compatibility table must make the device function equal NumPy's cos/sin
  (the host libm that the reference's external_source uses,
  ehss_state.py:45-49) on every listed word and on random words; a planted
  mistake in the table (a flipped direction, a dropped entry) must fail.
"""
import mpmath as mp
import numpy as np
import pytest

from ehss_amd import build as B
from ehss_amd import trig as T


@pytest.fixture(scope="module")
def table():
    if not T.HOST_LIB.exists():
        B.build_trig_host()
    B.build_trig_table()
    return T.load_table()


def _mp_cr(x):
    mp.mp.prec = 160
    v = mp.mpf(float(x))
    return float(mp.sin(v)), float(mp.cos(v))


def _structured_words():
    q = [0, 1, 2, 3, (1 << 30) - 1, (1 << 30) - 2]
    for k in range(1, 4):  # quadrant boundaries 2pi k/4 and their neighbours
        q += [k << 28, (k << 28) - 1, (k << 28) + 1]
    q += [1 << j for j in range(30)]
    q += [(1 << 30) // 8 * k for k in range(8)]
    return np.array(q, dtype=np.uint32)


def test_correctly_rounded_against_mpmath():
    rng = np.random.default_rng(11)
    q = np.concatenate([_structured_words(), rng.integers(0, 1 << 30, 1500).astype(np.uint32)])
    x = np.concatenate([T.angles(q), rng.uniform(-8, 8, 500), [0.0, 1e-300, 7.5, -7.9, np.pi / 4, np.pi / 2]])
    s, c, sure = T.sincos_cr(x)
    assert sure.all()
    for i, xi in enumerate(x):
        ms, mc = _mp_cr(xi)
        assert s[i] == ms and c[i] == mc, (xi, s[i], ms, c[i], mc)


def test_quadrant_words_give_tiny_exact_residues():
    """2pi q 2^-30 at q = 2^29 is the double nearest pi: sin is pi - fl(pi),
    the textbook residue 1.2246467991473532e-16."""
    s, c, _ = T.sincos_cr(T.angles(np.array([1 << 29], np.uint32)))
    assert s[0] == 1.2246467991473532e-16 and c[0] == -1.0


def test_table_matches_host_libm_on_every_listed_word(table):
    assert table.libm == T.host_libm()
    assert 0.45 < table.lookup_from < 0.5  # the libm errs only near rounding midpoints
    q, flags = table.words()
    assert len(q) > 0 and np.all(np.diff(q.astype(np.int64)) > 0)
    x = T.angles(q)
    s, c = T.sobol_sincos(q, table)
    np.testing.assert_array_equal(c, np.cos(x))
    np.testing.assert_array_equal(s, np.sin(x))
    # every listed word really differs from the correctly rounded value
    s0, c0 = T.sobol_sincos(q, None)
    cos_listed = (flags & 1) != 0
    sin_listed = (flags & 4) != 0
    assert np.all(cos_listed | sin_listed)
    np.testing.assert_array_equal(c0 != np.cos(x), cos_listed)
    np.testing.assert_array_equal(s0 != np.sin(x), sin_listed)


def test_table_matches_host_libm_on_random_and_structured_words(table):
    rng = np.random.default_rng(12)
    q = np.concatenate([_structured_words(), rng.integers(0, 1 << 30, 3_000_000).astype(np.uint32)])
    x = T.angles(q)
    s, c = T.sobol_sincos(q, table)
    np.testing.assert_array_equal(c, np.cos(x))
    np.testing.assert_array_equal(s, np.sin(x))
    frac = np.mean((T.sobol_sincos(q, None)[1] != np.cos(x)))
    assert 1e-4 < frac < 5e-3  # the host libm is not correctly rounded on ~0.15% of angles


def test_planted_table_mistakes_fail(table):
    q, flags = table.words()
    pick = np.arange(0, len(q), max(1, len(q) // 2000))
    x = T.angles(q[pick])
    # a flipped direction bit
    bad = T.CompatTable(table.offsets, table.entries.copy(), table.shift, table.libm, 0, table.lookup_from)
    bad.entries[pick] ^= np.uint16(2 | 8)
    s, c = T.sobol_sincos(q[pick], bad)
    assert np.sum(c != np.cos(x)) + np.sum(s != np.sin(x)) > len(pick) // 2
    # dropped entries (flags cleared)
    bad = T.CompatTable(table.offsets, table.entries.copy(), table.shift, table.libm, 0, table.lookup_from)
    bad.entries[pick] &= np.uint16(0xFFF0)
    s, c = T.sobol_sincos(q[pick], bad)
    assert np.sum(c != np.cos(x)) + np.sum(s != np.sin(x)) >= len(pick)


def test_table_format_guards(tmp_path, table):
    p = tmp_path / "t.bin"
    raw = bytearray(T.TABLE.read_bytes())
    p.write_bytes(bytes(raw[:-2]))
    with pytest.raises(ValueError, match="truncated"):
        T.load_table(p)
    raw[:8] = b"XXXXXXXX"
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="not a trig"):
        T.load_table(p)

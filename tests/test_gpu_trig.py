"""GPU: the device source-angle cos/sin equal NumPy's (the host libm's) bit for
bit -- on every word the compatibility table lists and on random words."""
import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd import trig as T
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def test_compat_table_is_loaded_and_matches_host_libm():
    info = L.trig_compat_info()
    say("trig table", info)
    table = T.load_table()
    assert info["entries"] == len(table.entries) > 0
    assert info["libm"] == table.libm == T.host_libm()


def test_device_sincos_equals_numpy_on_listed_and_random_words():
    table = T.load_table()
    q_listed, _ = table.words()
    rng = np.random.default_rng(21)
    for name, q in (("listed", q_listed), ("random", rng.integers(0, 1 << 30, 1 << 22).astype(np.uint32)),
                    ("quadrants", np.array([0, 1, (1 << 28) - 1, 1 << 28, 1 << 29, 3 << 28, (1 << 30) - 1],
                                           np.uint32))):
        say("device sincos", name, len(q))
        s, c = L.sobol_sincos(q)
        x = T.angles(q)
        np.testing.assert_array_equal(c, np.cos(x), err_msg=name)
        np.testing.assert_array_equal(s, np.sin(x), err_msg=name)
        hs, hc = T.sobol_sincos(q, table)  # host build of the same source
        np.testing.assert_array_equal(c, hc)
        np.testing.assert_array_equal(s, hs)
    with pytest.raises(L.EHSSError, match="30-bit"):
        L.sobol_sincos(np.array([1 << 30], np.uint32))

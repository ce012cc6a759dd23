"""The library's host scramble (ehss_scramble: SeedSequence, PCG64, Lemire
draws, LMS + shift) equals the NumPy restatement, which equals SciPy's points
(tests/test_sources.py, tests/test_oracle_sobol.py); here also directly against
SciPy through the first points of the net."""
import numpy as np
import pytest
from scipy.stats import qmc

from ehss_amd.engine import BITS, scramble, scramble_py

SEEDS = list(range(64)) + [98, 114, 881, 104729, 190003, 209458, 2 ** 31 - 1, 2 ** 32 - 1, 2 ** 32, 2 ** 32 + 1,
                          2 ** 40 + 7, 2 ** 63 + 5, 2 ** 64 - 1]


@pytest.mark.parametrize("seed", SEEDS)
def test_library_scramble_equals_numpy_restatement(seed):
    assert bytes(scramble(seed)) == bytes(scramble_py(seed))


@pytest.mark.parametrize("seed", [0, 1, 104729, 4294967295])
def test_library_scramble_gives_scipys_points(seed):
    s = scramble(seed)
    sv = np.array([[s.sv[d][j] for j in range(BITS)] for d in range(4)], dtype=np.uint64)
    shift = np.array([s.shift[d] for d in range(4)], dtype=np.uint64)
    n = 256
    i = np.arange(n, dtype=np.uint64)
    g = i ^ (i >> np.uint64(1))
    q = np.tile(shift, (n, 1))
    for j in range(8):
        q[((g >> np.uint64(j)) & np.uint64(1)).astype(bool)] ^= sv[:, j]
    np.testing.assert_array_equal(q.astype(np.float64) * 2.0 ** -BITS,
                                  qmc.Sobol(4, scramble=True, seed=seed).random_base2(8))


def test_seed_range():
    with pytest.raises(ValueError):
        scramble(-1)
    with pytest.raises(ValueError):
        scramble(2 ** 64)


def test_scramble_many_equals_one_at_a_time():
    import ctypes as ct
    from ehss_amd.engine import scramble_many
    rng = np.random.default_rng(5)
    seeds = np.concatenate([np.array([0, 1, 2**32 - 1, 2**63, 2**64 - 1], dtype=np.uint64),
                            rng.integers(0, 2**63, 700, dtype=np.int64).astype(np.uint64)])
    many = scramble_many(seeds)
    assert len(many) == len(seeds)
    for s, got in zip(seeds, many):
        assert bytes(got) == bytes(scramble(int(s))), int(s)
    assert len(scramble_many(np.zeros(0, np.uint64))) == 0
    assert ct.sizeof(many) == len(seeds) * ct.sizeof(type(many[0]))

"""Publication protocol logic (three pooled blocks, cap extension) on a fake
scene; the device version is in tests/test_gpu_serve.py."""
import numpy as np
import pytest

from ehss_amd import protocol


class FakeScene:
    """Unresolved weight halves with every cap doubling; Omega depends on seed."""

    def __init__(self, tail_at_512):
        self.tail_at_512 = tail_at_512
        self.calls = []

    def solve_many(self, sobols, power, cap, precision="fp32-fixup", pa=True, continuable=False):
        assert continuable  # the protocol extends its first attempt
        self.calls.append((tuple(sobols), power, cap))
        self.sobols = sobols
        return self._results(cap)

    def extend(self, new_cap, replicates=1):
        assert replicates == len(self.sobols) and new_cap > self.calls[-1][2]
        self.calls.append((tuple(self.sobols), self.calls[-1][1], new_cap))
        return self._results(new_cap)

    def _results(self, cap):
        tail = self.tail_at_512 * 512 / cap
        return [dict(omega=1000.0 + s % 7, pa=900.0, tail_bound=tail, unresolved=int(tail * 10)) for s in self.sobols]


def test_block_seeds_stride_and_wrap():
    assert protocol.block_seeds(5) == [5, 10012, 20019]
    assert protocol.block_seeds(2 ** 32 - 1) == [2 ** 32 - 1, 10006, 20013]


def test_cap_extension_stops_when_bound_is_small():
    sc = FakeScene(tail_at_512=2.0)  # 0.2 % of ~1000 at cap 512, 0.05 % at 2048
    res = protocol.publication_estimate(sc, 11, 12, scramble=lambda s: s)
    assert [a.cap for a in res.attempts] == [512, 1024, 2048]
    assert res.accepted and res.final.tail_bound <= 5e-4 * res.final.omega
    assert all(c[0] == (11, 10018, 20025) and c[1] == 12 for c in sc.calls)  # same rays, extended
    assert res.omega == pytest.approx(np.mean([1000 + 11 % 7, 1000 + 10018 % 7, 1000 + 20025 % 7]))


def test_unaccepted_after_last_cap_and_cap_validation():
    res = protocol.publication_estimate(FakeScene(tail_at_512=100.0), 1, 10, scramble=lambda s: s)
    assert [a.cap for a in res.attempts] == [512, 1024, 2048, 4096] and not res.accepted
    with pytest.raises(ValueError):
        protocol.publication_estimate(FakeScene(1.0), 1, 10, caps=(1024, 512), scramble=lambda s: s)

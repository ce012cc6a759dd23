"""GPU: continuation of unresolved rays at a larger cap (ehss_solve_extend).

Every extension must equal a full retrace of the same rays at the new cap bit
for bit -- totals, histograms, ledgers and per-ray records -- whichever kernel
and schedule made the first solve; misuse is refused.  The publication
protocol's pooled attempt at cap 512 matches the oracle on a small net."""
import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd import protocol
from ehss_amd.engine import GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

PHYSICAL = ("omega", "pa", "pa_from_ehss", "residual_mass", "tail_bound", "weight", "rays", "escaped",
            "unresolved", "pa_hits", "ehss_hits", "max_bounces")


def _same(a, b, key):
    for k in PHYSICAL:
        assert a[k] == b[k], (key, k, a[k], b[k])
    np.testing.assert_array_equal(a["histogram"], b["histogram"], err_msg=str(key))
    np.testing.assert_array_equal(a["ledger"], b["ledger"], err_msg=str(key))


@pytest.mark.parametrize("precision,variant", [("fp32-fixup", 0), ("fp64", 0), ("fp32-fixup", 1)])
def test_extension_equals_full_retrace(precision, variant):
    c, r = scenes.globule(650, 1)
    sob = scramble(21)
    with GPUScene(c, r) as sc, GPUScene(c, r) as fresh:
        first = sc.solve(sob, 14, 1, precision=precision, variant=variant, keep_rays=True, histogram=True,
                         continuable=True)
        assert first["unresolved"] > 1000  # cap 1 leaves many paths open
        for cap in (2, 4, 9, 64, 200):
            say("extend", precision, variant, cap)
            (ext,) = sc.extend(cap, histogram=True)
            ext_rays = sc.last_rays(1 << 14)
            full = fresh.solve(sob, 14, cap, precision=precision, variant=variant, keep_rays=True, histogram=True)
            full_rays = fresh.last_rays(1 << 14)
            _same(full, ext, (precision, variant, cap))
            for k in ("bounces", "escaped", "unresolved", "pa_hit", "contributions"):
                np.testing.assert_array_equal(full_rays[k], ext_rays[k], err_msg=k)


@pytest.mark.parametrize("reps,power", [(3, 12), (20, 11)])  # 20 x 2^11: replicates scheduled across lanes
def test_multi_replicate_extension_equals_retrace(reps, power):
    c, r = scenes.globule(300, 4)
    sobols = [scramble(500 + k) for k in range(reps)]
    with GPUScene(c, r) as sc, GPUScene(c, r) as fresh:
        sc.solve_many(sobols, power, 2, histogram=True, continuable=True)
        for cap in (3, 8, 64):
            say("extend many", reps, power, cap)
            ext = sc.extend(cap, replicates=reps, histogram=True)
            full = fresh.solve_many(sobols, power, cap, histogram=True)
            for k in range(reps):
                _same(full[k], ext[k], (reps, cap, k))


def test_extension_misuse_is_refused():
    c, r = scenes.c60()
    with GPUScene(c, r) as sc:
        with pytest.raises(L.EHSSError, match="no continuable solve"):
            sc.extend(8)
        sc.solve(scramble(1), 10, 2, continuable=True)
        with pytest.raises(L.EHSSError, match="new cap"):
            sc.extend(2)
        sc.extend(4)
        sc.solve(scramble(2), 10, 2)  # an ordinary solve overwrites the records
        with pytest.raises(L.EHSSError, match="no continuable solve"):
            sc.extend(8)
        sc.solve(scramble(3), 10, 2, continuable=True)
        sc.rebuild(*scenes.globule(120, 3))
        with pytest.raises(L.EHSSError, match="no continuable solve"):
            sc.extend(8)


def test_pooled_cap_ladder_attempt_matches_oracle(oracle):
    """A small net, pooled three blocks at cap 512 and its extensions (C3's
    2000-atom ellipsoid traps long paths), against the oracle's retraces."""
    c, r = scenes.globule(2000, 2, axes=(1.0, 1.0, 1.4))
    with GPUScene(c, r) as sc:
        ladder = protocol.publication_estimate(sc, 77, 10, caps=(2, 4, 16, 512), tail_fraction=0.0)
        at512 = protocol.publication_estimate(sc, 78, 10, caps=(512,))
    assert [a.cap for a in ladder.attempts][:3] == [2, 4, 16] and ladder.attempts[0].unresolved > 0
    tree = oracle.static_tree(c, r, 4)
    for seed, att in [(77, a) for a in ladder.attempts] + [(78, at512.attempts[0])]:
        say("attempt", att.cap, att.omega, att.unresolved)
        ref = []
        for sd in protocol.block_seeds(seed):
            src = oracle.external_source(c, r, 10, sd)
            tr = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=att.cap, record=False, threads=16)
            ref.append(oracle.readout(src, tr, att.cap)["omega"])
        np.testing.assert_allclose(att.block_omegas, ref, rtol=1e-12)

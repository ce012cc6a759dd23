"""Asynchronous solves and the request loop on the device: overlapped solves
equal synchronous ones, every request's result is the same alone and in a
mix, and the admission footprint bounds the device memory a request holds."""
import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd import molecules, serve
from ehss_amd.engine import GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

KEYS = ("omega", "pa", "pa_from_ehss", "escaped", "unresolved", "pa_hits", "ehss_hits", "max_bounces", "rays")


def test_async_solves_overlap_and_equal_synchronous():
    items = [scenes.globule(650, 1), scenes.SCENES["c60"](), scenes.globule(300, 4)]
    sc = [GPUScene(c, r) for c, r in items]
    try:
        sync = [s.solve(scramble(5 + k), 14, 64, histogram=True) for k, s in enumerate(sc)]
        say("async enqueue")
        for k, s in enumerate(sc):
            s.solve_async(scramble(5 + k), 14, 64)
        with pytest.raises(L.EHSSError, match="pending"):
            sc[0].solve_async(scramble(1), 10, 64)
        with pytest.raises(L.EHSSError, match="pending"):
            sc[0].last_rays(1 << 14)
        got = [s.solve_wait(histogram=True) for s in sc]
        for a, b in zip(sync, got):
            for k in KEYS:
                assert a[k] == b[k], k
            np.testing.assert_array_equal(a["histogram"], b["histogram"])
        with pytest.raises(L.EHSSError, match="no solve pending"):
            sc[1].solve_wait()
        assert sc[2].solve(scramble(7), 12, 64)["omega"] > 0  # usable again
    finally:
        for s in sc:
            s.close()


def test_footprint_bounds_device_memory():
    for n, seed in ((120, 3), (650, 1), (2000, 2)):
        c, r = scenes.globule(n, seed)
        with GPUScene(c, r) as s:
            for power, cap in ((10, 64), (14, 128)):
                s.solve(scramble(1), power, cap)
                assert s.device_bytes() <= serve.footprint_bytes(n, power, cap), (n, power, cap)
            s.solve_many([scramble(k) for k in range(24)], 12, 64)
            assert s.device_bytes() <= serve.footprint_bytes(n, 14, 128, 24)


def _bases():
    out = []
    for n, seed in ((150, 21), (400, 22), (900, 23), (2500, 24)):
        xyz, el = molecules.globule(n, seed)
        out.append((xyz, molecules.radii_for(el)))
    return out


def test_results_are_the_same_alone_and_in_the_mix():
    reqs = serve.poisson_stream(_bases(), rate_hz=400.0, count=16, seed=5)
    kw = dict(max_inflight=8, power=11, cap=64, target_rel=0.02, min_reps=4, max_reps=24)
    srv = serve.Server(1 << 34, **kw)
    say("mix of", len(reqs))
    mixed = {r.rid: r for r in srv.run(reqs)}
    stats = serve.latency_summary(mixed.values())
    say("mix", stats, "peak in flight", srv.peak_inflight)
    assert stats["served"] == len(reqs)
    assert srv.peak_inflight >= 2
    for r in mixed.values():
        assert r.error is None and (r.rel_se <= 0.02 or len(r.omegas) == 24)
    # 16 requests through at most 8 live scenes: later requests run on scenes
    # rebuilt in place by the worker threads, and those are compared too.
    assert srv.scenes_reused > 0
    for req in reqs:
        alone = serve.Server(1 << 34, **kw)
        (res,) = alone.run([serve.Request(req.rid, req.centers, req.radii, 0.0, req.seed)])
        assert res.omegas == mixed[req.rid].omegas, req.rid
        assert res.ccs == mixed[req.rid].ccs


@pytest.mark.parametrize("precision,variant", [("fp64", 0), ("fp32-fixup", 2), ("fp32-fixup", 3),
                                               ("fp32-fixup", 4)])
# 20 and 70: small nets scheduled one point of each replicate per lane; 2, 16
# and 64: 2^(6 - log2 reps) points of each replicate per wave from one cell.
@pytest.mark.parametrize("reps", [5, 20, 70, 2, 16, 64])
def test_replicates_in_one_launch_equal_separate_solves(precision, variant, reps):
    for name, (c, r) in (("globule650", scenes.globule(650, 1)), ("c60", scenes.SCENES["c60"]()),
                         ("single", (np.zeros((1, 3)), np.array([3.1])))):
        with GPUScene(c, r) as s:
            for power, cap in ((12, 64), (5, 16)) + (((16, 64),) if reps == 16 and name != "single" else ()):
                sobols = [scramble(300 + k) for k in range(reps)]
                say("replicates", name, precision, variant, power)
                many = s.solve_many(sobols, power, cap, precision=precision, histogram=True, variant=variant,
                                    keep_rays=True)
                first = s.last_rays(1 << power)
                for k, sob in enumerate(sobols):
                    one = s.solve(sob, power, cap, precision=precision, histogram=True, variant=variant,
                                  keep_rays=True)
                    for key in KEYS:
                        assert many[k][key] == one[key], (name, k, key)
                    np.testing.assert_array_equal(many[k]["histogram"], one["histogram"])
                    np.testing.assert_array_equal(many[k]["ledger"], one["ledger"])
                    if k == 0:
                        rays = s.last_rays(1 << power)
                        np.testing.assert_array_equal(first["contributions"], rays["contributions"])
                        np.testing.assert_array_equal(first["bounces"], rays["bounces"])


# 256 and more replicates of small nets: the launch's rays are ordered by the
# 4-D cell of their Sobol points before the packet kernel; 300 is not a whole
# number of waves of replicates, and 2^3 x 4096 leaves most cells empty.
@pytest.mark.parametrize("reps,power", [(256, 11), (300, 8), (4096, 3), (8192, 11)])
def test_sorted_schedule_equals_unsorted_and_separate_solves(reps, power):
    c, r = scenes.globule(650, 1)
    with GPUScene(c, r) as s:
        sobols = [scramble(7000 + k) for k in range(reps)]
        say("sorted schedule", reps, power)
        srt = s.solve_many(sobols, power, 64, histogram=True, keep_rays=True)
        rays_srt = s.last_rays(1 << power)  # the first replicate's records
        uns = s.solve_many(sobols, power, 64, histogram=True, keep_rays=True, unsorted=True)
        rays_uns = s.last_rays(1 << power)
        for k in range(reps):
            for key in KEYS:
                assert srt[k][key] == uns[k][key], (reps, power, k, key)
            np.testing.assert_array_equal(srt[k]["histogram"], uns[k]["histogram"])
            np.testing.assert_array_equal(srt[k]["ledger"], uns[k]["ledger"])
        for f in ("flags", "contributions", "bounces"):
            np.testing.assert_array_equal(rays_srt[f], rays_uns[f], err_msg=f)
        for k in (0, reps // 2, reps - 1):
            one = s.solve(sobols[k], power, 64, histogram=True)
            for key in KEYS:
                assert srt[k][key] == one[key], (reps, power, k, key)
            np.testing.assert_array_equal(srt[k]["histogram"], one["histogram"])


@pytest.mark.parametrize("power,reps", [(12, 20), (15, 16)])
def test_per_replicate_schedule_equals_separate_solves(power, reps):
    """EHSS_FLAG_PER_REPLICATE (and the automatic choice at 2^15 x 16, which is
    per replicate) give the same totals as separate solves."""
    c, r = scenes.globule(650, 1)
    with GPUScene(c, r) as s:
        sobols = [scramble(900 + k) for k in range(reps)]
        say("per-replicate schedule", power, reps)
        forced = s.solve_many(sobols, power, 64, histogram=True, per_replicate=True)
        auto = s.solve_many(sobols, power, 64, histogram=True)
        for k in (0, reps // 2, reps - 1):
            one = s.solve(sobols[k], power, 64, histogram=True)
            for key in KEYS:
                assert forced[k][key] == one[key] == auto[k][key], (power, reps, k, key)
            np.testing.assert_array_equal(forced[k]["histogram"], one["histogram"])


def test_array_totals_and_prebuilt_tables_equal_dicts():
    from ehss_amd.engine import sobol_table
    c, r = scenes.SCENES["c60"]()
    with GPUScene(c, r) as s:
        sobols = [scramble(40 + k) for k in range(300)]
        table = sobol_table(sobols)
        assert sobol_table(table) is table
        say("array totals")
        dicts = s.solve_many(sobols, 11, 64, pa=False)
        arr = s.solve_many(table, 11, 64, pa=False, as_array=True)
        assert len(arr) == 300
        for k in (0, 150, 299):
            for key in KEYS:
                assert arr[key][k] == dicts[k][key], (k, key)
        s.solve_async(table, 11, 64, pa=False)
        again = s.solve_wait(as_array=True)
        np.testing.assert_array_equal(again["omega"], arr["omega"])
        with pytest.raises(ValueError):
            s.solve_many(table, 11, 64, as_array=True, histogram=True)


def test_replicate_launch_async_and_limits():
    c, r = scenes.SCENES["c60"]()
    with GPUScene(c, r) as s:
        sobols = [scramble(k) for k in range(7)]
        s.solve_async(sobols, 11, 64)
        got = s.solve_wait()
        assert len(got) == 7
        assert [g["omega"] for g in got] == [s.solve(x, 11, 64)["omega"] for x in sobols]
        with pytest.raises(L.EHSSError, match="replicate count"):
            s.solve_many([], 10, 64)
        with pytest.raises(L.EHSSError, match="2\\^30"):
            s.solve_many(sobols[:4], 29, 64)


def test_publication_protocol_on_device():
    from ehss_amd import protocol
    c, r = scenes.globule(650, 1)
    with GPUScene(c, r) as s:
        res = protocol.publication_estimate(s, 4242, 12, caps=(1, 2, 4, 64), tail_fraction=1e-3)
        assert [a.cap for a in res.attempts][0] == 1 and len(res.attempts) >= 2
        for att in res.attempts:  # each attempt is the mean of the three blocks solved alone
            alone = [s.solve(scramble(x), 12, att.cap)["omega"] for x in protocol.block_seeds(4242)]
            assert att.block_omegas == alone
            assert att.omega == pytest.approx(np.mean(alone), rel=1e-15)
        tails = [a.tail_bound for a in res.attempts]
        assert all(b <= a for a, b in zip(tails, tails[1:]))


def test_batched_serving_equals_per_request_launches():
    """Batched mode (replicates of every ready request in one retargeted plan
    launch) gives every request the same replicate values as per-request
    launches."""
    reqs = serve.poisson_stream(_bases(), rate_hz=2000.0, count=24, seed=8)
    kw = dict(max_inflight=8, power=11, cap=64, target_rel=0.02, min_reps=4, max_reps=24)
    per = {r.rid: r for r in serve.Server(1 << 34, **kw).run(reqs)}
    srv = serve.Server(1 << 34, batch_plans=2, **kw)
    bat = {r.rid: r for r in srv.run(reqs)}
    say("batched", serve.latency_summary(bat.values()), "batches", srv.batches, "peak jobs", srv.peak_batch_jobs)
    assert srv.batches < sum(r.launches for r in bat.values()) and srv.scenes_reused > 0
    for rid, r in bat.items():
        assert r.error is None and r.omegas == per[rid].omegas, rid

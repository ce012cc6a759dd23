"""Prepared multi-structure plans (ehss_plan_*): many structures of uneven size,
scrambles, ray counts and caps as one launch sequence.  Every job's physical
totals, histogram and ledger must equal the same solve run alone, bit for bit,
in every kernel variant and drain schedule, with and without hipGraph replay."""
import numpy as np
import pytest

from ehss_amd.engine import GPUPlan, GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

# Totals that are physical results (work counters and the kernel-specific
# grazing diagnostic excluded).
PHYSICAL = ("omega", "pa", "pa_from_ehss", "residual_mass", "tail_bound", "weight", "rays", "escaped",
            "unresolved", "pa_hits", "ehss_hits", "max_bounces")


def _jobs():
    """(scene arrays, power, cap): uneven sizes, one job smaller than a wave."""
    return [
        (scenes.globule(650, 1), 14, 64),
        (scenes.SCENES["c60"](), 12, 64),
        (scenes.SCENES["f29"](), 10, 32),
        (scenes.SCENES["cross6"](), 8, 16),
        ((np.zeros((1, 3)), np.array([3.1])), 5, 8),
        (scenes.globule(300, 4), 13, 128),
    ]


def _same(alone, planned, key):
    for k in PHYSICAL:
        assert alone[k] == planned[k], (key, k, alone[k], planned[k])
    np.testing.assert_array_equal(alone["histogram"], planned["histogram"], err_msg=str(key))
    np.testing.assert_array_equal(alone["ledger"], planned["ledger"], err_msg=str(key))


CASES = [
    ("fp32-fixup", 4, None, None, False),  # library defaults: one drain, whole waves per ray at these sizes
    ("fp32-fixup", 4, 4, 2, False),        # two-stage drain: 4 lanes per ray, spill after 2 queries
    ("fp32-fixup", 4, 1, None, False),     # per-lane drain with cooperative tail
    ("fp32-fixup", 4, None, None, True),   # packet kernels on LDS-staged scenes
    ("fp32-fixup", 3, None, None, False),
    ("fp32-fixup", 2, None, None, True),
    ("fp32-fixup", 1, None, None, False),
    ("fp64", 0, None, None, False),
]


@pytest.mark.parametrize("precision,variant,group,stage1,lds", CASES)
def test_plan_equals_solves_alone(precision, variant, group, stage1, lds):
    jobs = _jobs()
    sc = [GPUScene(c, r) for (c, r), _, _ in jobs]
    sobols = [scramble(11 + j) for j in range(len(jobs))]
    powers = [p for _, p, _ in jobs]
    caps = [c for _, _, c in jobs]
    key = (precision, variant, group, stage1, lds)
    try:
        say("plan alone", *key)
        alone = [s.solve(sob, p, c, precision=precision, pa=True, histogram=True, variant=variant)
                 for s, sob, p, c in zip(sc, sobols, powers, caps)]
        say("plan", *key)
        with GPUPlan(sc, sobols, powers, caps, precision=precision, pa=True, variant=variant, group=group,
                     stage1=stage1, lds=lds) as plan:
            got = plan.run(histogram=True)
            again = plan.run(histogram=True)
        for j, (a, g, h) in enumerate(zip(alone, got, again)):
            _same(a, g, key + (j,))
            _same(a, h, key + (j, "relaunch"))
            assert g["histogram"].sum() == 1 << powers[j]
    finally:
        for s in sc:
            s.close()


def test_plan_default_two_stage_drain_at_size():
    """2^18 rays and more in one plan select the two-stage drain by default."""
    jobs = [(scenes.globule(650, 1), 17, 64), (scenes.globule(650, 2), 17, 64), (scenes.SCENES["c60"](), 12, 64)]
    sc = [GPUScene(c, r) for (c, r), _, _ in jobs]
    sobols = [scramble(7 + j) for j in range(len(jobs))]
    powers = [p for _, p, _ in jobs]
    caps = [c for _, _, c in jobs]
    try:
        say("plan two-stage alone")
        alone = [s.solve(sob, p, c, pa=True, histogram=True) for s, sob, p, c in zip(sc, sobols, powers, caps)]
        say("plan two-stage")
        with GPUPlan(sc, sobols, powers, caps, pa=True) as plan:
            got = plan.run(histogram=True)
        for j, (a, g) in enumerate(zip(alone, got)):
            _same(a, g, ("two-stage", j))
    finally:
        for s in sc:
            s.close()


def test_plan_graph_replays_follow_new_scrambles():
    """Capture once, replay with a different scramble per launch; eager solves
    and host synchronisation between replays; every replay equals the solves
    alone for its scrambles."""
    jobs = _jobs()[:5]
    sc = [GPUScene(c, r) for (c, r), _, _ in jobs]
    powers = [p for _, p, _ in jobs]
    caps = [c for _, _, c in jobs]
    try:
        with GPUPlan(sc, [scramble(100 + j) for j in range(len(jobs))], powers, caps, graph=True) as plan:
            for rep in range(4):
                sobols = [scramble(1000 * rep + j) for j in range(len(jobs))]
                for j, s in enumerate(sobols):
                    plan.set_sobol(j, s)
                say("plan graph replay", rep)
                plan.launch()
                got = plan.wait(histogram=True)
                alone = [s.solve(sob, p, c, pa=True, histogram=True)
                         for s, sob, p, c in zip(sc, sobols, powers, caps)]
                for j, (a, g) in enumerate(zip(alone, got)):
                    _same(a, g, ("graph", rep, j))
    finally:
        for s in sc:
            s.close()


def test_plan_async_query_and_argument_checks():
    c, r = scenes.SCENES["c60"]()
    with GPUScene(c, r) as s:
        say("plan async")
        with GPUPlan([s, s], [scramble(1), scramble(2)], [12, 12], [64, 64]) as plan:
            plan.launch()
            res = plan.wait()
            assert plan.done()
            assert res[0]["omega"] != res[1]["omega"]
            assert res[0]["omega"] == s.solve(scramble(1), 12, 64)["omega"]
            assert res[1]["omega"] == s.solve(scramble(2), 12, 64)["omega"]
        with pytest.raises(Exception, match="power"):
            GPUPlan([s], [scramble(1)], [27], [64])
        with pytest.raises(Exception, match="cap"):
            GPUPlan([s], [scramble(1)], [10], [0])
        with pytest.raises(ValueError):
            GPUPlan([s], [scramble(1)], [10, 11], [64])
        with pytest.raises(Exception, match="2\\^30"):
            GPUPlan([s] * 17, [scramble(1)] * 17, [26] * 17, [64] * 17)


def test_autotuned_plan_matches_solves_alone(tmp_path):
    """The tuner times real plans (its fingerprint check asserts that every
    configuration gives the same totals) and the winner equals the solves alone."""
    from ehss_amd import autotune

    structs = [scenes.globule(400, 5), scenes.SCENES["c60"](), scenes.globule(900, 6)]
    sobols = [scramble(50 + j) for j in range(len(structs))]
    space = {"group": (64, 4), "stage1": (2, 0), "coop": (8,), "lds": (False, True), "leaf_size": (4, 8)}
    say("autotune plan")
    res, scs, kw = autotune.tune_plan(structs, 14, 64, sobols, space=space, repeats=1, passes=1,
                                      cache_path=tmp_path / "tune.json")
    try:
        assert res.config["leaf_size"] in (4, 8) and len(res.trials) >= 4
        with GPUPlan(scs, sobols, [14] * 3, [64] * 3, **kw) as plan:
            got = plan.run(histogram=True)
        for j, (s, sob) in enumerate(zip(scs, sobols)):
            _same(s.solve(sob, 14, 64, pa=True, histogram=True), got[j], ("tuned", j))
        again, _, _ = autotune.tune_plan(structs, 14, 64, sobols, space=space, cache_path=tmp_path / "tune.json")
        assert again.cached and again.config == res.config
    finally:
        for s in scs:
            s.close()


def test_two_plans_in_flight_over_shared_scenes():
    """Two plans over the same resident scenes, launched together on their own
    streams, give the results of each run alone."""
    jobs = _jobs()[:4]
    sc = [GPUScene(c, r) for (c, r), _, _ in jobs]
    powers = [p for _, p, _ in jobs]
    caps = [c for _, _, c in jobs]
    try:
        a_sob = [scramble(500 + j) for j in range(len(jobs))]
        b_sob = [scramble(600 + j) for j in range(len(jobs))]
        with GPUPlan(sc, a_sob, powers, caps, graph=True) as a, GPUPlan(sc, b_sob, powers, caps) as b:
            alone_a, alone_b = a.run(histogram=True), b.run(histogram=True)
            say("two plans in flight")
            a.launch()
            b.launch()
            got_b, got_a = b.wait(histogram=True), a.wait(histogram=True)
        for j in range(len(jobs)):
            _same(alone_a[j], got_a[j], ("a", j))
            _same(alone_b[j], got_b[j], ("b", j))
    finally:
        for s in sc:
            s.close()


def test_plan_holds_its_scenes():
    """A plan copies its scenes' views: rebuilding or destroying a held scene is
    refused (EHSS_ERR_BUSY) until the plan is destroyed, then allowed."""
    from ehss_amd import _lib as L
    a, b = scenes.globule(300, 4), scenes.globule(650, 1)
    sc = GPUScene(*a)
    other = GPUScene(*b)
    try:
        plan = GPUPlan([sc, other, sc], [scramble(1), scramble(2), scramble(3)], [10] * 3, [64] * 3)
        before = plan.run()
        for fn in (lambda: sc.rebuild(*b), lambda: other.rebuild(*a)):
            with pytest.raises(L.EHSSError, match="held by"):
                fn()
        with pytest.raises(L.EHSSError, match="held by 2 plan job"):
            sc.close()
        assert plan.run()[0]["omega"] == before[0]["omega"]  # the plan still works, unchanged
        plan.close()
        sc.rebuild(*b)
        assert sc.info["atoms"] == len(b[1])
    finally:
        sc.close()
        other.close()


def test_retargeted_plan_equals_solves_alone_and_releases_scenes():
    """ehss_plan_retarget: the same plan object over other job lists (more and
    fewer jobs, other scenes) gives each job its solve-alone totals; an empty
    list releases the scenes for rebuild."""
    from ehss_amd import _lib as L
    structs = [scenes.globule(300, 4), scenes.globule(650, 1), scenes.SCENES["c60"]()]
    scs = [GPUScene(c, r) for c, r in structs]
    try:
        plan = GPUPlan([scs[0]], [scramble(1)], [12], [64])
        plan.run()
        for k, jobs in enumerate(([(0, 12, 64), (1, 14, 32), (2, 10, 64)], [(1, 11, 64)],
                                  [(2, 12, 16), (2, 12, 64), (0, 13, 64), (1, 9, 8)])):
            say("retarget", k, jobs)
            sob = [scramble(100 * k + j) for j in range(len(jobs))]
            plan.retarget([scs[i] for i, _, _ in jobs], sob, [p for _, p, _ in jobs], [c for _, _, c in jobs])
            got = plan.run(histogram=True)
            for (i, p, c), sb, g in zip(jobs, sob, got):
                alone = scs[i].solve(sb, p, c, histogram=True)
                for key in PHYSICAL:
                    assert alone[key] == g[key], (k, i, key)
                np.testing.assert_array_equal(alone["histogram"], g["histogram"])
        with pytest.raises(L.EHSSError, match="held by"):
            scs[1].rebuild(*structs[0])
        plan.release()
        scs[1].rebuild(*structs[0])
        with pytest.raises(L.EHSSError, match="no jobs"):
            plan.launch()
        plan.close()
    finally:
        for s in scs:
            s.close()

"""Request loop logic (admission by device-memory footprint, continuous batching,
adaptive replicate count, latency accounting) on a simulated device: a fake
scene whose solves take virtual time and return seeded values."""
import math

import numpy as np
import pytest

from ehss_amd import serve


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-6)


class FakeScene:
    """Solve time proportional to atoms; Omega = pi R^2 (1 + 0.03 z(seed))."""
    live = 0
    held = 0  # plans holding this scene (batched mode); closing a held scene is an error

    def __init__(self, clock, centers, radii):
        if len(radii) == 0:
            raise ValueError("Expected nonempty centers (N,3) and radii (N,).")
        self.clock, self.n = clock, len(radii)
        self.base = math.pi * (1.0 + float(np.max(np.linalg.norm(centers, axis=1)) + np.max(radii))) ** 2
        self.busy_until, self.seed, self.closed = None, None, False
        FakeScene.live += 1

    def solve_async(self, sobols, power, cap, precision="fp32-fixup", pa=True):
        assert self.busy_until is None, "one pending launch per scene"
        assert isinstance(sobols, list) and sobols
        self.seeds, self.rays = list(sobols), 1 << power
        self.busy_until = self.clock() + 1e-6 * self.n * (1 + 0.1 * len(sobols))

    def solve_done(self):
        return self.busy_until is not None and self.clock() >= self.busy_until

    def solve_wait(self):
        assert self.solve_done()
        self.busy_until = None
        out = []
        for seed in self.seeds:
            z = np.random.default_rng(seed).normal()
            out.append(dict(omega=self.base * (1 + 0.03 * z), pa=self.base, unresolved=0, rays=self.rays))
        return out

    def close(self):
        assert not self.closed and self.held == 0
        self.closed = True
        FakeScene.live -= 1


def make_server(clock, budget=1 << 40, **kw):
    return serve.Server(budget, scene_factory=lambda c, r: FakeScene(clock, c, r), scramble=lambda s: s,
                        replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, **kw)


def stream(count=40, rate=2000.0, seed=3):
    rng = np.random.default_rng(11)
    bases = []
    for n in (100, 400, 1500, 6000):
        c = rng.normal(scale=n ** (1 / 3), size=(n, 3))
        bases.append((c, np.full(n, 1.7)))
    return serve.poisson_stream(bases, rate, count, seed=seed)


def test_every_request_served_to_target_and_results_match_alone():
    reqs = stream()
    clock = Clock()
    srv = make_server(clock, max_inflight=8, power=10, target_rel=0.01, min_reps=4, max_reps=200)
    mixed = {r.rid: r for r in srv.run(reqs, clock=clock, sleep=clock.sleep)}
    assert len(mixed) == len(reqs) and FakeScene.live == 0
    assert srv.peak_inflight > 1  # solves of several requests were in flight together
    for r in mixed.values():
        assert r.error is None
        assert r.rel_se <= 0.01 and len(r.omegas) >= 4
        assert r.finished >= r.admitted >= r.arrival
        assert r.launches <= 4  # replicates are launched together, not one by one
    for req in reqs[:6]:
        c = Clock()
        alone = make_server(c, max_inflight=8, power=10, target_rel=0.01, min_reps=4, max_reps=200)
        (res,) = alone.run([serve.Request(req.rid, req.centers, req.radii, 0.0, req.seed)], clock=c, sleep=c.sleep)
        assert res.omegas == mixed[req.rid].omegas and res.ccs == mixed[req.rid].ccs


def test_admission_respects_the_device_budget():
    reqs = stream(count=30, rate=1e6)  # all arrive at once
    fps = sorted(serve.footprint_bytes(len(r.radii), 10, 64, 64) for r in reqs)
    budget = fps[-1] + fps[-2]  # room for the two largest at most
    clock = Clock()
    srv = make_server(clock, budget=budget, max_inflight=16, power=10)
    res = srv.run(reqs, clock=clock, sleep=clock.sleep)
    assert all(r.error is None for r in res)
    assert srv.peak_bytes <= budget
    assert srv.peak_inflight >= 2


def test_oversized_request_is_rejected_and_the_rest_served():
    reqs = stream(count=10)
    big = serve.Request(99, np.zeros((50000, 3)), np.ones(50000), arrival=0.0, seed=1)
    clock = Clock()
    budget = max(serve.footprint_bytes(len(r.radii), 10, 64, 64) for r in reqs) * 2
    srv = make_server(clock, budget=budget, power=10)
    res = {r.rid: r for r in srv.run(reqs + [big], clock=clock, sleep=clock.sleep)}
    assert res[99].error and "budget" in res[99].error
    assert all(res[r.rid].error is None for r in reqs)


def test_invalid_structure_reports_an_error():
    clock = Clock()
    srv = make_server(clock, power=10)
    (res,) = srv.run([serve.Request(0, np.zeros((0, 3)), np.zeros(0))], clock=clock, sleep=clock.sleep)
    assert res.error and "nonempty" in res.error


def test_max_reps_caps_a_noisy_request():
    clock = Clock()
    srv = make_server(clock, power=10, target_rel=1e-6, min_reps=3, max_reps=7)
    (res,) = srv.run(stream(count=1), clock=clock, sleep=clock.sleep)
    assert len(res.omegas) == 7 and res.rel_se > 1e-6


def test_latency_summary_and_stream():
    reqs = stream(count=50, rate=100.0)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    clock = Clock()
    res = make_server(clock, power=10).run(reqs, clock=clock, sleep=clock.sleep)
    s = serve.latency_summary(res)
    assert s["served"] == 50 and s["p50_s"] <= s["p99_s"] and s["ccs_per_s"] > 0


def test_limits_are_validated():
    with pytest.raises(ValueError):
        serve.Server(0, scene_factory=object, scramble=object, replicate_seed=object)
    with pytest.raises(ValueError):
        serve.Server(1 << 30, min_reps=5, max_reps=4, scene_factory=object, scramble=object,
                     replicate_seed=object)


class ReusableScene(FakeScene):
    """FakeScene with in-place rebuild (ehss_scene_rebuild) and a byte count."""
    rebuilds = 0

    def rebuild(self, centers, radii):
        assert self.busy_until is None and not self.closed, "rebuild of a busy or closed scene"
        if len(radii) == 0:
            raise ValueError("Expected nonempty centers (N,3) and radii (N,).")
        self.n = max(self.n, len(radii))
        self.base = math.pi * (1.0 + float(np.max(np.linalg.norm(centers, axis=1)) + np.max(radii))) ** 2
        ReusableScene.rebuilds += 1

    def device_bytes(self):
        return serve.footprint_bytes(self.n, 10, 64, 64)


def test_finished_scenes_are_rebuilt_for_later_requests_with_the_same_results():
    reqs = stream(count=40)
    clock = Clock()
    plain = {r.rid: r for r in make_server(clock, max_inflight=4, power=10).run(reqs, clock=clock, sleep=clock.sleep)}
    clock = Clock()
    ReusableScene.rebuilds = 0
    srv = serve.Server(1 << 40, scene_factory=lambda c, r: ReusableScene(clock, c, r), scramble=lambda s: s,
                       replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=4, power=10)
    got = {r.rid: r for r in srv.run(reqs, clock=clock, sleep=clock.sleep)}
    assert FakeScene.live == 0
    assert srv.scenes_reused > 0 and srv.scenes_built <= 4 and ReusableScene.rebuilds == srv.scenes_reused
    for rid, r in plain.items():
        assert got[rid].omegas == r.omegas and got[rid].ccs == r.ccs


def test_idle_scenes_give_way_to_the_budget():
    reqs = stream(count=30, rate=1e6)
    fps = sorted(serve.footprint_bytes(len(r.radii), 10, 64, 64) for r in reqs)
    budget = fps[-1] + fps[-2]
    clock = Clock()
    srv = serve.Server(budget, scene_factory=lambda c, r: ReusableScene(clock, c, r), scramble=lambda s: s,
                       replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=8, power=10)
    res = srv.run(reqs, clock=clock, sleep=clock.sleep)
    assert all(r.error is None for r in res) and len(res) == 30
    assert srv.peak_bytes <= budget and FakeScene.live == 0


def test_invalid_structure_with_an_idle_scene_keeps_serving():
    clock = Clock()
    srv = serve.Server(1 << 40, scene_factory=lambda c, r: ReusableScene(clock, c, r), scramble=lambda s: s,
                       replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=2, power=10)
    good = stream(count=3, rate=10.0)
    bad = serve.Request(77, np.zeros((0, 3)), np.zeros(0), arrival=good[-1].arrival + 1.0)
    after = serve.Request(78, good[0].centers, good[0].radii, arrival=bad.arrival + 0.1, seed=5)
    res = {r.rid: r for r in srv.run(good + [bad, after], clock=clock, sleep=clock.sleep)}
    assert res[77].error and "nonempty" in res[77].error
    assert res[78].error is None and FakeScene.live == 0


def test_rebuilds_on_worker_threads_and_in_the_loop_give_the_same_results():
    """rebuild_workers > 0 rebuilds reused scenes on threads, 0 in the loop; a
    request's results do not depend on which."""
    reqs = stream(count=40)
    got = {}
    for workers in (0, 3):
        clock = Clock()
        ReusableScene.rebuilds = 0
        srv = serve.Server(1 << 40, scene_factory=lambda c, r: ReusableScene(clock, c, r), scramble=lambda s: s,
                           replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=4, power=10,
                           rebuild_workers=workers)
        got[workers] = {r.rid: r for r in srv.run(reqs, clock=clock, sleep=clock.sleep)}
        assert srv.scenes_reused > 0 and ReusableScene.rebuilds == srv.scenes_reused and FakeScene.live == 0
    for rid, r in got[0].items():
        assert r.error is None and got[3][rid].omegas == r.omegas and got[3][rid].ccs == r.ccs
    with pytest.raises(ValueError):
        serve.Server(1 << 30, rebuild_workers=-1, scene_factory=object, scramble=object, replicate_seed=object)


def test_a_failed_launch_is_raised_not_reported_as_a_bad_structure():
    class FailingScene(ReusableScene):
        def solve_async(self, *a, **k):
            raise RuntimeError("device failure")

    before = FakeScene.live
    for workers in (0, 2):
        clock = Clock()
        srv = serve.Server(1 << 40, scene_factory=lambda c, r: FailingScene(clock, c, r), scramble=lambda s: s,
                           replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=2, power=10,
                           rebuild_workers=workers)
        with pytest.raises(RuntimeError, match="device failure"):
            srv.run(stream(count=3, rate=10.0), clock=clock, sleep=clock.sleep)
    FakeScene.live = before  # the aborted runs leave their scenes to the caller


class FakePlan:
    """A multi-structure launch on the simulated device: jobs of several scenes
    run together in less time than one launch each; it holds its scenes."""

    made = 0

    def __init__(self, clock, scenes, sobols, powers, caps):
        self.clock, self.scenes, self.busy_until = clock, [], None
        FakePlan.made += 1
        self.retarget(scenes, sobols, powers, caps)

    def retarget(self, scenes, sobols, powers, caps):
        assert self.busy_until is None, "retarget waits for the last launch"
        for s in self.scenes:
            s.held -= 1
        self.scenes, self.seeds, self.powers = list(scenes), list(sobols), list(powers)
        for s in self.scenes:
            assert not s.closed
            s.held += 1

    def launch(self):
        assert self.scenes and self.busy_until is None
        self.busy_until = self.clock() + 2e-7 * sum(s.n for s in set(self.scenes)) + 1e-6 * len(self.scenes)

    def done(self):
        return self.busy_until is not None and self.clock() >= self.busy_until

    def wait(self):
        assert self.done()
        self.busy_until = None
        return [dict(omega=s.base * (1 + 0.03 * np.random.default_rng(sd).normal()), pa=s.base, unresolved=0,
                     rays=1 << pw) for s, sd, pw in zip(self.scenes, self.seeds, self.powers)]

    def release(self):
        self.retarget([], [], [], [])

    def close(self):
        assert not self.scenes


def test_batched_mode_gives_the_same_results_as_per_request_launches():
    reqs = stream(count=60, rate=5000.0)
    kw = dict(max_inflight=12, power=10, target_rel=0.01, min_reps=4, max_reps=200)
    clock = Clock()
    FakeScene.held = 0
    alone = {r.rid: r for r in make_server(clock, **kw).run(reqs, clock=clock, sleep=clock.sleep)}
    clock = Clock()
    srv = make_server(clock, batch_plans=2, plan_factory=lambda *a: FakePlan(clock, *a), **kw)
    mixed = {r.rid: r for r in srv.run(reqs, clock=clock, sleep=clock.sleep)}
    assert len(mixed) == len(reqs) and FakeScene.live == 0
    assert srv.batches < sum(r.launches for r in mixed.values())  # requests shared launches
    assert srv.peak_batch_jobs > srv.min_reps
    for rid, r in mixed.items():
        assert r.error is None and r.omegas == alone[rid].omegas and r.ccs == alone[rid].ccs


def test_batched_mode_limits_jobs_per_launch():
    reqs = stream(count=30, rate=1e5)
    clock = Clock()
    FakeScene.held = 0
    srv = make_server(clock, max_inflight=30, power=10, target_rel=0.01, min_reps=4, max_reps=8, batch_plans=1,
                      plan_factory=lambda *a: FakePlan(clock, *a), max_batch_jobs=16)
    res = srv.run(reqs, clock=clock, sleep=clock.sleep)
    assert len(res) == 30 and srv.peak_batch_jobs <= 16 and FakeScene.live == 0
    with pytest.raises(ValueError):
        make_server(clock, max_reps=64, batch_plans=1, max_batch_jobs=32)


def test_warm_pool_serves_every_request_without_growing_a_scene():
    """warm_atoms: the pool is made before the stream (its clock starts after);
    every request reuses a pool scene already large enough, with the results
    of a cold server."""
    reqs = stream(count=40)
    clock = Clock()
    cold = {r.rid: r for r in make_server(clock, max_inflight=4, power=10).run(reqs, clock=clock, sleep=clock.sleep)}
    grown = []

    class Tracked(ReusableScene):
        def rebuild(self, centers, radii):
            if len(radii) > self.n:
                grown.append(len(radii))
            super().rebuild(centers, radii)

    clock = Clock()
    srv = serve.Server(1 << 40, scene_factory=lambda c, r: Tracked(clock, c, r), scramble=lambda s: s,
                       replicate_seed=lambda s, k: (s + 104729 * k) % 2 ** 32, max_inflight=4, power=10,
                       warm_atoms=6000)
    warm = {r.rid: r for r in srv.run(reqs, clock=clock, sleep=clock.sleep)}
    assert srv.scenes_built == 4 and srv.scenes_reused == len(reqs) and not grown and FakeScene.live == 0
    for rid, r in cold.items():
        assert warm[rid].omegas == r.omegas and warm[rid].ccs == r.ccs
    with pytest.raises(ValueError):
        serve.Server(1 << 30, warm_atoms=-1, scene_factory=object, scramble=object, replicate_seed=object)


def test_best_fit_reuse():
    idle = [(500, "a"), (100, "b"), (300, "c"), (900, "d")]
    assert idle[serve._best_fit(idle, 250)][1] == "c"   # the smallest that holds it
    assert idle[serve._best_fit(idle, 100)][1] == "b"
    assert idle[serve._best_fit(idle, 1000)][1] == "d"  # none does: the largest grows least

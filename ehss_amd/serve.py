"""Request loop for CCS serving on one GPU.

Each request is a structure (centres, radii).  The loop

* admits a request when it has arrived, a slot is free and its device-memory
  footprint (scene trees + solve workspace) fits the budget left -- the scene
  and its trees are built on the device at admission and stay resident until
  the request completes;
* keeps every admitted request's next solve in flight: either on the
  request's own stream (ehss_solve_async), so solves of different structures
  and sizes overlap on the device, or -- batched mode -- gathered with the next
  replicates of every other ready request into one multi-structure launch (a
  retargeted plan), a few batches in flight; either way a new request joins as
  soon as it is admitted (continuous batching, not a queue of whole jobs);
* chooses the ray count per request adaptively: independent scrambled-Sobol
  replicates of 2^power rays (seeds replicate_seed(seed, k), the reference's
  per-replicate rule) until the relative standard error of the mean CCS is at
  most target_rel (and at least min_reps replicates), capped at max_reps.  The
  first launch of a request runs min_reps replicates together; each later one
  runs as many as the spread so far predicts are still needed, so a request
  takes one or two launches rather than one per replicate;
* keeps the scenes of finished requests idle (up to max_inflight scenes in all)
  and rebuilds an idle one in place for the next request (ehss_scene_rebuild:
  the device arrays only grow), so that a request costs no device allocation
  or free -- a free synchronises the whole device.  The rebuilds run on worker
  threads (the library call releases the interpreter lock and each scene has
  its own stream), so the loop keeps launching and collecting meanwhile; a
  request's first launch follows its rebuild;
* records arrival, admission and completion times, from which p50/p99 latency
  and CCS/s follow.

A request's result depends only on its structure, seed and solve settings: the
kernels and the reduction are deterministic, so the same request gives the same
bits alone and in any mix.
"""
from __future__ import annotations

import collections
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

# Device bytes per atom of a scene (ehss_scene_create): centres 24, radii 8,
# FP64 spheres 32, groups 4, offsets 4, slot ids 4, FP64 tree <= 2 x 64, PA tree
# <= 2 x 64, PA ids 4, FP32 tree <= 2 x 32, FP32 spheres 16, slot-ordered FP64
# spheres 32, leaf boxes <= 32.
SCENE_BYTES_PER_ATOM = 24 + 8 + 32 + 4 + 4 + 4 + 128 + 128 + 4 + 64 + 16 + 32 + 32
SCENE_BYTES_FIXED = 64 * 2 + 4 * 8 * 64 + 4096  # PA root slack, frontier, small buffers
# Solve workspace per ray: record word 4, contribution 8, two queue entries 112; per
# replicate: scramble + ray map 660, counters 128, drain counter stripes 2048,
# histogram + ledger bins 16.
SOLVE_BYTES_PER_RAY = 4 + 8 + 2 * 112
SOLVE_BYTES_PER_REPLICATE = 660 + 128 + 2048
SOLVE_BYTES_FIXED = 64 * 1024


def footprint_bytes(n_atoms: int, power: int, cap: int, replicates: int = 1) -> int:
    """Upper bound on the device memory a scene holds with launches of up to
    `replicates` solves of 2^power rays."""
    rays = 1 << int(power)
    reduce_blocks = rays // 2048 + 1
    reps = int(replicates)
    return (SCENE_BYTES_PER_ATOM * int(n_atoms) + SCENE_BYTES_FIXED
            + reps * (SOLVE_BYTES_PER_RAY * rays + 64 * reduce_blocks + 16 * (int(cap) + 1)
                      + SOLVE_BYTES_PER_REPLICATE) + SOLVE_BYTES_FIXED)


@dataclass
class Request:
    rid: int
    centers: np.ndarray
    radii: np.ndarray
    arrival: float = 0.0  # seconds after the start of the stream
    seed: int = 0


@dataclass
class Result:
    rid: int
    atoms: int
    ccs: float = math.nan          # mean Omega over replicates (A^2)
    rel_se: float = math.nan       # standard error of the mean / mean
    omegas: list = field(default_factory=list)
    pa: list = field(default_factory=list)
    unresolved: int = 0
    rays: int = 0
    launches: int = 0
    arrival: float = 0.0
    admitted: float = math.nan
    finished: float = math.nan
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Live:
    req: Request
    res: Result
    scene: object
    footprint: int
    rebuild: object = None  # pending in-place rebuild (a Future), if any
    held: int = 0           # device bytes the reused scene held when idle


def _best_fit(idle, need: int) -> int:
    """Index of the idle (bytes, scene) to reuse for a request of `need` bytes:
    the smallest that holds it, else the largest (the one that grows least)."""
    fits = [k for k, (held, _) in enumerate(idle) if held >= need]
    if fits:
        return min(fits, key=lambda k: idle[k][0])
    return max(range(len(idle)), key=lambda k: idle[k][0])


def _rel_se(values) -> float:
    v = np.asarray(values, dtype=np.float64)
    if len(v) < 2 or v.mean() == 0:
        return math.inf
    return float(v.std(ddof=1) / math.sqrt(len(v)) / abs(v.mean()))


class Server:
    """Continuous-batching CCS server over device-resident scenes.

    scene_factory(centers, radii) must return an object with the GPUScene
    asynchronous interface (solve_async, solve_done, solve_wait, close); when
    it also has rebuild(centers, radii) and device_bytes(), finished scenes are
    reused.  scramble(seed) and replicate_seed(seed, k) give the per-replicate
    nets.  rebuild_workers: threads that rebuild reused scenes (0: in the loop
    itself).

    batch_plans > 0 switches to batched launches: the next replicates of every
    request that is ready are gathered into ONE multi-structure launch (a
    prepared plan, retargeted for each batch: ehss_plan_retarget), with up to
    batch_plans batches in flight.  One request's solve alone is a few dozen
    workgroups; a batch of many requests fills the device.  A job's totals in a
    plan are bit-identical to the same solve alone, so results do not depend on
    the mode or on the mix.  plan_factory(scenes, sobols, powers, caps) makes a
    plan (default: GPUPlan with the server's precision and PA setting).

    warm_atoms > 0 makes the scene pool before the stream starts: max_inflight
    scenes built for a lattice of warm_atoms atoms and solved once with
    max_reps replicates, so that their grow-only arrays already fit any request
    up to that size.  A finished request's scene goes back to the pool, and a new
    request takes the smallest idle scene whose arrays fit it (else the largest):
    a scene that grows frees an array, and a free synchronises the device."""

    def __init__(self, budget_bytes: int, max_inflight: int = 16, power: int = 12, cap: int = 64,
                 target_rel: float = 0.01, min_reps: int = 4, max_reps: int = 64, precision: str = "fp32-fixup",
                 pa: bool = True, scene_factory=None, scramble=None, replicate_seed=None, poll_s: float = 5e-5,
                 rebuild_workers: int = 12, batch_plans: int = 0, plan_factory=None, max_batch_jobs: int = 4096,
                 warm_atoms: int = 0):
        if budget_bytes <= 0 or max_inflight < 1 or min_reps < 2 or max_reps < min_reps or rebuild_workers < 0 \
                or batch_plans < 0 or max_batch_jobs < max_reps or warm_atoms < 0:
            raise ValueError("invalid server limits")
        self.warm_atoms = int(warm_atoms)
        self.rebuild_workers = int(rebuild_workers)
        if scene_factory is None or scramble is None or replicate_seed is None:
            from .engine import GPUScene, replicate_seed as rs, scramble as sc
            scene_factory = scene_factory or GPUScene
            scramble = scramble or sc
            replicate_seed = replicate_seed or rs
        if batch_plans and plan_factory is None:
            from .engine import GPUPlan

            def plan_factory(scenes, sobols, powers, caps):
                return GPUPlan(scenes, sobols, powers, caps, precision=precision, pa=pa)
        self.budget = int(budget_bytes)
        self.max_inflight = int(max_inflight)
        self.power, self.cap = int(power), int(cap)
        self.target_rel, self.min_reps, self.max_reps = float(target_rel), int(min_reps), int(max_reps)
        self.precision, self.pa = precision, bool(pa)
        self.scene_factory, self.scramble, self.replicate_seed = scene_factory, scramble, replicate_seed
        self.batch_plans, self.plan_factory, self.max_batch_jobs = int(batch_plans), plan_factory, int(max_batch_jobs)
        self.poll_s = poll_s
        self.peak_bytes = 0
        self.peak_inflight = 0
        self.scenes_built = 0
        self.scenes_reused = 0
        self.batches = 0
        self.peak_batch_jobs = 0

    def footprint(self, req: Request) -> int:
        return footprint_bytes(len(req.radii), self.power, self.cap, self.max_reps)

    def _next_count(self, res: Result) -> int:
        """Replicates of the next launch: min_reps first, then the predicted
        remainder ((sd / (target * mean))^2 - k, 20 % margin), at least one."""
        k = len(res.omegas)
        if k == 0:
            return self.min_reps
        v = np.asarray(res.omegas, dtype=np.float64)
        need = k + 1
        if k >= 2 and v.mean() != 0:
            need = math.ceil(1.2 * (v.std(ddof=1) / (self.target_rel * abs(v.mean()))) ** 2)
        return int(max(1, min(need - k, self.max_reps - k)))

    def _seeds(self, live) -> list:
        k = len(live.res.omegas)
        return [self.replicate_seed(live.req.seed, k + j) for j in range(self._next_count(live.res))]

    def _launch(self, live: _Live) -> None:
        sobols = [self.scramble(s) for s in self._seeds(live)]
        live.scene.solve_async(sobols, self.power, self.cap, precision=self.precision, pa=self.pa)
        live.res.launches += 1

    def _warm_scenes(self, sleep) -> list:
        """max_inflight scenes grown for warm_atoms atoms and max_reps replicates
        (a cubic lattice, spacing 3 A, radius 1.5 A: no two atoms alike)."""
        side = math.ceil(self.warm_atoms ** (1 / 3))
        axis = np.arange(side, dtype=np.float64) * 3.0
        grid = np.stack(np.meshgrid(axis, axis, axis, indexing="ij"), -1).reshape(-1, 3)[:self.warm_atoms]
        radii = np.full(len(grid), 1.5)
        sobols = [self.scramble(self.replicate_seed(0, k)) for k in range(self.max_reps)]
        scenes = []
        for _ in range(self.max_inflight):
            scene = self.scene_factory(grid, radii)
            scene.solve_async(sobols, self.power, self.cap, precision=self.precision, pa=self.pa)
            scenes.append(scene)
        for scene in scenes:
            while not scene.solve_done():
                sleep(self.poll_s)
            scene.solve_wait()
        return scenes

    def _converged(self, res: Result) -> bool:
        k = len(res.omegas)
        if k >= self.max_reps:
            return True
        return k >= self.min_reps and _rel_se(res.omegas) <= self.target_rel

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep) -> list[Result]:
        """Serve `requests` (arrival times relative to the call); returns one
        Result per request, in completion order."""
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting: collections.deque = collections.deque()
        live: dict[int, _Live] = {}
        done: list[Result] = []
        idle: list = []  # (device bytes, scene) of finished requests, kept for reuse
        used = 0         # footprints of live requests plus the bytes of idle scenes
        if self.warm_atoms:
            for scene in self._warm_scenes(sleep):
                held = int(scene.device_bytes())
                idle.append((held, scene))
                used += held
            self.scenes_built += len(idle)
        t0 = clock()
        pool = ThreadPoolExecutor(self.rebuild_workers) if self.rebuild_workers > 0 else None
        batched = self.batch_plans > 0
        ready: collections.deque = collections.deque()  # batched: requests whose next replicates wait
        plans_free: list = []
        plans_made = 0
        in_flight: list = []  # batched: (plan, [(live item, replicate count)])

        def take_scene(req, fp):
            """A scene holding req: an idle one rebuilt in place (on a worker
            thread when there is a pool: the pending Future is returned), or a
            new one.  Returns (scene, bytes accounted for it, Future or None,
            bytes the scene held when idle)."""
            if idle:
                held, scene = idle.pop()
                if pool is not None:
                    return scene, max(fp, held), pool.submit(scene.rebuild, req.centers, req.radii), held
                try:
                    scene.rebuild(req.centers, req.radii)
                except Exception:
                    idle.append((held, scene))
                    raise
                self.scenes_reused += 1
                return scene, max(fp, held), None, held
            scene = self.scene_factory(req.centers, req.radii)
            self.scenes_built += 1
            return scene, fp, None, 0

        def release_scene(scene):
            nonlocal used
            if hasattr(scene, "rebuild") and hasattr(scene, "device_bytes") and \
                    len(idle) + len(live) < self.max_inflight:
                held = int(scene.device_bytes())
                idle.append((held, scene))
                used += held
            else:
                scene.close()

        def start(item):
            """The request's next replicates: launched now (per-request mode) or
            queued for the next batch."""
            if batched:
                ready.append(item)
            else:
                self._launch(item)

        def flush_batches():
            """Batched mode: every ready request's next replicates, as jobs of one
            plan launch, while a plan is free."""
            nonlocal plans_made
            while ready and (plans_free or plans_made < self.batch_plans):
                scenes, sobols, assign = [], [], []
                while ready and len(scenes) + self._next_count(ready[0].res) <= self.max_batch_jobs:
                    item = ready.popleft()
                    seeds = self._seeds(item)
                    scenes += [item.scene] * len(seeds)
                    sobols += [self.scramble(sd) for sd in seeds]
                    assign.append((item, len(seeds)))
                    item.res.launches += 1
                powers, caps = [self.power] * len(scenes), [self.cap] * len(scenes)
                if plans_free:
                    plan = plans_free.pop()
                    plan.retarget(scenes, sobols, powers, caps)
                else:
                    plan = self.plan_factory(scenes, sobols, powers, caps)
                    plans_made += 1
                plan.launch()
                in_flight.append((plan, assign))
                self.batches += 1
                self.peak_batch_jobs = max(self.peak_batch_jobs, len(scenes))

        def finished():
            """(live item, list of replicate outputs) of every completed launch."""
            out = []
            if batched:
                for entry in list(in_flight):
                    plan, assign = entry
                    if not plan.done():
                        continue
                    in_flight.remove(entry)
                    outs = plan.wait()
                    plan.release()  # the scenes may be rebuilt for other requests now
                    plans_free.append(plan)
                    k = 0
                    for item, count in assign:
                        out.append((item, outs[k:k + count]))
                        k += count
            else:
                for item in list(live.values()):
                    if item.rebuild is None and item.scene.solve_done():
                        out.append((item, item.scene.solve_wait()))
            return out

        try:
            while pending or waiting or live:
                now = clock() - t0
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.popleft())
                # Admission: first come first served, by device-memory footprint.
                while waiting and len(live) < self.max_inflight:
                    req = waiting[0]
                    fp = self.footprint(req)
                    if fp > self.budget:
                        waiting.popleft()
                        done.append(Result(req.rid, len(req.radii), arrival=req.arrival, finished=clock() - t0,
                                           error=f"footprint {fp} B exceeds the device budget {self.budget} B"))
                        continue
                    # Memory held by idle scenes goes first when a request does not fit.
                    while used + fp > self.budget and idle:
                        held, scene = idle.pop(0)
                        scene.close()
                        used -= held
                    if used + fp > self.budget:
                        break
                    waiting.popleft()
                    res = Result(req.rid, len(req.radii), arrival=req.arrival, admitted=clock() - t0)
                    if idle:
                        idle.append(idle.pop(_best_fit(idle, fp)))  # the scene reused next goes last
                        used -= idle[-1][0]  # the scene about to be reused is accounted as live
                        held_before = idle[-1][0]
                    else:
                        held_before = 0
                    try:
                        scene, acct, rebuild, held = take_scene(req, fp)
                    except Exception as exc:  # invalid structure: report, keep serving
                        used += held_before   # the idle scene stays idle
                        res.error, res.finished = str(exc), clock() - t0
                        done.append(res)
                        continue
                    item = _Live(req, res, scene, acct, rebuild, held)
                    used += acct
                    live[req.rid] = item
                    if rebuild is None:
                        start(item)
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_inflight = max(self.peak_inflight, len(live))
                progressed = False
                # Rebuilds that finished: the request's first launch follows.
                for rid in list(live):
                    item = live[rid]
                    if item.rebuild is None or not item.rebuild.done():
                        continue
                    progressed = True
                    pending_rebuild, item.rebuild = item.rebuild, None
                    try:
                        pending_rebuild.result()
                    except Exception as exc:  # invalid structure: report; the scene stays idle
                        item.res.error, item.res.finished = str(exc), clock() - t0
                        used -= item.footprint
                        del live[rid]
                        idle.append((item.held, item.scene))
                        used += item.held
                        done.append(item.res)
                        continue
                    self.scenes_reused += 1
                    start(item)
                if batched:
                    flush_batches()
                # Completed launches: relaunch or complete each request.
                for item, outs in finished():
                    progressed = True
                    res = item.res
                    for o in outs:
                        res.omegas.append(o["omega"])
                        res.pa.append(o["pa"])
                        res.unresolved += int(o["unresolved"])
                        res.rays += int(o["rays"])
                    if self._converged(res):
                        res.ccs = float(np.mean(res.omegas))
                        res.rel_se = _rel_se(res.omegas)
                        res.finished = clock() - t0
                        used -= item.footprint
                        del live[res.rid]
                        release_scene(item.scene)
                        done.append(res)
                    else:
                        start(item)
                if batched:
                    flush_batches()
                if not progressed:
                    if not live and not waiting and pending:
                        sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.01)))
                    else:
                        sleep(self.poll_s)
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
            for plan, _ in in_flight:
                plan.wait()
                plan.release()
                plans_free.append(plan)
            for plan in plans_free:
                plan.close()
        for _, scene in idle:
            scene.close()
        return done


def latency_summary(results) -> dict:
    """p50/p99 latency (s), CCS/s over the stream, and counts."""
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
    span = (max(r.finished for r in ok) - min(r.arrival for r in ok)) if ok else 0.0
    return dict(requests=len(results), served=len(ok), rejected=len(results) - len(ok),
                p50_s=float(np.percentile(lat, 50)) if len(lat) else math.nan,
                p99_s=float(np.percentile(lat, 99)) if len(lat) else math.nan,
                mean_s=float(lat.mean()) if len(lat) else math.nan,
                ccs_per_s=len(ok) / span if span > 0 else math.nan,
                rays=int(sum(r.rays for r in ok)),
                max_rel_se=float(max(r.rel_se for r in ok)) if ok else math.nan)


def poisson_stream(bases, rate_hz: float, count: int, sigma: float = 0.3, seed: int = 0) -> list[Request]:
    """`count` requests with Poisson arrivals at `rate_hz`; each is a conformer
    (Gaussian jitter sigma, A) of a base structure drawn uniformly from `bases`
    (a list of (centers, radii))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
    out = []
    for k in range(count):
        c, r = bases[int(rng.integers(0, len(bases)))]
        jitter = rng.normal(scale=sigma, size=np.shape(c))
        out.append(Request(rid=k, centers=np.asarray(c) + jitter, radii=np.asarray(r), arrival=float(t[k]),
                           seed=int(rng.integers(0, 2 ** 31))))
    return out

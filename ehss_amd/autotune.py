"""Autotuner for the fused device solve: packet queries before the drain, lanes
per ray in the queue drain (rays per wavefront), queries of the first drain
stage, cooperative-tail threshold, LDS-staged scene for the packet kernels, and
BVH leaf size.

Every configuration returns the same physical results (the FP32 filter only
skips work whose FP64 outcome is decided, and the nearest hit does not depend
on the tree shape), so the tuner may pick any of them on speed alone; it still
checks that invariant on every configuration it times, through a caller-given
fingerprint of the results.

Search: coordinate descent from the library defaults.  Each parameter in turn
takes every value of its axis with the others held at the best so far; a
configuration's cost is the median of `repeats` timed runs.  Two passes at
most, stopping early when a pass changes nothing.  The winner is cached in a
JSON file under a key naming the device and the workload shape, so later runs
of the same shape skip the search.
"""
from __future__ import annotations

import json
import math
import os
import statistics
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable

DEFAULT_SPACE = {
    "packets": (3, 2, 4, 5),    # wave-uniform (packet) queries before the drain
    "group": (None, 64, 8, 4, 2, 1),  # lanes per ray in the first drain (None: chosen by the drain from its queue)
    "stage1": (6, 2, 4, 12, 0),  # queries before a ray spills to the second drain (0: one drain)
    "coop": (8, 4, 16),         # live rays per wave from which the tail is finished cooperatively
    "lds": (False, True),       # packet kernels read the scene staged in LDS
    "leaf_size": (4, 2, 8),     # spheres per BVH leaf group
}
DEFAULT_START = {"packets": None, "group": None, "stage1": None, "coop": None, "lds": False, "leaf_size": 4}


def default_cache_path() -> Path:
    env = os.environ.get("EHSS_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "ehss_amd" / "autotune.json"


def workload_key(device: str, atoms, rays_per_job: int, cap: int, precision: str, pa: bool) -> str:
    """Shape key: device, job count, log2 of the atom-count range, rays, cap."""
    atoms = [int(a) for a in atoms]
    lo = int(math.floor(math.log2(max(1, min(atoms)))))
    hi = int(math.ceil(math.log2(max(1, max(atoms)))))
    return (f"{device}|jobs={len(atoms)}|atoms=2^{lo}..2^{hi}|rays={int(rays_per_job)}|cap={int(cap)}"
            f"|{precision}|pa={int(bool(pa))}")


def _load(path: Path) -> dict:
    try:
        with open(path, encoding="utf-8") as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: Path, data: dict) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(path.suffix + ".tmp")
    with open(tmp, "w", encoding="utf-8") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _norm(cfg: dict) -> dict:
    """Drop settings that have no effect: stage1 does not matter with a first
    drain of 64 lanes per ray."""
    out = dict(cfg)
    if out.get("group") == 64 and "stage1" in out:
        out["stage1"] = 0
    return out


def _ckey(cfg: dict) -> str:
    return json.dumps(cfg, sort_keys=True)


@dataclass
class TuneResult:
    config: dict
    seconds: float
    default_seconds: float
    trials: list = field(default_factory=list)   # (config, seconds) of every configuration timed
    cached: bool = False

    @property
    def speedup(self) -> float:
        return self.default_seconds / self.seconds if self.seconds > 0 else float("nan")


def tune(measure: Callable[[dict], float], space: dict | None = None, start: dict | None = None,
         repeats: int = 3, passes: int = 2, fingerprint: Callable[[dict], object] | None = None,
         key: str | None = None, cache_path: Path | str | None = None, refresh: bool = False) -> TuneResult:
    """Coordinate-descent search.  measure(cfg) runs the workload once under cfg and
    returns seconds; fingerprint(cfg), when given, returns a comparable summary of
    the results under cfg, which must be the same for every configuration."""
    space = dict(DEFAULT_SPACE if space is None else space)
    start = _norm(dict(DEFAULT_START if start is None else start))
    if repeats < 1 or passes < 1:
        raise ValueError("repeats and passes must be positive")
    for name, axis in space.items():
        if not axis:
            raise ValueError(f"empty axis {name!r}")
    path = Path(cache_path) if cache_path is not None else None
    if key is not None and path is not None and not refresh:
        hit = _load(path).get(key)
        if hit and isinstance(hit.get("config"), dict):
            return TuneResult(config=hit["config"], seconds=float(hit.get("seconds", float("nan"))),
                              default_seconds=float(hit.get("default_seconds", float("nan"))), cached=True)

    timed: dict[str, float] = {}
    trials = []
    reference = [None]

    def cost(cfg: dict) -> float:
        cfg = _norm(cfg)
        k = _ckey(cfg)
        if k not in timed:
            if fingerprint is not None:
                fp = fingerprint(cfg)
                if reference[0] is None:
                    reference[0] = fp
                elif fp != reference[0]:
                    raise AssertionError(f"configuration {cfg} changed the results: {fp!r} != {reference[0]!r}")
            timed[k] = statistics.median(float(measure(cfg)) for _ in range(repeats))
            trials.append((cfg, timed[k]))
        return timed[k]

    best = dict(start)
    best_cost = cost(best)
    default_cost = best_cost
    for _ in range(passes):
        changed = False
        for name, axis in space.items():
            for value in axis:
                cand = _norm(dict(best, **{name: value}))
                if _ckey(cand) == _ckey(_norm(best)):
                    continue
                c = cost(cand)
                if c < best_cost:
                    best, best_cost, changed = cand, c, True
        if not changed:
            break
    best = _norm(best)
    result = TuneResult(config=best, seconds=best_cost, default_seconds=default_cost, trials=trials)
    if key is not None and path is not None:
        data = _load(path)
        data[key] = dict(config=best, seconds=best_cost, default_seconds=default_cost, trials=len(trials))
        _store(path, data)
    return result


def tune_plan(structures, power: int, cap: int, sobols, precision: str = "fp32-fixup", pa: bool = True,
              space: dict | None = None, repeats: int = 3, passes: int = 2, cache_path=None, refresh: bool = False,
              graph: bool = True):
    """Tunes a multi-structure plan of `structures` ((centers, radii) pairs), one
    2^power-ray job each.  Returns (TuneResult, scenes built with the winning leaf
    size, kwargs for GPUPlan)."""
    import time

    from . import _lib as L
    from .engine import GPUPlan, GPUScene

    structures = list(structures)
    scenes_by_leaf: dict[int, list] = {}

    def scenes_for(leaf: int):
        if leaf not in scenes_by_leaf:
            scenes_by_leaf[leaf] = [GPUScene(c, r, leaf_size=int(leaf)) for c, r in structures]
        return scenes_by_leaf[leaf]

    plans: dict[str, object] = {}  # the plan of the configuration being timed (one at a time)

    def plan_for(cfg: dict):
        k = _ckey(cfg)
        if k not in plans:
            for old in plans.values():
                old.close()
            plans.clear()
            plans[k] = GPUPlan(scenes_for(cfg.get("leaf_size", 4)), sobols, [power] * len(structures),
                               [cap] * len(structures), precision=precision, pa=pa, graph=graph,
                               coop=cfg.get("coop"), group=cfg.get("group"), stage1=cfg.get("stage1"),
                               lds=bool(cfg.get("lds", False)), packets=cfg.get("packets"))
            plans[k].run()  # first launch: captures the graph, uploads the job table
        return plans[k]

    def measure(cfg: dict) -> float:
        plan = plan_for(cfg)
        t0 = time.perf_counter()
        plan.run()
        return time.perf_counter() - t0

    def fingerprint(cfg: dict):
        return tuple((r["omega"], r["pa"], r["unresolved"], r["escaped"]) for r in plan_for(cfg).run())

    key = workload_key(L.device_info()["name"], [len(r) for _, r in structures], 1 << power, cap, precision, pa)
    # Start from the library's own choice for this many rays, spelled out, so
    # that the stage1 axis is explored around it.
    at_size = len(structures) << power >= 1 << 18
    start = {"packets": 3, "group": None if at_size else 64, "stage1": 6 if at_size else 0, "coop": 8, "lds": False,
             "leaf_size": 4}
    try:
        res = tune(measure, space=space, start=start, repeats=repeats, passes=passes, fingerprint=fingerprint, key=key,
                   cache_path=default_cache_path() if cache_path is None else cache_path, refresh=refresh)
    finally:
        for p in plans.values():
            p.close()
    leaf = int(res.config.get("leaf_size", 4))
    for lf, scs in scenes_by_leaf.items():
        if lf != leaf:
            for s in scs:
                s.close()
    kwargs = {k: res.config.get(k) for k in ("coop", "group", "stage1", "packets")}
    kwargs["lds"] = bool(res.config.get("lds", False))
    return res, scenes_for(leaf), kwargs

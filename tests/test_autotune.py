"""Autotuner search, invariant check and cache on a simulated cost model (no GPU)."""
import json

import pytest

from ehss_amd import autotune


def separable_cost(cfg):
    """Unique minimum at packets 4, group 4, stage1 2, coop 16, lds False, leaf 8."""
    c = {None: 0.3, 3: 0.3, 2: 0.6, 4: 0.1, 5: 0.2}[cfg["packets"]]
    c += {64: 3.0, 8: 2.0, 4: 1.0, 2: 1.5, 1: 4.0, None: 3.0}[cfg["group"]]
    c += {None: 0.5, 0: 0.7, 1: 0.4, 2: 0.1, 4: 0.3, 6: 0.2, 12: 0.35}[cfg["stage1"]]
    c += {None: 0.2, 8: 0.2, 4: 0.3, 16: 0.05}[cfg["coop"]]
    c += 0.0 if not cfg["lds"] else 0.25
    c += {4: 0.5, 2: 0.9, 8: 0.2}[cfg["leaf_size"]]
    return c


def test_coordinate_descent_finds_the_separable_minimum(tmp_path):
    calls = []

    def measure(cfg):
        calls.append(dict(cfg))
        return separable_cost(cfg)

    res = autotune.tune(measure, repeats=1, cache_path=tmp_path / "t.json", key="k")
    assert res.config == {"packets": 4, "group": 4, "stage1": 2, "coop": 16, "lds": False, "leaf_size": 8}
    assert res.seconds == pytest.approx(separable_cost(res.config))
    assert res.default_seconds == pytest.approx(separable_cost(autotune._norm(autotune.DEFAULT_START)))
    assert res.speedup > 1 and not res.cached
    # every configuration is timed once per repeat, never twice
    keys = [autotune._ckey(c) for c in calls]
    assert len(keys) == len(set(keys)) == len(res.trials)


def test_repeats_take_the_median():
    seq = iter([5.0, 1.0, 3.0] + [10.0] * 1000)
    res = autotune.tune(lambda cfg: next(seq), space={"group": (64,)}, start={"group": 64}, repeats=3)
    assert res.seconds == 3.0


def test_cache_hit_skips_the_search_and_refresh_redoes_it(tmp_path):
    path = tmp_path / "cache.json"
    autotune.tune(separable_cost, repeats=1, key="shape", cache_path=path)
    data = json.loads(path.read_text())
    assert data["shape"]["config"]["group"] == 4

    def never(cfg):
        raise AssertionError("measured despite a cached entry")

    hit = autotune.tune(never, key="shape", cache_path=path)
    assert hit.cached and hit.config == data["shape"]["config"]
    other = autotune.tune(lambda c: 1.0, space={"coop": (8, 4)}, start={"coop": 8}, repeats=1, key="other",
                          cache_path=path)
    assert not other.cached and set(json.loads(path.read_text())) == {"shape", "other"}
    redo = autotune.tune(separable_cost, repeats=1, key="shape", cache_path=path, refresh=True)
    assert not redo.cached


def test_a_configuration_that_changes_results_is_refused():
    def fingerprint(cfg):
        return 1.0 if cfg["group"] != 1 else 1.0000001  # a planted disagreement

    with pytest.raises(AssertionError, match="changed the results"):
        autotune.tune(separable_cost, repeats=1, fingerprint=fingerprint)


def test_stage1_is_ignored_with_a_whole_wave_first_drain():
    assert autotune._norm({"group": 64, "stage1": 2}) == {"group": 64, "stage1": 0}
    assert autotune._norm({"group": None, "stage1": 2}) == {"group": None, "stage1": 2}
    assert autotune._norm({"group": 4, "stage1": 2}) == {"group": 4, "stage1": 2}


def test_workload_key_and_validation():
    k = autotune.workload_key("MI355X", [200, 4999], 1 << 20, 64, "fp32-fixup", True)
    assert k == "MI355X|jobs=2|atoms=2^7..2^13|rays=1048576|cap=64|fp32-fixup|pa=1"
    with pytest.raises(ValueError):
        autotune.tune(separable_cost, repeats=0)
    with pytest.raises(ValueError):
        autotune.tune(separable_cost, space={"group": ()})


def test_unreadable_cache_is_treated_as_empty(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text("{not json")
    res = autotune.tune(lambda c: 1.0, space={"coop": (8,)}, start={"coop": 8}, repeats=1, key="k", cache_path=path)
    assert not res.cached and "k" in json.loads(path.read_text())

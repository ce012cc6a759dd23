"""GPU: alternative incidence sources generated, traced and reduced on the
device (ehss_solve_source, ehss_source_states) against the oracle.

* Reverse PA (ehss_state.py:55-69): sample_source's states, the inside-out
  blocked flags and the reflected states equal the oracle's bit for bit; the
  solve equals the oracle's trace + readout of the same states ray by ray
  (rays flagged grazing excepted) and in Omega; the reference's single-sphere
  test (test_ehss_reference.py:8-17) holds.
* Box (ehss_compact_source.py:56-70): states bit-identical to the oracle's; the
  solve equals the oracle's trace.
* Ellipsoid (ehss_compact_source.py:71-84): its rotation products go through
  BLAS in NumPy, so the states agree to rounding, and Omega within sampling of
  the impact disk's.
"""
import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd.engine import GPUScene, scramble, scramble5, source_params
from ehss_amd.sources import fit_centroid_ellipsoid
from ehss_amd.state import Spheres
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


def _cdf(r):
    w = r ** 2
    return np.cumsum(w) / w.sum()  # inside_out_pa.py:79


def _area(r):
    return float(4 * np.pi * np.dot(r, r))  # Spheres.area


def _box(c, r):
    """CompactBoundary.fit(kind='box') (ehss_compact_source.py:34-39)."""
    low = np.min(c - r[:, None], axis=0)
    high = np.max(c + r[:, None], axis=0)
    return (low + high) / 2, np.eye(3), (high - low) / 2 * (1 + 1e-9)


def _reverse(r):
    return source_params("reverse-pa", area=_area(r), cdf=_cdf(r))


STRUCTS = [("globule300", scenes.globule(300, 4)), ("c60", scenes.c60()), ("f37", scenes.SCENES["f37"]())]


def _check_solve_against_oracle(oracle, c, r, res, src, cap, keep=None, tag=""):
    """res: device solve with keep_rays; src: the oracle's source dict (kept rays only)."""
    tree = oracle.static_tree(c, r, 4)
    tr = oracle.trace(tree, c, r, src["origins"], src["directions"], last_atom=src["last_atom"],
                      initial_bounces=src["initial_bounces"], cap=cap, record=False, threads=8)
    out = oracle.readout(src, tr, cap)
    rays = res["rays_out"]
    sel = np.ones(len(rays["bounces"]), bool) if keep is None else keep
    ok = ~rays["grazing"][sel]
    for k, exp in (("bounces", tr["bounces"]), ("escaped", tr["escaped"]), ("unresolved", tr["unresolved"])):
        np.testing.assert_array_equal(rays[k][sel][ok], exp[ok], err_msg=f"{tag}/{k}")
    np.testing.assert_array_equal(rays["contributions"][sel][ok], out["contributions"][ok], err_msg=f"{tag}/contrib")
    differ = int(np.sum(rays["contributions"][sel] != out["contributions"]))
    wmax = float(np.max(src["weights"]))
    say(tag, "omega", res["omega"], out["omega"], "grazing", int((~ok).sum()), "differing", differ)
    assert abs(res["omega"] - out["omega"]) <= 2 * wmax * differ + 1e-12 * abs(out["omega"])
    hit = (rays["bounces"][sel] > 0) | rays["unresolved"][sel]
    hit_exp = (tr["bounces"] > 0) | tr["unresolved"]
    assert abs(res["pa_from_ehss"] - out["pa"]) <= wmax * int(np.sum(hit != hit_exp)) + 1e-13 * out["pa"]
    return out


@pytest.mark.parametrize("power,seed", [(12, 881), (14, 98)])
def test_reverse_pa_states_equal_the_oracle(oracle, power, seed):
    for name, (c, r) in STRUCTS:
        say("reverse-pa states", name, power, seed)
        with GPUScene(c, r) as sc:
            got = sc.source_states(_reverse(r), scramble5(seed, 5), power)
        smp = oracle.sample_source(c, r, power, seed)
        np.testing.assert_array_equal(got["origins"], smp["origins"], err_msg=name)
        np.testing.assert_array_equal(got["owners"], smp["owners"], err_msg=name)
        blocked = oracle.pa_brute(c, r, smp["origins"], smp["directions"], smp["owners"])
        np.testing.assert_array_equal(got["blocked"], blocked, err_msg=name)
        assert 0 < blocked.sum() < len(blocked)
        rev = oracle.reverse_pa_source(c, r, power, seed)
        keep = ~blocked
        for k in ("directions", "incoming", "weights"):
            np.testing.assert_array_equal(got[k][keep], rev[k], err_msg=f"{name}/{k}")
        assert np.all(got["weights"][blocked] == 0) and np.all(got["bounces"] == 1)


def test_reverse_pa_solve_equals_the_oracle_trace(oracle):
    for name, (c, r) in STRUCTS:
        say("reverse-pa solve", name)
        power, seed, cap = 14, 98, 64
        with GPUScene(c, r) as sc:
            src = _reverse(r)
            res = sc.solve_source(src, scramble5(seed, 5), power, cap, histogram=True, keep_rays=True)
            blocked = sc.source_states(src, scramble5(seed, 5), power)["blocked"]
        rev = oracle.reverse_pa_source(c, r, power, seed)
        _check_solve_against_oracle(oracle, c, r, res, rev, cap, keep=~blocked, tag=name)
        n = 1 << power
        assert res["pa_hits"] == int(blocked.sum())
        assert res["pa"] == _area(r) * float(np.mean(~blocked)) / 4  # inside_out_pa.py:66-67
        assert res["weight"] == pytest.approx(float(rev["weights"].sum()), rel=1e-14)
        assert res["rays"] == n and res["histogram"].sum() == int((~blocked).sum())


@pytest.mark.parametrize("radius", [.01, 1., 100.])
def test_single_sphere_reverse_pa(radius):
    """test_ehss_reference.py:8-17 on the device path: every reflected state
    escapes at once; Omega = PA = pi r^2."""
    c, r = np.array([[2., -1., 4.]]), np.array([radius])
    with GPUScene(c, r) as sc:
        res = sc.solve_source(_reverse(r), scramble5(98, 5), 14, 4, keep_rays=True)
    rays = res["rays_out"]
    assert res["escaped"] == 1 << 14 and res["unresolved"] == 0
    assert np.all(rays["bounces"] == 1) and rays["escaped"].all()
    assert res["omega"] == pytest.approx(np.pi * radius * radius, rel=1e-6)
    assert res["pa_from_ehss"] == pytest.approx(np.pi * radius * radius, rel=1e-7)
    assert res["pa"] == pytest.approx(np.pi * radius * radius, rel=1e-15)


def test_box_states_and_solve_equal_the_oracle(oracle):
    for name, (c, r) in STRUCTS:
        say("box", name)
        center, R, axes = _box(c, r)
        src = source_params("box", center, R, axes)
        with GPUScene(c, r) as sc:
            got = sc.source_states(src, scramble5(5, 5), 14)
            res = sc.solve_source(src, scramble5(5, 5), 14, 64, keep_rays=True)
        exp = oracle.compact_source("box", center, R, axes, 14, 5)
        for k in ("origins", "directions", "incoming", "weights"):
            np.testing.assert_array_equal(got[k], exp[k], err_msg=f"{name}/{k}")
        assert np.all(got["owners"] == -1) and np.all(got["bounces"] == 0) and not got["blocked"].any()
        _check_solve_against_oracle(oracle, c, r, res, exp, 64, tag=f"box/{name}")


def test_ellipsoid_states_and_physics(oracle):
    c, r = scenes.globule(300, 4)
    center, R, axes = fit_centroid_ellipsoid(Spheres(c, r))
    src = source_params("ellipsoid", center, R, axes)
    with GPUScene(c, r) as sc:
        got = sc.source_states(src, scramble5(9, 4), 14)
        res = sc.solve_source(src, scramble5(9, 4), 16, 64)
        disk = np.mean([sc.solve(scramble(100 + k), 16, 64)["omega"] for k in range(4)])
    exp = oracle.compact_source("ellipsoid", center, R, axes, 14, 9)
    np.testing.assert_array_equal(got["directions"], exp["directions"])  # no rotation involved
    np.testing.assert_allclose(got["origins"], exp["origins"], rtol=0, atol=1e-12 * np.max(np.abs(exp["origins"])))
    np.testing.assert_allclose(got["weights"], exp["weights"], rtol=1e-13)
    say("ellipsoid omega", res["omega"], "impact disk", disk)
    assert res["omega"] == pytest.approx(disk, rel=0.02)


def test_sources_agree_with_the_impact_disk():
    """Every isotropic source estimates the same cross-section."""
    c, r = scenes.globule(650, 1)
    with GPUScene(c, r) as sc:
        disk = np.mean([sc.solve(scramble(200 + k), 16, 64)["omega"] for k in range(4)])
        center, R, axes = _box(c, r)
        box = np.mean([sc.solve_source(source_params("box", center, R, axes), scramble5(k, 5), 16, 64)["omega"]
                       for k in range(4)])
        rev = np.mean([sc.solve_source(_reverse(r), scramble5(k, 5), 16, 64)["omega"] for k in range(4)])
    say("omega: disk", disk, "box", box, "reverse-pa", rev)
    assert box == pytest.approx(disk, rel=0.02) and rev == pytest.approx(disk, rel=0.02)


def test_source_arguments_are_checked():
    c, r = scenes.c60()
    with GPUScene(c, r) as sc:
        bad = source_params("box", np.zeros(3), np.eye(3), [1., 0., 1.])
        with pytest.raises(L.EHSSError, match="half-widths"):
            sc.solve_source(bad, scramble5(1, 5), 10)
        src = _reverse(r)
        src.cdf = None
        with pytest.raises(L.EHSSError, match="searchsorted"):
            sc.solve_source(src, scramble5(1, 5), 10)
        with pytest.raises(ValueError):
            source_params("sphere")


def test_product_entry_points_equal_the_c_abi(oracle):
    """CompactBoundary.solve and reverse_pa_solve (the package's API) run the
    same device path as ehss_solve_source, and reverse_pa_solve's totals equal
    the host form's states traced by the oracle."""
    from ehss_amd.sources import CompactBoundary, reverse_pa_solve
    c, r = scenes.globule(300, 4)
    spheres = Spheres(c, r)
    with GPUScene(c, r) as sc:
        box = CompactBoundary.fit(spheres, "box")
        a = box.solve(sc, 12, 3)
        b = sc.solve_source(source_params("box", *_box(c, r)), scramble5(3, 5), 12)
        assert a["omega"] == b["omega"] and a["escaped"] == b["escaped"]
        rev = reverse_pa_solve(sc, spheres, 12, 98)
    src = oracle.reverse_pa_source(c, r, 12, 98)
    assert rev["pa"] == pytest.approx(float(src["weights"].sum()), rel=1e-13)
    tr = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"],
                      last_atom=src["last_atom"], initial_bounces=src["initial_bounces"], cap=64, record=False)
    assert rev["omega"] == pytest.approx(oracle.readout(src, tr, 64)["omega"], rel=1e-3)

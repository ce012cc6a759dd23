"""GPU parity of the device scene (tree build), scene trace / PA query, the
device Sobol source and the fused solve, against the oracle."""
import numpy as np
import pytest

from ehss_amd import _lib as L
from ehss_amd.engine import GPUScene, scramble
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

BUILD_CASES = ["f29", "f37", "random47", "ties47", "coincident47", "single", "translated47", "pose4", "cross6",
               "f23", "c60"]


@pytest.mark.parametrize("leaf", [1, 4, 9])
def test_device_tree_build_equals_oracle(oracle, leaf):
    cases = [(n, scenes.SCENES[n]()) for n in BUILD_CASES]
    cases.append(("globule650", scenes.globule(650, 1)))
    cases.append(("ellipsoid2000", scenes.globule(2000, 2, (1., 1., 1.4))))
    for name, (c, r) in cases:
        say("build", name, "leaf", leaf)
        with GPUScene(c, r, leaf_size=leaf) as sc:
            got = sc.export_tree()
            exp = oracle.static_tree(c, r, leaf)
            for k in ("lo", "hi", "end", "leaf_group", "offsets", "ids", "groups"):
                np.testing.assert_array_equal(got[k], exp[k], err_msg=f"{name}/leaf{leaf}/{k}")
            gp, ep = sc.export_pa_tree(), oracle.pa_tree(c, r, leaf)
            for k in ("lo", "hi", "end", "first", "count", "ids"):
                np.testing.assert_array_equal(gp[k], ep[k], err_msg=f"{name}/leaf{leaf}/pa/{k}")
            center, radius = oracle.enclosing(c, r)
            assert sc.info["center"] == list(center) and sc.info["radius"] == radius
            assert sc.info["groups"] == int(exp["groups"].max()) + 1


def test_scene_trace_and_pa_match_oracle(oracle):
    for name in ("f29", "c60", "coincident47"):
        c, r = scenes.SCENES[name]()
        tree = oracle.static_tree(c, r, 4)
        ptree = oracle.pa_tree(c, r, 4)
        with GPUScene(c, r) as sc:
            for kind, src in (("external", oracle.external_source(c, r, 11, 5)),
                              ("reverse", oracle.reverse_pa_source(c, r, 10, 6))):
                exp = oracle.trace(tree, c, r, src["origins"], src["directions"], src["last_atom"],
                                   src["initial_bounces"], cap=64)
                for prec in ("fp64", "fp32-fixup"):
                    say("scene trace", name, kind, prec)
                    got = sc.trace(src["origins"], src["directions"], src["last_atom"], src["initial_bounces"], cap=64,
                                   record=True, precision=prec)
                    for k in ("positions", "outgoing", "escaped", "unresolved", "bounces", "collider_ids"):
                        np.testing.assert_array_equal(got[k], exp[k], err_msg=f"{name}/{kind}/{prec}/{k}")
                    if prec == "fp64":
                        assert (got["box_tests"], got["sphere_tests"]) == (exp["box_tests"], exp["sphere_tests"])
                owners = src["last_atom"]
                hits, nb, nt = sc.pa_query(src["origins"], src["directions"], owners)
                eh, eb, et = oracle.pa_query(ptree, c, r, src["origins"], src["directions"], owners)
                np.testing.assert_array_equal(hits, eh)
                assert (nb, nt) == (eb, et)


def test_device_source_matches_numpy_source(oracle):
    """Identical Sobol points give identical rays: u, directions and origins of
    the device source equal oracle.external_source (NumPy, ehss_state.py:42-52)
    bit for bit on every ray, at C1 (C60, 2^16 rays, seeds 0/104729/209458) and
    C2 (650-atom globule, 2^20 rays) sizes."""
    assert L.trig_compat_info()["entries"] > 0
    cases = [("c60", scenes.c60(), s, 16) for s in (0, 104729, 209458)]
    cases += [("globule650", scenes.globule(650, 1), s, p) for s, p in ((1, 20), (190003, 14), (4294967295, 10))]
    for name, (c, r), seed, power in cases:
        say("source", name, seed, power)
        with GPUScene(c, r) as sc:
            dev = sc.source_rays(scramble(seed), power)
        sv, shift = oracle.sobol_scramble(seed)
        np.testing.assert_array_equal(dev["u"], oracle.sobol_points(sv, shift, power))
        ref = oracle.external_source(c, r, power, seed)
        np.testing.assert_array_equal(dev["directions"], ref["directions"])
        np.testing.assert_array_equal(dev["origins"], ref["origins"])


def _solve_vs_oracle(oracle, c, r, seed, power, cap, prec):
    with GPUScene(c, r) as sc:
        res = sc.solve(scramble(seed), power, cap, precision=prec, pa=True, keep_rays=True, histogram=True)
        rays = sc.last_rays(1 << power)
    src = oracle.external_source(c, r, power, seed)
    tr = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=cap, record=False,
                      threads=8)
    ro = oracle.readout(src, tr, cap)
    hits, _, _ = oracle.pa_query(oracle.pa_tree(c, r, 4), c, r, src["origins"], src["directions"], threads=8)
    return res, rays, tr, ro, hits, src


@pytest.mark.parametrize("prec", ["fp64", "fp32-fixup"])
def test_fused_solve_matches_oracle_on_same_sobol_points(oracle, prec):
    """Same Sobol points: the device source is bit-identical to NumPy's, so every
    ray -- grazing or not -- has the oracle's bounce count, escape/unresolved
    and PA flags and contribution (bitwise); CCS, PA and tail agree to summation
    order (1e-12 relative; the stated bar is 1e-6)."""
    cases = [("c60", scenes.c60(), 0, 14, 64), ("f29", scenes.fixture_29(), 881, 12, 4),
             ("globule650", scenes.globule(650, 1), 1, 14, 64), ("cross6", scenes.fixture_cross(), 3, 12, 1)]
    # Small nets: below 2^6 rays the device map is the Gray code itself; at 2^6
    # every index bit is a lane bit; at 2^7 one bit selects the wave.
    cases += [("c60/2^3", scenes.c60(), 0, 3, 64), ("c60/2^5", scenes.c60(), 5, 5, 64),
              ("globule650/2^6", scenes.globule(650, 1), 2, 6, 64), ("globule650/2^7", scenes.globule(650, 1), 7, 7, 64)]
    for name, (c, r), seed, power, cap in cases:
        say("solve", name, prec, seed, power, cap)
        res, rays, tr, ro, hits, src = _solve_vs_oracle(oracle, c, r, seed, power, cap, prec)
        say(f"  grazing-flagged rays: {int(rays['grazing'].sum())} of {len(rays['grazing'])}")
        np.testing.assert_array_equal(rays["bounces"], tr["bounces"])
        np.testing.assert_array_equal(rays["escaped"], tr["escaped"])
        np.testing.assert_array_equal(rays["unresolved"], tr["unresolved"])
        np.testing.assert_array_equal(rays["pa_hit"], hits)
        np.testing.assert_array_equal(rays["contributions"], ro["contributions"])
        assert res["omega"] == pytest.approx(ro["omega"], rel=1e-12)
        assert res["pa"] == pytest.approx(float(src["weights"][hits].sum()), rel=1e-12, abs=1e-12)
        assert res["pa_from_ehss"] == pytest.approx(ro["pa"], rel=1e-12, abs=1e-12)
        assert res["tail_bound"] == pytest.approx(ro["tail_bound"], rel=1e-12, abs=1e-12)
        np.testing.assert_array_equal(res["histogram"], ro["histogram"])
        np.testing.assert_allclose(res["ledger"].sum(), res["omega"], rtol=1e-10)


@pytest.mark.parametrize("name", ["globule650", "c60", "f29", "cross6"])
def test_plain_and_fp32_fixup_kernels_are_bit_identical(oracle, name):
    c, r = scenes.globule(650, 1) if name == "globule650" else scenes.SCENES[name]()
    with GPUScene(c, r) as sc:
        runs = {}
        # coop None: default tail threshold; 0: tail mode off; 64: every wave
        # switches to cooperative queries as soon as it may.  group: lanes per
        # ray in variant 4's drain (1: per-lane traversal).
        # stage1: queries in the first drain before a ray moves to the whole-wave
        # second drain (0: one drain).
        kernels = [("fp64", 0, None, None, None)] + [("fp32-fixup", v, c, None, None) for v in (1, 2, 3, 4)
                                                     for c in (None, 0, 64)]
        kernels += [("fp32-fixup", 4, c, g, 0) for c in (0, 64) for g in (1, 2, 16, 64)]
        kernels += [("fp32-fixup", 4, None, g, q) for g, q in ((4, 1), (8, 2), (8, 4), (32, 14))]
        for prec, variant, coop, group, stage1 in kernels:
            say("kernels", name, prec, variant, coop, group, stage1)
            res = sc.solve(scramble(7), 16, 64, precision=prec, pa=True, keep_rays=True, histogram=True,
                           variant=variant, coop=coop, group=group, stage1=stage1)
            runs[(prec, variant, coop, group, stage1)] = (res, sc.last_rays(1 << 16))
        for packets in (1, 2, 5, 15):  # packet queries before variant 4's queue
            say("kernels", name, "packets", packets)
            res = sc.solve(scramble(7), 16, 64, pa=True, keep_rays=True, histogram=True, variant=4,
                           packets=packets)
            runs[("fp32-fixup", 4, "packets", packets, None)] = (res, sc.last_rays(1 << 16))
        for variant in (3, 4):  # packet phase reading an LDS-staged scene
            say("kernels", name, "lds packets", variant)
            res = sc.solve(scramble(7), 16, 64, pa=True, keep_rays=True, histogram=True, variant=variant, lds=True)
            runs[("fp32-fixup", variant, "lds", None, None)] = (res, sc.last_rays(1 << 16))
        a = runs[("fp64", 0, None, None, None)]
        for key, b in runs.items():
            assert a[0]["omega"] == b[0]["omega"] and a[0]["pa"] == b[0]["pa"], key
            np.testing.assert_array_equal(a[1]["contributions"], b[1]["contributions"], err_msg=str(key))
            phys = ~np.uint32(16)  # every flag bit except the kernel-specific grazing diagnostic
            np.testing.assert_array_equal(a[1]["flags"] & phys, b[1]["flags"] & phys, err_msg=str(key))
            np.testing.assert_array_equal(a[0]["histogram"], b[0]["histogram"], err_msg=str(key))
        b = runs[("fp32-fixup", 2, None, None, None)]
        # determinism: repeat
        again = sc.solve(scramble(7), 16, 64, precision="fp32-fixup", pa=True)
        assert again["omega"] == b[0]["omega"]


@pytest.mark.parametrize("leaf", [1, 2, 4, 8, "groups"])
def test_packet_tree_cuts_keep_results(leaf):
    """The packet queries walk the static tree with subtrees of about 10 atoms
    merged into leaves (here 10, 7, 3 and 1 groups per leaf for leaf sizes 1, 2,
    4 and 8 of the 650-atom globule, and 10 for one caller group per atom):
    any tree whose boxes enclose their spheres gives the plain FP64 kernel's
    per-ray records."""
    c, r = scenes.globule(650, 1)
    kw = dict(groups=np.arange(len(r))) if leaf == "groups" else dict(leaf_size=leaf)
    with GPUScene(c, r, **kw) as sc:
        info = sc.info
        assert (info["packet_nodes"] < info["nodes"]) == (leaf != 8), info
        ref = sc.solve(scramble(11), 16, 64, precision="fp64", pa=True, keep_rays=True, histogram=True)
        ref_rays = sc.last_rays(1 << 16)
        phys = ~np.uint32(16)  # every flag bit except the kernel-specific grazing diagnostic
        for opts in (dict(variant=4), dict(variant=4, lds=True), dict(variant=3), dict(variant=2)):
            say("packet tree", leaf, info["packet_nodes"], info["nodes"], opts)
            res = sc.solve(scramble(11), 16, 64, pa=True, keep_rays=True, histogram=True, **opts)
            rays = sc.last_rays(1 << 16)
            assert res["omega"] == ref["omega"] and res["pa"] == ref["pa"], opts
            np.testing.assert_array_equal(rays["contributions"], ref_rays["contributions"], err_msg=str(opts))
            np.testing.assert_array_equal(rays["flags"] & phys, ref_rays["flags"] & phys, err_msg=str(opts))
            np.testing.assert_array_equal(res["histogram"], ref["histogram"], err_msg=str(opts))


def test_large_scene_drain_walks_frontier_subtrees():
    """Above 16384 groups the drain's cooperative queries walk the frontier
    subtrees instead of scanning every leaf box; the answers still equal the
    plain FP64 kernel's bit for bit."""
    c, r = scenes.globule(20000, 3)
    with GPUScene(c, r, leaf_size=1) as sc:
        assert sc.info["groups"] > 16384
        runs = {}
        for key, kw in (("fp64", dict(precision="fp64")), ("default", {}), ("g4", dict(group=4, stage1=6)),
                        ("coop", dict(group=64, stage1=0))):
            say("large scene", key)
            res = sc.solve(scramble(5), 14, 64, pa=True, keep_rays=True, histogram=True, **kw)
            runs[key] = (res, sc.last_rays(1 << 14))
        a = runs["fp64"]
        phys = ~np.uint32(16)
        for key, b in runs.items():
            assert a[0]["omega"] == b[0]["omega"] and a[0]["pa"] == b[0]["pa"], key
            np.testing.assert_array_equal(a[1]["contributions"], b[1]["contributions"], err_msg=key)
            np.testing.assert_array_equal(a[1]["flags"] & phys, b[1]["flags"] & phys, err_msg=key)
            np.testing.assert_array_equal(a[0]["histogram"], b[0]["histogram"], err_msg=key)


def test_single_sphere_solve_is_pi_r2():
    c, r = np.zeros((1, 3)), np.array([3.1])
    with GPUScene(c, r) as sc:
        for prec in ("fp64", "fp32-fixup"):
            res = sc.solve(scramble(91), 12, 64, precision=prec)
            assert res["omega"] == pytest.approx(np.pi * 3.1 ** 2, rel=1e-6)
            assert res["pa"] == pytest.approx(np.pi * 3.1 ** 2, rel=2e-3)


@pytest.mark.parametrize("config", ["C1", "C2"])
def test_full_size_same_sobol_ccs_within_1e_6(oracle, config):
    """The stated correctness bar at full size: on identical Sobol points the
    fused solve's CCS is within 1e-6 relative of the oracle's FP64 value, and
    every ray's bounces, escape/unresolved and PA flags equal the oracle's.
    C1: C60, 2^16 rays x seeds 0/104729/209458, cap 64.  C2: 650-atom globule,
    2^20 rays, cap 64 (both kernels)."""
    if config == "C1":
        (c, r), runs = scenes.c60(), [(s, 16, "fp32-fixup") for s in (0, 104729, 209458)]
    else:
        (c, r), runs = scenes.globule(650, 1), [(1, 20, "fp64"), (1, 20, "fp32-fixup")]
    tree, ptree = oracle.static_tree(c, r, 4), oracle.pa_tree(c, r, 4)
    with GPUScene(c, r) as sc:
        for seed, power, prec in runs:
            say("full size", config, seed, power, prec)
            res = sc.solve(scramble(seed), power, 64, precision=prec, pa=True, keep_rays=True, histogram=True)
            rays = sc.last_rays(1 << power)
            src = oracle.external_source(c, r, power, seed)
            tr = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=64, record=False, threads=16)
            ro = oracle.readout(src, tr, 64)
            hits, _, _ = oracle.pa_query(ptree, c, r, src["origins"], src["directions"], threads=16)
            cpu = oracle.solve_cpu(c, r, seed, power, 64)
            rel = abs(res["omega"] - ro["omega"]) / ro["omega"]
            say(f"  omega {res['omega']:.12f} oracle {ro['omega']:.12f} rel {rel:.2e}; "
                f"C solve rel {abs(res['omega'] - cpu['omega']) / cpu['omega']:.2e}; "
                f"grazing-flagged {int(rays['grazing'].sum())}")
            assert rel <= 1e-6
            assert abs(res["omega"] - cpu["omega"]) <= 1e-6 * cpu["omega"]
            np.testing.assert_array_equal(rays["bounces"], tr["bounces"])
            np.testing.assert_array_equal(rays["escaped"], tr["escaped"])
            np.testing.assert_array_equal(rays["unresolved"], tr["unresolved"])
            np.testing.assert_array_equal(rays["pa_hit"], hits)
            assert abs(res["pa"] - cpu["pa"]) <= 1e-6 * cpu["pa"]
            assert res["unresolved"] == cpu["unresolved"]


def test_caller_groups_build(oracle):
    """PreparedStaticCPUEHSS(spheres, groups=...) (ehss_static.py:64-91) on the device."""
    c, r = scenes.fixture_29()
    for groups in (np.arange(29) // 3, np.arange(29)[::-1] % 5, np.zeros(29, dtype=int)):
        say("custom groups", groups[:6])
        with GPUScene(c, r, groups=groups) as sc:
            got = sc.export_tree()
            exp = oracle.static_tree(c, r, 4, groups=groups)
            for k in ("lo", "hi", "end", "leaf_group", "offsets", "ids"):
                np.testing.assert_array_equal(got[k], exp[k], err_msg=k)
            np.testing.assert_array_equal(got["groups"], groups)
            src = oracle.external_source(c, r, 10, 881)
            exp_t = oracle.trace(exp, c, r, src["origins"], src["directions"], cap=64)
            got_t = sc.trace(src["origins"], src["directions"], cap=64, record=True, precision="fp32-fixup")
            np.testing.assert_array_equal(got_t["outgoing"], exp_t["outgoing"])
    from ehss_amd import _lib as L
    for bad in ([0, 2], [-1, 0]):
        with pytest.raises(L.EHSSError, match="Contiguous"):
            GPUScene(np.array([[0., 0, 0], [2, 0, 0]]), np.ones(2), groups=bad)


def test_rebuild_in_place_equals_a_fresh_scene():
    """ehss_scene_rebuild: trees, PA tree and solves equal a scene created for
    the new structure, growing and shrinking; busy while a solve is pending."""
    structs = [scenes.globule(300, 4), scenes.globule(900, 6), scenes.SCENES["c60"](), scenes.globule(650, 1)]
    with GPUScene(*structs[0]) as sc:
        for k, (c, r) in enumerate(structs[1:]):
            say("rebuild", k, len(r))
            sc.rebuild(c, r)
            with GPUScene(c, r) as fresh:
                a, b = sc.export_tree(), fresh.export_tree()
                for key in a:
                    np.testing.assert_array_equal(a[key], b[key], err_msg=key)
                pa, pb = sc.export_pa_tree(), fresh.export_pa_tree()
                for key in pa:
                    np.testing.assert_array_equal(pa[key], pb[key], err_msg=key)
                assert sc.info["atoms"] == len(r) and sc.info["radius"] == fresh.info["radius"]
                x = sc.solve(scramble(5), 14, 64, histogram=True)
                y = fresh.solve(scramble(5), 14, 64, histogram=True)
                for key in ("omega", "pa", "escaped", "unresolved", "max_bounces"):
                    assert x[key] == y[key], (k, key)
                np.testing.assert_array_equal(x["histogram"], y["histogram"])
        sc.solve_async(scramble(1), 12, 64)
        with pytest.raises(L.EHSSError, match="pending"):
            sc.rebuild(*structs[0])
        sc.solve_wait()
        sc.solve(scramble(2), 12, 64, keep_rays=True)
        sc.rebuild(*structs[0])
        with pytest.raises(L.EHSSError, match="KEEP_RAYS"):
            sc.last_rays(1 << 12)
        with pytest.raises(ValueError):
            sc.rebuild(np.zeros((0, 3)), np.zeros(0))

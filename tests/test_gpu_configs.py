"""GPU tests of the configurations the benchmark runs beyond C1/C2:

C3  the 2000-atom ellipsoid (seed 2, axes 1:1:1.4) at cap 128: every ray equal
    to the oracle's at 2^14 rays (both kernels); at the full 2^22 rays the
    FP32+fixup kernel equals the plain FP64 kernel and the CCS is within 1e-6
    of the oracle's C solve.
C4  the 18-structure set (log-uniform 200-5000 atoms, seeds 3..20) at 2^20 rays
    each as one autotuned plan replayed from a hipGraph: every job equals its
    solve alone, on two replays with different scrambles; three jobs equal the
    oracle at 2^14 rays.
C5  a Poisson serving mix of conformers of bases from 100 to 20,000 atoms at a
    1 % target: every request's replicates are the same alone and in the mix,
    scenes are rebuilt in place on worker threads, and one request's
    replicates equal the oracle's solves of its conformer.
"""
import numpy as np
import pytest

from ehss_amd import autotune, molecules, serve
from ehss_amd.engine import GPUPlan, GPUScene, replicate_seed, scramble
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

PHYSICAL = ("omega", "pa", "pa_from_ehss", "residual_mass", "tail_bound", "weight", "rays", "escaped",
            "unresolved", "pa_hits", "ehss_hits", "max_bounces")


def _oracle_solve(oracle, c, r, seed, power, cap):
    src = oracle.external_source(c, r, power, seed)
    tr = oracle.trace(oracle.static_tree(c, r, 4), c, r, src["origins"], src["directions"], cap=cap, record=False,
                      threads=16)
    ro = oracle.readout(src, tr, cap)
    hits, _, _ = oracle.pa_query(oracle.pa_tree(c, r, 4), c, r, src["origins"], src["directions"], threads=16)
    return src, tr, ro, hits


def _same_as_oracle(res, ro, tr, hits):
    w = res["weight"]
    assert res["omega"] == pytest.approx(ro["omega"], rel=1e-12)
    assert res["unresolved"] == int(tr["unresolved"].sum())
    assert res["pa_hits"] == int(hits.sum())
    assert res["ehss_hits"] == int(((tr["bounces"] > 0) | tr["unresolved"]).sum())
    assert res["pa"] == pytest.approx(w * hits.sum(), rel=1e-12)


def test_c3_ellipsoid2000_cap128(oracle):
    c, r, _ = molecules.workload("ellipsoid2000")
    with GPUScene(c, r) as sc:
        src, tr, ro, hits = _oracle_solve(oracle, c, r, 2, 14, 128)
        for prec in ("fp64", "fp32-fixup"):
            say("C3 2^14 vs oracle", prec)
            res = sc.solve(scramble(2), 14, 128, precision=prec, pa=True, keep_rays=True, histogram=True)
            rays = sc.last_rays(1 << 14)
            np.testing.assert_array_equal(rays["bounces"], tr["bounces"])
            np.testing.assert_array_equal(rays["escaped"], tr["escaped"])
            np.testing.assert_array_equal(rays["unresolved"], tr["unresolved"])
            np.testing.assert_array_equal(rays["pa_hit"], hits)
            np.testing.assert_array_equal(rays["contributions"], ro["contributions"])
            np.testing.assert_array_equal(res["histogram"], ro["histogram"])
            _same_as_oracle(res, ro, tr, hits)
        say("C3 2^22, both kernels")
        a = sc.solve(scramble(2), 22, 128, precision="fp64", pa=True, histogram=True)
        b = sc.solve(scramble(2), 22, 128, precision="fp32-fixup", pa=True, histogram=True)
    for k in PHYSICAL:
        assert a[k] == b[k], k
    np.testing.assert_array_equal(a["histogram"], b["histogram"])
    say("C3 2^22 oracle C solve")
    cpu = oracle.solve_cpu(c, r, 2, 22, 128)
    rel = abs(b["omega"] - cpu["omega"]) / cpu["omega"]
    say(f"  omega {b['omega']:.10f} oracle {cpu['omega']:.10f} rel {rel:.2e}; unresolved {b['unresolved']}, "
        f"tail bound {b['tail_bound']:.4f} A^2, grazing-flagged {b['grazing_rays']}, "
        f"FP64 re-tests / sphere tests {b['fp64_retests'] / max(1, b['sphere_tests']):.4f}")
    assert rel <= 1e-6
    assert b["unresolved"] == cpu["unresolved"]
    assert abs(b["pa"] - cpu["pa"]) <= 1e-6 * cpu["pa"]


def test_c4_batch_set_tuned_plan_under_graph(oracle, tmp_path):
    structs = molecules.batch_set()
    jobs = len(structs)
    assert jobs == 18 and 200 <= min(len(r) for _, r in structs) and max(len(r) for _, r in structs) <= 5000

    def scrambles(k):
        return [scramble(replicate_seed(3 + j, k)) for j in range(jobs)]

    say("C4 tuning")
    tuned, scs, kw = autotune.tune_plan(structs, 20, 64, scrambles(0), repeats=1, passes=1,
                                        cache_path=tmp_path / "tune.json")
    say("  tuned", tuned.config)
    try:
        with GPUPlan(scs, scrambles(0), [20] * jobs, [64] * jobs, graph=True, **kw) as plan:
            for replay in (1, 2):
                for j, sob in enumerate(scrambles(replay)):
                    plan.set_sobol(j, sob)
                say("C4 replay", replay)
                got = plan.run(histogram=True)
                for j, sob in enumerate(scrambles(replay)):
                    alone = scs[j].solve(sob, 20, 64, pa=True, histogram=True)
                    for k in PHYSICAL:
                        assert alone[k] == got[j][k], (replay, j, k)
                    np.testing.assert_array_equal(alone["histogram"], got[j]["histogram"])
                    np.testing.assert_array_equal(alone["ledger"], got[j]["ledger"])
        pick = [0, 9, 11]  # 415, 4145 and 4241 atoms
        with GPUPlan([scs[j] for j in pick], [scramble(50 + j) for j in pick], [14] * 3, [64] * 3, graph=True,
                     **kw) as small:
            got = small.run()
        for k, j in enumerate(pick):
            say("C4 job vs oracle", j, len(structs[j][1]))
            c, r = structs[j]
            src, tr, ro, hits = _oracle_solve(oracle, c, r, 50 + j, 14, 64)
            _same_as_oracle(got[k], ro, tr, hits)
    finally:
        for s in scs:
            s.close()


def test_c5_serving_mix_100_to_20000_atoms(oracle):
    bases = []
    for n, seed in ((100, 31), (450, 32), (1900, 33), (8000, 34), (20000, 35)):
        xyz, el = molecules.globule(n, seed)
        bases.append((xyz, molecules.radii_for(el)))
    reqs = serve.poisson_stream(bases, rate_hz=300.0, count=24, seed=17)
    kw = dict(max_inflight=6, power=12, cap=64, target_rel=0.01, min_reps=4, max_reps=64, rebuild_workers=4)
    srv = serve.Server(1 << 36, **kw)
    say("C5 mix of", len(reqs))
    mixed = {r.rid: r for r in srv.run(reqs)}
    say("  mix", serve.latency_summary(mixed.values()), "built", srv.scenes_built, "rebuilt", srv.scenes_reused)
    assert len(mixed) == len(reqs) and all(r.error is None for r in mixed.values())
    assert srv.scenes_reused > 0 and srv.peak_inflight >= 2
    assert {len(r.radii) for r in reqs} == {100, 450, 1900, 8000, 20000}  # every base is served
    for r in mixed.values():
        assert r.rel_se <= 0.01 or len(r.omegas) == 64
    for req in reqs:
        alone = serve.Server(1 << 36, **kw)
        (res,) = alone.run([serve.Request(req.rid, req.centers, req.radii, 0.0, req.seed)])
        assert res.omegas == mixed[req.rid].omegas, req.rid
        assert res.ccs == mixed[req.rid].ccs
    req = min(reqs, key=lambda q: len(q.radii))
    say("C5 request vs oracle", req.rid, len(req.radii))
    for k in range(4):
        _, tr, ro, _ = _oracle_solve(oracle, req.centers, req.radii, replicate_seed(req.seed, k), 12, 64)
        assert mixed[req.rid].omegas[k] == pytest.approx(ro["omega"], rel=1e-12)

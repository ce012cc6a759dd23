#!/usr/bin/env python3
"""Benchmark of the EHSS/PA hot path on one MI355X.

Workload of one step: a fused device solve of the 650-atom synthetic globular
molecule (random walk, 1.5 A steps, 0.11 atoms/A^3, seed 1): 2^20 scrambled
Sobol rays -> impact-disk source -> BVH any-hit PA -> reflect-until-escape trace
(cap 64) -> FP64 compensated (1 - cos chi) reduction.  The scene and its trees
are resident in HBM before the timed region; each step uses its own replicate
seed.  Steps are issued STEPS_PER_LAUNCH at a time as one multi-replicate launch
(ehss_solve_many), whose per-step totals are bit-identical to separate solves,
with two launches in flight on two streams.  `value` is traced rays per second
of the whole timed region (host wall clock); the rate with one step per launch
is reported beside it.

Side measurements (not part of `value`): the plain FP64 kernel on the same
workload, CCS/s at 1 % relative RMSE (single stream, and with overlapped
solves), a Poisson serving stream (p50/p99 latency), the 2000-atom ellipsoid at
2^22 rays / cap 128 (FP32 error, re-test fraction, unresolved bound), the C60
case, 18 structures of 200-5000 atoms at 2^20 rays each as one autotuned plan
replayed from a hipGraph, and the reference CPU kernel on the host cores.

Prints one JSON line.  --traffic-probe runs only a fixed copy of known size and
a few solves (the command profiled for HBM traffic).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "EHSS traced rays/s + CCS/s at 1% RMSE on 1 MI355X; p50/p99 serve latency"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0
POWER, CAP, SEED = 20, 64, 1
STEPS_PER_LAUNCH = 64     # 2^20-ray steps per launch (2 in flight): 3.19 G rays/s at 32, 3.37 at 64
CCS_PER_LAUNCH = 8192     # 2,048-ray CCS estimates per launch
CCS_STREAMS = 3           # CCS launches in flight, each on its own scene copy and stream
COPY_BYTES = 256 << 20    # calibration copy of the traffic probe (bytes read = bytes written)
TRAFFIC_FILE = ROOT / "profiles" / "traffic_solve_globule650.txt"
VALU_FILE = ROOT / "profiles" / "valu_issue_solve_globule650.md"


def say(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def flops_per_ray(box, sph, refl, pa_box, pa_sph, rays):
    """Algorithmic FLOPs per traced ray of the reference traversal (leaf 4,
    stackless preorder): 12 per box test, 29 per sphere test, 47 per
    reflection, 12 per PA box test, 19 per PA sphere test, 7 for the readout."""
    return (12 * box + 29 * sph + 47 * refl + 12 * pa_box + 19 * pa_sph) / rays + 7


def read_traffic():
    """HBM bytes per launch of the trace kernels from the committed rocprofv3
    counter summary, or None."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def read_valu_issue():
    """Fraction of the packet kernel's time spent issuing VALU instructions,
    from the committed rocprofv3 SQ-counter summary, or None."""
    if not VALU_FILE.exists():
        return None
    for line in VALU_FILE.read_text().splitlines():
        if line.startswith("valu_issue_frac:"):
            return float(line.split(":", 1)[1])
    return None


def rays_per_ccs(scene, scramble, replicate_seed, reps=32, target=0.01):
    """Smallest power-of-two ray count whose relative SD of Omega over
    independent scrambles is <= target (one estimate = one CCS)."""
    rel = math.inf
    for power in range(6, 21):
        vals = [scene.solve(scramble(replicate_seed(100 + k, 0)), power, CAP, pa=False)["omega"]
                for k in range(reps)]
        rel = float(np.std(vals, ddof=1) / np.mean(vals))
        if rel <= target:
            return power, rel
    return 20, rel


def ccs_rate_sequential(scene, seeds, power, precision):
    """One CCS estimate per solve; each scramble is made inside the clock."""
    from ehss_amd.engine import scramble
    t = time.perf_counter()
    for sd in seeds:
        scene.solve(scramble(sd), power, CAP, precision=precision, pa=False)
    return len(seeds) / (time.perf_counter() - t)


def ccs_rate_overlapped(scenes, seeds, power, precision, per_launch=1):
    """CCS estimates with one launch of `per_launch` replicates in flight per
    scene copy (each on its own stream), refilled as they finish.  Each launch's
    scrambles are made (ehss_scramble_many) when it is issued, inside the clock,
    while the other launch runs on the device."""
    from ehss_amd.engine import scramble_many
    chunks = (seeds[k:k + per_launch] for k in range(0, len(seeds), per_launch))
    live = []
    count = 0
    t = time.perf_counter()
    for s in scenes:
        nxt = next(chunks, None)
        if nxt is None:
            break
        s.solve_async(scramble_many(nxt), power, CAP, precision=precision, pa=False)
        live.append(s)
    while live:
        for s in list(live):
            if s.solve_done():
                count += len(s.solve_wait(as_array=True))
                nxt = next(chunks, None)
                if nxt is None:
                    live.remove(s)
                else:
                    s.solve_async(scramble_many(nxt), power, CAP, precision=precision, pa=False)
        time.sleep(2e-5)
    return count / (time.perf_counter() - t)


def ccs_rate_replicated(scene, seeds, power, precision, per_launch):
    """CCS estimates as replicates of one launch (ehss_solve_many); scrambles
    made inside the clock."""
    from ehss_amd.engine import scramble_many
    t = time.perf_counter()
    count = 0
    for k in range(0, len(seeds), per_launch):
        count += len(scene.solve_many(scramble_many(seeds[k:k + per_launch]), power, CAP, precision=precision,
                                      pa=False, as_array=True))
    return count / (time.perf_counter() - t)


def batched18(molecules, scramble, replicate_seed, power=20, cap=64, launches=10):
    """18 globules (log-uniform 200..5000 atoms, seeds 3..20), 2^power rays each.
    Times: the 18 solves one after another (default settings), the plain FP64
    kernel as one plan, the default plan, and the autotuned plan replayed from a
    hipGraph with new scrambles every launch, alone and with two plans in
    flight."""
    import tempfile

    from ehss_amd import autotune
    from ehss_amd.engine import GPUPlan, GPUScene

    structs = molecules.batch_set()
    jobs = len(structs)
    rays = jobs << power

    def scrambles(k):
        return [scramble(replicate_seed(3 + j, k)) for j in range(jobs)]

    def plan_rate(plans, count):
        """Launches of new scrambles (made inside the clock), with len(plans)
        launches in flight, one per plan (each plan has its own stream and
        workspace; they share the resident scenes)."""
        for p in plans:
            p.run()
        t0 = time.perf_counter()
        live = []
        for k in range(count):
            p = plans[k % len(plans)]
            if len(live) == len(plans):
                live.pop(0).wait()
            for j, sob in enumerate(scrambles(k + 1)):
                p.set_sobol(j, sob)
            p.launch()
            live.append(p)
        for p in live:
            p.wait()
        return count * rays / (time.perf_counter() - t0)

    with tempfile.TemporaryDirectory() as tmp:
        tuned, scs, kw = autotune.tune_plan(structs, power, cap, scrambles(0), repeats=2, passes=2,
                                            cache_path=Path(tmp) / "tune.json")
    try:
        t0 = time.perf_counter()
        for k in range(2):
            for j, s in enumerate(scs):
                s.solve(scramble(replicate_seed(3 + j, k)), power, cap, pa=True)
        seq = 2 * rays / (time.perf_counter() - t0)
        with GPUPlan(scs, scrambles(0), [power] * jobs, [cap] * jobs, precision="fp64", graph=True) as plan:
            plain = plan_rate([plan], 2)
        with GPUPlan(scs, scrambles(0), [power] * jobs, [cap] * jobs, graph=True) as plan:
            default = plan_rate([plan], launches)
        with GPUPlan(scs, scrambles(0), [power] * jobs, [cap] * jobs, graph=True, **kw) as plan, \
                GPUPlan(scs, scrambles(0), [power] * jobs, [cap] * jobs, graph=True, **kw) as plan2:
            best = plan_rate([plan], launches)
            best2 = plan_rate([plan, plan2], launches)
            res = plan.run()
        return dict(
            structures=jobs, atoms=[len(r) for _, r in structs], rays_per_structure=1 << power, cap=cap,
            rays_per_launch=rays, tuned_config=tuned.config,
            tuner_trials=[dict(config=c, seconds=round(t, 5)) for c, t in tuned.trials],
            tuner_speedup_vs_default=tuned.speedup,
            rays_per_s_sequential_solves=seq, rays_per_s_plain_fp64_plan=plain,
            rays_per_s_default_plan_graph=default, rays_per_s_tuned_plan_graph=best,
            rays_per_s_tuned_plan_graph_2_in_flight=best2,
            tuned_plan_vs_plain_fp64_plan=max(best, best2) / plain,
            tuned_plan_vs_sequential=max(best, best2) / seq,
            ms_per_launch=1e3 * rays / best, omega_A2=[r["omega"] for r in res],
            unresolved_rays=int(sum(r["unresolved"] for r in res)))
    finally:
        for s in scs:
            s.close()


def cpu_host_info() -> dict:
    """The host CPU as lscpu reports it (model name, family/model/stepping,
    sockets, cores, clock), with /proc/cpuinfo's model name as a fallback."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Vendor ID": "vendor", "CPU family": "family", "Model": "model_number",
                "Stepping": "stepping", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "CPU max MHz": "max_mhz",
                "Hypervisor vendor": "hypervisor"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and v.strip():
                info.setdefault(keys[k.strip()], v.strip())
    except (OSError, subprocess.SubprocessError):
        pass
    if "model" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            info["model"] = "unknown"
    return info


# Per-config CPU samples: (config, workload, seeds, 2^power rays per seed, cap).
# C1 is the whole configuration; C2 and C3 are bounded samples of theirs.
CPU_CONFIGS = (("C1", "c60", (0, 104729, 209458), 16, 64),
               ("C2", "globule650", (SEED,), 21, 64),
               ("C3", "ellipsoid2000", (2,), 20, 128))
CPU_RUNS = 5


def cpu_baseline(molecules):
    """CPU baselines per config on the host cores, median of CPU_RUNS warm runs:
    the reference's own ehss_trace (src/native/ehss_cpu.cpp built from the
    reference sources into oracle/_ref) on the NumPy-path source of the same
    points (trace only), and the project's C restatement of the whole path
    (Sobol, source, PA, trees, trace, readout; kind "port").  The headline
    cpu_baseline value is C2's reference trace rate."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    host = cpu_host_info()
    host.update(cores_used=cores)
    ref = O.ref_lib()
    per = {}

    def median_time(fn):
        fn()  # warm
        ts = []
        for _ in range(CPU_RUNS):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    for name, wl, seeds, power, cap in CPU_CONFIGS:
        c, r, _ = molecules.workload(wl)
        rays = len(seeds) << power
        tree = O.static_tree(c, r, 4)
        row = dict(workload=wl, rays=rays, seeds=list(seeds), cap=cap, runs=CPU_RUNS, threads=cores)
        if ref is not None:
            srcs = [O.external_source(c, r, power, sd) for sd in seeds]

            def trace_all(threads=cores, m=None):
                for src in srcs:
                    O.trace(tree, c, r, src["origins"][:m], src["directions"][:m], cap=cap, record=False,
                            lib_handle=ref, threads=threads)
            row["reference_trace_rays_per_s"] = rays / median_time(trace_all)
            if name == "C2":
                m = 1 << (power - 4)
                row["reference_trace_rays_per_s_one_core"] = m / median_time(lambda: trace_all(1, m))
        row["whole_path_rays_per_s"] = rays / median_time(
            lambda: [O.solve_cpu(c, r, sd, power, cap, threads=cores) for sd in seeds])
        per[name] = row
        say(f"cpu {name}: {row}")
    c2 = per["C2"]
    if ref is not None:
        value, kind = c2["reference_trace_rays_per_s"], "reference"
        sample = (f"reference ehss_trace (src/native/ehss_cpu.cpp, g++ -O2 -fopenmp) on the first 2^21 rays of "
                  f"the globule650 net (seed {SEED}, cap 64), trace only (source, PA and readout excluded), "
                  f"median of {CPU_RUNS} warm runs with {cores} OpenMP threads; per_config has C1 (the whole "
                  f"config, 3 x 2^16 rays) and C3 (2^20 rays, cap 128) and the whole-path rates")
    else:
        value, kind = c2["whole_path_rays_per_s"], "port"
        sample = f"oracle C restatement of the whole path on 2^21 rays, median of {CPU_RUNS}"
    return dict(value=value, unit="rays/s", cores=cores, kind=kind, sample=sample, host=host, per_config=per)


def seam_rates(scene, power=20, runs=3):
    """PCIe-inclusive rates of the per-ray seams, on the first 2^power rays of
    the workload's net (host arrays in, host arrays out): ehss_trace (the
    reference's C signature: tree and rays uploaded, plain FP64 trace, outputs
    downloaded, every call), ehss_trace_status in FP32+fixup, and
    ehss_scene_trace against the resident scene."""
    import ctypes as ct
    from ehss_amd import _lib as L
    from ehss_amd.engine import scramble
    src = scene.source_rays(scramble(SEED), power)
    tree = scene.export_tree()
    c = np.ascontiguousarray(scene.centers, dtype=np.float64)
    r = np.ascontiguousarray(scene.radii, dtype=np.float64)
    n = 1 << power
    nn = len(tree["end"])

    def seam(precision):
        p, d = src["origins"].copy(), src["directions"].copy()
        lasts = np.full(n, -1, np.int64)
        counts = np.zeros(n, np.int64)
        esc, unr = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        ctr = np.zeros(2, np.int64)
        arrays = [p, d, lasts, counts, esc, unr, None, tree["lo"], tree["hi"], tree["end"], tree["leaf_group"],
                  tree["offsets"], tree["ids"], c, r, ctr]
        if precision is None:
            obs = ct.c_int(0)
            L.fn("ehss_trace")(n, nn, CAP, 16, 0, *(L.ptr(a) for a in arrays), ct.byref(obs))
            if obs.value < 0:
                raise RuntimeError(L.last_error())
        else:
            L.call("ehss_trace_status", n, nn, CAP, L.PRECISIONS[precision], 0, *(L.ptr(a) for a in arrays))

    def med(fn):
        fn()
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return n / float(np.median(ts))

    return dict(rays=n, bytes_per_ray_each_way=2 * 24 + 8 + 8 + 2,
                ehss_trace_fp64_rays_per_s=med(lambda: seam(None)),
                ehss_trace_status_fp32_fixup_rays_per_s=med(lambda: seam("fp32-fixup")),
                scene_trace_fp32_fixup_rays_per_s=med(lambda: scene.trace(src["origins"], src["directions"], cap=CAP)),
                note="host arrays in and out: PCIe transfers, uploads and the final synchronisation included")


def traffic_probe(args):
    """The command profiled for HBM traffic: one copy of COPY_BYTES (the
    calibration kernel), then a few solves of the workload."""
    import torch
    from ehss_amd import molecules
    from ehss_amd.engine import GPUScene, scramble, replicate_seed
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    centers, radii, _ = molecules.workload("globule650")
    with GPUScene(centers, radii) as scene:
        for k in range(args.steps):
            scene.solve_many([scramble(replicate_seed(SEED, STEPS_PER_LAUNCH * k + j))
                              for j in range(STEPS_PER_LAUNCH)], POWER, CAP, precision=args.precision, pa=True)
    say("traffic probe done")
    return 0


def ccs_probe(args):
    """The command profiled for the CCS-size launch: `--steps` rounds of one
    launch of CCS_PER_LAUNCH replicates of 2^11 rays (cap 64, no PA) in each of
    four forms: FP32+fixup sorted, FP32+fixup unsorted (EHSS_FLAG_UNSORTED),
    plain FP64 sorted, plain FP64 unsorted."""
    from ehss_amd import molecules
    from ehss_amd.engine import GPUScene, scramble_many, replicate_seed
    centers, radii, _ = molecules.workload("globule650")
    seeds = np.array([replicate_seed(7, k) for k in range(CCS_PER_LAUNCH)], dtype=np.uint64)
    with GPUScene(centers, radii) as scene:
        for _ in range(args.steps):
            for precision in ("fp32-fixup", "fp64"):
                for unsorted in (False, True):
                    t = time.perf_counter()
                    scene.solve_many(scramble_many(seeds), 11, CAP, precision=precision, pa=False, as_array=True,
                                     unsorted=unsorted)
                    say(f"ccs probe {precision} unsorted={unsorted}: {1e3 * (time.perf_counter() - t):.2f} ms")
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1024)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp32-fixup", choices=["fp32-fixup", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the side measurements")
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--ccs-probe", action="store_true")
    ap.add_argument("--hw-queues", type=int, default=12,
                    help="GPU_MAX_HW_QUEUES for this process (1..16)")
    args = ap.parse_args()
    if not 1 <= args.hw_queues <= 16:
        ap.error("--hw-queues must be within 1..16")
    # Hardware queues per process, read once when HIP initialises (nothing has
    # yet): the serving stream keeps up to 24 requests on their own streams, and
    # HIP's default of 4 queues runs their small launches four at a time.  12
    # measured best for the serving mix; 32 thrash.  Single-stream measurements
    # do not depend on it.
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the EHSS/PA path runs on a single MI355X here and its "
              "multi-GPU sharding is a design note only; run with --gpus 1.", file=sys.stderr)
        return 2
    if args.traffic_probe:
        return traffic_probe(args)
    if args.ccs_probe:
        return ccs_probe(args)

    from ehss_amd import _lib as L
    from ehss_amd import molecules, serve
    from ehss_amd.engine import GPUScene, scramble, replicate_seed

    centers, radii, meta = molecules.workload("globule650")
    dev = L.device_info()
    say(f"device {dev['name']} CUs {dev['compute_units']}")
    scene = GPUScene(centers, radii)
    # Warm-up launches are whole launches, at least one per stream, so that no
    # workspace grows inside the timed region.
    warm = max(args.warmup, 2 * STEPS_PER_LAUNCH)
    n = 1 << POWER

    # Warmup, then exactly K timed steps in launches of STEPS_PER_LAUNCH (the
    # last one takes the remainder), two launches in flight on two streams,
    # each over its own resident copy of the scene (the long paths of one
    # launch overlap the packet phase of the next).  The region ends when the
    # last launch has been collected, i.e. after device synchronisation.
    from ehss_amd.engine import sobol_table
    scene2 = GPUScene(centers, radii)

    def batches(first, count):
        """Each launch's scrambles (host PCG64 + LMS, ehss_scramble) and launch
        table, built when the launch is issued: inside the timed region."""
        for b in range(first, first + count, STEPS_PER_LAUNCH):
            yield sobol_table([scramble(replicate_seed(SEED, k)) for k in range(b, min(b + STEPS_PER_LAUNCH,
                                                                                          first + count))])

    def run_overlapped(tabs):
        pending, it, out = [], iter(tabs), None
        for sc in (scene, scene2):
            tb = next(it, None)
            if tb is None:
                break
            sc.solve_async(tb, POWER, CAP, precision=args.precision, pa=True)
            pending.append(sc)
        while pending:
            sc = pending.pop(0)
            out = sc.solve_wait()
            tb = next(it, None)
            if tb is not None:
                sc.solve_async(tb, POWER, CAP, precision=args.precision, pa=True)
                pending.append(sc)
        return out

    run_overlapped(batches(0, warm))
    t0 = time.perf_counter()
    res = run_overlapped(batches(warm, args.steps))[-1]
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n * args.steps / elapsed
    say(f"{args.steps} steps, {ms_per_step:.3f} ms/step, {value:.4g} rays/s")
    scene2.close()

    # Kernel time per launch for the roofline: the same launches one at a time
    # on one stream, HIP events around the trace kernels.
    timed = list(batches(warm, args.steps))
    sobols = [scramble(replicate_seed(SEED, k)) for k in range(max(20, 2 * STEPS_PER_LAUNCH))]
    kernel_ms, launch_rays = [], []
    for tb in timed:
        out = scene.solve_many(tb, POWER, CAP, precision=args.precision, pa=True, as_array=True)
        kernel_ms.append(float(out["kernel_ms"][0]))
        launch_rays.append(len(tb) * n)
    t1 = time.perf_counter()
    for k in range(20):
        scene.solve(sobols[k], POWER, CAP, precision=args.precision, pa=True)
    one_per_launch = 20 * n / (time.perf_counter() - t1)

    # Algorithmic work per ray from the plain FP64 kernel, whose box/sphere
    # counters equal the reference traversal's (pinned by the parity tests).
    plain = scene.solve(sobols[0], POWER, CAP, precision="fp64", pa=True, histogram=True)
    refl = float(np.dot(np.arange(CAP + 1), plain["histogram"]))
    fpr = flops_per_ray(plain["box_tests"], plain["sphere_tests"], refl, plain["pa_box_tests"],
                        plain["pa_sphere_tests"], n)
    # Per launch of STEPS_PER_LAUNCH steps (HIP events on the scene's stream
    # around the trace kernels; a shorter remainder launch is left out).
    rays_per_launch = max(launch_rays)
    full = [ms for ms, r in zip(kernel_ms, launch_rays) if r == rays_per_launch]
    avg_kernel_s = float(np.mean(full)) / 1e3
    achieved = fpr * rays_per_launch / avg_kernel_s / 1e12
    kernels = ("k_solve_x<4,1,0> + 2 x k_drain<1>" if args.precision == "fp32-fixup" else "k_solve_x<0,1,0>")
    roofline = dict(bound="valu", achieved=round(achieved, 4), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP32_PEAK_TFLOPS, 6), traffic=read_traffic(),
                    kernel=kernels, kernel_ms=round(float(np.mean(full)), 4), rays_per_launch=rays_per_launch,
                    flops_per_ray=round(fpr, 2),
                    per_ray=dict(box=plain["box_tests"] / n, sphere=plain["sphere_tests"] / n,
                                 reflections=refl / n, pa_box=plain["pa_box_tests"] / n,
                                 pa_sphere=plain["pa_sphere_tests"] / n))
    roofline["valu_issue_frac_profiled"] = read_valu_issue()
    if roofline["traffic"]:
        roofline["traffic_gbs"] = round(roofline["traffic"] / avg_kernel_s / 1e9, 2)
        roofline["traffic_frac_of_hbm_peak"] = round(roofline["traffic_gbs"] / HBM_PEAK_GBS, 6)

    extras = {}
    if not args.quick:
        say("plain FP64 kernel")
        t1 = time.perf_counter()
        for k in range(20):
            scene.solve(sobols[k], POWER, CAP, precision="fp64", pa=True)
        plain_rate = 20 * n / (time.perf_counter() - t1)
        t1 = time.perf_counter()
        for k in range(2):
            scene.solve_many(sobols[STEPS_PER_LAUNCH * k:STEPS_PER_LAUNCH * (k + 1)], POWER, CAP, precision="fp64",
                             pa=True)
        plain_batched = 2 * STEPS_PER_LAUNCH * n / (time.perf_counter() - t1)
        say("CCS at 1% RMSE")
        power_ccs, rel = rays_per_ccs(scene, scramble, replicate_seed)
        per = CCS_PER_LAUNCH
        ccs_seeds = np.array([replicate_seed(7, k) for k in range(12 * per)], dtype=np.uint64)
        seq_fast = ccs_rate_sequential(scene, ccs_seeds[:400], power_ccs, "fp32-fixup")
        seq_plain = ccs_rate_sequential(scene, ccs_seeds[:100], power_ccs, "fp64")
        ccs_rate_replicated(scene, ccs_seeds[:per], power_ccs, "fp32-fixup", per)
        rep_fast = ccs_rate_replicated(scene, ccs_seeds[:4 * per], power_ccs, "fp32-fixup", per)
        rep_plain = ccs_rate_replicated(scene, ccs_seeds[:per], power_ccs, "fp64", per)
        # Launches in flight on their own streams, over 12 launches (6 for the
        # plain kernel) so that the first and last launches' ramps weigh little.
        copies = [GPUScene(centers, radii) for _ in range(CCS_STREAMS)]
        ccs_rate_overlapped(copies, ccs_seeds[:CCS_STREAMS * per], power_ccs, "fp32-fixup", per)
        ovl_fast = ccs_rate_overlapped(copies, ccs_seeds, power_ccs, "fp32-fixup", per)
        ovl_plain = ccs_rate_overlapped(copies, ccs_seeds[:6 * per], power_ccs, "fp64", per)
        for c in copies:
            c.close()
        extras["globule650"] = dict(
            steps_per_launch=STEPS_PER_LAUNCH, rays_per_s_one_step_per_launch=one_per_launch,
            plain_fp64_rays_per_s=plain_rate, plain_fp64_rays_per_s_same_launches=plain_batched,
            speedup_vs_plain_fp64=value / plain_rate, speedup_vs_plain_fp64_same_launches=value / plain_batched,
            omega_A2=res["omega"], pa_A2=res["pa"], fp64_retests_per_ray=res["fp64_retests"] / n,
            grazing_rays=res["grazing_rays"], unresolved_rays=res["unresolved"], tail_bound_A2=res["tail_bound"],
            omega_fp32_fixup_minus_fp64=scene.solve(sobols[0], POWER, CAP, pa=True)["omega"] - plain["omega"])
        best_fast = max(rep_fast, ovl_fast)
        extras["ccs_1pct"] = dict(
            rays_per_ccs=1 << power_ccs, rel_sd_at_that_count=rel,
            ccs_per_s_plain_fp64_single_solves=seq_plain, ccs_per_s_fp32_fixup_single_solves=seq_fast,
            replicates_per_launch=per,
            ccs_per_s_plain_fp64_one_launch=rep_plain, ccs_per_s_fp32_fixup_one_launch=rep_fast,
            streams=CCS_STREAMS, ccs_per_s_plain_fp64_streams=ovl_plain, ccs_per_s_fp32_fixup_streams=ovl_fast,
            batched_fp32_fixup_vs_plain_fp64_kernel=best_fast / seq_plain,
            batched_fp32_fixup_vs_batched_plain_fp64=best_fast / max(rep_plain, ovl_plain),
            rays_per_s_at_ccs_size=best_fast * (1 << power_ccs))
        say("per-ray seams (PCIe-inclusive)")
        extras["seam"] = seam_rates(scene)
        scene.close()

        say("ellipsoid2000, 2^22 rays, cap 128")
        c3, r3, _ = molecules.workload("ellipsoid2000")
        with GPUScene(c3, r3) as s3:
            a = s3.solve(scramble(2), 22, 128, precision="fp64", pa=True)
            b = s3.solve(scramble(2), 22, 128, precision="fp32-fixup", pa=True)
            t = time.perf_counter()
            for k in range(5):
                s3.solve(scramble(replicate_seed(2, k)), 22, 128, pa=True)
            rate3 = 5 * (1 << 22) / (time.perf_counter() - t)
        extras["ellipsoid2000"] = dict(
            rays=1 << 22, cap=128, omega_fp64=a["omega"], omega_fp32_fixup=b["omega"],
            fp32_rel_error_vs_fp64=abs(b["omega"] - a["omega"]) / a["omega"],
            fp64_retest_fraction=b["fp64_retests"] / max(1, b["sphere_tests"]),
            fp64_retests_per_ray=b["fp64_retests"] / (1 << 22), grazing_rays=b["grazing_rays"],
            unresolved_rays=b["unresolved"], unresolved_bound_A2=b["tail_bound"], rays_per_s=rate3)

        say("C60, 3 x 65536 rays")
        c1, r1, _ = molecules.workload("c60")
        with GPUScene(c1, r1) as s1:
            vals = [s1.solve(scramble(replicate_seed(0, k)), 16, 64, pa=True) for k in range(3)]
        extras["c60"] = dict(omega_mean=float(np.mean([v["omega"] for v in vals])),
                             pa_mean=float(np.mean([v["pa"] for v in vals])),
                             ms_per_repeat=float(np.mean([v["kernel_ms"] for v in vals])))

        say("alternative sources on the device (globule650, 2^20 states, cap 64)")
        from ehss_amd.engine import scramble5, source_params
        lo_b = np.min(centers - radii[:, None], axis=0)
        hi_b = np.max(centers + radii[:, None], axis=0)
        w2 = radii ** 2
        alt = {}
        with GPUScene(centers, radii) as s_alt:
            for name, params, dim in (
                    ("box", source_params("box", (lo_b + hi_b) / 2, np.eye(3), (hi_b - lo_b) / 2 * (1 + 1e-9)), 5),
                    ("reverse_pa", source_params("reverse-pa", area=float(4 * np.pi * np.dot(radii, radii)),
                                                 cdf=np.cumsum(w2) / w2.sum()), 5)):
                s_alt.solve_source(params, scramble5(1, dim), POWER, CAP)  # warm-up (workspace)
                t = time.perf_counter()
                outs = [s_alt.solve_source(params, scramble5(1 + k, dim), POWER, CAP) for k in range(5)]
                el = time.perf_counter() - t
                alt[name] = dict(states_per_s=5 * (1 << POWER) / el, omega_mean=float(np.mean([o["omega"] for o in outs])),
                                 kernel_ms=float(np.mean([o["kernel_ms"] for o in outs])),
                                 pa=float(np.mean([o["pa"] for o in outs])) if name == "reverse_pa" else None)
        extras["alt_sources"] = alt

        say("C4: 18 structures x 2^20 rays, autotuned plan under hipGraph")
        extras["batched18"] = batched18(molecules, scramble, replicate_seed)

        say("serving stream")
        rng = np.random.default_rng(2024)
        sizes = np.exp(rng.uniform(math.log(100), math.log(20000), size=6)).astype(int)
        bases = []
        for k, na in enumerate(sizes):
            xyz, el = molecules.globule(int(na), 500 + k)
            bases.append((xyz, molecules.radii_for(el)))
        runs = {}
        # 5000 Hz offers more than the loop serves: its CCS/s is the capacity.
        for rate, count in ((200.0, 120), (2000.0, 600), (5000.0, 1500)):
            reqs = serve.poisson_stream(bases, rate_hz=rate, count=count, seed=9)
            srv = serve.Server(int(0.8 * dev["total_mem"]), max_inflight=24, power=12, cap=64, target_rel=0.01,
                               min_reps=4, max_reps=64, warm_atoms=int(sizes.max()))
            results = srv.run(reqs)
            runs[rate] = dict(serve.latency_summary(results), rate_hz=rate, requests=len(reqs),
                              peak_inflight=srv.peak_inflight, peak_admitted_bytes=srv.peak_bytes,
                              scenes_built=srv.scenes_built, scenes_rebuilt=srv.scenes_reused)
        extras["serving"] = dict(runs[200.0], base_atoms=[int(s) for s in sizes], rays_per_replicate=1 << 12,
                                 hw_queues=int(os.environ["GPU_MAX_HW_QUEUES"]), warm_pool_atoms=int(sizes.max()),
                                 saturated_2000_hz=runs[2000.0], capacity_5000_hz=runs[5000.0])
    else:
        scene.close()

    cpu = None if args.no_cpu_baseline else cpu_baseline(molecules)
    if cpu is not None:
        ratios = dict(rays_per_s_vs_cpu=value / cpu["value"])
        if "ccs_1pct" in extras:
            cc = extras["ccs_1pct"]
            cpu_ccs = cpu["value"] / cc["rays_per_ccs"]
            best = max(cc["ccs_per_s_fp32_fixup_one_launch"], cc["ccs_per_s_fp32_fixup_streams"])
            ratios.update(cpu_ccs_per_s=cpu_ccs, ccs_per_s_vs_cpu=best / cpu_ccs)
        extras["vs_cpu_baseline"] = ratios
    line = dict(metric=METRIC, value=value, unit="rays/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
                ms_per_step=ms_per_step, higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="f32+f64" if args.precision == "fp32-fixup" else "f64", data="synthetic",
                config=dict(workload="globule650: 650-atom synthetic globular molecule (seed 1), 2^20 scrambled-Sobol "
                                     "rays per step, cap 64, BVH any-hit PA, fused device solve",
                            atoms=len(radii), rays_per_step=n, cap=CAP, precision=args.precision,
                            steps_per_launch=STEPS_PER_LAUNCH, launches_in_flight=2,
                            parallelism="single GPU"),
                roofline=roofline, cpu_baseline=cpu, device=dev["name"], **extras)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

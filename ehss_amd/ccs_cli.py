"""`ccs.py` front end of the MI355X engine: PA and EHSS collision cross sections
of molecules in standard XYZ files.

Interface kept from the reference CLI (ccs_cli.py:235-358; docs/CCS_CLI.md):
the command-line flags and their validation limits, the replicate seeds
(seed + 104729 r) mod 2^32, the output directory with results.json (schema v1)
and summary.csv, and the exit codes (2 for invalid input, 130 when
interrupted).  Added: --backend (gpu) and --precision (fp32-fixup, the default,
or the plain fp64 kernel; both give the same results).

The work is organised for the device rather than per replicate:
  * every frame becomes one device-resident scene (trees built on the GPU);
  * all replicates of a frame run as ONE multi-replicate launch
    (ehss_solve_many), whose per-replicate totals, histograms and ledgers are
    bit-identical to separate solves; source, PA, trace and reduction are
    fused on the device, so no per-ray data crosses PCIe;
  * the finite-grid (adaptive) PA runs per replicate on the device
    (ehss_pa_grid).
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import math
import platform
import sys
import time
from collections import Counter
from dataclasses import dataclass, field
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

VERSION = "0.2.0"
TOOL = "ccs-xyz"
EXIT_INVALID, EXIT_INTERRUPTED = 2, 130

# Van der Waals radii (A), Mantina et al. (2009), Table 12, DOI
# 10.1021/jp8111556: the published values the reference's table uses
# (ccs_cli.py:26-31).  A probe radius is added to each.
MANTINA_RADII = (("H", 1.10), ("He", 1.40), ("C", 1.70), ("N", 1.55), ("O", 1.52), ("F", 1.47),
                 ("Si", 2.10), ("P", 1.80), ("S", 1.80), ("Cl", 1.75), ("Se", 1.90), ("Br", 1.83),
                 ("I", 1.98))
RADII_SOURCE = ("Mantina et al. 2009 Table 12, subset; DOI 10.1021/jp8111556; user overrides recorded in table")

# Power-of-two options and their ranges; positive integer options; other limits.
POW2_LIMITS = {"rays": (8, 1 << 22), "directions": (1, 1 << 16), "grid_size": (8, 8192)}
POSITIVE = ("repeats", "threads", "leaf_size", "max_collisions", "frame")
MAX_COLLISIONS = 100000


# ---------------------------------------------------------------------------
# Input: XYZ frames
# ---------------------------------------------------------------------------

@dataclass
class XYZFrame:
    number: int          # 1-based position in the file
    comment: str
    elements: tuple      # normalised symbols ("Cl", not "CL")
    coordinates: np.ndarray  # (n, 3) float64, file units


def element_symbol(token) -> str:
    """A one- or two-letter element symbol, normalised to title case."""
    if isinstance(token, str) and 1 <= len(token) <= 2 and token.isascii() and token.isalpha():
        return token.capitalize()
    raise ValueError(f"Expected an element symbol, got {token!r}")


def _number(token: str) -> float:
    """A coordinate; Fortran exponents (1D0, 2.5d-1) are accepted."""
    return float(token.translate(str.maketrans("Dd", "Ee")))


class _Lines:
    """Cursor over the lines of one file, for error messages with line numbers."""

    def __init__(self, path):
        self.path = path
        self.lines = Path(path).read_text(encoding="utf-8-sig").splitlines()
        self.pos = 0

    def fail(self, index: int, what: str):
        raise ValueError(f"{self.path}:{index + 1}: {what}")

    def skip_blank(self) -> bool:
        while self.pos < len(self.lines) and not self.lines[self.pos].strip():
            self.pos += 1
        return self.pos < len(self.lines)

    def frame(self, number: int) -> XYZFrame:
        head = self.pos
        token = self.lines[head].strip()
        if not (token.isascii() and token.isdigit() and token[0] != "0"):
            self.fail(head, "expected a positive atom count")
        count = int(token)
        if head + 1 + count >= len(self.lines):
            self.fail(head, f"incomplete XYZ frame (comment and {count} atom rows required)")
        comment = self.lines[head + 1]
        if "Lattice=" in comment or "Properties=" in comment:
            self.fail(head + 1, "extended/periodic XYZ is not supported; export a standard isolated molecule")
        symbols, xyz = [], np.empty((count, 3))
        for k in range(count):
            row = head + 2 + k
            fields = self.lines[row].split()
            if len(fields) != 4:
                self.fail(row, "expected exactly 'Element x y z'")
            try:
                symbols.append(element_symbol(fields[0]))
                xyz[k] = [_number(t) for t in fields[1:]]
            except ValueError:
                self.fail(row, "invalid element or coordinates")
            if not np.all(np.isfinite(xyz[k])):
                self.fail(row, "coordinates must be finite")
        self.pos = head + 2 + count
        return XYZFrame(number, comment, tuple(symbols), xyz)


def read_xyz(path) -> list[XYZFrame]:
    """Every frame of a standard XYZ file (a malformed later frame is an error
    even when only the first is used)."""
    cur = _Lines(path)
    frames = []
    while cur.skip_blank():
        frames.append(cur.frame(len(frames) + 1))
    if not frames:
        raise ValueError(f"{path}: no XYZ frames")
    return frames


# ---------------------------------------------------------------------------
# Radii and frame preparation
# ---------------------------------------------------------------------------

def radius_table(path=None) -> dict:
    """Atomic radii (A): the built-in table, with overrides or additions from a
    JSON object {"C": 1.70, ...} of positive finite numbers."""
    table = dict(MANTINA_RADII)
    if not path:
        return table
    raw = json.loads(Path(path).read_text(encoding="utf-8-sig"))
    if not isinstance(raw, dict) or not raw:
        raise ValueError("Radii JSON must be a nonempty object: {\"C\": 1.70, ...}")
    given = {}
    for key, value in raw.items():
        symbol = element_symbol(key)
        if symbol in given:
            raise ValueError(f"Duplicate normalized radius key: {symbol}")
        ok = isinstance(value, (int, float)) and not isinstance(value, bool) and math.isfinite(value) and value > 0
        if not ok:
            raise ValueError(f"Radius for {symbol} must be a positive finite number in angstrom")
        given[symbol] = float(value)
    table.update(given)
    return table


def prepare_frame(frame: XYZFrame, radii: dict, probe_radius: float, units: str = "angstrom",
                  exclude_hydrogen: bool = False):
    """Spheres of the selected atoms (radius + probe), recentred, plus the
    record's input description."""
    from .state import Spheres
    rows = np.array([i for i, e in enumerate(frame.elements) if not (exclude_hydrogen and e == "H")], dtype=int)
    if rows.size == 0:
        raise ValueError(f"Frame {frame.number}: atom selection is empty")
    elements = [frame.elements[i] for i in rows]
    missing = sorted({e for e in elements if e not in radii})
    if missing:
        raise ValueError(f"Frame {frame.number}: no radius for {', '.join(missing)}; supply --radii-json")
    xyz = frame.coordinates[rows] * (10.0 if units == "nm" else 1.0)
    if not np.all(np.isfinite(xyz)):
        raise ValueError("Coordinates overflow after unit conversion")
    order = np.lexsort(xyz.T[::-1])
    if np.any(np.all(np.diff(xyz[order], axis=0) == 0, axis=1)):
        raise ValueError(f"Frame {frame.number}: coincident selected atom centers; check duplicate atoms")
    # Recentring offset: first atom plus the mean displacement from it, the
    # expression of ccs_cli.py:116 -- it fixes the bits of every centre, and so
    # of the results, for a given input.
    offset = xyz[0] + np.mean(xyz - xyz[0], axis=0)
    effective = {e: radii[e] + probe_radius for e in set(elements)}
    spheres = Spheres(xyz - offset, np.array([effective[e] for e in elements]), deduplicate=False)
    info = dict(frame=frame.number, comment=frame.comment, input_atoms=len(frame.elements),
                selected_atoms=int(rows.size), selected_atom_rows_1based=(rows + 1).tolist(),
                element_counts=dict(Counter(elements)), recenter_offset_angstrom=offset.tolist(),
                effective_radii_angstrom={e: effective[e] for e in sorted(effective)})
    return spheres, info


# ---------------------------------------------------------------------------
# Device solve of one frame
# ---------------------------------------------------------------------------

def _stats(values) -> dict:
    """Mean, sample SD and SEM over replicates (SD/SEM null for one)."""
    v = np.asarray(values, dtype=np.float64)
    if len(v) < 2:
        return dict(mean=float(v.mean()), sd=None, sem=None, replicates=len(v))
    sd = float(np.std(v, ddof=1))
    return dict(mean=float(v.mean()), sd=sd, sem=sd / math.sqrt(len(v)), replicates=len(v))


def _ehss_row(res: dict, seconds: float) -> dict:
    omega, tail = res["omega"], res["tail_bound"]
    return dict(ehss_angstrom2=omega, pa_from_same_rays_angstrom2=res["pa_from_ehss"], tail_bound_angstrom2=tail,
                sampled_interval_angstrom2=[omega, omega + tail], unresolved_rays=int(res["unresolved"]),
                query_s=seconds, readout_s=0.0, collision_histogram=[int(x) for x in res["histogram"]],
                collision_contribution_angstrom2=[float(x) for x in res["ledger"]],
                box_tests=int(res["box_tests"]), sphere_tests=int(res["sphere_tests"]),
                device_kernel_ms=res["kernel_ms"], fp64_retests=int(res["fp64_retests"]),
                grazing_rays=int(res["grazing_rays"]))


def solve_frame(spheres, args, progress: bool = True) -> dict:
    """The replicates of one frame (the role of ccs_cli.py:170-232): one device
    scene, all EHSS / BVH-PA replicates in one launch, the finite-grid PA per
    replicate.  Returns {summary, runs, timings}."""
    from .adaptive_pa import adaptive_pa
    from .engine import GPUScene, replicate_seed, scramble_many
    want_ehss = args.method in ("ehss", "both")
    want_pa = args.method in ("pa", "both")
    grid_pa = want_pa and args.pa_method == "adaptive"
    bvh_pa = want_pa and not grid_pa
    seeds = [replicate_seed(args.seed, r) for r in range(args.repeats)]
    power = args.rays.bit_length() - 1
    t_start = time.perf_counter()
    runs = [dict(repeat=r + 1, seed=s) for r, s in enumerate(seeds)]
    setup_s = 0.0
    if want_ehss or bvh_pa:
        t = time.perf_counter()
        scene = GPUScene(spheres.centers, spheres.radii, leaf_size=args.leaf_size)
        setup_s = time.perf_counter() - t
        try:
            if progress:
                print(f"  {len(seeds)} replicate(s) of {args.rays} rays in one device launch", file=sys.stderr,
                      flush=True)
            t = time.perf_counter()
            out = scene.solve_many(scramble_many(np.array(seeds, dtype=np.uint64)), power, args.max_collisions,
                                   precision=args.precision, pa=bvh_pa, histogram=True)
            share = (time.perf_counter() - t) / len(seeds)  # one launch: its time split evenly
        finally:
            scene.close()
        for run, res in zip(runs, out):
            run["impact_source_s"] = 0.0  # generated inside the fused device solve
            if bvh_pa:
                run["pa"] = dict(pa_angstrom2=res["pa"], query_s=share, readout_s=0.0,
                                 box_tests=int(res["pa_box_tests"]), sphere_tests=int(res["pa_sphere_tests"]))
            if want_ehss:
                run["ehss"] = _ehss_row(res, share)
                if progress and res["unresolved"]:
                    print(f"  WARNING: replicate {run['repeat']}: {int(res['unresolved'])} unresolved rays; omitted "
                          f"contribution <= {res['tail_bound']:.6g} A^2. Increase --max-collisions.",
                          file=sys.stderr)
    if grid_pa:
        for run in runs:
            run["pa"] = adaptive_pa(spheres, args.directions, args.grid_size, run["seed"])
    summary = {}
    if want_pa:
        summary["pa_angstrom2"] = _stats([r["pa"]["pa_angstrom2"] for r in runs])
    if want_ehss:
        summary["ehss_angstrom2"] = _stats([r["ehss"]["ehss_angstrom2"] for r in runs])
        summary["ehss_tail_bound_mean_angstrom2"] = float(np.mean([r["ehss"]["tail_bound_angstrom2"] for r in runs]))
        summary["ehss_unresolved_rays_total"] = int(sum(r["ehss"]["unresolved_rays"] for r in runs))
    timings = dict(ehss_setup_s=setup_s if want_ehss else 0.0, pa_setup_s=setup_s if bvh_pa else 0.0,
                   all_replicates_with_setup_s=time.perf_counter() - t_start)
    return dict(summary=summary, runs=runs, timings=timings)


# ---------------------------------------------------------------------------
# Command line
# ---------------------------------------------------------------------------

def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="PA/EHSS for isolated molecules in standard XYZ files (results in A^2).")
    add = p.add_argument
    add("xyz", nargs="+", help="XYZ paths; quoted wildcard patterns also work")
    add("--version", action="version", version=f"{TOOL} {VERSION}")
    add("--method", choices=["pa", "ehss", "both"], default="both")
    add("--pa-method", choices=["adaptive", "bvh"], default="adaptive")
    add("--rays", type=int, default=16384, help="EHSS/BVH PA rays per replicate, power of two (default: 16384)")
    add("--directions", type=int, default=64, help="Adaptive PA directions per replicate, power of two (default: 64)")
    add("--grid-size", type=int, default=512, help="Adaptive PA square midpoint grid, power of two (default: 512)")
    add("--repeats", type=int, default=3, help="Independent scrambled Sobol replicates (default: 3)")
    add("--seed", type=int, default=190003)
    add("--threads", type=int, default=1, help="Accepted for compatibility; the GPU engine ignores it")
    add("--leaf-size", type=int, default=4, help="Maximum spatial group size for XYZ BVH (default: 4)")
    add("--max-collisions", type=int, default=64)
    add("--probe-radius", type=float, default=1.4, help="Added probe radius in angstrom (default: 1.4, He scenario)")
    add("--radii-json", type=Path, help="Overrides/additions to atomic VDW radii in angstrom, never effective radii")
    add("--units", choices=["angstrom", "nm"], default="angstrom", help="Input coordinate units only")
    add("--exclude-hydrogen", action="store_true", help="Explicitly omit H; default includes every supplied atom")
    frames = p.add_mutually_exclusive_group()
    frames.add_argument("--frame", type=int, default=1, help="1-based frame selection (default: first frame)")
    frames.add_argument("--all-frames", action="store_true",
                        help="Calculate every frame independently, no MD cache reuse")
    add("--output", type=Path, default=Path("ccs_results"), help="New result directory; existing directory is preserved")
    add("--backend", choices=["gpu"], default="gpu", help="Compute backend (MI355X HIP engine)")
    add("--precision", choices=["fp32-fixup", "fp64"], default="fp32-fixup",
        help="Trace kernel: FP32 filter with exact FP64 re-test (default) or the plain FP64 kernel; both return "
             "identical results")
    return p


def validate_args(args) -> None:
    for name, (low, high) in POW2_LIMITS.items():
        v = getattr(args, name)
        if v < low or v > high or (v & (v - 1)):
            raise ValueError(f"--{name.replace('_', '-')} must be a power of two in {low}..{high}")
    for name in POSITIVE:
        if getattr(args, name) <= 0:
            raise ValueError(f"--{name.replace('_', '-')} must be positive")
    if args.max_collisions > MAX_COLLISIONS:
        raise ValueError(f"--max-collisions must be <= {MAX_COLLISIONS}")
    if args.seed < 0 or args.seed >= 1 << 32:
        raise ValueError("--seed must be in 0..4294967295")
    if not (math.isfinite(args.probe_radius) and args.probe_radius >= 0):
        raise ValueError("--probe-radius must be finite and nonnegative")
    if args.output.exists():
        raise ValueError(f"Output already exists: {args.output}; choose a new --output directory")


@dataclass
class _Job:
    spheres: object
    info: dict


@dataclass
class _Inputs:
    jobs: list = field(default_factory=list)

    @classmethod
    def collect(cls, args, table) -> "_Inputs":
        """Expand the XYZ arguments (paths or glob patterns; each file once), read
        and validate every selected frame before any device work starts."""
        out, seen = cls(), set()
        for pattern in args.xyz:
            paths = [pattern] if Path(pattern).is_file() else sorted(glob.glob(pattern))
            if not paths:
                raise ValueError(f"No XYZ files match: {pattern}")
            for name in paths:
                path = Path(name).resolve()
                if path in seen:
                    continue
                seen.add(path)
                frames = read_xyz(path)
                if args.all_frames:
                    chosen = frames
                elif args.frame <= len(frames):
                    chosen = [frames[args.frame - 1]]
                    if len(frames) > 1:
                        print(f"{path.name}: using frame {args.frame}/{len(frames)}; use --all-frames for all.",
                              file=sys.stderr)
                else:
                    raise ValueError(f"{path.name}: --frame {args.frame} exceeds {len(frames)} frames")
                digest = hashlib.sha256(path.read_bytes()).hexdigest()
                for fr in chosen:
                    spheres, info = prepare_frame(fr, table, args.probe_radius, args.units, args.exclude_hydrogen)
                    out.jobs.append(_Job(spheres, dict(input_file=str(path), input_sha256=digest,
                                                       total_frames=len(frames), **info)))
        return out


NOTES = (
    "Static elastic hard-sphere model; atomic VDW radius + probe radius. Not TM or fitted IMoS/EHSSrot parameters.",
    "XYZ spatial groups, no inferred bonds/residues; no support response reuse.",
    "Independent Sobol scrambles across replicates. SD/SEM measure replicate variation, not total error or a "
    "guaranteed confidence interval.",
    "Adaptive PA has finite-grid and directional error; replicate SD does not estimate grid bias.",
    "EHSS sampled interval covers omitted cap survivors only, not sampling error or physical model error.",
    "Warmup/import/IO excluded from per-frame compute timing; setup is once per frame, shared across replicates.",
    "For both + BVH PA, the impact source is shared with EHSS and generated on the device.",
    "GPU engine: source, PA, trace and reduction fused on the device; all replicates of a frame run in one "
    "launch (per-replicate query_s is the launch time split evenly); collision ledger in 2^-40 fixed point.",
)

CSV_COLUMNS = ("input_file", "frame", "atoms", "pa_mean_A2", "pa_sd_A2", "ehss_mean_A2", "ehss_sd_A2",
               "ehss_tail_bound_mean_A2", "unresolved_rays_total", "compute_all_replicates_s")


def _csv_row(record: dict) -> dict:
    s = record["summary"]
    pa, eh = s.get("pa_angstrom2") or {}, s.get("ehss_angstrom2") or {}
    values = (record["input_file"], record["frame"], record["selected_atoms"], pa.get("mean"), pa.get("sd"),
              eh.get("mean"), eh.get("sd"), s.get("ehss_tail_bound_mean_angstrom2"),
              s.get("ehss_unresolved_rays_total"), record["timings"]["all_replicates_with_setup_s"])
    return dict(zip(CSV_COLUMNS, values))


def _write_outputs(directory: Path, payload: dict) -> None:
    directory.mkdir(parents=True, exist_ok=False)
    text = json.dumps(payload, ensure_ascii=False, indent=2, allow_nan=False) + "\n"
    (directory / "results.json").write_text(text, encoding="utf-8")
    with (directory / "summary.csv").open("w", newline="", encoding="utf-8-sig") as f:
        w = csv.DictWriter(f, fieldnames=list(CSV_COLUMNS))
        w.writeheader()
        w.writerows(_csv_row(r) for r in payload["records"])


def _report(record: dict) -> None:
    for key in ("pa_angstrom2", "ehss_angstrom2"):
        if key in record["summary"]:
            v = record["summary"][key]
            sd = "n/a" if v["sd"] is None else f"{v['sd']:.6f}"
            print(f"  {key.split('_')[0].upper()}: {v['mean']:.6f} A^2 (replicate SD {sd})")
    print(f"  Compute: {record['timings']['all_replicates_with_setup_s']:.4f} s (setup + all replicates; "
          f"warmup excluded)")


def execute(args) -> dict:
    """Validate everything, warm the device up, solve every frame, write the
    outputs; returns the results.json payload."""
    validate_args(args)
    import scipy

    from . import _lib
    from .state import Spheres
    t0 = time.perf_counter()
    table = radius_table(args.radii_json)
    inputs = _Inputs.collect(args, table)
    print(f"Validated {len(inputs.jobs)} frame(s). Preparing the GPU engine...", file=sys.stderr, flush=True)
    tiny = argparse.Namespace(**{**vars(args), "rays": 8, "directions": 1, "grid_size": 8, "repeats": 1})
    t = time.perf_counter()
    solve_frame(Spheres(np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]]), np.array([1.0, 1.0])), tiny, progress=False)
    warmup_s = time.perf_counter() - t
    records = []
    for job in inputs.jobs:
        print(f"{Path(job.info['input_file']).name}, frame {job.info['frame']}, {job.info['selected_atoms']} atoms",
              file=sys.stderr, flush=True)
        record = {**job.info, **solve_frame(job.spheres, args)}
        _report(record)
        records.append(record)
    device = _lib.device_info()
    settings = {k: (str(v) if isinstance(v, Path) else v) for k, v in vars(args).items()}
    environment = dict(python=platform.python_version(), platform=platform.platform(), numpy=np.__version__,
                       scipy=scipy.__version__, backend="hip-gfx950", device=device["name"],
                       compute_units=device["compute_units"], source_trig=_lib.trig_compat_info()["libm"] or
                       "correctly rounded")
    payload = dict(schema_version=1, tool=TOOL, version=VERSION, created_utc=datetime.now(timezone.utc).isoformat(),
                   settings=settings, atomic_radii_angstrom=table, radii_source=RADII_SOURCE,
                   environment=environment, warmup_s=warmup_s, elapsed_before_output_s=time.perf_counter() - t0,
                   notes=list(NOTES), records=records)
    _write_outputs(args.output, payload)
    for name in ("summary.csv", "results.json"):
        print(f"Saved {args.output.resolve() / name}")
    return payload


def main(argv=None) -> int:
    p = parser()
    args = p.parse_args(argv)
    try:
        execute(args)
    except (ValueError, OSError) as exc:
        p.exit(EXIT_INVALID, f"Error: {exc}\n")
    except KeyboardInterrupt:
        p.exit(EXIT_INTERRUPTED, "Interrupted.\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

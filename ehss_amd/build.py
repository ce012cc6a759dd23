"""Builds the package's native parts in-tree:

  ehss_amd/libehss_hip.so       the gfx950 HIP library (hipcc; one object per
                                source, compiled in parallel, then linked);
  ehss_amd/libehss_trighost.so  host build of the source-angle trig (g++);
  ehss_amd/trig_compat.bin      the host-libm compatibility table of the Sobol
                                angles, generated with libehss_trighost.so from
                                this host's libm (regenerated when the trig
                                source or the libm changes).

Usage: python -m ehss_amd.build   (or ehss_amd.build.build() from Python)
"""
from __future__ import annotations

import ctypes as ct
import hashlib
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libehss_hip.so"
TRIG_LIB = PKG / "libehss_trighost.so"
TRIG_TABLE = PKG / "trig_compat.bin"
OBJ = PKG / "_build"

SOURCES = ["capi.hip", "trace.hip", "build.hip", "build_shaped.hip", "scene.hip", "solve.hip", "sources.hip",
           "multi.hip", "pagrid.hip", "scramble.hip"]
TRIG_SOURCES = ["trig_host.cpp", "ehss_trig.h", "ehss_trig_table.h"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Exact FP64 semantics: no FMA contraction, IEEE division and square root.
    "-ffp-contract=off",
    "-fno-fast-math",
    # No packed FP32 instructions anywhere in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]

HOST_FLAGS = ["-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c++17", "-Wall",
              f"-I{CSRC}"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def sources() -> list[Path]:
    return [CSRC / s for s in SOURCES if (CSRC / s).exists()]


def _run(cmd: list[str], what: str) -> str:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"{what} failed ({proc.returncode})")
    return proc.stderr


def _noise(stderr: str) -> list[str]:
    return [ln for ln in stderr.splitlines()
            if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]


def build_hip(verbose: bool = False) -> Path:
    """Compile every HIP source (in parallel) and link LIB (always recompiles)."""
    srcs = sources()
    OBJ.mkdir(exist_ok=True)
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    cc = hipcc()
    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        errs = list(pool.map(lambda so: _run([cc, *FLAGS, "-c", "-o", str(so[1]), str(so[0])], f"hipcc {so[0].name}"),
                             zip(srcs, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    errs.append(_run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], "hipcc link"))
    os.replace(tmp, LIB)
    noise = [ln for e in errs for ln in _noise(e)]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return LIB


def build_trig_host() -> Path:
    tmp = TRIG_LIB.with_suffix(".so.tmp")
    _run(["g++", *HOST_FLAGS, "-o", str(tmp), str(CSRC / "trig_host.cpp"), "-lm"], "g++ trig_host.cpp")
    os.replace(tmp, TRIG_LIB)
    return TRIG_LIB


def _trig_stamp() -> bytes:
    """Identifies a table: the trig sources and the host libm's version."""
    h = hashlib.sha256()
    for name in TRIG_SOURCES:
        h.update((CSRC / name).read_bytes())
    lib = ct.CDLL(str(TRIG_LIB))
    buf = ct.create_string_buffer(64)
    lib.ehss_trig_libm_version(buf, ct.c_size_t(len(buf)))
    h.update(buf.value)
    return h.hexdigest().encode()


def build_trig_table(verbose: bool = False) -> Path:
    """Generate TRIG_TABLE from this host's libm unless an identical one exists."""
    stamp_path = TRIG_TABLE.with_suffix(".stamp")
    stamp = _trig_stamp()
    if TRIG_TABLE.exists() and stamp_path.exists() and stamp_path.read_bytes() == stamp:
        return TRIG_TABLE
    lib = ct.CDLL(str(TRIG_LIB))
    lib.ehss_trig_compat_generate.restype = ct.c_int
    lib.ehss_trig_compat_generate.argtypes = [ct.c_char_p, ct.c_int, ct.c_void_p]
    stats = (ct.c_int64 * 8)()
    tmp = TRIG_TABLE.with_suffix(".tmp")
    t0 = time.time()
    rc = lib.ehss_trig_compat_generate(str(tmp).encode(), min(16, os.cpu_count() or 1), stats)
    if rc != 0:
        raise RuntimeError(f"trig compatibility table generation failed ({rc}); stats {list(stats)[:6]}")
    os.replace(tmp, TRIG_TABLE)
    stamp_path.write_bytes(stamp)
    if verbose:
        print(f"generated {TRIG_TABLE.relative_to(ROOT)}: {stats[0]} angles where the host libm differs from the "
              f"correctly rounded value ({stats[1]} cos, {stats[2]} sin); the device looks up "
              f"{stats[6] / 2**30:.2%} of angles; {time.time() - t0:.1f}s")
    return TRIG_TABLE


def build(verbose: bool = False) -> Path:
    t0 = time.time()
    build_trig_host()
    build_trig_table(verbose)
    build_hip(verbose)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(sources())} sources in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

"""Prepared static EHSS engine on the GPU: the operator the CLI calls.

Mirrors (reference file:line):
  static_spatial_groups    ehss_static.py:53-59
  PreparedStaticCPUEHSS    ehss_static.py:62-101  -> PreparedStaticGPUEHSS
  CPUEHSSTransport.trace   ehss_cpu.py:118-142    -> GPUEHSSTransport.trace
  check_threads            ehss_cpu.py:107-109

The static tree and the PA tree are built on the device; `.model.tree` exports
the reference-identical arrays on first access.
"""
from __future__ import annotations

import time
from types import SimpleNamespace

import numpy as np

from .engine import GPUScene
from .sources import CompactBoundary
from .state import EHSSResult, Spheres


def check_threads(threads):
    """The GPU engine accepts the CLI's --threads (host threads are not used)."""
    if type(threads) is not int or threads < 1:
        raise ValueError("Threads must be a positive integer")


def static_spatial_groups(spheres: Spheres, leaf_size: int = 4) -> np.ndarray:
    """Median-split groups of ehss_static.py:53-59, computed by the device build."""
    if type(leaf_size) is not int or leaf_size < 1:
        raise ValueError("Positive integer leaf size required")
    with GPUScene(spheres.centers, spheres.radii, leaf_size=leaf_size) as scene:
        return scene.export_tree()["groups"]


class _LazyTree(SimpleNamespace):
    """Tree arrays exported from the device on first attribute access."""

    def __init__(self, scene: GPUScene):
        super().__init__()
        object.__setattr__(self, "_scene", scene)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        arrays = self._scene.export_tree()
        for k, v in arrays.items():
            object.__setattr__(self, k, v)
        if name not in arrays:
            raise AttributeError(name)
        return arrays[name]


class GPUEHSSTransport:
    """engine.trace(source, cap, record_history) on the device scene."""

    def __init__(self, model, scene: GPUScene, threads: int = 1, precision: str = "fp32-fixup"):
        check_threads(threads)
        self.model, self.scene = model, scene
        self.threads, self.precision = threads, precision
        self.certified_centers = model.spheres.centers.copy()
        self.certified_radii = model.spheres.radii.copy()
        self.prepare_s = 0.0

    def trace(self, source, cap=64, record_history=False) -> EHSSResult:
        if type(cap) is not int or cap < 1 or cap > np.iinfo(np.int32).max or np.any(source.initial_bounces > 1):
            raise ValueError("Positive cap and zero/one initial collision required")
        s = self.model.spheres
        if np.any(source.last_atom < -1) or np.any(source.last_atom >= len(s.radii)):
            raise ValueError("Invalid initial physical identity")
        start = time.perf_counter()
        if not (np.array_equal(s.centers, self.certified_centers) and np.array_equal(s.radii, self.certified_radii)):
            raise ValueError("Geometry changed: bind a new transport")
        out = self.scene.trace(source.origins, source.directions, source.last_atom, source.initial_bounces, cap=cap,
                               record=bool(record_history), precision=self.precision)
        history = out["collider_ids"] if record_history else np.full((len(source.weights), 0), -1, np.int32)
        return EHSSResult(out["positions"], out["outgoing"], out["escaped"], out["unresolved"], out["bounces"],
                          history, source,
                          dict(query_s=time.perf_counter() - start, sphere_tests=out["sphere_tests"],
                               box_tests=out["box_tests"], fp64_retests=out["fp64_retests"],
                               backend=f"HIP gfx950 ({self.precision})", threads=self.threads, scalar=True,
                               fastmath=False, history_recorded=bool(record_history), response_reuse=False))


class PreparedStaticGPUEHSS:
    """Owned geometry for repeated sampling of one unchanged structure, resident
    on the GPU (ehss_static.py:62-101)."""

    def __init__(self, spheres, groups=None, leaf_size=4, threads=1, boundary="sphere", simplify=False,
                 coverage_depth=4, precision="fp32-fixup"):
        check_threads(threads)
        if simplify:
            raise ValueError("Atom simplification is not part of the static path")
        start = time.perf_counter()
        s = Spheres(spheres.centers, spheres.radii, deduplicate=False)  # own the coordinates
        if groups is not None:
            groups = np.asarray(groups)
            if groups.shape != (len(s.radii),) or not np.issubdtype(groups.dtype, np.integer):
                raise ValueError("One integer group per atom required")
            unique = np.unique(groups)
            if not np.array_equal(unique, np.arange(len(unique))):
                raise ValueError("Contiguous nonempty groups required")
        tick = time.perf_counter()
        self.scene = GPUScene(s.centers, s.radii, leaf_size=leaf_size, groups=groups)
        build_s = time.perf_counter() - tick
        self.model = SimpleNamespace(spheres=s, tree=_LazyTree(self.scene))
        tick = time.perf_counter()
        try:
            self.boundary = CompactBoundary.fit(s, boundary)
        except ValueError:
            self.scene.close()
            raise
        boundary_s = time.perf_counter() - tick
        self.engine = GPUEHSSTransport(self.model, self.scene, threads=threads, precision=precision)
        self.setup = dict(total_s=time.perf_counter() - start, group_s=0.0, bounds_s=build_s, boundary_s=boundary_s,
                          engine_s=0.0, device_build_s=build_s, threads=threads, geometry_threads=1,
                          preparation="static (device build)", response_metadata=False,
                          exclusivity_computed=False)

"""Sphere model, source states and the EHSS readout: the host-side data types of
the reference estimator API, with the same fields, validation and properties.

Mirrors (reference file:line):
  Spheres          inside_out_pa.py:17-42
  Source           inside_out_pa.py:61-69
  unit_sphere      inside_out_pa.py:45-49   (host helper; the device path generates rays itself)
  tangent_frames   inside_out_pa.py:52-58
  EHSSSource       ehss_state.py:8-39
  external_source  ehss_state.py:42-52      (rays generated on the GPU)
  EHSSResult       ehss_state.py:72-108
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Spheres:
    centers: np.ndarray
    radii: np.ndarray
    deduplicate: bool = True

    def __post_init__(self):
        c = np.ascontiguousarray(self.centers, dtype=np.float64)
        r = np.ascontiguousarray(self.radii, dtype=np.float64)
        if c.ndim != 2 or c.shape[1] != 3 or r.shape != (len(c),) or len(c) == 0:
            raise ValueError("Expected nonempty centers (N,3) and radii (N,).")
        if not np.isfinite(c).all() or not np.isfinite(r).all() or np.any(r <= 0):
            raise ValueError("Finite centers and positive finite radii required.")
        if self.deduplicate:  # exact duplicates need a single surface owner; keep the first
            _, ids = np.unique(np.column_stack((c, r)), axis=0, return_index=True)
            ids.sort()
        else:
            ids = np.arange(len(c))
        object.__setattr__(self, "centers", c[ids].copy())
        object.__setattr__(self, "radii", r[ids].copy())

    @property
    def area(self):
        return float(4 * np.pi * np.dot(self.radii, self.radii))


def unit_sphere(uv):
    z = 1 - 2 * uv[:, 0]
    phi = 2 * np.pi * uv[:, 1]
    rho = np.sqrt(np.maximum(0, 1 - z * z))
    return np.column_stack((rho * np.cos(phi), rho * np.sin(phi), z))


def tangent_frames(normals):
    axes = np.zeros_like(normals)
    axes[:, 2] = 1
    axes[np.abs(normals[:, 2]) > .9] = [1, 0, 0]
    t = np.cross(axes, normals)
    t /= np.linalg.norm(t, axis=1)[:, None]
    return t, np.cross(normals, t)


@dataclass
class Source:
    origins: np.ndarray
    directions: np.ndarray
    owners: np.ndarray
    area: float

    def pa(self, blocked):
        return self.area * float(np.mean(~blocked)) / 4


@dataclass
class EHSSSource:
    origins: np.ndarray
    directions: np.ndarray
    incoming: np.ndarray
    weights: np.ndarray
    last_atom: np.ndarray
    initial_bounces: np.ndarray
    convention: str

    def __post_init__(self):
        n = len(self.weights)
        for name in ["origins", "directions", "incoming"]:
            value = np.ascontiguousarray(getattr(self, name), dtype=float)
            if value.shape != (n, 3) or not np.isfinite(value).all():
                raise ValueError("Finite (N,3) positions and directions required")
            setattr(self, name, value)
        for name in ["directions", "incoming"]:
            if not np.allclose(np.linalg.norm(getattr(self, name), axis=1), 1., atol=1e-12, rtol=1e-12):
                raise ValueError("Unit physical directions required")
        self.weights = np.ascontiguousarray(self.weights, dtype=float)
        if self.weights.shape != (n,) or not np.isfinite(self.weights).all() or np.any(self.weights < 0):
            raise ValueError("Nonnegative finite source weights required")
        for name in ["last_atom", "initial_bounces"]:
            value = np.ascontiguousarray(getattr(self, name), dtype=np.int64)
            if value.shape != (n,):
                raise ValueError("One initial identity and bounce count per state required")
            setattr(self, name, value)
        if np.any(self.initial_bounces < 0):
            raise ValueError("Nonnegative initial bounce count required")
        if np.any((self.initial_bounces == 1) & (self.last_atom < 0)):
            raise ValueError("First-reflected source states require their collider identity")


def external_source(spheres: Spheres, power: int, seed: int) -> EHSSSource:
    """Independent isotropic incidence times a uniform enclosing impact disk
    (ehss_state.py:42-52): 2^power scrambled Sobol points mapped to rays on the
    GPU (ehss_source_rays), returned as host arrays."""
    from .engine import GPUScene, scramble
    if type(power) is not int or not 0 <= power <= 26:
        raise ValueError("power must be an integer in 0..26")
    with GPUScene(spheres.centers, spheres.radii) as scene:
        rays = scene.source_rays(scramble(seed), power)
        radius = scene.info["radius"]
    n = 1 << power
    d = rays["directions"]
    return EHSSSource(rays["origins"], d, d.copy(), np.full(n, np.pi * radius * radius / n),
                      np.full(n, -1), np.zeros(n, dtype=np.int64), "external impact disk")


@dataclass
class EHSSResult:
    positions: np.ndarray
    outgoing: np.ndarray
    escaped: np.ndarray
    unresolved: np.ndarray
    bounces: np.ndarray
    collider_ids: np.ndarray
    source: EHSSSource
    metrics: dict

    @property
    def contributions(self):
        # Numerical clamp only at the analytic [0,2] bounds.
        angle = np.clip(1 - np.sum(self.source.incoming * self.outgoing, axis=1), 0., 2.)
        return self.source.weights * angle * self.escaped

    @property
    def omega(self):
        return float(self.contributions.sum())

    @property
    def residual_mass(self):
        return float(self.source.weights[self.unresolved].sum())

    @property
    def tail_bound(self):
        return 2 * self.residual_mass

    @property
    def pa(self):
        return float(self.source.weights[(self.bounces > 0) | self.unresolved].sum())

    def order_ledger(self):
        c = self.contributions
        return [dict(order=k, count=int(np.count_nonzero(self.escaped & (self.bounces == k))),
                     mass=float(self.source.weights[self.escaped & (self.bounces == k)].sum()),
                     omega=float(c[self.bounces == k].sum()))
                for k in range(int(self.bounces.max(initial=0)) + 1)]

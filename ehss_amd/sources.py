"""Alternative incidence sources of the EHSS estimator, for the device engine.

Two forms.  The host form generates the source states with NumPy (as the
reference does) for the device to trace (PreparedStaticGPUEHSS.engine.trace)
and, for the PA sources, test (SphereIndex.query).  The fused form
(CompactBoundary.solve, reverse_pa_solve) generates the same states on the
device and traces and reduces them there.

Mirrors (reference file:line):
  ball_metric_bounds       pa_ellipsoidal.py:15-66
  fit_centroid_ellipsoid   pa_ellipsoidal.py:69-89
  CompactBoundary          ehss_compact_source.py:16-87 (sphere, ellipsoid, box)
  sample_source            inside_out_pa.py:72-90
  inside_out               inside_out_pa.py:115-118 (any-hit on the device BVH instead
                           of the all-pairs loop; same predicate, same flags)
  reverse_pa_source        ehss_state.py:55-69
  CompactBoundary.solve,   the same sources generated, traced and reduced on the
  reverse_pa_solve         device (ehss_solve_source), nothing per ray crossing PCIe

Sobol points come from engine.sobol_points, which reproduces
scipy.stats.qmc.Sobol(d, scramble=True, seed).random_base2(m) bit for bit
(checked against SciPy for d = 2, 4, 5).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .engine import sobol_points
from .state import EHSSSource, Source, Spheres, tangent_frames, unit_sphere


def ball_metric_bounds(points, radii, axes):
    """Conservative bounds on sqrt(sum((p + u)^2 / a^2)) over ||u|| <= r, per
    ball (pa_ellipsoidal.py:15-66): the dual bounds
      max q <= base + lam r^2 + sum(b^2 / (lam - D)), lam > max(D),
      min q >= base - lam r^2 - sum(b^2 / (D + lam)), lam >= 0,
    with D = diag(1/a^2), b = D p, base = p'Dp, tightened by 64 bisection steps
    on lam, and outward rounding margins.  Vectorised over the balls; each
    ball's bisection takes the same steps as the scalar loop."""
    points = np.asarray(points, dtype=np.float64)
    radii = np.asarray(radii, dtype=np.float64)
    axes = np.asarray(axes, dtype=np.float64)
    dd = 1.0 / (axes * axes)
    dmax = float(np.max(dd))
    b = dd[None, :] * points
    base = (points[:, 0] * b[:, 0] + points[:, 1] * b[:, 1]) + points[:, 2] * b[:, 2]
    bnorm = np.sqrt((b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1]) + b[:, 2] * b[:, 2])
    r = radii
    n = len(points)
    low = np.zeros(n)
    high = np.empty(n)
    zero = bnorm == 0.0
    high[zero] = r[zero] * np.sqrt(dmax) * (1.0 + 1e-12)
    act = ~zero

    def length2(lam, sign):
        out = np.zeros_like(lam)
        for k in range(3):
            out = out + (b[:, k] / (lam - dd[k] if sign < 0 else dd[k] + lam)) ** 2
        return out

    # Upper bound: lam in (dmax, dmax + |b|/r].
    left = np.full(n, dmax)
    right = np.nextafter(dmax + np.where(act, bnorm, 0.0) / r, np.inf)
    live = act.copy()
    for _ in range(64):
        mid = left + (right - left) / 2.0
        stop = (mid <= dmax) | (mid == left) | (mid == right)
        live &= ~stop
        if not live.any():
            break
        big = length2(np.where(live, mid, dmax + 1.0), -1) > r * r
        left = np.where(live & big, mid, left)
        right = np.where(live & ~big, mid, right)
    lam = np.maximum(right, np.nextafter(dmax, np.inf))
    upper = base + lam * r * r
    for k in range(3):
        upper = upper + b[:, k] * b[:, k] / (lam - dd[k])
    high[act] = np.sqrt(np.maximum(0.0, upper + 1e-12 * np.maximum(1.0, upper)))[act]
    # Lower bound, for balls not containing the origin of the metric.
    pp = (points[:, 0] * points[:, 0] + points[:, 1] * points[:, 1]) + points[:, 2] * points[:, 2]
    need = act & ~(pp <= r * r)
    left = np.zeros(n)
    right = np.where(need, bnorm / r, 0.0)
    for _ in range(64):
        mid = (left + right) / 2.0
        big = length2(mid, +1) > r * r
        left = np.where(need & big, mid, left)
        right = np.where(need & ~big, mid, right)
    lam = right
    corr = lam * r * r
    for k in range(3):
        corr = corr + b[:, k] * b[:, k] / (dd[k] + lam)
    lo = np.sqrt(np.maximum(0.0, base - corr - 1e-12 * np.maximum(1.0, base + corr)))
    low[need] = lo[need]
    return low, high


def fit_centroid_ellipsoid(spheres: Spheres):
    """PCA shape with a sphere-radius floor, uniformly scaled to enclose every
    ball (pa_ellipsoidal.py:69-89): (center, rotation, axes)."""
    center = spheres.centers.mean(axis=0)
    centered = spheres.centers - center
    covariance = centered.T @ centered / len(centered) + np.eye(3) * np.mean(spheres.radii ** 2) / 5.0
    eigenvalues, rotation = np.linalg.eigh(covariance)
    axes = np.sqrt(eigenvalues)
    points = np.ascontiguousarray(centered @ rotation)
    _, maximum = ball_metric_bounds(points, spheres.radii, axes)
    axes = axes * (np.max(maximum) * (1.0 + 1e-10))
    return center, rotation, axes


@dataclass
class CompactBoundary:
    """Isotropic incidence on a compact convex boundary enclosing the molecule
    (ehss_compact_source.py:16-87): directions uniform on S2, positions uniform
    in the boundary's projection along the direction, each ray weighted by
    projected area / sample count."""
    kind: str
    center: np.ndarray
    rotation: np.ndarray
    axes: np.ndarray
    prepare_s: float = 0.0

    @classmethod
    def fit(cls, spheres: Spheres, kind: str = "ellipsoid") -> "CompactBoundary":
        start = time.perf_counter()
        center = spheres.centers.mean(axis=0)
        rotation = np.eye(3)
        if kind == "sphere":
            radius = np.max(np.linalg.norm(spheres.centers - center, axis=1) + spheres.radii) * (1 + 1e-9)
            axes = np.full(3, radius)
        elif kind == "ellipsoid":
            center, rotation, axes = fit_centroid_ellipsoid(spheres)
        elif kind == "box":
            low = np.min(spheres.centers - spheres.radii[:, None], axis=0)
            high = np.max(spheres.centers + spheres.radii[:, None], axis=0)
            center = (low + high) / 2
            axes = (high - low) / 2 * (1 + 1e-9)
        else:
            raise ValueError("Boundary must be sphere, ellipsoid, or box")
        return cls(kind, center, rotation, axes, time.perf_counter() - start)

    def source(self, power, seed) -> EHSSSource:
        if not isinstance(power, (int, np.integer)) or power < 0:
            raise ValueError("Nonnegative integer Sobol power required")
        u = sobol_points(seed, 5 if self.kind == "box" else 4, int(power))
        d = unit_sphere(u[:, :2])
        n = len(d)
        if self.kind == "sphere":
            t, b = tangent_frames(d)
            radius = self.axes[0]
            rho, phi = radius * np.sqrt(u[:, 2]), 2 * np.pi * u[:, 3]
            origins = self.center - radius * d + rho[:, None] * (np.cos(phi)[:, None] * t + np.sin(phi)[:, None] * b)
            area = np.full(n, np.pi * radius * radius)
        elif self.kind == "box":
            local_d = d @ self.rotation
            face_areas = 4 * np.prod(self.axes) / self.axes
            projected = np.abs(local_d) * face_areas
            area = projected.sum(axis=1)
            face = (u[:, 2, None] * area[:, None] >= np.cumsum(projected, axis=1)).sum(axis=1)
            p = np.empty((n, 3))
            for axis in range(3):
                sel = face == axis
                p[sel, axis] = -np.sign(local_d[sel, axis]) * self.axes[axis]
                other = [k for k in range(3) if k != axis]
                p[np.ix_(sel, other)] = (2 * u[sel, 3:5] - 1) * self.axes[other]
            # Start just outside the boundary, upstream on the same line.
            origins = self.center + p @ self.rotation.T - 1e-10 * max(1.0, self.axes.max()) * d
        else:
            t, b = tangent_frames(d)
            lt, lb = (t @ self.rotation) * self.axes, (b @ self.rotation) * self.axes
            g00 = np.sum(lt * lt, axis=1)
            g01 = np.sum(lt * lb, axis=1)
            g11 = np.sum(lb * lb, axis=1)
            l00 = np.sqrt(g00)
            l10 = g01 / l00
            l11 = np.sqrt(np.maximum(0.0, g11 - l10 * l10))
            rho, phi = np.sqrt(u[:, 2]), 2 * np.pi * u[:, 3]
            x, y = rho * np.cos(phi), rho * np.sin(phi)
            offset = (l00 * x)[:, None] * t + (l10 * x + l11 * y)[:, None] * b
            extent = np.linalg.norm((d @ self.rotation) * self.axes, axis=1)
            origins = self.center + offset - (extent * (1 + 1e-10))[:, None] * d
            area = np.pi * l00 * l11
        return EHSSSource(origins, d, d.copy(), area / n, np.full(n, -1), np.zeros(n, dtype=np.int64),
                          f"isotropic projected {self.kind}")

    def solve(self, scene, power: int, seed: int, cap: int = 64, keep_rays: bool = False) -> dict:
        """This source generated, traced and reduced on the device in one call
        (ehss_solve_source): the states of source(power, seed), traced against
        `scene` (a GPUScene of the same structure), with the EHSSResult readout
        and per-ray weights.  Nothing crosses PCIe but the totals.  The 'sphere'
        kind is the fused solve's impact disk (the same rays as
        external_source when the boundary is the fitted one)."""
        from .engine import scramble, scramble5, source_params
        if self.kind == "sphere":
            return scene.solve(scramble(seed), power, cap, pa=False)
        params = source_params(self.kind, self.center, self.rotation, self.axes)
        return scene.solve_source(params, scramble5(seed, 5 if self.kind == "box" else 4), power, cap,
                                  keep_rays=keep_rays)


def sample_source(spheres: Spheres, power: int, seed: int) -> Source:
    """Area-selected sphere surface points with cosine-hemisphere directions
    (inside_out_pa.py:72-90): 5-D scrambled Sobol points."""
    u = sobol_points(seed, 5, int(power))
    weights = spheres.radii ** 2
    ids = np.searchsorted(np.cumsum(weights) / weights.sum(), u[:, 0], side="right")
    normals = unit_sphere(u[:, 1:3])
    t, b = tangent_frames(normals)
    phi = 2 * np.pi * u[:, 4]
    rho = np.sqrt(u[:, 3])
    directions = (np.sqrt(1 - u[:, 3])[:, None] * normals
                  + (rho * np.cos(phi))[:, None] * t
                  + (rho * np.sin(phi))[:, None] * b)
    origins = spheres.centers[ids] + spheres.radii[ids, None] * normals
    return Source(np.ascontiguousarray(origins), np.ascontiguousarray(directions), ids.astype(np.int64),
                  spheres.area)


def inside_out(spheres: Spheres, source: Source, index=None):
    """(PA, blocked, sphere tests) of an inside-out source: a state is blocked
    when any sphere other than its owner is hit ahead of it (rays starting
    inside a sphere count as hits).  The any-hit runs on the device PA tree."""
    from .pa import SphereIndex
    own = index is None
    idx = SphereIndex(spheres) if own else index
    try:
        blocked, _, tests = idx.query(source)
    finally:
        if own:
            idx.scene.close()
    blocked = np.asarray(blocked, dtype=bool)
    return source.pa(blocked), blocked, int(tests)


def reverse_pa_source(spheres: Spheres, power: int, seed: int, index=None) -> EHSSSource:
    """Reverse-PA escape states with the first reflection already applied
    (ehss_state.py:55-69); the weights keep the pre-filter sample measure, so
    they sum to the inside-out PA."""
    source = sample_source(spheres, power, seed)
    pa, blocked, _ = inside_out(spheres, source, index)
    keep = ~blocked
    owners = source.owners[keep]
    origins = source.origins[keep]
    incoming = -source.directions[keep]
    normals = (origins - spheres.centers[owners]) / spheres.radii[owners, None]
    normals /= np.linalg.norm(normals, axis=1)[:, None]
    outgoing = incoming - 2 * np.sum(incoming * normals, axis=1)[:, None] * normals
    outgoing /= np.linalg.norm(outgoing, axis=1)[:, None]
    result = EHSSSource(origins, outgoing, incoming, np.full(len(owners), source.area / (4 * len(source.owners))),
                        owners, np.ones(len(owners), dtype=np.int64),
                        "PA reverse, first reflection already applied")
    if not np.isclose(result.weights.sum(), pa, rtol=1e-13, atol=1e-13):
        raise AssertionError("PA source measure changed during filtering")
    return result


def reverse_pa_solve(scene, spheres: Spheres, power: int, seed: int, cap: int = 64, keep_rays: bool = False) -> dict:
    """reverse_pa_source (ehss_state.py:55-69) on the device in one call: the
    area-selected surface samples, their inside-out any-hit test, the reflected
    unblocked states, their trace against `scene` (a GPUScene of `spheres`) and
    the readout (ehss_solve_source).  pa is the inside-out PA
    (inside_out_pa.py:66-67); pa_from_ehss the result's PA."""
    from .engine import scramble5, source_params
    w = spheres.radii ** 2
    cdf = np.cumsum(w) / w.sum()  # inside_out_pa.py:79
    params = source_params("reverse-pa", area=spheres.area, cdf=cdf)
    return scene.solve_source(params, scramble5(seed, 5), power, cap, keep_rays=keep_rays)

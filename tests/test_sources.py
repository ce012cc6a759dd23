"""Alternative incidence sources on the host (no GPU): compact boundaries
(ehss_compact_source.py:16-87), the ellipsoid fit (pa_ellipsoidal.py:15-89) and
the inside-out sample source (inside_out_pa.py:72-90).  Pinned by SciPy's own
Sobol points, closed forms (Cauchy's projected-area formula, sphere and box
limits) and the enclosure the fit guarantees."""
import math

import numpy as np
import pytest
from scipy.stats import qmc

from ehss_amd import sources as S
from ehss_amd.engine import sobol_points
from ehss_amd.state import Spheres
from tests import scenes


@pytest.mark.parametrize("dim", [2, 4, 5])
@pytest.mark.parametrize("seed", [0, 98, 114, 881, 190003])
def test_sobol_points_equal_scipy(dim, seed):
    np.testing.assert_array_equal(sobol_points(seed, dim, 9), qmc.Sobol(dim, scramble=True, seed=seed).random_base2(9))


def _molecule():
    c, r = scenes.globule(300, 4)
    return Spheres(c, r, deduplicate=False)


def _ball_metric_brute(points, radii, axes, samples=20000, seed=0):
    """max and min of sqrt(sum((p+u)^2/a^2)) over sampled u on the ball's surface."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(samples, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    out = []
    for p, r in zip(points, radii):
        q = np.sqrt((((p[None, :] + r * u) / axes) ** 2).sum(axis=1))
        out.append((q.min(), q.max()))
    return np.array(out)


def test_ball_metric_bounds_bracket_the_sampled_extremes():
    rng = np.random.default_rng(3)
    points = rng.normal(scale=4.0, size=(40, 3))
    points[0] = 0.0                     # centred ball: lower bound 0
    points[1] = [0.1, 0.0, 0.0]         # ball containing the origin
    radii = rng.uniform(0.5, 2.0, size=40)
    axes = np.array([3.0, 5.0, 9.0])
    low, high = S.ball_metric_bounds(points, radii, axes)
    ext = _ball_metric_brute(points, radii, axes)
    assert np.all(low <= ext[:, 0] + 1e-12) and np.all(high >= ext[:, 1] - 1e-12)
    # tight: within 1e-3 relative of the sampled extremes (the bisection converges)
    assert np.all(high <= ext[:, 1] * (1 + 2e-3))
    inside = np.linalg.norm(points, axis=1) > radii
    assert np.all(low[inside] >= ext[inside, 0] - 2e-3 * ext[inside, 1])
    assert low[0] == 0.0 and low[1] == 0.0
    # a sphere metric has the closed form | |p| -+ r | / a
    la, ha = S.ball_metric_bounds(points, radii, np.full(3, 2.0))
    d = np.linalg.norm(points, axis=1)
    np.testing.assert_allclose(ha, (d + radii) / 2.0, rtol=1e-9)
    np.testing.assert_allclose(la[inside], (d[inside] - radii[inside]) / 2.0, rtol=1e-9, atol=1e-12)


def test_fitted_ellipsoid_encloses_every_ball():
    s = _molecule()
    center, rot, axes = S.fit_centroid_ellipsoid(s)
    np.testing.assert_allclose(rot.T @ rot, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(center, s.centers.mean(axis=0))
    rng = np.random.default_rng(1)
    u = rng.normal(size=(500, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    for c, r in zip(s.centers, s.radii):
        p = (c + r * u - center) @ rot
        assert np.all(np.sqrt(((p / axes) ** 2).sum(axis=1)) <= 1.0)


@pytest.mark.parametrize("kind", ["sphere", "ellipsoid", "box"])
def test_compact_sources_are_isotropic_with_cauchy_mean_area(kind):
    """Directions are unit and isotropic; origins lie on (just outside) the
    boundary's upstream surface; the mean projected area over directions is a
    quarter of the boundary's surface area (Cauchy)."""
    s = _molecule()
    b = S.CompactBoundary.fit(s, kind)
    src = b.source(14, 881)
    n = 1 << 14
    d = src.directions
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert np.all(np.abs(d.mean(axis=0)) < 5e-3)
    area = src.weights * n
    a0, a1, a2 = b.axes
    if kind == "sphere":
        surface = 4 * math.pi * a0 * a0
        np.testing.assert_allclose(area, math.pi * a0 * a0, rtol=1e-15)
        r = np.linalg.norm(s.centers - s.centers.mean(axis=0), axis=1) + s.radii
        assert a0 == pytest.approx(r.max() * (1 + 1e-9), rel=1e-15)
    elif kind == "box":
        surface = 8 * (a0 * a1 + a1 * a2 + a2 * a0)
        p = src.origins - b.center + 1e-10 * max(1.0, b.axes.max()) * d
        on_face = np.isclose(np.abs(p), b.axes[None, :], rtol=1e-12, atol=1e-9)
        assert np.all(on_face.sum(axis=1) >= 1)
        assert np.all(np.abs(p) <= b.axes[None, :] * (1 + 1e-12) + 1e-9)
        # the face is an upstream face: the ray enters the box through it
        face = np.argmax(on_face, axis=1)
        assert np.all(p[np.arange(n), face] * d[np.arange(n), face] < 0)
    else:
        # Knud Thomsen's approximation of the ellipsoid surface (|error| < 1.1 %)
        pp = 1.6075
        surface = 4 * math.pi * (((a0 * a1) ** pp + (a1 * a2) ** pp + (a2 * a0) ** pp) / 3) ** (1 / pp)
        # every origin is upstream of the ellipsoid: the segment back from it misses the interior
        loc = (src.origins - b.center) @ b.rotation
        assert np.all(np.sqrt(((loc / b.axes) ** 2).sum(axis=1)) >= 1.0)
    tol = 0.012 if kind == "ellipsoid" else 3e-3
    assert area.mean() == pytest.approx(surface / 4, rel=tol)


def test_sphere_boundary_source_equals_the_external_formula():
    """The 'sphere' boundary is the enclosing-sphere impact disk of the
    external source, on a 4-D net."""
    s = _molecule()
    src = S.CompactBoundary.fit(s, "sphere").source(10, 5)
    u = qmc.Sobol(4, scramble=True, seed=5).random_base2(10)
    z = 1 - 2 * u[:, 0]
    np.testing.assert_allclose(src.directions[:, 2], z, rtol=0, atol=1e-15)
    assert np.all(src.last_atom == -1) and np.all(src.initial_bounces == 0)


def test_box_source_uses_five_dimensions_and_face_weights():
    s = Spheres(np.zeros((1, 3)), np.array([1.0]))
    b = S.CompactBoundary.fit(s, "box")
    np.testing.assert_allclose(b.axes, np.full(3, 1.0 + 1e-9))
    src = b.source(12, 3)
    u = qmc.Sobol(5, scramble=True, seed=3).random_base2(12)
    d = src.directions
    proj = np.abs(d) * (4 * np.prod(b.axes) / b.axes)
    np.testing.assert_allclose(src.weights * 4096, proj.sum(axis=1), rtol=1e-15)
    face = (u[:, 2, None] * proj.sum(axis=1)[:, None] >= np.cumsum(proj, axis=1)).sum(axis=1)
    counts = np.bincount(face, minlength=3) / 4096
    expected = proj.mean(axis=0) / proj.sum(axis=1).mean()
    np.testing.assert_allclose(counts, expected, atol=0.02)


def test_invalid_boundary_and_power():
    s = _molecule()
    with pytest.raises(ValueError, match="sphere, ellipsoid, or box"):
        S.CompactBoundary.fit(s, "cylinder")
    with pytest.raises(ValueError, match="Nonnegative integer"):
        S.CompactBoundary.fit(s, "box").source(-1, 0)


def test_sample_source_is_area_selected_cosine_hemisphere():
    c = np.array([[0.0, 0, 0], [5.0, 0, 0], [0, 7.0, 0]])
    r = np.array([1.0, 2.0, 3.0])
    s = Spheres(c, r)
    src = S.sample_source(s, 14, 7)
    np.testing.assert_allclose(np.bincount(src.owners, minlength=3) / (1 << 14), r ** 2 / (r ** 2).sum(), atol=2e-3)
    normal = (src.origins - c[src.owners]) / r[src.owners, None]
    np.testing.assert_allclose(np.linalg.norm(normal, axis=1), 1.0, atol=1e-12)
    cos = np.sum(normal * src.directions, axis=1)
    assert np.all(cos >= -1e-12)
    assert cos.mean() == pytest.approx(2 / 3, abs=5e-3)  # E[cos] under cosine weighting
    assert src.area == pytest.approx(4 * math.pi * (r ** 2).sum())

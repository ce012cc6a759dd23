"""CPU pins of the oracle's compact-boundary sources (ehss_compact_source.py:44-87).

Closed forms per ray fix what a plausible mistake would break: the box's
projected area sum_k |d_k| x face area and the face it enters through (the one
facing the ray, point inside its rectangle); the ellipsoid's projected area
pi abc |diag(1/axes) R^T d| under a general rotation (a transposed rotation or
swapped axes fail it); every ray enters the body; the mean weight is the
surface area / 4 (Cauchy's formula)."""
import numpy as np
import pytest


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _box_surface(axes):
    a, b, c = 2 * axes
    return 2 * (a * b + b * c + a * c)


def _ellipsoid_surface(axes, m=400):
    """Numerical surface area (midpoint rule on the parametric sphere)."""
    th = (np.arange(m) + 0.5) * np.pi / m
    ph = (np.arange(2 * m) + 0.5) * np.pi / m
    T, P = np.meshgrid(th, ph, indexing="ij")
    a, b, c = axes
    # |d/dth x d/dph| of (a sin th cos ph, b sin th sin ph, c cos th)
    nx = b * c * np.sin(T) ** 2 * np.cos(P)
    ny = a * c * np.sin(T) ** 2 * np.sin(P)
    nz = a * b * np.sin(T) * np.cos(T)
    return float(np.sum(np.sqrt(nx * nx + ny * ny + nz * nz)) * (np.pi / m) ** 2)


def test_box_closed_forms(oracle):
    center, axes = np.array([1.5, -2.0, 0.25]), np.array([3.0, 5.0, 2.0])
    src = oracle.compact_source("box", center, np.eye(3), axes, 14, 7)
    d, o, w = src["directions"], src["origins"], src["weights"]
    n = len(d)
    face_areas = 4 * np.prod(axes) / axes
    np.testing.assert_allclose(w * n, np.abs(d) @ face_areas, rtol=1e-15)
    p = o + 1e-10 * max(1., axes.max()) * d - center  # the point on the box surface
    on_face = np.isclose(np.abs(p), axes, rtol=0, atol=1e-12)
    assert np.all(on_face.sum(axis=1) >= 1)
    face = np.argmax(on_face, axis=1)
    # the ray enters through the face that faces it, inside the face's rectangle
    assert np.all(p[np.arange(n), face] * d[np.arange(n), face] <= 0)
    assert np.all(np.abs(p) <= axes + 1e-12)
    # faces are chosen in proportion to their projected areas
    share = np.array([np.sum(w[face == k]) for k in range(3)]) / w.sum()
    expect = (np.abs(d) * face_areas).sum(axis=0) / (np.abs(d) @ face_areas).sum()
    np.testing.assert_allclose(share, expect, atol=0.01)
    # Cauchy: mean projected area = surface / 4
    assert w.sum() == pytest.approx(_box_surface(axes) / 4, rel=2e-3)


def test_ellipsoid_closed_forms(oracle):
    center, axes, R = np.array([0.5, 1.0, -1.0]), np.array([4.0, 2.5, 1.5]), _rotation(3)
    src = oracle.compact_source("ellipsoid", center, R, axes, 14, 11)
    d, o, w = src["directions"], src["origins"], src["weights"]
    n = len(d)
    local = d @ R  # direction in the ellipsoid's frame
    expect_area = np.pi * np.prod(axes) * np.linalg.norm(local / axes, axis=1)
    np.testing.assert_allclose(w * n, expect_area, rtol=1e-12)
    # every ray meets the ellipsoid: |A^-1 (o + t d - c)| <= 1 at the closest approach
    A_inv = np.diag(1 / axes) @ R.T
    y0, yd = (o - center) @ A_inv.T, d @ A_inv.T
    t = -np.sum(y0 * yd, axis=1) / np.sum(yd * yd, axis=1)
    closest = np.linalg.norm(y0 + t[:, None] * yd, axis=1)
    assert np.all(closest <= 1 + 1e-9) and np.all(t > 0)
    assert w.sum() == pytest.approx(_ellipsoid_surface(axes) / 4, rel=2e-3)


def test_planted_mistakes_fail(oracle):
    center, axes, R = np.zeros(3), np.array([4.0, 2.5, 1.5]), _rotation(5)
    d = oracle.compact_source("ellipsoid", center, R, axes, 10, 2)["directions"]
    good = np.pi * np.prod(axes) * np.linalg.norm((d @ R) / axes, axis=1)
    transposed = np.pi * np.prod(axes) * np.linalg.norm((d @ R.T) / axes, axis=1)
    swapped = np.pi * np.prod(axes) * np.linalg.norm((d @ R) / axes[[1, 0, 2]], axis=1)
    assert np.max(np.abs(transposed / good - 1)) > 1e-3 and np.max(np.abs(swapped / good - 1)) > 1e-3
    # face areas of half-width products without the factor 4 miss Cauchy's mean
    src = oracle.compact_source("box", center, np.eye(3), axes, 14, 3)
    assert src["weights"].sum() == pytest.approx(_box_surface(axes) / 4, rel=2e-3)
    wrong = np.abs(src["directions"]) @ (np.prod(axes) / axes)
    assert wrong.mean() != pytest.approx(_box_surface(axes) / 4, rel=2e-2)


def test_single_sphere_omega_through_the_oracle_trace(oracle):
    """Any isotropic source gives a lone sphere Omega = pi r^2."""
    c, r = np.array([[0.3, -0.2, 0.1]]), np.array([1.7])
    tree = oracle.static_tree(c, r, 4)
    for kind, axes in (("box", np.array([2.5, 3.0, 2.0])), ("ellipsoid", np.array([2.5, 3.0, 2.0]))):
        src = oracle.compact_source(kind, c[0], _rotation(7), axes, 16, 5)
        tr = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=8, record=False)
        out = oracle.readout(src, tr, 8)
        assert out["omega"] == pytest.approx(np.pi * r[0] ** 2, rel=3e-3), kind

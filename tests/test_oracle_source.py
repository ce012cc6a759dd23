"""Pins of the oracle's source mapping (oracle.source_from_u, restating
ehss_state.py:42-52 with inside_out_pa.py:45-58) and of its whole-path C solve.

The source is pinned per ray by closed forms that fix which Sobol coordinate
drives which geometric quantity:
  d_z = 1 - 2 u0 exactly; the azimuth of d is 2 pi u1;
  the impact offset o - c + R d has length R sqrt(u2), lies in the plane
  orthogonal to d, and its azimuth in the (t, b) frame is 2 pi u3;
  t is a unit vector orthogonal to d, b = d x t; sum of weights = pi R^2.
A planted swap of u2 and u3 (or of cos and sin) must fail them.

oracle_solve_cpu (the C whole-path solve) must equal the NumPy path
(external_source -> trace -> readout) on the same points.
"""
import numpy as np
import pytest

from tests import scenes


def _closed_form_violations(u, src, center, radius):
    d, o = src["directions"], src["origins"]
    bad = {}
    bad["dz"] = int(np.sum(d[:, 2] != 1 - 2 * u[:, 0]))
    az = np.arctan2(d[:, 1], d[:, 0]) % (2 * np.pi)
    rho = np.hypot(d[:, 0], d[:, 1])
    diff = np.abs((az - 2 * np.pi * u[:, 1] + np.pi) % (2 * np.pi) - np.pi)
    bad["phi"] = int(np.sum((rho > 1e-6) & (diff > 1e-12 / np.maximum(rho, 1e-300))))
    bad["unit"] = int(np.sum(np.abs(np.linalg.norm(d, axis=1) - 1) > 4e-16))
    # frame from the directions (the mapping's tangent_frames rule)
    axes = np.zeros_like(d)
    axes[:, 2] = 1
    axes[np.abs(d[:, 2]) > .9] = [1, 0, 0]
    t = np.cross(axes, d)
    t /= np.linalg.norm(t, axis=1)[:, None]
    b = np.cross(d, t)
    off = o - center + radius * d
    ln = np.linalg.norm(off, axis=1)
    bad["radius"] = int(np.sum(np.abs(ln - radius * np.sqrt(u[:, 2])) > 1e-13 * radius))
    bad["plane"] = int(np.sum(np.abs(np.einsum("ij,ij->i", off, d)) > 1e-13 * radius))
    psi = np.arctan2(np.einsum("ij,ij->i", off, b), np.einsum("ij,ij->i", off, t)) % (2 * np.pi)
    dpsi = np.abs((psi - 2 * np.pi * u[:, 3] + np.pi) % (2 * np.pi) - np.pi)
    bad["psi"] = int(np.sum((ln > 1e-6 * radius) & (dpsi > 1e-12 * radius / np.maximum(ln, 1e-300))))
    bad["orth"] = int(np.sum(np.abs(np.einsum("ij,ij->i", t, d)) > 1e-15))
    bad["weights"] = int(not np.isclose(src["weights"].sum(), np.pi * radius * radius, rtol=1e-12, atol=0))
    return bad


@pytest.mark.parametrize("seed,power", [(0, 12), (104729, 10), (4294967295, 9)])
def test_source_mapping_closed_forms(oracle, seed, power):
    c, r = scenes.globule(650, 1)
    sv, shift = oracle.sobol_scramble(seed)
    u = oracle.sobol_points(sv, shift, power)
    src = oracle.source_from_u(u, c, r)
    center, radius = oracle.enclosing(c, r)
    bad = _closed_form_violations(u, src, center, radius)
    assert all(v == 0 for v in bad.values()), bad
    # the mapping's own frame equals the frame rebuilt from d (tangent_frames, inside_out_pa.py:52-58)
    t, b = oracle.tangent_frames(src["directions"])
    np.testing.assert_allclose(np.einsum("ij,ij->i", t, t), 1, atol=4e-16)
    np.testing.assert_allclose(np.cross(src["directions"], t), b, atol=0)


def test_planted_source_mistakes_fail_the_closed_forms(oracle):
    c, r = scenes.globule(650, 1)
    sv, shift = oracle.sobol_scramble(7)
    u = oracle.sobol_points(sv, shift, 11)
    center, radius = oracle.enclosing(c, r)
    swapped = u[:, [0, 1, 3, 2]]  # u2 <-> u3 inside the mapping
    bad = _closed_form_violations(u, oracle.source_from_u(swapped, c, r), center, radius)
    assert bad["radius"] > 0 and bad["psi"] > 0, bad
    swapped = u[:, [1, 0, 2, 3]]  # u0 <-> u1
    bad = _closed_form_violations(u, oracle.source_from_u(swapped, c, r), center, radius)
    assert bad["dz"] > 0 and bad["phi"] > 0, bad
    # cos and sin exchanged in the impact azimuth: o' reflects psi -> pi/2 - psi
    src = oracle.source_from_u(u, c, r)
    off = src["origins"] - center + radius * src["directions"]
    t, b = oracle.tangent_frames(src["directions"])
    a, bb = np.einsum("ij,ij->i", off, t), np.einsum("ij,ij->i", off, b)
    planted = dict(src, origins=center - radius * src["directions"] + bb[:, None] * t + a[:, None] * b)
    assert _closed_form_violations(u, planted, center, radius)["psi"] > 0


@pytest.mark.parametrize("config", ["C1", "C2"])
def test_c_solve_equals_numpy_path(oracle, config):
    """oracle_solve_cpu (ehss_oracle.c) == external_source -> trace -> readout on
    the same points: omega to summation order, PA, tail and counters exactly."""
    c, r = scenes.c60() if config == "C1" else scenes.globule(650, 1)
    # C2 at its full 2^20 rays: a one-ulp source difference shows up there as a
    # changed long path (it did when the compiler fused cos/sin into sincos).
    seed, power, cap = (0, 16, 64) if config == "C1" else (1, 20, 64)
    cpu = oracle.solve_cpu(c, r, seed, power, cap, threads=8)
    src = oracle.external_source(c, r, power, seed)
    tree = oracle.static_tree(c, r, 4)
    tr = oracle.trace(tree, c, r, src["origins"], src["directions"], cap=cap, record=False, threads=8)
    ro = oracle.readout(src, tr, cap)
    hits, _, _ = oracle.pa_query(oracle.pa_tree(c, r, 4), c, r, src["origins"], src["directions"], threads=8)
    assert cpu["omega"] == pytest.approx(ro["omega"], rel=1e-12)
    # PA and masses are sums of equal weights: compare the counts they imply
    # (the C solve adds ~10^5 weights in sequence per thread, 1e-12 off).
    w = src["weights"][0]
    assert round(cpu["pa"] / w) == int(hits.sum())
    assert round(cpu["pa_from_ehss"] / w) == int(((tr["bounces"] > 0) | tr["unresolved"]).sum())
    assert round(cpu["residual_mass"] / w) == int(tr["unresolved"].sum())
    assert cpu["pa"] == pytest.approx(float(src["weights"][hits].sum()), rel=1e-10)
    assert cpu["unresolved"] == int(tr["unresolved"].sum())
    assert cpu["box_tests"] == tr["box_tests"] and cpu["sphere_tests"] == tr["sphere_tests"]

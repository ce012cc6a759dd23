"""Alternative incidence sources traced on the device: the reference's own
reverse-PA tests (tests/test_ehss_reference.py:8-66, restated for the device
engine) and the compact-boundary cross-check of the reference's box comparison
(ehss_compact_source.py; supporting_information.md:31): every boundary gives
the same CCS as the enclosing-sphere impact disk within its sampling error."""
import math

import numpy as np
import pytest

from ehss_amd import sources as S
from ehss_amd.engine import GPUScene, scramble, replicate_seed
from ehss_amd.state import Spheres, external_source
from ehss_amd.static import PreparedStaticGPUEHSS
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("radius", [0.01, 1.0, 100.0])
def test_single_sphere_analytic_and_reverse_initialization(radius):
    s = Spheres(np.array([[2.0, -1.0, 4.0]]), np.array([radius]))
    prepared = PreparedStaticGPUEHSS(s)
    try:
        for factory in (external_source, S.reverse_pa_source):
            say("single sphere", radius, factory.__name__)
            source = factory(s, 14, 98)
            r = prepared.engine.trace(source, 4)
            assert r.escaped.all() and not r.unresolved.any()
            assert np.all(r.bounces == 1)
            assert r.omega == pytest.approx(math.pi * radius * radius, rel=1e-6)
            assert r.pa == pytest.approx(math.pi * radius * radius, rel=1e-7)
            assert np.allclose(np.linalg.norm(r.outgoing, axis=1), 1.0, atol=1e-14)
    finally:
        prepared.scene.close()


def test_multi_scattering_and_cap_bound():
    s = Spheres(np.array([[2.0, 0, 0], [-2.0, 0, 0], [0, 2.0, 0], [0, -2.0, 0], [0, 0, 2.0], [0, 0, -2.0]]),
                np.full(6, 1.25))
    prepared = PreparedStaticGPUEHSS(s)
    try:
        say("reverse PA, six spheres")
        source = S.reverse_pa_source(s, 14, 114)
        small = prepared.engine.trace(source, 1)
        full = prepared.engine.trace(source, 64)
        assert np.count_nonzero(full.bounces >= 2) > 100
        assert full.omega >= small.omega - 1e-12
        assert full.omega <= small.omega + small.tail_bound + 1e-12
        assert np.all(small.escaped | small.unresolved)
        assert not np.any(small.escaped & small.unresolved)
        assert full.source.weights.sum() == pytest.approx(full.source.weights[full.escaped].sum() + full.residual_mass)
    finally:
        prepared.scene.close()


def test_pa_filter_does_not_renormalize_exposed_states_and_overlap():
    s = Spheres(np.array([[0.0, 0, 0], [1.0, 0, 0]]), np.ones(2))
    prepared = PreparedStaticGPUEHSS(s)
    try:
        say("reverse PA, overlapping pair")
        source = S.reverse_pa_source(s, 14, 1)
        assert len(source.weights) < 2 ** 14
        assert np.all(source.weights == s.area / (4 * 2 ** 14))
        result = prepared.engine.trace(source)
        assert result.escaped.all()
        assert np.isfinite(result.omega)
    finally:
        prepared.scene.close()


def test_inside_out_pa_equals_the_external_pa():
    """Inside-out PA (area-weighted exposed surface / 4) and the external
    any-hit PA estimate the same projected area."""
    c, r = scenes.SCENES["c60"]()
    s = Spheres(c, r)
    say("inside-out PA")
    pa_in = np.mean([S.inside_out(s, S.sample_source(s, 16, replicate_seed(5, k)))[0] for k in range(4)])
    with GPUScene(c, r) as sc:
        pa_out = np.mean([sc.solve(scramble(replicate_seed(9, k)), 16, 64)["pa"] for k in range(4)])
    assert pa_in == pytest.approx(pa_out, rel=4e-3)


@pytest.mark.parametrize("kind", ["sphere", "ellipsoid", "box"])
def test_compact_boundaries_agree_with_the_impact_disk(kind):
    c, r = scenes.SCENES["c60"]()
    s = Spheres(c, r)
    prepared = PreparedStaticGPUEHSS(s, boundary=kind)
    try:
        assert prepared.boundary.kind == kind
        say("compact boundary", kind)
        vals = [prepared.engine.trace(prepared.boundary.source(16, replicate_seed(3, k)), 64).omega for k in range(6)]
        with GPUScene(c, r) as sc:
            ref = [sc.solve(scramble(replicate_seed(4, k)), 18, 64)["omega"] for k in range(4)]
        se = math.sqrt(np.var(vals, ddof=1) / len(vals) + np.var(ref, ddof=1) / len(ref))
        assert abs(np.mean(vals) - np.mean(ref)) <= 5 * se + 1e-3 * np.mean(ref), (kind, np.mean(vals), np.mean(ref))
        assert np.std(vals, ddof=1) / np.mean(vals) < 0.01
    finally:
        prepared.scene.close()


def test_compact_boundary_on_one_sphere_gives_pi_r2():
    s = Spheres(np.zeros((1, 3)), np.array([3.1]))
    for kind in ("ellipsoid", "box"):
        prepared = PreparedStaticGPUEHSS(s, boundary=kind)
        try:
            say("one sphere", kind)
            res = prepared.engine.trace(prepared.boundary.source(16, 11), 8)
            assert res.omega == pytest.approx(math.pi * 3.1 ** 2, rel=5e-3)
            assert res.pa == pytest.approx(math.pi * 3.1 ** 2, rel=5e-3)
            assert res.escaped.all() and np.all(res.bounces <= 1)
        finally:
            prepared.scene.close()

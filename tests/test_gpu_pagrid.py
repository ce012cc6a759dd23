"""Finite-grid (adaptive) PA on the device against the oracle: the covered-pixel
count bit-exact for the same projected disks, and the whole adaptive PA on the
device's own projection."""
import numpy as np
import pytest

from ehss_amd import adaptive_pa as G
from tests import scenes
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [8, 64, 512, 2048])
def test_grid_count_is_exact(oracle, n):
    rng = np.random.default_rng(n)
    for trial in range(6):
        k = int(rng.integers(1, 400))
        xy = rng.uniform(-12, 12, size=(k, 2))
        r = rng.uniform(0.0, 4.0, size=k)
        if trial == 1:
            xy[:] = xy[0]  # coincident disks
        if trial == 2:
            r[:] = 0.0  # degenerate disks still cover a midpoint they pass through exactly
        lower = np.array([-14.0, -14.0])
        sp = 28.0 / n
        say("grid count", n, trial, k)
        assert G.grid_count(xy, r, lower, sp, n) == oracle.grid_count(xy, r, lower, sp, n)[0]


def test_grid_count_on_projected_molecules(oracle):
    for name, (c, r) in (("globule650", scenes.globule(650, 1)), ("c60", scenes.SCENES["c60"]())):
        for d in oracle.adaptive_directions(8, 190003):
            xy, lower, sp, *_ = oracle.project_grid(c, r, d, 512)
            assert G.grid_count(xy, r, lower, sp, 512) == oracle.grid_count(xy, r, lower, sp, 512)[0], name


def test_adaptive_pa_matches_oracle(oracle):
    """The device projects with its own dot products, so a boundary pixel may
    flip where the projections differ in the last ulp: counts agree within a
    few pixels per direction and PA within 2e-5 relative."""
    class S:
        pass
    for name, (c, r) in (("globule650", scenes.globule(650, 1)), ("c60", scenes.SCENES["c60"]()),
                         ("ellipsoid2000", scenes.globule(2000, 2, (1., 1., 1.4)))):
        sph = S()
        sph.centers, sph.radii = c, r
        got = G.adaptive_pa(sph, 64, 512, 190003)
        exp, values = oracle.adaptive_pa(c, r, 64, 512, 190003)
        ds = oracle.adaptive_directions(64, 190003)
        counts, spacing = G.grid_counts(c, r, ds, 512)
        ref = []
        for d in ds:
            xy, lower, sp, *_ = oracle.project_grid(c, r, d, 512)
            ref.append(oracle.grid_count(xy, r, lower, sp, 512)[0])
        ref = np.array(ref)
        say(name, got["pa_angstrom2"], exp, "max |dcount|", int(np.abs(counts - ref).max()))
        assert np.abs(counts - ref).max() <= 8
        assert got["pa_angstrom2"] == pytest.approx(exp, rel=2e-5)


def test_single_sphere_adaptive_pa_is_pi_r2():
    class S:
        pass
    sph = S()
    sph.centers, sph.radii = np.zeros((1, 3)), np.array([3.1])
    got = G.adaptive_pa(sph, 16, 2048, 5)["pa_angstrom2"]
    assert got == pytest.approx(np.pi * 3.1 ** 2, rel=5e-3)


def test_grid_arguments_are_checked():
    from ehss_amd import _lib as L
    with pytest.raises(L.EHSSError, match="power of two"):
        G.grid_count(np.zeros((1, 2)), np.ones(1), np.array([-2.0, -2.0]), 0.5, 12)
    with pytest.raises(L.EHSSError, match="finite"):
        G.grid_count(np.array([[np.nan, 0.0]]), np.ones(1), np.array([-2.0, -2.0]), 0.5, 16)

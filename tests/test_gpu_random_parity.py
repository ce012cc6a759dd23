"""Randomised parity sweep: the FP32-filter kernel (default schedule, with one
packet query, with one lane per drained ray, and with the packet queries reading
an LDS-staged scene) against the plain FP64 kernel,
bit for bit (Omega, PA, every ray's contribution and physical flags, the bounce
histogram), on random molecules across scales, sizes and degenerate placements:
clouds at scales 1e-3 to 1e3, tangent-sphere lattices, coincident centres,
radii spread over four decades, and stacks of a dozen identical spheres (every
ray through a stack has more equal-entry candidates than a packet query's
candidate list holds, so those rays are handed to the queue drain, and the tie
rule picks the lowest id).  The FP32 filter's margins are relative to the
scene scale, so every scale must give the plain kernel's bits."""
import numpy as np
import pytest

from ehss_amd.engine import GPUScene, scramble
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

PHYS = ~np.uint32(16)  # every flag bit except the kernel-specific grazing diagnostic


def molecule(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "cloud":
        n = int(rng.integers(1, 2500))
        scale = (1e-3, 0.1, 1.0, 100.0, 1000.0)[seed % 5]
        c = rng.normal(scale=scale * n ** (1 / 3), size=(n, 3))
        r = rng.uniform(0.5, 2.0, size=n) * scale
    elif kind == "lattice":  # touching spheres: exact tangencies between neighbours
        k = int(rng.integers(2, 9))
        g = np.arange(k, dtype=np.float64) * 2.0
        c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        r = np.ones(len(c))
    elif kind == "coincident":  # repeated centres with different radii
        n = int(rng.integers(10, 400))
        base = rng.normal(scale=3.0, size=(n, 3))
        c = np.concatenate([base, base[: n // 2]])
        r = np.concatenate([rng.uniform(1.0, 2.0, n), rng.uniform(0.5, 2.5, n // 2)])
    elif kind == "stack":  # a cloud with identical spheres stacked at a few centres
        n = int(rng.integers(20, 600))
        c = rng.normal(scale=6.0, size=(n, 3))
        r = rng.uniform(1.0, 2.0, size=n)
        k = 12 + seed % 5
        tops = rng.normal(scale=4.0, size=(3, 3))
        c = np.concatenate([c] + [np.repeat(t[None], k, axis=0) for t in tops])
        r = np.concatenate([r, np.full(3 * k, 1.7)])
    else:  # "radii": four decades of radii
        n = int(rng.integers(50, 1500))
        c = rng.normal(scale=8.0, size=(n, 3))
        r = 10.0 ** rng.uniform(-3, 1, size=n)
    return c, r


CASES = [(kind, seed) for kind in ("cloud", "lattice", "coincident", "radii", "stack") for seed in range(5)]


@pytest.mark.parametrize("kind,seed", CASES)
def test_fp32_filter_kernels_equal_the_plain_kernel(kind, seed):
    c, r = molecule(kind, 1000 + seed)
    say("random parity", kind, seed, len(r), "atoms")
    power = 13
    with GPUScene(c, r) as sc:
        runs = {}
        for key, kw in (("fp64", dict(precision="fp64")), ("default", {}), ("packets1", dict(packets=1)),
                        ("lane", dict(group=1, stage1=0)), ("lds", dict(lds=True))):
            res = sc.solve(scramble(seed + 17), power, 64, pa=True, keep_rays=True, histogram=True, **kw)
            runs[key] = (res, sc.last_rays(1 << power))
        a = runs["fp64"]
        for key, b in runs.items():
            assert a[0]["omega"] == b[0]["omega"] and a[0]["pa"] == b[0]["pa"], (kind, seed, key)
            np.testing.assert_array_equal(a[1]["contributions"], b[1]["contributions"], err_msg=f"{kind}/{seed}/{key}")
            np.testing.assert_array_equal(a[1]["flags"] & PHYS, b[1]["flags"] & PHYS, err_msg=f"{kind}/{seed}/{key}")
            np.testing.assert_array_equal(a[0]["histogram"], b[0]["histogram"], err_msg=f"{kind}/{seed}/{key}")


@pytest.mark.parametrize("atoms", [550, 650, 750, 850])
def test_lds_staged_packets_near_the_staging_limit(atoms):
    """The packet kernels stage the scene in LDS beside their candidate lists
    (static LDS) when the scene is small enough; scenes around that limit
    (about 700 atoms) run staged or fall back, and either way equal the default
    kernel bit for bit."""
    from ehss_amd import molecules
    xyz, el = molecules.globule(atoms, 40 + atoms)
    r = molecules.radii_for(el)
    say("lds staging", atoms, "atoms")
    with GPUScene(xyz, r) as sc:
        a = sc.solve(scramble(3), 14, 64, pa=True, keep_rays=True, histogram=True)
        ra = sc.last_rays(1 << 14)
        b = sc.solve(scramble(3), 14, 64, pa=True, keep_rays=True, histogram=True, lds=True)
        rb = sc.last_rays(1 << 14)
    assert a["omega"] == b["omega"] and a["pa"] == b["pa"]
    np.testing.assert_array_equal(ra["contributions"], rb["contributions"])
    np.testing.assert_array_equal(ra["flags"] & PHYS, rb["flags"] & PHYS)
    np.testing.assert_array_equal(a["histogram"], b["histogram"])

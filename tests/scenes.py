"""Sphere sets shared by the tests: the reference tests' own fixtures (same
seeds and generators as the reference's tests/), analytic geometries, and
the synthetic benchmark molecules at reduced ray counts."""
from __future__ import annotations

import numpy as np


def fixture_29():
    """tests/test_ehss_cpu.py:86-88."""
    rng = np.random.default_rng(88271)
    return rng.normal(size=(29, 3)) * 3, rng.uniform(.8, 1.5, 29)


def fixture_37():
    """tests/test_ehss_native_ablation.py:10-12."""
    rng = np.random.default_rng(317)
    return rng.normal(size=(37, 3)) * 3, rng.uniform(.7, 1.5, 37)


def fixture_static(kind):
    """tests/test_ehss_static.py:10-19."""
    rng = np.random.default_rng(936)
    xyz = rng.normal(size=(47, 3)) * 4
    if kind == "ties":
        xyz = np.round(xyz)
    if kind == "coincident":
        xyz[:] = 0
    if kind == "single":
        xyz = xyz[:1]
    if kind == "translated":
        xyz += 1e6
    return xyz, rng.uniform(.8, 2.5, len(xyz))


def fixture_pose():
    """tests/test_ehss_publication_invariance.py:10-11."""
    return np.array([[0., 0., 0.], [4., 0., 0.], [1., 5., 0.], [.5, 1., 6.]]), np.full(4, 3.1)


def fixture_cross():
    """tests/test_ehss_reference.py:47-48 (six-sphere cross)."""
    return (np.array([[2., 0., 0.], [-2., 0., 0.], [0., 2., 0.], [0., -2., 0.], [0., 0., 2.], [0., 0., -2.]]),
            np.full(6, 1.25))


def fixture_23():
    """tests/test_ccs_cli.py:76-77."""
    rng = np.random.default_rng(667)
    return rng.normal(size=(23, 3)) * 2, np.full(23, 1.2)


def c60():
    from ehss_amd.molecules import c60 as make, radii_for
    xyz, el = make()
    return xyz, radii_for(el)


def globule(n, seed, axes=(1., 1., 1.)):
    from ehss_amd.molecules import globule as make, radii_for
    xyz, el = make(n, seed, axes)
    return xyz, radii_for(el)


SCENES = {
    "f29": fixture_29,
    "f37": fixture_37,
    "random47": lambda: fixture_static("random"),
    "ties47": lambda: fixture_static("ties"),
    "coincident47": lambda: fixture_static("coincident"),
    "single": lambda: fixture_static("single"),
    "translated47": lambda: fixture_static("translated"),
    "pose4": fixture_pose,
    "cross6": fixture_cross,
    "f23": fixture_23,
    "c60": c60,
}

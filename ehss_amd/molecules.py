"""Seeded synthetic molecules for the benchmark workloads and tests.

* ``c60()``: truncated icosahedron, 60 C atoms, bond length 1.40 A, centred.
* ``globule(n, seed, axes)``: random-walk chain of 1.5 A steps confined to an
  ellipsoid holding about 0.11 atoms/A^3, minimum pair distance 1.0 A, element
  mix H 49 / C 31 / O 9.5 / N 8.5 / S 1 % by largest remainder.
* ``conformer(base, seed, sigma)``: Gaussian coordinate jitter of a structure.

Radii are van der Waals radii plus the 1.4 A He probe, as ccs_cli.py:29-31,104-126
assigns them (H 2.50, C 3.10, N 2.95, O 2.92, S 3.20 A).
"""
from __future__ import annotations

import math

import numpy as np

VDW = {"H": 1.10, "He": 1.40, "C": 1.70, "N": 1.55, "O": 1.52, "F": 1.47, "Si": 2.10, "P": 1.80,
       "S": 1.80, "Cl": 1.75, "Se": 1.90, "Br": 1.83, "I": 1.98}
PROBE = 1.4
MIX = (("H", 0.49), ("C", 0.31), ("O", 0.095), ("N", 0.085), ("S", 0.01))
DENSITY = 0.11  # atoms / A^3


def radii_for(elements, probe: float = PROBE) -> np.ndarray:
    return np.array([VDW[e] + probe for e in elements], dtype=np.float64)


def c60(bond: float = 1.40):
    """Buckminsterfullerene: even permutations of (0, +-1, +-3phi), (+-1, +-(2+phi), +-2phi),
    (+-phi, +-2, +-phi^3), scaled so that the edge length is `bond`."""
    phi = (1 + math.sqrt(5)) / 2
    base = [(0.0, 1.0, 3 * phi), (1.0, 2 + phi, 2 * phi), (phi, 2.0, phi ** 3)]
    pts = set()
    for a, b, c in base:
        for sa in (-1, 1):
            for sb in (-1, 1):
                for sc in (-1, 1):
                    v = (sa * a, sb * b, sc * c)
                    for k in range(3):  # cyclic (even) permutations
                        pts.add(tuple(round(x, 12) for x in (v[k % 3], v[(k + 1) % 3], v[(k + 2) % 3])))
    xyz = np.array(sorted(pts), dtype=np.float64) * (bond / 2)
    if len(xyz) != 60:
        raise AssertionError("C60 construction must give 60 atoms")
    xyz -= xyz.mean(axis=0)
    return xyz, ["C"] * 60


def _element_counts(n: int):
    # The stated shares sum to 99 %; they are normalised before the split.
    total = sum(f for _, f in MIX)
    raw = [(e, f / total * n) for e, f in MIX]
    counts = {e: int(math.floor(v)) for e, v in raw}
    rest = n - sum(counts.values())
    order = sorted(raw, key=lambda ev: (-(ev[1] - math.floor(ev[1])), [e for e, _ in MIX].index(ev[0])))
    for e, _ in order[:rest]:
        counts[e] += 1
    return counts


def globule(n: int, seed: int, axes=(1.0, 1.0, 1.0), step: float = 1.5, min_dist: float = 1.0):
    """Random walk confined to an ellipsoid of semi-axes a*axes with n / volume = DENSITY."""
    rng = np.random.default_rng(seed)
    ax = np.asarray(axes, dtype=np.float64)
    a = (3 * n / (4 * math.pi * DENSITY * float(np.prod(ax)))) ** (1 / 3)
    semi = a * ax
    cell = min_dist
    grid: dict = {}
    xyz = np.empty((n, 3))

    def free(p):
        key = tuple(np.floor(p / cell).astype(int))
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.sum((xyz[j] - p) ** 2) < min_dist * min_dist:
                            return False
        return True

    def inside(p):
        return float(np.sum((p / semi) ** 2)) <= 1.0

    def place(i, p):
        xyz[i] = p
        grid.setdefault(tuple(np.floor(p / cell).astype(int)), []).append(i)

    place(0, np.zeros(3))
    cur = 0
    i = 1
    while i < n:
        for _ in range(64):
            v = rng.normal(size=3)
            p = xyz[cur] + step * v / np.linalg.norm(v)
            if inside(p) and free(p):
                place(i, p)
                cur = i
                i += 1
                break
        else:  # stuck: branch from a random earlier atom
            cur = int(rng.integers(0, i))
    counts = _element_counts(n)
    elements = [e for e, _ in MIX for _ in range(counts[e])]
    elements = [elements[k] for k in rng.permutation(n)]
    return xyz, elements


def conformer(xyz: np.ndarray, seed: int, sigma: float = 0.3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return xyz + rng.normal(scale=sigma, size=xyz.shape)


def workload(name: str):
    """(centers, radii, meta) of a named benchmark molecule."""
    if name == "c60":
        xyz, el = c60()
    elif name == "globule650":
        xyz, el = globule(650, 1)
    elif name == "ellipsoid2000":
        xyz, el = globule(2000, 2, axes=(1.0, 1.0, 1.4))
    else:
        raise ValueError(f"unknown workload {name}")
    return xyz, radii_for(el), dict(name=name, atoms=len(el))


def batch_set(count: int = 18, lo: int = 200, hi: int = 5000, first_seed: int = 3):
    """`count` globules with log-uniform atom counts in [lo, hi], seeds first_seed.."""
    rng = np.random.default_rng(12345)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(int)
    out = []
    for k, n in enumerate(sizes):
        xyz, el = globule(int(n), first_seed + k)
        out.append((xyz, radii_for(el)))
    return out


def write_xyz(path, xyz, elements, comment=""):
    with open(path, "w", encoding="utf-8") as f:
        f.write(f"{len(elements)}\n{comment}\n")
        for e, (x, y, z) in zip(elements, xyz):
            f.write(f"{e} {x:.6f} {y:.6f} {z:.6f}\n")

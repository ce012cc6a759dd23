"""Finite-grid ("adaptive") PA of the CLI, ccs_cli.py:142-160, on the device.

Directions: unit_sphere (inside_out_pa.py:45-49) of the first `directions`
points of a 2-D scrambled Sobol net with the replicate seed.  For each
direction the projection frame is the one of project_grid
(pa_walsh_geometry.py:171-180); the device projects the centred atoms, sizes the
grid and counts the midpoint pixels covered by the union of the projected
disks (ehss_pa_grid).  PA = mean over directions of count x spacing^2.
"""
from __future__ import annotations

import ctypes as ct
import time

import numpy as np

from . import _lib as L
from .engine import sobol_points


def unit_sphere(uv):
    """inside_out_pa.py:45-49."""
    z = 1 - 2 * uv[:, 0]
    phi = 2 * np.pi * uv[:, 1]
    rho = np.sqrt(np.maximum(0, 1 - z * z))
    return np.column_stack((rho * np.cos(phi), rho * np.sin(phi), z))


def projection_frame(direction):
    """pa_walsh_geometry.py:173-176: (u, v) of the plane orthogonal to the
    normalised direction."""
    direction = np.asarray(direction, float)
    direction = direction / np.linalg.norm(direction)
    axis = np.array([0., 0., 1.]) if abs(direction[2]) < .9 else np.array([1., 0., 0.])
    u = np.cross(axis, direction)
    u /= np.linalg.norm(u)
    v = np.cross(direction, u)
    return u, v


def grid_counts(centers, radii, directions, grid_size):
    """Covered-pixel counts and pixel sizes for each direction (device)."""
    c = np.ascontiguousarray(centers, dtype=np.float64)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    centered = np.ascontiguousarray(c - c.mean(axis=0))
    frames = np.ascontiguousarray(np.array([np.concatenate(projection_frame(d)) for d in directions]))
    counts = np.zeros(len(frames), np.int64)
    spacing = np.zeros(len(frames), np.float64)
    L.call("ehss_pa_grid", len(r), L.ptr(centered), L.ptr(r), len(frames), L.ptr(frames), int(grid_size),
           L.ptr(counts), L.ptr(spacing))
    return counts, spacing


def grid_count(xy, radii, lower, spacing, grid_size) -> int:
    """Covered-pixel count of caller-projected disks (device)."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    out = ct.c_int64(0)
    L.call("ehss_pa_grid_count", len(r), L.ptr(xy), L.ptr(r), L.ptr(lo), float(spacing), int(grid_size),
           ct.byref(out))
    return int(out.value)


def adaptive_pa(spheres, directions, grid_size, seed):
    """Same arguments and result fields as the reference's adaptive_pa."""
    tick = time.perf_counter()
    ds = unit_sphere(sobol_points(seed, 2, int(directions).bit_length() - 1))
    source_s = time.perf_counter() - tick
    tick = time.perf_counter()
    counts, spacing = grid_counts(spheres.centers, spheres.radii, ds, grid_size)
    query_s = time.perf_counter() - tick
    tick = time.perf_counter()
    values = [float(n) * float(s) ** 2 for n, s in zip(counts, spacing)]
    pa = float(np.mean(values))
    readout_s = time.perf_counter() - tick
    return dict(pa_angstrom2=pa, source_s=source_s, projection_s=0.0, query_s=query_s, readout_s=readout_s)

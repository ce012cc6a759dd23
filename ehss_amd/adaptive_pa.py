"""Adaptive grid PA (ccs_cli.py:142-160), the CLI's default --pa-method.

Not yet available on the GPU engine: select --pa-method bvh (any-hit BVH PA on
the EHSS rays) until the device rasteriser lands.
"""
from __future__ import annotations


def adaptive_pa(spheres, directions, grid_size, seed):
    raise ValueError("--pa-method adaptive is not available on the GPU backend yet; use --pa-method bvh")

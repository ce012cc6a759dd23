"""Any-hit PA index on the GPU.

Mirrors SphereIndex (pa_spatial.py:157-170) with kind 'bvh': the PA tree of
build_tree (pa_spatial.py:38-75) is built on the device and queried by the
any-hit kernel (pa_spatial.py:98-142); `.tree` exports the reference-identical
arrays (lo, hi, end, first, count, ids).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .engine import GPUScene
from .state import Spheres


@dataclass
class Tree:
    lo: np.ndarray
    hi: np.ndarray
    end: np.ndarray
    first: np.ndarray
    count: np.ndarray
    ids: np.ndarray

    @property
    def arrays(self):
        return (self.lo, self.hi, self.end, self.first, self.count, self.ids)

    @property
    def nbytes(self):
        return sum(x.nbytes for x in self.arrays)


class SphereIndex:
    def __init__(self, spheres: Spheres, kind: str = "bvh", leaf_size: int = 4):
        if kind != "bvh" or type(leaf_size) is not int or leaf_size < 1:
            raise ValueError("Valid kind ('bvh'), positive leaf size and nonempty primitives required.")
        self.spheres = spheres
        self.scene = GPUScene(spheres.centers, spheres.radii, leaf_size=leaf_size)
        self._tree = None

    @property
    def tree(self) -> Tree:
        if self._tree is None:
            self._tree = Tree(**self.scene.export_pa_tree())
        return self._tree

    def query(self, source):
        """(hits, box tests, sphere tests) for Source(origins, directions, owners, area)."""
        return self.scene.pa_query(source.origins, source.directions, source.owners)

    def refit(self, spheres):
        if len(spheres.radii) != len(self.spheres.radii):
            raise ValueError("Refit requires fixed primitive identities.")
        self.spheres = spheres
        self.scene.close()
        self.scene = GPUScene(spheres.centers, spheres.radii, leaf_size=self.scene.info["leaf_size"])
        self._tree = None

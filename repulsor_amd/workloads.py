"""Synthetic workloads of the tangent-point benchmarks (seeded, deterministic).

Each workload is a mesh generator call plus the energy exponents and the tree settings:
q=6, p=12 for surfaces and q=2, p=4 for the curve, far-field theta 0.5, near-field theta 10,
leaf size (split_threshold) 16, max_refinement 30 (the reference library defaults,
src/Settings.hpp:20-54, with split_threshold raised to 16).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import SimplicialMesh, make_icosphere, make_torus, make_trefoil


@dataclass(frozen=True)
class Workload:
    name: str
    kind: str          # "torus", "icosphere", "trefoil"
    args: tuple
    q: float
    p: float
    leaf: int = 16
    far_theta: float = 0.5
    near_theta: float = 10.0

    def geometry(self):
        if self.kind == "torus":
            return make_torus(*self.args)
        if self.kind == "icosphere":
            return make_icosphere(*self.args)
        if self.kind == "trefoil":
            return make_trefoil(*self.args)
        raise ValueError(self.kind)

    def mesh(self, threads=None, coords=None):
        X, S = self.geometry()
        if coords is not None:
            X = coords
        M = SimplicialMesh(X, S, thread_count=threads)
        M.cluster_tree_settings.split_threshold = self.leaf
        M.block_cluster_tree_settings.far_field_separation_parameter = self.far_theta
        M.adaptivity_settings.theta = self.near_theta
        return M


WORKLOADS = {
    # torus R=1, r=0.5, 60x20 grid, normal jitter sigma=1e-3 (seed 1): 2,400 triangles
    "torus_2400": Workload("torus_2400", "torus", (60, 20, 1.0, 0.5, 1e-3, 1), 6.0, 12.0),
    # icosphere level 6, spherical-harmonic radial noise l<=8, amplitude 0.05 (seed 2): 81,920 triangles
    "sphere_81920": Workload("sphere_81920", "icosphere", (6, 8, 0.05, 2), 6.0, 12.0),
    # trefoil + Fourier modes 1..16, amplitude 0.02 (seed 3): 200,000 edges
    "trefoil_200000": Workload("trefoil_200000", "trefoil", (200000, 16, 0.02, 3), 2.0, 4.0),
    # torus R=1, r=0.4, 320x240 grid: 153,600 triangles
    "torus_153600": Workload("torus_153600", "torus", (320, 240, 1.0, 0.4, 0.0, 4), 6.0, 12.0),
    # icosphere level 8, spherical-harmonic noise l<=16, amplitude 0.03 (seed 5): 1,310,720 triangles
    "sphere_1310720": Workload("sphere_1310720", "icosphere", (8, 16, 0.03, 5), 6.0, 12.0),
    # small cases for parity tests
    "sphere_1280": Workload("sphere_1280", "icosphere", (3, 4, 0.05, 7), 6.0, 12.0),
    "sphere_5120": Workload("sphere_5120", "icosphere", (4, 8, 0.05, 2), 6.0, 12.0),
    "trefoil_3000": Workload("trefoil_3000", "trefoil", (3000, 16, 0.02, 3), 2.0, 4.0),
}


def random_field(V, seed=11, nrhs=3):
    """X ~ N(0,1), V x nrhs (the MultiplyMetric input of the benchmarks)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((V, nrhs))

"""Cluster tree and block cluster tree built on the GPU (rp_build_topology) against the host build.

The host construction (include/repulsor/ClusterTree.hpp, BlockClusterTree.hpp) restates the
reference's ClusterTree split/serialisation and BlockClusterTree::ComputeBlocks; the device build must
reproduce it element for element:

* tree arrays (ordering, children, ranges, depths), cluster boxes and tree-ordered primitive data equal;
* far CSR, leaf blocks, very-near and intersecting pairs, near tiles (columns and row masks) equal;
* energy, differential and metric product of a device-built mesh equal the host-built mesh's bit for bit;
* surfaces and curves, leaf sizes 1..32, meshes with very-near and intersecting pairs, a fan whose every
  split takes the median fallback (with tied coordinates), and the 1.31M-triangle sphere;
* a fallback on more than 2048 primitives is refused with RP_EUNSUPPORTED.
"""

import time

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import vf_circles, vf_shells

pytestmark = pytest.mark.gpu


def _setup(M, leaf, far_theta=0.5, near_theta=10.0):
    M.cluster_tree_settings.split_threshold = leaf
    M.block_cluster_tree_settings.far_field_separation_parameter = far_theta
    M.adaptivity_settings.theta = near_theta


def _twins(X, S, leaf, far_theta=0.5, near_theta=10.0):
    Mh = R.SimplicialMesh(X, S)
    Md = R.SimplicialMesh(X, S)
    _setup(Mh, leaf, far_theta, near_theta)
    _setup(Md, leaf, far_theta, near_theta)
    Md.cluster_tree_settings.device_build = True
    t0 = time.perf_counter()
    bh = Mh.build()
    t1 = time.perf_counter()
    bd = Md.build()
    t2 = time.perf_counter()
    return Mh, Md, bh, bd, t1 - t0, t2 - t1


def _same(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.dtype.kind == "f":
        # doubles compare by value (the two builds may differ only in the sign of a zero)
        bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))).ravel())
    else:
        bad = np.flatnonzero((a != b).ravel())
    assert bad.size == 0, f"{what}: {bad.size} of {a.size} entries differ, first at {bad[:5]}"


def _compare(Mh, Md, bh, bd, name, full=True):
    assert bh == bd, (name, bh, bd)
    th, td = Mh.tree(), Md.tree()
    for k in th:
        _same(th[k], td[k], f"{name} tree {k}")
    dh, dd = Mh.tree_data(), Md.tree_data()
    for k in ("C_serialized", "P_serialized", "pre_cols", "P_near", "P_far", "C_far", "pre_vals"):
        _same(dh[k], dd[k], f"{name} {k}")
    for a, b, k in zip(Mh.pattern("far"), Md.pattern("far"), ("rp", "ci", "diag")):
        _same(a, b, f"{name} far {k}")
    for a, b, k in zip(Mh.leaf_blocks(), Md.leaf_blocks(), ("leaf blocks", "very near", "inter")):
        _same(a, b, f"{name} {k}")
    for a, b, k in zip(Mh.tiles(), Md.tiles(), ("col_ptr", "col_prim", "col_mask", "val_ptr")):
        _same(a, b, f"{name} tiles {k}")
    if full:
        for a, b, k in zip(Mh.pattern("verynear"), Md.pattern("verynear"), ("rp", "ci", "diag")):
            _same(a, b, f"{name} very near {k}")


def _energy_check(Mh, Md, q, p):
    E = R.TangentPointEnergy0(q, p)
    Yh, eh = E.Differential(Mh, return_energy=True)
    Yd, ed = E.Differential(Md, return_energy=True)
    assert eh == ed
    _same(Yh, Yd, "differential")
    A = R.TangentPointMetric0(q, p)
    X = random_field(Mh.V)
    Zh = np.zeros_like(X)
    Zd = np.zeros_like(X)
    A.MultiplyMetric(Mh, 1.0, X, 0.0, Zh)
    A.MultiplyMetric(Md, 1.0, X, 0.0, Zd)
    _same(Zh, Zd, "metric product")


@pytest.mark.parametrize("name,leaf", [("sphere_1280", 16), ("sphere_1280", 1), ("sphere_1280", 4), ("sphere_1280", 32),
                                       ("torus_2400", 16), ("trefoil_3000", 16), ("trefoil_3000", 1), ("trefoil_3000", 32)])
def test_device_build_small(name, leaf):
    wl = WORKLOADS[name]
    X, S = wl.geometry()
    Mh, Md, bh, bd, th, td = _twins(X, S, leaf, wl.far_theta, wl.near_theta)
    print(f"{name} leaf {leaf}: {bd} host {th * 1e3:.1f} ms, device build {td * 1e3:.1f} ms "
          f"(kernels {Md.build_device_ms():.2f} ms)", flush=True)
    _compare(Mh, Md, bh, bd, f"{name}/{leaf}")
    _energy_check(Mh, Md, wl.q, wl.p)


@pytest.mark.parametrize("kind,leaf", [("shells", 1), ("shells", 8), ("circles", 4)])
def test_device_build_very_near(kind, leaf):
    M0 = vf_shells(3, 0.004) if kind == "shells" else vf_circles(400, 3e-4)
    X, S = M0.VertexCoordinates().copy(), M0.simplices
    Mh, Md, bh, bd, _, _ = _twins(X, S, leaf)
    print(f"{kind} leaf {leaf}: {bd}", flush=True)
    assert bd["verynear_nnz"] > 0
    _compare(Mh, Md, bh, bd, f"{kind}/{leaf}")


def _two_spheres(level=3, shift=0.6):
    X, S = R.make_icosphere(level, 0, 0.0, 3)
    S = np.asarray(S, dtype=np.int32)
    Y = X + np.array([shift, 0.1, 0.05])
    return np.vstack([X, Y]), np.vstack([S, S + len(X)]).astype(np.int32)


def test_device_build_intersecting():
    X, S = _two_spheres()
    Mh, Md, bh, bd, _, _ = _twins(X, S, 8)
    print(f"two intersecting spheres: {bd}", flush=True)
    assert bd["inter_nnz"] > 0
    _compare(Mh, Md, bh, bd, "intersecting")


def _fan(n, jitter, seed=5):
    """n triangles around a far apex: every interior point lies below the box center along the long
    axis, so every split of more than split_threshold primitives takes the median fallback."""
    rng = np.random.default_rng(seed)
    t = 2 * np.pi * np.arange(n) / n
    ring = np.stack([jitter * rng.standard_normal(n), np.cos(t), np.sin(t)], 1)
    X = np.vstack([ring, [[10.0, 0.0, 0.0]]])
    S = np.stack([np.full(n, n), np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    return X, S


@pytest.mark.parametrize("jitter", [0.0, 1e-3])
def test_device_build_median_fallback(jitter):
    X, S = _fan(1500, jitter)
    Mh, Md, bh, bd, _, _ = _twins(X, S, 16)
    print(f"fan jitter {jitter}: {bd}", flush=True)
    _compare(Mh, Md, bh, bd, f"fan/{jitter}")


def test_device_build_large_fallback_refused():
    X, S = _fan(3000, 1e-3)
    M = R.SimplicialMesh(X, S)
    _setup(M, 16)
    M.cluster_tree_settings.device_build = True
    with pytest.raises(R.RepulsorError, match="2048"):
        M.build()


@pytest.mark.parametrize("name", ["sphere_81920", "trefoil_200000", "torus_153600", "sphere_1310720"])
def test_device_build_workloads(name):
    wl = WORKLOADS[name]
    X, S = wl.geometry()
    Mh, Md, bh, bd, th, td = _twins(X, S, wl.leaf, wl.far_theta, wl.near_theta)
    print(f"{name}: {bd}; host build {th:.2f} s, device build {td:.2f} s (kernels {Md.build_device_ms():.1f} ms)", flush=True)
    _compare(Mh, Md, bh, bd, name, full=False)
    if bd["T"] <= 200000:
        _energy_check(Mh, Md, wl.q, wl.p)

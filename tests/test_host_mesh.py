"""Host layer (CPU): generators, per-simplex data, cluster tree, block cluster tree.

Pins (none retypes the implementation's formula):
* generators: counts, closedness/manifoldness, Euler characteristic, exact torus implicit
  equation, icosphere radii and amplitude, trefoil perturbation amplitude;
* simplex data: area = |e1 x e2|/2 (|e| for edges), normal projector n n^T (I - t t^T), barycenter,
  and the surface-gradient operator applied to the coordinate functions = tangent projector;
* cluster tree: depth-first numbering, range partition, leaf sizes, tight AABBs, the
  longest-axis split rule, cluster data = area-weighted sums computed directly with numpy;
* block cluster tree: brute-force exact cover of all unordered primitive pairs by
  near/very-near/inter pairs and far blocks; AABB MAC on far blocks; leaf blocks fail it;
  near/far equivalence with the reference's BinaryMatrixCSR conventions.
"""

import numpy as np
import pytest

import oracle as O
import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS


# ---------------------------------------------------------------------------------------------- generators
def _edges(S):
    n1 = S.shape[1]
    e = []
    for a in range(n1):
        for b in range(a + 1, n1):
            e.append(np.sort(S[:, [a, b]], axis=1))
    return np.concatenate(e)


def test_torus_generator():
    X, S = R.make_torus(60, 20, 1.0, 0.5, 0.0, 1)
    assert X.shape == (1200, 3) and S.shape == (2400, 3)
    rho = np.sqrt(X[:, 0] ** 2 + X[:, 1] ** 2)
    assert np.max(np.abs((rho - 1.0) ** 2 + X[:, 2] ** 2 - 0.25)) < 1e-14
    E = np.unique(_edges(S), axis=0)
    _, cnt = np.unique(_edges(S), axis=0, return_counts=True)
    assert np.all(cnt == 2)                        # closed 2-manifold
    assert 1200 - len(E) + 2400 == 0               # Euler characteristic of the torus
    Xj, _ = R.make_torus(60, 20, 1.0, 0.5, 1e-3, 1)
    d = np.linalg.norm(Xj - X, axis=1)
    assert 0 < np.max(d) < 6e-3 and np.std(d) < 2e-3
    Xj2, _ = R.make_torus(60, 20, 1.0, 0.5, 1e-3, 1)
    assert np.array_equal(Xj, Xj2)                 # seeded


def test_icosphere_generator():
    X, S = R.make_icosphere(3, 4, 0.0, 0)
    assert X.shape == (642, 3) and S.shape == (1280, 3)
    assert np.max(np.abs(np.linalg.norm(X, axis=1) - 1.0)) < 1e-14
    _, cnt = np.unique(_edges(S), axis=0, return_counts=True)
    assert np.all(cnt == 2)
    assert 642 - len(cnt) + 1280 == 2              # sphere
    # outward orientation: signed volume close to 4 pi / 3
    v = np.einsum("ij,ij->i", X[S[:, 0]], np.cross(X[S[:, 1]], X[S[:, 2]])).sum() / 6
    assert 4.0 < v < 4 * np.pi / 3
    Xp, _ = R.make_icosphere(3, 4, 0.05, 7)
    r = np.linalg.norm(Xp, axis=1)
    assert abs(np.max(np.abs(r - 1.0)) - 0.05) < 1e-14   # max |delta r| = amplitude
    # radial only
    assert np.max(np.linalg.norm(np.cross(Xp, X), axis=1)) < 1e-14


def test_trefoil_generator():
    N = 3000
    X, S = R.make_trefoil(N, 16, 0.02, 3)
    X0, _ = R.make_trefoil(N, 0, 0.0, 3)
    t = 2 * np.pi * np.arange(N) / N
    ref = np.stack([np.sin(t) + 2 * np.sin(2 * t), np.cos(t) - 2 * np.cos(2 * t), -np.sin(3 * t)], axis=1)
    assert np.max(np.abs(X0 - ref)) < 1e-14
    assert abs(np.max(np.linalg.norm(X - ref, axis=1)) - 0.02) < 1e-14
    assert np.array_equal(S[:, 1], np.roll(S[:, 0], -1))


# ---------------------------------------------------------------------------------------------- simplex data
def _closed_forms(X, S):
    n = S.shape[1] - 1
    P0 = X[S[:, 0]]
    if n == 2:
        nu = np.cross(X[S[:, 1]] - P0, X[S[:, 2]] - P0)
        area = 0.5 * np.linalg.norm(nu, axis=1)
        nh = nu / np.linalg.norm(nu, axis=1)[:, None]
        P = nh[:, :, None] * nh[:, None, :]          # projector onto the normal line
    else:
        tau = X[S[:, 1]] - P0
        area = np.linalg.norm(tau, axis=1)
        th = tau / area[:, None]
        P = np.eye(3)[None] - th[:, :, None] * th[:, None, :]
    center = X[S].mean(axis=1)
    packed = P[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    return area, center, packed, P


@pytest.mark.parametrize("name", ["sphere_1280", "trefoil_3000"])
def test_simplex_data_closed_forms(name):
    wl = WORKLOADS[name]
    X, S = wl.geometry()
    n = S.shape[1] - 1
    area, center, packed, P = _closed_forms(X, S)
    near, far, dop, srt = O.simplex_data(X, S)
    assert np.max(np.abs(near[:, 0] - area) / area) < 1e-13
    assert np.max(np.abs(far[:, 1:4] - center)) < 1e-14
    assert np.max(np.abs(far[:, 4:] - packed)) < 1e-12
    assert np.array_equal(near[:, 1:1 + 3 * (n + 1)], X[S].reshape(len(S), -1))   # hull in input order
    # DiffOp applied to the coordinate functions = tangent projector I - P
    G = np.einsum("tkl,tlc->tkc", dop, X[srt])
    assert np.max(np.abs(G - (np.eye(3)[None] - P))) < 1e-12
    # product host data (tree order) equals the oracle rows permuted by P_ordering
    M = wl.mesh()
    M.build()
    td, tr = M.tree_data(), M.tree()
    perm = tr["P_ordering"]
    assert np.max(np.abs(td["P_near"] - near[perm])) <= 1e-13 * np.max(np.abs(near))
    assert np.max(np.abs(td["P_far"] - far[perm])) <= 1e-13 * np.max(np.abs(far))
    # mixed pre-operator rows: a/(n+1) averaging and a * DiffOp on the sorted simplex
    a = near[perm, 0]
    assert np.array_equal(td["pre_cols"], srt[perm])
    assert np.allclose(td["pre_vals"][:, 0, :], (a / (n + 1))[:, None], rtol=1e-14, atol=0)
    assert np.allclose(td["pre_vals"][:, 1:, :], a[:, None, None] * dop[perm], rtol=1e-12, atol=1e-14 * np.max(np.abs(a[:, None, None] * dop[perm])))


# ---------------------------------------------------------------------------------------------- cluster tree
@pytest.fixture(scope="module")
def torus_small():
    X, S = R.make_torus(24, 12, 1.0, 0.5, 1e-3, 1)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 4
    M.block_cluster_tree_settings.far_field_separation_parameter = 0.5
    M.build()
    return X, S, M


def test_cluster_tree_structure(torus_small):
    X, S, M = torus_small
    tr, td = M.tree(), M.tree_data()
    L, Rt, B, E = tr["C_left"], tr["C_right"], tr["C_begin"], tr["C_end"]
    Cn = len(L)
    T = len(S)
    assert B[0] == 0 and E[0] == T
    assert sorted(tr["P_ordering"]) == list(range(T))
    assert np.array_equal(tr["P_inverse_ordering"][tr["P_ordering"]], np.arange(T))
    desc = np.ones(Cn, dtype=np.int64)
    for c in range(Cn - 1, -1, -1):
        if L[c] >= 0:
            assert (Rt[c] >= 0) and L[c] == c + 1
            desc[c] = 1 + desc[L[c]] + desc[Rt[c]]
            assert Rt[c] == c + 1 + desc[L[c]]
            assert B[L[c]] == B[c] and E[L[c]] == B[Rt[c]] and E[Rt[c]] == E[c]
            assert B[c] < E[L[c]] < E[c]
        else:
            assert 1 <= E[c] - B[c] <= 4
    assert desc[0] == Cn
    leaves = np.where(L < 0)[0]
    assert np.array_equal(tr["leaf_clusters"], leaves)
    # tight AABBs of the hull points (frozen coordinates), [r^2, center, half extents]
    Pser = td["P_serialized"]
    hull = Pser[:, 4:].reshape(T, 3, 3)
    assert np.array_equal(hull, X[S[tr["P_ordering"]]])
    box = td["C_serialized"]
    for c in range(Cn):
        pts = hull[B[c]:E[c]].reshape(-1, 3)
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        assert np.allclose(box[c, 1:4], 0.5 * (lo + hi), rtol=0, atol=1e-15)
        assert np.allclose(box[c, 4:7], 0.5 * (hi - lo), rtol=0, atol=1e-15)
        assert abs(box[c, 0] - np.sum((0.5 * (hi - lo)) ** 2)) < 1e-15
    # split rule: longest axis of the parent box, primitives' interior point < box center go left
    centers = Pser[:, 1:4]
    for c in range(Cn):
        if L[c] < 0:
            continue
        k = int(np.argmax(box[c, 4:7]))
        mid = box[c, 1 + k]
        left = centers[B[c]:E[L[c]], k]
        right = centers[B[Rt[c]]:E[c], k]
        median_fallback = np.all(centers[B[c]:E[c], k] < mid) or np.all(centers[B[c]:E[c], k] >= mid)
        if not median_fallback:
            assert np.all(left < mid) and np.all(right >= mid)


def _reference_split_order(hull, centers, leaf):
    """The reference's tree construction order (AABB_LongestAxisSplit::split,
    src/BoundingVolumes/AABB_LongestAxisSplit.hpp:134-253): per range, the box of the hull points, the
    longest axis, then the in-place scan that swaps every primitive whose interior point lies below the
    box center to the front; if all lie on one side, the range is ordered by that coordinate (stably,
    the order this build fixes for the nth_element fallback) and split at the middle."""
    order = np.arange(len(centers))
    stack = [(0, len(centers))]
    while stack:
        b, e = stack.pop()
        if e - b <= leaf:
            continue
        pts = hull[order[b:e]].reshape(-1, 3)
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        diff = 0.5 * (hi - lo)
        mid_all = lo + diff
        k = 0
        for a in (1, 2):
            if diff[a] > diff[k]:
                k = a
        mid = mid_all[k]
        s = b
        for i in range(b, e):
            if centers[order[i], k] < mid:
                order[s], order[i] = order[i], order[s]
                s += 1
        if s == b or s == e:
            s = b + (e - b) // 2
            seg = order[b:e]
            order[b:e] = seg[np.argsort(centers[seg, k], kind="stable")]
        stack.append((s, e))
        stack.append((b, s))
    return order


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_cluster_tree_order_is_the_reference_partition(leaf):
    """P_ordering equals the order the reference's swap scan produces (restated loop by loop in Python)."""
    X, S = R.make_torus(30, 12, 1.0, 0.4, 1e-3, 3)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = leaf
    tr, td = M.tree(), M.tree_data()
    T = len(S)
    inv = tr["P_inverse_ordering"]
    hull = td["P_serialized"][:, 4:].reshape(T, 3, 3)[inv]       # original order
    centers = td["P_serialized"][:, 1:4][inv]
    assert np.array_equal(tr["P_ordering"], _reference_split_order(hull, centers, leaf))


def test_cluster_tree_median_fallback_order():
    """A fan of triangles around a far apex: every interior point lies below the box center along the
    long axis, so every split is the median fallback, and the tree order is the stable order of the
    interior points' coordinate along that axis (ties keep the input order)."""
    n = 600
    rng = np.random.default_rng(5)
    t = 2 * np.pi * np.arange(n) / n
    for jitter in (0.0, 1e-3):
        ring = np.stack([jitter * rng.standard_normal(n), np.cos(t), np.sin(t)], 1)
        X = np.vstack([ring, [[10.0, 0.0, 0.0]]])
        S = np.stack([np.full(n, n), np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
        M = R.SimplicialMesh(X, S)
        M.cluster_tree_settings.split_threshold = 16
        tr, td = M.tree(), M.tree_data()
        bx = td["P_serialized"][:, 1][tr["P_inverse_ordering"]]
        assert np.array_equal(tr["P_ordering"], np.argsort(bx, kind="stable"))
        L, B, E = tr["C_left"], tr["C_begin"], tr["C_end"]
        for c in np.flatnonzero(L >= 0):
            assert E[L[c]] == B[c] + (E[c] - B[c]) // 2


def test_cluster_data(torus_small):
    X, S, M = torus_small
    tr, td = M.tree(), M.tree_data()
    far = td["P_far"]            # tree order
    for c in range(len(tr["C_left"])):
        rows = far[tr["C_begin"][c]:tr["C_end"][c]]
        a = rows[:, 0]
        A = a.sum()
        assert abs(td["C_far"][c, 0] - A) <= 1e-13 * A
        assert np.allclose(td["C_far"][c, 1:], (a[:, None] * rows[:, 1:]).sum(axis=0) / A, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------------------------- block cluster tree
def _aabb_mac(box, c, d, theta2):
    P, Q = box[c], box[d]
    lo = np.maximum(P[1:4] - P[4:7], Q[1:4] - Q[4:7])
    hi = np.minimum(P[1:4] + P[4:7], Q[1:4] + Q[4:7])
    d2 = np.sum(np.maximum(0.0, lo - hi) ** 2)
    return max(P[0], Q[0]) < theta2 * d2


@pytest.mark.parametrize("far_theta,leaf", [(0.5, 4), (0.25, 1), (0.5, 16)])
def test_bct_exact_cover(far_theta, leaf):
    X, S = R.make_torus(24, 12, 1.0, 0.5, 1e-3, 1)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = leaf
    M.block_cluster_tree_settings.far_field_separation_parameter = far_theta
    b = M.build()
    T = b["T"]
    tr, td = M.tree(), M.tree_data()
    cover = np.zeros((T, T), dtype=np.int32)
    nrp, nci, ndg = M.pattern("near")
    rows = np.repeat(np.arange(T), np.diff(nrp))
    cover[rows, nci] += 1
    assert np.all(nci[ndg[rows] <= np.arange(len(nci))] > rows[ndg[rows] <= np.arange(len(nci))])
    assert np.array_equal(cover, cover.T)      # symmetric pattern
    frp, fci, fdg = M.pattern("far")
    Bc, Ec = tr["C_begin"], tr["C_end"]
    box = td["C_serialized"]
    for c in range(len(frp) - 1):
        for k in range(frp[c], frp[c + 1]):
            d = fci[k]
            assert d != c
            assert _aabb_mac(box, c, d, far_theta ** 2)
            cover[Bc[c]:Ec[c], Bc[d]:Ec[d]] += 1
    blocks, vf, inter = M.leaf_blocks()
    for p in list(vf) + list(inter):
        cover[p[0], p[1]] += 1
        cover[p[1], p[0]] += 1
    off = ~np.eye(T, dtype=bool)
    assert np.all(cover[off] == 1)             # every unordered pair covered exactly once
    assert np.all(np.diag(cover) == 0)
    lc = tr["leaf_clusters"]
    for I, J in blocks:
        if I != J:
            assert not _aabb_mac(box, lc[I], lc[J], far_theta ** 2)
    assert b["near_nnz"] == len(nci) and b["far_nnz"] == len(fci)


def _qp_distance(P, Q):
    """Distance between two triangles by minimising |sum l_i P_i - sum m_j Q_j| over the two
    probability simplices (SLSQP), independent of the library's closed-point queries."""
    from scipy.optimize import minimize

    def f(z):
        d = z[:3] @ P - z[3:] @ Q
        return d @ d

    cons = [{"type": "eq", "fun": lambda z: z[:3].sum() - 1}, {"type": "eq", "fun": lambda z: z[3:].sum() - 1}]
    best = np.inf
    for z0 in (np.full(6, 1 / 3), np.array([1, 0, 0, 1, 0, 0.0]), np.array([0, 1, 0, 0, 0, 1.0])):
        r = minimize(f, z0, bounds=[(0, 1)] * 6, constraints=cons, method="SLSQP", options={"ftol": 1e-16, "maxiter": 200})
        best = min(best, r.fun)
    return max(best, 0.0)


def test_near_pairs_adjacency_or_mac():
    """Near pairs either share a vertex or satisfy the primitive MAC max(r^2) < theta_near^2 dist^2
    (reference BlockSplit_Kernel.hpp:186-192), with dist from an independent QP on a sample."""
    X, S = R.make_torus(24, 12, 1.0, 0.5, 1e-3, 1)
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 8
    M.build()
    tr, td = M.tree(), M.tree_data()
    nrp, nci, ndg = M.pattern("near")
    Pser = td["P_serialized"]
    St = S[tr["P_ordering"]]
    T = len(St)
    rows = np.repeat(np.arange(T), np.diff(nrp))
    adj = np.array([len(set(St[i]) & set(St[j])) > 0 for i, j in zip(rows, nci)])
    assert adj.any() and (~adj).any()
    idx = np.where(~adj)[0]
    # closest non-adjacent pairs are the hardest MAC decisions
    cdist = np.linalg.norm(Pser[rows[idx], 1:4] - Pser[nci[idx], 1:4], axis=1)
    pick = idx[np.argsort(cdist)[:60]]
    for k in pick:
        i, j = rows[k], nci[k]
        d2 = _qp_distance(Pser[i, 4:].reshape(3, 3), Pser[j, 4:].reshape(3, 3))
        assert max(Pser[i, 0], Pser[j, 0]) < 100.0 * d2
    assert M.build()["verynear_nnz"] == 0


@pytest.mark.parametrize("name", ["torus_2400", "sphere_5120", "trefoil_3000"])
def test_config_meshes_have_no_very_near_field(name):
    wl = WORKLOADS[name]
    M = wl.mesh()
    b = M.build()
    assert b["verynear_nnz"] == 0 and b["inter_nnz"] == 0
    # tiles carry exactly the near pairs (masked count) and one slot per (row, column)
    assert b["tile_slots"] >= b["near_nnz"]

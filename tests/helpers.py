"""Shared helpers of the test suite: run the oracle on the same tree and patterns as the product."""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def oracle_state(M):
    """Tree, tree-ordered data and reference-layout patterns of a built mesh."""
    M.build()
    tr = M.tree()
    td = M.tree_data()
    npat = M.pattern("near")
    fpat = M.pattern("far")
    return tr, td, npat, fpat


def oracle_vf(M, q, p, flags):
    """Oracle very-near field of M (None if its VeryNear pattern is empty) and the pattern."""
    if M.build()["verynear_nnz"] == 0:
        return None, None
    pat = M.pattern("verynear")
    td = M.tree_data()
    return O.vf_compute(M.dom_dim, td["P_near"], pat, M.adaptivity_settings.theta,
                        M.adaptivity_settings.max_refinement, q, p, flags), pat


def oracle_energy_diff(M, q, p, X=None, charges=None):
    """Oracle energy and vertex differential (V x 3) on M's hierarchy (very-near field included)."""
    tr, td, npat, fpat = oracle_state(M)
    n = M.dom_dim
    o = O.tp0_compute(n, td["P_near"], td["C_far"], npat, fpat, q, p, 3)
    vf, _ = oracle_vf(M, q, p, 3)
    if vf is not None:
        o["energy"] += vf["energy"]
        o["P_D_near_tree"] = o["P_D_near_tree"] + vf["P_D_near_tree"]
    V = M.VertexCoordinates() if X is None else X
    Y, N, F = O.differential(n, V, M.simplices, tr, o["P_D_near_tree"], o["C_D_far"], charges)
    return o["energy"], Y, N, F


def oracle_metric(M, q, p):
    tr, td, npat, fpat = oracle_state(M)
    o = O.tp0_compute(M.dom_dim, td["P_near"], td["C_far"], npat, fpat, q, p, 4)
    return o, tr, td, npat, fpat


def with_verynear_metric(M, q, p, o, npat):
    """Near pattern and values with the very-near blocks merged in (same block form), and the near
    diagonal plus the very-near diagonal, for the oracle's metric product."""
    vf, vpat = oracle_vf(M, q, p, 4)
    if vf is None:
        return npat, o["NF"], o["NF_diag"]
    pat, vals = O.merge_patterns(npat, o["NF"], vpat, vf["VF"])
    return pat, vals, o["NF_diag"] + vf["VF_diag"]


def oracle_multiply(M, q, p, X, alpha=1.0, beta=0.0, Y=None, nf=True, ff=True, vf=True):
    o, tr, td, npat, fpat = oracle_metric(M, q, p)
    Y = np.zeros_like(X) if Y is None else Y.copy()
    if vf and nf:
        npat, nvals, ndiag = with_verynear_metric(M, q, p, o, npat)
    elif vf:
        v, vpat = oracle_vf(M, q, p, 4)
        npat, nvals, ndiag = (vpat, v["VF"], v["VF_diag"]) if v is not None else (npat, o["NF"], np.zeros_like(o["NF_diag"]))
        nf = v is not None
    else:
        nvals, ndiag = o["NF"], o["NF_diag"]
    O.multiply_metric(M.dom_dim, M.V, tr, td["pre_cols"], td["pre_vals"], npat, nvals, ndiag,
                      fpat, o["FF"], o["FF_diag"], alpha, X, beta, Y, nf, ff)
    return Y


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def h1_matrix(M):
    """The host-assembled H^1 metric as a scipy CSR matrix."""
    rp, ci, v = M.h1_csr()
    return sp.csr_matrix((v, ci, rp), shape=(M.V, M.V))


class OracleSolve:
    """A (oracle metric), P0 = H^-1 L H^-1 (scipy LU of H + oracle pseudo-Laplacian)."""

    def __init__(self, M, q, p):
        self.M = M
        n = M.dom_dim
        s = (p - n) / q
        self.o, self.tr, self.td, self.npat, self.fpat = oracle_metric(M, q, p)
        self.pl = O.pl_compute(n, self.td["P_near"], self.td["C_far"], self.npat, self.fpat, 2.0 - s)
        # very-near blocks merged into the near pattern of both operators
        self.mpat, self.mvals, self.mdiag = with_verynear_metric(M, q, p, self.o, self.npat)
        self.plpat = self.npat
        if M.build()["verynear_nnz"] > 0:
            vpat = M.pattern("verynear")
            v = O.pl_vf(n, self.td["P_near"], vpat, M.adaptivity_settings.theta, 20, 2.0 - s)
            self.plpat, nfv = O.merge_patterns(self.npat, self.pl["NF"], vpat, v["VF"])
            self.pl = dict(self.pl, NF=nfv, NF_diag=self.pl["NF_diag"] + v["VF_diag"])
        self.lu = spla.splu(h1_matrix(M).tocsc())

    def A(self, X):
        M, td = self.M, self.td
        Y = np.zeros_like(X)
        return O.multiply_metric(M.dom_dim, M.V, self.tr, td["pre_cols"], td["pre_vals"], self.mpat, self.mvals, self.mdiag,
                                 self.fpat, self.o["FF"], self.o["FF_diag"], 1.0, X, 0.0, Y)

    def P(self, X):
        M, td = self.M, self.td
        Z = np.ascontiguousarray(self.lu.solve(np.ascontiguousarray(X)))
        Y = np.zeros(Z.shape)
        O.pl_multiply(M.dom_dim, M.V, self.tr, td["pre_cols"], td["pre_vals"], self.plpat, self.pl, self.fpat, 1.0, Z, 0.0, Y)
        return np.ascontiguousarray(self.lu.solve(Y))

    def diagonal(self):
        """diag(A) by probing: A applied to the unit vector of every vertex."""
        V = self.M.V
        d = np.zeros(V)
        for v in range(V):
            e = np.zeros((V, 3)); e[v, 0] = 1.0
            d[v] = self.A(e)[v, 0]
        return d

    def with_diagonal_term(self, w, diag):
        """Switches P to P0 + w diag^-1."""
        P0 = self.P
        self.P = lambda X: P0(X) + w * X / diag[:, None]
        return self

    def pcg(self, B, tol=1e-6, max_iter=500, measure="l2"):
        """Textbook PCG per column (all columns iterate together); measure "l2" stops on
        |r|_2 <= tol |b|_2, "pc" on <r, P r> <= tol^2 <b, P b>."""
        X = np.zeros_like(B); Rr = B.copy(); Z = self.P(Rr); Pd = Z.copy()
        rho = np.sum(Rr * Z, axis=0); bb = np.sum(B * B, axis=0); rho0 = rho.copy()
        for it in range(1, max_iter + 1):
            AP = self.A(Pd)
            al = rho / np.sum(Pd * AP, axis=0)
            X += al * Pd; Rr -= al * AP
            if measure == "l2" and np.all(np.sum(Rr * Rr, axis=0) <= tol * tol * bb):
                return X, it
            Z = self.P(Rr)
            rn = np.sum(Rr * Z, axis=0)
            if measure == "pc" and np.all(rn <= tol * tol * rho0):
                return X, it
            Pd = Z + (rn / rho) * Pd
            rho = rn
        return X, max_iter

    def gmres(self, B, tol=1e-6, max_iter=500, restart=100):
        """Textbook left-preconditioned restarted GMRES (modified Gram-Schmidt, Givens rotations), one
        column at a time; stops on |P r|_2 <= tol |P b|_2. Returns X and the largest step count."""
        X = np.zeros_like(B)
        its = 0
        for k in range(B.shape[1]):
            b = B[:, k]
            def op(v):
                Z = np.zeros_like(B); Z[:, k] = v
                return self.P(self.A(Z))[:, k]
            def pc(v):
                Z = np.zeros_like(B); Z[:, k] = v
                return self.P(Z)[:, k]
            x = np.zeros(len(b)); r = pc(b); beta0 = np.linalg.norm(r); it = 0
            while it < max_iter and beta0 > 0:
                beta = np.linalg.norm(r)
                Vb = [r / beta]; Hm = np.zeros((restart + 1, restart)); g = np.zeros(restart + 1); g[0] = beta
                cs = np.zeros(restart); sn = np.zeros(restart); j = 0; done = False
                while j < restart and it < max_iter:
                    w = op(Vb[j])
                    for i in range(j + 1):
                        Hm[i, j] = Vb[i] @ w; w = w - Hm[i, j] * Vb[i]
                    Hm[j + 1, j] = np.linalg.norm(w); Vb.append(w / Hm[j + 1, j] if Hm[j + 1, j] > 0 else 0 * w)
                    for i in range(j):
                        a, c = Hm[i, j], Hm[i + 1, j]
                        Hm[i, j], Hm[i + 1, j] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
                    d = np.hypot(Hm[j, j], Hm[j + 1, j]); cs[j], sn[j] = Hm[j, j] / d, Hm[j + 1, j] / d
                    Hm[j, j] = d; Hm[j + 1, j] = 0; g[j + 1] = -sn[j] * g[j]; g[j] = cs[j] * g[j]
                    j += 1; it += 1
                    if abs(g[j]) <= tol * beta0:
                        done = True
                        break
                y = np.linalg.solve(np.triu(Hm[:j, :j]), g[:j])
                x = x + np.array(Vb[:j]).T @ y
                if done:
                    break
                Z = np.zeros_like(B); Z[:, k] = x
                r = pc(b - self.A(Z)[:, k])
            X[:, k] = x
            its = max(its, it)
        return X, its


# ---------------------------------------------------------------------------------------------
# Meshes with a non-empty very-near field: primitives of different components closer than a tenth
# of their radius.

def vf_shells(level=3, gap=0.004, seed=7):
    """Two nested icospheres (unit sphere and a radially jittered copy at distance ~gap, reversed
    orientation): every triangle has very-near partners on the other shell."""
    import repulsor_amd as R
    X, S = R.make_icosphere(level, 0, 0.0, seed)
    X = X / np.linalg.norm(X, axis=1)[:, None]
    rng = np.random.default_rng(seed)
    Y = X * (1.0 + gap * (1.0 + 0.5 * rng.uniform(-1, 1, len(X))))[:, None]
    S = np.asarray(S, dtype=np.int32)
    return R.SimplicialMesh(np.vstack([X, Y]), np.vstack([S, S[:, ::-1] + len(X)]).astype(np.int32))


def vf_circles(N=400, gap=3e-4, seed=8):
    """Two unit circles with N edges each, the second lifted by ~gap along z (jittered)."""
    import repulsor_amd as R
    t = 2 * np.pi * np.arange(N) / N
    rng = np.random.default_rng(seed)
    A = np.stack([np.cos(t), np.sin(t), np.zeros(N)], 1)
    B = A.copy()
    B[:, 2] = gap * (1 + 0.5 * rng.uniform(-1, 1, N))
    E = np.stack([np.arange(N), (np.arange(N) + 1) % N], 1)
    return R.SimplicialMesh(np.vstack([A, B]), np.vstack([E, E + N]).astype(np.int32))

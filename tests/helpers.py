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


def oracle_energy_diff(M, q, p, X=None, charges=None):
    """Oracle energy and vertex differential (V x 3) on M's hierarchy."""
    tr, td, npat, fpat = oracle_state(M)
    n = M.dom_dim
    o = O.tp0_compute(n, td["P_near"], td["C_far"], npat, fpat, q, p, 3)
    V = M.VertexCoordinates() if X is None else X
    Y, N, F = O.differential(n, V, M.simplices, tr, o["P_D_near_tree"], o["C_D_far"], charges)
    return o["energy"], Y, N, F


def oracle_metric(M, q, p):
    tr, td, npat, fpat = oracle_state(M)
    o = O.tp0_compute(M.dom_dim, td["P_near"], td["C_far"], npat, fpat, q, p, 4)
    return o, tr, td, npat, fpat


def oracle_multiply(M, q, p, X, alpha=1.0, beta=0.0, Y=None, nf=True, ff=True):
    o, tr, td, npat, fpat = oracle_metric(M, q, p)
    Y = np.zeros_like(X) if Y is None else Y.copy()
    O.multiply_metric(M.dom_dim, M.V, tr, td["pre_cols"], td["pre_vals"], npat, o["NF"], o["NF_diag"],
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
        self.lu = spla.splu(h1_matrix(M).tocsc())

    def A(self, X):
        M, td = self.M, self.td
        Y = np.zeros_like(X)
        return O.multiply_metric(M.dom_dim, M.V, self.tr, td["pre_cols"], td["pre_vals"], self.npat, self.o["NF"], self.o["NF_diag"],
                                 self.fpat, self.o["FF"], self.o["FF_diag"], 1.0, X, 0.0, Y)

    def P(self, X):
        M, td = self.M, self.td
        Z = np.ascontiguousarray(self.lu.solve(np.ascontiguousarray(X)))
        Y = np.zeros(Z.shape)
        O.pl_multiply(M.dom_dim, M.V, self.tr, td["pre_cols"], td["pre_vals"], self.npat, self.pl, self.fpat, 1.0, Z, 0.0, Y)
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

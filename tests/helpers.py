"""Shared helpers of the test suite: run the oracle on the same tree and patterns as the product."""

import numpy as np

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

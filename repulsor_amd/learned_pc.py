"""Learned preconditioner model: parameter layout, initialisation and model files.

P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2}; a 2-layer MLP of width 64 per cluster-tree node maps the node's
features (include/repulsor/LearnedPreconditioner.hpp) to s_c = softplus(w2 . tanh(W1 f + b1) + b2);
log d_v averages, over the primitives containing v, the sums of log s_c along their leaf-to-root
paths, smoothed by 16 steps of vertex means of primitive means. Parameters (float32): W1 (64 x F, row-major), b1 (64), w2 (64), b2 (1), log w (1); the weight w
is exp(log w) times the base weight (a fixed positive setting, else the estimate lambda_max(P0 A) / 4).

Model file: b"RPLP", int32 F, int32 parameter count, float32 parameters (the C++ reader is
ReadLearnedPreconditioner).
"""

from __future__ import annotations

import os
import struct

import numpy as np

HIDDEN = 64
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "learned_pc.rplp")
TRAINED_MODEL = os.path.join(HERE, "models", "learned_pc_trained.rplp")
# the packaged model is the trained one with its output layer shrunk toward the identity by this
# factor, chosen on held-out meshes (icosphere level 8 with other noise than the benchmark's): PCG
# iterations to 1e-6 for factors 0 / 0.25 / 0.5 / 1 were 14/14/14/14, 14/14/15/15 and 13/13/13/13
PACKAGED_SHRINK = 0.25


def shrink_output_layer(params: np.ndarray, F: int, alpha: float) -> np.ndarray:
    """The model whose node outputs o = w2 . h + b2 move alpha of the way from the identity's
    (w2 = 0, b2 = softplus^-1(1): every s_c = 1) to the trained ones; the hidden layer and log w stay."""
    p = np.array(params, dtype=np.float32, copy=True)
    p0 = init_params(F)
    o = HIDDEN * F + HIDDEN
    p[o:o + HIDDEN] = alpha * p[o:o + HIDDEN]
    p[o + HIDDEN] = p0[o + HIDDEN] + alpha * (p[o + HIDDEN] - p0[o + HIDDEN])
    return p


def feature_count() -> int:
    from . import H
    return int(H.rpm_learned_feature_count())


def param_count(F: int) -> int:
    return HIDDEN * F + 2 * HIDDEN + 2


def init_params(F: int, seed: int = 0, log_w: float = 0.0) -> np.ndarray:
    """Scalings start at exactly 1 (w2 = 0, b2 = softplus^-1(1)) and log w at 0, so the initial model
    is the two-level preconditioner P0 + w diag(A)^-1 with the estimated weight."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((HIDDEN, F)) * (1.0 / np.sqrt(F))
    b1 = np.zeros(HIDDEN)
    w2 = np.zeros(HIDDEN)
    b2 = np.array([np.log(np.expm1(1.0))])
    return np.concatenate([W1.ravel(), b1, w2, b2, [log_w]]).astype(np.float32)


def save_model(path: str, params: np.ndarray, F: int) -> None:
    params = np.ascontiguousarray(params, dtype=np.float32)
    assert params.size == param_count(F)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(b"RPLP")
        f.write(struct.pack("<ii", F, params.size))
        f.write(params.tobytes())


def load_model(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        magic = f.read(4)
        if magic != b"RPLP":
            raise ValueError(f"{path}: not a learned-preconditioner model")
        F, n = struct.unpack("<ii", f.read(8))
        if n != param_count(F):
            raise ValueError(f"{path}: parameter count {n} does not match {F} features")
        p = np.frombuffer(f.read(4 * n), dtype=np.float32).copy()
    if p.size != n:
        raise ValueError(f"{path}: truncated")
    return p


def node_scalings(params: np.ndarray, feats: np.ndarray) -> np.ndarray:
    """s_c of every node (numpy FP64 evaluation of the same MLP; used by the tests)."""
    F = feats.shape[1]
    p = params.astype(np.float64)
    W1 = p[: HIDDEN * F].reshape(HIDDEN, F)
    b1 = p[HIDDEN * F: HIDDEN * F + HIDDEN]
    w2 = p[HIDDEN * F + HIDDEN: HIDDEN * F + 2 * HIDDEN]
    b2 = p[HIDDEN * F + 2 * HIDDEN]
    o = np.tanh(feats.astype(np.float64) @ W1.T + b1) @ w2 + b2
    return np.log1p(np.exp(o))

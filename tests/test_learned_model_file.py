"""The learned-preconditioner model files and their shrinkage toward the identity (CPU only).

* The packaged model equals the trained model with its output layer shrunk by PACKAGED_SHRINK.
* shrink_output_layer with 0 gives every node the scaling 1 (the two-level preconditioner with the
  learned weight), with 1 the trained parameters, and in between node outputs o = w2 . h + b2 that move
  linearly from the identity's to the trained ones (checked through node_scalings).
* Model files round-trip and a wrong magic or truncated file is refused.
"""

import os

import numpy as np
import pytest

from repulsor_amd import learned_pc as LP

F = 9   # features per node (LEARNED_PC_FEATURES)


def test_packaged_model_is_the_shrunk_trained_model():
    p = LP.load_model(LP.TRAINED_MODEL)
    q = LP.load_model(LP.DEFAULT_MODEL)
    assert np.array_equal(q, LP.shrink_output_layer(p, F, LP.PACKAGED_SHRINK))
    assert q[-1] == p[-1]          # the learned weight factor is kept


def test_shrink_moves_node_outputs_linearly():
    p = LP.load_model(LP.TRAINED_MODEL)
    feats = np.random.default_rng(3).standard_normal((200, F))
    inv_softplus = lambda s: np.log(np.expm1(s))
    o1 = inv_softplus(LP.node_scalings(p, feats))
    o0 = inv_softplus(LP.node_scalings(LP.shrink_output_layer(p, F, 0.0), feats))
    assert np.allclose(LP.node_scalings(LP.shrink_output_layer(p, F, 0.0), feats), 1.0, rtol=1e-6)
    assert np.array_equal(LP.shrink_output_layer(p, F, 1.0), p)
    for a in (0.25, 0.5):
        oa = inv_softplus(LP.node_scalings(LP.shrink_output_layer(p, F, a), feats))
        assert np.allclose(oa, o0 + a * (o1 - o0), rtol=0, atol=1e-5)


def test_model_file_round_trip_and_errors(tmp_path):
    p = LP.init_params(F, seed=4, log_w=0.3)
    path = os.path.join(tmp_path, "m.rplp")
    LP.save_model(path, p, F)
    assert np.array_equal(LP.load_model(path), p)
    raw = open(path, "rb").read()
    open(path, "wb").write(b"XXXX" + raw[4:])
    with pytest.raises(ValueError):
        LP.load_model(path)
    open(path, "wb").write(raw[:-8])
    with pytest.raises(ValueError):
        LP.load_model(path)

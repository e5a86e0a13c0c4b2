"""Failure paths on the GPU: non-finite coordinates are refused with RP_EINVAL and a message, before any
kernel runs, whether they arrive with a SemiStaticUpdate (rp_update_coordinates) or for a device tree
build (rp_build_topology); the mesh keeps working with finite coordinates afterwards."""

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS

pytestmark = pytest.mark.gpu


def _mesh(device_build=False, X=None):
    wl = WORKLOADS["torus_2400"]
    X0, S = wl.geometry()
    M = R.SimplicialMesh(X0 if X is None else X, S)
    M.cluster_tree_settings.split_threshold = wl.leaf
    M.cluster_tree_settings.device_build = device_build
    return wl, M, X0


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_semistatic_update_rejects_nonfinite(bad):
    wl, M, X0 = _mesh()
    M.build()
    E = R.TangentPointEnergy0(wl.q, wl.p)
    e0 = E.Value(M)
    X = X0.copy()
    X[17, 1] = bad
    with pytest.raises(R.RepulsorError) as ei:
        M.SemiStaticUpdate(X)
    print(f"refused: {ei.value}", flush=True)
    assert ei.value.status == R.RP_EINVAL and ("NaN" in str(ei.value) or "Inf" in str(ei.value))
    M.SemiStaticUpdate(X0)
    assert E.Value(M) == e0


def test_device_build_rejects_nan():
    X0 = WORKLOADS["torus_2400"].geometry()[0].copy()
    X0[3, 0] = np.nan
    wl, M, _ = _mesh(device_build=True, X=X0)
    with pytest.raises(R.RepulsorError) as ei:
        M.build()
    print(f"refused: {ei.value}", flush=True)
    assert ei.value.status == R.RP_EINVAL and "NaN" in str(ei.value)

"""Device-resident SemiStaticUpdate (rp_update_coordinates) against the host path.

* the near rows and cluster rows computed on the GPU from new coordinates equal the host's rows for
  the same coordinates bit for bit (surface and curve meshes, with and without vertex charges);
* energy, differential and metric product after a device-pointer update equal, bit for bit, those of a
  mesh whose rows were computed on the host and uploaded;
* the host reads the coordinates back only when it needs them (VertexCoordinates, H^1), and a solve
  after a device update reaches the oracle-measured residual.
"""

import numpy as np
import pytest

import repulsor_amd as R
from repulsor_amd.workloads import WORKLOADS, random_field
from helpers import OracleSolve, oracle_energy_diff

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _meshes(name, charges=False):
    wl = WORKLOADS[name]
    X, S = wl.geometry()
    q = None
    if charges:
        q = 1.0 + 0.5 * np.random.default_rng(4).random(len(X))
    M1 = R.SimplicialMesh(X, S, charges=q)
    M3 = R.SimplicialMesh(X, S, charges=q)
    for M in (M1, M3):
        M.cluster_tree_settings.split_threshold = wl.leaf
        M.block_cluster_tree_settings.far_field_separation_parameter = wl.far_theta
        M.adaptivity_settings.theta = wl.near_theta
        M.build()
    rng = np.random.default_rng(9)
    X2 = X + 0.01 * rng.standard_normal(X.shape) * np.linalg.norm(X[1] - X[0])
    return wl, M1, M3, X2


@pytest.mark.parametrize("name,charges", [("sphere_1280", False), ("torus_2400", True), ("trefoil_3000", False)])
def test_device_update_rows_are_bitwise_host_rows(name, charges):
    wl, M1, M3, X2 = _meshes(name, charges)
    R.TangentPointEnergy0(wl.q, wl.p).Value(M1)          # device set up with the old state
    dX = torch.tensor(X2, dtype=torch.float64, device="cuda").contiguous()
    torch.cuda.synchronize()                              # the copy runs on torch's stream
    M1.SemiStaticUpdate(device_ptr=dX.data_ptr())
    N_dev, C_dev = M1.device_state_rows()
    assert np.array_equal(M1.VertexCoordinates(), X2)     # read back from the device
    td = M1.tree_data()                                   # host rows recomputed from those coordinates
    assert np.array_equal(N_dev, td["P_near"])
    assert np.array_equal(C_dev, td["C_far"])
    # with the device set up, the host-pointer update computes the same rows on the device
    R.TangentPointEnergy0(wl.q, wl.p).Value(M3)
    M3.SemiStaticUpdate(X2)
    N3, C3 = M3.device_state_rows()
    assert np.array_equal(N3, td["P_near"]) and np.array_equal(C3, td["C_far"])


@pytest.mark.parametrize("name", ["sphere_1280", "trefoil_3000"])
def test_device_update_results_equal_uploaded_rows(name):
    wl, M1, M3, X2 = _meshes(name)
    E = R.TangentPointEnergy0(wl.q, wl.p)
    A = R.TangentPointMetric0(wl.q, wl.p)
    E.Value(M1)
    dX = torch.tensor(X2, dtype=torch.float64, device="cuda").contiguous()
    torch.cuda.synchronize()                              # the copy runs on torch's stream
    M1.SemiStaticUpdate(device_ptr=dX.data_ptr())
    M3.SemiStaticUpdate(X2)                               # no device yet: host rows, uploaded later
    Y1, e1 = E.Differential(M1, return_energy=True)
    Y3, e3 = E.Differential(M3, return_energy=True)
    assert e1 == e3
    assert np.array_equal(Y1, Y3)
    X = random_field(M1.V, seed=3)
    P1 = np.zeros_like(X); A.MultiplyMetric(M1, 1.0, X, 0.0, P1)
    P3 = np.zeros_like(X); A.MultiplyMetric(M3, 1.0, X, 0.0, P3)
    assert np.array_equal(P1, P3)


def test_solve_after_device_update():
    wl, M1, _, X2 = _meshes("sphere_1280")
    E = R.TangentPointEnergy0(wl.q, wl.p)
    E.Value(M1)
    dX = torch.tensor(X2, dtype=torch.float64, device="cuda").contiguous()
    torch.cuda.synchronize()                              # the copy runs on torch's stream
    M1.SemiStaticUpdate(device_ptr=dX.data_ptr())
    B = E.Differential(M1)
    A = R.TangentPointMetric0(wl.q, wl.p)
    A.SetPreconditionerDiagonalWeight(-1.0)
    X = np.zeros_like(B)
    A.Solve(M1, 1.0, B, 0.0, X, max_iter=200, tolerance=1e-6)
    S = OracleSolve(M1, wl.q, wl.p)
    _, Bo, _, _ = oracle_energy_diff(M1, wl.q, wl.p)
    true = np.linalg.norm(Bo - S.A(X), axis=0) / np.linalg.norm(Bo, axis=0)
    assert np.all(true <= 1.5e-6), true

"""The reference's user call sequence (Example2/main.cpp:37-147) as a C++ program compiled against
include/repulsor (tests/cpp/reference_call_sequence.cpp): factories, Make_FromFile on the reference's own
mesh file Meshes/TorusMesh_00002400T.txt (a copy in tests/golden/), Differential(M, ptr),
Solve(..., 100, 1e-5) with the default solver, MaximumSafeStepSize, SemiStaticUpdate, Value.

Checked: energy and differential against the oracle on the same mesh and hierarchy (relative 1e-10);
the solve equals the Python mirror's GMRES with P0 on the same mesh (same iterations, gradient within
1e-12); the step size lies in (0, 1]; the energy after the step equals the oracle's at the moved
coordinates (1e-10) and is lower than before, as Example2 checks."""

import os
import subprocess

import numpy as np
import pytest

import repulsor_amd as R
from helpers import oracle_energy_diff, rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "_build", "reference_call_sequence")
MESH = os.path.join(ROOT, "tests", "golden", "TorusMesh_00002400T.txt")


def read_mesh_file(path):
    """The reference's text format: dom_dim n amb_dim m vertex_count V simplex_count T, coordinates, simplices."""
    tok = open(path).read().split()
    n, m, V, T = int(tok[1]), int(tok[3]), int(tok[5]), int(tok[7])
    vals = tok[8:]
    X = np.array(vals[:V * m], dtype=np.float64).reshape(V, m)
    S = np.array(vals[V * m:V * m + T * (n + 1)], dtype=np.int64).reshape(T, n + 1).astype(np.int32)
    return X, S


def _run(tmp_path):
    if not os.path.exists(PROG):
        pytest.skip("tests/cpp/_build/reference_call_sequence is not built (run __graft_entry__.build())")
    out = tmp_path / "out.bin"
    p = subprocess.run([PROG, MESH, str(out)], capture_output=True, text=True, timeout=300)
    print(p.stdout + p.stderr, flush=True)
    return p, out


def _parse(out):
    raw = out.read_bytes()
    V, it = np.frombuffer(raw[:8], dtype=np.int32)
    d = np.frombuffer(raw[8:], dtype=np.float64)
    relres, en, t, en_new = d[:3], d[3], d[4], d[5]
    diff = d[6:6 + 3 * V].reshape(V, 3)
    grad = d[6 + 3 * V:6 + 6 * V].reshape(V, 3)
    return int(it), relres, en, t, en_new, diff, grad


def _mesh(X, S):
    M = R.SimplicialMesh(X, S)
    M.cluster_tree_settings.split_threshold = 2
    M.block_cluster_tree_settings.far_field_separation_parameter = 0.25
    M.adaptivity_settings.theta = 10.0
    M.build()
    return M


@pytest.mark.gpu
def test_reference_call_sequence(tmp_path):
    p, out = _run(tmp_path)
    assert p.returncode == 0
    it, relres, en, t, en_new, diff, grad = _parse(out)
    X, S = read_mesh_file(MESH)
    M = _mesh(X, S)
    en_o, G_o, _, _ = oracle_energy_diff(M, 6.0, 12.0)
    print(f"energy {en!r} oracle {en_o!r}; differential rel err {rel(diff, G_o):.2e}; GMRES {it} iterations, "
          f"residuals {relres}; t = {t}; energy after the step {en_new!r}", flush=True)
    assert abs(en - en_o) <= 1e-10 * abs(en_o)
    assert rel(diff, G_o) <= 1e-10
    # the same solve through the Python mirror (reference GMRES, P0, 100 iterations, 1e-5)
    A = R.TangentPointMetric0(6.0, 12.0)
    Xs = np.zeros_like(diff)
    A.Solve(M, 1.0, diff, 0.0, Xs, max_iter=100, tolerance=1e-5, method="gmres", raise_on_stagnation=False)
    assert A.Solver_IterationCount() == it and 1 <= it <= 100
    assert np.allclose(A.Solver_RelativeResiduals(), relres, rtol=1e-10, atol=0)
    assert rel(grad, Xs) <= 1e-12
    assert 0.0 < t <= 1.0
    Xn = X - t * grad
    M.SemiStaticUpdate(Xn)
    en_n, _, _, _ = oracle_energy_diff(M, 6.0, 12.0, X=Xn)
    assert abs(en_new - en_n) <= 1e-10 * abs(en_n)
    assert en_new < en


def test_reference_call_sequence_reports_errors_without_device(tmp_path):
    """Without a device every call prints an error and returns 0; nothing throws across the API."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("device present: covered by test_reference_call_sequence")
    p, out = _run(tmp_path)
    assert p.returncode == 0
    assert "ERROR" in p.stderr and "rp_ctx_create" in p.stderr
    it, relres, en, t, en_new, _, _ = _parse(out)
    assert en == 0.0 and t == 0.0 and it == 0

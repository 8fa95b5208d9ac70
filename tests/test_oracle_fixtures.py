"""The oracle reproduces the committed golden fixtures of the transport configurations
(tests/golden/make_fixtures.py), and their sizes match the closed forms: the fixtures the GPU
tests compare against stay tied to the pinned oracle."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as MF  # noqa: E402
from dsv import workloads  # noqa: E402


def _load(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name,builder,exact,nnz", [
    ("cube20_single", lambda: workloads.cube_single(20, seed=0), True, 128581),
    ("cube48_jitter", lambda: workloads.cube_jitter(48, seed=0), False, 1707697),
])
def test_oracle_reproduces_fixture(name, builder, exact, nnz):
    fx = _load(name)
    assert int(fx["nnz"]) == nnz  # closed-form Kuhn-grid count V + 2E
    new = MF.config_fixture(builder(), exact)
    assert str(new["pattern_sha256"]) == str(fx["pattern_sha256"])
    for k in ("values_fro", "values_sum", "rhs_norm", "x_norm"):
        assert abs(float(new[k]) - float(fx[k])) <= 1e-12 * abs(float(fx[k])), k
    tol = 1e-13 if exact else 1e-10
    assert np.abs(new["x"] - fx["x"]).max() <= tol * np.abs(fx["x"]).max()


def test_cube20_fixture_solution_is_a_backward_euler_step():
    """Residual of the fixture's solution in the independently re-assembled system, and the
    Dirichlet face holds c = 1."""
    from oracle import assembly as OA
    w = workloads.cube_single(20, seed=0)
    s = w["species"][0]
    fx = _load("cube20_single")
    ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], 1.0, 100.0, s["alpha"], 1, s["bc_dofs"])
    A = OA.to_scipy(ip, ix, data, len(w["xyz"]))
    b = OA.assemble_rhs(w["conn"], w["xyz"], s["u0"], 100.0, None, 1, s["bc_dofs"], 1.0)
    x = fx["x"]
    assert np.linalg.norm(b - A @ x) <= 1e-12 * np.linalg.norm(b)
    assert np.all(x[s["bc_dofs"]] == 1.0)


def test_disc3_fixture_matches_an_independent_cg_solve():
    """The disc3 step-1 fixture (sparse LU) against the oracle's C Jacobi-PCG of the same
    re-assembled systems to rtol 1e-12, at the fixture's vertices: within 1e-9 of max |x|; the
    mesh is the one the generator builds (digest)."""
    from oracle import assembly as OA
    from oracle import solve as OS
    fx = _load("disc3_step1")
    w = workloads.disc3()
    assert MF.mesh_digest(w["conn"], w["xyz"]) == str(fx["mesh_sha256"])
    stride = int(fx["stride"])
    inv_dt = 1.0 / w["dt"]
    for s in w["species"]:
        ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], s["D_value"], inv_dt, s["alpha"],
                                          1, s["bc_dofs"])
        b = OA.assemble_rhs(w["conn"], w["xyz"], s["u0"], inv_dt, s["f"], 1, s["bc_dofs"],
                            s["bc_value"])
        x0 = np.array(s["u0"], float)
        if s["bc_dofs"] is not None:
            x0[s["bc_dofs"]] = s["bc_value"]
        x, it, rel = OS.pcg_jacobi_c(ip, ix, data, b, x0, 1e-12, 20000, 8)
        assert rel <= 1e-12
        ref = fx["x_" + s["name"]]
        assert np.abs(x[::stride] - ref).max() <= 1e-9 * float(fx["max_" + s["name"]]), s["name"]
        assert abs(np.abs(x).max() - float(fx["max_" + s["name"]])) <= 1e-9

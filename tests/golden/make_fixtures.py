"""Writes the golden fixtures of the transport configurations (run from the repository root:
`python tests/golden/make_fixtures.py`).

The reference ships no vectors for this path and cannot be imported here (it needs cupy,
petsc4py and dolfinx), so the fixtures hold the oracle's answers (oracle/, itself pinned by the
closed-form and brute-force tests of tests/test_oracle_*.py) on the seeded inputs of
dsv.workloads:

  cube20_single.npz  -- unit cube 20^3 Kuhn, 48,000 tets, one species, D = 1, k = 0.1 where
                        z < 0.5, c = 1 on x = 0, dt = 0.01: pattern digest, matrix and RHS
                        norms, and the exact (sparse LU) step solution;
  cube48_jitter.npz  -- 48^3 Kuhn, 663,552 tets, interior jitter 0.2 h, D lognormal per cell,
                        k = 1 where |x - c| < 0.3, dt = 0.01: pattern digest, norms, and the
                        step solution (C Jacobi-PCG to rtol 1e-13) at every 97th vertex plus
                        its norm;
  disc3_step1.npz    -- the 805,903-tet graded disc, three species (O2 with the ORR sink and the
                        bottom Dirichlet, Cl, OH with its source), dt = 0.005: a digest of the
                        mesh, and per species the first backward-Euler step from u0 solved by
                        sparse LU (scipy splu, one refinement step), at every 11th vertex plus its norm and max.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dissolve-2.0_amd"))

from dsv import workloads  # noqa: E402
from oracle import assembly as OA  # noqa: E402
from oracle import solve as OS  # noqa: E402

STRIDE = 97


def pattern_digest(indptr, indices):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(indptr, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(indices, dtype=np.int32).tobytes())
    return h.hexdigest()


def config_fixture(w, exact):
    s = w["species"][0]
    inv_dt = 1.0 / w["dt"]
    De = s["D"] if s.get("D") is not None else s["D_value"]
    ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], De, inv_dt, s.get("alpha"), 1,
                                      s.get("bc_dofs"))
    b = OA.assemble_rhs(w["conn"], w["xyz"], s["u0"], inv_dt, s.get("f"), 1, s.get("bc_dofs"),
                        s.get("bc_value", 0.0))
    x0 = np.array(s["u0"], float)
    if s.get("bc_dofs") is not None and len(s["bc_dofs"]):
        x0[s["bc_dofs"]] = s["bc_value"]
    if exact:
        x = OS.direct(OA.to_scipy(ip, ix, data, len(w["xyz"])), b)
    else:
        x, it, rel = OS.pcg_jacobi_c(ip, ix, data, b, x0, 1e-13, 20000)
        assert rel <= 1e-13, rel
    out = dict(nnz=np.int64(len(ix)), pattern_sha256=np.array(pattern_digest(ip, ix)),
               values_fro=np.float64(np.linalg.norm(data)), values_sum=np.float64(data.sum()),
               rhs_norm=np.float64(np.linalg.norm(b)), x_norm=np.float64(np.linalg.norm(x)),
               stride=np.int64(1 if exact else STRIDE))
    out["x"] = x if exact else x[::STRIDE]
    return out


DISC_STRIDE = 11


def mesh_digest(conn, xyz):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(conn, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(xyz, dtype=np.float64).tobytes())
    return h.hexdigest()


def disc3_fixture():
    w = workloads.disc3()
    inv_dt = 1.0 / w["dt"]
    out = dict(mesh_sha256=np.array(mesh_digest(w["conn"], w["xyz"])),
               stride=np.int64(DISC_STRIDE), names=np.array([s["name"] for s in w["species"]]))
    for s in w["species"]:
        ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], s["D_value"], inv_dt, s["alpha"],
                                          1, s["bc_dofs"])
        A = OA.to_scipy(ip, ix, data, len(w["xyz"]))
        b = OA.assemble_rhs(w["conn"], w["xyz"], s["u0"], inv_dt, s["f"], 1, s["bc_dofs"],
                            s["bc_value"])
        import scipy.sparse.linalg as spla
        lu = spla.splu(A.tocsc())
        x = lu.solve(b)
        x += lu.solve(b - A @ x)  # one step of iterative refinement on the LU factors
        rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert rel <= 1e-13, rel
        out["x_" + s["name"]] = x[::DISC_STRIDE]
        out["norm_" + s["name"]] = np.float64(np.linalg.norm(x))
        out["max_" + s["name"]] = np.float64(np.abs(x).max())
        print("disc3", s["name"], rel, float(np.abs(x).max()))
    return out


def main(which=("cube20_single", "cube48_jitter", "disc3_step1")):
    for name, make, exact in (("cube20_single", lambda: workloads.cube_single(20, seed=0), True),
                              ("cube48_jitter", lambda: workloads.cube_jitter(48, seed=0), False)):
        if name not in which:
            continue
        fx = config_fixture(make(), exact)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **fx)
        print(name, int(fx["nnz"]), str(fx["pattern_sha256"])[:16], float(fx["x_norm"]))
    if "disc3_step1" in which:
        np.savez_compressed(os.path.join(HERE, "disc3_step1.npz"), **disc3_fixture())


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("cube20_single", "cube48_jitter", "disc3_step1"))

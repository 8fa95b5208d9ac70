"""GPU parity at the full sizes of the transport configurations, through the C-ABI.

* 20^3 cube, one species (48,000 tets) and 48^3 jittered cube, one species, per-cell D
  (663,552 tets): pattern digest bit-exact, matrix and RHS norms within 1e-12, and the step
  solution of every method at rtol 1e-10 within 1e-8 of the golden fixture (oracle direct /
  rtol-1e-13 solve, tests/golden/).
* 806k-tet graded disc, three species: three batched steps of the chosen path (FP32 CG in FP64
  refinement, Galerkin warm start); the last step's solutions satisfy the independently
  assembled oracle system to rtol 1e-8 and agree with the plain FP64 CSR path.
* 14.7M-tet gyroid, three species: the stored-geometry FP32 apply agrees with the FP64 and the
  recomputing FP32 apply, and one batched step converges.
"""
import os

import numpy as np
import pytest

from dsv import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = {
    "cube20_single": lambda: workloads.cube_single(20, seed=0),
    "cube48_jitter": lambda: workloads.cube_jitter(48, seed=0),
}
_CACHE = {}


def _config(name):
    from dsv.core import Mesh
    if name not in _CACHE:
        w = CONFIGS[name]()
        _CACHE[name] = (w, Mesh(w["conn"], w["xyz"]),
                        np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False))
    return _CACHE[name]


def _digest(mesh):
    import hashlib
    from dsv.core import VIEW_INDICES, VIEW_INDPTR
    h = hashlib.sha256()
    h.update(mesh.view(VIEW_INDPTR).cpu().numpy().astype(np.int32).tobytes())
    h.update(mesh.view(VIEW_INDICES).cpu().numpy().astype(np.int32).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", list(CONFIGS))
def test_pattern_matrix_and_rhs(gpu, name):
    from dsv.gpu_assembler import GPUSparseAssembler
    w, mesh, fx = _config(name)
    s = w["species"][0]
    print(f"[config] {name}: {len(w['conn'])} tets", flush=True)
    assert mesh.nnz == int(fx["nnz"])
    assert _digest(mesh) == str(fx["pattern_sha256"])
    asm = GPUSparseAssembler(mesh)
    De = s["D"] if s.get("D") is not None else np.full(len(w["xyz"]), s["D_value"])
    inv_dt = 1.0 / w["dt"]
    asm.assemble(De, inv_dt, alpha=s.get("alpha"), bc_dofs=s.get("bc_dofs"))
    data = asm.d_data.cpu().numpy()
    assert abs(np.linalg.norm(data) - float(fx["values_fro"])) <= 1e-12 * float(fx["values_fro"])
    assert abs(data.sum() - float(fx["values_sum"])) <= 1e-12 * abs(float(fx["values_sum"]))
    b = asm.assemble_rhs(s["u0"], inv_dt, f=s.get("f"), bc_dofs=s.get("bc_dofs"),
                         bc_value=s.get("bc_value", 0.0))
    rn = float(np.linalg.norm(b.cpu().numpy()))
    assert abs(rn - float(fx["rhs_norm"])) <= 1e-12 * float(fx["rhs_norm"])


@pytest.mark.parametrize("method", ["csr64", "mf64", "mf32_ir64"])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_step_solution(gpu, name, method):
    import torch
    from dsv.gpu_assembler import GPUSparseAssembler
    from dsv.solvers import solve_batched
    w, mesh, fx = _config(name)
    s = w["species"][0]
    inv_dt = 1.0 / w["dt"]
    print(f"[config] {name} {method}", flush=True)
    asm = GPUSparseAssembler(mesh)
    b = asm.assemble_rhs(s["u0"], inv_dt, f=s.get("f"), bc_dofs=s.get("bc_dofs"),
                         bc_value=s.get("bc_value", 0.0)).clone()
    sp = dict(D=s.get("D"), D_value=s["D_value"], alpha=s.get("alpha"), bc_dofs=s.get("bc_dofs"),
              bc_value=s.get("bc_value", 0.0), b=b,
              x=torch.tensor(s["u0"], dtype=torch.float64, device=gpu))
    st = solve_batched(mesh, [sp], inv_dt, rtol=1e-10, method=method)
    assert st["converged"][0] and st["rel_residual"][0] <= 1e-10, st
    x = sp["x"].cpu().numpy()[::int(fx["stride"])]
    ref = fx["x"]
    assert np.abs(x - ref).max() <= 1e-8 * np.abs(ref).max(), (method, st)


def test_disc3_batched_steps_satisfy_the_oracle_system(gpu):
    from dsv.core import Mesh
    from dsv.transport import Stepper
    from oracle import assembly as OA
    w = workloads.disc3()
    print(f"[config] disc3: {len(w['conn'])} tets", flush=True)
    mesh = Mesh(w["conn"], w["xyz"])
    chosen = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=w["rtol"])
    chosen.set_warm_start("extrapolate")
    plain = Stepper(mesh, w["species"], w["dt"], method="csr64", rtol=w["rtol"])
    for _ in range(3):
        hc, hp = chosen.step(), plain.step()
        assert all(hc["converged"]) and all(hp["converged"]), (hc, hp)
    inv_dt = 1.0 / w["dt"]
    for k, s in enumerate(w["species"]):
        u_old = chosen.u_prev[k].cpu().numpy()
        x = chosen.u[k].cpu().numpy()
        ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], s["D_value"], inv_dt, s["alpha"],
                                          1, s["bc_dofs"])
        A = OA.to_scipy(ip, ix, data, len(w["xyz"]))
        b = OA.assemble_rhs(w["conn"], w["xyz"], u_old, inv_dt, s["f"], 1, s["bc_dofs"],
                            s["bc_value"])
        rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert rel <= 1.01e-8, (s["name"], rel)
        xp = plain.u[k].cpu().numpy()
        assert np.abs(x - xp).max() <= 1e-5 * np.abs(xp).max(), s["name"]


def test_gyroid_apply_forms_and_one_step(gpu):
    import torch
    from dsv.ceed_operators import apply_batched
    from dsv.core import Mesh
    from dsv.transport import Stepper
    w = workloads.gyroid3()
    print(f"[config] gyroid3: {len(w['conn'])} tets", flush=True)
    mesh = Mesh(w["conn"], w["xyz"])
    sp = w["species"]
    inv_dt = 1.0 / w["dt"]
    rng = np.random.default_rng(2)
    xs = [torch.tensor(rng.uniform(0, 1, len(w["xyz"])), device=gpu) for _ in sp]
    y64 = apply_batched(mesh, sp, inv_dt, xs, precision=64)
    y32s = apply_batched(mesh, sp, inv_dt, xs, precision=32)
    mesh.set_operator_form(Mesh.FORM_RECOMPUTE)
    y32r = apply_batched(mesh, sp, inv_dt, xs, precision=32)
    mesh.set_operator_form(Mesh.FORM_STORED)
    for a, b, c in zip(y64, y32s, y32r):
        scale = float(a.abs().max())
        assert float((b - a).abs().max()) <= 2e-5 * scale
        assert float((c - a).abs().max()) <= 2e-5 * scale
    st = Stepper(mesh, sp, w["dt"], method="mf32_ir64", rtol=w["rtol"])
    h = st.step()
    assert all(h["converged"]) and max(h["rel_residual"]) <= w["rtol"], h

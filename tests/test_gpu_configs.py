"""GPU parity at the full sizes of the transport configurations, through the C-ABI.

* 20^3 cube, one species (48,000 tets) and 48^3 jittered cube, one species, per-cell D
  (663,552 tets): pattern digest bit-exact, matrix and RHS norms within 1e-12, and the step
  solution of every method at rtol 1e-10 within 1e-8 of the golden fixture (oracle direct /
  rtol-1e-13 solve, tests/golden/).
* 806k-tet graded disc, three species: step 1 of every method at rtol 1e-10 within 1e-8 of the
  golden sparse-LU solution (tests/golden/disc3_step1.npz); five steps of the FP32 path
  (Galerkin warm start and previous-solution start) within 1e-8 at every step of the plain
  FP64 path's rtol-1e-13 solution of the same system, the last step satisfying the
  independently assembled oracle system.
* 14.7M-tet gyroid, three species: the stored-geometry FP32 apply agrees with the FP64 and the
  recomputing FP32 apply, and three batched steps of the chosen path (two warm-started) lie
  within 1e-8 of the plain FP64 path's rtol-1e-13 solutions.
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


_DISC = {}


def _disc3():
    from dsv.core import Mesh
    if not _DISC:
        import hashlib
        w = workloads.disc3()
        fx = np.load(os.path.join(HERE, "golden", "disc3_step1.npz"), allow_pickle=False)
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(w["conn"], dtype=np.int32).tobytes())
        h.update(np.ascontiguousarray(w["xyz"], dtype=np.float64).tobytes())
        assert h.hexdigest() == str(fx["mesh_sha256"]), "disc3 mesh differs from the fixture's"
        _DISC.update(w=w, fx=fx, mesh=Mesh(w["conn"], w["xyz"]))
    return _DISC["w"], _DISC["mesh"], _DISC["fx"]


def _fixture_error(fx, k, name, x):
    """max over the fixture's vertices of |x - x*| / max |x*| (x* the direct solve)."""
    ref = fx["x_" + name]
    return float(np.abs(x[::int(fx["stride"])] - ref).max() / float(fx["max_" + name]))


@pytest.mark.parametrize("method", ["csr64", "mf64", "mf32_ir64"])
def test_disc3_step1_matches_direct_solve(gpu, method):
    """The first three-species step of the 806k-tet disc at rtol 1e-10 against the golden
    fixture (sparse-LU solve of the oracle system): every species within 1e-8 of max |x*|.  The
    same step at rtol 1e-8 misses that bar (O2 about 9e-7, Cl 2e-8 on the CPU), which is why the
    bar is tested at 1e-10."""
    from dsv.transport import Stepper
    w, mesh, fx = _disc3()
    print(f"[config] disc3 step 1 {method}: {len(w['conn'])} tets", flush=True)
    st = Stepper(mesh, w["species"], w["dt"], method=method, rtol=1e-10)
    h = st.step()
    assert all(h["converged"]), h
    for k, s in enumerate(w["species"]):
        err = _fixture_error(fx, k, s["name"], st.u[k].cpu().numpy())
        print(f"  {s['name']}: error vs direct solve {err:.2e}", flush=True)
        assert err <= 1e-8, (s["name"], err, h)


@pytest.mark.parametrize("warm", ["extrapolate", "previous"])
def test_disc3_steps_within_1e8_of_fp64(gpu, warm):
    """Five steps of the FP32 path (Galerkin warm start from the second step, or the previous
    solution) at rtol 1e-10: at every step each species' solution is within 1e-8 of the same
    system solved by the plain FP64 path to rtol 1e-13, which is itself within 1e-9 of the
    sparse-LU fixture at step 1; the last step satisfies the independently assembled oracle
    system to rtol 1e-10."""
    from dsv.transport import Stepper
    from oracle import assembly as OA
    w, mesh, fx = _disc3()
    print(f"[config] disc3 5 steps {warm}", flush=True)
    st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=1e-10)
    if warm != "previous":
        st.set_warm_start(warm)
    for n in range(5):
        h = st.step()
        assert all(h["converged"]), (n, h)
        refs, hr = st.fp64_reference(1e-13)
        assert all(hr["converged"]), hr
        if n == 0:
            for k, s in enumerate(w["species"]):
                assert _fixture_error(fx, k, s["name"], refs[k].cpu().numpy()) <= 1e-9
        errs = st.step_error(refs)
        print(f"  step {n + 1}: iterations {h['iterations']} error vs FP64 {errs}", flush=True)
        assert max(errs) <= 1e-8, (n, errs, h)
    inv_dt = 1.0 / w["dt"]
    for k, s in enumerate(w["species"]):
        u_old = st.u_prev[k].cpu().numpy()
        x = st.u[k].cpu().numpy()
        ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], s["D_value"], inv_dt, s["alpha"],
                                          1, s["bc_dofs"])
        A = OA.to_scipy(ip, ix, data, len(w["xyz"]))
        b = OA.assemble_rhs(w["conn"], w["xyz"], u_old, inv_dt, s["f"], 1, s["bc_dofs"],
                            s["bc_value"])
        rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert rel <= 1.01e-10, (s["name"], rel)


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
    # three steps of the chosen path at rtol 1e-10, the last two from the Galerkin warm start:
    # each within 1e-8 of the plain FP64 path's rtol-1e-13 solution of the same systems
    st = Stepper(mesh, sp, w["dt"], method="mf32_ir64", rtol=1e-10)
    st.set_warm_start("extrapolate")
    for n in range(3):
        h = st.step()
        assert all(h["converged"]) and max(h["rel_residual"]) <= 1e-10, (n, h)
        refs, hr = st.fp64_reference(1e-13)
        errs = st.step_error(refs)
        print(f"  gyroid step {n + 1}: iterations {h['iterations']} error vs FP64 {errs}",
              flush=True)
        assert all(hr["converged"]) and max(errs) <= 1e-8, (n, errs)
    assert len(st.warm_log) == 2 and any(any(e["taken"]) for e in st.warm_log)

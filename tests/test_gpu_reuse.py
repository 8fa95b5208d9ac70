"""Operator reuse (dsv_set_operator_reuse): a stepper that keeps its operator across steps is
bit-identical to one that rebuilds it, also after refreshed coefficients; the per-call APIs
never serve a stale operator."""
import numpy as np
import pytest

from dsv import meshes

pytestmark = pytest.mark.gpu


def _setup(gpu, seed=4):
    import torch
    from dsv.core import Mesh
    from oracle import assembly as OA
    conn, xyz = meshes.graded_disc(seed=1, scale=6.0)
    rng = np.random.default_rng(seed)
    V = len(xyz)
    bc = np.nonzero(xyz[:, 2] <= xyz[:, 2].min() + 1e-9)[0].astype(np.int32)
    t = lambda a, dt=torch.float64: torch.tensor(a, dtype=dt, device=gpu)
    alpha = t(rng.uniform(0, 30, V))
    u0 = [rng.uniform(0, 1, V) for _ in range(3)]
    bs = [t(OA.assemble_rhs(conn, xyz, u, 200.0, None, 1, bc if k == 0 else None, 1.0))
          for k, u in enumerate(u0)]
    sps = [dict(D_value=2.0, alpha=alpha, bc_dofs=t(bc, torch.int32), bc_value=1.0),
           dict(D_value=2.0), dict(D_value=5.3, alpha_value=0.5)]
    return Mesh(conn, xyz), sps, u0, bs, t


def _solve(mesh, sps, u0, bs, t, method):
    from dsv.solvers import solve_batched
    sp = [dict(s, b=b, x=t(u)) for s, u, b in zip(sps, u0, bs)]
    st = solve_batched(mesh, sp, 200.0, rtol=1e-10, method=method)
    assert all(st["converged"]), st
    return [s["x"].cpu().numpy() for s in sp], st


def test_stepper_reuse_is_bit_identical_and_follows_refresh(gpu):
    """A stepper that keeps its operator across steps (reuse on) against one that rebuilds it
    every call (reuse off, its own mesh): the same iterations and bit-identical solutions, also
    after the coefficients are refreshed in place, with the warm start's A d of the shifted
    history taken from the previous step under reuse."""
    from dsv import workloads
    from dsv.core import Mesh
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    steppers = []
    for reuse in (True, False):
        st = Stepper(Mesh(w["conn"], w["xyz"]), w["species"], w["dt"], method="mf32_ir64",
                     rtol=w["rtol"], reuse_operator=reuse)
        st.set_warm_start("extrapolate")
        steppers.append(st)
    a, b = steppers
    reused, gram = [], []
    for n in range(9):
        if n == 5:
            for st in steppers:
                st.refresh_coefficients(alpha=[st.alpha_dev[0] * 2.5, None, None])
        ha, hb = a.step(), b.step()
        assert ha["iterations"] == hb["iterations"], (n, ha, hb)
        for ua, ub in zip(a.u, b.u):
            assert np.array_equal(ua.cpu().numpy(), ub.cpu().numpy()), n
        assert hb["dirs_reused"] == 0 and hb["gram_reused"] == 0
        reused.append(ha["dirs_reused"])
        gram.append(ha["gram_reused"])
    # the shifted history's A d_k (k >= 1, the start read from the history) come from the
    # previous step while the operator stands: none on the first steps (too short a history)
    # and right after the refresh
    assert reused[:2] == [0, 0] and reused[2] == 1 and reused[5] == 0, reused
    assert reused[3] > 1 and reused[4] > 1 and reused[8] > 1, reused
    # with them, the Gram entries of the shifted directions (bit-identical solutions above)
    assert gram[3] == 1 and gram[8] == 1 and gram[5] == 0, gram


def test_per_call_apis_never_reuse_a_stale_operator(gpu):
    """With reuse on (e.g. left on by a Stepper), solve_batched builds its operator afresh, so an
    array changed in place at the same address -- what a freed and reallocated array looks
    like -- is never served from the cache."""
    mesh, sps, u0, bs, t = _setup(gpu, seed=6)
    mesh.set_operator_reuse(True)
    _solve(mesh, sps, u0, bs, t, "mf32_ir64")
    sps[0]["alpha"].mul_(3.0)  # same address, new values
    got, _ = _solve(mesh, sps, u0, bs, t, "mf32_ir64")
    mesh.set_operator_reuse(False)
    ref, _ = _solve(mesh, sps, u0, bs, t, "mf32_ir64")
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_stepper_start_from_history_is_bit_identical(gpu):
    """The stepper's Galerkin solves with the start read from the history
    (dsv_solve_warm_history) against the start copied into x (dsv_solve_warm): the same
    directions, so bit-identical solutions and iterations, with one more direction's A d
    reused per step (d_1, the previous step's d_0)."""
    from dsv import workloads
    from dsv.core import Mesh
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    steppers = []
    for from_hist in (True, False):
        st = Stepper(Mesh(w["conn"], w["xyz"]), w["species"], w["dt"], method="mf32_ir64",
                     rtol=w["rtol"])
        st.set_warm_start("extrapolate")
        st.start_from_history = from_hist
        steppers.append(st)
    a, b = steppers
    for n in range(9):
        if n == 5:
            for st in steppers:
                st.refresh_coefficients(alpha=[st.alpha_dev[0] * 2.5, None, None])
        ha, hb = a.step(), b.step()
        assert ha["iterations"] == hb["iterations"] and ha["outer"] == hb["outer"], (n, ha, hb)
        for ua, ub in zip(a.u, b.u):
            assert np.array_equal(ua.cpu().numpy(), ub.cpu().numpy()), n
        assert ha["gram_reused"] == hb["gram_reused"], n
        extra = 0 if n < 2 or n == 5 else 1
        assert ha["dirs_reused"] == hb["dirs_reused"] + extra, (n, ha["dirs_reused"],
                                                                 hb["dirs_reused"])


@pytest.mark.parametrize("between", [False, True])
def test_direction_from_a_written_solution_is_not_reused(gpu, between):
    """A direction whose arrays include a solution a solve wrote since is formed anew.
    A solve with the start in x (X) records d_0 = X - H0; a history-start solve whose history
    then names X as step n-1 has d_1 = X - H0 from the same two arrays, but X now holds that
    solve's solution.  Against the same calls on a mesh without reuse: bit-identical, and only
    d_2 = H0 - H1 (arrays nobody wrote) taken from the previous call -- none at all when a
    plain solve (into another array) ran in between, which drops the recorded directions."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    from dsv import workloads
    w = workloads.disc3(seed=1, scale=6.0)
    V = len(w["xyz"])
    rng = np.random.default_rng(21)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    u0 = [np.asarray(s["u0"], dtype=np.float64) for s in w["species"]]
    H0 = [t(u - rng.normal(0, 1e-3, V)) for u in u0]
    H1 = [t(u - rng.normal(0, 2e-3, V)) for u in u0]
    Y = [t(u + rng.normal(0, 1e-3, V)) for u in u0]
    runs = []
    for reuse in (True, False):
        mesh = Mesh(w["conn"], w["xyz"])
        mesh.set_operator_reuse(reuse)
        keep = []
        X = [t(u) for u in u0]
        bs = [t(np.ones(V)) for _ in u0]
        inv_dt = 1.0 / w["dt"]
        # device coefficient arrays made once: the operator key names their addresses
        dev = [dict(D=None if s.get("D") is None else t(s["D"]),
                    alpha=None if s.get("alpha") is None else t(s["alpha"]),
                    bc=None if s.get("bc_dofs") is None else
                    torch.tensor(s["bc_dofs"], dtype=torch.int32, device=gpu))
               for s in w["species"]]

        def structs(xs):
            return species_array([species_struct(
                mesh, D=d["D"], D_value=s.get("D_value", 0.0), alpha=d["alpha"],
                alpha_value=s.get("alpha_value", 0.0), bc_dofs=d["bc"],
                bc_value=s.get("bc_value", 0.0), b=b, x=x, keep=keep)
                for s, d, b, x in zip(w["species"], dev, bs, xs)])

        def arr(rows):
            return (ctypes.c_void_p * (3 * len(rows)))(*[ptr(v).value for r in rows for v in r])
        st1, st2 = _lib.Stats(), _lib.Stats()
        _lib.call("dsv_solve_warm", mesh.handle, 3, structs(X), inv_dt, 1,
                  _lib.METHODS["mf32_ir64"], 1e-9, 5000, 2, arr([H0, H1]), 0, None,
                  stream_ptr(gpu), st1, None, None, None)
        if between:
            W = [t(u) for u in u0]
            _lib.call("dsv_solve", mesh.handle, 3, structs(W), inv_dt, 1,
                      _lib.METHODS["mf32_ir64"], 1e-9, 5000, stream_ptr(gpu), _lib.Stats())
        Z = [torch.full((V,), float("nan"), dtype=torch.float64, device=gpu) for _ in u0]
        _lib.call("dsv_solve_warm_history", mesh.handle, 3, structs(Z), inv_dt, 1,
                  _lib.METHODS["mf32_ir64"], 1e-9, 5000, 3, arr([Y, X, H0, H1]), 0, None,
                  stream_ptr(gpu), st2, None, None, None)
        torch.cuda.synchronize()
        runs.append(([z.cpu().numpy() for z in Z], st2.as_dict(3)))
    (za, da), (zb, db) = runs
    assert all(da["converged"]) and all(db["converged"]), (da, db)
    assert da["iterations"] == db["iterations"], (da, db)
    for x, y in zip(za, zb):
        assert np.isfinite(x).all() and np.array_equal(x, y)
    assert da["dirs_reused"] == (0 if between else 1) and db["dirs_reused"] == 0, (da, db)


def test_stepper_reuse_is_bit_identical_on_a_large_mesh(gpu):
    """The same comparison on 614,125 vertices, where the warm start's Gram kernels run on more
    than 256 workgroups (their grid grows with the mesh): reuse on and off still give the same
    iterations and bit-identical solutions, with the shifted Gram rows taken under reuse."""
    from dsv import workloads
    from dsv.core import Mesh
    from dsv.transport import Stepper
    conn, xyz = meshes.kuhn_cube(84)
    conn, xyz = meshes.locality_order(conn, xyz)
    assert len(xyz) > 2048 * 256
    bc = np.nonzero(xyz[:, 2] <= 1e-12)[0].astype(np.int32)
    w = workloads.three_species(conn, xyz, xyz[:, 2] - 0.8, bc, 0.005, 0.03)
    steppers = []
    for reuse in (True, False):
        st = Stepper(Mesh(w["conn"], w["xyz"]), w["species"], w["dt"], method="mf32_ir64",
                     rtol=w["rtol"], reuse_operator=reuse)
        st.set_warm_start("extrapolate")
        steppers.append(st)
    a, b = steppers
    gram = []
    for n in range(6):
        ha, hb = a.step(), b.step()
        assert ha["iterations"] == hb["iterations"], (n, ha, hb)
        for ua, ub in zip(a.u, b.u):
            assert np.array_equal(ua.cpu().numpy(), ub.cpu().numpy()), n
        assert hb["gram_reused"] == 0
        gram.append(ha["gram_reused"])
    assert gram[4] == 1 and gram[5] == 1, gram

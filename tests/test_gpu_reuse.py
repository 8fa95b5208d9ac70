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
    # the shifted history's A d_k (k >= 2) come from the previous step while the operator
    # stands: none on the first steps (too short a history) and right after the refresh
    assert reused[:3] == [0, 0, 0] and reused[5] == 0, reused
    assert reused[3] > 0 and reused[4] > 0 and reused[8] > 0, reused
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

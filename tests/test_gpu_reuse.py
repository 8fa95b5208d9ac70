"""Operator reuse (dsv_set_operator_reuse) through the C-ABI: with reuse on, solves and
residuals are bit-identical to rebuilding the operator every call; a coefficient array at a new
address, and an array changed in place followed by a new reuse call, are picked up."""
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


@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
def test_reuse_is_bit_identical_and_tracks_changes(gpu, method):
    mesh, sps, u0, bs, t = _setup(gpu)
    ref, st_ref = _solve(mesh, sps, u0, bs, t, method)
    mesh.set_operator_reuse(True)
    for _ in range(2):  # the first call builds, the second reuses
        got, st = _solve(mesh, sps, u0, bs, t, method)
        assert st["iterations"] == st_ref["iterations"]
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    # a new alpha array (new address): picked up without a reuse call
    sps2 = [dict(sps[0], alpha=sps[0]["alpha"] * 3.0), sps[1], sps[2]]
    mesh.set_operator_reuse(False)
    ref2, _ = _solve(mesh, sps2, u0, bs, t, method)
    mesh.set_operator_reuse(True)
    _solve(mesh, sps, u0, bs, t, method)           # cache the first operator
    got2, _ = _solve(mesh, sps2, u0, bs, t, method)
    for a, b in zip(got2, ref2):
        assert np.array_equal(a, b)
    # the same array changed in place, then a new reuse call
    sps[0]["alpha"].mul_(3.0)
    mesh.set_operator_reuse(True)
    got3, _ = _solve(mesh, sps, u0, bs, t, method)
    for a, b in zip(got3, ref2):
        assert np.array_equal(a, b)
    mesh.set_operator_reuse(False)

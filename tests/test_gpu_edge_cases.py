"""Edge cases of the hot path on the GPU, through the C-ABI, against the oracle: the smallest
meshes (one tetrahedron, two sharing a face, a ragged strip whose vertices have very different
degrees), a species whose every vertex is a Dirichlet row, and a species with no Dirichlet rows
next to one with many in the same batched solve.

Bars as in the parity tests: pattern and scatter map bit-exact, FP64 entries within 1e-12
relative, every solve within 1e-8 of scipy splu, gradients of a linear field exact to 1e-11."""
import numpy as np
import pytest

from oracle import assembly as OA
from oracle import coefficients as OC
from oracle import solve as OS

pytestmark = pytest.mark.gpu


def _one_tet():
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    return np.array([[0, 1, 2, 3]], dtype=np.int32), xyz


def _two_tets():
    xyz = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
                    [0.4, 0.5, -0.9]])
    return np.array([[0, 1, 2, 3], [1, 0, 2, 4]], dtype=np.int32), xyz


def _fan():
    """Tetrahedra fanned around one axis: the two axis vertices touch every cell, the rim
    vertices two cells each (rows of length 4 next to rows of length n + 2)."""
    n = 40
    t = 2 * np.pi * np.arange(n) / n
    rim = np.stack([np.cos(t), np.sin(t), 0.1 * np.sin(3 * t)], axis=1)
    xyz = np.concatenate([[[0.0, 0.0, -1.0], [0.0, 0.0, 1.0]], rim])
    conn = np.array([[0, 1, 2 + k, 2 + (k + 1) % n] for k in range(n)], dtype=np.int32)
    return conn, xyz


CASES = {"one_tet": _one_tet(), "two_tets": _two_tets(), "fan": _fan()}


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _csr(conn, xyz, D, alpha, inv_dt, lumped, bc):
    import scipy.sparse as sp
    V = len(xyz)
    ip, ix, vals = OA.assemble_values(conn, xyz, D, inv_dt, alpha, lumped, bc)
    return sp.csr_matrix((vals, ix, ip), shape=(V, V))


@pytest.mark.parametrize("name", list(CASES))
def test_smallest_meshes(gpu, name):
    import torch
    from dsv.coefficients import nodal_gradient
    from dsv.core import VIEW_INDICES, VIEW_INDPTR, VIEW_SCATTER_MAP, Mesh
    from dsv.gpu_assembler import GPUSparseAssembler
    from dsv.solvers import solve_batched
    conn, xyz = CASES[name]
    C, V = len(conn), len(xyz)
    m = Mesh(conn, xyz, vertices_per_block=64)
    ip, ix = OA.pattern(conn, V)
    np.testing.assert_array_equal(m.view(VIEW_INDPTR).cpu().numpy(), ip)
    np.testing.assert_array_equal(m.view(VIEW_INDICES).cpu().numpy(), ix)
    np.testing.assert_array_equal(m.view(VIEW_SCATTER_MAP).cpu().numpy(),
                                  OA.scatter_map(conn, ip, ix))
    rng = np.random.default_rng(5)
    D, al = rng.lognormal(0, 0.5, C), rng.uniform(0, 2, V)
    bc = np.array([0], dtype=np.int32)
    for lumped in (1, 0):
        asm = GPUSparseAssembler(m)
        asm.lumped = lumped
        asm.assemble(D, 50.0, alpha=al, bc_dofs=bc)
        _, _, ref = OA.assemble_values(conn, xyz, D, 50.0, al, lumped, bc)
        assert np.abs(asm.d_data.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    # one solve per method: a field-D species with a Dirichlet vertex, a scalar-D species
    # without Dirichlet rows, and a third with a reaction everywhere
    u0 = [rng.uniform(0, 1, V) for _ in range(3)]
    spec = [dict(D=D, alpha=al, bc_dofs=bc, bc_value=1.0), dict(D_value=2.0),
            dict(D_value=5.3, alpha_value=0.7)]
    refs, bs = [], []
    for s, u in zip(spec, u0):
        A = _csr(conn, xyz, s.get("D", np.full(C, s.get("D_value", 0.0))),
                 s.get("alpha", np.full(V, s.get("alpha_value", 0.0))), 200.0, 1,
                 s.get("bc_dofs"))
        b = OA.assemble_rhs(conn, xyz, u, 200.0, None, 1, s.get("bc_dofs"), s.get("bc_value", 0.0))
        refs.append(OS.direct(A, b))
        bs.append(b)
    for method in ("csr64", "mf64", "mf32_ir64"):
        sp = [dict(s, b=torch.tensor(b, device=m.device), x=torch.tensor(u, device=m.device))
              for s, b, u in zip(spec, bs, u0)]
        st = solve_batched(m, sp, 200.0, rtol=1e-11, method=method)
        for k in range(3):
            assert st["converged"][k], (method, st)
            assert _relerr(sp[k]["x"].cpu().numpy(), refs[k]) < 1e-8, (method, k)
    # the edge-stencil gradient: the oracle's cell average, and exact for a linear field
    u = np.sin(2 * xyz[:, 0]) + xyz[:, 1] * xyz[:, 2]
    g = nodal_gradient(m, u).cpu().numpy()
    ref = OC.nodal_gradient(conn, xyz, u)
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()
    a = np.array([0.3, -1.1, 2.5])
    gl = nodal_gradient(m, xyz @ a - 0.2).cpu().numpy()
    assert np.abs(gl - a).max() < 1e-11


def test_every_vertex_dirichlet_and_mixed_batch(gpu):
    """A species whose every vertex is a Dirichlet row solves to its boundary value; batched
    with species that have no Dirichlet rows, the others still match splu."""
    import torch
    from dsv import meshes
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz = meshes.kuhn_cube(5)
    C, V = len(conn), len(xyz)
    m = Mesh(conn, xyz, vertices_per_block=64)
    rng = np.random.default_rng(9)
    everything = np.arange(V, dtype=np.int32)
    spec = [dict(D_value=1.0, bc_dofs=everything, bc_value=0.37), dict(D_value=2.0),
            dict(D=rng.lognormal(0, 0.5, C), alpha_value=1.5)]
    u0 = [rng.uniform(0, 1, V) for _ in spec]
    refs, bs = [], []
    for s, u in zip(spec, u0):
        A = _csr(conn, xyz, s.get("D", np.full(C, s.get("D_value", 0.0))),
                 np.full(V, s.get("alpha_value", 0.0)), 200.0, 1, s.get("bc_dofs"))
        b = OA.assemble_rhs(conn, xyz, u, 200.0, None, 1, s.get("bc_dofs"), s.get("bc_value", 0.0))
        refs.append(OS.direct(A, b))
        bs.append(b)
    np.testing.assert_allclose(refs[0], 0.37, rtol=0, atol=1e-15)
    for method in ("csr64", "mf64", "mf32_ir64"):
        sp = [dict(s, b=torch.tensor(b, device=m.device), x=torch.tensor(u, device=m.device))
              for s, b, u in zip(spec, bs, u0)]
        st = solve_batched(m, sp, 200.0, rtol=1e-11, method=method)
        assert all(st["converged"]), (method, st)
        assert np.abs(sp[0]["x"].cpu().numpy() - 0.37).max() <= 1e-12, method
        for k in (1, 2):
            assert _relerr(sp[k]["x"].cpu().numpy(), refs[k]) < 1e-8, (method, k)

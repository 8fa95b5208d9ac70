"""Pins the oracle's solve: exact backward-Euler fixed points, splu / scipy cg cross-checks,
and the C restatement against the numpy one (CPU)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import assembly as A
from oracle import solve as S
from dsv import meshes


def _mesh(n=5, seed=0):
    conn, xyz = meshes.kuhn_cube(n)
    rng = np.random.default_rng(seed)
    xyz = meshes.jitter(xyz, 1.0 / n, 0.2, rng, ~meshes.boundary_vertex_mask(xyz))
    return conn, xyz


@pytest.mark.parametrize("lumped", [0, 1])
def test_fixed_point_linear_field_with_matching_dirichlet(lumped):
    conn, xyz = _mesh()
    u = xyz @ np.array([1.0, -0.5, 0.25]) + 2.0
    bnd = np.nonzero(meshes.boundary_vertex_mask(xyz))[0]
    # alpha = 0 and u_old linear: u_old solves the step exactly when the boundary holds u_old
    inv_dt = 1.0 / 0.01
    ip, ix, data = A.assemble_values(conn, xyz, 1.7, inv_dt, None, lumped)
    Am = A.to_scipy(ip, ix, data, len(xyz)).tolil()
    b = A.assemble_rhs(conn, xyz, u, inv_dt, None, lumped)
    for r in bnd:
        Am.rows[r] = [r]
        Am.data[r] = [1.0]
        b[r] = u[r]
    x = S.direct(Am.tocsr(), b)
    np.testing.assert_allclose(x, u, rtol=1e-12)


@pytest.mark.parametrize("lumped", [0, 1])
def test_fixed_point_constant_field_uniform_reaction(lumped):
    conn, xyz = _mesh()
    dt, al = 0.01, 3.0
    u = np.full(len(xyz), 0.8)
    x = S.be_step(conn, xyz, u, dt, De=2.0, alpha=np.full(len(xyz), al), lumped=lumped)
    np.testing.assert_allclose(x, u / (1.0 + al * dt), rtol=1e-12)


def test_pcg_matches_direct_and_scipy_cg():
    conn, xyz = _mesh(6)
    rng = np.random.default_rng(7)
    V = len(xyz)
    bc = np.nonzero(xyz[:, 0] == 0.0)[0]
    inv_dt = 100.0
    ip, ix, data = A.assemble_values(conn, xyz, rng.lognormal(0, 0.5, len(conn)), inv_dt,
                                     rng.uniform(0, 1, V), 1, bc)
    Am = A.to_scipy(ip, ix, data, V)
    u0 = rng.uniform(0, 1, V)
    b = A.assemble_rhs(conn, xyz, u0, inv_dt, None, 1, bc, 1.0)
    xd = S.direct(Am, b)
    x, it, ok, rel = S.pcg_jacobi(Am, b, u0, 1e-12, bc_dofs=bc, bc_value=1.0)
    assert ok and rel < 1e-11
    np.testing.assert_allclose(x, xd, rtol=1e-9)
    # scipy's CG on the symmetrically lifted system agrees
    free = np.setdiff1d(np.arange(V), bc)
    g = np.zeros(V)
    g[bc] = 1.0
    Aff = Am[free][:, free]
    rhs = b[free] - (Am[free] @ g)
    xs, info = spla.cg(Aff, rhs, rtol=1e-13, maxiter=5000)
    assert info == 0
    np.testing.assert_allclose(xs, xd[free], rtol=1e-9)


def test_c_pcg_matches_numpy_pcg():
    conn, xyz = _mesh(6, seed=2)
    V = len(xyz)
    rng = np.random.default_rng(8)
    ip, ix, data = A.assemble_values(conn, xyz, 1.0, 50.0, None, 1)
    Am = A.to_scipy(ip, ix, data, V)
    b = A.assemble_rhs(conn, xyz, rng.uniform(0, 1, V), 50.0, None, 1)
    x0 = np.zeros(V)
    xn, itn, okn, reln = S.pcg_jacobi(Am, b, x0, 1e-10)
    xc, itc, relc = S.pcg_jacobi_c(ip, ix, data, b, x0, 1e-10, threads=2)
    assert abs(itc - itn) <= 1
    np.testing.assert_allclose(xc, xn, rtol=1e-9)
    assert relc <= 1e-10


def test_under_relaxation():
    a, b = np.array([1.0, 2.0]), np.array([3.0, 6.0])
    np.testing.assert_allclose(S.under_relax(a, b, 0.3), 0.3 * a + 0.7 * b)


@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("lumped", [0, 1])
def test_c_step_matches_numpy_oracle(threads, lumped):
    """oracle/assemble.c (assembly through the scatter map, RHS) against the numpy oracle, per
    cell and per vertex coefficients, with Dirichlet rows; then its PCG step against splu."""
    conn, xyz = _mesh(6, seed=3)
    rng = np.random.default_rng(9)
    C, V = len(conn), len(xyz)
    cs = S.CStep(conn, xyz)
    bc = np.arange(0, V, 11, dtype=np.int32)
    for D, al in ((rng.lognormal(0, 0.5, C), rng.uniform(0, 2, V)), (1.3, None)):
        vals = cs.assemble(D, 50.0, al, lumped, bc, threads).copy()
        ip, ix, ref = A.assemble_values(conn, xyz, D, 50.0, al, lumped, bc)
        assert np.abs(vals - ref).max() <= 1e-12 * np.abs(ref).max()
    u, f = rng.uniform(0, 1, V), rng.normal(size=V)
    b = cs.rhs(u, 50.0, f, lumped, bc, 0.4, threads)
    bref = A.assemble_rhs(conn, xyz, u, 50.0, f, lumped, bc, 0.4)
    assert np.abs(b - bref).max() <= 1e-12 * np.abs(bref).max()
    x0 = u.copy()
    x0[bc] = 0.4
    x, it, rel = cs.solve(b, x0, 1e-12, threads=threads)
    xd = S.direct(A.to_scipy(ip, ix, ref, V), bref)
    assert rel <= 1e-12 and np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()

"""Pins the CPU oracle of the Galerkin warm start (oracle/warmstart.py) on a small jittered
cube: exact answers, the Galerkin orthogonality, energy minimality by brute force, and the
dropping and fallback rules."""
import numpy as np
import pytest

from oracle import assembly as A
from oracle import solve as S
from oracle import warmstart as W
from dsv import meshes


@pytest.fixture(scope="module")
def system():
    conn, xyz = meshes.kuhn_cube(5)
    rng = np.random.default_rng(0)
    xyz = meshes.jitter(xyz, 0.2, 0.2, rng, ~meshes.boundary_vertex_mask(xyz))
    V = len(xyz)
    bc = np.nonzero(xyz[:, 0] == 0.0)[0]
    inv_dt = 50.0
    ip, ix, data = A.assemble_values(conn, xyz, 1.3, inv_dt, rng.uniform(0, 2, V), 1, bc)
    Am = A.to_scipy(ip, ix, data, V)
    u0 = rng.uniform(0, 1, V)
    b = A.assemble_rhs(conn, xyz, u0, inv_dt, rng.uniform(0, 1, V), 1, bc, 0.7)
    xs = S.direct(Am, b)
    free = np.setdiff1d(np.arange(V), bc)
    return Am, b, u0, xs, bc, free, rng


def _energy(Am, free, e):
    ef = e[free]
    return float(ef @ (Am[free][:, free] @ ef))


def test_exact_increment_gives_exact_solution(system):
    Am, b, u0, xs, bc, free, rng = system
    x0 = u0.copy()
    x0[bc] = 0.7
    x, c, n0, n1 = W.galerkin_start(Am, b, u0, [xs - x0], bc, 0.7)
    assert abs(c[0] - 1.0) < 1e-10
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-12 * np.abs(xs).max())
    assert n1 < 1e-10 * n0


def test_galerkin_orthogonality_and_energy_minimality(system):
    Am, b, u0, xs, bc, free, rng = system
    dirs = [rng.normal(0, 1, len(u0)) for _ in range(3)]
    x, c, n0, n1 = W.galerkin_start(Am, b, u0, dirs, bc, 0.7)
    assert n1 < n0 and np.all(c != 0)
    r = b - Am @ x
    for d in dirs:
        dd = d.copy()
        dd[bc] = 0
        assert abs(dd @ r) <= 1e-10 * np.linalg.norm(dd) * n0
    # no other point of the affine space has a smaller A-energy error
    x0 = u0.copy()
    x0[bc] = 0.7
    Dm = np.array(dirs).T
    Dm[bc] = 0
    best = _energy(Am, free, x - xs)
    for _ in range(200):
        cc = c + rng.normal(0, 0.05, 3) * np.abs(c).max()
        assert _energy(Am, free, x0 + Dm @ cc - xs) >= best * (1 - 1e-12)


def test_dependent_and_zero_directions_are_dropped(system):
    Am, b, u0, xs, bc, free, rng = system
    d = rng.normal(0, 1, len(u0))
    x1, c1, _, _ = W.galerkin_start(Am, b, u0, [d], bc, 0.7)
    x2, c2, _, _ = W.galerkin_start(Am, b, u0, [d, 2.0 * d, np.zeros_like(d)], bc, 0.7)
    np.testing.assert_allclose(x2, x1, rtol=1e-12, atol=1e-14)
    assert c2[1] == 0.0 and c2[2] == 0.0 and abs(c2[0] - c1[0]) <= 1e-12 * abs(c1[0])


def test_direction_orthogonal_to_residual_keeps_start(system):
    Am, b, u0, xs, bc, free, rng = system
    x0 = u0.copy()
    x0[bc] = 0.7
    r0 = b - Am @ x0
    d = rng.normal(0, 1, len(u0))
    d[bc] = 0
    d -= (d @ r0) / (r0 @ r0) * r0
    x, c, n0, n1 = W.galerkin_start(Am, b, u0, [d], bc, 0.7)
    assert abs(c[0]) < 1e-12 and n1 <= n0 * (1 + 1e-14)
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-12)


def test_solve_gram_matches_numpy():
    rng = np.random.default_rng(3)
    M = rng.normal(0, 1, (3, 3))
    G = M @ M.T + 0.1 * np.eye(3)
    g = rng.normal(0, 1, 3)
    np.testing.assert_allclose(W.solve_gram(G, g), np.linalg.solve(G, g), rtol=1e-12)


def test_nan_direction_is_dropped_and_the_others_used(system):
    # a direction holding NaN gets c = 0 and the result equals the start over the others
    Am, b, u0, xs, bc, free, rng = system
    x0 = u0.copy()
    x0[bc] = 0.7
    good = [0.5 * (xs - x0) + rng.normal(0, 1e-2, len(u0)), rng.normal(0, 1, len(u0))]
    bad = rng.normal(0, 1, len(u0))
    bad[free[:3]] = np.nan
    x, c, n0, n1 = W.galerkin_start(Am, b, u0, [good[0], bad, good[1]], bc, 0.7)
    xr, cr, m0, m1 = W.galerkin_start(Am, b, u0, good, bc, 0.7)
    assert c[1] == 0.0 and np.all(np.isfinite(x))
    np.testing.assert_allclose([c[0], c[2]], cr, rtol=1e-10)
    np.testing.assert_allclose(x, xr, rtol=0, atol=1e-12 * np.abs(xr).max())
    assert n1 < n0 and abs(n1 - m1) <= 1e-10 * n0

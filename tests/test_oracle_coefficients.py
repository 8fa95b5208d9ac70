"""Pins the CPU oracle of the coefficient fields (oracle/coefficients.py) by closed forms and a
brute-force loop."""
import numpy as np

from dsv import meshes
from oracle import coefficients as OC


def _meshes():
    rng = np.random.default_rng(0)
    c1, x1 = meshes.kuhn_cube(5)
    x1 = meshes.jitter(x1, 0.2, 0.25, rng, ~meshes.boundary_vertex_mask(x1))
    c2, x2 = meshes.graded_disc(seed=2, scale=8.0)
    return [(c1, x1), (c2, x2)]


def test_gradient_of_linear_and_constant_fields():
    a = np.array([0.7, -1.3, 2.1])
    for conn, xyz in _meshes():
        g = OC.nodal_gradient(conn, xyz, xyz @ a + 0.4)
        assert np.abs(g - a).max() < 1e-11
        g0 = OC.nodal_gradient(conn, xyz, np.full(len(xyz), 3.0))
        assert np.abs(g0).max() < 1e-11
        n = OC.grad_norm(np.tile(a, (len(xyz), 1)))
        assert np.allclose(n, np.sqrt(a @ a + 1e-12))


def test_gradient_matches_bruteforce_loop():
    conn, xyz = meshes.kuhn_cube(2)
    rng = np.random.default_rng(1)
    u = rng.normal(0, 1, len(xyz))
    acc = np.zeros((len(xyz), 3))
    ws = np.zeros(len(xyz))
    for c in conn:
        P = xyz[c]
        J = np.column_stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]])
        gref = np.array([u[c[1]] - u[c[0]], u[c[2]] - u[c[0]], u[c[3]] - u[c[0]]])
        gc = np.linalg.solve(J.T, gref)          # J^-T g_ref
        w = abs(np.linalg.det(J)) / 6.0
        for v in c:
            acc[v] += w * gc
            ws[v] += w
    assert np.allclose(OC.nodal_gradient(conn, xyz, u), acc / ws[:, None], rtol=1e-12, atol=1e-14)


def test_delta_eps_closed_forms():
    eps = 0.013
    phi = np.linspace(-2 * eps, 2 * eps, 400001)
    d = OC.delta_eps(phi, eps)
    assert abs(np.trapezoid(d, phi) - 1.0) < 1e-8
    assert np.isclose(OC.delta_eps(np.array([0.0]), eps)[0], 1.0 / eps)
    assert OC.delta_eps(np.array([eps, -eps, 3 * eps]), eps).max() == 0.0
    assert np.allclose(d, d[::-1])


def test_transport_coefficients_closed_forms():
    phi = np.array([-0.5, 0.0, 0.002, 0.5])
    gn = np.array([1.0, 2.0, 1.5, 1.0])
    F = np.array([0.0, 0.5, 2.0, -1.0])
    Cl = np.array([0.3, 0.4, 0.5, 0.6])
    O2 = np.array([1.0, 0.8, 0.6, 0.4])
    eps, k, kd, Fmax, floor = 0.01, 3.0, 0.2, 1.0, 0.0625
    al, f, film, med = OC.transport_coefficients(phi, gn, F, Cl, O2, eps, k, kd, Fmax, floor)
    dl = OC.delta_eps(phi, eps) * gn
    assert np.allclose(al, dl * k) and al[0] == 0.0 and al[3] == 0.0
    assert np.isclose(al[1], 2.0 * k / eps)  # delta_eps(0) = 1 / eps
    assert np.allclose(f, [0.0, kd * 0.5 * 0.16 + 4 * dl[1] * k * 0.8, 4 * dl[2] * k * 0.6, 0.0])
    assert np.allclose(film, [1.0, 0.5 + 0.5 * floor, floor, 1.0])
    assert np.allclose(med, film * [1, 1, 0, 0])

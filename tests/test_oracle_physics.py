"""Pins the CPU oracle of the per-step physics (oracle/physics.py) by closed forms, fixed points
and brute force (CPU)."""
import numpy as np
import pytest

from dsv import meshes
from oracle import assembly as OA
from oracle import physics as OP
from oracle import solve as OS


def _meshes():
    rng = np.random.default_rng(0)
    c1, x1 = meshes.kuhn_cube(6)
    x1 = meshes.jitter(x1, 1 / 6, 0.2, rng, ~meshes.boundary_vertex_mask(x1))
    c2, x2 = meshes.graded_disc(seed=2, scale=8.0)
    return [(c1, x1), (c2, x2)]


def test_film_update_solves_its_equation_and_clips():
    rng = np.random.default_rng(1)
    n = 2000
    phi = rng.uniform(-1, 1, n)
    Zn, Cl, F0 = rng.uniform(0, 2, n), rng.uniform(0, 2, n), rng.uniform(0, 1.2, n)
    kf, kd, Fmax, dt = 3.0, 0.7, 1.0, 0.05
    F = OP.film_update(phi, Zn, Cl, F0, kf, kd, Fmax, dt)
    med = phi <= 0
    # unclipped vertices satisfy (F - F0)/dt = kf Zn - (kf Zn / Fmax + kd Cl^2) F in the medium,
    # F = F0 in the scaffold
    free = (F > 0) & (F < Fmax)
    res = (F - F0) / dt - np.where(med, kf * Zn - (kf * Zn / Fmax + kd * Cl ** 2) * F, 0.0)
    assert np.abs(res[free]).max() < 1e-10
    assert np.all((F >= 0) & (F <= Fmax))
    np.testing.assert_allclose(F[~med], np.clip(F0[~med], 0, Fmax), rtol=1e-15)
    # no Zn and no Cl: the film holds; F_old above Fmax clips to Fmax
    np.testing.assert_allclose(OP.film_update(phi, 0 * Zn, 0 * Cl, F0, kf, kd, Fmax, dt),
                               np.clip(F0, 0, Fmax), rtol=1e-15)


def test_zinc_coefficients_and_effective_diffusion_closed_forms():
    phi = np.array([-1.0, -0.1, 0.0, 0.2, 3.0])
    F = np.array([0.0, 0.5, 1.0, 0.3, 2.0])
    Cl = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    a, s, scaffold = OP.zinc_coefficients(phi, F, Cl, kf=4.0, kd=0.5, Fmax=1.0)
    np.testing.assert_allclose(a, [4.0, 2.0, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(s, [0.0, 0.5 * 0.5 * 4.0, 0.5 * 1.0 * 9.0, 0.0, 0.0])
    assert scaffold.tolist() == [3, 4]
    D0 = (0.7, 2.0, 5.3, 2.0)
    floor = 0.25 * 1.0 / 4.0
    dz, dc, doh, do2 = OP.effective_diffusion(phi, F, 1.0, floor, D0)
    f = np.array([1.0, 0.5 + 0.5 * floor, floor, 0.7 + 0.3 * floor, floor])
    np.testing.assert_allclose(dz, 0.7 * f)
    med = (phi <= 0).astype(float)
    for d, d0 in ((dc, 2.0), (doh, 5.3), (do2, 2.0)):
        np.testing.assert_allclose(d, d0 * f * med)


@pytest.mark.parametrize("k", [0, 1])
def test_interface_velocity_closed_form_for_linear_fields(k):
    conn, xyz = _meshes()[k]
    V = len(xyz)
    nrm = np.array([0.3, -0.4, 0.866])
    nrm = nrm / np.linalg.norm(nrm)
    c0 = xyz.mean(0)
    phi = 1.7 * (xyz - c0) @ nrm                 # |grad phi| = 1.7, normal nrm
    gzn, go2 = np.array([1.0, 2.0, -0.5]), np.array([-0.3, 0.1, 0.8])
    Zn = 0.2 + xyz @ gzn
    O2 = 0.5 + xyz @ go2
    rng = np.random.default_rng(3)
    DeZn, DeO2 = rng.uniform(-0.2, 2, V), rng.uniform(-0.2, 2, V)
    h, k_orr, zs, zsat = 0.05, 0.8, 1.1e-4, 1e-8
    v = OP.interface_velocity(conn, xyz, phi, Zn, O2, DeZn, DeO2, h, k_orr, zs, zsat)
    g = 1.7
    n_eff = g * nrm / np.sqrt(g * g + 1e-12)
    v_zn = -np.maximum(DeZn, 0) * abs(gzn @ n_eff) / (zs - zsat)
    v_o2 = -2 * np.maximum(DeO2, 0) * abs(go2 @ n_eff) / zs
    v_orr = -(2 / zs) * k_orr * np.maximum(O2, 0)
    ref = np.where(np.abs(phi) < 3 * h,
                   np.minimum(np.maximum(np.maximum(v_zn, v_orr), v_o2), 0.0), 0.0)
    assert np.abs(v - ref).max() <= 1e-9 * np.abs(ref).max()
    assert (np.abs(phi) < 3 * h).sum() > 10 and np.all(v <= 0)
    assert np.all(v[np.abs(phi) >= 3 * h] == 0)
    # no band: all zero
    assert np.all(OP.interface_velocity(conn, xyz, phi + 100, Zn, O2, DeZn, DeO2, h, k_orr) == 0)
    # left-out rates: without Zn the max runs over the ORR and O2-diffusion rates, without both
    # it is the ORR law alone (the same closed forms)
    band = np.abs(phi) < 3 * h
    v2 = OP.interface_velocity(conn, xyz, phi, None, O2, None, DeO2, h, k_orr, zs, zsat)
    ref2 = np.where(band, np.minimum(np.maximum(v_orr, v_o2), 0.0), 0.0)
    assert np.abs(v2 - ref2).max() <= 1e-9 * np.abs(ref2).max()
    v3 = OP.interface_velocity(conn, xyz, phi, None, O2, None, None, h, k_orr, zs, zsat)
    assert np.array_equal(v3, np.where(band, np.minimum(v_orr, 0.0), 0.0))
    v4 = OP.interface_velocity(conn, xyz, phi, Zn, O2, DeZn, None, h, k_orr, zs, zsat)
    ref4 = np.where(band, np.minimum(np.maximum(v_zn, v_orr), 0.0), 0.0)
    assert np.abs(v4 - ref4).max() <= 1e-9 * np.abs(ref4).max()


def test_h_avg_bruteforce():
    conn, xyz = _meshes()[0]
    L = [max(np.linalg.norm(xyz[c[a]] - xyz[c[b]]) for a in range(4) for b in range(a + 1, 4))
         for c in conn]
    assert abs(OP.h_avg(conn, xyz) - np.mean(L)) < 1e-14


@pytest.mark.parametrize("k", [0, 1])
def test_redistance_fixed_point_and_one_step_closed_form(k):
    conn, xyz = _meshes()[k]
    h = OP.h_avg(conn, xyz)
    z0 = 0.5 * (xyz[:, 2].min() + xyz[:, 2].max())
    d = xyz[:, 2] - z0
    # a signed distance is a fixed point up to the 1e-12 regularisation of |grad phi|
    out = OP.redistance(conn, xyz, d, 3, h)
    assert np.abs(out - d).max() < 1e-11
    # phi = 2 d: |grad phi| = 2, so each iteration moves phi by (h/4) sign (1 - 2) off the band
    phi = 2.0 * d
    one = OP.redistance(conn, xyz, phi, 1, h)
    sign = phi / np.sqrt(phi ** 2 + (1.5 * h) ** 2)
    ref = np.where(np.abs(phi) < 3 * h, phi, phi - 0.25 * h * sign)
    assert np.abs(one - ref).max() < 1e-11
    # the frozen band and the zero level stay
    assert np.array_equal(one[np.abs(phi) < 3 * h], phi[np.abs(phi) < 3 * h])


def _coupled_case():
    conn, xyz = meshes.graded_disc(seed=1, scale=8.0)
    conn, xyz = meshes.locality_order(conn, xyz)
    from dsv import workloads
    w = workloads.coupled_disc(conn, xyz)
    return conn, xyz, w


def test_coupled_step_equals_species_steps_in_order():
    """coupled_step against the species' backward-Euler steps written out one by one, each
    solved by splu: the order (Zn, Cl, film, O2 + omega, OH) and the couplings."""
    conn, xyz, w = _coupled_case()
    p, st, dt = w["params"], w["state"], w["dt"]
    new = OP.coupled_step(conn, xyz, st, p, dt)
    V = len(xyz)
    inv_dt = 1 / dt
    De = OP.effective_diffusion(st["phi"], st["F"], p["Fmax"], p["film_floor"], p["D0"])

    def step(u, D, alpha=None, f=None, bc=None, bcv=0.0):
        ip, ix, data = OA.assemble_values(conn, xyz, D, inv_dt, alpha, 1, bc)
        b = OA.assemble_rhs(conn, xyz, u, inv_dt, f, 1, bc, bcv)
        return OS.direct(OA.to_scipy(ip, ix, data, V), b)

    med = (st["phi"] <= 0).astype(float)
    Zn = step(st["Zn"], De[0], med * p["kf"] * (1 - st["F"] / p["Fmax"]),
              med * p["kd"] * st["F"] * st["Cl"] ** 2, np.nonzero(st["phi"] > 0)[0], p["Znbc"])
    Cl = step(st["Cl"], De[1])
    alpha_f = med * (p["kf"] * Zn / p["Fmax"] + p["kd"] * Cl ** 2)
    F = np.clip((st["F"] / dt + med * p["kf"] * Zn) / (1 / dt + alpha_f), 0, p["Fmax"])
    from oracle import coefficients as OC
    band = OC.delta_eps(st["phi"], p["eps"]) * st["grad_phi"] * p["k_orr"]
    O2n = step(st["O2"], De[3], band, None, st["o2_bc"], p["o2_bc_value"])
    O2 = p["omega"] * O2n + (1 - p["omega"]) * st["O2"]
    OH = step(st["OH"], De[2], None, med * p["kd"] * F * Cl ** 2 + 4 * band * O2)
    for name, ref in (("Zn", Zn), ("Cl", Cl), ("F", F), ("O2", O2), ("OH", OH)):
        assert np.abs(new[name] - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-30), name
    assert np.all(new["Zn"][st["phi"] > 0] == p["Znbc"])
    assert new["F"].max() > 0  # the film forms where Zn and Cl meet


def _hull_fraction(P, v):
    """Exact fraction of the tet P (4 x 3) where the linear field with vertex values v is > 0:
    the convex hull volume of the positive corners and the edge crossings (scipy qhull)."""
    from scipy.spatial import ConvexHull
    pts = [P[k] for k in range(4) if v[k] > 0]
    for a in range(4):
        for b in range(a + 1, 4):
            if (v[a] > 0) != (v[b] > 0):
                t = v[a] / (v[a] - v[b])
                pts.append(P[a] + t * (P[b] - P[a]))
    if len(pts) < 4:
        return 0.0
    tet = abs(np.linalg.det(np.column_stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))) / 6
    return ConvexHull(np.array(pts)).volume / tet


def test_solid_fraction_against_convex_hull_for_every_sign_pattern():
    rng = np.random.default_rng(7)
    for trial in range(400):
        P = rng.normal(size=(4, 3))
        v = rng.normal(size=4)
        npos = trial % 5                       # force every count of positive vertices
        v = np.abs(v)
        v[rng.permutation(4)[:4 - npos]] *= -1
        f = OP.solid_fraction(v[None, :])[0]
        assert abs(f - _hull_fraction(P, v)) < 1e-9, (v, f)
    # one-signed cells
    assert OP.solid_fraction(np.array([[1.0, 2, 3, 4], [-1.0, -2, -3, -4]])).tolist() == [1.0, 0.0]


def test_exact_and_smoothed_volumes_of_a_plane():
    conn, xyz = meshes.kuhn_cube(5)
    for c in (0.37, 0.5, 0.91):
        phi = xyz[:, 2] - c                     # the plane cut is exact on every cell
        vol, frac = OP.scaffold_volume_exact(conn, xyz, phi)
        assert abs(vol - (1.0 - c)) < 1e-13
        assert np.all((frac >= 0) & (frac <= 1))
    # |phi| > eps at every quadrature point: H is the indicator, the smoothed volume the exact
    # volume of the positive cells
    phi = xyz[:, 2] - 0.5 + 0.01
    eps = 1e-3
    pv = phi[conn]
    pq = (OP.KEAST_A - OP.KEAST_B) * pv + OP.KEAST_B * pv.sum(1, keepdims=True)
    assert np.abs(pq).min() > eps
    vs, cell_h = OP.scaffold_volume_smoothed(conn, xyz, phi, eps)
    ref = ((pq > 0).mean(1) * OP.cell_volumes(conn, xyz)).sum()
    assert abs(vs - ref) < 1e-14
    # with eps large, H = (1 + phi/eps + sin(pi phi/eps)/pi)/2 = 1/2 + phi/eps + O((phi/eps)^3),
    # whose integral over the unit cube the Keast rule gives exactly: 1/2 + (1/2 - 0.49)/eps
    eps = 1e5
    vs, _ = OP.scaffold_volume_smoothed(conn, xyz, phi, eps)
    assert abs(vs - (0.5 + 0.01 / eps)) < 1e-14

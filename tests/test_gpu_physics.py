"""The per-step physics kernels and the reference-order coupled step on the GPU, through the
C-ABI, against the oracle (oracle/physics.py).

Bars: elementwise fields within 1e-14 relative (rounding of the same formulas), the scaffold
list bit-exact; gradient-based fields (velocity, redistancing) within 1e-12 relative to the
field's max; h_avg within 1e-13; the coupled step's fields within 1e-8 of the oracle's direct
solves after three steps."""
import numpy as np
import pytest

from dsv import meshes, workloads
from oracle import physics as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    from dsv.core import Mesh
    conn, xyz = meshes.graded_disc(seed=1, scale=8.0)
    conn, xyz = meshes.locality_order(conn, xyz)
    w = workloads.coupled_disc(conn, xyz)
    return Mesh(conn, xyz), conn, xyz, w


def _rel(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def _t(a, dev):
    import torch
    return torch.tensor(np.asarray(a, dtype=np.float64), device=dev)


def test_zinc_film_and_under_relaxation(case):
    from dsv import physics
    mesh, conn, xyz, w = case
    rng = np.random.default_rng(2)
    V = len(xyz)
    phi = w["state"]["phi"] + rng.normal(0, 0.01, V)
    F, Cl, Zn = rng.uniform(0, 1.3, V), rng.uniform(0, 2, V), rng.uniform(0, 2, V)
    kf, kd, Fmax, dt = 5.0, 2.0, 1.0, 0.005
    a, s, sc = physics.zinc_coefficients(mesh, phi, F, Cl, kf, kd, Fmax)
    ra, rs, rsc = OP.zinc_coefficients(phi, F, Cl, kf, kd, Fmax)
    assert _rel(a.cpu().numpy(), ra) <= 1e-14 and _rel(s.cpu().numpy(), rs) <= 1e-14
    assert np.array_equal(sc.cpu().numpy(), rsc) and len(rsc) > 0
    Fn = physics.film_update(mesh, phi, Zn, Cl, F, kf, kd, Fmax, dt)
    assert _rel(Fn.cpu().numpy(), OP.film_update(phi, Zn, Cl, F, kf, kd, Fmax, dt)) <= 1e-14
    # in place
    Ft = _t(F, mesh.device)
    physics.film_update(mesh, phi, Zn, Cl, Ft, kf, kd, Fmax, dt, out=Ft)
    assert _rel(Ft.cpu().numpy(), OP.film_update(phi, Zn, Cl, F, kf, kd, Fmax, dt)) <= 1e-14
    u = physics.under_relax(_t(Zn, mesh.device), _t(Cl, mesh.device), 0.3)
    assert _rel(u.cpu().numpy(), 0.3 * Zn + 0.7 * Cl) <= 1e-15
    # an empty scaffold: all phi <= 0
    _, _, sc0 = physics.zinc_coefficients(mesh, -np.abs(phi), F, Cl, kf, kd, Fmax)
    assert sc0.numel() == 0
    with pytest.raises(ValueError):
        physics.zinc_coefficients(mesh, phi[:-1], F, Cl, kf, kd, Fmax)


def test_interface_velocity_advance_h_avg_and_redistance(case):
    from dsv import physics
    mesh, conn, xyz, w = case
    rng = np.random.default_rng(4)
    V = len(xyz)
    h = OP.h_avg(conn, xyz)
    assert abs(physics.mesh_h_avg(mesh) - h) <= 1e-13 * h
    phi = (xyz[:, 2] - 0.1) * 1.3 + 0.02 * np.sin(7 * xyz[:, 0])
    Zn = np.exp(-20 * (xyz[:, 2] - 0.12) ** 2) + rng.uniform(0, 0.01, V)
    O2 = 1.0 - xyz[:, 2] * 2 + rng.uniform(-0.01, 0.01, V)
    DeZn, DeO2 = rng.uniform(0, 0.7, V), rng.uniform(-0.1, 2, V)
    args = (h, 1.0, 2.0, 1e-8)
    v = physics.interface_velocity(mesh, phi, Zn, O2, DeZn, DeO2, *args).cpu().numpy()
    ref = OP.interface_velocity(conn, xyz, phi, Zn, O2, DeZn, DeO2, *args)
    assert (ref != 0).sum() > 50 and _rel(v, ref) <= 1e-12
    # the rates left out: no Zn term, no O2-diffusion term, the ORR law alone
    for zn, dezn, deo2 in ((None, None, DeO2), (Zn, DeZn, None), (None, None, None)):
        got = physics.interface_velocity(mesh, phi, zn, O2, dezn, deo2, *args).cpu().numpy()
        want = OP.interface_velocity(conn, xyz, phi, zn, O2, dezn, deo2, *args)
        assert (want != 0).sum() > 50 and _rel(got, want) <= 1e-12, (zn is None, deo2 is None)
    with pytest.raises(ValueError):
        physics.interface_velocity(mesh, phi, Zn, O2, None, DeO2, *args)
    p2 = physics.levelset_advance(mesh, phi, v, 0.005).cpu().numpy()
    assert _rel(p2, phi + v * 0.005) <= 1e-15
    for iters in (1, 2, 5):
        t = _t(phi, mesh.device)
        physics.redistance_device(mesh, t, iters)
        assert _rel(t.cpu().numpy(), OP.redistance(conn, xyz, phi, iters, h)) <= 1e-12, iters


@pytest.mark.parametrize("method", ["mf32_ir64", "csr64"])
def test_coupled_step_matches_oracle(case, method):
    """Three steps of the reference-order four-species step (Zn, Cl and O2 batched, O2
    under-relaxed, the film, then OH from the new fields) against the oracle's coupled_step with
    direct solves: every field within 1e-8 of max |field| after each step."""
    from dsv.coupled import CoupledStepper
    mesh, conn, xyz, w = case
    warm = "extrapolate" if method != "csr64" else "previous"
    cs = CoupledStepper(mesh, w, method=method, rtol=1e-12, warm=warm)
    ref = dict(w["state"])
    for n in range(3):
        h = cs.step()
        assert all(h["converged"]), (n, h)
        ref = OP.coupled_step(conn, xyz, ref, w["params"], w["dt"])
        for name in ("Zn", "Cl", "O2", "OH", "F"):
            got = cs.field(name).cpu().numpy()
            assert _rel(got, ref[name]) <= 1e-8, (n, name, _rel(got, ref[name]))
    assert float(cs.F.max()) > 0.0


def test_coupled_step_with_interface_update(case):
    """With the level-set update on, one step's phi (velocity from the fields at the start of
    the step, advance, one redistancing, |grad phi|) against the oracle from the same fields;
    the solves still converge and the interface recedes (v <= 0)."""
    from dsv.coupled import CoupledStepper
    from oracle import coefficients as OC
    mesh, conn, xyz, w = case
    cs = CoupledStepper(mesh, w, rtol=1e-10, levelset=True, redistance_interval=1,
                        redistance_iters=5)
    for n in range(3):
        pre = {k: cs.field(k).cpu().numpy() for k in ("Zn", "O2", "F", "phi")}
        h = cs.step()
        assert all(h["converged"]), (n, h)
        p = w["params"]
        De = OP.effective_diffusion(pre["phi"], pre["F"], p["Fmax"], p["film_floor"], p["D0"])
        v = OP.interface_velocity(conn, xyz, pre["phi"], pre["Zn"], pre["O2"], De[0], De[3],
                                  p["h"], p["k_orr"], cs.velocity["zn_solid"],
                                  cs.velocity["zn_sat"])
        assert np.all(v <= 0)
        phi = OP.redistance(conn, xyz, pre["phi"] + v * w["dt"], 5, p["h"])
        assert _rel(cs.phi.cpu().numpy(), phi) <= 1e-12, n
        gref = OC.grad_norm(OC.nodal_gradient(conn, xyz, phi))
        assert _rel(cs.grad_phi.cpu().numpy(), gref) <= 1e-12
    assert float(np.abs(cs.phi.cpu().numpy() - w["state"]["phi"]).max()) > 0


def test_scaffold_volumes(case):
    """dsv_scaffold_volume against the oracle: the smoothed-Heaviside quadrature volume
    (GPUSparseAssembler.scaffold_volume) and the exact marching-tetrahedra volume
    (compute_scaffold_volume_exact) of a wavy level set, per cell and in total."""
    from dsv import fem, physics
    from dsv.gpu_assembler import GPUSparseAssembler
    mesh, conn, xyz, w = case
    phi = xyz[:, 2] - 0.11 + 0.03 * np.sin(9 * xyz[:, 0]) * np.cos(7 * xyz[:, 1])
    eps = 0.02
    vol, frac = physics.scaffold_volume(mesh, phi, eps, with_fraction=True)
    rvol, rfrac = OP.scaffold_volume_smoothed(conn, xyz, phi, eps)
    assert abs(vol - rvol) <= 1e-12 * rvol and _rel(frac.cpu().numpy(), rfrac) <= 1e-12
    assert abs(GPUSparseAssembler(mesh).scaffold_volume(phi, eps) - rvol) <= 1e-12 * rvol
    vol, frac = physics.scaffold_volume(mesh, phi, exact=True, with_fraction=True)
    rvol, rfrac = OP.scaffold_volume_exact(conn, xyz, phi)
    f = frac.cpu().numpy()
    assert abs(vol - rvol) <= 1e-12 * rvol and np.abs(f - rfrac).max() <= 1e-12
    assert ((f > 0) & (f < 1)).sum() > 100  # cut cells of every kind
    Vs = fem.FunctionSpace(mesh)
    pf = fem.Function(Vs)
    pf.x.array[:] = phi
    assert abs(physics.compute_scaffold_volume_exact(pf, mesh) - rvol) <= 1e-12 * rvol
    with pytest.raises(ValueError):
        physics.scaffold_volume(mesh, phi, eps=0.0)


def test_receding_interface_driver_matches_oracle(gpu):
    """The receding-interface series' interface step: the ORR law alone in the |phi| < 3 h band
    (libdsv's velocity with the Zn and O2-diffusion rates left out), phi += v dt, then
    |grad phi| and the refreshed ORR sink and OH source, against the oracle on the same fields."""
    from dsv.core import Mesh
    from dsv.transport import Stepper
    from oracle import coefficients as OC
    w = workloads.disc3(seed=1, scale=6.0)
    conn, xyz = w["conn"], w["xyz"]
    mesh = Mesh(conn, xyz)
    st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=1e-10, phi=w["phi"],
                 grad_phi=w["grad_norm"], eps=w["eps"])
    st.step()  # O2 away from its initial field
    drv = workloads.RecedingInterface(st, w, speed=0.5)
    phi0, o2 = st.phi.cpu().numpy(), st.u[0].cpu().numpy()
    drv()
    h = OP.h_avg(conn, xyz)
    v = OP.interface_velocity(conn, xyz, phi0, None, o2, None, None, h, w["k_orr"],
                              2.0 * w["k_orr"] / 0.5, 0.0)
    band = np.abs(phi0) < 3 * h
    assert band.sum() > 100 and np.all(v[~band] == 0.0) and v.min() < 0.0
    phi1 = phi0 + w["dt"] * v
    assert _rel(st.phi.cpu().numpy(), phi1) <= 1e-12
    gn = OC.grad_norm(OC.nodal_gradient(conn, xyz, phi1))
    assert _rel(st.grad_phi.cpu().numpy(), gn) <= 1e-12
    alpha, f_oh, _, _ = OC.transport_coefficients(phi1, gn, None, None, o2, w["eps"], w["k_orr"],
                                                  0.0, 1.0, 1.0)
    assert _rel(st.alpha_dev[0].cpu().numpy(), alpha) <= 1e-12
    assert _rel(st.f[2].cpu().numpy(), f_oh) <= 1e-12

"""CPU oracle of the per-step physics around the transport solves (test infrastructure only,
like the rest of oracle/): numpy restatements of

* `zinc_coefficients` -- alpha_zn = [phi <= 0] kf (1 - F/Fmax), src_zn = [phi <= 0] kd F Cl^2 and
  the scaffold Dirichlet vertices {phi > 0} of the Zn block (`dissolve.py:740-778`);
* `film_update` -- the lumped pointwise film solution clip((F_old/dt + src)/(1/dt + alpha), 0,
  Fmax) of the film block (`dissolve.py:797-833`);
* `effective_diffusion` -- D_Zn0 f and D_i0 f [phi <= 0] for Cl, OH, O2, with the film factor f
  (`physics/interface_velocity.py:90-155`);
* `interface_velocity` -- compute_interface_velocity_paper, default branch (nodal O2, no
  exact-ORR override, no probe, no R_diff) (`physics/interface_velocity.py:622-785`);
* `h_avg` -- the mean longest cell edge (`domain/mesh_utils.py:559-575`, DOLFINx mesh.h);
* `redistance` -- Sussman pseudo-time redistancing (`domain/mesh_utils.py:450-542`);
* `scaffold_volume_smoothed` -- the smoothed-Heaviside volume at the 4-point degree-2 rule
  (`gpu/gpu_assembler.py:604-643`);
* `solid_fraction`, `scaffold_volume_exact` -- the exact per-cell fraction of {phi > 0} by
  marching tetrahedra and the exact volume (`analysis/mechanics_fe.py:30-82`,
  `domain/mesh_utils.py:608-630`);
* `coupled_step` -- one time step of the reference's species blocks in their order
  (`dissolve.py:740-954`): Zn, Cl, film, O2 (under-relaxed), OH from the new fields.

Pinning (tests/test_oracle_physics.py): the film update solves its own backward-Euler equation
where it is not clipped and sits at 0 / Fmax where it is; alpha_zn vanishes at F = Fmax; the
effective diffusivities are D0 without film, D0 times the floor at saturation and 0 in the
scaffold; for linear fields on jittered and graded meshes the velocity equals the closed form
of its three rates (the nodal gradients are exact), with the band and the sign cap; a
signed-distance plane is a fixed point of redistancing and phi = 2 (z - c) moves by exactly
-(h/4) sign per iteration outside the band; h_avg against a per-cell loop; the exact fraction
against the convex-hull volume of the cut polytope (scipy) for every sign pattern, and a plane
through a Kuhn cube gives the exact volume above it; the smoothed volume is that of the
positive quadrature points when |phi| > eps at all of them, and the exact integral of a linear
field when eps is large; the coupled step against the species' own backward-Euler steps done
one by one.
"""
import numpy as np

from . import coefficients as OC
from . import solve as OS

ZN_SOLID = 7.14e-3 / 65.38   # rho_Zn / M_Zn (`physics/interface_velocity.py:49-52`)


def zinc_coefficients(phi, F, Cl, kf, kd, Fmax):
    phi = np.asarray(phi, float)
    medium = np.where(phi <= 0.0, 1.0, 0.0)
    F = np.zeros_like(phi) if F is None else np.asarray(F, float)
    Cl = np.zeros_like(phi) if Cl is None else np.asarray(Cl, float)
    alpha = medium * kf * (1.0 - F / Fmax)
    src = medium * kd * F * Cl ** 2
    return alpha, src, np.nonzero(phi > 0.0)[0].astype(np.int32)


def film_update(phi, Zn, Cl, F_old, kf, kd, Fmax, dt):
    phi = np.asarray(phi, float)
    inv_dt = 1.0 / dt
    medium = np.where(phi <= 0.0, 1.0, 0.0)
    alpha = medium * (kf * Zn / Fmax + kd * Cl ** 2)
    src = medium * kf * Zn
    return np.clip((F_old * inv_dt + src) / (inv_dt + alpha), 0.0, Fmax)


def effective_diffusion(phi, F, Fmax, film_floor, D0):
    """D0 = (D_Zn0, D_Cl0, D_OH0, D_O20) -> (DeZn, DeCl, DeOH, DeO2) per vertex."""
    _, _, film, medium = OC.transport_coefficients(phi, np.zeros(len(phi)), F, None, None, 1.0,
                                                   0.0, 0.0, Fmax, film_floor)
    return D0[0] * film, D0[1] * medium, D0[2] * medium, D0[3] * medium


def interface_velocity(conn, xyz, phi, Zn, O2, DeZn, DeO2, h, k_orr, zn_solid=ZN_SOLID,
                       zn_sat=0.0):
    phi = np.asarray(phi, float)
    denom = zn_solid - zn_sat
    if abs(denom) < 1e-30:
        denom = zn_solid
    band = np.abs(phi) < 3.0 * h
    if not band.any():
        return np.zeros_like(phi)
    gphi = OC.nodal_gradient(conn, xyz, phi)
    gmag = np.sqrt(gphi[:, 0] ** 2 + gphi[:, 1] ** 2 + gphi[:, 2] ** 2 + 1e-12)
    n = gphi / gmag[:, None]
    # a term whose fields are None (Zn and DeZn, or DeO2) is left out of the max
    v_zn = v_o2d = np.full(len(phi), -np.inf)
    if Zn is not None:
        dzn = (OC.nodal_gradient(conn, xyz, Zn) * n).sum(1)
        v_zn = -np.maximum(DeZn, 0.0) * np.abs(dzn) / denom
    if DeO2 is not None:
        do2 = (OC.nodal_gradient(conn, xyz, O2) * n).sum(1)
        v_o2d = -2.0 * np.maximum(DeO2, 0.0) * np.abs(do2) / zn_solid
    v_orr = -(2.0 / zn_solid) * k_orr * np.maximum(O2, 0.0)
    v = np.minimum(np.maximum(v_zn, np.maximum(v_orr, v_o2d)), 0.0)
    return np.where(band, v, 0.0)


def h_avg(conn, xyz):
    p = np.asarray(xyz, float)[np.asarray(conn)]
    L = np.zeros(len(conn))
    for a in range(4):
        for b in range(a + 1, 4):
            L = np.maximum(L, np.linalg.norm(p[:, a] - p[:, b], axis=1))
    return float(L.mean())


def redistance(conn, xyz, phi, num_iters, h):
    phi0 = np.array(phi, float)
    phi = phi0.copy()
    frozen = np.abs(phi0) < 3.0 * h
    sign = phi0 / np.sqrt(phi0 ** 2 + (1.5 * h) ** 2)
    for _ in range(num_iters):
        gm = OC.grad_norm(OC.nodal_gradient(conn, xyz, phi))
        dtau = 0.5 * h / np.maximum(gm, 0.5)
        phi = phi + np.where(frozen, 0.0, dtau * sign * (1.0 - gm))
    return phi


def coupled_step(conn, xyz, st, p, dt, rtol=None):
    """One step of the species blocks in the reference's order on the state dict st (phi,
    grad_phi, Zn, Cl, O2, OH, F: per-vertex arrays; o2_bc: O2 Dirichlet vertices) with the
    parameters p (kf, kd, Fmax, film_floor, D0, Znbc, o2_bc_value, eps, k_orr, omega).
    Returns the new state.  Direct solves (rtol None) or Jacobi-PCG from the current fields."""
    phi = st["phi"]
    DeZn, DeCl, DeOH, DeO2 = effective_diffusion(phi, st["F"], p["Fmax"], p["film_floor"], p["D0"])
    # Zn: F and Cl of the previous step, scaffold Dirichlet rows (`dissolve.py:740-778`)
    a_zn, s_zn, scaffold = zinc_coefficients(phi, st["F"], st["Cl"], p["kf"], p["kd"], p["Fmax"])
    Zn = OS.be_step(conn, xyz, st["Zn"], dt, DeZn, alpha=a_zn, f=s_zn, bc_dofs=scaffold,
                    bc_value=p["Znbc"], rtol=rtol)
    # Cl (`:781-791`)
    Cl = OS.be_step(conn, xyz, st["Cl"], dt, DeCl, rtol=rtol)
    # film from the new Zn and Cl (`:797-833`)
    F = film_update(phi, Zn, Cl, st["F"], p["kf"], p["kd"], p["Fmax"], dt)
    # O2 with the ORR sink, bottom Dirichlet and under-relaxation (`:863-903`)
    a_o2, _, _, _ = OC.transport_coefficients(phi, st["grad_phi"], None, None, None, p["eps"],
                                              p["k_orr"], 0.0, 1.0, 1.0)
    O2n = OS.be_step(conn, xyz, st["O2"], dt, DeO2, alpha=a_o2, bc_dofs=st["o2_bc"],
                     bc_value=p["o2_bc_value"], rtol=rtol)
    O2 = OS.under_relax(O2n, st["O2"], p["omega"])
    # OH with the source from the new F, Cl and O2 (`:919-951`)
    _, f_oh, _, _ = OC.transport_coefficients(phi, st["grad_phi"], F, Cl, O2, p["eps"],
                                              p["k_orr"], p["kd"], p["Fmax"], 1.0)
    OH = OS.be_step(conn, xyz, st["OH"], dt, DeOH, f=f_oh, rtol=rtol)
    out = dict(st)
    out.update(Zn=Zn, Cl=Cl, F=F, O2=O2, OH=OH, O2_raw=O2n)
    return out


KEAST_A = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
KEAST_B = (5.0 - np.sqrt(5.0)) / 20.0


def smoothed_heaviside(p, eps):
    return np.where(p < -eps, 0.0, np.where(p > eps, 1.0,
                    0.5 * (1.0 + p / eps + np.sin(np.pi * p / eps) / np.pi)))


def cell_volumes(conn, xyz):
    p = np.asarray(xyz, float)[np.asarray(conn)]
    return np.abs(np.einsum("ij,ij->i", p[:, 1] - p[:, 0],
                            np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))) / 6.0


def scaffold_volume_smoothed(conn, xyz, phi, eps):
    pv = np.asarray(phi, float)[np.asarray(conn)]
    pq = (KEAST_A - KEAST_B) * pv + KEAST_B * pv.sum(axis=1, keepdims=True)
    cell_h = smoothed_heaviside(pq, eps).mean(axis=1)
    return float((cell_h * cell_volumes(conn, xyz)).sum()), cell_h


def _corner(v, a):
    r = np.ones(len(v))
    for k in range(4):
        m = a != k
        r[m] *= v[m, a[m]] / (v[m, a[m]] - v[m, k])
    return r


def _tetvol(a, b, c, d):
    return np.abs(np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))) / 6.0


def solid_fraction(phiv):
    """Per-cell fraction where the linear field with vertex values phiv (n x 4) is > 0."""
    v = np.asarray(phiv, float)
    n = len(v)
    pos = v > 0.0
    npos = pos.sum(axis=1)
    f = np.zeros(n)
    f[npos == 4] = 1.0
    m = npos == 1
    if m.any():
        f[m] = _corner(v[m], np.argmax(pos[m], axis=1))
    m = npos == 3
    if m.any():
        f[m] = 1.0 - _corner(v[m], np.argmin(pos[m], axis=1))
    m = npos == 2
    if m.any():
        w = v[m]
        order = np.argsort(~(w > 0.0), axis=1, kind="stable")
        r = np.arange(len(w))
        pa, pb, na, nb = (w[r, order[:, k]] for k in range(4))
        z = np.zeros(len(w))
        A = np.stack([z, z, z], 1)
        B = np.stack([z + 1, z, z], 1)
        q1 = np.stack([z, pa / (pa - na), z], 1)
        q2 = np.stack([z, z, pa / (pa - nb)], 1)
        q3 = np.stack([1 - pb / (pb - na), pb / (pb - na), z], 1)
        q4 = np.stack([1 - pb / (pb - nb), z, pb / (pb - nb)], 1)
        f[m] = 6.0 * (_tetvol(A, q1, q2, B) + _tetvol(q1, q2, B, q3) + _tetvol(q2, B, q3, q4))
    return np.clip(f, 0.0, 1.0)


def scaffold_volume_exact(conn, xyz, phi):
    frac = solid_fraction(np.asarray(phi, float)[np.asarray(conn)])
    return float((frac * cell_volumes(conn, xyz)).sum()), frac

"""The transport workloads: meshes, coefficients and initial states (host numpy).

Each builder returns a dict with the mesh (conn, xyz), dt, and a list of species dicts
(name, D (array or None) / D_value, alpha (per-vertex array or None), f (per-vertex source or
None), bc_dofs, bc_value, u0) plus the level set phi.  The species coefficients follow the
caller's builders in the reference: the ORR sink alpha_O2 = delta_eps(phi) |grad phi| k_orr
and the OH source 4 delta_eps |grad phi| k_orr O2 (`dissolve.py:866-875,923-936`), with the
cosine delta of `physics/weak_forms.py:18-24`.

* cube_single    -- unit cube 20^3 Kuhn (48,000 tets), one species, FP64, D=1, alpha=0.1 where
                    phi = z-0.5 < 0, dt=0.01, u0 ~ U(0,1), Dirichlet c=1 on x=0, rtol 1e-10.
* cube_jitter    -- 48^3 Kuhn (663,552 tets) with interior jitter U(+-0.2h); D_e ~
                    lognormal(0, 0.5) per cell; alpha=1 where |x-c|-0.3 < 0; rtol 1e-8.
* disc3          -- graded Delaunay disc, r=1, height 0.2, ~800k tets; O2 / Cl / OH with
                    D = 2.0 / 2.0 / 5.3, dt = 0.005, ORR sink on O2, OH source on phi < 0,
                    O2 Dirichlet (bulk value) on the bottom face.
* gyroid3        -- gyroid lattice 160^3 Kuhn (14.7M tets), same three species.
* RecedingInterface -- the disc3 time series with the interface receding by the ORR law in
                    the |phi| < 3 h band (speed proportional to local O2), coefficients
                    recomputed on the GPU every step.
"""
from __future__ import annotations

import numpy as np

from . import meshes

SPECIES3 = ("O2", "Cl", "OH")
D3 = (2.0, 2.0, 5.3)


def delta_eps(phi, eps):
    """Cosine-regularised Dirac (`physics/weak_forms.py:18-24`)."""
    return np.where(np.abs(phi) < eps, (1.0 / (2.0 * eps)) * (1.0 + np.cos(np.pi * phi / eps)), 0.0)


def smoothed_heaviside(phi, eps):
    """`dissolve.py:345-348`."""
    return np.where(phi > eps, 1.0, np.where(phi < -eps, 0.0,
                    0.5 * (1.0 + phi / eps + np.sin(np.pi * phi / eps) / np.pi)))


def nodal_grad_norm(conn, xyz, phi):
    """|grad phi| at vertices: volume-weighted average of the cell gradients, then
    sqrt(g.g + 1e-12) (`physics/interface_velocity.py:65-87`, `domain/mesh_utils.py:22-78`)."""
    p = xyz[conn]
    J = (p[:, 1:, :] - p[:, 0:1, :]).transpose(0, 2, 1)
    det = np.linalg.det(J)
    JinvT = np.linalg.inv(J).transpose(0, 2, 1)
    ref = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    g_ref = np.einsum("ca,ad->cd", phi[conn], ref)
    g = np.einsum("cij,cj->ci", JinvT, g_ref)
    w = np.abs(det) / 6.0
    acc = np.zeros((len(xyz), 3))
    ws = np.zeros(len(xyz))
    for k in range(4):
        np.add.at(acc, conn[:, k], w[:, None] * g)
        np.add.at(ws, conn[:, k], w)
    g = acc / np.maximum(ws, 1e-300)[:, None]
    return np.sqrt((g * g).sum(1) + 1e-12)


def h_avg(conn, xyz):
    """Mean over cells of the longest edge (restatement of the cell size used for the ORR band
    width, `dissolve.py:325-328,371`)."""
    p = xyz[conn]
    L = np.zeros(len(conn))
    for a in range(4):
        for b in range(a + 1, 4):
            L = np.maximum(L, np.linalg.norm(p[:, a] - p[:, b], axis=1))
    return float(L.mean())


def cube_single(n=20, seed=0):
    conn, xyz = meshes.kuhn_cube(n)
    rng = np.random.default_rng(seed)
    phi = xyz[:, 2] - 0.5
    alpha = np.where(phi < 0.0, 0.1, 0.0)
    bc = np.nonzero(xyz[:, 0] == 0.0)[0].astype(np.int32)
    u0 = rng.uniform(0.0, 1.0, len(xyz))
    sp = dict(name="c", D=None, D_value=1.0, alpha=alpha, f=None, bc_dofs=bc, bc_value=1.0, u0=u0)
    return dict(name=f"cube{n}_single", conn=conn, xyz=xyz, dt=0.01, phi=phi, species=[sp],
                rtol=1e-10)


def cube_jitter(n=48, seed=0, order=True):
    conn, xyz = meshes.kuhn_cube(n)
    rng = np.random.default_rng(seed)
    interior = ~meshes.boundary_vertex_mask(xyz)
    xyz = meshes.jitter(xyz, 1.0 / n, 0.2, rng, interior)
    if order:
        conn, xyz = meshes.locality_order(conn, xyz)
    D = rng.lognormal(0.0, 0.5, len(conn))
    phi = np.linalg.norm(xyz - 0.5, axis=1) - 0.3
    alpha = np.where(phi < 0.0, 1.0, 0.0)
    u0 = rng.uniform(0.0, 1.0, len(xyz))
    sp = dict(name="c", D=D, D_value=0.0, alpha=alpha, f=None, bc_dofs=None, bc_value=0.0, u0=u0)
    return dict(name=f"cube{n}_jitter", conn=conn, xyz=xyz, dt=0.01, phi=phi, species=[sp],
                rtol=1e-8)


def three_species(conn, xyz, phi, bc_dofs, dt, eps, k_orr=1.0, c0=(1.0, 1.0, 0.1), name=""):
    """O2 / Cl / OH on a mesh with level set phi (phi > 0 metal, phi <= 0 medium)."""
    ng = nodal_grad_norm(conn, xyz, phi)
    dl = delta_eps(phi, eps) * ng
    H = smoothed_heaviside(phi, eps)
    medium = 1.0 - H
    o2 = c0[0] * medium
    cl = c0[1] * medium
    oh = c0[2] * medium
    alpha_o2 = dl * k_orr
    f_oh = np.where(phi < 0.0, 4.0 * alpha_o2 * o2, 0.0)
    sp = [
        dict(name="O2", D=None, D_value=D3[0], alpha=alpha_o2, f=None, bc_dofs=bc_dofs,
             bc_value=c0[0], u0=o2),
        dict(name="Cl", D=None, D_value=D3[1], alpha=None, f=None, bc_dofs=None, bc_value=0.0,
             u0=cl),
        dict(name="OH", D=None, D_value=D3[2], alpha=None, f=f_oh, bc_dofs=None, bc_value=0.0,
             u0=oh),
    ]
    return dict(name=name, conn=conn, xyz=xyz, dt=dt, phi=phi, species=sp, rtol=1e-8, eps=eps,
                k_orr=k_orr, grad_norm=ng)


# nondimensional parameters of the four-species coupled step (same roles as the reference's
# `config/parameters.py` kf, kd, Fmax, film porosity / constrictivity / tortuosity, D_i0, Znbc)
COUPLED_PARAMS = dict(kf=5.0, kd=2.0, Fmax=1.0, film_floor=0.25 * 1.0 / 2.0 ** 2,
                      D0=(0.7, 2.0, 5.3, 2.0), Znbc=1.0, o2_bc_value=1.0, k_orr=1.0, omega=0.3,
                      orr_eps_factor=2.0)


def coupled_disc(conn, xyz, z_if=0.19, dt=0.005, name="coupled"):
    """The four-species step of the reference's species blocks (`dissolve.py:740-954`) on a disc
    mesh: Zn (scaffold Dirichlet at phi > 0), Cl, the film, O2 (ORR sink, bottom Dirichlet,
    under-relaxation 0.3) and OH (source from the new film, Cl and O2), with the effective
    diffusivities of the film (`physics/interface_velocity.py:90-155`).  phi = z - z_if (the
    metal layer on top), the ORR band eps = 2 h_avg (`dissolve.py:371`), no film at the start.
    Returns dict(conn, xyz, dt, params, state) with state = phi, grad_phi, Zn, Cl, O2, OH, F
    (per-vertex arrays) and o2_bc (the bottom-face vertices)."""
    p = dict(COUPLED_PARAMS)
    phi = xyz[:, 2] - z_if
    h = h_avg(conn, xyz)
    p["eps"] = p["orr_eps_factor"] * h
    p["h"] = h
    medium = 1.0 - smoothed_heaviside(phi, p["eps"])
    st = dict(phi=phi, grad_phi=nodal_grad_norm(conn, xyz, phi),
              Zn=np.where(phi > 0.0, p["Znbc"], 0.0), Cl=medium.copy(), O2=medium.copy(),
              OH=0.1 * medium, F=np.zeros(len(xyz)),
              o2_bc=np.nonzero(xyz[:, 2] <= 1e-12)[0].astype(np.int32))
    return dict(name=name, conn=conn, xyz=xyz, dt=dt, params=p, state=st)


def disc3(seed=0, scale=1.47, xy_factor=0.94, z_if=0.19):
    """seed 0: 805,903 tets, 165,863 vertices."""
    conn, xyz = meshes.graded_disc(seed=seed, scale=scale, xy_factor=xy_factor)
    conn, xyz = meshes.locality_order(conn, xyz)
    phi = xyz[:, 2] - z_if                      # metal layer on top, medium below
    eps = 1.5 * 0.005 * scale                   # band half-width ~1.5 h at the interface
    bc = np.nonzero(xyz[:, 2] <= 1e-12)[0].astype(np.int32)   # the whole bottom layer
    return three_species(conn, xyz, phi, bc, 0.005, eps, name="disc3")


def gyroid3(n=160, level=0.3):
    conn, xyz = meshes.gyroid_lattice(n, level)
    conn, xyz = meshes.locality_order(conn, xyz)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    g = np.sin(x) * np.cos(y) + np.sin(y) * np.cos(z) + np.sin(z) * np.cos(x)
    h = 4.0 * np.pi / n
    phi = g                                     # metal shell 0 < g < level, medium g <= 0
    bc = np.nonzero(x <= 1e-12)[0].astype(np.int32)
    return three_species(conn, xyz, phi, bc, 0.005, 1.5 * h, name="gyroid3")


class RecedingInterface:
    """The receding-interface time series on a three-species workload: before each step the
    level set recedes with the ORR law alone, v = -(2 / zn_solid) k_orr max(O2, 0) inside the
    band |phi| < 3 h (h the mean cell size) and 0 outside (libdsv dsv_interface_velocity with
    the Zn-dissolution and O2-diffusion rates left out, `physics/interface_velocity.py:741,749`),
    zn_solid chosen so that v = -speed O2 in the band; then phi += v dt (dsv_levelset_advance,
    `dissolve.py:684-690`), |grad phi| (dsv_nodal_gradient), and the step's coefficients: the
    ORR sink alpha_O2 = delta_eps(phi) |grad phi| k_orr and the OH source
    4 delta_eps |grad phi| k_orr O2 (dsv_transport_coefficients, the species blocks'
    `dissolve.py:866-875,923-936` with no film).  Every field stays on the device and every
    update is a libdsv kernel.

    Call the instance before each stepper.step(); it updates stepper.phi and stepper.grad_phi
    in place (the learned warm start's features) and refreshes the stepper's alpha_O2 and f_OH.
    """

    def __init__(self, stepper, w, speed=0.008):
        import torch
        from .physics import mesh_h_avg
        self.st = stepper
        self.w = w
        if stepper.phi is None or stepper.grad_phi is None:
            raise ValueError("the stepper needs phi and grad_phi")
        if not speed > 0.0:
            raise ValueError("speed must be positive")
        self.k_orr = float(w["k_orr"])
        self.zn_solid = 2.0 * self.k_orr / float(speed)
        self.h = mesh_h_avg(stepper.mesh)
        self.v = torch.empty_like(stepper.phi)

    def __call__(self):
        from .coefficients import grad_norm, transport_coefficients
        from .physics import interface_velocity, levelset_advance
        st, w = self.st, self.w
        mesh = st.mesh
        o2 = st.u[0]
        interface_velocity(mesh, st.phi, None, o2, None, None, self.h, self.k_orr, self.zn_solid,
                           0.0, out=self.v)
        levelset_advance(mesh, st.phi, self.v, st.dt, out=st.phi)
        grad_norm(mesh, st.phi, out=st.grad_phi)
        c = transport_coefficients(mesh, st.phi, st.grad_phi, w["eps"], w["k_orr"], O2=o2)
        st.refresh_coefficients(alpha=[c["alpha_o2"], None, None], f=[None, None, c["f_oh"]])

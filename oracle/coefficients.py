"""CPU oracle of the upstream coefficient fields of the transport step (test infrastructure only,
like the rest of oracle/).

* `nodal_gradient` -- the volume-weighted average of the cells' constant P1 gradients, weights
  |det J| / 6, scatter-add then divide, as the reference's GPU kernel
  (`gpu/gpu_assembler.py:251-307`, wrapper `:646-671`, J^-T from the inverse of J as in
  `:379-393`);
* `grad_norm` -- sqrt(gx^2 + gy^2 + gz^2 + 1e-12) (`physics/interface_velocity.py:65-87`);
* `delta_eps`, `transport_coefficients` -- the ORR sink alpha_O2 = delta_eps(phi) |grad phi| k_orr,
  the OH source [phi <= 0] kd F Cl^2 + 4 delta_eps |grad phi| k_orr O2 (`dissolve.py:866-875,
  923-936`) and the film / medium factors of the effective diffusivities
  (`physics/interface_velocity.py:90-155`).

Pinning (tests/test_oracle_coefficients.py): the gradient of a linear field is exact at every
vertex and that of a constant is 0, on jittered and graded meshes; a per-cell brute-force loop;
delta_eps integrates to 1 over phi, peaks at 1/eps and vanishes at |phi| >= eps; the film factor
is 1 without film and the floor at and beyond saturation.
"""
import numpy as np


def nodal_gradient(conn, xyz, u):
    conn = np.asarray(conn, dtype=np.int64)
    p = np.asarray(xyz, float)[conn]                     # C x 4 x 3
    J = (p[:, 1:, :] - p[:, 0:1, :]).transpose(0, 2, 1)  # columns e1, e2, e3
    detJ = np.abs(np.linalg.det(J))
    JinvT = np.linalg.inv(J).transpose(0, 2, 1)
    uc = np.asarray(u, float)[conn]
    g_ref = uc[:, 1:] - uc[:, 0:1]                        # sum_a u_a grad N_ref[a]
    g = np.einsum("cij,cj->ci", JinvT, g_ref)
    w = detJ / 6.0
    V = len(xyz)
    acc = np.zeros((V, 3))
    ws = np.zeros(V)
    for a in range(4):
        np.add.at(acc, conn[:, a], w[:, None] * g)
        np.add.at(ws, conn[:, a], w)
    return acc / ws[:, None]


def grad_norm(g):
    g = np.asarray(g, float)
    return np.sqrt(g[:, 0] ** 2 + g[:, 1] ** 2 + g[:, 2] ** 2 + 1e-12)


def delta_eps(phi, eps):
    phi = np.asarray(phi, float)
    return np.where(np.abs(phi) < eps, (1.0 / (2.0 * eps)) * (1.0 + np.cos(np.pi * phi / eps)), 0.0)


def transport_coefficients(phi, gnorm, F, Cl, O2, eps, k_orr, kd, Fmax, film_floor):
    """(alpha_o2, f_oh, film_factor, medium_factor) per vertex; F, Cl, O2 may be None (0)."""
    phi = np.asarray(phi, float)
    V = len(phi)
    z = np.zeros(V)
    F = z if F is None else np.asarray(F, float)
    Cl = z if Cl is None else np.asarray(Cl, float)
    O2 = z if O2 is None else np.asarray(O2, float)
    dl = delta_eps(phi, eps) * np.asarray(gnorm, float)
    medium = np.where(phi <= 0.0, 1.0, 0.0)
    alpha = dl * k_orr
    f_oh = medium * kd * F * Cl ** 2 + 4.0 * dl * k_orr * O2
    sat = np.clip(F / max(Fmax, 1e-20), 0.0, 1.0)
    film = (1.0 - sat) + sat * film_floor
    return alpha, f_oh, film, film * medium

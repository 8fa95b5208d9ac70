"""The per-step physics around the transport solves, on the GPU (libdsv physics kernels), with
the reference's function names where it has them:

* `zinc_coefficients(mesh, phi, F_film, Cl, kf, kd, Fmax)` -- alpha_zn, src_zn and the
  scaffold Dirichlet vertices of the Zn block (`dissolve.py:740-778`);
* `film_update(mesh, phi, Zn, Cl, F_old, kf, kd, Fmax, dt)` -- the film block's pointwise
  solution (`dissolve.py:797-833`);
* `under_relax(u_new, u_prev, omega)` -- solve_transport's under-relaxation
  (`numerics/solvers.py:278-279`);
* `interface_velocity(mesh, ...)` and `compute_interface_velocity_paper(phi, Zn, O2, DeZn, DeO2,
  v_field, params, h, t, gpu_asm)` -- `physics/interface_velocity.py:622-785` (default branch);
* `levelset_advance(mesh, phi_old, v, dt)` -- `dissolve.py:684-690`;
* `mesh_h_avg(mesh)` and `redistance(phi, domain, num_iters, dt_redist, h, gpu_asm)` --
  `domain/mesh_utils.py:450-575`;
* `scaffold_volume(mesh, phi, eps)`, `compute_scaffold_volume(phi, domain, epsilon, gpu_asm)`
  and `compute_scaffold_volume_exact(phi, domain)` -- the smoothed and the exact (marching
  tetrahedra) volume of the scaffold (`gpu/gpu_assembler.py:604-643`,
  `domain/mesh_utils.py:578-630`, `analysis/mechanics_fe.py:45-82`).

Fields are torch tensors on the mesh's device; the reference-signature functions also take
objects with `.x.array` (DOLFINx Functions, dsv.fem.Function), copied in and out like the
reference's cupy round trips.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .core import ptr, stream_ptr, to_device


def _dev(mesh, a, name="field"):
    import torch
    if a is None:
        return None
    if hasattr(a, "x"):
        a = a.x.array
    t = to_device(a, mesh.device, torch.float64)
    if t.numel() != mesh.n_dofs:
        raise ValueError(f"{name} length {t.numel()} != n_dofs {mesh.n_dofs}")
    return t


def _empty(mesh):
    import torch
    return torch.empty(mesh.n_dofs, dtype=torch.float64, device=mesh.device)


def _put(field, t):
    if hasattr(field, "x"):
        field.x.array[:] = t.cpu().numpy()
        field.x.scatter_forward()
    else:
        field.copy_(t)


def zinc_coefficients(mesh, phi, F_film, Cl, kf, kd, Fmax, scaffold=True, out=None):
    """(alpha_zn, src_zn, scaffold vertices or None) as device tensors; `out` may give the
    (alpha_zn, src_zn) tensors to fill."""
    import torch
    d_phi, d_F, d_Cl = _dev(mesh, phi, "phi"), _dev(mesh, F_film, "F"), _dev(mesh, Cl, "Cl")
    a, s = out if out is not None else (_empty(mesh), _empty(mesh))
    p = _lib.FilmParams(float(kf), float(kd), float(Fmax))
    sc, n = None, ctypes.c_int32(0)
    if scaffold:
        sc = torch.empty(mesh.n_dofs, dtype=torch.int32, device=mesh.device)
    _lib.call("dsv_zinc_coefficients", mesh.handle, ptr(d_phi), ptr(d_F), ptr(d_Cl),
              ctypes.byref(p), ptr(a), ptr(s), ptr(sc), ctypes.byref(n) if scaffold else None,
              stream_ptr(mesh.device))
    return a, s, (sc[:n.value] if scaffold else None)


def film_update(mesh, phi, Zn, Cl, F_old, kf, kd, Fmax, dt, out=None):
    """F_new (device tensor; `out` may be F_old itself)."""
    d = [_dev(mesh, a, n) for a, n in ((phi, "phi"), (Zn, "Zn"), (Cl, "Cl"), (F_old, "F_old"))]
    out = _empty(mesh) if out is None else out
    p = _lib.FilmParams(float(kf), float(kd), float(Fmax))
    _lib.call("dsv_film_update", mesh.handle, *[ptr(t) for t in d], ctypes.byref(p),
              1.0 / float(dt), ptr(out), stream_ptr(mesh.device))
    return out


def under_relax(u_new, u_prev, omega, out=None):
    """omega u_new + (1 - omega) u_prev for FP64 device tensors of equal length."""
    import torch
    if u_new.numel() != u_prev.numel() or u_new.dtype != torch.float64 or \
            u_prev.dtype != torch.float64 or not (u_new.is_contiguous() and u_prev.is_contiguous()):
        raise ValueError("under_relax takes contiguous FP64 device tensors of equal length")
    out = torch.empty_like(u_new) if out is None else out
    _lib.call("dsv_under_relax", int(u_new.numel()), ptr(u_new), ptr(u_prev), float(omega),
              ptr(out), stream_ptr(u_new.device))
    return out


def mesh_h_avg(mesh):
    h = ctypes.c_double()
    _lib.call("dsv_mesh_h_avg", mesh.handle, ctypes.byref(h))
    return h.value


def interface_velocity(mesh, phi, Zn, O2, DeZn, DeO2, h, k_orr, zn_solid, zn_sat, out=None):
    """The recession velocity per vertex (device tensor).  Zn and DeZn None (together) leave the
    Zn-dissolution rate out, DeO2 None the O2-diffusion rate: with both, the ORR law alone."""
    if (Zn is None) != (DeZn is None):
        raise ValueError("Zn and DeZn are given together or not at all")
    d = [_dev(mesh, a, n) for a, n in ((phi, "phi"), (Zn, "Zn"), (O2, "O2"), (DeZn, "DeZn"),
                                       (DeO2, "DeO2"))]
    out = _empty(mesh) if out is None else out
    p = _lib.VelocityParams(float(h), float(k_orr), float(zn_solid), float(zn_sat))
    _lib.call("dsv_interface_velocity", mesh.handle, *[ptr(t) for t in d], ctypes.byref(p),
              ptr(out), stream_ptr(mesh.device))
    return out


def compute_interface_velocity_paper(phi, Zn, O2, DeZn, DeO2, v_field, params, h, t,
                                     gpu_asm=None, exact_O2_override=None):
    """Reference signature (`physics/interface_velocity.py:622-785`): v_field <- the velocity.
    Only the default branch is provided: nodal O2, no exact-ORR override, no probe distance,
    no legacy R_diff term."""
    if exact_O2_override is not None or float(getattr(params, "orr_probe_dist", 0.0) or 0.0) > 0 \
            or getattr(params, "legacy_vorr_r_diff", False):
        raise NotImplementedError("only the nodal-O2 branch of the velocity law is provided")
    from .coefficients import _mesh_of
    mesh = _mesh_of(phi, gpu_asm)
    zn_solid = getattr(params, "Znsolid", None) or getattr(params, "zn_solid")
    v = interface_velocity(mesh, phi, Zn, O2, DeZn, DeO2, h, getattr(params, "k_orr", 0.015),
                           zn_solid, getattr(params, "Znbc", 0.0))
    _put(v_field, v)


def levelset_advance(mesh, phi_old, v, dt, out=None):
    d_old, d_v = _dev(mesh, phi_old, "phi_old"), _dev(mesh, v, "v")
    out = _empty(mesh) if out is None else out
    _lib.call("dsv_levelset_advance", mesh.handle, ptr(d_old), ptr(d_v), float(dt), ptr(out),
              stream_ptr(mesh.device))
    return out


def redistance_device(mesh, phi, num_iters, h=None):
    """In place on the FP64 device tensor phi; h None -> the mesh's h_avg."""
    _lib.call("dsv_redistance", mesh.handle, ptr(phi), int(num_iters),
              float(h) if h is not None else 0.0, stream_ptr(mesh.device))
    return phi


def redistance(phi, domain=None, num_iters: int = 8, dt_redist: float = None, h: float = None,
               gpu_asm=None):
    """Reference signature (`domain/mesh_utils.py:450-542`).  As in the reference, the band,
    the sign smoothing and the pseudo time step all use h_avg; `h` is accepted and unused.  A
    fixed dt_redist is not provided."""
    if dt_redist is not None:
        raise NotImplementedError("a fixed pseudo time step is not provided")
    from .coefficients import _mesh_of
    mesh = gpu_asm.mesh if gpu_asm is not None else (
        getattr(domain, "dsv_mesh", domain) if domain is not None else _mesh_of(phi))
    t = _dev(mesh, phi, "phi").clone()
    redistance_device(mesh, t, num_iters)
    _put(phi, t)


def scaffold_volume(mesh, phi, eps=None, exact=False, with_fraction=False):
    """Volume of {phi > 0}: the smoothed-Heaviside quadrature with half-width eps, or the exact
    marching-tetrahedra volume (exact=True); with_fraction also returns the per-cell values."""
    import torch
    if not exact and not (eps is not None and eps > 0):
        raise ValueError("the smoothed volume needs eps > 0")
    d_phi = _dev(mesh, phi, "phi")
    frac = torch.empty(mesh.n_cells, dtype=torch.float64, device=mesh.device) \
        if with_fraction else None
    vol = ctypes.c_double()
    _lib.call("dsv_scaffold_volume", mesh.handle, ptr(d_phi), 1 if exact else 0,
              float(eps) if eps is not None else 0.0, ptr(frac), ctypes.byref(vol),
              stream_ptr(mesh.device))
    return (vol.value, frac) if with_fraction else vol.value


def compute_scaffold_volume(phi, domain=None, epsilon: float = None, gpu_asm=None) -> float:
    """Reference signature (`domain/mesh_utils.py:578-605`); epsilon defaults to 0.5 h_min, h_min
    the smallest cell diameter (longest edge, `domain/mesh_utils.py:545-556`)."""
    from .coefficients import _mesh_of
    mesh = gpu_asm.mesh if gpu_asm is not None else (
        getattr(domain, "dsv_mesh", domain) if domain is not None else _mesh_of(phi))
    if epsilon is None:
        epsilon = 0.5 * mesh_h_min(mesh)
    return scaffold_volume(mesh, phi, epsilon)


def compute_scaffold_volume_exact(phi, domain=None) -> float:
    """Reference signature (`domain/mesh_utils.py:608-630`)."""
    from .coefficients import _mesh_of
    mesh = getattr(domain, "dsv_mesh", domain) if domain is not None else _mesh_of(phi)
    return scaffold_volume(mesh, phi, exact=True)


def mesh_h_min(mesh):
    """Smallest longest-edge over the cells (host, from the mesh's copy of the coordinates)."""
    import numpy as np
    p = mesh.xyz[mesh.conn]
    L = np.zeros(len(mesh.conn))
    for a in range(4):
        for b in range(a + 1, 4):
            L = np.maximum(L, np.linalg.norm(p[:, a] - p[:, b], axis=1))
    return float(L.min())

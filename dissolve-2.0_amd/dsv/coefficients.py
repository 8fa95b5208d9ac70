"""Upstream coefficient fields of the transport step on the GPU (libdsv dsv_nodal_gradient,
dsv_transport_coefficients), with the reference's function names:

* `nodal_gradient(mesh, u)` -- GPUSparseAssembler.nodal_gradient (`gpu/gpu_assembler.py:646-671`);
* `compute_norm_grad_phi(phi, normgradphi, gpu_asm)` -- `physics/interface_velocity.py:65-87`;
* `compute_effective_diffusion(phi, F_film, params, DeZn, DeCl, DeOH, DeO2)` --
  `physics/interface_velocity.py:90-155`, with its NaN / negative checks;
* `orr_sink_and_oh_source(...)` -- the O2 sink alpha_O2 and the OH source of the species blocks
  (`dissolve.py:866-875,923-936`).

Fields are torch tensors on the mesh's device, or objects with `.x.array` (DOLFINx Functions,
dsv.fem.Function), whose host arrays are copied in and out like the reference's cupy round trips.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .core import ptr, stream_ptr, to_device


def _dev(mesh, a):
    import torch
    if a is None:
        return None
    if hasattr(a, "x"):
        a = a.x.array
    t = to_device(a, mesh.device, torch.float64)
    if t.numel() != mesh.n_dofs:
        raise ValueError(f"field length {t.numel()} != n_dofs {mesh.n_dofs}")
    return t


def _mesh_of(field, gpu_asm=None):
    if gpu_asm is not None:
        return gpu_asm.mesh
    m = field.function_space.mesh
    return getattr(m, "dsv_mesh", m)


def nodal_gradient(mesh, u, with_norm=False):
    """grad u (n_dofs x 3) of a P1 field, and |grad u| (with the 1e-12 regularisation) when
    with_norm."""
    import torch
    d_u = _dev(mesh, u)
    g = torch.empty((mesh.n_dofs, 3), dtype=torch.float64, device=mesh.device)
    n = torch.empty(mesh.n_dofs, dtype=torch.float64, device=mesh.device) if with_norm else None
    _lib.call("dsv_nodal_gradient", mesh.handle, ptr(d_u), ptr(g), ptr(n), stream_ptr(mesh.device))
    return (g, n) if with_norm else g


def grad_norm(mesh, u, out=None):
    """|grad u| per vertex into `out` (a device tensor, allocated when None)."""
    import torch
    d_u = _dev(mesh, u)
    if out is None:
        out = torch.empty(mesh.n_dofs, dtype=torch.float64, device=mesh.device)
    _lib.call("dsv_nodal_gradient", mesh.handle, ptr(d_u), None, ptr(out), stream_ptr(mesh.device))
    return out


def compute_norm_grad_phi(phi, normgradphi, gpu_asm=None):
    """normgradphi <- |grad phi| at the vertices (reference signature)."""
    mesh = _mesh_of(phi, gpu_asm)
    n = grad_norm(mesh, phi)
    if hasattr(normgradphi, "x"):
        normgradphi.x.array[:] = n.cpu().numpy()
        normgradphi.x.scatter_forward()
    else:
        normgradphi.copy_(n)


def transport_coefficients(mesh, phi, gnorm, eps, k_orr=0.0, kd=0.0, F_film=None, Cl=None,
                           O2=None, Fmax=1.0, film_floor=1.0, outputs=("alpha_o2", "f_oh"),
                           out=None):
    """Dict of the requested per-vertex fields (device tensors): alpha_o2, f_oh, film_factor,
    medium_factor (libdsv dsv_transport_coefficients); `out` may map some of the requested
    names to contiguous FP64 device tensors of length n_dofs to fill instead."""
    import torch
    names = ("alpha_o2", "f_oh", "film_factor", "medium_factor")
    bad = [o for o in outputs if o not in names]
    if bad:
        raise ValueError(f"unknown outputs {bad}")
    given = dict(out or {})
    for o, t in given.items():
        if o not in outputs or not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or \
                t.numel() != mesh.n_dofs or not t.is_contiguous() or t.device != mesh.device:
            raise ValueError(f"out[{o!r}] must be a requested output held in a contiguous FP64 "
                             "device tensor of length n_dofs")
    out = {o: given[o] if o in given else
           torch.empty(mesh.n_dofs, dtype=torch.float64, device=mesh.device) for o in outputs}
    p = _lib.CoefParams(float(eps), float(k_orr), float(kd), float(Fmax), float(film_floor))
    keep = [_dev(mesh, a) for a in (phi, gnorm, F_film, Cl, O2)]
    _lib.call("dsv_transport_coefficients", mesh.handle, *[ptr(t) for t in keep], ctypes.byref(p),
              *[ptr(out.get(n)) for n in names], stream_ptr(mesh.device))
    return out


def orr_sink_and_oh_source(mesh, phi, normgradphi, epsilon, k_orr, kd=0.0, F_film=None, Cl=None,
                           O2=None):
    """(alpha_O2, OH source) per vertex as the species blocks build them."""
    d = transport_coefficients(mesh, phi, normgradphi, epsilon, k_orr, kd, F_film, Cl, O2)
    return d["alpha_o2"], d["f_oh"]


def effective_diffusion(mesh, phi, F_film, Fmax, film_floor, D0, out=None):
    """(DeZn, DeCl, DeOH, DeO2) device tensors (libdsv dsv_effective_diffusion): D0 = (D_Zn0,
    D_Cl0, D_OH0, D_O20), DeZn = D_Zn0 f, the others D_i0 f [phi <= 0], with the film factor
    f = (1 - s) + s film_floor, s = clip(F / Fmax, 0, 1); `out` may give the four tensors."""
    import torch
    out = out if out is not None else tuple(
        torch.empty(mesh.n_dofs, dtype=torch.float64, device=mesh.device) for _ in range(4))
    p = _lib.CoefParams(1.0, 0.0, 0.0, float(Fmax), float(film_floor))
    d0 = (ctypes.c_double * 4)(*[float(x) for x in D0])
    keep = [_dev(mesh, a) for a in (phi, F_film)]
    _lib.call("dsv_effective_diffusion", mesh.handle, ptr(keep[0]), ptr(keep[1]),
              ctypes.byref(p), d0, *[ptr(t) for t in out], stream_ptr(mesh.device))
    return out


def compute_effective_diffusion(phi, F_film, params, DeZn, DeCl, DeOH, DeO2):
    """D_Zn = D_Zn0 f, D_i = D_i0 f [phi <= 0] for Cl, OH, O2 with the film factor
    f = (1 - s) + s eps delta / tau^2, s = clip(F / Fmax, 0, 1) (reference signature and checks;
    the checks run on the host copies)."""
    mesh = _mesh_of(phi)
    floor = params.film_porosity * params.film_constrictivity / (params.film_tortuosity ** 2)
    fields = (DeZn, DeCl, DeOH, DeO2)
    out = effective_diffusion(mesh, phi, F_film, params.Fmax, floor,
                              (params.diff_zn, params.diff_cl, params.diff_oh, params.diff_o2))
    host = [t.cpu().numpy() for t in out]
    for fn, arr in zip(fields, host):
        name = getattr(fn, "name", "De")
        if not np.all(np.isfinite(arr)):
            raise RuntimeError(f"[bounds] {name} contains NaN/Inf after effective diffusion update.")
        if np.any(arr < 0):
            raise RuntimeError(f"[bounds] {name} has {int((arr < 0).sum())} negative values.")
    for fn, arr, t in zip(fields, host, out):
        if hasattr(fn, "x"):
            fn.x.array[:] = arr
            fn.x.scatter_forward()
        else:
            fn.copy_(t)

"""Matrix-free operator with the call surface of the reference's `gpu/ceed_operators.py`.

The reference's libCEED path (`gpu/ceed_operators.py:86-108,247-353`) is not called by its
driver and computes a different operator (consistent mass, no reaction, no Dirichlet rows,
and a diffusion flux without the quadrature weight).  CeedDiffusionMassOp here applies the
operator the reference actually solves -- the assembled one of `gpu/gpu_assembler.py:60-165`
-- matrix-free in libdsv: y = A x with (1/dt + alpha_e) M_e + De_e K_e recomputed per cell
from the coordinates, lumped or consistent mass, Dirichlet identity rows, for one to three
species in one kernel.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .core import Mesh, ptr, species_array, species_struct, stream_ptr, to_device

# degree-2 Keast rule on the unit tet (4 points, weight 1/24 each); basix's ordering of the
# same rule is irrelevant to every quantity computed from it here
_A = (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
_B = (5.0 - np.sqrt(5.0)) / 20.0
Q_PTS = np.array([[_B, _B, _B], [_A, _B, _B], [_B, _A, _B], [_B, _B, _A]])
Q_WTS = np.full(4, 1.0 / 24.0)


def libceed_available():
    return True


def make_ceed(use_gpu: bool = True):
    """Returns (context, backend name) like `gpu/ceed_operators.py:44-61`; the context is the
    device index libdsv runs on."""
    import torch
    if not (use_gpu and torch.cuda.is_available()):
        raise _lib.DsvLibraryError("the matrix-free operator runs on the GPU only")
    _lib.load()
    return torch.cuda.current_device(), "HIP-GFX950"


def extract_ceed_mesh_data(domain, V=None):
    """dict(elem_nodes, coords, num_cells, num_dofs, num_qpts, q_pts, q_wts)
    (`gpu/ceed_operators.py:125-156`) from a dsv Mesh."""
    m = domain if isinstance(domain, Mesh) else domain.dsv_mesh
    return {"elem_nodes": m.conn, "coords": m.xyz, "num_cells": m.n_cells,
            "num_dofs": m.n_dofs, "num_qpts": 4, "q_pts": Q_PTS, "q_wts": Q_WTS, "mesh": m}


def compute_geometric_factors(coords, elem_nodes, q_wts):
    """[wdetJ, G00..G22] per quadrature point, flattened (`gpu/ceed_operators.py:159-199`);
    host numpy like the reference (set-up only, not on the solve path)."""
    p = coords[elem_nodes]
    J = (p[:, 1:, :] - p[:, 0:1, :]).transpose(0, 2, 1)
    detJ = np.abs(np.linalg.det(J))
    Jinv = np.linalg.inv(J)
    G = (Jinv @ Jinv.transpose(0, 2, 1)).reshape(len(elem_nodes), 9)
    nq = len(q_wts)
    geom = np.empty((len(elem_nodes), nq, 10))
    for q in range(nq):
        geom[:, q, 0] = q_wts[q] * detJ
        geom[:, q, 1:] = G
    return geom.reshape(-1)


class CeedDiffusionMassOp:
    """y = ((1/dt + alpha) M + De K) x, matrix-free on the GPU.

    De_vals / alpha: length n_cells or n_dofs (cell-averaged like the assembler) or None;
    bc_dofs: Dirichlet rows (identity); precision 64 or 32 (FP32 geometry and arithmetic)."""

    def __init__(self, ceed, mesh_data, De_vals, inv_dt, num_comp=1, alpha=None, bc_dofs=None,
                 lumped=1, precision=64):
        if num_comp != 1:
            raise ValueError("num_comp must be 1 (batch species with apply_batched)")
        self.ceed = ceed
        self.mesh_data = mesh_data
        self.mesh = mesh_data["mesh"]
        self.alpha = alpha
        self.bc_dofs = bc_dofs
        self.lumped = int(lumped)
        self.precision = int(precision)
        self.update_coeffs(De_vals, inv_dt)

    def update_coeffs(self, De_vals, inv_dt):
        """Refresh coefficients (`gpu/ceed_operators.py:345-353`)."""
        import torch
        n = int(np.shape(De_vals)[0])
        if n not in (self.mesh.n_cells, self.mesh.n_dofs):
            raise ValueError(f"array length {n} matches neither n_cells={self.mesh.n_cells} "
                             f"nor n_dofs={self.mesh.n_dofs}")
        self.De_vals = to_device(De_vals, self.mesh.device, torch.float64)
        self.inv_dt = float(inv_dt)

    def apply(self, u_vec, out_vec=None):
        """out = A u (device tensor); `out_vec`, if given, receives the values too."""
        import torch
        m = self.mesh
        x = to_device(u_vec, m.device, torch.float64)
        y = torch.empty_like(x)
        m.fresh_operator()
        keep = []
        s = species_struct(m, D=self.De_vals, alpha=self.alpha, bc_dofs=self.bc_dofs, keep=keep)
        import ctypes
        xs = (ctypes.c_void_p * 1)(ptr(x).value)
        ys = (ctypes.c_void_p * 1)(ptr(y).value)
        _lib.call("dsv_apply", m.handle, 1, species_array([s]), self.inv_dt, self.lumped, xs, ys,
                  self.precision, stream_ptr(m.device))
        if out_vec is not None:
            if isinstance(out_vec, torch.Tensor):
                out_vec.copy_(y)
            else:
                out_vec[:] = y.cpu().numpy()
        return y


def apply_batched(mesh, species, inv_dt, xs, lumped=1, precision=32):
    """y_s = A_s x_s for up to three species in one fused kernel; species are dicts like
    dsv.solvers.solve_batched's.  Returns a list of device tensors."""
    import ctypes
    import torch
    mesh.fresh_operator()
    keep = []
    structs = [species_struct(mesh, D=sp.get("D"), D_value=sp.get("D_value", 0.0),
                              alpha=sp.get("alpha"), alpha_value=sp.get("alpha_value", 0.0),
                              bc_dofs=sp.get("bc_dofs"), keep=keep) for sp in species]
    xt = [to_device(x, mesh.device, torch.float64) for x in xs]
    yt = [torch.empty_like(x) for x in xt]
    xp = (ctypes.c_void_p * len(xt))(*[ptr(x).value for x in xt])
    yp = (ctypes.c_void_p * len(yt))(*[ptr(y).value for y in yt])
    _lib.call("dsv_apply", mesh.handle, len(structs), species_array(structs), float(inv_dt),
              int(lumped), xp, yp, int(precision), stream_ptr(mesh.device))
    return yt

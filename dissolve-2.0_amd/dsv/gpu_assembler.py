"""GPUSparseAssembler with the call surface of the reference's `gpu/gpu_assembler.py:311-603`.

Assembles A = (1/dt + alpha_e) M_e + De_e K_e for P1 tetrahedra into CSR on the GPU, with the
same pattern (sorted-column union of the 4x4 cell cliques), the same lumped / consistent mass
rule (`LUMP_MASS`, default 1, env `DISSOLVE_LUMP_MASS`, `gpu/gpu_assembler.py:28-37`), the
same cell averaging of per-vertex coefficients (`:452-470`), the same row-only Dirichlet
elimination (`:147-165`) and the same RHS (`:176-242`).  All work runs in libdsv kernels; the
matrix values stay on the device (`d_data`), and `mat` exposes them in PETSc AIJ layout.

`domain` is a :class:`dsv.core.Mesh` (or a (conn, xyz) pair); `V` is accepted for signature
compatibility (a :class:`dsv.fem.FunctionSpace` or None).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from .core import (Mesh, VIEW_CONN, VIEW_DETJ, VIEW_GFLAT, VIEW_INDICES, VIEW_INDPTR,
                   VIEW_JINVT, VIEW_SCATTER_MAP, ptr, species_struct, stream_ptr, to_device)

LUMP_MASS = int(os.environ.get("DISSOLVE_LUMP_MASS", "1"))


class CSRMatrix:
    """PETSc-AIJ-compatible view of an assembled matrix: `getValuesCSR()` returns
    (indptr, indices, values) host arrays like petsc4py's Mat.getValuesCSR."""

    def __init__(self, asm):
        self._asm = asm

    def getSize(self):
        n = self._asm.n_dofs
        return (n, n)

    def getValuesCSR(self):
        a = self._asm
        data = a.d_data  # assembles a deferred operator first
        return (a.d_indptr.cpu().numpy(), a.d_indices.cpu().numpy(), data.cpu().numpy())

    def to_scipy(self):
        import scipy.sparse as sp
        ip, ix, v = self.getValuesCSR()
        return sp.csr_matrix((v, ix, ip), shape=self.getSize())

    def to_petsc(self, comm=None):
        """A petsc4py AIJ Mat with the same values (only when petsc4py is importable)."""
        from petsc4py import PETSc
        ip, ix, v = self.getValuesCSR()
        return PETSc.Mat().createAIJ(size=self.getSize(), csr=(ip, ix, v), comm=comm)


def _as_mesh(domain):
    if isinstance(domain, Mesh):
        return domain
    if hasattr(domain, "dsv_mesh"):
        return domain.dsv_mesh
    if isinstance(domain, (tuple, list)) and len(domain) == 2:
        return Mesh(domain[0], domain[1])
    raise TypeError("domain must be a dsv.core.Mesh, an object with .dsv_mesh, or (conn, xyz)")


class GPUSparseAssembler:
    """Assembles the P1 (M/dt + De K + alpha M) matrix entirely on the GPU.

    Per step:  asm.assemble(De, inv_dt, alpha, bc_dofs); asm.assemble_rhs(u_old, inv_dt, f, ...)
    """

    def __init__(self, domain, V=None, comm=None, ref_asm=None):
        import torch
        self.comm = comm
        if ref_asm is not None:
            # borrow the geometry / pattern of the primary assembler (`gpu/gpu_assembler.py:346-367`)
            self.mesh = ref_asm.mesh
        else:
            self.mesh = _as_mesh(domain)
        m = self.mesh
        self.n_cells = m.n_cells
        self.n_dofs = m.n_dofs
        self._nnz = m.nnz
        self._d_data = torch.zeros(self._nnz, dtype=torch.float64, device=m.device)
        self._pending = None
        self.d_b = torch.zeros(self.n_dofs, dtype=torch.float64, device=m.device)
        self.mat = CSRMatrix(self)
        self.lumped = LUMP_MASS

    @property
    def d_data(self):
        """CSR values (nnz, FP64) of the last assembled operator; an operator deferred by a
        solve (defer_assembly) is assembled here, on first read."""
        if self._pending is not None:
            De, inv_dt, alpha, bc_dofs = self._pending
            self.assemble(De, inv_dt, alpha=alpha, bc_dofs=bc_dofs)
        return self._d_data

    def defer_assembly(self, De, inv_dt: float, alpha=None, bc_dofs=None):
        """Records the coefficients of an operator that a solve used without assembling it
        (the matrix-free and warm-started paths); `d_data` and `mat` assemble it when read."""
        import torch

        def snap(a):
            if a is None:
                return None
            return a.clone() if isinstance(a, torch.Tensor) else np.array(a, copy=True)
        self._pending = (snap(De), float(inv_dt), snap(alpha), snap(bc_dofs))

    # shared read-only device arrays (`gpu/gpu_assembler.py:388-438`)
    @property
    def d_indptr(self):
        return self.mesh.view(VIEW_INDPTR)

    @property
    def d_indices(self):
        return self.mesh.view(VIEW_INDICES)

    @property
    def d_scatter_map(self):
        return self.mesh.view(VIEW_SCATTER_MAP)

    @property
    def d_detJ(self):
        return self.mesh.view(VIEW_DETJ)

    @property
    def d_Gflat(self):
        return self.mesh.view(VIEW_GFLAT)

    @property
    def d_JinvT(self):
        return self.mesh.view(VIEW_JINVT)

    @property
    def d_elem_nodes(self):
        return self.mesh.view(VIEW_CONN)

    def _check_len(self, arr):
        if arr is None:
            return
        n = int(arr.shape[0])
        if n != self.n_cells and n != self.n_dofs:
            raise ValueError(f"array length {n} matches neither n_cells={self.n_cells} "
                             f"nor n_dofs={self.n_dofs}")

    def species(self, De, alpha=None, bc_dofs=None, bc_value=0.0, keep=None):
        self._check_len(De)
        self._check_len(alpha)
        return species_struct(self.mesh, D=De, alpha=alpha, bc_dofs=bc_dofs, bc_value=bc_value,
                              keep=keep)

    def assemble(self, De, inv_dt: float, alpha=None, bc_dofs=None):
        """Fill d_data (and mat) with (1/dt + alpha_e) M_e + De_e K_e, then Dirichlet rows."""
        keep = []
        s = self.species(De, alpha, bc_dofs, keep=keep)
        self._pending = None
        _lib.call("dsv_assemble", self.mesh.handle, s, float(inv_dt), int(self.lumped),
                  ptr(self._d_data), stream_ptr(self.mesh.device))

    def assemble_rhs(self, u_old, inv_dt: float, f=None, bc_dofs=None, bc_value: float = 0.0):
        """b = M u_old/dt + M f on the GPU; b[bc_dofs] = bc_value.  Returns d_b (torch)."""
        import torch
        if int(u_old.shape[0]) != self.n_dofs:
            raise ValueError(f"u_old length {u_old.shape[0]} != n_dofs {self.n_dofs}")
        dev = self.mesh.device
        d_u = to_device(u_old, dev, torch.float64)
        d_f = None
        if f is not None:
            n = int(f.shape[0])
            if n != self.n_dofs and n != self.n_cells:
                raise ValueError(f"f length {f.shape[0]} unrecognised")
            # length V: per vertex; length C: the reference's per-cell branch
            # (`gpu/gpu_assembler.py:560-571`), mapped to vertices inside dsv_rhs
            d_f = to_device(f, dev, torch.float64)
        bc = None
        if bc_dofs is not None and len(bc_dofs) > 0:
            bc = to_device(bc_dofs, dev, torch.int32)
        _lib.call("dsv_rhs", self.mesh.handle, ptr(d_u), ptr(d_f),
                  int(d_f.numel()) if d_f is not None else 0, float(inv_dt), int(self.lumped),
                  ptr(bc), int(bc.numel()) if bc is not None else 0, float(bc_value),
                  ptr(self.d_b), stream_ptr(dev))
        return self.d_b

    def nodal_gradient(self, u_vert):
        """Per-vertex grad u (torch, n_dofs x 3) of a P1 field: the cells' constant gradients
        averaged with weights |det J| / 6 (`gpu/gpu_assembler.py:646-671`), on the GPU."""
        from .coefficients import nodal_gradient
        if hasattr(u_vert, "x"):
            u_vert = u_vert.x.array
        self._check_len(u_vert)
        return nodal_gradient(self.mesh, u_vert)

    def scaffold_volume(self, phi_arr, epsilon):
        """Integral of the smoothed Heaviside H_eps(phi) at the 4-point degree-2 rule, on the GPU
        (`gpu/gpu_assembler.py:604-643`); a Python float."""
        from .physics import scaffold_volume
        if hasattr(phi_arr, "x"):
            phi_arr = phi_arr.x.array
        return scaffold_volume(self.mesh, phi_arr, float(epsilon))

"""Batched backward-Euler stepper for up to three species on one mesh (device resident).

One step, for every species s: b_s = M (u_s / dt + f_s) with b_s[bc] = bc value
(`gpu/gpu_assembler.py:534-603`), then A_s x_s = b_s solved for all species in the same
kernels by libdsv (`numerics/solvers.py:287-372`), warm-started from the previous solution
(`numerics/solvers.py:329,366-367`) or from a predictor.  Method "csr64" re-assembles the
FP64 CSR of every species each step (the reference's operator path); "mf32_ir64" applies the
operator matrix-free in FP32 inside FP64 refinement.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .core import Mesh, ptr, species_array, species_struct, stream_ptr, to_device


class Stepper:
    def __init__(self, mesh: Mesh, species, dt, method="mf32_ir64", rtol=1e-8, max_it=20000,
                 lumped=1):
        import torch
        self.mesh = mesh
        self.dt = float(dt)
        self.method = method
        self.rtol = float(rtol)
        self.max_it = int(max_it)
        self.lumped = int(lumped)
        dev = mesh.device
        self.S = len(species)
        self.names = [s.get("name", str(i)) for i, s in enumerate(species)]
        self._keep = []
        self.defs = species
        self.u = [to_device(s["u0"], dev, torch.float64).clone() for s in species]
        self.u_prev = [u.clone() for u in self.u]
        self.b = [torch.zeros_like(u) for u in self.u]
        self.x = [torch.zeros_like(u) for u in self.u]
        self.f = [to_device(s["f"], dev, torch.float64) if s.get("f") is not None else None
                  for s in species]
        self.bc = [to_device(s["bc_dofs"], dev, torch.int32) if s.get("bc_dofs") is not None
                   and len(s["bc_dofs"]) > 0 else None for s in species]
        self._structs = None
        self.set_coefficients(species)
        self.history = []

    def set_coefficients(self, species):
        """(Re)bind D / alpha / f / bc of every species (device copies)."""
        keep = []
        structs = []
        for k, s in enumerate(species):
            st = species_struct(self.mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                                bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                                b=self.b[k], x=self.x[k], keep=keep)
            structs.append(st)
            if "f" in s:
                import torch
                self.f[k] = (to_device(s["f"], self.mesh.device, torch.float64)
                             if s["f"] is not None else None)
        self._keep = keep
        self._arr = species_array(structs)
        self.defs = species

    def rhs(self):
        m = self.mesh
        inv_dt = 1.0 / self.dt
        for k, s in enumerate(self.defs):
            f = self.f[k]
            bc = self.bc[k]
            _lib.call("dsv_rhs", m.handle, ptr(self.u[k]), ptr(f), int(f.numel()) if f is not None else 0,
                      inv_dt, self.lumped, ptr(bc), int(bc.numel()) if bc is not None else 0,
                      float(s.get("bc_value", 0.0)), ptr(self.b[k]), stream_ptr(m.device))

    def step(self, warm=None):
        """One backward-Euler step of every species; `warm` (list of tensors) replaces the
        previous-solution initial guess.  Returns the solver statistics."""
        t0 = time.perf_counter()
        self.rhs()
        for k in range(self.S):
            self.x[k].copy_(self.u[k] if warm is None else warm[k])
        st = _lib.Stats()
        _lib.call("dsv_solve", self.mesh.handle, self.S, self._arr, 1.0 / self.dt, self.lumped,
                  _lib.METHODS[self.method], self.rtol, self.max_it, stream_ptr(self.mesh.device), st)
        for k in range(self.S):
            self.u_prev[k].copy_(self.u[k])
            self.u[k].copy_(self.x[k])
        d = st.as_dict(self.S)
        d["step_s"] = time.perf_counter() - t0
        self.history.append(d)
        return d

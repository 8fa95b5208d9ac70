"""Batched backward-Euler stepper for up to three species on one mesh (device resident).

One step, for every species s: b_s = M (u_s / dt + f_s) with b_s[bc] = bc value
(`gpu/gpu_assembler.py:534-603`), then A_s x_s = b_s solved for all species in the same
kernels by libdsv (`numerics/solvers.py:287-372`), warm-started from the previous solution
(`numerics/solvers.py:329,366-367`) or from the learned predictor (dsv.mlp).  Method "csr64"
re-assembles the FP64 CSR of every species each step (the reference's operator path);
"mf32_ir64" applies the operator in FP32 inside FP64 refinement.

Learned warm start: `train_warm_start(n)` runs n steps from the previous-solution start while
collecting (features at step n, increment to step n+1) pairs, fits the MLP on them, and from
then on every step starts from u_n + sigma f(X) for every species whose initial FP64 residual
is smaller than the previous solution's (the previous solution stays the fallback and the
judge); `warm_log` records the choice and both residuals per step.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .core import Mesh, ptr, species_array, species_struct, stream_ptr, to_device


class Stepper:
    def __init__(self, mesh: Mesh, species, dt, method="mf32_ir64", rtol=1e-8, max_it=20000,
                 lumped=1, phi=None, grad_phi=None, eps=None, mlp=None):
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
        # learned warm start
        self.mlp = mlp
        self.phi = to_device(phi, dev, torch.float64) if phi is not None else None
        self.grad_phi = to_device(grad_phi, dev, torch.float64) if grad_phi is not None else None
        self.eps = float(eps) if eps is not None else None
        self.bc_mask = []
        for k, s in enumerate(species):
            mask = None
            if self.bc[k] is not None:
                mask = torch.zeros(self.u[k].numel(), dtype=torch.uint8, device=dev)
                mask[self.bc[k].long()] = 1
            self.bc_mask.append(mask)
        self.use_mlp = False
        self.collect = False
        self.pairs = []
        self.warm_log = []
        self.n_steps = 0

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

    def _features(self, k):
        s = self.defs[k]
        al = s.get("alpha")
        alpha = to_device(al, self.mesh.device, self.u[k].dtype) if al is not None else None
        return self.mlp.features(self.u[k], self.u_prev[k], self.phi, self.grad_phi,
                                 s.get("D_value", 1.0) if s.get("D") is None else float(np.mean(s["D"])),
                                 self.dt, self.eps, alpha=alpha,
                                 alpha_value=s.get("alpha_value", 0.0), bc_mask=self.bc_mask[k])

    def residual_norms(self):
        """||b - A x|| per species for the current initial guesses self.x (FP64)."""
        rn = (ctypes.c_double * self.S)()
        _lib.call("dsv_residual_norms", self.mesh.handle, self.S, self._arr, 1.0 / self.dt,
                  self.lumped, rn, None, stream_ptr(self.mesh.device))
        return [rn[k] for k in range(self.S)]

    def step(self, warm=None):
        """One backward-Euler step of every species; `warm` (list of tensors) replaces the
        initial guess.  Returns the solver statistics."""
        t0 = time.perf_counter()
        self.rhs()
        feats = None
        want = self.mlp is not None and (self.collect or self.use_mlp) and self.n_steps >= 1
        if want:
            feats = [self._features(k) for k in range(self.S)]
        if warm is not None:
            for k in range(self.S):
                self.x[k].copy_(warm[k])
        elif self.use_mlp and feats is not None:
            cand = [self.u[k] + feats[k][1].double() * self.mlp.forward(feats[k][0]).double()
                    for k in range(self.S)]
            for k in range(self.S):
                self.x[k].copy_(cand[k])
            r_mlp = self.residual_norms()
            for k in range(self.S):
                self.x[k].copy_(self.u[k])
            r_prev = self.residual_norms()
            use = [r_mlp[k] < r_prev[k] for k in range(self.S)]
            for k in range(self.S):
                if use[k]:
                    self.x[k].copy_(cand[k])
            self.warm_log.append({"step": self.n_steps, "mlp": use, "r_mlp": r_mlp,
                                  "r_prev": r_prev})
        else:
            for k in range(self.S):
                self.x[k].copy_(self.u[k])
        st = _lib.Stats()
        _lib.call("dsv_solve", self.mesh.handle, self.S, self._arr, 1.0 / self.dt, self.lumped,
                  _lib.METHODS[self.method], self.rtol, self.max_it, stream_ptr(self.mesh.device), st)
        if self.collect and feats is not None:
            for k in range(self.S):
                X, sigma = feats[k]
                y = ((self.x[k] - self.u[k]) / sigma.double()).float()
                self.pairs.append((X, y))
        for k in range(self.S):
            self.u_prev[k].copy_(self.u[k])
            self.u[k].copy_(self.x[k])
        self.n_steps += 1
        d = st.as_dict(self.S)
        d["step_s"] = time.perf_counter() - t0
        self.history.append(d)
        return d

    def train_warm_start(self, n_steps, epochs=20, batch=16384):
        """Run n_steps steps from the previous-solution start collecting training pairs, fit
        the MLP on them, and switch to the learned warm start.  Returns the loss per epoch."""
        import torch
        if self.mlp is None or self.phi is None or self.grad_phi is None or self.eps is None:
            raise ValueError("the learned warm start needs mlp, phi, grad_phi and eps")
        self.collect = True
        for _ in range(n_steps):
            self.step()
        self.collect = False
        if not self.pairs:
            raise ValueError("no training pairs (needs at least 2 steps)")
        X = torch.cat([p[0] for p in self.pairs])
        y = torch.cat([p[1] for p in self.pairs])
        hist = self.mlp.fit(X, y, epochs=epochs, batch=batch)
        self.pairs = []
        self.use_mlp = True
        return hist

"""Batched backward-Euler stepper for up to three species on one mesh (device resident).

One step, for every species s: b_s = M (u_s / dt + f_s) with b_s[bc] = bc value
(`gpu/gpu_assembler.py:534-603`), then A_s x_s = b_s solved for all species in the same
kernels by libdsv (`numerics/solvers.py:287-372`), warm-started from the previous solution
(`numerics/solvers.py:329,366-367`) or from a Galerkin warm start.  Method "csr64"
re-assembles the FP64 CSR of every species each step (the reference's operator path);
"mf32_ir64" applies the operator in FP32 inside FP64 refinement.

Warm starts (no counterpart in the reference, whose solves start from the previous
solution): "previous" starts every solve from u_n; "extrapolate" and "learned" start from the
Galerkin warm start of libdsv over u_n + span{directions}, folded into the solve
(dsv_solve_warm_history, which reads u_n and the past solutions in place), per species --
the combination minimising the A-energy norm of the error, taken only where its FP64
residual is below that of u_n (the previous solution stays the fallback and the yardstick).
"extrapolate" uses the last history_dirs (default 8, at most 8) increments u_n - u_{n-1},
u_{n-1} - u_{n-2}, ...; "learned" uses at most seven of them plus the per-vertex MLP's
prediction of the part of the next increment that extrapolation misses,
(u_{n+1} - 2 u_n + u_{n-1}) / (sigma tau), sigma = max |u_n - u_{n-1}| and tau the training
targets' RMS.  `train_warm_start(n)` runs n steps collecting (features, target) pairs, fits
the MLP on the GPU and switches to "learned"; `warm_log` records per step and species the
coefficients and both residual norms.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from .core import Mesh, ptr, species_array, species_struct, stream_ptr, to_device
from .mlp import NIN


class Stepper:
    MAX_HIST = 8  # past solutions kept (libdsv dsv_solve_warm takes up to 8 directions)

    def __init__(self, mesh: Mesh, species, dt, method="mf32_ir64", rtol=1e-8, max_it=20000,
                 lumped=1, phi=None, grad_phi=None, eps=None, mlp=None, reuse_operator=True):
        import torch
        self.mesh = mesh
        self.reuse_operator = bool(reuse_operator)
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
        # u_{n-1}, u_{n-2}, ..., u_{n-MAX_HIST}
        self.u_hist = [[u.clone() for u in self.u] for _ in range(self.MAX_HIST)]
        self.b = [torch.zeros_like(u) for u in self.u]
        self.x = [torch.zeros_like(u) for u in self.u]
        # device addresses of the rotating buffers, kept beside them so that a step builds its
        # pointer arrays without asking torch again
        self._p_u = [t.data_ptr() for t in self.u]
        self._p_hist = [[t.data_ptr() for t in h] for h in self.u_hist]
        self._p_x = [t.data_ptr() for t in self.x]
        self._p_b = (ctypes.c_void_p * self.S)(*[t.data_ptr() for t in self.b])
        self.f = [to_device(s["f"], dev, torch.float64).clone() if s.get("f") is not None else None
                  for s in species]
        self.bc = [to_device(s["bc_dofs"], dev, torch.int32) if s.get("bc_dofs") is not None
                   and len(s["bc_dofs"]) > 0 else None for s in species]
        self._structs = None
        self.set_coefficients(species)
        self.history = []
        # learned warm start
        self.mlp = mlp
        self.phi = to_device(phi, dev, torch.float64).clone() if phi is not None else None
        self.grad_phi = (to_device(grad_phi, dev, torch.float64).clone() if grad_phi is not None
                         else None)
        self.eps = float(eps) if eps is not None else None
        self.bc_mask = []
        for k, s in enumerate(species):
            mask = None
            if self.bc[k] is not None:
                mask = torch.zeros(self.u[k].numel(), dtype=torch.uint8, device=dev)
                mask[self.bc[k].long()] = 1
            self.bc_mask.append(mask)
        self.warm = "previous"
        self.history_dirs = self.MAX_HIST
        # the Galerkin solves read the start from the history arrays (dsv_solve_warm_history):
        # no copy of u_n, and every past direction's A d comes from the previous step
        self.start_from_history = True
        self.fuse_rhs = True  # dsv_step_warm_history for the Galerkin steps from the history
        self.collect = False
        self.pairs = []
        self.tau = 1.0
        self._warm_raw = []   # per Galerkin step: (step, mode, nd, coef, r0, r1) ctypes arrays
        self._warm_log = []
        self.n_steps = 0

    @property
    def warm_log(self):
        """Per Galerkin-started step: {"step", "mode", "coef" [direction][species],
        "r_previous" (residual norm of u_n), "r_start" (of the start taken), "taken"}; built from
        the library's outputs when read."""
        S = self.S
        for n, mode, nd, coef, r0, r1 in self._warm_raw[len(self._warm_log):]:
            self._warm_log.append({
                "step": n, "mode": mode,
                "coef": [[coef[k * S + s] for s in range(S)] for k in range(nd)],
                "r_previous": list(r0), "r_start": list(r1),
                "taken": [r1[k] < r0[k] for k in range(S)]})
        return self._warm_log

    @property
    def u_prev(self):
        """The previous step's solutions u_{n-1}."""
        return self.u_hist[0]

    def set_warm_start(self, mode, history_dirs=None):
        """"previous", "extrapolate" or "learned" (needs a trained MLP); history_dirs
        (1..MAX_HIST) is the number of past increments u_n - u_{n-1}, u_{n-1} - u_{n-2}, ...
        among the directions (at most MAX_HIST - 1 are used beside the learned direction)."""
        if mode not in ("previous", "extrapolate", "learned"):
            raise ValueError(f"unknown warm start {mode!r}")
        if mode != "previous" and self.method == "csr64":
            raise ValueError("the Galerkin warm starts need a matrix-free method (mf32_ir64, mf64)")
        if mode == "learned" and (self.mlp is None or not self.mlp.trained):
            raise ValueError("the learned warm start needs a trained MLP (train_warm_start)")
        if history_dirs is not None:
            if not 1 <= int(history_dirs) <= self.MAX_HIST:
                raise ValueError(f"history_dirs must be in 1..{self.MAX_HIST}")
            self.history_dirs = int(history_dirs)
        self.warm = mode

    def set_coefficients(self, species):
        """(Re)bind D / alpha / f / bc of every species (device copies)."""
        import torch
        keep = []
        structs = []
        self.alpha_dev = []
        for k, s in enumerate(species):
            al = s.get("alpha")
            al = to_device(al, self.mesh.device, torch.float64).clone() if al is not None else None
            self.alpha_dev.append(al)
            st = species_struct(self.mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                alpha=al, alpha_value=s.get("alpha_value", 0.0),
                                bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                                b=self.b[k], x=self.x[k], keep=keep)
            structs.append(st)
            if "f" in s:
                self.f[k] = (self._vertex_field(s["f"]).clone() if s["f"] is not None else None)
        self._keep = keep
        self._arr = species_array(structs)
        self.defs = species
        # the coefficient arrays above stay unchanged until the next call: the library may
        # keep the operator it builds from them (this call discards any earlier one)
        self.mesh.set_operator_reuse(self.reuse_operator)

    def _vertex_field(self, a):
        """A per-vertex FP64 source on the device; other lengths raise ValueError (the
        batched RHS reads V values of each)."""
        import torch
        t = to_device(a, self.mesh.device, torch.float64)
        if t.numel() != self.mesh.n_dofs:
            raise ValueError(f"f length {t.numel()} unrecognised (expected n_dofs={self.mesh.n_dofs})")
        return t

    def refresh_coefficients(self, alpha=None, f=None):
        """Overwrite, in place, the per-vertex reaction coefficients (alpha[k]) and sources
        (f[k]) of the species (lists with None for the ones left unchanged; a species' alpha must
        already be a per-vertex array), and discard the operator the library built from the old
        values."""
        import torch
        dev = self.mesh.device
        for k in range(self.S):
            if alpha is not None and alpha[k] is not None:
                if self.alpha_dev[k] is None:
                    raise ValueError(f"species {self.names[k]} has no per-vertex alpha to refresh")
                self.alpha_dev[k].copy_(to_device(alpha[k], dev, torch.float64))
            if f is not None and f[k] is not None:
                fk = self._vertex_field(f[k])
                if self.f[k] is None or self.f[k].numel() != fk.numel():
                    self.f[k] = fk.clone()
                else:
                    self.f[k].copy_(fk)
        self.mesh.set_operator_reuse(self.reuse_operator)

    def rhs(self, stream=None):
        """b_s = M (u_s / dt + f_s), b_s[bc] = bc value, all species in one call."""
        P = ctypes.c_void_p * self.S
        _lib.call("dsv_rhs_batch", self.mesh.handle, self.S, self._arr, P(*self._p_u),
                  P(*[f.data_ptr() if f is not None else None for f in self.f]),
                  1.0 / self.dt, self.lumped, self._p_b,
                  stream if stream is not None else stream_ptr(self.mesh.device))

    def _features(self):
        """The MLP features of every species in one (S V, 7) buffer, so that one forward
        serves them all: (X, sigma [S], [(X_k, sigma_k) per species])."""
        import torch
        V = self.u[0].numel()
        X = torch.empty((self.S * V, NIN), dtype=torch.float32, device=self.mesh.device)
        sig = torch.empty(self.S, dtype=torch.float32, device=self.mesh.device)
        per = [(X[k * V:(k + 1) * V], sig[k:k + 1]) for k in range(self.S)]
        for k, s in enumerate(self.defs):
            D = s.get("D_value", 1.0) if s.get("D") is None else float(np.mean(s["D"]))
            self.mlp.features(self.u[k], self.u_hist[0][k], self.phi, self.grad_phi, D, self.dt,
                              self.eps, alpha=self.alpha_dev[k],
                              alpha_value=s.get("alpha_value", 0.0), bc_mask=self.bc_mask[k],
                              X=per[k][0], sigma=per[k][1])
        return X, sig, per

    def build_low_modes(self, n_modes=32, degree=80, passes=4, seed=0):
        """Low-mode subspace of the current operator (libdsv dsv_build_low_modes): n_modes
        vectors per species approximating the eigenvectors of the lowest eigenvalues of D^-1 A,
        which every later Galerkin-started step applies as a correction of its start while the
        coefficients stay bound (set_coefficients / refresh_coefficients drop it).  Needs the
        mf32_ir64 method and operator reuse.  Returns {"seconds", "ritz" [species][mode]}."""
        import torch
        if self.method != "mf32_ir64" or not self.reuse_operator:
            raise ValueError("the low-mode subspace needs method mf32_ir64 with operator reuse")
        ritz = (ctypes.c_double * max(1, self.S * int(n_modes)))()
        torch.cuda.synchronize(self.mesh.device)
        t = time.perf_counter()
        _lib.call("dsv_build_low_modes", self.mesh.handle, self.S, self._arr, 1.0 / self.dt,
                  self.lumped, int(n_modes), int(degree), int(passes), int(seed), ritz,
                  stream_ptr(self.mesh.device))
        torch.cuda.synchronize(self.mesh.device)
        return {"seconds": time.perf_counter() - t,
                "ritz": [[ritz[s * n_modes + k] for k in range(n_modes)] for s in range(self.S)]}

    def residual_norms(self):
        """||b - A x|| per species for the current initial guesses self.x (FP64)."""
        rn = (ctypes.c_double * self.S)()
        _lib.call("dsv_residual_norms", self.mesh.handle, self.S, self._arr, 1.0 / self.dt,
                  self.lumped, rn, None, stream_ptr(self.mesh.device))
        return [rn[k] for k in range(self.S)]

    def galerkin_start(self, dirs):
        """Replaces self.x (the start points) by the Galerkin warm start over
        self.x + span{dirs[k][s]} where that lowers the FP64 residual; returns
        (coefficients [k][s], residual norms before, after)."""
        nd = len(dirs)
        keep = [d for dk in dirs for d in dk]
        arr = (ctypes.c_void_p * (nd * self.S))(*[ptr(d).value for d in keep])
        coef = (ctypes.c_double * (nd * self.S))()
        r0 = (ctypes.c_double * self.S)()
        r1 = (ctypes.c_double * self.S)()
        _lib.call("dsv_warm_start", self.mesh.handle, self.S, self._arr, 1.0 / self.dt,
                  self.lumped, nd, arr, coef, r0, r1, stream_ptr(self.mesh.device))
        return ([[coef[k * self.S + s] for s in range(self.S)] for k in range(nd)],
                list(r0), list(r1))

    def _mlp_directions(self, X, sig):
        """sigma_k tau f(X_k) per species (FP64 device vectors, libdsv dsv_mlp_directions), from
        one forward over all species."""
        import torch
        V = self.u[0].numel()
        if getattr(self, "_dirs", None) is None:
            self._dirs = [torch.empty(V, dtype=torch.float64, device=self.mesh.device)
                          for _ in range(self.S)]
        out = self.mlp.forward(X)
        P = ctypes.c_void_p * self.S
        _lib.call("dsv_mlp_directions", self.S, V, ptr(out), ptr(sig), float(self.tau),
                  P(*[d.data_ptr() for d in self._dirs]), stream_ptr(self.mesh.device))
        return self._dirs

    def step(self, warm=None):
        """One backward-Euler step of every species; `warm` (list of tensors) replaces the
        initial guess.  Returns the solver statistics."""
        t0 = time.perf_counter()
        stream = stream_ptr(self.mesh.device)
        galerkin = warm is None and self.warm != "previous" and self.n_steps >= 1
        # the Galerkin step from the history forms its right-hand sides inside the solve call
        # (dsv_step_warm_history: the same bits as dsv_rhs_batch, one call and two kernels less)
        fused = galerkin and self.start_from_history and self.lumped and self.fuse_rhs
        if not fused:
            self.rhs(stream)
        feats = None
        learned = self.warm == "learned" and self.n_steps >= 1
        if self.mlp is not None and (self.collect or learned) and self.n_steps >= 1:
            X_all, sig_all, feats = self._features()
        if not (galerkin and self.start_from_history):
            for k in range(self.S):
                self.x[k].copy_(self.u[k] if warm is None else warm[k])
        st = _lib.Stats()
        if galerkin:
            # Galerkin warm start over past increments (and the MLP direction), folded into
            # the solve: the directions are formed from the stored past solutions in libdsv
            nh = min(self.history_dirs, self.MAX_HIST - (1 if learned else 0), self.n_steps)
            extra = list(self._mlp_directions(X_all, sig_all)) if learned else []
            nd = nh + (1 if learned else 0)
            if self.start_from_history:  # u_n, u_{n-1}, ..., u_{n-nh}
                steps, fn = [self._p_u] + self._p_hist[:nh], "dsv_solve_warm_history"
            else:  # u_{n-1}, ..., u_{n-nh}; the start u_n was copied into x
                steps, fn = self._p_hist[:nh], "dsv_solve_warm"
            hist = (ctypes.c_void_p * max(1, len(steps) * self.S))(
                *[p for a in steps for p in a])
            ext = (ctypes.c_void_p * max(1, len(extra)))(*[e.data_ptr() for e in extra])
            coef = (ctypes.c_double * (nd * self.S))()
            r0 = (ctypes.c_double * self.S)()
            r1 = (ctypes.c_double * self.S)()
            head = (self.mesh.handle, self.S, self._arr)
            if fused:
                fn = "dsv_step_warm_history"
                head += ((ctypes.c_void_p * self.S)(
                    *[f.data_ptr() if f is not None else None for f in self.f]),)
            _lib.call(fn, *head, 1.0 / self.dt,
                      self.lumped, _lib.METHODS[self.method], self.rtol, self.max_it, nh, hist,
                      1 if learned else 0, ext, stream, st, coef, r0, r1)
            self._warm_raw.append((self.n_steps, self.warm, nd, coef, r0, r1))
        else:
            _lib.call("dsv_solve", self.mesh.handle, self.S, self._arr, 1.0 / self.dt,
                      self.lumped, _lib.METHODS[self.method], self.rtol, self.max_it,
                      stream, st)
        if self.collect and feats is not None:
            for k in range(self.S):
                X, sigma = feats[k]
                # sigma = max |u_n - u_{n-1}| is 0 for a species left unchanged by its previous
                # solve: such a pair has no target scale and is dropped when the pairs are fitted
                y = ((self.x[k] - 2.0 * self.u[k] + self.u_hist[0][k]) /
                     sigma.double().clamp_min(1e-30)).float()
                self.pairs.append((X, y, sigma))
        # rotate the buffers instead of copying them: the solution becomes u, u becomes
        # u_{n-1}, and the oldest history buffer holds the next step's start
        oldest, p_oldest = self.u_hist[-1], self._p_hist[-1]
        self.u_hist = [self.u] + self.u_hist[:-1]
        self._p_hist = [self._p_u] + self._p_hist[:-1]
        self.u, self.x = self.x, oldest
        self._p_u, self._p_x = self._p_x, p_oldest
        for k in range(self.S):
            self._arr[k].x = self._p_x[k]
        self.n_steps += 1
        d = st.as_dict(self.S)
        d["step_s"] = time.perf_counter() - t0
        self.history.append(d)
        return d

    def fp64_reference(self, rtol=1e-13, max_it=100000):
        """The last step's systems solved again by the plain FP64 path (assembled FP64 CSR,
        FP64 Jacobi-PCG) to `rtol`, starting from the step's solutions: a yardstick for the
        solutions of the faster paths (their error is max |u - ref| / max |ref| per species).
        Returns (list of FP64 device tensors, solver statistics)."""
        import torch
        if self.n_steps == 0:
            raise ValueError("no step taken yet")
        keep = []
        refs, structs = [], []
        for k, s in enumerate(self.defs):
            x = self.u[k].clone()
            b = self.b[k].clone()
            keep += [x, b]
            refs.append(x)
            structs.append(species_struct(self.mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                          alpha=self.alpha_dev[k],
                                          alpha_value=s.get("alpha_value", 0.0),
                                          bc_dofs=self.bc[k], bc_value=s.get("bc_value", 0.0),
                                          b=b, x=x, keep=keep))
        st = _lib.Stats()
        _lib.call("dsv_solve", self.mesh.handle, self.S, species_array(structs), 1.0 / self.dt,
                  self.lumped, _lib.METHODS["csr64"], float(rtol), int(max_it),
                  stream_ptr(self.mesh.device), st)
        torch.cuda.synchronize(self.mesh.device)
        return refs, st.as_dict(self.S)

    def step_error(self, refs):
        """max |u_s - ref_s| / max |ref_s| per species for the last step's solutions."""
        return [float((u - r).abs().max() / r.abs().max().clamp_min(1e-300))
                for u, r in zip(self.u, refs)]

    def train_warm_start(self, n_steps, epochs=20, batch=16384, before_step=None):
        """Run n_steps steps (with the current warm start; `before_step()` is called before each
        one, e.g. to move the interface) collecting training pairs, fit the MLP on them, and
        switch to the learned warm start.  Returns the loss per epoch (relative to the targets'
        mean square, which is 1 after scaling by tau)."""
        import torch
        if self.mlp is None or self.phi is None or self.grad_phi is None or self.eps is None:
            raise ValueError("the learned warm start needs mlp, phi, grad_phi and eps")
        self.collect = True
        for _ in range(n_steps):
            if before_step is not None:
                before_step()
            self.step()
        self.collect = False
        if not self.pairs:
            raise ValueError("no training pairs (needs at least 2 steps)")
        sig = torch.cat([p[2] for p in self.pairs]).cpu()
        use = [p for p, sg in zip(self.pairs, sig.tolist())
               if sg > 0.0 and np.isfinite(sg) and bool(torch.isfinite(p[1]).all())]
        self.pairs = []
        if not use:
            raise ValueError("no usable training pairs (every species unchanged or non-finite)")
        X = torch.cat([p[0] for p in use])
        y = torch.cat([p[1] for p in use])
        self.tau = max(float(torch.sqrt(torch.mean(y.double() ** 2))), 1e-30)
        hist = self.mlp.fit(X, (y / self.tau).float(), epochs=epochs, batch=batch)
        if not (np.isfinite(self.tau) and np.all(np.isfinite(hist))):
            self.mlp.trained = False
            raise ValueError(f"the warm-start fit diverged (tau {self.tau}, loss {hist[-1]})")
        self.warm = "learned"
        return hist

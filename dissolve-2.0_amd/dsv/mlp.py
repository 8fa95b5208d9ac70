"""Learned CG warm-start direction: a per-vertex MLP (7 -> 64 -> 64 -> 1, tanh) trained on the
GPU by libdsv's forward, gradient and Adam kernels on time-step pairs the solver itself
produces.

Features per vertex (libdsv dsv_mlp_features): u_n, (u_n - u_{n-1})/sigma, clamp(phi/eps)/4,
min(|grad phi|, 4)/4, log(D)/2, alpha dt, Dirichlet flag, with sigma = max |u_n - u_{n-1}|.
dsv.transport.Stepper trains it on (u_{n+1} - 2 u_n + u_{n-1}) / (sigma tau) -- the part of the
next increment that linear extrapolation misses, tau the targets' RMS -- and uses
sigma tau f(X) as one direction of the Galerkin warm start, whose result is taken only when it
lowers the FP64 residual of the previous-solution start (the reference's,
`numerics/solvers.py:329,366-367`); every solve still converges to the same tolerance.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .core import ptr, stream_ptr, to_device

NIN, NH = 7, 64
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1  # 4737, DSV_MLP_NPARAM


def glorot_params(seed=0):
    """Glorot-uniform weights, zero biases, packed W1, b1, W2, b2, w3, b3 (host float32)."""
    rng = np.random.default_rng(seed)

    def glorot(n_out, n_in):
        a = np.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, (n_out, n_in))

    p = np.concatenate([glorot(NH, NIN).ravel(), np.zeros(NH), glorot(NH, NH).ravel(),
                        np.zeros(NH), glorot(1, NH)[0], [0.0]])
    return p.astype(np.float32)


class WarmStartMLP:
    def __init__(self, device, seed=0, lr=2e-3):
        import torch
        self.device = device
        self.params = torch.as_tensor(glorot_params(seed), device=device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0
        self.lr = float(lr)
        self.trained = False
        self._ws = None

    # ---- kernels ----
    def features(self, u, u_prev, phi, grad_phi, D, dt, eps, alpha=None, alpha_value=0.0,
                 bc_mask=None, X=None, sigma=None):
        """(X [n, 7] float32, sigma [1] float32) on the device; X and sigma may be given
        (contiguous views, e.g. one species' rows of a batch)."""
        import torch
        n = int(u.numel())
        if X is None:
            X = torch.empty((n, NIN), dtype=torch.float32, device=self.device)
        if sigma is None:
            sigma = torch.empty(1, dtype=torch.float32, device=self.device)
        if tuple(X.shape) != (n, NIN) or not X.is_contiguous() or sigma.numel() != 1:
            raise ValueError("features: X must be a contiguous (n, 7) and sigma a 1-element tensor")
        _lib.call("dsv_mlp_features", n, ptr(u), ptr(u_prev), ptr(phi), ptr(grad_phi), ptr(alpha),
                  float(alpha_value), float(D), float(dt), float(eps), ptr(bc_mask), ptr(X),
                  ptr(sigma), stream_ptr(self.device))
        return X, sigma

    def forward(self, X):
        import torch
        out = torch.empty(X.shape[0], dtype=torch.float32, device=self.device)
        _lib.call("dsv_mlp_forward", ptr(self.params), ptr(X), int(X.shape[0]), ptr(out),
                  stream_ptr(self.device))
        return out

    def grad(self, X, y):
        """(dL/dparams, L) of L = mean (f(X) - y)^2."""
        import torch
        n = int(X.shape[0])
        need = int(_lib.load().dsv_mlp_grad_workspace(n))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        g = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        _lib.call("dsv_mlp_grad", ptr(self.params), ptr(X), ptr(y), n, ptr(g), ptr(loss),
                  ptr(self._ws), ctypes.c_size_t(self._ws.numel()), stream_ptr(self.device))
        return g, loss

    def adam(self, g, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        _lib.call("dsv_mlp_adam", ptr(self.params), ptr(g), ptr(self.m), ptr(self.v),
                  self.step_count, ctypes.c_float(self.lr), ctypes.c_float(b1), ctypes.c_float(b2),
                  ctypes.c_float(eps), stream_ptr(self.device))

    # ---- training ----
    def fit(self, X, y, epochs=20, batch=16384, seed=0):
        """Mini-batch Adam over (X, y) (device tensors); returns the loss after each epoch."""
        import torch
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        n = int(X.shape[0])
        hist = []
        for _ in range(epochs):
            order = torch.randperm(n, device=self.device, generator=gen)
            for k in range(0, n, batch):
                idx = order[k:k + batch]
                g, _ = self.grad(X[idx].contiguous(), y[idx].contiguous())
                self.adam(g)
            _, loss = self.grad(X, y)
            hist.append(float(loss.item()))
        self.trained = True
        return hist

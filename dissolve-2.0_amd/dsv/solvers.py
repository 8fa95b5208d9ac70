"""solve_transport with the call surface of the reference's `numerics/solvers.py:206-372`.

The GPU branch of the reference (`_solve_gpu_asm`, taken when use_gpu and de_field are given
and pc_type is None / bjacobi / jacobi) assembles the operator on the GPU and solves it with
PETSc GMRES(50) + bjacobi, warm-started from the current solution.  Here the same operator
(same assembly, same RHS, same row-only Dirichlet rows) is solved by libdsv's Jacobi-PCG,
all on the GPU: the matrix is symmetric positive definite on the free rows once the
Dirichlet values are placed in the initial guess, so CG replaces GMRES; `ksp_type` and
`pc_type` are validated for compatibility but the Krylov method is always Jacobi-PCG.
Convergence: ||b - A x||_2 <= rtol ||b||_2 measured in FP64 (PETSc's default test is relative
to the preconditioned ||b||).

Every other branch of the reference (libCEED MatShell, DOLFINx CPU/GPU standard path) needs
DOLFINx forms; they are not part of this library and raise NotImplementedError.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .core import ptr, species_array, species_struct, stream_ptr, to_device
from .gpu_assembler import GPUSparseAssembler

_SHARED_GEOM = None
_GPU_ASM_CACHE = {}


def set_shared_geometry(asm) -> None:
    """Register the primary GPUSparseAssembler as the shared geometry source
    (`numerics/solvers.py:44-47`)."""
    global _SHARED_GEOM
    _SHARED_GEOM = asm


def _mesh_of(de_field):
    V = de_field.function_space
    m = V.mesh
    return getattr(m, "dsv_mesh", m)


def _arr(f):
    if f is None:
        return None
    return f.x.array if hasattr(f, "x") else f


class WarmStartHistory:
    """Device history of one species' solve starts, newest first (the Galerkin warm start's
    directions are the differences of consecutive starts).  It stands in the reference's
    per-name cache as entry "x" (`numerics/solvers.py:358-360`), whose ``set(0.0)`` the driver
    calls after redistancing because the warm start has gone stale (`dissolve.py:710-719`):
    here ``set`` discards the history, so the next solve starts from ``u`` alone."""

    def __init__(self, n_dofs, device, depth):
        import torch
        self.bufs = [torch.empty(n_dofs, dtype=torch.float64, device=device)
                     for _ in range(depth + 1)]
        self.n = 0

    def set(self, value=0.0):
        self.n = 0

    def push(self, u_host):
        """Copies the new start into the oldest buffer and makes it the newest."""
        import torch
        new = self.bufs.pop()
        new.copy_(torch.as_tensor(np.asarray(u_host), dtype=torch.float64))
        self.bufs.insert(0, new)
        self.n = min(self.n + 1, len(self.bufs))
        return new


def solve_transport(name: str, a_form, L_form, u, bcs: list = None, ksp_type: str = "gmres",
                    pc_type: str = "bjacobi", rtol: float = 1e-8, max_it: int = 1000,
                    under_relax: float = None, comm=None, extra_petsc_opts: dict = None,
                    ceed_op=None, use_gpu: bool = False, de_field=None, alpha_field=None,
                    u_old=None, source_field=None, bc_value: float = 0.0,
                    method: str = "mf32_ir64", warm_start: str = "extrapolate",
                    history_dirs: int = 8) -> dict:
    """Solve one species' backward-Euler step a(u, v) = L(v) on the GPU.

    Returns {"converged", "iterations", "time_s"} like the reference, plus "rel_residual".
    `method` selects the libdsv solver: "mf32_ir64" (default: FP32 matrix-free CG inside FP64
    refinement), "mf64" or "csr64" (assembled FP64 CSR, the reference's operator path).
    `warm_start` "extrapolate" (default, matrix-free methods) starts from the Galerkin warm
    start over the last `history_dirs` increments of this species' starts (u itself as the
    fallback); "previous" starts from u, like the reference."""
    t0 = time.time()
    if bcs is None:
        bcs = []
    u_prev = None
    if under_relax is not None and under_relax < 1.0:
        u_prev = u.x.array.copy()
        u_old = u_prev   # `numerics/solvers.py:246-247` rebinds u_old to the current u
    if ceed_op is not None:
        raise NotImplementedError("the libCEED MatShell path needs DOLFINx forms; use de_field")
    if not (use_gpu and de_field is not None and pc_type in (None, "bjacobi", "jacobi")):
        raise NotImplementedError("only the GPU assembled path (use_gpu=True, de_field given, "
                                  "pc_type in {None, 'bjacobi', 'jacobi'}) is provided")
    if ksp_type not in ("gmres", "fgmres", "cg", "bcgs"):
        raise ValueError(f"unsupported ksp_type {ksp_type!r}")
    if method not in _lib.METHODS:
        raise ValueError(f"unknown method {method!r}")
    if warm_start not in ("extrapolate", "previous"):
        raise ValueError(f"unknown warm_start {warm_start!r}")
    try:
        converged, iterations, rel = _solve_gpu_asm(name, de_field, L_form, u, bcs, ksp_type,
                                                    pc_type or "bjacobi", rtol, max_it, comm,
                                                    alpha_field=alpha_field, u_old=u_old,
                                                    source_field=source_field, bc_value=bc_value,
                                                    method=method, warm_start=warm_start,
                                                    history_dirs=history_dirs)
    except Exception as e:
        # as `numerics/solvers.py:264-271`: report and drop the species' cached assembler; there
        # is no CPU path to fall back to here, so the error propagates
        print(f"  [GPU-ASM] {name}: failed ({e})")
        _GPU_ASM_CACHE.pop(name, None)
        raise
    if u_prev is not None:
        u.x.array[:] = under_relax * u.x.array[:] + (1.0 - under_relax) * u_prev
    u.x.scatter_forward()
    return {"converged": converged, "iterations": iterations, "time_s": time.time() - t0,
            "rel_residual": rel}


def _solve_gpu_asm(name, de_field, L_form, u, bcs, ksp_type, pc_type, rtol, max_it, comm,
                   alpha_field=None, u_old=None, source_field=None, bc_value: float = 0.0,
                   method: str = "mf32_ir64", warm_start: str = "extrapolate",
                   history_dirs: int = 8):
    """Per-name cached assembler and warm-start history; RHS, start from u (or the Galerkin
    warm start over the history), solve, read back (`numerics/solvers.py:287-372`)."""
    import ctypes
    import torch
    mesh = _mesh_of(de_field)
    inv_dt = float(getattr(de_field, "_solver_inv_dt", 1.0))
    bc_dofs = getattr(de_field, "_solver_bc_dofs", None)
    alpha_arr = _arr(alpha_field)
    depth = max(1, min(int(history_dirs), 8))
    cache = _GPU_ASM_CACHE.get(name)
    if cache is None:
        asm = GPUSparseAssembler(mesh, comm=comm, ref_asm=_SHARED_GEOM)
        cache = {"asm": asm, "x": WarmStartHistory(asm.n_dofs, mesh.device, depth),
                 "sol": torch.zeros(asm.n_dofs, dtype=torch.float64, device=mesh.device)}
        _GPU_ASM_CACHE[name] = cache
    elif len(cache["x"].bufs) != depth + 1:
        cache["x"] = WarmStartHistory(cache["asm"].n_dofs, mesh.device, depth)
    asm, hist, sol = cache["asm"], cache["x"], cache["sol"]
    if u_old is None:
        raise NotImplementedError("the DOLFINx RHS branch needs L_form; pass u_old")
    d_b = asm.assemble_rhs(_arr(u_old), inv_dt, f=_arr(source_field), bc_dofs=bc_dofs,
                           bc_value=bc_value)
    start = hist.push(u.x.array)
    keep = []
    mesh.fresh_operator()
    s = asm.species(_arr(de_field), alpha_arr, bc_dofs, bc_value, keep=keep)
    s.b = ptr(d_b).value
    s.x = ptr(sol).value
    arr = species_array([s])
    st = _lib.Stats()
    n_hist = hist.n - 1 if (warm_start == "extrapolate" and method != "csr64") else 0
    stream = stream_ptr(mesh.device)
    if n_hist > 0:
        steps = hist.bufs[:n_hist + 1]
        harr = (ctypes.c_void_p * len(steps))(*[ptr(t).value for t in steps])
        _lib.call("dsv_solve_warm_history", mesh.handle, 1, arr, inv_dt, int(asm.lumped),
                  _lib.METHODS[method], float(rtol), int(max_it), n_hist, harr, 0, None, stream,
                  st, None, None, None)
    else:
        sol.copy_(start)
        _lib.call("dsv_solve", mesh.handle, 1, arr, inv_dt, int(asm.lumped),
                  _lib.METHODS[method], float(rtol), int(max_it), stream, st)
    # asm.mat is assembled from these coefficients only if the caller reads it
    asm.defer_assembly(_arr(de_field), inv_dt, alpha=alpha_arr, bc_dofs=bc_dofs)
    u.x.array[:] = sol.cpu().numpy()
    return bool(st.converged[0]), int(st.iterations[0]), float(st.rel_residual[0])


def solve_batched(mesh, species, inv_dt, rtol=1e-8, max_it=5000, method="mf32_ir64",
                  lumped=1, stream=None):
    """Solve up to three species sharing `mesh` in the same kernels.

    species: list of dicts with keys D / D_value, alpha / alpha_value, bc_dofs, bc_value,
    b (torch f64 device), x (torch f64 device, warm start in, solution out).
    Returns the libdsv statistics as a dict."""
    mesh.fresh_operator()
    keep = []
    structs = []
    for sp in species:
        s = species_struct(mesh, D=sp.get("D"), D_value=sp.get("D_value", 0.0),
                           alpha=sp.get("alpha"), alpha_value=sp.get("alpha_value", 0.0),
                           bc_dofs=sp.get("bc_dofs"), bc_value=sp.get("bc_value", 0.0),
                           b=sp["b"], x=sp["x"], keep=keep)
        structs.append(s)
    st = _lib.Stats()
    _lib.call("dsv_solve", mesh.handle, len(structs), species_array(structs), float(inv_dt),
              int(lumped), _lib.METHODS[method], float(rtol), int(max_it),
              stream if stream is not None else stream_ptr(mesh.device), st)
    return st.as_dict(len(structs))


def galerkin_warm_start(mesh, species, inv_dt, dirs, lumped=1, stream=None):
    """Galerkin warm start of up to three species (libdsv dsv_warm_start): replaces each
    species' x by the point of x + span{dirs[k][s]} minimising the A-energy norm of the error
    when that lowers the FP64 residual (x stays the fallback).

    species: dicts as for solve_batched (b and x required); dirs: list (1..3) of per-species
    lists of device FP64 vectors (None = zero).  Returns (coefficients [k][s], residual norm
    of the given x, residual norm of the start left in x) per species."""
    import ctypes
    mesh.fresh_operator()
    keep = []
    structs = []
    for sp in species:
        structs.append(species_struct(mesh, D=sp.get("D"), D_value=sp.get("D_value", 0.0),
                                      alpha=sp.get("alpha"), alpha_value=sp.get("alpha_value", 0.0),
                                      bc_dofs=sp.get("bc_dofs"), bc_value=sp.get("bc_value", 0.0),
                                      b=sp["b"], x=sp["x"], keep=keep))
    S, nd = len(structs), len(dirs)
    if any(len(dk) != S for dk in dirs):
        raise ValueError("every direction needs one vector per species")
    x0 = species[0]["x"]
    for dk in dirs:
        for d in dk:
            if d is not None and (d.numel() != mesh.n_dofs or d.dtype != x0.dtype
                                  or d.device != x0.device or not d.is_contiguous()):
                raise ValueError("directions must be contiguous FP64 device vectors of length V")
    arr = (ctypes.c_void_p * (nd * S))(*[ptr(d).value for dk in dirs for d in dk])
    coef = (ctypes.c_double * (nd * S))()
    r0 = (ctypes.c_double * S)()
    r1 = (ctypes.c_double * S)()
    _lib.call("dsv_warm_start", mesh.handle, S, species_array(structs), float(inv_dt), int(lumped),
              nd, arr, coef, r0, r1, stream if stream is not None else stream_ptr(mesh.device))
    return ([[coef[k * S + s] for s in range(S)] for k in range(nd)], list(r0), list(r1))


class MatrixFreeOperator:
    """Matrix-free operator handle with the reference's `mult(A, x, y)` hook
    (`numerics/solvers.py:172-203`); x and y are host or device arrays of length n_dofs."""

    def __init__(self, ceed_op, ndofs: int, comm=None):
        self.ceed_op = ceed_op
        self.ndofs = ndofs
        self.comm = comm
        self.mat = self

    def mult(self, A, x, y):
        import torch
        out = self.ceed_op.apply(x, None)
        if isinstance(y, torch.Tensor):
            y.copy_(out)
        else:
            y[:] = out.cpu().numpy()

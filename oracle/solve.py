"""Oracle: Jacobi-PCG, direct solve and one backward-Euler species step (numpy/scipy, FP64).

Test infrastructure only (see ``oracle/__init__.py``).

The reference solves each species with PETSc KSP (GMRES(50) + bjacobi, nonzero initial guess =
previous solution, `numerics/solvers.py:322-336,366-369`) on the row-eliminated matrix of
`gpu/gpu_assembler.py:147-165`.  The matrix is SPD on the free rows once the Dirichlet values
are fixed in the initial guess, so the oracle's iterative solve is Jacobi-PCG with the
stopping test ||b - A x||_2 <= rtol ||b||_2, and its ground truth is ``scipy.sparse.linalg.splu``.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from . import assembly


def pcg_jacobi(A, b, x0, rtol, max_it=10000, bc_dofs=None, bc_value=0.0):
    """Jacobi-preconditioned CG on the row-eliminated CSR ``A`` (scipy), FP64.

    x0[bc] is set to bc_value first, so r[bc] = 0 and the iteration stays on the free rows.
    Returns (x, iterations, converged, final relative residual)."""
    x = np.array(x0, dtype=np.float64)
    if bc_dofs is not None and len(bc_dofs) > 0:
        x[np.asarray(bc_dofs)] = bc_value
    dinv = 1.0 / A.diagonal()
    r = b - A @ x
    bn = np.linalg.norm(b)
    tol = rtol * bn
    rn = np.linalg.norm(r)
    if rn <= tol:
        return x, 0, True, rn / bn if bn > 0 else 0.0
    z = dinv * r
    p = z.copy()
    rz = r @ z
    it = 0
    while it < max_it:
        q = A @ p
        alpha = rz / (p @ q)
        x += alpha * p
        r -= alpha * q
        it += 1
        rn = np.linalg.norm(r)
        if rn <= tol:
            break
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    true_rel = np.linalg.norm(b - A @ x) / bn if bn > 0 else 0.0
    return x, it, rn <= tol, true_rel


def direct(A, b):
    return spla.splu(A.tocsc()).solve(b)


def be_step(conn, xyz, u_old, dt, De, alpha=None, f=None, bc_dofs=None, bc_value=0.0,
            lumped=1, rtol=None, x0=None, max_it=10000):
    """One backward-Euler step of one species: assemble A and b as the GPU assembler does,
    then solve (direct when rtol is None, else Jacobi-PCG from x0 / u_old)."""
    inv_dt = 1.0 / dt
    indptr, indices, data = assembly.assemble_values(conn, xyz, De, inv_dt, alpha, lumped,
                                                     bc_dofs)
    A = assembly.to_scipy(indptr, indices, data, len(xyz))
    b = assembly.assemble_rhs(conn, xyz, u_old, inv_dt, f, lumped, bc_dofs, bc_value)
    if rtol is None:
        return direct(A, b)
    x, it, ok, rel = pcg_jacobi(A, b, u_old if x0 is None else x0, rtol, max_it,
                                bc_dofs, bc_value)
    return x


def under_relax(u_new, u_prev, omega):
    """u = omega u_new + (1 - omega) u_prev (`numerics/solvers.py:278-279`)."""
    return omega * u_new + (1.0 - omega) * u_prev


def _c_lib():
    import ctypes
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
    lib = ctypes.CDLL(path)
    lib.oracle_pcg.restype = ctypes.c_int
    P = ctypes.c_void_p
    I, Dbl = ctypes.c_int, ctypes.c_double
    lib.oracle_pcg.argtypes = [I, P, P, P, P, P, Dbl, I, I, ctypes.POINTER(Dbl)]
    lib.oracle_assemble.restype = None
    lib.oracle_assemble.argtypes = [I, P, P, P, P, Dbl, I, P, P, P, I, P, I, P, I]
    lib.oracle_rhs.restype = None
    lib.oracle_rhs.argtypes = [I, I, P, P, P, P, Dbl, I, P, I, Dbl, P, I]
    return lib


def pcg_jacobi_c(indptr, indices, data, b, x0, rtol, max_it=10000, threads=0):
    """The C restatement (oracle/pcg.c) of pcg_jacobi; returns (x, iterations, rel)."""
    import ctypes
    lib = _c_lib()
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    a = np.ascontiguousarray(data, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    rel = ctypes.c_double()
    c = lambda t: t.ctypes.data_as(ctypes.c_void_p)
    it = lib.oracle_pcg(len(bb), c(ip), c(ix), c(a), c(bb), c(x), float(rtol), int(max_it),
                        int(threads), ctypes.byref(rel))
    return x, it, rel.value


class CStep:
    """One species' backward-Euler step in the oracle's C restatement (oracle/assemble.c and
    oracle/pcg.c): assembly through the scatter map, RHS, Jacobi-PCG.  The pattern and the
    scatter map are built once (numpy), like the reference's assembler set-up
    (`gpu/gpu_assembler.py:369-450`); assemble / rhs / solve are the per-step work."""

    def __init__(self, conn, xyz):
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.C, self.V = len(conn), len(xyz)
        self.indptr, self.indices = assembly.pattern(self.conn, self.V)
        self.smap = np.ascontiguousarray(assembly.scatter_map(self.conn, self.indptr, self.indices),
                                         dtype=np.int32)
        self.values = np.zeros(len(self.indices))
        self.lib = _c_lib()

    @staticmethod
    def _p(a):
        import ctypes
        return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

    def assemble(self, De, inv_dt, alpha=None, lumped=1, bc_dofs=None, threads=1):
        De_c = np.ascontiguousarray(assembly.cell_average(De, self.conn, self.C, self.V))
        al_c = np.ascontiguousarray(assembly.cell_average(alpha, self.conn, self.C, self.V))
        bc = None if bc_dofs is None or len(bc_dofs) == 0 else \
            np.ascontiguousarray(bc_dofs, dtype=np.int32)
        p = self._p
        self.lib.oracle_assemble(self.C, p(self.conn), p(self.xyz), p(De_c), p(al_c),
                                 float(inv_dt), int(lumped), p(self.smap), p(self.indptr),
                                 p(self.indices), len(self.indices), p(bc),
                                 0 if bc is None else len(bc), p(self.values), int(threads))
        return self.values

    def rhs(self, u_old, inv_dt, f=None, lumped=1, bc_dofs=None, bc_value=0.0, threads=1):
        u = np.ascontiguousarray(u_old, dtype=np.float64)
        fv = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        if fv is not None and len(fv) != self.V:
            raise ValueError("CStep.rhs takes a per-vertex f")
        bc = None if bc_dofs is None or len(bc_dofs) == 0 else \
            np.ascontiguousarray(bc_dofs, dtype=np.int32)
        b = np.zeros(self.V)
        p = self._p
        self.lib.oracle_rhs(self.C, self.V, p(self.conn), p(self.xyz), p(u), p(fv), float(inv_dt),
                            int(lumped), p(bc), 0 if bc is None else len(bc), float(bc_value),
                            p(b), int(threads))
        return b

    def solve(self, b, x0, rtol, max_it=20000, threads=1):
        return pcg_jacobi_c(self.indptr, self.indices, self.values, b, x0, rtol, max_it, threads)

"""CPU oracle of the Galerkin warm start (test infrastructure only, like the rest of oracle/).

No counterpart in the reference, whose solves start from the previous solution
(`numerics/solvers.py:329,366-367`).  For one species with operator A (the assembled matrix
with Dirichlet identity rows, `gpu/gpu_assembler.py:147-165`), right-hand side b, start x0
and directions d_k: x0[bc] = bc_value, d_k[bc] = 0, r0 = b - A x0,
(d_j . A d_k) c = d_j . r0 solved by LDL^T with a direction dropped (c_k = 0) when its pivot
falls below 1e-10 of d_k . A d_k, x = x0 + sum_k c_k d_k, kept only if ||b - A x|| < ||r0||.

Pinning (tests/test_oracle_warmstart.py): the exact solution increment as the only direction
gives the exact solution; the residual of the result is orthogonal to every direction
(Galerkin condition); the A-energy error of the result is below that of any other point of
x0 + span{d_k} (brute force over random combinations); duplicated and zero directions are
dropped without changing the result; a direction orthogonal to r0 leaves x0 in place; a
direction holding NaN is dropped and the others are still used.
"""
import numpy as np


def solve_gram(G, g, rel_tol=1e-10):
    """LDL^T solve of the small SPD system G c = g, dropping numerically dependent rows."""
    n = len(g)
    L = np.zeros((n, n))
    Dg = np.zeros(n)
    keep = np.zeros(n, bool)
    for j in range(n):
        dj = G[j, j] - sum(L[j, k] ** 2 * Dg[k] for k in range(j) if keep[k])
        keep[j] = G[j, j] > 0 and np.isfinite(dj) and dj > rel_tol * G[j, j]
        if not keep[j]:
            continue
        Dg[j] = dj
        for i in range(j + 1, n):
            L[i, j] = (G[i, j] - sum(L[i, k] * L[j, k] * Dg[k] for k in range(j) if keep[k])) / dj
    y = np.zeros(n)
    for j in range(n):
        if keep[j]:
            y[j] = g[j] - sum(L[j, k] * y[k] for k in range(j) if keep[k])
    c = np.zeros(n)
    for j in range(n - 1, -1, -1):
        if keep[j]:
            c[j] = y[j] / Dg[j] - sum(L[i, j] * c[i] for i in range(j + 1, n) if keep[i])
    return c


def galerkin_start(A, b, x0, dirs, bc_dofs=None, bc_value=0.0):
    """(x, c, ||r0||, ||r(x)||) for scipy sparse A; dirs is a list of vectors."""
    x0 = np.array(x0, float)
    D = np.array([np.array(d, float) for d in dirs]).T  # V x n
    if bc_dofs is not None and len(bc_dofs):
        x0[bc_dofs] = bc_value
        D[bc_dofs, :] = 0.0
    r0 = b - A @ x0
    AD = A @ D
    G = D.T @ AD
    G = 0.5 * (G + G.T)
    c = solve_gram(G, D.T @ r0)
    used = c != 0.0  # a dropped direction (c = 0) adds nothing, even if it holds NaN or Inf
    x = x0 + D[:, used] @ c[used]
    n0, n1 = np.linalg.norm(r0), np.linalg.norm(b - A @ x)
    if not (np.isfinite(n1) and n1 < n0):
        return x0, np.zeros_like(c), n0, n0
    return x, c, n0, n1

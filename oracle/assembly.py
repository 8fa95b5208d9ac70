"""Oracle: P1-tet assembly of A = (1/dt + alpha_e) M_e + De_e K_e into CSR (numpy, FP64).

Test infrastructure only (see ``oracle/__init__.py``).  Each function names the reference
lines it restates.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# reference-tet basis gradients (`gpu/gpu_assembler.py:41-46`)
REF_GRAD = np.array([[-1.0, -1.0, -1.0],
                     [1.0, 0.0, 0.0],
                     [0.0, 1.0, 0.0],
                     [0.0, 0.0, 1.0]])
# consistent unit mass (1 + delta_ij)/120 (`gpu/gpu_assembler.py:48-51`)
M_UNIT = (np.ones((4, 4)) + np.eye(4)) / 120.0


def geometry(conn, xyz):
    """|detJ|, G = J^-1 J^-T and J^-T per cell, J = [x1-x0 | x2-x0 | x3-x0]
    (`gpu/gpu_assembler.py:380-386`)."""
    p = xyz[conn]
    edges = p[:, 1:, :] - p[:, 0:1, :]
    J = edges.transpose(0, 2, 1)
    detJ = np.abs(np.linalg.det(J))
    Jinv = np.linalg.inv(J)
    G = Jinv @ Jinv.transpose(0, 2, 1)
    JinvT = Jinv.transpose(0, 2, 1)
    return detJ, G, JinvT


def pattern(conn, n_dofs):
    """Sorted-column CSR pattern = union over cells of the 4x4 vertex cliques
    (`gpu/gpu_assembler.py:395-410`), vectorised with a sort/unique of (row, col) keys."""
    c = conn.astype(np.int64)
    rows = np.repeat(c, 4, axis=1).ravel()
    cols = np.tile(c, (1, 4)).ravel()
    keys = np.unique(rows * n_dofs + cols)
    r = keys // n_dofs
    indices = (keys % n_dofs).astype(np.int32)
    indptr = np.zeros(n_dofs + 1, dtype=np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    return indptr, indices


def pattern_bruteforce(conn, n_dofs):
    """Literal set-union pattern, for tiny meshes (`gpu/gpu_assembler.py:396-407`)."""
    adj = [set() for _ in range(n_dofs)]
    for e in conn:
        for r in e:
            adj[int(r)].update(int(v) for v in e)
    indptr = np.zeros(n_dofs + 1, dtype=np.int32)
    for r in range(n_dofs):
        indptr[r + 1] = indptr[r] + len(adj[r])
    indices = np.empty(indptr[-1], dtype=np.int32)
    for r in range(n_dofs):
        indices[indptr[r]:indptr[r + 1]] = sorted(adj[r])
    return indptr, indices


def scatter_map(conn, indptr, indices):
    """CSR position of every (cell, i, j) entry, flattened cell*16 + i*4 + j
    (`gpu/gpu_assembler.py:423-438`)."""
    n = len(indptr) - 1
    c = conn.astype(np.int64)
    rows = np.repeat(c, 4, axis=1).ravel()
    cols = np.tile(c, (1, 4)).ravel()
    row_of = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    csr_keys = row_of * n + indices.astype(np.int64)       # sorted ascending
    pos = np.searchsorted(csr_keys, rows * n + cols)
    assert np.array_equal(csr_keys[pos], rows * n + cols)
    return pos.astype(np.int32)


def cell_average(arr, conn, n_cells, n_dofs):
    """Per-cell coefficient: None -> zeros, length C -> as is, length V -> mean of the 4
    vertex values, else ValueError (`gpu/gpu_assembler.py:452-470`)."""
    if arr is None:
        return np.zeros(n_cells)
    arr = np.asarray(arr, dtype=np.float64)
    if arr.ndim == 0:                       # a scalar coefficient (build extension)
        return np.full(n_cells, float(arr))
    n = int(arr.shape[0])
    if n == n_cells:
        return arr.copy()
    if n == n_dofs:
        return arr[conn].reshape(n_cells, 4).mean(axis=1)
    raise ValueError(f"array length {n} matches neither n_cells={n_cells} nor n_dofs={n_dofs}")


def element_matrices(detJ, G, De_c, alpha_c, inv_dt, lumped):
    """A_e[i,j] = (dJ De/6) g_i^T G g_j + (inv_dt + alpha) M_ij with lumped M = dJ/24 I or
    consistent M = dJ (1+delta)/120 (`gpu/gpu_assembler.py:92-132`).  Shape (C, 4, 4)."""
    K = np.einsum("ia,cab,jb->cij", REF_GRAD, G, REF_GRAD) * (detJ * De_c / 6.0)[:, None, None]
    if lumped:
        M = (detJ * 5.0 / 120.0)[:, None, None] * np.eye(4)[None]
    else:
        M = detJ[:, None, None] * M_UNIT[None]
    return K + (inv_dt + alpha_c)[:, None, None] * M


def assemble_values(conn, xyz, De, inv_dt, alpha=None, lumped=1, bc_dofs=None,
                    indptr=None, indices=None):
    """CSR values of A in the pattern of ``pattern`` with Dirichlet rows applied
    (`gpu/gpu_assembler.py:472-532`)."""
    n_cells, n_dofs = len(conn), len(xyz)
    if indptr is None:
        indptr, indices = pattern(conn, n_dofs)
    detJ, G, _ = geometry(conn, xyz)
    De_c = cell_average(De, conn, n_cells, n_dofs)
    al_c = cell_average(alpha, conn, n_cells, n_dofs)
    Ae = element_matrices(detJ, G, De_c, al_c, inv_dt, lumped)
    smap = scatter_map(conn, indptr, indices)
    data = np.zeros(len(indices))
    np.add.at(data, smap, Ae.reshape(-1))
    if bc_dofs is not None and len(bc_dofs) > 0:
        apply_dirichlet_rows(np.asarray(bc_dofs), indptr, indices, data)
    return indptr, indices, data


def assemble_dense_bruteforce(conn, xyz, De_c, alpha_c, inv_dt, lumped):
    """Dense A by a literal per-cell double loop (tiny meshes only)."""
    n = len(xyz)
    A = np.zeros((n, n))
    for e, nodes in enumerate(conn):
        p = xyz[nodes]
        J = (p[1:] - p[0]).T
        dJ = abs(np.linalg.det(J))
        Jinv = np.linalg.inv(J)
        Gm = Jinv @ Jinv.T
        for i in range(4):
            for j in range(4):
                K = REF_GRAD[i] @ Gm @ REF_GRAD[j] * dJ * De_c[e] / 6.0
                if lumped:
                    M = dJ * 5.0 / 120.0 if i == j else 0.0
                else:
                    M = dJ * (2.0 if i == j else 1.0) / 120.0
                A[nodes[i], nodes[j]] += K + (inv_dt + alpha_c[e]) * M
    return A


def apply_dirichlet_rows(bc_dofs, indptr, indices, data):
    """Row-only elimination: zero the row, 1 on the diagonal (`gpu/gpu_assembler.py:147-165`)."""
    for row in np.asarray(bc_dofs, dtype=np.int64):
        rs, re = indptr[row], indptr[row + 1]
        seg = indices[rs:re]
        data[rs:re] = np.where(seg == row, 1.0, 0.0)
    return data


def searchsorted_left_per_key(a, keys):
    """CuPy's ``searchsorted(a, keys)`` (side 'left', array taken as increasing), the call of
    the reference's per-cell RHS branch (`gpu/gpu_assembler.py:565-569`; CuPy is
    ``cupy-cuda12x``, version unpinned, `INSTALL.md:373,839`).  CuPy searches every key on its
    own: len(a) when key > a[-1], else the left bisection over [0, len(a) - 1]
    (mid = lo + (hi - lo) // 2; a[mid] < key -> lo = mid + 1, else hi = mid; answer hi).  On a
    sorted array that is the usual insertion point; on the unsorted flat connectivity it is
    whatever the bisection path lands on.  numpy's ``searchsorted`` differs there: for
    increasing keys it starts each bisection from the previous key's answer.  Vectorised over
    the keys (every key takes the same number of halvings of its own interval)."""
    a = np.asarray(a)
    keys = np.asarray(keys)
    n = len(a)
    lo = np.zeros(len(keys), dtype=np.int64)
    hi = np.full(len(keys), n - 1, dtype=np.int64)
    active = lo < hi
    while active.any():
        mid = lo + (hi - lo) // 2
        go_right = active & (a[np.minimum(mid, n - 1)] < keys)
        go_left = active & ~go_right
        lo = np.where(go_right, mid + 1, lo)
        hi = np.where(go_left, mid, hi)
        active = lo < hi
    return np.where(keys > a[n - 1], n, hi)


def assemble_rhs(conn, xyz, u_old, inv_dt, f=None, lumped=1, bc_dofs=None, bc_value=0.0):
    """b = M (inv_dt u_old + f), lumped (dJ/24 per vertex) or consistent, then b[bc] = bc_value
    (`gpu/gpu_assembler.py:176-242,534-603`).  A per-cell f follows the reference's
    per-cell branch (`:560-571`): f_vert = f[searchsorted(conn.ravel(), arange(V))] with CuPy's
    per-key search (``searchsorted_left_per_key``) and out-of-range indices wrapped as CuPy's
    fancy indexing does."""
    n_cells, n_dofs = len(conn), len(xyz)
    u_old = np.asarray(u_old, dtype=np.float64)
    if u_old.shape[0] != n_dofs:
        raise ValueError(f"u_old length {u_old.shape[0]} != n_dofs {n_dofs}")
    if f is None:
        fv = np.zeros(n_dofs)
    elif len(f) == n_dofs:
        fv = np.asarray(f, dtype=np.float64)
    elif len(f) == n_cells:
        pos = searchsorted_left_per_key(conn.ravel(), np.arange(n_dofs, dtype=np.int32))
        fv = np.asarray(f, dtype=np.float64)[pos % n_cells]
    else:
        raise ValueError(f"f length {len(f)} unrecognised")
    detJ, _, _ = geometry(conn, xyz)
    b = np.zeros(n_dofs)
    u = u_old[conn]
    fe = fv[conn]
    if lumped:
        Ld = detJ * 5.0 / 120.0
        contrib = Ld[:, None] * (inv_dt * u + fe)
    else:
        base = detJ / 120.0
        contrib = (base * inv_dt)[:, None] * (u.sum(1, keepdims=True) + u) + \
            base[:, None] * (fe.sum(1, keepdims=True) + fe)
    np.add.at(b, conn.ravel(), contrib.ravel())
    if bc_dofs is not None and len(bc_dofs) > 0:
        b[np.asarray(bc_dofs, dtype=np.int64)] = bc_value
    return b


def to_scipy(indptr, indices, data, n):
    return sp.csr_matrix((data, indices, indptr), shape=(n, n))


def lumped_mass(conn, xyz):
    detJ, _, _ = geometry(conn, xyz)
    m = np.zeros(len(xyz))
    np.add.at(m, conn.ravel(), np.repeat(detJ / 24.0, 4))
    return m

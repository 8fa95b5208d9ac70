"""Device mesh handle and tensor plumbing around libdsv.

PyTorch provides device memory and the current HIP stream; every computation on the path is
a libdsv kernel.  All device arrays handed to the library are contiguous torch tensors on the
mesh's device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

VIEW_INDPTR, VIEW_INDICES, VIEW_SCATTER_MAP, VIEW_DETJ, VIEW_GFLAT, VIEW_JINVT, VIEW_CONN = range(7)


def _torch():
    import torch
    return torch


def stream_ptr(device=None):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def to_device(arr, device, dtype):
    """numpy / torch (any device) -> contiguous torch tensor on `device` with `dtype`."""
    torch = _torch()
    if isinstance(arr, torch.Tensor):
        if arr.dtype == dtype and arr.device == device and arr.is_contiguous():
            return arr
        t = arr.to(device=device, dtype=dtype)
    else:
        t = torch.as_tensor(np.asarray(arr), device=device, dtype=dtype)
    return t.contiguous()


class Mesh:
    """A P1 tetrahedral mesh resident on one GPU: connectivity and coordinates, CSR pattern,
    scatter map, per-cell geometry and the matrix-free partition (built by libdsv).

    conn : (C, 4) int vertex ids, xyz : (V, 3) float64 coordinates, both host arrays.
    """

    def __init__(self, conn, xyz, vertices_per_block=0, device=None):
        torch = _torch()
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise _lib.DsvLibraryError("dsv.Mesh needs a GPU (torch.cuda.is_available() is False)")
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if conn.ndim != 2 or conn.shape[1] != 4:
            raise ValueError("conn must have shape (n_cells, 4)")
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("xyz must have shape (n_verts, 3)")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                   else torch.device(device).index or 0)
        torch.cuda.set_device(self.device)
        h = ctypes.c_void_p()
        _lib.check(lib.dsv_mesh_create(ctypes.byref(h), conn.shape[0], xyz.shape[0],
                                       conn.ctypes.data_as(ctypes.c_void_p),
                                       xyz.ctypes.data_as(ctypes.c_void_p),
                                       int(vertices_per_block)))
        self._h = h
        self.conn = conn
        self.xyz = xyz
        self._views = {}
        self.refresh_info()

    def refresh_info(self):
        info = (ctypes.c_int64 * 8)()
        _lib.call("dsv_mesh_info", self._h, info)
        self.n_cells, self.n_dofs, self.nnz = int(info[0]), int(info[1]), int(info[2])
        self.n_blocks, self.vertices_per_block, self.n_entries = int(info[3]), int(info[4]), int(info[5])
        self.n_halo, self.max_local = int(info[6]), int(info[7])

    @property
    def handle(self):
        return self._h

    def view(self, which):
        """Device copy (torch) of a mesh array: indptr, indices, scatter map, detJ, Gflat,
        JinvT or elem_nodes (flattened like the reference's d_* attributes)."""
        torch = _torch()
        if which in self._views:
            return self._views[which]
        n, dt = {
            VIEW_INDPTR: (self.n_dofs + 1, torch.int32), VIEW_INDICES: (self.nnz, torch.int32),
            VIEW_SCATTER_MAP: (16 * self.n_cells, torch.int32), VIEW_DETJ: (self.n_cells, torch.float64),
            VIEW_GFLAT: (9 * self.n_cells, torch.float64), VIEW_JINVT: (9 * self.n_cells, torch.float64),
            VIEW_CONN: (4 * self.n_cells, torch.int32),
        }[which]
        t = torch.empty(n, dtype=dt, device=self.device)
        _lib.call("dsv_mesh_copy_view", self._h, int(which), ptr(t), stream_ptr(self.device))
        self._views[which] = t
        return t

    def set_partition(self, vertices_per_block, block_threads, lanes_per_vertex=4):
        _lib.call("dsv_mesh_set_partition", self._h, int(vertices_per_block), int(block_threads),
                  int(lanes_per_vertex))
        self.refresh_info()

    def set_solver_options(self, inner_rtol=1e-3, max_outer=12, graph_chunk=8, use_graphs=False):
        """Inner tolerance and sweep limit of mf32_ir64, CG iterations per polled chunk, and
        whether chunks are replayed from captured hipGraphs (the defaults launch directly:
        on the 806k-tet disc that is about 5% faster per step than graph replay)."""
        _lib.call("dsv_set_solver_options", self._h, float(inner_rtol), int(max_outer),
                  int(graph_chunk), int(bool(use_graphs)))

    def set_species_narrowing(self, enable=True):
        """Continue a batched inner CG whose species finish at different iterations with the
        kernel of fewer species (libdsv dsv_set_species_narrowing; default on)."""
        _lib.call("dsv_set_species_narrowing", self._h, 1 if enable else 0)

    FORM_RECOMPUTE, FORM_STORED = 0, 1

    def set_operator_form(self, form):
        """Operator of the FP32 inner CG (scalar D, lumped mass): FORM_STORED (default) or
        FORM_RECOMPUTE (element Jacobians recomputed on the fly)."""
        _lib.call("dsv_set_operator_form", self._h, int(form))

    def set_operator_reuse(self, reuse=True):
        """Reuse the operator data built from unchanged coefficient arrays across calls
        (libdsv dsv_set_operator_reuse); every call discards what was built."""
        _lib.call("dsv_set_operator_reuse", self._h, 1 if reuse else 0)
        self._reuse = bool(reuse)

    def fresh_operator(self):
        """Discard any reusable operator before a call whose coefficient arrays are new
        (allocated per call, so their addresses may repeat those of freed arrays)."""
        if getattr(self, "_reuse", False):
            _lib.call("dsv_set_operator_reuse", self._h, 1)

    def close(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dsv_mesh_destroy(self._h)
            finally:
                self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def species_struct(mesh, D=None, D_value=0.0, alpha=None, alpha_value=0.0, bc_dofs=None,
                   bc_value=0.0, b=None, x=None, keep=None):
    """Fill a ctypes Species from torch device tensors (or None).  `keep` collects the
    tensors that must stay alive while the struct is in use."""
    torch = _torch()
    s = _lib.Species()
    dev = mesh.device

    def dev_f64(a):
        if a is None:
            return None
        t = to_device(a, dev, torch.float64)
        if keep is not None:
            keep.append(t)
        return t

    Dt = dev_f64(D)
    At = dev_f64(alpha)
    s.D = ptr(Dt).value if Dt is not None else None
    s.D_len = int(Dt.numel()) if Dt is not None else 0
    s.D_value = float(D_value)
    s.alpha = ptr(At).value if At is not None else None
    s.alpha_len = int(At.numel()) if At is not None else 0
    s.alpha_value = float(alpha_value)
    if bc_dofs is not None and len(bc_dofs) > 0:
        bt = to_device(bc_dofs, dev, torch.int32)
        if keep is not None:
            keep.append(bt)
        s.bc_dofs = ptr(bt).value
        s.n_bc = int(bt.numel())
    else:
        s.bc_dofs = None
        s.n_bc = 0
    s.bc_value = float(bc_value)
    s.b = ptr(b).value if b is not None else None
    s.x = ptr(x).value if x is not None else None
    return s


def species_array(structs):
    arr = (_lib.Species * len(structs))()
    for i, s in enumerate(structs):
        arr[i] = s
    return arr

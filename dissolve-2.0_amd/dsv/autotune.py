"""Autotuner of the matrix-free launch shape.

The fused apply kernel runs one workgroup per block of `vertices_per_block` consecutive
(Morton-ordered) vertices; each owned vertex gets `lanes_per_vertex` lanes that share its
incident cells, and the workgroup size sets how many lanes run at once.  Small blocks give more
workgroups but stage more halo vertices per owned one; more lanes per vertex shorten the
longest per-lane list at the price of a cross-lane sum.  Blocks whose owned + halo vertices
exceed 1024 are rejected by the library and skipped.  The tuner times the CG form of the fused
kernel (HIP events, libdsv dsv_time_apply) for each candidate on the real coefficients and
keeps the fastest.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .core import species_array, species_struct

DEFAULT_NV = (64, 128, 256, 512)
DEFAULT_THREADS = (256, 512)
DEFAULT_LANES = (1, 2, 4)


def time_apply(mesh, species, inv_dt, precision=32, mode=1, reps=30, lumped=1, stream=None):
    """Average ms per launch of the fused kernel for `species` (list of dicts)."""
    import torch
    mesh.fresh_operator()
    keep = []
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              keep=keep) for s in species]
    arr = species_array(structs)
    if stream is None:
        stream = torch.cuda.Stream(device=mesh.device)
    ms = ctypes.c_double()
    _lib.call("dsv_time_apply", mesh.handle, len(structs), arr, float(inv_dt), int(lumped),
              int(precision), int(mode), int(reps), ctypes.c_void_p(stream.cuda_stream),
              ctypes.byref(ms))
    return ms.value


def tune(mesh, species, inv_dt, nv_list=DEFAULT_NV, threads_list=DEFAULT_THREADS,
         lanes_list=DEFAULT_LANES, reps=30, precision=32):
    """Try every (vertices_per_block, threads, lanes_per_vertex) triple, leave the mesh on the
    fastest, and return (best, table) with entries (nv, threads, lanes, entries, ms)."""
    table = []
    best = None
    for nv in nv_list:
        for lanes in lanes_list:
            for th in threads_list:
                try:
                    mesh.set_partition(nv, th, lanes)
                    # the FP64 residual kernel must fit the same partition (LDS per workgroup)
                    time_apply(mesh, species, inv_dt, precision=64, mode=0, reps=1)
                    ms = time_apply(mesh, species, inv_dt, precision=precision, reps=reps)
                except ValueError:
                    continue
                table.append((nv, th, lanes, mesh.n_entries, ms))
                if best is None or ms < best[4]:
                    best = (nv, th, lanes, mesh.n_entries, ms)
    if best is None:
        raise RuntimeError("no launch shape fits this device")
    mesh.set_partition(best[0], best[1], best[2])
    return best, table

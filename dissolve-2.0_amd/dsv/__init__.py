"""dsv -- MI355X (gfx950) transport hot path of Dissolve 2.0.

P1-tet backward-Euler diffusion-reaction assembly into CSR, the matrix-free apply of the same
operator and the Jacobi-PCG solve for O2, Cl- and OH-, in hand-written HIP kernels
(libdsv.so, C-ABI in include/dsv.h) driven from Python through ctypes.

Modules mirroring the reference's call surface:
  dsv.gpu_assembler   GPUSparseAssembler            (`gpu/gpu_assembler.py`)
  dsv.ceed_operators  CeedDiffusionMassOp, ...      (`gpu/ceed_operators.py`)
  dsv.solvers         solve_transport, ...          (`numerics/solvers.py`)
New:
  dsv.core            Mesh (device mesh handle)
  dsv.transport       batched three-species backward-Euler stepper
  dsv.mlp             learned CG warm start direction (per-vertex MLP, HIP kernels)
  dsv.autotune        launch-shape autotuner of the matrix-free kernels
  dsv.workloads       the transport workloads (meshes, coefficients, initial states)
  dsv.meshes          mesh generators of the transport workloads
"""
from ._lib import DsvError, DsvLibraryError, load as load_library  # noqa: F401

__all__ = ["DsvError", "DsvLibraryError", "load_library"]

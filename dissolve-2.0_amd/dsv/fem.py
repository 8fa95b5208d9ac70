"""Minimal stand-ins for the DOLFINx objects the reference's solve_transport touches.

`numerics/solvers.py:206-372` reads `u.x.array`, `u.x.scatter_forward()`,
`de_field.x.array`, `de_field.function_space.mesh` and the side channels
`de_field._solver_inv_dt` / `de_field._solver_bc_dofs`.  These classes provide exactly that on
a :class:`dsv.core.Mesh`, so the reference's species blocks (`dissolve.py:781-954`) can call
:func:`dsv.solvers.solve_transport` unchanged apart from the imports.  With DOLFINx present,
its own Function objects work too (anything with `.x.array` and a function space whose mesh
carries a `dsv_mesh` attribute).
"""
from __future__ import annotations

import numpy as np


class FunctionSpace:
    def __init__(self, mesh):
        self.mesh = mesh


class _Vector:
    def __init__(self, n):
        self.array = np.zeros(n, dtype=np.float64)

    def scatter_forward(self):
        """Single-rank: nothing to exchange."""


class Function:
    def __init__(self, V, name=""):
        self.function_space = V
        self.name = name
        self.x = _Vector(V.mesh.n_dofs)

"""CPU oracle for the three-species transport path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy (and plain C for the Jacobi-PCG, ``pcg.c``), the
reference algorithm of Dissolve 2.0's GPU transport path:

* ``assembly``  -- geometry, CSR pattern, scatter map, element matrices, CSR assembly,
  Dirichlet rows, RHS (`gpu/gpu_assembler.py:60-242,369-603`);
* ``solve``     -- Jacobi-PCG, direct solve, one backward-Euler step, under-relaxation
  (`numerics/solvers.py:206-284,287-372`);
* ``mlp``       -- forward/backward/Adam and features of the per-vertex warm-start network;
* ``warmstart`` -- the Galerkin warm start over past increments and the learned direction.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline; the product path under
``dissolve-2.0_amd/`` never imports it.

Pinning (each function is checked by something other than a retyping of its formula):
closed-form unit-tet element matrices, K.1 = 0, symmetry/PSD, translation/rotation/scaling
invariance, a brute-force dense per-element assembly, a brute-force set-union pattern,
closed-form Kuhn-grid nnz, exact backward-Euler fixed points, scipy ``splu``/``cg``
cross-checks, central finite differences for the network gradients, and for the Galerkin
warm start exact answers, the Galerkin orthogonality and brute-force energy minimality.  The reference
itself ships no golden vectors or tests for this path and cannot be imported here (it needs
cupy, petsc4py and dolfinx), so nothing is pinned against reference-produced vectors; the
reference's own code comments (`gpu/gpu_assembler.py:15-20,49-51,122-131`) are the formula
text the closed forms check.
"""

"""The reference's time step in its own order on the GPU: Zn2+, Cl-, the film, O2 and OH-
(`dissolve.py:740-954`), optionally preceded by the interface update (`dissolve.py:660-730`).

Per step (every array device resident, every computation a libdsv kernel):

1. effective diffusivities from phi and the previous film F (`physics/interface_velocity.py:
   90-155`): D_Zn = D_Zn0 f, D_i = D_i0 f [phi <= 0] for Cl, OH, O2;
2. with `levelset` on: the recession velocity from phi, Zn, O2 and the diffusivities, the
   algebraic advance phi += v dt, Sussman redistancing every `redistance_interval` steps, and
   |grad phi| of the new phi (`dissolve.py:660-730`);
3. Zn (alpha_zn = [phi <= 0] kf (1 - F/Fmax), source [phi <= 0] kd F Cl^2 from the previous Cl,
   Dirichlet Znbc on the scaffold phi > 0), Cl (no source) and O2 (ORR sink delta_eps(phi)
   |grad phi| k_orr, Dirichlet on the bottom face) in ONE batched solve: none of them needs
   another's new value (`dissolve.py:740-791,863-903`);
4. O2 under-relaxed with omega = 0.3 (`dissolve.py:895-903`, `numerics/solvers.py:278-279`);
5. the film from the new Zn and Cl (`dissolve.py:797-833`);
6. OH with the source [phi <= 0] kd F Cl^2 + 4 delta_eps |grad phi| k_orr O2 from the new film,
   Cl and the under-relaxed O2 (`dissolve.py:919-951`), in a second solve.

The solves start from the Galerkin warm start over the last `history_dirs` increments of each
field (u_n itself as the fallback), or from u_n ("previous", the reference's start).  The
diffusivity fields change with the film every step, so no operator is reused across steps.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .coefficients import effective_diffusion, transport_coefficients, grad_norm
from .core import ptr, species_array, species_struct, stream_ptr, to_device
from .physics import (film_update, interface_velocity, mesh_h_avg, redistance_device,
                      under_relax, zinc_coefficients, levelset_advance)

NAMES = ("Zn", "Cl", "O2", "OH")


class _History:
    """Ring of the last H + 1 values of each species of a solve group, newest first; the solve
    writes into separate output buffers."""

    def __init__(self, fields, depth):
        import torch
        self.S = len(fields)
        self.ring = [[f.clone() for f in fields]] + \
            [[torch.empty_like(f) for f in fields] for _ in range(depth)]
        self.ptrs = [[f.data_ptr() for f in r] for r in self.ring]  # rotates with the ring
        self.out = [torch.empty_like(f) for f in fields]
        self.n = 1  # valid ring entries

    @property
    def current(self):
        return self.ring[0]

    def push(self):
        """The oldest ring entry becomes the newest; returns it (to be filled by the caller)."""
        oldest = self.ring.pop()
        self.ring.insert(0, oldest)
        self.ptrs.insert(0, self.ptrs.pop())
        self.n = min(self.n + 1, len(self.ring))
        return oldest


class CoupledStepper:
    def __init__(self, mesh, w, method="mf32_ir64", rtol=1e-8, max_it=20000, warm="extrapolate",
                 history_dirs=8, levelset=False, redistance_interval=1, redistance_iters=5,
                 velocity=None):
        import torch
        if warm not in ("extrapolate", "previous"):
            raise ValueError(f"unknown warm start {warm!r}")
        if warm != "previous" and method == "csr64":
            raise ValueError("the Galerkin warm start needs a matrix-free method")
        self.mesh, self.dt, self.method = mesh, float(w["dt"]), method
        self.rtol, self.max_it, self.warm = float(rtol), int(max_it), warm
        self.p = dict(w["params"])
        self.levelset = bool(levelset)
        self.redistance_interval = int(redistance_interval)
        self.redistance_iters = int(redistance_iters)
        # velocity law parameters (dsv_velocity_params); the default zn_solid keeps the
        # nondimensional recession speed near a fifth of a cell per step
        self.velocity = dict(zn_solid=2.0, zn_sat=self.p["Znbc"] * 1e-8) if velocity is None \
            else dict(velocity)
        dev = mesh.device
        t = lambda a, dt=torch.float64: to_device(a, dev, dt).clone()
        st = w["state"]
        self.phi = t(st["phi"])
        self.grad_phi = t(st["grad_phi"])
        self.F = t(st["F"])
        self.o2_bc = t(st["o2_bc"], torch.int32)
        depth = max(1, min(int(history_dirs), 8))
        self.g1 = _History([t(st["Zn"]), t(st["Cl"]), t(st["O2"])], depth)
        self.g2 = _History([t(st["OH"])], depth)
        V = mesh.n_dofs
        e = lambda: torch.empty(V, dtype=torch.float64, device=dev)
        self.De = [e() for _ in range(4)]
        self.alpha_zn, self.src_zn, self.alpha_o2, self.f_oh, self.vel = e(), e(), e(), e(), e()
        self.b = [e() for _ in range(4)]
        self.scaffold = None
        self.n_steps = 0
        self.history = []
        # the two solves' species records, built once: every array they name keeps its address
        # (the scaffold list is re-pointed when the interface moves)
        p = self.p
        self._keep = []
        self._arr1 = species_array([
            species_struct(mesh, D=self.De[0], alpha=self.alpha_zn, b=self.b[0],
                           x=self.g1.out[0], bc_value=p["Znbc"], keep=self._keep),
            species_struct(mesh, D=self.De[1], b=self.b[1], x=self.g1.out[1], keep=self._keep),
            species_struct(mesh, D=self.De[3], alpha=self.alpha_o2, bc_dofs=self.o2_bc,
                           bc_value=p["o2_bc_value"], b=self.b[2], x=self.g1.out[2],
                           keep=self._keep)])
        self._arr2 = species_array([
            species_struct(mesh, D=self.De[2], b=self.b[3], x=self.g2.out[0], keep=self._keep)])
        P3, P1 = ctypes.c_void_p * 3, ctypes.c_void_p * 1
        self._rhs1 = (P3(*self.g1.ptrs[0]), P3(self.src_zn.data_ptr(), None, None),
                      P3(*[b.data_ptr() for b in self.b[:3]]))
        self._rhs2 = (P1(self.f_oh.data_ptr()), P1(self.b[3].data_ptr()))
        self._interface_changed()

    # ---- fields by name
    def field(self, name):
        return {"Zn": self.g1.current[0], "Cl": self.g1.current[1], "O2": self.g1.current[2],
                "OH": self.g2.current[0], "F": self.F, "phi": self.phi}[name]

    def _interface_changed(self):
        """Quantities that depend on phi alone: the ORR sink and the scaffold list."""
        p = self.p
        transport_coefficients(self.mesh, self.phi, self.grad_phi, p["eps"], p["k_orr"],
                               outputs=("alpha_o2",), out={"alpha_o2": self.alpha_o2})
        _, _, self.scaffold = zinc_coefficients(self.mesh, self.phi, None, None, p["kf"], p["kd"],
                                                p["Fmax"], scaffold=True,
                                                out=(self.alpha_zn, self.src_zn))
        n = int(self.scaffold.numel())
        self._arr1[0].bc_dofs = self.scaffold.data_ptr() if n > 0 else None
        self._arr1[0].n_bc = n

    def _effective_diffusion(self):
        p = self.p
        effective_diffusion(self.mesh, self.phi, self.F, p["Fmax"], p["film_floor"], p["D0"],
                            out=self.De)

    def _interface_step(self):
        """Velocity, advance, redistancing and |grad phi| (`dissolve.py:660-730`)."""
        v = self.velocity
        interface_velocity(self.mesh, self.phi, self.g1.current[0], self.g1.current[2],
                           self.De[0], self.De[3], self.p["h"], self.p["k_orr"], v["zn_solid"],
                           v["zn_sat"], out=self.vel)
        levelset_advance(self.mesh, self.phi, self.vel, self.dt, out=self.phi)
        if self.redistance_interval > 0 and (self.n_steps + 1) % self.redistance_interval == 0:
            redistance_device(self.mesh, self.phi, self.redistance_iters, self.p["h"])
        grad_norm(self.mesh, self.phi, out=self.grad_phi)
        self._interface_changed()

    def _solve(self, grp, arr, stream):
        S = grp.S
        st = _lib.Stats()
        nh = grp.n - 1 if self.warm == "extrapolate" else 0
        if nh > 0:
            hist = (ctypes.c_void_p * ((nh + 1) * S))(*[q for a in grp.ptrs[:nh + 1] for q in a])
            _lib.call("dsv_solve_warm_history", self.mesh.handle, S, arr, 1.0 / self.dt, 1,
                      _lib.METHODS[self.method], self.rtol, self.max_it, nh, hist, 0, None,
                      stream, st, None, None, None)
        else:
            for k in range(S):
                grp.out[k].copy_(grp.current[k])
            _lib.call("dsv_solve", self.mesh.handle, S, arr, 1.0 / self.dt, 1,
                      _lib.METHODS[self.method], self.rtol, self.max_it, stream, st)
        return st.as_dict(S)

    def step(self):
        p, mesh = self.p, self.mesh
        stream = stream_ptr(mesh.device)
        # the coefficient arrays keep their addresses but change every step: nothing built from
        # them may be reused
        mesh.set_operator_reuse(False)
        self._effective_diffusion()
        if self.levelset:
            self._interface_step()
        zn, cl, o2 = self.g1.current
        # Zn coefficients from the previous film and Cl
        zinc_coefficients(mesh, self.phi, self.F, cl, p["kf"], p["kd"], p["Fmax"], scaffold=False,
                          out=(self.alpha_zn, self.src_zn))
        rhs_u, rhs_f, rhs_b = self._rhs1
        for k in range(3):
            rhs_u[k] = self.g1.ptrs[0][k]
        _lib.call("dsv_rhs_batch", mesh.handle, 3, self._arr1, rhs_u, rhs_f, 1.0 / self.dt, 1,
                  rhs_b, stream)
        r1 = self._solve(self.g1, self._arr1, stream)
        # new Zn, Cl; O2 under-relaxed against its value at the start of the step
        new = self.g1.push()
        old = self.g1.ring[1]
        new[0].copy_(self.g1.out[0])
        new[1].copy_(self.g1.out[1])
        under_relax(self.g1.out[2], old[2], p["omega"], out=new[2])
        # film from the new Zn and Cl
        film_update(mesh, self.phi, new[0], new[1], self.F, p["kf"], p["kd"], p["Fmax"], self.dt,
                    out=self.F)
        # OH source from the new film, Cl and O2
        transport_coefficients(mesh, self.phi, self.grad_phi, p["eps"], p["k_orr"], p["kd"],
                               F_film=self.F, Cl=new[1], O2=new[2], Fmax=p["Fmax"],
                               outputs=("f_oh",), out={"f_oh": self.f_oh})
        P1 = ctypes.c_void_p * 1
        _lib.call("dsv_rhs_batch", mesh.handle, 1, self._arr2, P1(self.g2.ptrs[0][0]),
                  self._rhs2[0], 1.0 / self.dt, 1, self._rhs2[1], stream)
        r2 = self._solve(self.g2, self._arr2, stream)
        self.g2.push()[0].copy_(self.g2.out[0])
        self.n_steps += 1
        d = {"iterations": r1["iterations"] + r2["iterations"],
             "converged": r1["converged"] + r2["converged"],
             "rel_residual": r1["rel_residual"] + r2["rel_residual"], "names": list(NAMES)}
        self.history.append(d)
        return d

    def h_avg(self):
        return mesh_h_avg(self.mesh)

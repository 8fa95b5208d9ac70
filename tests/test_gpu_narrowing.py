"""Species narrowing of the batched matrix-free inner CG (dsv_set_species_narrowing): once some
species are done, the rest continue with the kernel of fewer species.  Against the same solve
with narrowing off: every species converges, to the direct solution (oracle), with the same
CG iteration counts up to the order of partial sums (within 2 per species), and the narrowed
runs were taken (dsv_stats.narrowings).  Bars: solutions within 1e-7 of max|x_direct| at
rtol 1e-10, the two runs within 1e-7 of each other."""
import numpy as np
import pytest

from dsv import meshes
from oracle import assembly as OA
from oracle import solve as OS

pytestmark = pytest.mark.gpu


def _case(seed=3):
    """Three species of different difficulty on a small graded disc: a lognormal D field with
    a reaction field and Dirichlet rows, D = 2, and D = 5.3 with a reaction constant."""
    rng = np.random.default_rng(seed)
    conn, xyz = meshes.graded_disc(seed=1, scale=6.0)
    C, V = len(conn), len(xyz)
    bc = np.nonzero(xyz[:, 2] <= xyz[:, 2].min() + 1e-9)[0].astype(np.int32)
    dt = 0.005
    sps = [dict(D=rng.lognormal(0, 0.5, C), alpha=rng.uniform(0, 50, V), bc_dofs=bc, bc_value=1.0),
           dict(D_value=2.0),
           dict(D_value=5.3, alpha_value=0.3)]
    u0 = [rng.uniform(0, 1, V) for _ in sps]
    mats, bs = [], []
    for s, u in zip(sps, u0):
        D = s["D"] if "D" in s else s["D_value"]
        al = s["alpha"] if "alpha" in s else s.get("alpha_value", 0.0)
        ip, ix, data = OA.assemble_values(conn, xyz, D, 1 / dt, al, 1, s.get("bc_dofs"))
        mats.append(OA.to_scipy(ip, ix, data, V))
        bs.append(OA.assemble_rhs(conn, xyz, u, 1 / dt, rng.uniform(0, 1, V), 1, s.get("bc_dofs"),
                                  s.get("bc_value", 0.0)))
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    return conn, xyz, sps, u0, bs, exact, 1 / dt


def _solve(gpu, mesh, sps, starts, bs, inv_dt, method, narrow, lumped=1):
    import torch
    from dsv.solvers import solve_batched
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    mesh.set_species_narrowing(narrow)
    sp = [dict(s, b=t(b), x=t(u)) for s, u, b in zip(sps, starts, bs)]
    st = solve_batched(mesh, sp, inv_dt, rtol=1e-10, method=method, lumped=lumped)
    return [s["x"].cpu().numpy() for s in sp], st


def _compare(on, off, exact, sel):
    (xa, sa), (xb, sb) = on, off
    assert all(sa["converged"]) and all(sb["converged"]), (sa, sb)
    for s in sel:
        scale = np.abs(exact[s]).max()
        assert np.abs(xa[s] - exact[s]).max() <= 1e-7 * scale, s
        assert np.abs(xa[s] - xb[s]).max() <= 1e-7 * scale, s
        assert abs(sa["iterations"][s] - sb["iterations"][s]) <= 2, (s, sa, sb)
    assert sa["narrowings"] >= 1 and sb["narrowings"] == 0
    # narrowing lowers the cost of a launch, not their number; each narrowed run's last chunk
    # may add a few no-op launches
    assert 0 < sa["cg_launches"] <= 1.25 * sb["cg_launches"] + 8, (sa, sb)


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
def test_narrowing_between_chunks(gpu, method, graphs):
    """All three species start; the polls every 2 iterations see them finish one by one."""
    from dsv.core import Mesh
    conn, xyz, sps, u0, bs, exact, inv_dt = _case()
    mesh = Mesh(conn, xyz)
    mesh.set_solver_options(graph_chunk=2, use_graphs=graphs)
    on = _solve(gpu, mesh, sps, u0, bs, inv_dt, method, True)
    off = _solve(gpu, mesh, sps, u0, bs, inv_dt, method, False)
    print(f"[narrow] {method} graphs={graphs}: on {on[1]['iterations']} in "
          f"{on[1]['cg_launches']} launches ({on[1]['narrowings']} narrowings), off "
          f"{off[1]['iterations']} in {off[1]['cg_launches']}", flush=True)
    _compare(on, off, exact, range(3))


@pytest.mark.parametrize("done", [[1], [0, 2]])
@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
def test_narrowing_before_the_first_iteration(gpu, method, done):
    """Species that start at their solution are not active: the inner solve runs the others'
    kernel from the start, and the finished species keep their start exactly."""
    from dsv.core import Mesh
    conn, xyz, sps, u0, bs, exact, inv_dt = _case(seed=5)
    mesh = Mesh(conn, xyz)
    starts = [exact[s] if s in done else u0[s] for s in range(3)]
    on = _solve(gpu, mesh, sps, starts, bs, inv_dt, method, True)
    off = _solve(gpu, mesh, sps, starts, bs, inv_dt, method, False)
    print(f"[narrow] {method} done={done}: on {on[1]['iterations']} outer {on[1]['outer']} in "
          f"{on[1]['cg_launches']} launches ({on[1]['narrowings']} narrowings), off "
          f"{off[1]['iterations']} in {off[1]['cg_launches']}", flush=True)
    for s in done:
        assert on[1]["iterations"][s] == 0 and np.array_equal(on[0][s], off[0][s])
        assert np.abs(on[0][s] - exact[s]).max() <= 1e-13 * np.abs(exact[s]).max()
    _compare(on, off, exact, [s for s in range(3) if s not in done])


def test_narrowing_with_consistent_mass(gpu):
    """Consistent mass runs the recomputing generic kernels (per-cell D, no stored geometry):
    narrowing hands over their coefficient arrays the same way."""
    from dsv.core import Mesh
    conn, xyz, sps, u0, _, _, inv_dt = _case(seed=17)
    rng = np.random.default_rng(17)
    V = len(xyz)
    bs, exact = [], []
    for s, u in zip(sps, u0):
        D = s["D"] if "D" in s else s["D_value"]
        al = s["alpha"] if "alpha" in s else s.get("alpha_value", 0.0)
        ip, ix, data = OA.assemble_values(conn, xyz, D, inv_dt, al, 0, s.get("bc_dofs"))
        b = OA.assemble_rhs(conn, xyz, u, inv_dt, rng.uniform(0, 1, V), 0, s.get("bc_dofs"),
                            s.get("bc_value", 0.0))
        bs.append(b)
        exact.append(OS.direct(OA.to_scipy(ip, ix, data, V), b))
    mesh = Mesh(conn, xyz)
    mesh.set_solver_options(graph_chunk=2)
    for method in ("mf32_ir64", "mf64"):
        on = _solve(gpu, mesh, sps, u0, bs, inv_dt, method, True, lumped=0)
        off = _solve(gpu, mesh, sps, u0, bs, inv_dt, method, False, lumped=0)
        _compare(on, off, exact, range(3))


def test_csr_path_never_narrows(gpu):
    """The plain FP64 CSR path has its own kernels and vectors: with narrowing on it solves as
    before and takes no narrowed run."""
    from dsv.core import Mesh
    conn, xyz, sps, u0, bs, exact, inv_dt = _case(seed=11)
    mesh = Mesh(conn, xyz)
    mesh.set_solver_options(graph_chunk=2)
    x, st = _solve(gpu, mesh, sps, u0, bs, inv_dt, "csr64", True)
    assert all(st["converged"]) and st["narrowings"] == 0, st
    assert len(set(st["iterations"])) == 3  # the species finish at different iterations
    for s in range(3):
        assert np.abs(x[s] - exact[s]).max() <= 1e-7 * np.abs(exact[s]).max()


def test_two_species_narrow_to_one(gpu):
    from dsv.core import Mesh
    conn, xyz, sps, u0, bs, exact, inv_dt = _case(seed=7)
    mesh = Mesh(conn, xyz)
    mesh.set_solver_options(graph_chunk=2)
    pick = [0, 2]
    args = ([sps[s] for s in pick], [u0[s] for s in pick], [bs[s] for s in pick], inv_dt,
            "mf32_ir64")
    on = _solve(gpu, mesh, *args[:4], args[4], True)
    off = _solve(gpu, mesh, *args[:4], args[4], False)
    _compare(on, off, [exact[s] for s in pick], range(2))


def test_stepper_with_narrowing(gpu):
    """The benchmark's stepper (Galerkin warm start, FP32 CG in FP64 refinement) on a small
    disc: with narrowing on and off, the same converged trajectory to rounding."""
    from dsv import workloads
    from dsv.core import Mesh
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    runs = []
    for narrow in (True, False):
        mesh = Mesh(w["conn"], w["xyz"])
        mesh.set_species_narrowing(narrow)
        st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=w["rtol"])
        st.set_warm_start("extrapolate")
        hist = [st.step() for _ in range(6)]
        assert all(all(h["converged"]) for h in hist)
        runs.append((st, hist))
    (a, ha), (b, hb) = runs
    assert sum(h["narrowings"] for h in ha) >= 1
    assert sum(h["narrowings"] for h in hb) == 0
    for ua, ub in zip(a.u, b.u):
        ua, ub = ua.cpu().numpy(), ub.cpu().numpy()
        assert np.abs(ua - ub).max() <= 1e-6 * np.abs(ub).max()

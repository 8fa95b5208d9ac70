"""GPU Galerkin warm start (dsv_warm_start) against the CPU oracle (oracle/warmstart.py),
through the C-ABI.  Bars: coefficients within 1e-8 relative, start points within 1e-10
relative to max|x|, residual norms within 1e-9 relative; the exact increment reproduces the
direct solution to 1e-9."""
import numpy as np
import pytest

from dsv import meshes
from oracle import assembly as OA
from oracle import solve as OS
from oracle import warmstart as OW

pytestmark = pytest.mark.gpu


def _case(name, seed=3):
    rng = np.random.default_rng(seed)
    if name == "disc_small":
        conn, xyz = meshes.graded_disc(seed=1, scale=6.0)
    else:
        conn, xyz = meshes.kuhn_cube(7)
        xyz = meshes.jitter(xyz, 1 / 7, 0.2, rng, ~meshes.boundary_vertex_mask(xyz))
        conn, xyz = meshes.locality_order(conn, xyz)
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
    return conn, xyz, sps, u0, mats, bs, 1 / dt, rng


def _run(dev, mesh, sps, u0, bs, inv_dt, dirs):
    import torch
    from dsv.solvers import galerkin_warm_start
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    sp = []
    for s, u, b in zip(sps, u0, bs):
        d = dict(s)
        d["b"], d["x"] = t(b), t(u)
        sp.append(d)
    tdirs = [[t(d) if d is not None else None for d in dk] for dk in dirs]
    coef, r0, r1 = galerkin_warm_start(mesh, sp, inv_dt, tdirs)
    return [d["x"].cpu().numpy() for d in sp], coef, r0, r1


@pytest.mark.parametrize("name", ["kuhn7_jitter_morton", "disc_small"])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_matches_oracle(gpu, name, nd):
    from dsv.core import Mesh
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case(name)
    mesh = Mesh(conn, xyz)
    V = len(xyz)
    # smooth-ish directions (increments of the exact solution plus noise) and pure noise
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    dirs = []
    for k in range(nd):
        dirs.append([(exact[s] - u0[s]) * rng.uniform(0.5, 1.5) + rng.normal(0, 0.1, V)
                     if k == 0 else rng.normal(0, 1, V) for s in range(3)])
    x, coef, r0, r1 = _run(gpu, mesh, sps, u0, bs, inv_dt, dirs)
    for s in range(3):
        xr, cr, n0, n1 = OW.galerkin_start(mats[s], bs[s], u0[s], [d[s] for d in dirs],
                                           sps[s].get("bc_dofs"), sps[s].get("bc_value", 0.0))
        assert abs(r0[s] - n0) <= 1e-9 * n0
        assert abs(r1[s] - n1) <= 1e-9 * n0
        for k in range(nd):
            assert abs(coef[k][s] - cr[k]) <= 1e-8 * max(np.abs(cr).max(), 1e-30), (s, k)
        assert np.abs(x[s] - xr).max() <= 1e-10 * np.abs(xr).max()
        assert n1 < n0


def test_exact_increment_and_dropped_directions(gpu):
    from dsv.core import Mesh
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton", seed=5)
    mesh = Mesh(conn, xyz)
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    inc = []
    for s in range(3):
        x0 = u0[s].copy()
        if sps[s].get("bc_dofs") is not None:
            x0[sps[s]["bc_dofs"]] = sps[s]["bc_value"]
        inc.append(exact[s] - x0)
    # the exact increment, a duplicate of it and an absent (NULL) direction
    x, coef, r0, r1 = _run(gpu, mesh, sps, u0, bs, inv_dt,
                           [inc, [2.0 * d for d in inc], [None, None, None]])
    for s in range(3):
        assert abs(coef[0][s] - 1.0) < 1e-9 and coef[1][s] == 0.0 and coef[2][s] == 0.0
        assert np.abs(x[s] - exact[s]).max() <= 1e-9 * np.abs(exact[s]).max()
        assert r1[s] < 1e-8 * r0[s]


def test_argument_errors(gpu):
    import torch
    from dsv.core import Mesh
    from dsv.solvers import galerkin_warm_start
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton")
    mesh = Mesh(conn, xyz)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    sp = [dict(sps[0], b=t(bs[0]), x=t(u0[0]))]
    d = t(np.ones(len(xyz)))
    for bad in ([], [[d]] * 4):
        with pytest.raises(ValueError):
            galerkin_warm_start(mesh, sp, inv_dt, bad)
    with pytest.raises(ValueError):
        galerkin_warm_start(mesh, sp, inv_dt, [[d[:-1]]])


@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
def test_solve_warm_matches_separate_warm_start_and_solve(gpu, method):
    """dsv_solve_warm (directions from past solutions, folded into the solve) against
    dsv_warm_start with the same directions followed by dsv_solve: the same coefficients, and
    solutions within the solve tolerance's reach."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    from dsv.solvers import galerkin_warm_start, solve_batched
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("disc_small", seed=9)
    mesh = Mesh(conn, xyz)
    V = len(xyz)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    # a synthetic history: past solutions drifting toward the exact one
    hist = [[t(u0[s] - (j + 1) * 0.3 * (exact[s] - u0[s]) + rng.normal(0, 1e-3, V))
             for s in range(3)] for j in range(2)]
    dirs = [[t(u0[s]) - hist[0][s] for s in range(3)], [hist[0][s] - hist[1][s] for s in range(3)]]
    # separate: warm start, then solve
    sp_a = [dict(s, b=t(b), x=t(u)) for s, u, b in zip(sps, u0, bs)]
    coef_a, r0_a, r1_a = galerkin_warm_start(mesh, sp_a, inv_dt, dirs)
    st_a = solve_batched(mesh, sp_a, inv_dt, rtol=1e-10, method=method)
    # fused
    keep = []
    sp_b = [dict(s, b=t(b), x=t(u)) for s, u, b in zip(sps, u0, bs)]
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              b=s["b"], x=s["x"], keep=keep) for s in sp_b]
    harr = (ctypes.c_void_p * 6)(*[ptr(hist[j][s]).value for j in range(2) for s in range(3)])
    coef = (ctypes.c_double * 6)()
    r0, r1 = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    st = _lib.Stats()
    _lib.call("dsv_solve_warm", mesh.handle, 3, species_array(structs), inv_dt, 1,
              _lib.METHODS[method], 1e-10, 5000, 2, harr, 0, None, stream_ptr(gpu), st, coef, r0, r1)
    d = st.as_dict(3)
    assert all(d["converged"]) and all(st_a["converged"])
    for s in range(3):
        assert abs(r0[s] - r0_a[s]) <= 1e-10 * r0_a[s]
        for k in range(2):
            assert abs(coef[k * 3 + s] - coef_a[k][s]) <= 1e-10 * max(abs(coef_a[k][s]), 1e-300)
        a, b = sp_a[s]["x"].cpu().numpy(), sp_b[s]["x"].cpu().numpy()
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
        assert np.abs(b - exact[s]).max() <= 1e-7 * np.abs(exact[s]).max()
    # a direction that does not help is dropped: the start stays and the solve still converges
    sp_c = [dict(s, b=t(b), x=t(u)) for s, u, b in zip(sps, u0, bs)]
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              b=s["b"], x=s["x"], keep=keep) for s in sp_c]
    same = [t(u0[s]) for s in range(3)]  # the start itself as u_{n-1}: d_0 = x - x = 0
    keep.extend(same)
    zero = (ctypes.c_void_p * 3)(*[ptr(v).value for v in same])
    _lib.call("dsv_solve_warm", mesh.handle, 3, species_array(structs), inv_dt, 1,
              _lib.METHODS[method], 1e-10, 5000, 1, zero, 0, None, stream_ptr(gpu), st, coef, r0, r1)
    d = st.as_dict(3)
    assert all(d["converged"]) and all(r1[s] == r0[s] for s in range(3))
    with pytest.raises(ValueError):
        _lib.call("dsv_solve_warm", mesh.handle, 3, species_array(structs), inv_dt, 1,
                  _lib.METHODS["csr64"], 1e-10, 5000, 1, zero, 0, None, stream_ptr(gpu), st,
                  None, None, None)


@pytest.mark.parametrize("n_hist,n_extra", [(8, 0), (7, 1), (6, 0), (5, 1), (4, 0)])
def test_solve_warm_many_directions_match_oracle(gpu, n_hist, n_extra):
    """dsv_solve_warm with up to eight directions (past solutions and explicit ones): the
    coefficients and both residual norms of its Galerkin start against the oracle's
    galerkin_start over the same directions (the start's FP64 projection is the oracle's;
    bars as above), and the solve converges to the direct solution."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton", seed=13)
    mesh = Mesh(conn, xyz)
    V = len(xyz)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    # past solutions drifting smoothly toward the exact one, plus noise
    past = [[u0[s] - (j + 1) * 0.2 * (exact[s] - u0[s]) * (1 + 0.1 * j) + rng.normal(0, 1e-3, V)
             for s in range(3)] for j in range(n_hist)]
    extra = [[rng.normal(0, 1, V) for _ in range(3)] for _ in range(n_extra)]
    keep = [t(v) for row in past + extra for v in row]
    hist_t = keep[:3 * n_hist]
    extra_t = keep[3 * n_hist:]
    sp = [dict(s, b=t(b), x=t(u)) for s, u, b in zip(sps, u0, bs)]
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              b=s["b"], x=s["x"], keep=keep) for s in sp]
    harr = (ctypes.c_void_p * (3 * n_hist))(*[ptr(v).value for v in hist_t])
    earr = (ctypes.c_void_p * max(1, 3 * n_extra))(*[ptr(v).value for v in extra_t])
    nd = n_hist + n_extra
    coef = (ctypes.c_double * (3 * nd))()
    r0, r1 = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    st = _lib.Stats()
    _lib.call("dsv_solve_warm", mesh.handle, 3, species_array(structs), inv_dt, 1,
              _lib.METHODS["mf64"], 1e-10, 5000, n_hist, harr, n_extra, earr, stream_ptr(gpu), st,
              coef, r0, r1)
    d = st.as_dict(3)
    assert all(d["converged"]), d
    for s in range(3):
        seq = [u0[s]] + [past[j][s] for j in range(n_hist)]
        dirs = [seq[j] - seq[j + 1] for j in range(n_hist)] + [extra[k][s] for k in range(n_extra)]
        xr, cr, n0, n1 = OW.galerkin_start(mats[s], bs[s], u0[s], dirs, sps[s].get("bc_dofs"),
                                           sps[s].get("bc_value", 0.0))
        assert abs(r0[s] - n0) <= 1e-9 * n0
        assert abs(r1[s] - n1) <= 1e-9 * n0
        for k in range(nd):
            assert abs(coef[k * 3 + s] - cr[k]) <= 1e-8 * max(np.abs(cr).max(), 1e-30), (s, k)
        x = sp[s]["x"].cpu().numpy()
        assert np.abs(x - exact[s]).max() <= 1e-7 * np.abs(exact[s]).max()


@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
@pytest.mark.parametrize("n_hist,n_extra", [(3, 0), (2, 1), (0, 1)])
def test_solve_warm_history_matches_solve_warm(gpu, method, n_hist, n_extra):
    """dsv_solve_warm_history (start = hist_0, directions hist_j - hist_(j+1)) against
    dsv_solve_warm with x = a copy of hist_0 and the history hist_1 ...: the same start and
    directions, so the same bits -- solution, coefficients, residual norms, iterations.  The
    solution array's content on entry (NaN here) is not read; a solution array named in the
    history and a NULL history are refused."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton", seed=17)
    mesh = Mesh(conn, xyz)
    V = len(xyz)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    start = [t(u0[s]) for s in range(3)]
    past = [[t(u0[s] - (j + 1) * 0.25 * (exact[s] - u0[s]) + rng.normal(0, 1e-3, V))
             for s in range(3)] for j in range(n_hist)]
    extra = [[t(rng.normal(0, 1, V)) for _ in range(3)] for _ in range(n_extra)]
    keep = [v for row in past + extra for v in row] + start
    earr = (ctypes.c_void_p * max(1, 3 * n_extra))(*[ptr(v).value for r in extra for v in r])
    bt = [t(b) for b in bs]
    outs = []
    for fn in ("dsv_solve_warm", "dsv_solve_warm_history"):
        if fn == "dsv_solve_warm":
            x = [s.clone() for s in start]
            rows = past
        else:
            x = [torch.full((V,), float("nan"), dtype=torch.float64, device=gpu) for _ in range(3)]
            rows = [start] + past
        keep.extend(x)
        structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                  alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                                  bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                                  b=bt[k], x=x[k], keep=keep) for k, s in enumerate(sps)]
        harr = (ctypes.c_void_p * max(1, 3 * len(rows)))(*[ptr(v).value for r in rows for v in r])
        nd = n_hist + n_extra
        coef = (ctypes.c_double * (3 * nd))()
        r0, r1 = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        st = _lib.Stats()
        _lib.call(fn, mesh.handle, 3, species_array(structs), inv_dt, 1, _lib.METHODS[method],
                  1e-10, 5000, n_hist, harr if rows else None, n_extra, earr, stream_ptr(gpu), st,
                  coef, r0, r1)
        torch.cuda.synchronize()
        outs.append(([v.cpu().numpy() for v in x], list(coef), list(r0), list(r1), st.as_dict(3)))
    (xa, ca, r0a, r1a, da), (xb, cb, r0b, r1b, db) = outs
    assert all(da["converged"]) and all(db["converged"]), (da, db)
    assert ca == cb and r0a == r0b and r1a == r1b
    assert da["iterations"] == db["iterations"] and da["outer"] == db["outer"]
    for s in range(3):
        assert np.array_equal(xa[s], xb[s])
        assert np.abs(xb[s] - exact[s]).max() <= 1e-7 * np.abs(exact[s]).max()
    # refusals: a solution array in the history; no history at all
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              b=bt[k], x=start[k], keep=keep) for k, s in enumerate(sps)]
    harr = (ctypes.c_void_p * 3)(*[ptr(v).value for v in start])
    for h in (harr, None):
        with pytest.raises(ValueError):
            _lib.call("dsv_solve_warm_history", mesh.handle, 3, species_array(structs), inv_dt, 1,
                      _lib.METHODS[method], 1e-10, 5000, 0, h, 0, None, stream_ptr(gpu), None,
                      None, None, None)


@pytest.mark.parametrize("pick", [[2], [0, 2]])
def test_solve_warm_history_fewer_species(gpu, pick):
    """dsv_solve_warm_history for one and two species against dsv_solve_warm with the start
    copied into x: the same bits."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton", seed=23)
    mesh = Mesh(conn, xyz)
    V, S = len(xyz), len(pick)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    exact = [OS.direct(mats[s], bs[s]) for s in pick]
    start = [t(u0[s]) for s in pick]
    past = [[t(u0[s] - (j + 1) * 0.3 * (exact[i] - u0[s]) + rng.normal(0, 1e-3, V))
             for i, s in enumerate(pick)] for j in range(3)]
    bt = [t(bs[s]) for s in pick]
    keep = [v for row in past for v in row] + start + bt
    outs = []
    for fn in ("dsv_solve_warm", "dsv_solve_warm_history"):
        if fn == "dsv_solve_warm":
            x, rows = [v.clone() for v in start], past
        else:
            x = [torch.full((V,), float("nan"), dtype=torch.float64, device=gpu) for _ in pick]
            rows = [start] + past
        keep.extend(x)
        structs = [species_struct(mesh, D=sps[s].get("D"), D_value=sps[s].get("D_value", 0.0),
                                  alpha=sps[s].get("alpha"),
                                  alpha_value=sps[s].get("alpha_value", 0.0),
                                  bc_dofs=sps[s].get("bc_dofs"), bc_value=sps[s].get("bc_value", 0.0),
                                  b=bt[i], x=x[i], keep=keep) for i, s in enumerate(pick)]
        harr = (ctypes.c_void_p * (S * len(rows)))(*[ptr(v).value for r in rows for v in r])
        coef = (ctypes.c_double * (3 * S))()
        st = _lib.Stats()
        _lib.call(fn, mesh.handle, S, species_array(structs), inv_dt, 1, _lib.METHODS["mf32_ir64"],
                  1e-10, 5000, 3, harr, 0, None, stream_ptr(gpu), st, coef, None, None)
        torch.cuda.synchronize()
        outs.append(([v.cpu().numpy() for v in x], list(coef), st.as_dict(S)))
    (xa, ca, da), (xb, cb, db) = outs
    assert all(da["converged"]) and all(db["converged"]) and ca == cb, (da, db)
    assert da["iterations"] == db["iterations"]
    for i in range(S):
        assert np.array_equal(xa[i], xb[i])
        assert np.abs(xb[i] - exact[i]).max() <= 1e-7 * np.abs(exact[i]).max()


@pytest.mark.parametrize("method", ["mf32_ir64", "mf64"])
def test_nan_direction_keeps_start_finite(gpu, method):
    """An explicit direction holding NaN for one species (and one of Inf for another) is
    dropped: the coefficients and residual norms of the start match the oracle's galerkin_start
    over the same directions (which drops them too and still uses the history), the start and
    the solution stay finite, and the solve converges to the direct solution."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, species_array, species_struct, stream_ptr, ptr
    conn, xyz, sps, u0, mats, bs, inv_dt, rng = _case("kuhn7_jitter_morton", seed=29)
    mesh = Mesh(conn, xyz)
    V = len(xyz)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=gpu)
    exact = [OS.direct(A, b) for A, b in zip(mats, bs)]
    n_hist = 2
    past = [[u0[s] - (j + 1) * 0.3 * (exact[s] - u0[s]) + rng.normal(0, 1e-3, V)
             for s in range(3)] for j in range(n_hist)]
    extra = [rng.normal(0, 1, V) for _ in range(3)]
    extra[1][[3, 50, 200]] = np.nan
    extra[2][7] = np.inf
    rows = [[t(u0[s]) for s in range(3)]] + [[t(v) for v in r] for r in past]
    ext = [t(v) for v in extra]
    bt = [t(b) for b in bs]
    x = [torch.full((V,), float("nan"), dtype=torch.float64, device=gpu) for _ in range(3)]
    keep = [v for r in rows for v in r] + ext + bt + x
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), alpha_value=s.get("alpha_value", 0.0),
                              bc_dofs=s.get("bc_dofs"), bc_value=s.get("bc_value", 0.0),
                              b=bt[k], x=x[k], keep=keep) for k, s in enumerate(sps)]
    harr = (ctypes.c_void_p * (3 * len(rows)))(*[ptr(v).value for r in rows for v in r])
    earr = (ctypes.c_void_p * 3)(*[ptr(v).value for v in ext])
    nd = n_hist + 1
    coef = (ctypes.c_double * (3 * nd))()
    r0, r1 = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    st = _lib.Stats()
    _lib.call("dsv_solve_warm_history", mesh.handle, 3, species_array(structs), inv_dt, 1,
              _lib.METHODS[method], 1e-10, 5000, n_hist, harr, 1, earr, stream_ptr(gpu), st,
              coef, r0, r1)
    torch.cuda.synchronize()
    d = st.as_dict(3)
    assert all(d["converged"]), d
    for s in range(3):
        seq = [u0[s]] + [past[j][s] for j in range(n_hist)]
        dirs = [seq[j] - seq[j + 1] for j in range(n_hist)] + [extra[s]]
        xr, cr, n0, n1 = OW.galerkin_start(mats[s], bs[s], u0[s], dirs, sps[s].get("bc_dofs"),
                                           sps[s].get("bc_value", 0.0))
        assert np.isfinite(r1[s]) and abs(r0[s] - n0) <= 1e-9 * n0
        assert abs(r1[s] - n1) <= 1e-9 * n0
        for k in range(nd):
            assert abs(coef[k * 3 + s] - cr[k]) <= 1e-8 * max(np.abs(cr).max(), 1e-30), (s, k)
        if s > 0:
            assert coef[n_hist * 3 + s] == 0.0 and cr[n_hist] == 0.0
            assert n1 < n0  # the history directions are still taken
        xs = x[s].cpu().numpy()
        assert np.all(np.isfinite(xs))
        assert np.abs(xs - exact[s]).max() <= 1e-7 * np.abs(exact[s]).max()


def test_fused_step_is_bit_identical(gpu):
    """dsv_step_warm_history (right-hand sides formed inside the solve call) against
    dsv_rhs_batch + dsv_solve_warm_history: the same right-hand sides and solutions, bit for
    bit, over Galerkin steps with a Dirichlet species, a source and a reaction field; and its
    argument checks."""
    import ctypes
    import numpy as np
    import pytest
    import torch
    from dsv import _lib, workloads
    from dsv.core import Mesh
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    mesh = Mesh(w["conn"], w["xyz"], vertices_per_block=128)
    runs = []
    for fuse in (False, True):
        st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=1e-8)
        st.set_warm_start("extrapolate", 6)
        st.fuse_rhs = fuse
        out = []
        for _ in range(5):
            h = st.step()
            assert all(h["converged"]), h
            out.append([t.cpu().numpy().copy() for t in st.b + st.u])
        runs.append(out)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    # b may not be read by the step: a right-hand side that is also a history array is refused
    P = ctypes.c_void_p * 3
    hist = (ctypes.c_void_p * 6)(*[st.b[0].data_ptr()] + [u.data_ptr() for u in st.u] + [0, 0])
    with pytest.raises(ValueError, match="history"):
        _lib.call("dsv_step_warm_history", mesh.handle, 3, st._arr, P(None, None, None),
                  1.0 / w["dt"], 1, 1, 1e-8, 100, 1, hist, 0, P(), None, _lib.Stats(), None,
                  None, None)
    with pytest.raises(ValueError, match="lumped"):
        _lib.call("dsv_step_warm_history", mesh.handle, 3, st._arr, P(None, None, None),
                  1.0 / w["dt"], 0, 1, 1e-8, 100, 1, hist, 0, P(), None, _lib.Stats(), None,
                  None, None)

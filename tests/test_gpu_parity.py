"""GPU parity of libdsv against the oracle, through the C-ABI (ctypes).

Bars: CSR pattern and scatter map bit-exact; FP64 matrix / RHS entries within 1e-12
relative (to the largest entry of the row set); FP64 matrix-free apply within 1e-12; FP32
apply within 2e-5 relative to max|y|; every solve within 1e-8 relative of scipy splu.
"""
import ctypes

import numpy as np
import pytest

from dsv import _lib, meshes, workloads
from oracle import assembly as OA
from oracle import solve as OS

pytestmark = pytest.mark.gpu


def _mesh_cases():
    out = {}
    conn, xyz = meshes.kuhn_cube(6)
    out["kuhn6"] = (conn, xyz)
    rng = np.random.default_rng(11)
    c2, x2 = meshes.kuhn_cube(7)
    x2 = meshes.jitter(x2, 1 / 7, 0.2, rng, ~meshes.boundary_vertex_mask(x2))
    out["kuhn7_jitter_morton"] = meshes.locality_order(c2, x2)
    out["disc_small"] = meshes.graded_disc(seed=1, scale=6.0)
    # random cell order + random vertex numbering (no locality at all)
    c3, x3 = meshes.kuhn_cube(5)
    p = rng.permutation(len(x3))
    inv = np.empty_like(p)
    inv[p] = np.arange(len(p))
    out["kuhn5_shuffled"] = (inv[c3][rng.permutation(len(c3))].astype(np.int32), x3[p])
    return out


CASES = _mesh_cases()


@pytest.fixture(scope="module")
def torch_dev(gpu):
    return gpu


def _m(name, **kw):
    from dsv.core import Mesh
    conn, xyz = CASES[name]
    return Mesh(conn, xyz, **kw), conn, xyz


def _relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", list(CASES))
def test_pattern_and_scatter_map_bit_exact(torch_dev, name):
    from dsv.core import VIEW_INDICES, VIEW_INDPTR, VIEW_SCATTER_MAP, VIEW_DETJ
    m, conn, xyz = _m(name)
    ip, ix = OA.pattern(conn, len(xyz))
    assert m.nnz == len(ix)
    np.testing.assert_array_equal(m.view(VIEW_INDPTR).cpu().numpy(), ip)
    np.testing.assert_array_equal(m.view(VIEW_INDICES).cpu().numpy(), ix)
    np.testing.assert_array_equal(m.view(VIEW_SCATTER_MAP).cpu().numpy(), OA.scatter_map(conn, ip, ix))
    detJ, G, JinvT = OA.geometry(conn, xyz)
    assert _relerr(m.view(VIEW_DETJ).cpu().numpy(), detJ) < 1e-12


@pytest.mark.parametrize("name", ["kuhn7_jitter_morton", "disc_small", "kuhn5_shuffled"])
@pytest.mark.parametrize("lumped", [1, 0])
@pytest.mark.parametrize("coef", ["cell", "vertex", "scalar"])
def test_assembly_values(torch_dev, name, lumped, coef):
    from dsv.gpu_assembler import GPUSparseAssembler
    m, conn, xyz = _m(name)
    rng = np.random.default_rng(3)
    C, V = len(conn), len(xyz)
    if coef == "scalar":
        D, al = np.full(V, 1.7), None
    else:
        n = C if coef == "cell" else V
        D, al = rng.lognormal(0, 0.5, n), rng.uniform(0, 2, n)
    bc = np.sort(rng.choice(V, size=max(1, V // 20), replace=False)).astype(np.int32)
    asm = GPUSparseAssembler(m)
    asm.lumped = lumped
    asm.assemble(D, 200.0, alpha=al, bc_dofs=bc)
    ip, ix, ref = OA.assemble_values(conn, xyz, D, 200.0, al, lumped, bc)
    got = asm.d_data.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # the exported CSR has the reference layout
    ip2, ix2, v2 = asm.mat.getValuesCSR()
    np.testing.assert_array_equal(ip2, ip)
    np.testing.assert_array_equal(ix2, ix)


@pytest.mark.parametrize("name", ["kuhn7_jitter_morton", "disc_small", "kuhn5_shuffled"])
def test_assembly_scalar_coefficients(torch_dev, name):
    """D and alpha as scalars take the streaming form D K1 + (inv_dt + alpha) M1, and a scalar D
    with a per-vertex alpha under lumped mass the edge-regrouped reaction mass: the oracle's
    entries, and the row assembly of the same values as arrays, within 1e-12 / 1e-13, for
    both lumpings (each alternated with the other, so that both unit mass matrices are kept),
    with and without Dirichlet rows."""
    import torch
    from dsv.core import species_struct, stream_ptr
    m, conn, xyz = _m(name)
    C, V = len(conn), len(xyz)
    rng = np.random.default_rng(21)
    bc = np.sort(rng.choice(V, size=max(1, V // 20), replace=False)).astype(np.int32)
    out = torch.empty(m.nnz, dtype=torch.float64, device=m.device)
    for lumped in (1, 0, 1):
        for D, al, b in ((1.7, 0.0, None), (5.3, 0.45, bc), (0.0, 2.0, None)):
            keep = []
            s = species_struct(m, D_value=D, alpha_value=al, bc_dofs=b, keep=keep)
            _lib.call("dsv_assemble", m.handle, s, 200.0, lumped, ctypes.c_void_p(out.data_ptr()),
                      stream_ptr(m.device))
            got = out.cpu().numpy()
            _, _, ref = OA.assemble_values(conn, xyz, np.full(C, D), 200.0, np.full(C, al),
                                           lumped, b)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (lumped, D, al)
            s2 = species_struct(m, D=np.full(C, D), alpha=np.full(C, al), bc_dofs=b, keep=keep)
            _lib.call("dsv_assemble", m.handle, s2, 200.0, lumped, ctypes.c_void_p(out.data_ptr()),
                      stream_ptr(m.device))
            rows = out.cpu().numpy()
            assert np.abs(got - rows).max() <= 1e-13 * np.abs(rows).max(), (lumped, D, al)
        # scalar D with a per-vertex reaction field: the edge-regrouped lumped reaction mass
        # (lumped), the row assembly (consistent); the same bars
        for D, b in ((2.0, None), (5.3, bc)):
            keep = []
            al = rng.uniform(0.0, 3.0, V)
            s = species_struct(m, D_value=D, alpha=al, bc_dofs=b, keep=keep)
            _lib.call("dsv_assemble", m.handle, s, 200.0, lumped, ctypes.c_void_p(out.data_ptr()),
                      stream_ptr(m.device))
            got = out.cpu().numpy()
            _, _, ref = OA.assemble_values(conn, xyz, np.full(C, D), 200.0, al, lumped, b)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (lumped, D, "vertex")
            s2 = species_struct(m, D=np.full(C, D), alpha=al, bc_dofs=b, keep=keep)
            _lib.call("dsv_assemble", m.handle, s2, 200.0, lumped, ctypes.c_void_p(out.data_ptr()),
                      stream_ptr(m.device))
            rows = out.cpu().numpy()
            assert np.abs(got - rows).max() <= 1e-13 * np.abs(rows).max(), (lumped, D, "vertex")


@pytest.mark.parametrize("name", ["disc_small", "kuhn5_shuffled"])
@pytest.mark.parametrize("lumped", [1, 0])
def test_rhs(torch_dev, lumped, name):
    from dsv.gpu_assembler import GPUSparseAssembler
    m, conn, xyz = _m(name)
    rng = np.random.default_rng(4)
    V = len(xyz)
    u, f = rng.uniform(0, 1, V), rng.uniform(-1, 1, V)
    bc = np.arange(0, V, 17, dtype=np.int32)
    asm = GPUSparseAssembler(m)
    asm.lumped = lumped
    got = asm.assemble_rhs(u, 200.0, f=f, bc_dofs=bc, bc_value=0.3).cpu().numpy()
    ref = OA.assemble_rhs(conn, xyz, u, 200.0, f, lumped, bc, 0.3)
    assert _relerr(got, ref) < 1e-12
    # per-cell f: the reference's lossy branch (CuPy's per-key search of every vertex in the
    # unsorted flat connectivity, on the GPU), with and without Dirichlet rows
    fc = rng.uniform(0, 1, len(conn))
    got = asm.assemble_rhs(u, 200.0, f=fc).cpu().numpy()
    assert _relerr(got, OA.assemble_rhs(conn, xyz, u, 200.0, fc, lumped)) < 1e-12
    got = asm.assemble_rhs(u, 200.0, f=fc, bc_dofs=bc, bc_value=0.3).cpu().numpy()
    assert _relerr(got, OA.assemble_rhs(conn, xyz, u, 200.0, fc, lumped, bc, 0.3)) < 1e-12
    with pytest.raises(ValueError):
        asm.assemble_rhs(u[:-1], 200.0)
    with pytest.raises(ValueError):
        asm.assemble_rhs(u, 200.0, f=np.ones(V + 3))


@pytest.mark.parametrize("lumped", [1, 0])
def test_rhs_batch_matches_per_species(torch_dev, lumped):
    """dsv_rhs_batch (three species in one call) against dsv_rhs per species: the same bits,
    with a source on two species and Dirichlet lists on two; NULL outputs are refused."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import Mesh, ptr, species_array, species_struct, stream_ptr
    conn, xyz = meshes.graded_disc(seed=1, scale=6.0)
    m = Mesh(conn, xyz)
    rng = np.random.default_rng(8)
    V = len(xyz)
    t = lambda a, dt=torch.float64: torch.tensor(a, dtype=dt, device=torch_dev)
    u = [t(rng.uniform(0, 1, V)) for _ in range(3)]
    f = [t(rng.uniform(-1, 1, V)), None, t(rng.uniform(0, 2, V))]
    bc = [t(np.arange(0, V, 13, dtype=np.int32), torch.int32), None,
          t(np.arange(5, V, 29, dtype=np.int32), torch.int32)]
    bcv = [0.7, 0.0, -1.25]
    keep = []
    sp = species_array([species_struct(m, D_value=1.0, bc_dofs=bc[k], bc_value=bcv[k], keep=keep)
                        for k in range(3)])
    ref = [torch.zeros(V, dtype=torch.float64, device=torch_dev) for _ in range(3)]
    for k in range(3):
        _lib.call("dsv_rhs", m.handle, ptr(u[k]), ptr(f[k]), V if f[k] is not None else 0, 200.0,
                  lumped, ptr(bc[k]), int(bc[k].numel()) if bc[k] is not None else 0, bcv[k],
                  ptr(ref[k]), stream_ptr(torch_dev))
    got = [torch.full((V,), float("nan"), dtype=torch.float64, device=torch_dev) for _ in range(3)]
    P = ctypes.c_void_p * 3
    _lib.call("dsv_rhs_batch", m.handle, 3, sp, P(*[x.data_ptr() for x in u]),
              P(*[x.data_ptr() if x is not None else None for x in f]), 200.0, lumped,
              P(*[x.data_ptr() for x in got]), stream_ptr(torch_dev))
    torch.cuda.synchronize()
    for a, r in zip(got, ref):
        assert torch.equal(a, r)
    with pytest.raises(ValueError):
        _lib.call("dsv_rhs_batch", m.handle, 3, sp, P(*[x.data_ptr() for x in u]), None, 200.0,
                  lumped, P(got[0].data_ptr(), None, got[2].data_ptr()), stream_ptr(torch_dev))
    with pytest.raises(ValueError):
        _lib.call("dsv_rhs_batch", m.handle, 4, sp, P(*[x.data_ptr() for x in u]), None, 200.0,
                  lumped, P(*[x.data_ptr() for x in got]), stream_ptr(torch_dev))


def _species_csr(conn, xyz, sp, inv_dt, lumped):
    D = sp["D"] if sp.get("D") is not None else sp.get("D_value", 0.0)
    al = sp["alpha"] if sp.get("alpha") is not None else sp.get("alpha_value", 0.0)
    ip, ix, data = OA.assemble_values(conn, xyz, D, inv_dt, al, lumped, sp.get("bc_dofs"))
    return OA.to_scipy(ip, ix, data, len(xyz))


@pytest.mark.parametrize("name", ["kuhn7_jitter_morton", "disc_small", "kuhn5_shuffled"])
@pytest.mark.parametrize("lumped", [1, 0])
def test_matrix_free_apply_three_species(torch_dev, name, lumped):
    import torch
    from dsv.ceed_operators import apply_batched
    m, conn, xyz = _m(name, vertices_per_block=64)
    rng = np.random.default_rng(5)
    C, V = len(conn), len(xyz)
    bc = np.arange(0, V, 11, dtype=np.int32)
    species = [dict(D=rng.lognormal(0, 0.5, C), alpha=rng.uniform(0, 3, V), bc_dofs=bc),
               dict(D_value=2.0),
               dict(D=rng.uniform(1, 5, V), alpha_value=0.5)]
    xs = [rng.uniform(-1, 1, V) for _ in species]
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        ys = apply_batched(m, species, 200.0, xs, lumped=lumped, precision=prec)
        for s, x, y in zip(species, xs, ys):
            A = _species_csr(conn, xyz, s, 200.0, lumped)
            ref = A @ x
            assert _relerr(y.cpu().numpy(), ref) < tol, (prec, name)


@pytest.mark.parametrize("nv,threads,lanes", [(64, 256, 1), (64, 512, 2), (128, 256, 4),
                                              (128, 512, 4), (256, 512, 1), (256, 256, 4),
                                              (512, 512, 2)])
def test_matrix_free_apply_launch_shapes(torch_dev, nv, threads, lanes):
    """Every (block size, workgroup size, lanes per vertex) computes the same operator: slices
    with padded lanes, workgroups that loop over several slices and the cross-lane sums."""
    from dsv.ceed_operators import apply_batched
    m, conn, xyz = _m("disc_small")
    m.set_partition(nv, threads, lanes)
    assert m.vertices_per_block == nv
    rng = np.random.default_rng(9)
    C, V = len(conn), len(xyz)
    species = [dict(D=rng.lognormal(0, 0.5, C), alpha=rng.uniform(0, 3, V),
                    bc_dofs=np.arange(0, V, 7, dtype=np.int32)),
               dict(D_value=2.0, alpha_value=1.5),
               dict(D=rng.uniform(1, 5, V))]
    xs = [rng.uniform(-1, 1, V) for _ in species]
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        ys = apply_batched(m, species, 200.0, xs, lumped=1, precision=prec)
        for s, x, y in zip(species, xs, ys):
            ref = _species_csr(conn, xyz, s, 200.0, 1) @ x
            assert _relerr(y.cpu().numpy(), ref) < tol, (prec, nv, threads, lanes)
    with pytest.raises(ValueError):
        m.set_partition(96, 256, 4)     # not a multiple of 64
    with pytest.raises(ValueError):
        m.set_partition(128, 256, 3)    # lanes per vertex must be 1, 2 or 4
    with pytest.raises(ValueError):
        m.set_partition(512, 256, 4)    # at most 1024 lane slots per block
    with pytest.raises(ValueError):
        m.set_partition(128, 128, 4)    # workgroups of 256 or 512 threads


def _solve_case(name, lumped=1, seed=6):
    conn, xyz = CASES[name]
    rng = np.random.default_rng(seed)
    C, V = len(conn), len(xyz)
    bc = np.nonzero(xyz[:, 2] <= xyz[:, 2].min() + 1e-9)[0].astype(np.int32)
    dt = 0.005
    sps = [dict(D=rng.lognormal(0, 0.5, C), alpha=rng.uniform(0, 50, V), bc_dofs=bc, bc_value=1.0),
           dict(D_value=2.0),
           dict(D_value=5.3, f=rng.uniform(0, 1, V))]
    u0 = [rng.uniform(0, 1, V) for _ in sps]
    refs, bs = [], []
    for s, u in zip(sps, u0):
        A = _species_csr(conn, xyz, s, 1 / dt, lumped)
        b = OA.assemble_rhs(conn, xyz, u, 1 / dt, s.get("f"), lumped, s.get("bc_dofs"),
                            s.get("bc_value", 0.0))
        refs.append(OS.direct(A, b))
        bs.append(b)
    return conn, xyz, sps, u0, bs, refs, 1 / dt


@pytest.mark.parametrize("method", ["csr64", "mf64", "mf32_ir64"])
@pytest.mark.parametrize("name", ["kuhn7_jitter_morton", "disc_small"])
@pytest.mark.parametrize("lumped", [1, 0])
def test_batched_solve_matches_direct(torch_dev, method, name, lumped):
    import torch
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz, sps, u0, bs, refs, inv_dt = _solve_case(name, lumped)
    m = Mesh(conn, xyz, vertices_per_block=64)
    sp = []
    for s, u, b in zip(sps, u0, bs):
        d = dict(s)
        d.pop("f", None)
        d["b"] = torch.tensor(b, dtype=torch.float64, device=torch_dev)
        d["x"] = torch.tensor(u, dtype=torch.float64, device=torch_dev)
        sp.append(d)
    st = solve_batched(m, sp, inv_dt, rtol=1e-11, method=method, lumped=lumped)
    for k in range(3):
        assert st["converged"][k], st
        assert st["rel_residual"][k] <= 1e-11
        err = _relerr(sp[k]["x"].cpu().numpy(), refs[k])
        assert err < 1e-8, (method, k, err, st)


@pytest.mark.parametrize("nv,threads,lanes", [(64, 512, 2), (256, 256, 4), (128, 256, 1)])
@pytest.mark.parametrize("method", ["mf64", "mf32_ir64"])
def test_matrix_free_solve_launch_shapes(torch_dev, nv, threads, lanes, method):
    """The fused CG iteration (several staging and lane rounds per thread, workgroups with
    several partition blocks) converges to the direct solution on every launch shape."""
    import torch
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz, sps, u0, bs, refs, inv_dt = _solve_case("disc_small", 1, seed=8)
    m = Mesh(conn, xyz)
    m.set_partition(nv, threads, lanes)
    sp = []
    for s, u, b in zip(sps, u0, bs):
        d = dict(s)
        d.pop("f", None)
        d["b"] = torch.tensor(b, dtype=torch.float64, device=torch_dev)
        d["x"] = torch.tensor(u, dtype=torch.float64, device=torch_dev)
        sp.append(d)
    st = solve_batched(m, sp, inv_dt, rtol=1e-11, method=method, lumped=1)
    for k in range(3):
        assert st["converged"][k], st
        assert _relerr(sp[k]["x"].cpu().numpy(), refs[k]) < 1e-8, (k, st)


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("nv,threads,lanes", [(64, 512, 2), (256, 256, 4), (128, 256, 1)])
def test_inner_operator_forms(torch_dev, form, nv, threads, lanes):
    """Scalar D and lumped mass (the transport workload's case): the FP32 inner CG runs on the
    stored-geometry operator (form 1) or recomputes the Jacobians (form 0); both converge to
    the direct FP64 solution of the assembled operator."""
    import torch
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz = CASES["disc_small"]
    rng = np.random.default_rng(12)
    V = len(xyz)
    bc = np.nonzero(xyz[:, 2] <= xyz[:, 2].min() + 1e-9)[0].astype(np.int32)
    inv_dt = 200.0
    sps = [dict(D_value=2.0, alpha=rng.uniform(0, 40, V), bc_dofs=bc, bc_value=1.0),
           dict(D_value=2.0),
           dict(D_value=5.3, alpha_value=0.7)]
    m = Mesh(conn, xyz)
    m.set_partition(nv, threads, lanes)
    m.set_operator_form(form)
    sp, refs = [], []
    for s in sps:
        u = rng.uniform(0, 1, V)
        b = OA.assemble_rhs(conn, xyz, u, inv_dt, None, 1, s.get("bc_dofs"), s.get("bc_value", 0.0))
        refs.append(OS.direct(_species_csr(conn, xyz, s, inv_dt, 1), b))
        d = dict(s)
        d["b"] = torch.tensor(b, dtype=torch.float64, device=torch_dev)
        d["x"] = torch.tensor(u, dtype=torch.float64, device=torch_dev)
        sp.append(d)
    st = solve_batched(m, sp, inv_dt, rtol=1e-11, method="mf32_ir64", lumped=1)
    for k in range(3):
        assert st["converged"][k], st
        assert _relerr(sp[k]["x"].cpu().numpy(), refs[k]) < 1e-8, (k, form, st)
    with pytest.raises(ValueError):
        m.set_operator_form(2)


def test_single_species_and_warm_start(torch_dev):
    import torch
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz, sps, u0, bs, refs, inv_dt = _solve_case("kuhn7_jitter_morton")
    m = Mesh(conn, xyz)
    s = dict(sps[0])
    s["b"] = torch.tensor(bs[0], device=torch_dev)
    s["x"] = torch.tensor(u0[0], device=torch_dev)
    st1 = solve_batched(m, [s], inv_dt, rtol=1e-10, method="csr64")
    assert st1["converged"][0]
    # warm start from the solution: converged before any iteration
    st2 = solve_batched(m, [s], inv_dt, rtol=1e-10, method="csr64")
    assert st2["converged"][0] and st2["iterations"][0] == 0
    # iteration cap reports non-convergence instead of raising
    s["x"] = torch.tensor(u0[0], device=torch_dev)
    st3 = solve_batched(m, [s], inv_dt, rtol=1e-12, method="mf64", max_it=2)
    assert not st3["converged"][0]


def test_fixed_points_on_gpu(torch_dev):
    import torch
    from dsv.core import Mesh
    from dsv.solvers import solve_batched
    conn, xyz = CASES["kuhn7_jitter_morton"]
    V = len(xyz)
    m = Mesh(conn, xyz)
    dt, al = 0.01, 3.0
    u = np.full(V, 0.8)
    b = OA.assemble_rhs(conn, xyz, u, 1 / dt, None, 1)
    for method in ("csr64", "mf64", "mf32_ir64"):
        s = dict(D_value=2.0, alpha_value=al, b=torch.tensor(b, device=torch_dev),
                 x=torch.zeros(V, dtype=torch.float64, device=torch_dev))
        st = solve_batched(m, [s], 1 / dt, rtol=1e-12, method=method)
        assert _relerr(s["x"].cpu().numpy(), u / (1 + al * dt)) < 1e-10, (method, st)


def test_argument_errors(torch_dev):
    import torch
    from dsv.core import Mesh
    from dsv.gpu_assembler import GPUSparseAssembler
    from dsv.solvers import solve_batched
    conn, xyz = CASES["kuhn6"]
    m = Mesh(conn, xyz)
    asm = GPUSparseAssembler(m)
    with pytest.raises(ValueError, match="matches neither"):
        asm.assemble(np.ones(len(xyz) + 1), 1.0)
    V = len(xyz)
    s = dict(D_value=1.0, bc_dofs=np.array([V + 5], dtype=np.int32),
             b=torch.zeros(V, dtype=torch.float64, device=torch_dev),
             x=torch.zeros(V, dtype=torch.float64, device=torch_dev))
    with pytest.raises(ValueError, match="out of range"):
        solve_batched(m, [s], 1.0)
    bad = conn.copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(_lib.DsvError):
        Mesh(bad, xyz)
    with pytest.raises(ValueError):
        Mesh(np.array([[0, 1, 2, 99]], dtype=np.int32), xyz[:4])

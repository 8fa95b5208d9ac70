"""The reference's drop-in call surface on the GPU, through libdsv (ctypes): solve_transport as
the species blocks of `dissolve.py:781-954` call it (`numerics/solvers.py:206-372`),
CeedDiffusionMassOp / MatrixFreeOperator (`gpu/ceed_operators.py:247-353`,
`numerics/solvers.py:172-203`) and the shared-geometry assembler (`gpu/gpu_assembler.py:346-367`),
each against the oracle.

Bars: every solve within 1e-8 (relative to max |x|) of the oracle's direct solve of the same
step (be_step, then under_relax for O2); FP64 apply within 1e-12 and FP32 apply within 2e-5 of
the oracle's CSR product; CSR values within 1e-12."""
import numpy as np
import pytest

from dsv import meshes, workloads
from oracle import assembly as OA
from oracle import solve as OS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def disc(gpu):
    from dsv.core import Mesh
    conn, xyz = meshes.graded_disc(seed=1, scale=6.0)
    conn, xyz = meshes.locality_order(conn, xyz)
    phi = xyz[:, 2] - 0.12
    bc = np.nonzero(xyz[:, 2] <= 1e-12)[0].astype(np.int32)
    w = workloads.three_species(conn, xyz, phi, bc, 0.005, 0.05)
    return Mesh(conn, xyz), w


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("method", ["mf32_ir64", "mf64", "csr64"])
def test_solve_transport_species_blocks(disc, method):
    """Three time steps of Cl, O2 (omega = 0.3, Dirichlet through _solver_bc_dofs) and OH (source
    built from the new O2) through solve_transport, in the reference's order, against the
    oracle's direct step plus under-relaxation; from the second step on the matrix-free
    methods start from the Galerkin warm start over the cached history."""
    from dsv import fem, solvers
    mesh, w = disc
    solvers._GPU_ASM_CACHE.clear()
    conn, xyz, dt = w["conn"], w["xyz"], w["dt"]
    V = fem.FunctionSpace(mesh)
    o2s, cls, ohs = w["species"]
    fields = {}
    for s in (o2s, cls, ohs):
        u, u_old, De = fem.Function(V), fem.Function(V), fem.Function(V)
        u.x.array[:] = s["u0"]
        De.x.array[:] = s["D_value"]
        De._solver_inv_dt = 1.0 / dt
        fields[s["name"]] = (u, u_old, De)
    fields["O2"][2]._solver_bc_dofs = o2s["bc_dofs"]
    ref = {n: np.array(fields[n][0].x.array) for n in fields}
    alpha_o2 = o2s["alpha"]
    band = workloads.delta_eps(w["phi"], w["eps"]) * w["grad_norm"] * w["k_orr"]
    for step in range(3):
        # Cl: no source, no Dirichlet
        u, u_old, De = fields["Cl"]
        u_old.x.array[:] = u.x.array
        r = solvers.solve_transport("Cl", None, None, u, rtol=1e-11, use_gpu=True, de_field=De,
                                    u_old=u_old, method=method)
        ref["Cl"] = OS.be_step(conn, xyz, ref["Cl"], dt, cls["D_value"])
        assert r["converged"] and _rel(u.x.array, ref["Cl"]) <= 1e-8, (step, r)
        # O2: ORR sink, Dirichlet bulk value, under-relaxation 0.3 (u_old rebound to u)
        u, u_old, De = fields["O2"]
        u_old.x.array[:] = u.x.array
        r = solvers.solve_transport("O2", None, None, u, rtol=1e-11, under_relax=0.3,
                                    use_gpu=True, de_field=De, alpha_field=alpha_o2,
                                    u_old=u_old, bc_value=o2s["bc_value"], method=method)
        new = OS.be_step(conn, xyz, ref["O2"], dt, o2s["D_value"], alpha=alpha_o2,
                         bc_dofs=o2s["bc_dofs"], bc_value=o2s["bc_value"])
        ref["O2"] = OS.under_relax(new, ref["O2"], 0.3)
        assert r["converged"] and _rel(u.x.array, ref["O2"]) <= 1e-8, (step, r)
        # OH: source 4 delta |grad phi| k_orr O2 from the new (blended) O2
        u, u_old, De = fields["OH"]
        u_old.x.array[:] = u.x.array
        f = np.where(w["phi"] < 0.0, 4.0 * band * fields["O2"][0].x.array, 0.0)
        r = solvers.solve_transport("OH", None, None, u, rtol=1e-11, use_gpu=True, de_field=De,
                                    u_old=u_old, source_field=f, method=method)
        ref["OH"] = OS.be_step(conn, xyz, ref["OH"], dt, ohs["D_value"], f=f)
        assert r["converged"] and _rel(u.x.array, ref["OH"]) <= 1e-8, (step, r)
    hist = solvers._GPU_ASM_CACHE["O2"]["x"]
    assert hist.n == 3
    # the cached operator, assembled only when read, is the O2 operator of the last step
    ip, ix, vals = solvers._GPU_ASM_CACHE["O2"]["asm"].mat.getValuesCSR()
    _, _, ref_vals = OA.assemble_values(conn, xyz, o2s["D_value"], 1.0 / dt, alpha_o2, 1,
                                        o2s["bc_dofs"])
    assert np.abs(vals - ref_vals).max() <= 1e-12 * np.abs(ref_vals).max()
    # the driver's stale-warm-start reset after redistancing (`dissolve.py:715-719`)
    for entry in solvers._GPU_ASM_CACHE.values():
        entry["x"].set(0.0)
    assert hist.n == 0
    solvers._GPU_ASM_CACHE.clear()


def test_solve_transport_argument_errors(disc):
    from dsv import fem, solvers
    mesh, w = disc
    V = fem.FunctionSpace(mesh)
    u, De = fem.Function(V), fem.Function(V)
    De.x.array[:] = 1.0
    with pytest.raises(ValueError):
        solvers.solve_transport("X", None, None, u, use_gpu=True, de_field=De, u_old=u,
                                method="nope")
    with pytest.raises(ValueError):
        solvers.solve_transport("X", None, None, u, use_gpu=True, de_field=De, u_old=u,
                                warm_start="nope")
    with pytest.raises(ValueError):
        solvers.solve_transport("X", None, None, u, use_gpu=True, de_field=De,
                                u_old=np.zeros(3))
    solvers._GPU_ASM_CACHE.clear()


@pytest.mark.parametrize("lumped", [1, 0])
def test_ceed_operator_apply_and_update(disc, lumped):
    """CeedDiffusionMassOp.apply / update_coeffs and MatrixFreeOperator.mult against the
    oracle's assembled operator times x (same operator: alpha and Dirichlet rows included)."""
    import torch
    from dsv.ceed_operators import CeedDiffusionMassOp, extract_ceed_mesh_data, make_ceed
    from dsv.solvers import MatrixFreeOperator
    mesh, w = disc
    conn, xyz = w["conn"], w["xyz"]
    rng = np.random.default_rng(7)
    C, V = len(conn), len(xyz)
    ceed, name = make_ceed(True)
    md = extract_ceed_mesh_data(mesh)
    D1, al = rng.lognormal(0, 0.5, C), rng.uniform(0, 3, V)
    bc = np.arange(0, V, 23, dtype=np.int32)
    x = rng.uniform(-1, 1, V)
    for prec, tol in ((64, 1e-12), (32, 2e-5)):
        op = CeedDiffusionMassOp(ceed, md, D1, 200.0, alpha=al, bc_dofs=bc, lumped=lumped,
                                 precision=prec)
        ip, ix, data = OA.assemble_values(conn, xyz, D1, 200.0, al, lumped, bc)
        ref = OA.to_scipy(ip, ix, data, V) @ x
        assert _rel(op.apply(x).cpu().numpy(), ref) <= tol
        # new coefficients: a per-vertex D and another 1/dt
        D2 = rng.lognormal(0, 0.3, V)
        op.update_coeffs(D2, 50.0)
        ip, ix, data = OA.assemble_values(conn, xyz, D2, 50.0, al, lumped, bc)
        ref = OA.to_scipy(ip, ix, data, V) @ x
        out_host = np.zeros(V)
        y = op.apply(torch.tensor(x, device=mesh.device), out_host)
        assert _rel(y.cpu().numpy(), ref) <= tol and _rel(out_host, ref) <= tol
        # the reference's MatShell hook: mult(A, x, y)
        mf = MatrixFreeOperator(op, V)
        y_host = np.zeros(V)
        mf.mult(None, x, y_host)
        assert _rel(y_host, ref) <= tol
        with pytest.raises(ValueError):
            op.update_coeffs(np.ones(V + 1), 50.0)


def test_shared_geometry_assembler(disc):
    """GPUSparseAssembler(ref_asm=primary) borrows the primary's mesh arrays (no new
    geometry) and assembles its own values; solve_transport's cached assemblers borrow from
    the geometry registered with set_shared_geometry."""
    from dsv import fem, solvers
    from dsv.gpu_assembler import GPUSparseAssembler
    mesh, w = disc
    conn, xyz = w["conn"], w["xyz"]
    V = len(xyz)
    primary = GPUSparseAssembler(mesh)
    field = GPUSparseAssembler(None, ref_asm=primary)
    assert field.mesh is primary.mesh and field.d_indptr.data_ptr() == primary.d_indptr.data_ptr()
    assert field.d_data.data_ptr() != primary.d_data.data_ptr()
    rng = np.random.default_rng(3)
    D = rng.lognormal(0, 0.5, V)
    primary.assemble(np.ones(V), 100.0)
    field.assemble(D, 200.0, alpha=np.full(V, 0.5))
    _, _, ref = OA.assemble_values(conn, xyz, D, 200.0, np.full(V, 0.5), 1)
    assert np.abs(field.d_data.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    _, _, ref1 = OA.assemble_values(conn, xyz, np.ones(V), 100.0, None, 1)
    assert np.abs(primary.d_data.cpu().numpy() - ref1).max() <= 1e-12 * np.abs(ref1).max()
    solvers._GPU_ASM_CACHE.clear()
    solvers.set_shared_geometry(primary)
    try:
        Vs = fem.FunctionSpace(mesh)
        u, De = fem.Function(Vs), fem.Function(Vs)
        u.x.array[:] = rng.uniform(0, 1, V)
        De.x.array[:] = 2.0
        De._solver_inv_dt = 200.0
        u_old = np.array(u.x.array)
        r = solvers.solve_transport("Cl", None, None, u, rtol=1e-11, use_gpu=True, de_field=De,
                                    u_old=u_old)
        assert solvers._GPU_ASM_CACHE["Cl"]["asm"].mesh is primary.mesh
        assert r["converged"] and _rel(u.x.array, OS.be_step(conn, xyz, u_old, 0.005, 2.0)) <= 1e-8
    finally:
        solvers.set_shared_geometry(None)
        solvers._GPU_ASM_CACHE.clear()


def test_load_mesh_from_xdmf_hdf5(gpu):
    """dsv.meshio.load_mesh on the reference's XDMF + HDF5 layout (a file the HDF5 library
    wrote): the device mesh's CSR pattern is bit-exact and its assembly within 1e-12 of the
    oracle on the arrays of the file; the cell tags are the file's region labels."""
    import os
    from dsv import meshio
    from dsv.core import VIEW_INDICES, VIEW_INDPTR
    from dsv.gpu_assembler import GPUSparseAssembler
    h5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")
    exp = np.load(os.path.join(h5, "mesh_expected.npz"), allow_pickle=False)
    mesh, tags = meshio.load_mesh(os.path.join(h5, "mesh_default.xdmf"))
    conn, xyz = exp["cells"].astype(np.int32), exp["points"]
    ip, ix = OA.pattern(conn, len(xyz))
    assert np.array_equal(mesh.view(VIEW_INDPTR).cpu().numpy(), ip)
    assert np.array_equal(mesh.view(VIEW_INDICES).cpu().numpy(), ix)
    assert np.array_equal(tags, exp["cell_data"])
    asm = GPUSparseAssembler(mesh)
    asm.assemble(np.full(len(xyz), 1.5), 40.0)
    _, _, ref = OA.assemble_values(conn, xyz, 1.5, 40.0, None, 1)
    assert np.abs(asm.d_data.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()

"""Pins the oracle's assembly against closed forms, invariants and brute force (CPU)."""
import numpy as np
import pytest

from oracle import assembly as A
from dsv import meshes

UNIT = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
CONN1 = np.array([[0, 1, 2, 3]], dtype=np.int32)


def _elem(xyz, De=1.0, alpha=0.0, inv_dt=0.0, lumped=1):
    detJ, G, _ = A.geometry(CONN1, xyz)
    return A.element_matrices(detJ, G, np.array([De]), np.array([alpha]), inv_dt, lumped)[0]


def test_unit_tet_stiffness_closed_form():
    K = _elem(UNIT, De=1.0)
    ref = np.array([[3.0, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6.0
    np.testing.assert_allclose(K, ref, atol=1e-15)
    np.testing.assert_allclose(_elem(UNIT, De=2.5), 2.5 * ref, atol=1e-15)


def test_unit_tet_mass_closed_forms():
    # A = inv_dt * M when De = 0
    Mc = _elem(UNIT, De=0.0, inv_dt=1.0, lumped=0)
    np.testing.assert_allclose(Mc, (np.ones((4, 4)) + np.eye(4)) / 120.0, atol=1e-16)
    Ml = _elem(UNIT, De=0.0, inv_dt=1.0, lumped=1)
    np.testing.assert_allclose(Ml, np.eye(4) / 24.0, atol=1e-16)
    # reaction adds to inv_dt
    np.testing.assert_allclose(_elem(UNIT, De=0.0, alpha=3.0, inv_dt=1.0), 4.0 * np.eye(4) / 24.0)
    # 1^T M 1 = volume for both masses
    assert abs(Mc.sum() - 1.0 / 6.0) < 1e-15 and abs(Ml.sum() - 1.0 / 6.0) < 1e-15


def test_stiffness_invariants_random_tets():
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = rng.normal(size=(4, 3))
        K = _elem(x)
        np.testing.assert_allclose(K @ np.ones(4), 0.0, atol=1e-10 * np.abs(K).max())
        np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())
        assert np.linalg.eigvalsh(K).min() > -1e-10 * np.abs(K).max()
        # translation and rotation invariance, scaling K ~ s, M ~ s^3
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x2 = x @ Q.T + rng.normal(size=3)
        np.testing.assert_allclose(_elem(x2), K, rtol=1e-9, atol=1e-12 * np.abs(K).max())
        np.testing.assert_allclose(_elem(2.0 * x), 2.0 * K, rtol=1e-9, atol=1e-12)
        M = _elem(x, De=0.0, inv_dt=1.0, lumped=0)
        np.testing.assert_allclose(_elem(2.0 * x, De=0.0, inv_dt=1.0, lumped=0), 8.0 * M, rtol=1e-12)


def test_vertex_order_and_orientation_do_not_matter():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 3))
    K = _elem(x)
    perm = [2, 0, 3, 1]
    K2 = _elem(x[perm])
    np.testing.assert_allclose(K2, K[np.ix_(perm, perm)], rtol=1e-9, atol=1e-12)


def _tiny():
    conn, xyz = meshes.kuhn_box(2, 1, 2)
    rng = np.random.default_rng(3)
    xyz = xyz + rng.uniform(-0.05, 0.05, xyz.shape)
    return conn, xyz


@pytest.mark.parametrize("lumped", [0, 1])
@pytest.mark.parametrize("coef", ["cell", "vertex"])
def test_assembly_matches_bruteforce_dense(lumped, coef):
    conn, xyz = _tiny()
    rng = np.random.default_rng(4)
    C, V = len(conn), len(xyz)
    D = rng.uniform(0.5, 2.0, C if coef == "cell" else V)
    al = rng.uniform(0.0, 1.0, C if coef == "cell" else V)
    ip, ix, data = A.assemble_values(conn, xyz, D, 7.0, al, lumped)
    dense = A.to_scipy(ip, ix, data, V).toarray()
    Dc = A.cell_average(D, conn, C, V)
    ac = A.cell_average(al, conn, C, V)
    ref = A.assemble_dense_bruteforce(conn, xyz, Dc, ac, 7.0, lumped)
    np.testing.assert_allclose(dense, ref, rtol=1e-13, atol=1e-15)


def test_cell_average_rules():
    conn, xyz = _tiny()
    C, V = len(conn), len(xyz)
    assert np.all(A.cell_average(None, conn, C, V) == 0.0)
    v = np.arange(V, dtype=float)
    np.testing.assert_allclose(A.cell_average(v, conn, C, V), v[conn].mean(1))
    c = np.arange(C, dtype=float)
    np.testing.assert_array_equal(A.cell_average(c, conn, C, V), c)
    with pytest.raises(ValueError, match="matches neither"):
        A.cell_average(np.ones(V + 1), conn, C, V)


def test_pattern_bruteforce_and_scatter_map():
    conn, xyz = _tiny()
    V = len(xyz)
    ip, ix = A.pattern(conn, V)
    ip2, ix2 = A.pattern_bruteforce(conn, V)
    np.testing.assert_array_equal(ip, ip2)
    np.testing.assert_array_equal(ix, ix2)
    smap = A.scatter_map(conn, ip, ix)
    rows = np.repeat(np.arange(V), np.diff(ip))
    for e in range(len(conn)):
        for i in range(4):
            for j in range(4):
                p = smap[16 * e + 4 * i + j]
                assert rows[p] == conn[e, i] and ix[p] == conn[e, j]


@pytest.mark.parametrize("n,nnz", [(20, 128581), (48, 1707697)])
def test_kuhn_nnz_closed_form(n, nnz):
    conn, xyz = meshes.kuhn_cube(n)
    ip, ix = A.pattern(conn, len(xyz))
    E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    assert int(ip[-1]) == nnz == len(xyz) + 2 * E


def test_patch_test_linear_field():
    # (K u)_i = 0 at interior vertices for a linear u (P1 reproduces linears)
    conn, xyz = meshes.kuhn_cube(4)
    rng = np.random.default_rng(5)
    interior = ~meshes.boundary_vertex_mask(xyz)
    xyz = meshes.jitter(xyz, 0.25, 0.2, rng, interior)
    ip, ix, data = A.assemble_values(conn, xyz, 1.3, 0.0, None, 1)
    K = A.to_scipy(ip, ix, data, len(xyz))
    u = xyz @ np.array([0.3, -1.2, 2.0]) + 0.7
    np.testing.assert_allclose((K @ u)[interior], 0.0, atol=1e-12)


def test_dirichlet_rows_and_rhs():
    conn, xyz = _tiny()
    V = len(xyz)
    bc = np.array([0, 3], dtype=np.int32)
    ip, ix, data = A.assemble_values(conn, xyz, 1.0, 5.0, None, 1, bc_dofs=bc)
    Ad = A.to_scipy(ip, ix, data, V).toarray()
    for r in bc:
        e = np.zeros(V)
        e[r] = 1.0
        np.testing.assert_array_equal(Ad[r], e)
    # lumped RHS: b = m (inv_dt u + f); consistent RHS of a constant u = M_c 1 u / dt = m u / dt
    u = np.full(V, 2.0)
    f = np.linspace(0, 1, V)
    m = A.lumped_mass(conn, xyz)
    np.testing.assert_allclose(A.assemble_rhs(conn, xyz, u, 5.0, f, 1), m * (5.0 * u + f), rtol=1e-14)
    np.testing.assert_allclose(A.assemble_rhs(conn, xyz, u, 5.0, None, 0), m * 10.0, rtol=1e-13)
    b = A.assemble_rhs(conn, xyz, u, 5.0, f, 1, bc, 9.0)
    assert np.all(b[bc] == 9.0)
    with pytest.raises(ValueError):
        A.assemble_rhs(conn, xyz, u[:-1], 5.0)
    # total mass = volume
    assert abs(m.sum() - np.abs(meshes.tet_signed_volume(conn, xyz)).sum()) < 1e-14


def _jittered(n=4, seed=11):
    conn, xyz = meshes.kuhn_cube(n)
    rng = np.random.default_rng(seed)
    interior = ~meshes.boundary_vertex_mask(xyz)
    return conn, meshes.jitter(xyz, 1.0 / n, 0.2, rng, interior), rng


def test_rhs_equals_mass_matrix_times_fields():
    # b(u, f) = M (inv_dt u + f) with M the consistent (lumped) mass matrix assembled by
    # assemble_values(De=0, inv_dt=1, alpha=None), whose element matrices the unit-tet closed
    # forms above pin; non-linear u and f exercise every term of both RHS branches
    conn, xyz, rng = _jittered()
    V = len(xyz)
    u = np.sin(3.0 * xyz[:, 0]) * np.exp(xyz[:, 1]) + rng.uniform(0, 1, V)
    f = xyz[:, 2] ** 3 - 2.0 * xyz[:, 0] * xyz[:, 1] + rng.normal(size=V)
    inv_dt = 7.0
    for lumped in (0, 1):
        ip, ix, data = A.assemble_values(conn, xyz, 0.0, 1.0, None, lumped)
        M = A.to_scipy(ip, ix, data, V)
        for uu, ff in ((u, f), (np.zeros(V), f), (u, None)):
            ref = M @ (inv_dt * uu + (ff if ff is not None else 0.0))
            b = A.assemble_rhs(conn, xyz, uu, inv_dt, ff, lumped)
            np.testing.assert_allclose(b, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max())
    # the lumped mass is the row sum of the consistent one
    ip, ix, data = A.assemble_values(conn, xyz, 0.0, 1.0, None, 0)
    Mc = A.to_scipy(ip, ix, data, V)
    np.testing.assert_allclose(np.asarray(Mc.sum(axis=1)).ravel(), A.lumped_mass(conn, xyz),
                               rtol=1e-13)


def _cupy_searchsorted_stdlib(a, key):
    """CuPy's per-key search written with the standard library's bisect_left (CPython's
    textbook bisection: mid = (lo + hi) // 2, a[mid] < key -> lo = mid + 1, else hi = mid)."""
    import bisect
    if key > a[-1]:
        return len(a)
    return bisect.bisect_left(a, key, 0, len(a) - 1)


def test_per_key_searchsorted_matches_stdlib_bisect():
    conn, xyz, rng = _jittered(3, seed=2)
    keys = np.arange(len(xyz), dtype=np.int32)
    # sorted input: the usual insertion point
    srt = np.sort(conn.ravel())
    np.testing.assert_array_equal(A.searchsorted_left_per_key(srt, keys), np.searchsorted(srt, keys))
    # unsorted flat connectivity, as the reference passes it, in several orders
    for trial in range(4):
        c = conn[rng.permutation(len(conn))] if trial else conn
        flat = c.ravel()
        lst = flat.tolist()
        ref = np.array([_cupy_searchsorted_stdlib(lst, int(k)) for k in keys])
        np.testing.assert_array_equal(A.searchsorted_left_per_key(flat, keys), ref)
    # keys above every entry go to len(a), as do keys above the last entry of an unsorted array
    a = np.array([5, 1, 9, 3])
    np.testing.assert_array_equal(A.searchsorted_left_per_key(a, np.array([10, 4, 0])),
                                  [_cupy_searchsorted_stdlib(a.tolist(), k) for k in (10, 4, 0)])


def test_per_cell_rhs_source_uses_per_key_search():
    # the per-cell branch maps vertex v to f[p mod C] with p the per-key search of v
    conn, xyz, rng = _jittered(3, seed=4)
    conn = conn[rng.permutation(len(conn))]
    C, V = len(conn), len(xyz)
    fc = rng.normal(size=C)
    lst = conn.ravel().tolist()
    fv = np.array([fc[_cupy_searchsorted_stdlib(lst, v) % C] for v in range(V)])
    u = rng.uniform(size=V)
    for lumped in (0, 1):
        np.testing.assert_array_equal(A.assemble_rhs(conn, xyz, u, 3.0, fc, lumped),
                                      A.assemble_rhs(conn, xyz, u, 3.0, fv, lumped))

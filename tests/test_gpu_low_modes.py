"""The low-mode subspace of the Galerkin-started solves (dsv_build_low_modes), through the C-ABI.

Bars: the lowest Rayleigh-Ritz values of every species within 1e-4 relative of the lowest
eigenvalues of D^-1 A on the non-Dirichlet vertices (scipy, shift-invert); every step solved with
the correction within 1e-8 of scipy splu of the same system; fewer CG iterations than the same
steps without it; the correction dropped as soon as the operator's epoch changes; argument
errors refused."""
import numpy as np
import pytest

from dsv import workloads
from oracle import assembly as OA

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    from dsv.core import Mesh
    w = workloads.disc3(seed=1, scale=6.0)
    return Mesh(w["conn"], w["xyz"], vertices_per_block=128), w


def _matrix(w, s):
    D = s["D"] if s.get("D") is not None else s.get("D_value", 0.0)
    al = s["alpha"] if s.get("alpha") is not None else s.get("alpha_value", 0.0)
    ip, ix, data = OA.assemble_values(w["conn"], w["xyz"], D, 1.0 / w["dt"], al, 1,
                                      s.get("bc_dofs"))
    return OA.to_scipy(ip, ix, data, len(w["xyz"]))


def _stepper(mesh, w, rtol=1e-8):
    from dsv.transport import Stepper
    st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=rtol)
    st.set_warm_start("extrapolate", 6)
    return st


def test_ritz_values_are_the_lowest_eigenvalues(case):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    mesh, w = case
    st = _stepper(mesh, w)
    info = st.build_low_modes(16, 80, 4)
    assert info["seconds"] > 0
    for k, s in enumerate(w["species"]):
        A = _matrix(w, s).tocsr()
        keep = np.ones(A.shape[0], bool)
        if s.get("bc_dofs") is not None:
            keep[s["bc_dofs"]] = False
        Aii = A[keep][:, keep]
        d = 1.0 / np.sqrt(Aii.diagonal())
        As = (sp.diags(d) @ Aii @ sp.diags(d)).tocsc()
        ev = np.sort(sla.eigsh(As, k=4, sigma=0, which="LM", return_eigenvectors=False))
        ritz = np.asarray(info["ritz"][k])
        assert np.all(np.diff(ritz) >= 0), k
        assert np.abs(ritz[:4] - ev).max() <= 1e-4 * ev.max(), (k, ritz[:4], ev)


def test_steps_match_direct_solves_with_fewer_iterations(case):
    import scipy.sparse.linalg as sla
    mesh, w = case
    lus = [sla.splu(_matrix(w, s).tocsc()) for s in w["species"]]
    totals = []
    for with_modes in (False, True):
        st = _stepper(mesh, w, rtol=1e-11)
        if with_modes:
            st.build_low_modes(32, 80, 4)
        total = 0
        for n in range(6):
            h = st.step()
            assert all(h["converged"]), h
            total += sum(h["iterations"])
            if with_modes and n >= 1:  # a Galerkin start from the second step on
                assert h["low_modes"] > 0, h
            for k, lu in enumerate(lus):
                ref = lu.solve(st.b[k].cpu().numpy())
                got = st.u[k].cpu().numpy()
                assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max(), (with_modes, n, k)
        totals.append(total)
    assert totals[1] < 0.9 * totals[0], totals


def test_dropped_when_the_operator_changes(case):
    mesh, w = case
    st = _stepper(mesh, w)
    st.build_low_modes(8, 40, 2)
    st.step()
    assert st.step()["low_modes"] > 0
    st.refresh_coefficients()  # a new operator epoch: the subspace belongs to the old one
    h = st.step()
    assert h["low_modes"] == 0 and all(h["converged"])


def test_argument_errors(case):
    import ctypes
    from dsv import _lib
    from dsv.core import species_array, species_struct, stream_ptr
    from dsv.transport import Stepper
    mesh, w = case
    st = _stepper(mesh, w)
    for n in (6, 68, -4):
        with pytest.raises(ValueError):
            st.build_low_modes(n)
    with pytest.raises(ValueError):
        st.build_low_modes(8, degree=0)
    with pytest.raises(ValueError):
        Stepper(mesh, w["species"], w["dt"], method="csr64").build_low_modes(8)
    # the library refuses a build while operator reuse is off
    mesh.set_operator_reuse(False)
    keep = []
    arr = species_array([species_struct(mesh, D_value=2.0, keep=keep)])
    with pytest.raises(ValueError, match="reuse"):
        _lib.call("dsv_build_low_modes", mesh.handle, 1, arr, 1.0 / w["dt"], 1, 8, 40, 2, 0,
                  (ctypes.c_double * 8)(), stream_ptr(mesh.device))


@pytest.mark.parametrize("sel", [[0], [1, 2]])
def test_fewer_species(case, sel):
    """The one- and two-species kernels of the subspace (its FP16 storage holds every species of
    a vertex group together): steps within 1e-8 of scipy splu, the correction taken from the
    second step on."""
    import scipy.sparse.linalg as sla
    from dsv.transport import Stepper
    mesh, w = case
    species = [w["species"][k] for k in sel]
    lus = [sla.splu(_matrix(w, s).tocsc()) for s in species]
    st = Stepper(mesh, species, w["dt"], method="mf32_ir64", rtol=1e-11)
    st.set_warm_start("extrapolate", 6)
    info = st.build_low_modes(16, 80, 4)
    assert len(info["ritz"]) == len(sel)
    for n in range(4):
        h = st.step()
        assert all(h["converged"]), h
        if n >= 1:
            assert h["low_modes"] > 0, h
        for k, lu in enumerate(lus):
            ref = lu.solve(st.b[k].cpu().numpy())
            got = st.u[k].cpu().numpy()
            assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max(), (sel, n, k)

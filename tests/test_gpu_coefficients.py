"""GPU coefficient fields (dsv_nodal_gradient, dsv_transport_coefficients) against the CPU oracle
(oracle/coefficients.py) through the C-ABI.  Bars: nodal gradient within 1e-12 of the largest
component (the gather order differs from the reference's atomics only by rounding), exact for a
linear field to 1e-11; elementwise fields within 1e-13 relative."""
import types

import numpy as np
import pytest

from dsv import meshes
from oracle import coefficients as OC

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(7)
    c1, x1 = meshes.kuhn_cube(7)
    x1 = meshes.jitter(x1, 1 / 7, 0.2, rng, ~meshes.boundary_vertex_mask(x1))
    c2, x2 = meshes.graded_disc(seed=1, scale=6.0)
    c3, x3 = meshes.kuhn_cube(5)
    p = rng.permutation(len(x3))
    inv = np.empty_like(p)
    inv[p] = np.arange(len(p))
    return {"kuhn7_jitter": (c1, x1), "disc_small": (c2, x2),
            "kuhn5_shuffled": (inv[c3][rng.permutation(len(c3))].astype(np.int32), x3[p])}


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_nodal_gradient_matches_oracle(gpu, name):
    from dsv.coefficients import nodal_gradient
    from dsv.core import Mesh
    from dsv.gpu_assembler import GPUSparseAssembler
    conn, xyz = CASES[name]
    m = Mesh(conn, xyz)
    rng = np.random.default_rng(3)
    u = np.sin(3 * xyz[:, 0]) * np.cos(2 * xyz[:, 1]) + xyz[:, 2] ** 2 + rng.normal(0, 0.01, len(xyz))
    g, n = nodal_gradient(m, u, with_norm=True)
    ref = OC.nodal_gradient(conn, xyz, u)
    assert np.abs(g.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(n.cpu().numpy() - OC.grad_norm(ref)).max() <= 1e-12 * OC.grad_norm(ref).max()
    a = np.array([0.7, -1.3, 2.1])
    gl = GPUSparseAssembler(m).nodal_gradient(xyz @ a + 0.4).cpu().numpy()
    assert np.abs(gl - a).max() < 1e-11


def test_transport_coefficients_match_oracle(gpu):
    from dsv.coefficients import transport_coefficients
    from dsv.core import Mesh
    conn, xyz = CASES["disc_small"]
    m = Mesh(conn, xyz)
    rng = np.random.default_rng(5)
    V = len(xyz)
    phi = rng.uniform(-0.03, 0.03, V)
    phi[::7] = 0.0
    gn = rng.uniform(0.5, 2.0, V)
    F, Cl, O2 = rng.uniform(-0.2, 1.5, V), rng.uniform(0, 1, V), rng.uniform(0, 1, V)
    prm = dict(eps=0.011, k_orr=2.5, kd=0.3, Fmax=1.2, film_floor=0.0625)
    names = ("alpha_o2", "f_oh", "film_factor", "medium_factor")
    got = transport_coefficients(m, phi, gn, prm["eps"], prm["k_orr"], prm["kd"], F, Cl, O2,
                                 prm["Fmax"], prm["film_floor"], outputs=names)
    ref = OC.transport_coefficients(phi, gn, F, Cl, O2, **prm)
    for k, r in zip(names, ref):
        assert np.abs(got[k].cpu().numpy() - r).max() <= 1e-13 * max(np.abs(r).max(), 1e-30), k
    # the same into caller tensors (out=); a tensor of the wrong length is refused
    import torch
    buf = {"f_oh": torch.full((len(xyz),), np.nan, dtype=torch.float64, device=m.device)}
    got1 = transport_coefficients(m, phi, gn, prm["eps"], prm["k_orr"], prm["kd"], F, Cl, O2,
                                  prm["Fmax"], prm["film_floor"], outputs=names, out=buf)
    assert got1["f_oh"] is buf["f_oh"]
    for k in names:
        assert torch.equal(got1[k], got[k]), k
    with pytest.raises(ValueError):
        transport_coefficients(m, phi, gn, prm["eps"], outputs=("f_oh",),
                               out={"f_oh": buf["f_oh"][:-1]})
    # absent film / chloride / oxygen are zero
    got0 = transport_coefficients(m, phi, gn, prm["eps"], prm["k_orr"], prm["kd"], outputs=names)
    ref0 = OC.transport_coefficients(phi, gn, None, None, None, **prm)
    for k, r in zip(names, ref0):
        assert np.abs(got0[k].cpu().numpy() - r).max() <= 1e-13 * max(np.abs(r).max(), 1e-30), k


def test_reference_surface_functions(gpu):
    from dsv import fem
    from dsv.coefficients import compute_effective_diffusion, compute_norm_grad_phi
    from dsv.core import Mesh
    conn, xyz = CASES["kuhn7_jitter"]
    m = Mesh(conn, xyz)
    Vs = fem.FunctionSpace(m)
    phi, ng, F = fem.Function(Vs, "phi"), fem.Function(Vs, "ng"), fem.Function(Vs, "F")
    phi.x.array[:] = xyz[:, 2] - 0.45
    F.x.array[:] = np.linspace(0, 2, len(xyz))
    compute_norm_grad_phi(phi, ng)
    assert np.abs(ng.x.array - OC.grad_norm(OC.nodal_gradient(conn, xyz, phi.x.array))).max() < 1e-12
    params = types.SimpleNamespace(Fmax=1.0, film_tortuosity=2.0, film_porosity=0.25,
                                   film_constrictivity=1.0, diff_zn=1.1, diff_cl=2.0,
                                   diff_oh=5.3, diff_o2=2.0)
    out = [fem.Function(Vs, n) for n in ("DeZn", "DeCl", "DeOH", "DeO2")]
    compute_effective_diffusion(phi, F, params, *out)
    _, _, film, med = OC.transport_coefficients(phi.x.array, np.zeros(len(xyz)), F.x.array, None,
                                                None, 1.0, 0.0, 0.0, 1.0, 0.0625)
    for fn, d0, f in zip(out, (1.1, 2.0, 5.3, 2.0), (film, med, med, med)):
        assert np.allclose(fn.x.array, d0 * f, rtol=1e-14, atol=0)
    # the device form against the oracle's effective_diffusion, with and without a film
    from dsv.coefficients import effective_diffusion
    from oracle import physics as OP
    D0 = (0.7, 2.0, 5.3, 2.0)
    for Fa in (F.x.array, None):
        got = effective_diffusion(m, phi.x.array, Fa, 1.0, 0.0625, D0)
        ref = OP.effective_diffusion(phi.x.array, np.zeros(len(xyz)) if Fa is None else Fa, 1.0,
                                     0.0625, D0)
        for g, r in zip(got, ref):
            assert np.array_equal(g.cpu().numpy(), r)
    # the reference's bounds check: a non-finite film gives non-finite diffusivities
    F.x.array[3] = np.nan
    with pytest.raises(RuntimeError, match="NaN/Inf"):
        compute_effective_diffusion(phi, F, params, *out)


def test_argument_errors(gpu):
    from dsv import _lib
    from dsv.coefficients import transport_coefficients
    from dsv.core import Mesh
    conn, xyz = CASES["kuhn7_jitter"]
    m = Mesh(conn, xyz)
    with pytest.raises(ValueError):
        transport_coefficients(m, np.zeros(len(xyz)), np.ones(len(xyz)), 0.0)  # eps <= 0
    with pytest.raises(ValueError):
        transport_coefficients(m, np.zeros(len(xyz) - 1), np.ones(len(xyz)), 0.1)
    with pytest.raises(ValueError):
        _lib.call("dsv_nodal_gradient", m.handle, None, None, None, None)

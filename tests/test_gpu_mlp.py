"""GPU kernels of the learned warm start against the CPU oracle (oracle/mlp.py), and the gain
of the trained network on a solver-generated time series.

Bars: forward within 1e-5 of the FP64 oracle (absolute, outputs O(1)); gradient within 1e-4 of
the largest gradient component; Adam within FP32 rounding; features within FP32 rounding.
"""
import numpy as np
import pytest

from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _net(seed=0, n=1000):
    rng = np.random.default_rng(seed)
    p = OM.init(seed)
    p[OM.OFF_B1:OM.OFF_W2] = rng.normal(0, 0.3, OM.NH)
    p[OM.OFF_B2:OM.OFF_W3] = rng.normal(0, 0.3, OM.NH)
    X = rng.normal(0, 1, (n, OM.NIN)).astype(np.float32)
    y = rng.normal(0, 1, n).astype(np.float32)
    return p.astype(np.float32), X, y


@pytest.fixture(scope="module")
def dev(gpu):
    return gpu


def _mlp(dev, p):
    import torch
    from dsv.mlp import WarmStartMLP
    m = WarmStartMLP(dev)
    m.params.copy_(torch.as_tensor(p))
    return m


@pytest.mark.parametrize("n", [1, 63, 1000, 70001])
def test_forward_and_gradient_match_oracle(dev, n):
    import torch
    p, X, y = _net(1, n)
    m = _mlp(dev, p)
    Xt, yt = torch.as_tensor(X, device=dev), torch.as_tensor(y, device=dev)
    out = m.forward(Xt).cpu().numpy()
    p64, X64, y64 = p.astype(np.float64), X.astype(np.float64), y.astype(np.float64)
    assert np.abs(out - OM.forward(p64, X64)).max() < 1e-5
    g, loss = m.grad(Xt, yt)
    gref = OM.grad(p64, X64, y64)
    assert np.abs(g.cpu().numpy() - gref).max() <= 1e-4 * np.abs(gref).max()
    assert abs(loss.item() - OM.loss(p64, X64, y64)) <= 1e-5 * OM.loss(p64, X64, y64)
    # deterministic: a second evaluation is bit-identical
    g2, _ = m.grad(Xt, yt)
    assert torch.equal(g, g2)


def test_adam_matches_oracle(dev):
    import torch
    p, X, y = _net(2, 500)
    m = _mlp(dev, p)
    Xt, yt = torch.as_tensor(X, device=dev), torch.as_tensor(y, device=dev)
    pr, mr, vr = p.astype(np.float64), np.zeros(OM.NPARAM), np.zeros(OM.NPARAM)
    for step in range(1, 4):
        g, _ = m.grad(Xt, yt)
        gn = g.cpu().numpy().astype(np.float64)
        m.adam(g)
        pr, mr, vr = OM.adam_step(pr, gn, mr, vr, step, m.lr)
        assert np.abs(m.params.cpu().numpy() - pr).max() < 1e-6


def test_features_match_oracle(dev):
    import torch
    from dsv.mlp import WarmStartMLP
    rng = np.random.default_rng(3)
    n = 5000
    u, up = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    phi, gp, al = rng.normal(0, 0.05, n), rng.uniform(0, 6, n), rng.uniform(0, 40, n)
    bc = (rng.uniform(0, 1, n) < 0.1).astype(np.uint8)
    t = lambda a, dt=torch.float64: torch.as_tensor(a, dtype=dt, device=dev)
    m = WarmStartMLP(dev)
    X, sigma = m.features(t(u), t(up), t(phi), t(gp), 2.0, 0.005, 0.011, alpha=t(al),
                          bc_mask=t(bc, torch.uint8))
    Xr, sr = OM.features(u, up, phi, gp, al, 2.0, 0.005, 0.011, bc)
    assert abs(sigma.item() - sr) <= 1e-7 * sr
    assert np.abs(X.cpu().numpy() - Xr).max() < 2e-6
    X2, _ = m.features(t(u), t(up), t(phi), t(gp), 2.0, 0.005, 0.011, alpha_value=1.5)
    assert np.allclose(X2.cpu().numpy()[:, 5], 1.5 * 0.005) and X2.cpu().numpy()[:, 6].max() == 0


def test_training_reduces_loss(dev):
    import torch
    from dsv.mlp import WarmStartMLP
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (20000, OM.NIN)).astype(np.float32)
    y = (0.8 * X[:, 1] + 0.2 * np.tanh(X[:, 2]) * X[:, 0]).astype(np.float32)
    m = WarmStartMLP(dev, lr=3e-3)
    hist = m.fit(torch.as_tensor(X, device=dev), torch.as_tensor(y, device=dev), epochs=15,
                 batch=2048)
    assert hist[-1] < 0.02 * float(np.mean(y * y)), hist


def test_learned_warm_start_saves_iterations(dev):
    """Train on the first steps of a 3-species graded-disc run, then compare the following steps
    against the previous-solution start: fewer CG iterations, every step's solutions (rtol
    1e-10) within 1e-8 of the plain FP64 path's rtol-1e-13 solution of the same system, the
    learned direction used by the Galerkin combination."""
    import torch
    from dsv import workloads
    from dsv.core import Mesh
    from dsv.mlp import WarmStartMLP
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    mesh = Mesh(w["conn"], w["xyz"])
    kw = dict(method="mf32_ir64", rtol=1e-10, phi=w["phi"], grad_phi=w["grad_norm"],
              eps=w["eps"])
    plain = Stepper(mesh, w["species"], w["dt"], **kw)
    learned = Stepper(mesh, w["species"], w["dt"], mlp=WarmStartMLP(mesh.device), **kw)
    n_train, n_eval = 8, 6
    for _ in range(n_train):
        plain.step()
    learned.set_warm_start("extrapolate", history_dirs=2)
    hist = learned.train_warm_start(n_train, epochs=30)
    assert hist[-1] < hist[0] and learned.warm == "learned"
    it_plain = it_learned = 0
    for _ in range(n_eval):
        it_plain += sum(plain.step()["iterations"])
        it_learned += sum(learned.step()["iterations"])
        for st in (plain, learned):
            refs, hr = st.fp64_reference(1e-13)
            assert all(hr["converged"]) and max(st.step_error(refs)) <= 1e-8, st.step_error(refs)
    assert all(all(h["converged"]) for h in learned.history)
    log = learned.warm_log[-n_eval:]
    assert all(e["mode"] == "learned" and len(e["coef"]) == 3 for e in log)
    assert any(any(e["taken"]) for e in log)
    assert any(c != 0.0 for e in log for c in e["coef"][-1])  # the MLP direction is used
    assert it_learned < it_plain, (it_learned, it_plain, log[-1])
    # the learned directions come from one forward over every species' rows: the same bits as
    # one forward per species, sigma_k tau f(X_k)
    X_all, sig_all, per = learned._features()
    batched = learned._mlp_directions(X_all, sig_all)
    for k, (X, sigma) in enumerate(per):
        one = (sigma.double() * learned.tau) * learned.mlp.forward(X.clone()).double()
        assert torch.equal(batched[k], one), k


def _start_from(st, n_hist, extra, rtol, max_it=20000):
    """Solve the stepper's current systems (right-hand sides already formed) from the Galerkin
    start over its last n_hist increments plus the explicit directions `extra`, into fresh
    output arrays.  With rtol = 1 the start itself meets the tolerance and comes back unchanged
    by any CG iteration.  Returns (solutions, stats, coefficients per species)."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import ptr, species_array, species_struct, stream_ptr
    keep, structs, xs = [], [], []
    for k, s in enumerate(st.defs):
        x = torch.empty_like(st.u[k])
        xs.append(x)
        structs.append(species_struct(st.mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                      alpha=st.alpha_dev[k], alpha_value=s.get("alpha_value", 0.0),
                                      bc_dofs=st.bc[k], bc_value=s.get("bc_value", 0.0),
                                      b=st.b[k], x=x, keep=keep))
    S = st.S
    steps = [st.u] + st.u_hist[:n_hist]
    hist = (ctypes.c_void_p * (len(steps) * S))(*[ptr(a[k]).value for a in steps for k in range(S)])
    ext = (ctypes.c_void_p * max(1, len(extra)))(*[ptr(e).value for e in extra])
    nd = n_hist + (1 if extra else 0)
    coef = (ctypes.c_double * (nd * S))()
    stats = _lib.Stats()
    _lib.call("dsv_solve_warm_history", st.mesh.handle, S, species_array(structs), 1.0 / st.dt,
              st.lumped, _lib.METHODS[st.method], float(rtol), int(max_it), n_hist, hist,
              1 if extra else 0, ext, stream_ptr(st.mesh.device), stats, coef, None, None)
    torch.cuda.synchronize(st.mesh.device)
    return xs, stats.as_dict(S), [[coef[j * S + s] for j in range(nd)] for s in range(S)]


@pytest.mark.parametrize("n_hist", [2, 7])
def test_mlp_direction_isolated_on_the_same_system(dev, n_hist):
    """What the trained network adds, on one and the same system of a learned trajectory: the
    Galerkin start over n_hist increments with and without the MLP direction.  Galerkin
    minimality over nested subspaces fixes the order of their A-energy errors (with the
    direction never worse, to the Gram solve's rounding); both then converge to the bar.  The
    iteration counts of the two solves are the network's own gain (printed; the bench reports
    the same comparison over whole trajectories)."""
    import torch
    from dsv import workloads
    from dsv.ceed_operators import apply_batched
    from dsv.core import Mesh
    from dsv.mlp import WarmStartMLP
    from dsv.transport import Stepper
    w = workloads.disc3(seed=1, scale=6.0)
    mesh = Mesh(w["conn"], w["xyz"])
    st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=1e-10, phi=w["phi"],
                 grad_phi=w["grad_norm"], eps=w["eps"], mlp=WarmStartMLP(mesh.device))
    st.set_warm_start("extrapolate", history_dirs=n_hist)
    st.train_warm_start(9, epochs=30)
    st.set_warm_start("learned", history_dirs=n_hist)
    for _ in range(2):
        st.step()
    # the next step's systems, untouched by the comparison below
    st.rhs()
    X_all, sig_all, _ = st._features()
    extra = list(st._mlp_directions(X_all, sig_all))
    with_mlp, _, c_with = _start_from(st, n_hist, extra, 1.0)
    without, _, c_without = _start_from(st, n_hist, [], 1.0)
    # the tight FP64 solution of the same systems, from u_n
    refs = [u.clone() for u in st.u]
    from dsv import _lib
    from dsv.core import species_array, species_struct, stream_ptr
    keep, structs = [], []
    for k, s in enumerate(st.defs):
        structs.append(species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                                      alpha=st.alpha_dev[k], alpha_value=s.get("alpha_value", 0.0),
                                      bc_dofs=st.bc[k], bc_value=s.get("bc_value", 0.0),
                                      b=st.b[k], x=refs[k], keep=keep))
    stats = _lib.Stats()
    _lib.call("dsv_solve", mesh.handle, st.S, species_array(structs), 1.0 / st.dt, st.lumped,
              _lib.METHODS["csr64"], 1e-13, 100000, stream_ptr(mesh.device), stats)
    torch.cuda.synchronize()
    assert all(stats.as_dict(st.S)["converged"])
    # A-energy errors of the two starts
    spec = [dict(D=s.get("D"), D_value=s.get("D_value", 0.0), alpha=st.alpha_dev[k],
                 alpha_value=s.get("alpha_value", 0.0), bc_dofs=st.bc[k])
            for k, s in enumerate(st.defs)]
    e_with = [a - r for a, r in zip(with_mlp, refs)]
    e_without = [a - r for a, r in zip(without, refs)]
    Ae_with = apply_batched(mesh, spec, 1.0 / st.dt, e_with, lumped=st.lumped, precision=64)
    Ae_without = apply_batched(mesh, spec, 1.0 / st.dt, e_without, lumped=st.lumped,
                               precision=64)
    E_with = [float(torch.dot(e, a)) for e, a in zip(e_with, Ae_with)]
    E_without = [float(torch.dot(e, a)) for e, a in zip(e_without, Ae_without)]
    for k in range(st.S):
        assert E_with[k] >= 0.0 and E_without[k] >= 0.0
        if any(c != 0.0 for c in c_with[k]) and any(c != 0.0 for c in c_without[k]):
            assert E_with[k] <= E_without[k] * (1.0 + 1e-6) + 1e-300, (k, E_with, E_without)
    # the two solves of the same systems, to the bar
    sol_with, s_with, _ = _start_from(st, n_hist, extra, 1e-10)
    sol_without, s_without, _ = _start_from(st, n_hist, [], 1e-10)
    for sols, s in ((sol_with, s_with), (sol_without, s_without)):
        assert all(s["converged"])
        for k in range(st.S):
            err = float((sols[k] - refs[k]).abs().max() / refs[k].abs().max())
            assert err <= 1e-8, (k, err)
    print(f"n_hist={n_hist}: start energy with/without the MLP direction per species "
          f"{[round(a / b, 4) if b > 0 else None for a, b in zip(E_with, E_without)]}; "
          f"CG iterations with {s_with['iterations']} without {s_without['iterations']}")

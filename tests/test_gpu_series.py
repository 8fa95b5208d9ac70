"""The learned warm start over the receding-interface time series on the full 806k-tet disc:
200 backward-Euler steps of O2 / Cl / OH while the interface recedes by the ORR law in the
|phi| < 3 h band, at a speed proportional to local O2 (coefficients recomputed on the GPU every
step); the MLP is trained on the GPU on steps
1-100 of the solver's own outputs and used on steps 101-200.  Against the plain
previous-solution start over steps 101-200: far fewer CG iterations; at rtol 1e-10 every
checked step's solutions lie within 1e-8 of the plain FP64 path's rtol-1e-13 solution of the
same system."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(mesh, w, warm, n=200, epochs=5):
    from dsv import workloads
    from dsv.mlp import WarmStartMLP
    from dsv.transport import Stepper
    st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=1e-10, phi=w["phi"],
                 grad_phi=w["grad_norm"], eps=w["eps"],
                 mlp=WarmStartMLP(mesh.device) if warm == "learned" else None)
    drv = workloads.RecedingInterface(st, w)
    half = n // 2
    hist = None
    if warm == "learned":
        hist = st.train_warm_start(half, epochs=epochs, before_step=drv)
        st.set_warm_start("learned", history_dirs=2)
    else:
        for _ in range(half):
            drv()
            st.step()
    its, errs = [], []
    for n_done in range(n - half):
        drv()
        its.append(st.step()["iterations"])
        if n_done % 20 == 19:
            refs, hr = st.fp64_reference(1e-13)
            assert all(hr["converged"]), hr
            errs.append(max(st.step_error(refs)))
    return st, its, hist, errs


def test_learned_warm_start_over_receding_series(gpu):
    from dsv import workloads
    from dsv.core import Mesh
    w = workloads.disc3()
    mesh = Mesh(w["conn"], w["xyz"])
    print("[series] previous-solution start", flush=True)
    prev, it_prev, _, err_prev = _run(mesh, w, "previous")
    print("[series] learned start", flush=True)
    lrn, it_lrn, hist, err_lrn = _run(mesh, w, "learned")
    assert hist[-1] < hist[0]
    for st in (prev, lrn):
        assert all(all(h["converged"]) for h in st.history)
        assert float(st.phi.max()) < float(np.max(w["phi"]))  # the interface moved
    tot_prev = sum(sum(i) for i in it_prev)
    tot_lrn = sum(sum(i) for i in it_lrn)
    print(f"[series] CG iterations, steps 101-200: previous start {tot_prev}, learned {tot_lrn}")
    assert tot_lrn < 0.5 * tot_prev, (tot_lrn, tot_prev)
    log = lrn.warm_log[-100:]
    assert all(e["mode"] == "learned" for e in log) and any(any(e["taken"]) for e in log)
    assert any(c != 0.0 for e in log for c in e["coef"][-1])  # the MLP direction takes part
    print(f"[series] error vs FP64 at every 20th step: previous {err_prev}, learned {err_lrn}")
    assert max(err_prev) <= 1e-8 and max(err_lrn) <= 1e-8

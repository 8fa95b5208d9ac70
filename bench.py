"""Benchmark: three-species backward-Euler steps/s on the ~800k-tet graded disc, rtol 1e-8.

One step = RHS assembly + batched solve of O2, Cl-, OH- to ||b - A x|| <= 1e-8 ||b|| (FP64
true residual).  Four GPU paths advance the same trajectory from the same initial state (W
untimed warm-up steps, then K timed steps):
  plain    -- FP64 CSR re-assembled every step + FP64 Jacobi-PCG (csr64), previous-solution
              start (the reference's operator path and warm start)
  fp32     -- fused FP32 3-species Jacobi-PCG inside FP64 refinement (mf32_ir64), previous start
  chosen   -- the fp32 path started from the Galerkin warm start over the last H solution
              increments (--history-dirs, default 8; previous solution as fallback)
  learned  -- the fp32 path started from the Galerkin warm start over H - 1 increments and the
              learned MLP direction, trained on the GPU on the W warm-up steps
`value` is the chosen path's steps/s; the timed steps include every warm-start computation
(features, inference, projection).  The roofline entry is the fused FP32 3-species CG
iteration kernel (the solve's dominant kernel), timed with HIP events on its own stream; the
CPU baseline is the oracle (numpy assembly + C OpenMP Jacobi-PCG) on one 3-species step of
the same mesh.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload disc3]
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dissolve-2.0_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _coef_bytes(a, C, V):
    """4 bytes per value of a per-element (length C) or per-vertex (length V) field, 0 for a
    scalar."""
    if a is None:
        return 0
    n = len(a)
    return 4 * C if n == C else 4 * V


def apply_bytes(C, V, species):
    """Algorithmic bytes of one FP32 multi-species matrix-free apply (the minimal-input
    formulation, geometry recomputed): connectivity 16 C, coordinates 12 V, and per species x
    read and y written (8 V) plus its coefficient fields."""
    return 16 * C + 12 * V + sum(8 * V + _coef_bytes(s.get("D"), C, V) +
                                 _coef_bytes(s.get("alpha"), C, V) for s in species)


def cg_iteration_bytes(C, V, species):
    """Algorithmic bytes of one fused FP32 Jacobi-PCG iteration: the apply plus the vector
    passes of the iteration, 9 FP32 words per vertex and species."""
    return apply_bytes(C, V, species) + 36 * V * len(species)


def measured_traffic(workload="disc3"):
    """HBM bytes per launch of the fused CG iteration from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel factors), or
    None."""
    path = os.path.join(ROOT, "profiles", f"k_cg_{workload}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return {"bytes_per_launch": int(line.split(":")[1]),
                            "source": os.path.relpath(path, ROOT)}
    except OSError:
        return None
    return None


def load_workload(name):
    from dsv import workloads
    if name == "disc3":
        return workloads.disc3()
    if name == "cube_jitter":
        return workloads.cube_jitter()
    if name == "gyroid3":
        return workloads.gyroid3()
    raise SystemExit(f"unknown workload {name}")


def run_path(mesh, w, method, steps, warmup, rtol, warm="previous", epochs=30, history_dirs=8,
             after_step=None, low_modes=None):
    """W warm-up steps (the learned path trains on them), then K timed steps.  after_step(st),
    if given, runs after every timed step inside the timed region (only the untimed error
    replays pass one).  low_modes = (n_modes, degree, passes): the operator's low-mode subspace
    is built once before the warm-up (Stepper.build_low_modes; set-up, not timed) and corrects
    every Galerkin start."""
    import torch
    from dsv.transport import Stepper
    mlp = None
    if warm == "learned":
        from dsv.mlp import WarmStartMLP
        mlp = WarmStartMLP(mesh.device)
    st = Stepper(mesh, w["species"], w["dt"], method=method, rtol=rtol, phi=w.get("phi"),
                 grad_phi=w.get("grad_norm"), eps=w.get("eps"), mlp=mlp)
    lm = None
    if low_modes and low_modes[0] > 0 and warm != "previous":
        n, deg, passes = low_modes
        info = st.build_low_modes(n, deg, passes)
        lm = {"n_modes": n, "degree": deg, "passes": passes,
              "setup_seconds": round(info["seconds"], 3),
              "lowest_ritz": [round(r[0], 6) for r in info["ritz"]],
              "largest_ritz": [round(r[-1], 6) for r in info["ritz"]]}
    train = None
    if warm == "learned":
        st.set_warm_start("extrapolate", history_dirs=min(history_dirs, st.MAX_HIST - 1))
        torch.cuda.synchronize()
        t = time.perf_counter()
        loss = st.train_warm_start(warmup, epochs=epochs)
        torch.cuda.synchronize()
        train = {"pairs_from_steps": warmup, "epochs": epochs,
                 "seconds_incl_steps": round(time.perf_counter() - t, 2),
                 "loss_first_last": [round(loss[0], 5), round(loss[-1], 5)]}
        log(f"warm start trained on {warmup} steps in {train['seconds_incl_steps']}s, "
            f"loss {loss[0]:.3g} -> {loss[-1]:.3g}")
    else:
        if warm != "previous":
            st.set_warm_start(warm, history_dirs=history_dirs)
        for i in range(warmup):
            st.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for i in range(steps):
        d = st.step()
        its.append(d)
        if after_step is not None:
            after_step(st)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    conv = all(all(h["converged"]) for h in its)
    log_tail = st.warm_log[-steps:] if warm != "previous" else []
    return {
        "steps_per_s": steps / t, "ms_per_step": 1e3 * t / steps, "converged": conv,
        "iterations_per_step": [h["iterations"] for h in its],
        "outer_per_step": [h["outer"] for h in its],
        "rel_residual_last": its[-1]["rel_residual"],
        "train": train,
        "low_modes": lm,
        "taken_per_step": [e["taken"] for e in log_tail],
        "coef_last_step": log_tail[-1]["coef"] if log_tail else None,
    }, st


def pick_low_modes(mesh, w, rtol, low_modes, history_dirs=None):
    """Whether the Galerkin paths take the low-mode correction, by a short trial from the
    initial state: 2 untimed and 5 timed steps with it (6 increments) and without it (8), the
    faster wins.  Its per-step cost (two passes over the subspace and one FP64 apply) pays
    where the warm-started solves still spend their iterations in the few smoothest modes (the
    disc); on a mesh whose slow spectrum those modes do not cover (the gyroid lattice) it does
    not."""
    on, _ = run_path(mesh, w, "mf32_ir64", 5, 2, rtol, warm="extrapolate",
                     history_dirs=history_dirs or 6, low_modes=low_modes)
    off, _ = run_path(mesh, w, "mf32_ir64", 5, 2, rtol, warm="extrapolate",
                      history_dirs=history_dirs or 8)
    trial = {"steps": "3-7 of the path, each way", "with_steps_per_s": round(on["steps_per_s"], 2),
             "without_steps_per_s": round(off["steps_per_s"], 2),
             "setup_seconds": on["low_modes"]["setup_seconds"] if on["low_modes"] else None}
    return on["steps_per_s"] > off["steps_per_s"], trial


def error_replay(mesh, w, method, steps, warmup, rtol, warm, history_dirs, final_u, every=1,
                 epochs=30, low_modes=None):
    """Replays a path untimed and solves each checked step's three systems again with the plain
    FP64 path to rtol 1e-13 (Stepper.fp64_reference, started from the step's solution).  The
    solver is deterministic (no atomics, fixed reduction orders), so the replay retraces the
    timed run: its final solutions are compared with the timed run's bit for bit.  Returns
    {"per_step": [[err per species] per checked timed step], "max_per_species", "replay_bit_identical"}."""
    import torch
    errs = []
    count = [0]

    def check(st):
        count[0] += 1
        if count[0] % every == 0 or count[0] == steps:
            refs, hr = st.fp64_reference(1e-13)
            if not all(hr["converged"]):
                raise RuntimeError(f"FP64 reference solve did not converge: {hr}")
            errs.append([float(f"{e:.3e}") for e in st.step_error(refs)])

    r, st = run_path(mesh, w, method, steps, warmup, rtol, warm=warm, epochs=epochs,
                     history_dirs=history_dirs, after_step=check, low_modes=low_modes)
    same = all(torch.equal(a, b) for a, b in zip(st.u, final_u))
    S = len(st.u)
    return {"per_step": errs, "steps_checked": len(errs),
            "max_per_species": {w["species"][k]["name"]: max(e[k] for e in errs) for k in range(S)},
            "replay_bit_identical": same}


def run_coupled(mesh, w, steps, warmup, rtol, levelset=False):
    """The reference-order four-species step (dsv.coupled.CoupledStepper: Zn, Cl and O2
    batched, O2 under-relaxed, the film, OH from the new fields; the diffusivities follow the
    film) on the same mesh: plain FP64 path vs the chosen path, K timed steps after W warm-up
    steps each."""
    import torch
    from dsv import workloads
    from dsv.coupled import CoupledStepper
    cw = workloads.coupled_disc(w["conn"], w["xyz"])
    out = {"workload": f"{w['name']} mesh, Zn/Cl/film/O2/OH in the order of the reference's "
                       "species blocks, film-dependent diffusivities, O2 omega = 0.3"
                       + (", interface velocity + level-set advance + redistancing every step"
                          if levelset else ""),
           "rtol": rtol}
    for name, method, warm in (("plain_fp64", "csr64", "previous"),
                               ("chosen", "mf32_ir64", "extrapolate")):
        cs = CoupledStepper(mesh, cw, method=method, rtol=rtol, warm=warm, levelset=levelset)
        for _ in range(warmup):
            cs.step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        hs = [cs.step() for _ in range(steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out[name] = {"steps_per_s": round(steps / dt, 3),
                     "cg_iterations_per_step": round(sum(sum(h["iterations"]) for h in hs) / steps, 1),
                     "iterations_last_step": dict(zip(hs[-1]["names"], hs[-1]["iterations"])),
                     "converged": all(all(h["converged"]) for h in hs)}
        log(f"coupled {name}{' + level set' if levelset else ''}: {out[name]}")
        del cs
    out["chosen_vs_plain"] = round(out["chosen"]["steps_per_s"] / out["plain_fp64"]["steps_per_s"], 3)
    return out


def cpu_baseline(w, all_threads):
    """The oracle's C restatement of one 3-species step of the same mesh from u0 (bounded
    sample): per species assembly through the scatter map, RHS, Jacobi-PCG to the same rtol,
    on 1 core and on `all_threads` cores, with the three parts timed apart.  The pattern and
    scatter map are built once beforehand, untimed, like the GPU mesh."""
    import numpy as np
    from oracle.solve import CStep
    t = time.perf_counter()
    cs = CStep(w["conn"], w["xyz"])
    setup_s = time.perf_counter() - t
    inv_dt = 1.0 / w["dt"]
    legs = {}
    for threads in (1, all_threads):
        ta = tr = tsolve = 0.0
        its = []
        for s in w["species"]:
            De = s["D"] if s.get("D") is not None else s["D_value"]
            t = time.perf_counter()
            cs.assemble(De, inv_dt, s.get("alpha"), 1, s.get("bc_dofs"), threads)
            ta += time.perf_counter() - t
            t = time.perf_counter()
            b = cs.rhs(s["u0"], inv_dt, s.get("f"), 1, s.get("bc_dofs"), s.get("bc_value", 0.0),
                       threads)
            tr += time.perf_counter() - t
            x0 = np.array(s["u0"], dtype=np.float64)
            if s.get("bc_dofs") is not None and len(s["bc_dofs"]):
                x0[s["bc_dofs"]] = s["bc_value"]
            t = time.perf_counter()
            _, it, rel = cs.solve(b, x0, w["rtol"], 20000, threads)
            tsolve += time.perf_counter() - t
            its.append(int(it))
        tot = ta + tr + tsolve
        legs[f"{threads}_cores"] = {"steps_per_s": round(1.0 / tot, 4), "step_s": round(tot, 3),
                                    "assembly_s": round(ta, 3), "rhs_s": round(tr, 3),
                                    "solve_s": round(tsolve, 3), "cg_iterations": its}
        log(f"cpu baseline on {threads} cores: {tot:.2f}s per step (assembly {ta:.2f}, rhs "
            f"{tr:.2f}, solve {tsolve:.2f}), iterations {its}")
    return legs, setup_s


def workload_text(w, C, V, rtol):
    sp = "/".join(s["name"] for s in w["species"])
    D = "/".join(f"{s['D_value']:g}" for s in w["species"])
    if w["name"] == "disc3":
        geo = ("graded Delaunay disc r=1 height 0.2, h graded 0.0074 (top) to 0.059 (bottom) "
               "(the specified 0.005-0.04 scaled by 1.47 to reach ~800k tets with this "
               "generator), seed 0")
    elif w["name"] == "gyroid3":
        geo = "gyroid lattice sin x cos y + sin y cos z + sin z cos x < 0.3 on [0,4pi]^3, 160^3 Kuhn"
    else:
        geo = w["name"]
    return (f"{w['name']}: {geo}, {C} tets, {V} vertices, {sp} D={D}, dt={w['dt']}, "
            f"rtol={rtol:g}")


def assembly_kernel(C, V, s):
    """The kernel dsv_assemble takes for species s under lumped mass: 'scalar' (scalar D and
    alpha), 'valpha' (scalar D, per-vertex alpha) or 'rows' (the row assembly)."""
    a = s.get("alpha")
    if s.get("D") is None and a is None:
        return "scalar"
    if s.get("D") is None and len(a) == V and V != C:
        return "valpha"
    return "rows"


def assembly_bytes(C, V, nnz, s):
    """Algorithmic bytes of one lumped FP64 CSR assembly of species s.  'scalar' (D K1 +
    (inv_dt + alpha) M1): K1 and M1 read, the values written, 24 nnz.  'valpha': K1, the
    consistent unit mass and the column indices read, the values written (28 nnz), and per
    vertex the row pointer, diagonal position, lumped mass and alpha (24 V).  'rows' (from the
    geometry): connectivity 16 C, FP64 coordinates 24 V, the values written 8 nnz, and 8 bytes
    per value of a coefficient field."""
    def coef(a):
        return 0 if a is None else 8 * len(a)
    kind = assembly_kernel(C, V, s)
    if kind == "scalar":
        return 24 * nnz
    if kind == "valpha":
        return 28 * nnz + 24 * V
    return 16 * C + 24 * V + 8 * nnz + coef(s.get("D")) + coef(s.get("alpha"))


def time_assembly(mesh, w, reps, s=None):
    """ms per launch of dsv_assemble's kernel for species s (default: the first species with a
    reaction field, else the first; no Dirichlet rows), HIP events around `reps`
    graph-replayed launches."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import species_array, species_struct
    if s is None:
        s = next((x for x in w["species"] if x.get("alpha") is not None), w["species"][0])
    keep = []
    st = species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0), alpha=s.get("alpha"),
                        alpha_value=s.get("alpha_value", 0.0), keep=keep)
    stream = torch.cuda.Stream()
    ms = ctypes.c_double()
    _lib.call("dsv_time_apply", mesh.handle, 1, species_array([st]), 1.0 / w["dt"], 1, 64, 2,
              reps, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(ms))
    return ms.value, s


def copy_floor(nbytes, reps):
    """A plain device copy moving the same bytes as one apply (nbytes/2 read, nbytes/2 written),
    `reps` copies captured in one hipGraph and replayed between two HIP events: what a kernel
    of this size reaches on this GPU when it does nothing but stream its bytes."""
    import torch
    n = max(1, nbytes // 8)
    x = torch.rand(n, device="cuda")
    y = torch.empty_like(x)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(3):
            y.copy_(x)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(reps):
                y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g.replay()
        e0.record(s)
        g.replay()
        e1.record(s)
        e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    gbs = 8 * n / (us * 1e-6) / 1e9
    return {"bytes_moved": 8 * n, "us": round(us, 2), "gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def time_kernels(mesh, w, reps, stream=None):
    """(ms per fused CG iteration, ms per apply) of the FP32 3-species kernels, HIP events
    around `reps` back-to-back launches on one stream."""
    import ctypes
    import torch
    from dsv import _lib
    from dsv.core import species_array, species_struct
    mesh.fresh_operator()
    keep = []
    structs = [species_struct(mesh, D=s.get("D"), D_value=s.get("D_value", 0.0),
                              alpha=s.get("alpha"), bc_dofs=s.get("bc_dofs"),
                              bc_value=s.get("bc_value", 0.0), keep=keep) for s in w["species"]]
    arr = species_array(structs)
    stream = stream or torch.cuda.Stream()
    out = []
    for mode in (1, 0):
        ms = ctypes.c_double()
        _lib.call("dsv_time_apply", mesh.handle, len(structs), arr, 1.0 / w["dt"], 1, 32, mode,
                  reps, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(ms))
        out.append(ms.value)
    return out[0], out[1]


def profile_run(mesh, w, reps):
    import torch
    n = 64 * 1024 * 1024  # 256 MiB of float32
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy reference: 3 x {4 * n} bytes read and written")
    ms_iter, ms_apply = time_kernels(mesh, w, reps)
    torch.cuda.synchronize()
    log(f"fused CG iteration {ms_iter * 1e3:.1f} us, apply {ms_apply * 1e3:.1f} us "
        f"({reps} launches each plus 3 warm-up)")
    import numpy as np
    C, V = len(w["conn"]), len(w["xyz"])
    # the first species with a reaction field (else the first), a species with scalar D and
    # alpha, and a per-cell D field: the three assembly kernels
    scalar = [x for x in w["species"] if x.get("D") is None and x.get("alpha") is None][:1]
    cases = [None] + scalar + [{"name": "per-cell D field (lognormal, seed 7)",
                                "D": np.random.default_rng(7).lognormal(0.0, 0.5, C)}]
    for s in cases:
        ms_asm, s_asm = time_assembly(mesh, w, reps, s)
        torch.cuda.synchronize()
        log(f"FP64 CSR assembly ({s_asm['name']}, {assembly_kernel(C, V, s_asm)}) "
            f"{ms_asm * 1e3:.1f} us, {assembly_bytes(C, V, mesh.nnz, s_asm)} algorithmic bytes "
            f"({reps} launches plus 3 warm-up)")


def kernels_at_scale(shapes=((256, 256, 1), (256, 256, 2))):
    """The fused CG iteration and the apply on the 14.7M-tet gyroid lattice, where the working
    set (about 1 GB per CG iteration) is far beyond the 256 MB MALL: HIP-event times and
    fractions of HBM peak of the algorithmic bytes, best of the given launch shapes."""
    from dsv import workloads
    from dsv.core import Mesh
    t = time.perf_counter()
    w = workloads.gyroid3()
    mesh = Mesh(w["conn"], w["xyz"])
    C, V = len(w["conn"]), len(w["xyz"])
    nb, ab = cg_iteration_bytes(C, V, w["species"]), apply_bytes(C, V, w["species"])
    best = None
    for nv, th, lanes in shapes:
        mesh.set_partition(nv, th, lanes)
        it, ap = time_kernels(mesh, w, 30)
        if best is None or it < best[0]:
            best = (it, ap, (nv, th, lanes))
    it, ap, shape = best
    out = {"workload": f"gyroid3: {C} tets, {V} vertices",
           "launch_shape": {"vertices_per_block": shape[0], "threads": shape[1],
                            "lanes_per_vertex": shape[2]},
           "cg_iteration": {"us": round(it * 1e3, 2), "algorithmic_bytes": nb,
                            "achieved_gbs": round(nb / (it * 1e-3) / 1e9, 1),
                            "frac": round(nb / (it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "apply": {"us": round(ap * 1e3, 2), "algorithmic_bytes": ab,
                     "achieved_gbs": round(ab / (ap * 1e-3) / 1e9, 1),
                     "frac": round(ab / (ap * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    log(f"at scale ({time.perf_counter() - t:.0f}s incl. mesh): {out}")
    mesh.close()
    return out


def run_series(mesh, w, n=200, epochs=5, history_dirs=8):
    """Receding-interface series: the previous-solution start, the Galerkin extrapolation and
    the learned start (trained on steps 1..n/2) over steps n/2+1..n; prints one JSON line."""
    import torch
    from dsv import workloads
    from dsv.mlp import WarmStartMLP
    from dsv.transport import Stepper
    out = {"metric": "receding-interface series, 3-species BE steps/s and CG iterations over "
                     f"steps {n // 2 + 1}-{n}, rtol {w['rtol']:g}",
           "workload": f"{w['name']}: {len(w['conn'])} tets, interface receding by the ORR law "
                       "v = -0.008 O2 inside |phi| < 3 h (libdsv velocity and level-set advance), "
                       "coefficients recomputed on the GPU every step", "paths": {}}
    for name, method, warm in (("plain_fp64", "csr64", "previous"),
                               ("previous", "mf32_ir64", "previous"),
                               ("extrapolate", "mf32_ir64", "extrapolate"),
                               ("learned", "mf32_ir64", "learned")):
        st = Stepper(mesh, w["species"], w["dt"], method=method, rtol=w["rtol"],
                     phi=w["phi"], grad_phi=w["grad_norm"], eps=w["eps"],
                     mlp=WarmStartMLP(mesh.device) if warm == "learned" else None)
        drv = workloads.RecedingInterface(st, w)
        half = n // 2
        train = None
        if warm == "learned":
            t = time.perf_counter()
            loss = st.train_warm_start(half, epochs=epochs, before_step=drv)
            train = {"epochs": epochs, "seconds_incl_steps": round(time.perf_counter() - t, 2),
                     "loss_first_last": [round(loss[0], 4), round(loss[-1], 4)]}
            st.set_warm_start("learned", history_dirs=min(history_dirs, st.MAX_HIST - 1))
        else:
            if warm == "extrapolate":
                st.set_warm_start("extrapolate", history_dirs=history_dirs)
            for _ in range(half):
                drv()
                st.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(n - half):
            drv()
            its.append(st.step()["iterations"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out["paths"][name] = {"steps_per_s": round((n - half) / dt, 2),
                              "cg_iterations": sum(sum(i) for i in its),
                              "converged": all(all(h["converged"]) for h in st.history),
                              "training": train}
        log(f"series {name}: {out['paths'][name]}")
    base = out["paths"]["previous"]["cg_iterations"]
    for p in out["paths"].values():
        p["cg_iterations_saved_pct"] = round(100 * (1 - p["cg_iterations"] / base), 1)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps of every path; the learned warm start trains on them")
    ap.add_argument("--workload", default="disc3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the kernel timings on the 14.7M-tet gyroid lattice")
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--history-dirs", type=int, default=None,
                    help="past solution increments in the Galerkin warm start (1..8); default 6 "
                         "with the low-mode correction, 8 without it (the series, the leg "
                         "without it), each its best")
    ap.add_argument("--low-modes", type=int, default=64,
                    help="low-mode subspace per species of the Galerkin paths (0: none; a "
                         "multiple of 4, at most 64)")
    ap.add_argument("--low-mode-degree", type=int, default=80)
    ap.add_argument("--low-mode-passes", type=int, default=4)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--series", action="store_true",
                    help="instead of the headline: the receding-interface time series on the "
                         "disc (200 steps; the MLP trains on steps 1-100), one JSON line")
    ap.add_argument("--paths",
                    default="plain,fp32,chosen,chosen_no_low_modes,learned,extrapolate_no_mlp,"
                            "plain_bar,chosen_bar",
                    help="comma list of the paths to run (the JSON line needs the first four)")
    ap.add_argument("--bar-rtol", type=float, default=5e-11,
                    help="rtol of the *_bar paths: the tolerance at which every species' step "
                         "solution is within 1e-8 of the FP64 solution")
    ap.add_argument("--no-coupled", action="store_true",
                    help="skip the reference-order four-species step legs")
    ap.add_argument("--coupled-only", action="store_true",
                    help="instead of the headline: only the reference-order four-species step "
                         "legs (static, then with the interface update), one JSON line")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the untimed error replays against the FP64 solutions")
    ap.add_argument("--error-every", type=int, default=1,
                    help="check every N-th timed step in the error replays")
    ap.add_argument("--partition", default=None,
                    help="NV,THREADS,LANES launch shape of the matrix-free kernels (skips tuning)")
    ap.add_argument("--profile-run", action="store_true",
                    help="kernel-only run for rocprofv3: a 256 MiB device copy (the counter "
                         "reference), then the fused CG iteration and the apply, --apply-reps "
                         "launches each, on the default launch shape")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    from dsv.core import Mesh

    t = time.perf_counter()
    w = load_workload(a.workload)
    C, V = len(w["conn"]), len(w["xyz"])
    log(f"workload {w['name']}: {C} tets, {V} vertices, built in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    mesh = Mesh(w["conn"], w["xyz"])
    log(f"device mesh: nnz={mesh.nnz} blocks={mesh.n_blocks} entries={mesh.n_entries} "
        f"({time.perf_counter() - t:.1f}s)")
    rtol = w["rtol"]
    if a.partition:
        nv, th, lanes = (int(t) for t in a.partition.split(","))
        mesh.set_partition(nv, th, lanes)
        a.no_tune = True
        log(f"launch shape: vertices_per_block={nv} threads={th} lanes_per_vertex={lanes}")
    if a.profile_run:
        profile_run(mesh, w, a.apply_reps)
        return
    # Python's collector: everything the set-up created (torch, numpy, the mesh) goes to the
    # permanent generation, so that a full collection during the timed steps scans only the
    # stepping's own objects (a full collection of the whole process pauses 20-50 ms)
    gc.collect()
    gc.freeze()
    tuned = None
    if a.partition:
        tuned = {"vertices_per_block": nv, "threads": th, "lanes_per_vertex": lanes}
    if not a.no_tune:
        from dsv.autotune import tune
        best, table = tune(mesh, w["species"], 1.0 / w["dt"])
        for row in table:
            log(f"tune nv={row[0]} threads={row[1]} lanes={row[2]} entries={row[3]} "
                f"{row[4] * 1e3:.1f} us")
        tuned = {"vertices_per_block": best[0], "threads": best[1], "lanes_per_vertex": best[2]}
        log(f"tuned: {tuned}")
    if a.series:
        run_series(mesh, w, history_dirs=a.history_dirs or 8)
        return
    if a.coupled_only:
        print(json.dumps({"static": run_coupled(mesh, w, a.steps, a.warmup, w["rtol"]),
                          "with_interface": run_coupled(mesh, w, a.steps, a.warmup, w["rtol"],
                                                        levelset=True)}), flush=True)
        return
    res, finals = {}, {}
    paths = a.paths.split(",")
    lm_arg = (a.low_modes, a.low_mode_degree, a.low_mode_passes)
    use_lm, lm_trial = False, None
    if a.low_modes > 0:
        use_lm, lm_trial = pick_low_modes(mesh, w, rtol, lm_arg, a.history_dirs)
        log(f"low-mode trial: {lm_trial} -> {'with' if use_lm else 'without'}")
    hd = a.history_dirs or (6 if use_lm else 8)
    PATHS = (("plain", "csr64", "previous", rtol), ("fp32", "mf32_ir64", "previous", rtol),
             ("chosen", "mf32_ir64", "extrapolate", rtol),
             # the same Galerkin start without the low-mode correction
             ("chosen_no_low_modes", "mf32_ir64", "extrapolate", rtol),
             ("learned", "mf32_ir64", "learned", rtol),
             # the learned path's increments without its MLP direction: the network's own gain
             ("extrapolate_no_mlp", "mf32_ir64", "extrapolate", rtol),
             # the tolerance at which every species' step solution meets 1e-8 (see error legs)
             ("plain_bar", "csr64", "previous", a.bar_rtol),
             ("chosen_bar", "mf32_ir64", "extrapolate", a.bar_rtol))
    for name, method, warm, rt in PATHS:
        if name not in paths:
            continue
        h = min(hd, 7) if name in ("learned", "extrapolate_no_mlp") else \
            ((a.history_dirs or 8) if name == "chosen_no_low_modes" else hd)
        lm = lm_arg if use_lm and name != "chosen_no_low_modes" else None
        r, st = run_path(mesh, w, method, a.steps, a.warmup, rt, warm=warm, history_dirs=h,
                         low_modes=lm)
        r.update(method=method, warm=warm, rtol=rt, history_dirs=h, low_modes_arg=lm)
        res[name] = r
        finals[name] = [u.clone() for u in st.u]
        log(f"{name} ({method}, {warm} start, rtol {rt:g}): {r['steps_per_s']:.2f} steps/s, "
            f"iterations/step {r['iterations_per_step']}, converged={r['converged']}")
        del st

    needed = ("plain", "fp32", "chosen", "learned")
    if any(p not in res for p in needed):
        log("not every path ran: no JSON line")
        return
    # error of every step's solutions against the plain FP64 path's rtol-1e-13 solution of the
    # same systems, from untimed replays of the timed runs
    errors = {}
    for name in ("plain", "chosen", "learned", "plain_bar", "chosen_bar"):
        if name not in res or a.no_errors:
            continue
        r = res[name]
        t = time.perf_counter()
        errors[name] = error_replay(mesh, w, r["method"], a.steps, a.warmup, r["rtol"], r["warm"],
                                    r["history_dirs"], finals[name], every=a.error_every,
                                    low_modes=r["low_modes_arg"])
        errors[name]["rtol"] = r["rtol"]
        log(f"error replay {name}: max per species {errors[name]['max_per_species']}, "
            f"bit-identical replay {errors[name]['replay_bit_identical']} "
            f"({time.perf_counter() - t:.1f}s)")

    coupled = None
    if not a.no_coupled and w["name"] == "disc3":
        coupled = {"static": run_coupled(mesh, w, a.steps, a.warmup, rtol),
                   "with_interface": run_coupled(mesh, w, a.steps, a.warmup, rtol, levelset=True)}

    # roofline of the fused FP32 3-species CG iteration (the chosen path's dominant kernel)
    ms_iter, ms_apply = time_kernels(mesh, w, a.apply_reps)
    nbytes = cg_iteration_bytes(C, V, w["species"])
    abytes = apply_bytes(C, V, w["species"])
    achieved = nbytes / (ms_iter * 1e-3) / 1e9
    apply_gbs = abytes / (ms_apply * 1e-3) / 1e9
    log(f"fused FP32 CG iteration: {ms_iter * 1e3:.1f} us/launch, {achieved:.0f} GB/s; "
        f"matrix-free apply alone {ms_apply * 1e3:.1f} us, {apply_gbs:.0f} GB/s")
    traffic = measured_traffic(w["name"])
    copy_ref = copy_floor(abytes, a.apply_reps)
    ms_asm, s_asm = time_assembly(mesh, w, a.apply_reps)
    asm_bytes = assembly_bytes(C, V, mesh.nnz, s_asm)
    asm_gbs = asm_bytes / (ms_asm * 1e-3) / 1e9
    log(f"FP64 CSR assembly ({s_asm['name']}): {ms_asm * 1e3:.1f} us/launch, {asm_gbs:.0f} GB/s")
    # the row assembly from the geometry, as a per-cell diffusivity field takes it
    import numpy as np
    s_fld = {"name": "per-cell D field (lognormal, seed 7)",
             "D": np.random.default_rng(7).lognormal(0.0, 0.5, C)}
    ms_fld, _ = time_assembly(mesh, w, a.apply_reps, s_fld)
    fld_bytes = assembly_bytes(C, V, mesh.nnz, s_fld)
    fld_gbs = fld_bytes / (ms_fld * 1e-3) / 1e9
    log(f"FP64 CSR assembly ({s_fld['name']}): {ms_fld * 1e3:.1f} us/launch, {fld_gbs:.0f} GB/s")
    # the streaming form, as a species with scalar D and alpha takes it
    s_sc = next((x for x in w["species"] if x.get("D") is None and x.get("alpha") is None), None)
    sc_entry = None
    if s_sc is not None:
        ms_sc, _ = time_assembly(mesh, w, a.apply_reps, s_sc)
        sc_bytes = assembly_bytes(C, V, mesh.nnz, s_sc)
        sc_gbs = sc_bytes / (ms_sc * 1e-3) / 1e9
        log(f"FP64 CSR assembly ({s_sc['name']}, scalar coefficients): {ms_sc * 1e3:.1f} us/launch, "
            f"{sc_gbs:.0f} GB/s")
        sc_entry = {"kernel": "k_assemble_scalar (D K1 + (1/dt + alpha) M1 from the unit stiffness "
                              "and mass kept on the mesh)",
                    "species": s_sc["name"], "algorithmic_bytes_per_launch": sc_bytes,
                    "us_per_launch": round(ms_sc * 1e3, 2), "achieved_gbs": round(sc_gbs, 1),
                    "frac": round(sc_gbs / HBM_PEAK_GBS, 4)}

    scale = None
    if not a.no_scale and a.workload == "disc3":
        scale = kernels_at_scale()

    cpu = None
    if not a.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        legs, setup_s = cpu_baseline(w, threads)
        top = legs[f"{threads}_cores"]
        cpu = {"value": top["steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"one 3-species BE step of {w['name']} from u0: the oracle's C "
                         f"assembly (scatter map, OpenMP atomics), RHS and OpenMP Jacobi-PCG to "
                         f"rtol {rtol:g}; {top['step_s']}s on {threads} cores, "
                         f"{legs['1_cores']['step_s']}s on 1 core (pattern and scatter map "
                         f"built once beforehand in {setup_s:.1f}s, untimed)",
               "legs": legs}

    chosen, plain, fp32, learned = res["chosen"], res["plain"], res["fp32"], res["learned"]

    def its(r):
        return round(sum(sum(i) for i in r["iterations_per_step"]) / a.steps, 1)

    def saved(r):
        return round(100 * (1 - its(r) / its(fp32)), 1)

    def leg(name):
        r = res[name]
        out = {"steps_per_s": round(r["steps_per_s"], 3), "cg_iterations_per_step": its(r),
               "rtol": r["rtol"]}
        if name in errors:
            out["max_error_vs_fp64"] = errors[name]["max_per_species"]
        if r.get("low_modes"):
            out["low_modes"] = r["low_modes"]
        return out

    mlp = None
    if "extrapolate_no_mlp" in res:
        e7 = res["extrapolate_no_mlp"]
        nh = min(hd, 7)
        mlp = {"compared": f"learned ({nh} increments + the MLP direction) vs the same {nh} "
                           "increments alone, each on its own trajectory (both with the "
                           "low-mode correction)",
               "steps_per_s_with_mlp": round(learned["steps_per_s"], 3),
               "steps_per_s_without": round(e7["steps_per_s"], 3),
               "cg_iterations_per_step_with_mlp": its(learned),
               "cg_iterations_per_step_without": its(e7),
               "cg_iterations_saved_by_mlp_pct": round(100 * (1 - its(learned) / its(e7)), 1),
               "steps_per_s_ratio": round(learned["steps_per_s"] / e7["steps_per_s"], 3)}
    lm_gain = None
    if "chosen_no_low_modes" in res:
        c0 = res["chosen_no_low_modes"]
        lm_gain = {"compared": f"chosen ({hd} increments"
                               + (" + the low-mode correction" if use_lm else "") +
                               ") vs the Galerkin start over 8 increments alone, each on its "
                               "own trajectory",
                   "steps_per_s_without": round(c0["steps_per_s"], 3),
                   "cg_iterations_per_step_without": its(c0),
                   "cg_iterations_saved_pct": round(100 * (1 - its(chosen) / its(c0)), 1),
                   "steps_per_s_ratio": round(chosen["steps_per_s"] / c0["steps_per_s"], 3)}
    bar = None
    if "chosen_bar" in res and "plain_bar" in res:
        bar = {"rtol": a.bar_rtol, "chosen": leg("chosen_bar"), "plain_fp64": leg("plain_bar"),
               "chosen_vs_plain": round(res["chosen_bar"]["steps_per_s"] /
                                        res["plain_bar"]["steps_per_s"], 3)}
    window = f"steps {a.warmup + 1}-{a.warmup + a.steps} of each path ({a.warmup} untimed " \
             f"warm-up steps first, from the same initial state); Python's collector on, with " \
             f"the set-up's objects frozen (gc.freeze)"
    metric = "3-species BE steps/s at 800k tets, rtol 1e-8: chosen vs plain FP64; HBM %"
    if w["name"] != "disc3":
        metric = f"3-species BE steps/s on {w['name']} ({C} tets), rtol {rtol:g}: chosen vs " \
                 "plain FP64; HBM %"
    out = {
        "metric": metric,
        "value": round(chosen["steps_per_s"], 3), "unit": "steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(chosen["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload_text(w, C, V, rtol),
                   "timed_window": window,
                   "method": "FP32 fused 3-species Jacobi-PCG (Chronopoulos-Gear, one kernel per "
                             "iteration, stored FP32 geometry) in FP64 refinement, FP64 "
                             f"Galerkin warm start over the last {hd} solution "
                             "increments (previous solution as fallback), "
                             + (f"then the Galerkin correction over {a.low_modes} low modes per "
                                "species of D^-1 A (built once per operator before the warm-up, "
                                "Chebyshev-filtered subspace iteration), " if use_lm else
                                "no low-mode correction (the trial found it slower here), ")
                             + "species narrowing, 8-iteration polled chunks of direct launches",
                   "low_mode_trial": lm_trial,
                   "step": "the three species batched in one solve with static coefficients "
                           "(f_OH from the initial O2, no O2 under-relaxation)",
                   "parallelism": "1 GPU", "launch_shape": tuned},
        "plain_fp64": leg("plain"),
        "fp32_previous_start": leg("fp32"),
        "chosen_galerkin_extrapolation": dict(leg("chosen"),
                                              cg_iterations_saved_pct_vs_previous_start=saved(chosen)),
        "learned_warm_start": dict(leg("learned"),
                                   cg_iterations_saved_pct_vs_previous_start=saved(learned),
                                   training=learned["train"],
                                   mlp_coef_last_step=learned["coef_last_step"][-1]
                                   if learned["coef_last_step"] else None,
                                   taken_last_step=learned["taken_per_step"][-1]
                                   if learned["taken_per_step"] else None),
        "mlp_direction_gain": mlp,
        "low_mode_gain": lm_gain,
        "chosen_vs_plain": round(chosen["steps_per_s"] / plain["steps_per_s"], 3),
        "error_vs_fp64": {
            "definition": "per timed step and species, max_i |u_i - u*_i| / max_i |u*_i| with u* "
                          "the same step's system solved by the plain FP64 path to rtol 1e-13 "
                          "(untimed replay of the timed run)",
            "bar": 1e-8, "legs": errors},
        "at_error_bar": bar,
        "reference_order_step": coupled,
        "iterations_last_step": chosen["iterations_per_step"][-1],
        "outer_last_step": chosen["outer_per_step"][-1],
        "converged": all(r["converged"] for r in res.values()),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": "k_cg_st<3> (fused 3-species FP32 Jacobi-PCG iteration, "
                               "stored geometry)",
                     "algorithmic_bytes_per_launch": nbytes,
                     "us_per_launch": round(ms_iter * 1e3, 2),
                     "traffic_source": traffic["source"] if traffic else None},
        "apply_hbm": {"kernel": "k_mf<float,3,APPLY,stored> (fused 3-species FP32 apply)",
                      "algorithmic_bytes_per_launch": abytes,
                      "us_per_launch": round(ms_apply * 1e3, 2), "achieved_gbs": round(apply_gbs, 1),
                      "frac": round(apply_gbs / HBM_PEAK_GBS, 4),
                      "copy_of_same_bytes": copy_ref},
        "assembly_hbm": {"kernel": {"scalar": "k_assemble_scalar (D K1 + (1/dt + alpha) M1)",
                                    "valpha": "k_assemble_scalar_valpha (D K1, lumped mass and "
                                              "the per-vertex reaction by graph edge)",
                                    "rows": "k_assemble_seg (row assembly from the geometry)"}[
                                        assembly_kernel(C, V, s_asm)]
                                   + ": the FP64 CSR assembly of the plain path's per-step "
                                     "operator, this species",
                         "species": s_asm["name"], "algorithmic_bytes_per_launch": asm_bytes,
                         "us_per_launch": round(ms_asm * 1e3, 2), "achieved_gbs": round(asm_gbs, 1),
                         "frac": round(asm_gbs / HBM_PEAK_GBS, 4),
                         "field": {"kernel": "k_assemble_seg (FP64 CSR row assembly from the "
                                             "geometry, per-cell D)",
                                   "coefficients": s_fld["name"],
                                   "algorithmic_bytes_per_launch": fld_bytes,
                                   "us_per_launch": round(ms_fld * 1e3, 2),
                                   "achieved_gbs": round(fld_gbs, 1),
                                   "frac": round(fld_gbs / HBM_PEAK_GBS, 4)},
                         "scalar": sc_entry},
        "kernels_at_scale": scale,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

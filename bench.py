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


def measured_traffic(path=os.path.join(ROOT, "profiles", "k_cg_disc3_traffic.md")):
    """HBM bytes per launch of the fused CG iteration from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel factors), or None."""
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


def run_path(mesh, w, method, steps, warmup, rtol, warm="previous", epochs=30, history_dirs=8):
    import torch
    from dsv.transport import Stepper
    mlp = None
    if warm == "learned":
        from dsv.mlp import WarmStartMLP
        mlp = WarmStartMLP(mesh.device)
    st = Stepper(mesh, w["species"], w["dt"], method=method, rtol=rtol, phi=w.get("phi"),
                 grad_phi=w.get("grad_norm"), eps=w.get("eps"), mlp=mlp)
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
        "taken_per_step": [e["taken"] for e in log_tail],
        "coef_last_step": log_tail[-1]["coef"] if log_tail else None,
    }, st


def cpu_baseline(w, threads):
    """Oracle CPU path on one 3-species step of the same mesh (bounded sample)."""
    import numpy as np
    from oracle import assembly as OA
    from oracle import solve as OS
    conn, xyz = w["conn"], w["xyz"]
    inv_dt = 1.0 / w["dt"]
    t0 = time.perf_counter()
    for s in w["species"]:
        De = s["D"] if s.get("D") is not None else s["D_value"]
        ip, ix, data = OA.assemble_values(conn, xyz, De, inv_dt, s.get("alpha"), 1, s.get("bc_dofs"))
        b = OA.assemble_rhs(conn, xyz, s["u0"], inv_dt, s.get("f"), 1, s.get("bc_dofs"),
                            s.get("bc_value", 0.0))
        x0 = np.array(s["u0"])
        if s.get("bc_dofs") is not None and len(s["bc_dofs"]):
            x0[s["bc_dofs"]] = s["bc_value"]
        OS.pcg_jacobi_c(ip, ix, data, b, x0, w["rtol"], 20000, threads)
    return time.perf_counter() - t0


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
           "workload": f"{w['name']}: {len(w['conn'])} tets, interface speed 0.008 O2, "
                       "coefficients recomputed on the GPU every step", "paths": {}}
    for warm in ("previous", "extrapolate", "learned"):
        st = Stepper(mesh, w["species"], w["dt"], method="mf32_ir64", rtol=w["rtol"],
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
        out["paths"][warm] = {"steps_per_s": round((n - half) / dt, 2),
                              "cg_iterations": sum(sum(i) for i in its),
                              "converged": all(all(h["converged"]) for h in st.history),
                              "training": train}
        log(f"series {warm}: {out['paths'][warm]}")
    base = out["paths"]["previous"]["cg_iterations"]
    for p in out["paths"].values():
        p["cg_iterations_saved_pct"] = round(100 * (1 - p["cg_iterations"] / base), 1)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=8,
                    help="untimed steps of every path; the learned warm start trains on them")
    ap.add_argument("--workload", default="disc3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the kernel timings on the 14.7M-tet gyroid lattice")
    ap.add_argument("--apply-reps", type=int, default=50)
    ap.add_argument("--history-dirs", type=int, default=8,
                    help="past solution increments in the Galerkin warm start (1..8)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--series", action="store_true",
                    help="instead of the headline: the receding-interface time series on the "
                         "disc (200 steps; the MLP trains on steps 1-100), one JSON line")
    ap.add_argument("--paths", default="plain,fp32,chosen,learned",
                    help="comma list of the paths to run (the JSON line needs all four)")
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
    if a.series:
        run_series(mesh, w, history_dirs=a.history_dirs)
        return
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
    res = {}
    ref_u = None
    paths = a.paths.split(",")
    for name, method, warm in (("plain", "csr64", "previous"), ("fp32", "mf32_ir64", "previous"),
                               ("chosen", "mf32_ir64", "extrapolate"),
                               ("learned", "mf32_ir64", "learned")):
        if name not in paths:
            continue
        r, st = run_path(mesh, w, method, a.steps, a.warmup, rtol, warm=warm,
                         history_dirs=a.history_dirs)
        res[name] = r
        log(f"{name} ({method}, {warm} start): {r['steps_per_s']:.2f} steps/s, "
            f"iterations/step {r['iterations_per_step']}, converged={r['converged']}")
        if name == "fp32":
            ref_u = [u.clone() for u in st.u]
        elif ref_u is not None:
            dev = max(float((u - v).abs().max() / v.abs().max().clamp_min(1e-300))
                      for u, v in zip(st.u, ref_u))
            r["max_rel_diff_vs_fp32_previous_start"] = dev
            log(f"{name} vs fp32 previous-start solutions after the timed steps: "
                f"max rel diff {dev:.2e}")
        del st

    if any(p not in res for p in ("plain", "fp32", "chosen", "learned")):
        log("not every path ran: no JSON line")
        return
    # roofline of the fused FP32 3-species CG iteration (the chosen path's dominant kernel)
    ms_iter, ms_apply = time_kernels(mesh, w, a.apply_reps)
    nbytes = cg_iteration_bytes(C, V, w["species"])
    abytes = apply_bytes(C, V, w["species"])
    achieved = nbytes / (ms_iter * 1e-3) / 1e9
    apply_gbs = abytes / (ms_apply * 1e-3) / 1e9
    log(f"fused FP32 CG iteration: {ms_iter * 1e3:.1f} us/launch, {achieved:.0f} GB/s; "
        f"matrix-free apply alone {ms_apply * 1e3:.1f} us, {apply_gbs:.0f} GB/s")
    traffic = measured_traffic()

    scale = None
    if not a.no_scale and a.workload == "disc3":
        scale = kernels_at_scale()

    cpu = None
    if not a.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        t_cpu = cpu_baseline(w, threads)
        cpu = {"value": 1.0 / t_cpu, "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": f"one 3-species BE step of {w['name']} (assembly numpy + C OpenMP "
                         f"Jacobi-PCG to rtol {rtol:g}), {t_cpu:.1f}s"}
        log(f"cpu baseline: {t_cpu:.1f}s per step on {threads} threads")

    chosen, plain, fp32, learned = res["chosen"], res["plain"], res["fp32"], res["learned"]

    def its(r):
        return round(sum(sum(i) for i in r["iterations_per_step"]) / a.steps, 1)

    def saved(r):
        return round(100 * (1 - its(r) / its(fp32)), 1)

    out = {
        "metric": "3-species BE steps/s at 800k tets, rtol 1e-8: chosen vs plain FP64; HBM %",
        "value": round(chosen["steps_per_s"], 3), "unit": "steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(chosen["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{w['name']}: graded Delaunay disc r=1 h=0.2, {C} tets, {V} "
                               f"vertices, O2/Cl/OH D=2.0/2.0/5.3, dt={w['dt']}, rtol={rtol:g}",
                   "method": "FP32 fused 3-species Jacobi-PCG (Chronopoulos-Gear, one kernel per "
                             "iteration, stored FP32 geometry) in FP64 refinement, FP64 "
                             f"Galerkin warm start over the last {a.history_dirs} solution "
                             "increments (previous solution as fallback), species narrowing, "
                             "8-iteration polled chunks of direct launches",
                   "parallelism": "1 GPU", "launch_shape": tuned},
        "plain_fp64": {"steps_per_s": round(plain["steps_per_s"], 3),
                       "cg_iterations_per_step": its(plain)},
        "fp32_previous_start": {"steps_per_s": round(fp32["steps_per_s"], 3),
                                "cg_iterations_per_step": its(fp32)},
        "chosen_galerkin_extrapolation": {"steps_per_s": round(chosen["steps_per_s"], 3),
                                          "cg_iterations_per_step": its(chosen),
                                          "cg_iterations_saved_pct_vs_previous_start": saved(chosen),
                                          "max_rel_diff_vs_fp32_previous_start":
                                              chosen.get("max_rel_diff_vs_fp32_previous_start")},
        "learned_warm_start": {"steps_per_s": round(learned["steps_per_s"], 3),
                               "cg_iterations_per_step": its(learned),
                               "cg_iterations_saved_pct_vs_previous_start": saved(learned),
                               "max_rel_diff_vs_fp32_previous_start":
                                   learned.get("max_rel_diff_vs_fp32_previous_start"),
                               "training": learned["train"],
                               "mlp_coef_last_step": learned["coef_last_step"][-1]
                               if learned["coef_last_step"] else None,
                               "taken_last_step": learned["taken_per_step"][-1]
                               if learned["taken_per_step"] else None},
        "chosen_vs_plain": round(chosen["steps_per_s"] / plain["steps_per_s"], 3),
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
                      "frac": round(apply_gbs / HBM_PEAK_GBS, 4)},
        "kernels_at_scale": scale,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""SIMP iterations per second of the MI355X SIMP compliance loop.

Workload (default): the 256x128x128 voxel hexa8 cantilever (4,194,304
elements, 12,780,288 free DOFs): x = 0 clamped, unit total -z load on the
bottom edge of x = 256, E0 = 1, Emin = 1e-9, nu = 0.3, p = 3, rmin = 1.5,
initial design the seeded Gaussian field (32 blobs, sigma 4-12, mean 0.3,
seed 3).  One step is one SIMP design iteration on the GPU: density filter ->
equilibrium solve to the verified FP64 relative residual -> compliance and
sensitivities -> sensitivity filter -> OC update.  Inputs are resident in HBM
when the timed region starts.

Tolerance: 1e-8 (the metric's) where FP64 reaches it; 2e-8 on the
256x128x128 field, where 47% of the filtered densities lie below 0.01 and
the displacements of near-void regions reach 3.9e7: rounding the solution
itself to FP64 leaves a relative residual of 1.4-1.9e-8 on the first design,
so no FP64 iterate verifies at 1e-8, and 2e-8 is the tightest tolerance every
solve of the loop reaches (the chosen value is in config.workload).

The default method solves with PCG preconditioned by a geometric multigrid
V-cycle (FP64, same verified stop).  Beside the value, on the same workload
and tolerance: the plain method of the reference (FP64 Jacobi-PCG: its SIMP
loop when it reaches the tolerance within --plain-maxit iterations, else the
iterations it ran, their rate and the residual reached), one solve per
precision variant (multigrid with the FP32 fine level, Jacobi-PCG with the
FP32 and the BF16 MFMA contraction inside FP64 refinement), the dominant
kernel against its roofline, the whole iteration's fine-level bytes against
HBM, and the CPU baseline.

Prints one JSON line (the benchmark contract)."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    "c1": dict(dims=(60, 20, 20), rmin=1.5, field=None),
    "c2": dict(dims=(192, 64, 64), rmin=1.5, field=2),
    "c3": dict(dims=(256, 128, 128), rmin=1.5, field=3),
    "c4": dict(dims=(128, 64, 64), rmin=1.5, field=4),
    # 24 independent problems of uneven size (icarat2_amd.workloads)
    "c5": dict(dims=None, rmin=None, field=None),
}
# tolerance of each workload: the metric's 1e-8 where FP64 reaches it
CONFIG_TOL = {"c1": 1e-8, "c2": 1e-8, "c3": 2e-8, "c4": 1e-8, "c5": 1e-8}
METHODS = {
    # name: (precision, preconditioner)
    "fp64": (0, 0),
    "fp32": (1, 0),
    "bf16": (2, 0),
    "learned": (1, 1),
    # geometric multigrid V-cycle as the PCG preconditioner (FP64)
    "mg": (0, 2),
    # the same with the V-cycle's fine level in FP32 (vectors, K.u, transfers)
    "mg32": (1, 2),
}
MI355X_PEAK = {"fp64_tflops": 78.6, "fp32_mfma_tflops": 157.3, "bf16_mfma_tflops": 2500.0,
               "hbm_gbs": 8000.0}
# dominant kernel of each method: (timing-hook id, kernel, compute peak key,
# key of its PMC traffic line)
DOMINANT = {
    "fp64": (0, "k_stencil64z<PCG, cube, fused beta> (FP64 VALU, KE0 orbit values in SGPRs)",
             "fp64_tflops", "k_stencil64z<true,2,true,0>"),
    "learned": (0, "k_stencil64z<PCG, cube, fused beta> (FP64 VALU, KE0 orbit values in SGPRs)",
                "fp64_tflops", "k_stencil64z<true,2,true,0>"),
    # the kernel with the largest time share of a multigrid PCG iteration
    "mg": (10, "k_stencil64z<PCG, cube> of the multigrid solve (p = z + beta p_old, q = K p; "
               "FP64 VALU, KE0 orbit values in SGPRs)", "fp64_tflops",
           "k_stencil64z<true,2,false,0>"),
    "mg32": (10, "k_stencil64z<PCG, cube> of the multigrid solve (p = z + beta p_old, q = K p; "
                 "FP64 VALU, KE0 orbit values in SGPRs)", "fp64_tflops",
             "k_stencil64z<true,2,false,0>"),
    "fp32": (5, "k_stencil_mfma<FP32> (v_mfma_f32_16x16x4_f32, FP64 vectors)", "fp32_mfma_tflops",
             "k_stencil_mfma<0>"),
    "bf16": (6, "k_stencil_mfma<BF16> (v_mfma_f32_16x16x32_bf16, FP64 vectors)", "bf16_mfma_tflops",
             "k_stencil_mfma<1>"),
}


# the label of a compute roofline by the peak it is priced against
BOUND_OF_PEAK = {"fp64_tflops": "fp64-valu", "fp32_mfma_tflops": "mfma",
                 "bf16_mfma_tflops": "mfma"}


def traffic_from_profiles(key, workload):
    """HBM bytes per launch of a kernel on a workload from the committed PMC
    summaries (profiles/pmc_traffic_*.txt: FETCH_SIZE + WRITE_SIZE scaled by
    the copy kernel's calibration; lines `TRAFFIC_BYTES <kernel> <bytes>` after
    a line `WORKLOAD <name>`), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not (name.startswith("pmc_traffic_") and name.endswith(".txt")):
            continue
        cur = None
        with open(os.path.join(d, name)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and parts[0] == "WORKLOAD":
                    cur = parts[1]
                if (len(parts) == 3 and parts[0] == "TRAFFIC_BYTES" and parts[1] == key and
                        cur == workload):
                    try:
                        return float(parts[2])
                    except ValueError:
                        return None
    return None


def mg_fine_bytes(numnp, nelx):
    """Compulsory HBM bytes of one multigrid PCG iteration on the fine level
    (coarse levels excluded), V = one FP64 lattice vector (3 numnp doubles):
    PCG stencil z, p_old in, p, q out (4V); update x, its compensation part xlo,
    p, r, q, dinv in, x, xlo, r, z0 out (10V); residual K z0 (2V); restriction
    r, K z0 in (2V); prolongation x in and out (2V); fused last sweep x, b,
    dinv, q in, z out (5V); young and the free mask for each of the three fine
    K.u passes."""
    V = 3.0 * numnp * 8
    return 25.0 * V + 3.0 * (8.0 * nelx + numnp)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_problem(A, cfg, helm=False):
    nx, ny, nz = cfg["dims"]
    c = A.cantilever(nx, ny, nz)
    if cfg["field"] is None:
        import numpy as np
        x0 = np.full(c.nelx, 0.3)
    else:
        x0, _ = A.gaussian_field(nx, ny, nz, cfg["field"])
    c.filter_setup(cfg["rmin"])
    if helm:
        c.helm_setup(cfg["rmin"])
    return c, x0


def run_gpu(A, cfg, method, steps, warmup, tol, helm=False):
    prec, pre = METHODS[method]
    c, x0 = make_problem(A, cfg, helm)
    kw = dict(tol=tol, maxit=100000, precision=prec, precond=pre, helmholtz=helm)
    if pre == 1:
        # the network is trained online on the Jacobi solves of the warm-up
        # iterations (at least one); the timed iterations use the trained network
        c.learned_setup(train_iters=max(warmup, 1), train_steps=400 * max(warmup, 1))
        warmup = max(warmup, 1)
    if warmup > 0:
        log(f"[bench] warmup {warmup} SIMP iteration(s), method {method}")
        c.simp_run(x0, warmup, **kw)
    log(f"[bench] timed {steps} SIMP iteration(s), method {method}")
    t0 = time.perf_counter()
    x, hist, st = c.simp_run(x0, steps, **kw)
    dt = time.perf_counter() - t0
    return c, x0, x, hist, st, dt


def cpu_baseline(cfg, pcg_iters_per_solve, n_cg=60, helm=False):
    """Times the FP64 CPU oracle (assembled CSR + Jacobi-PCG + filter + OC) on a
    bounded sample of one SIMP iteration and scales the PCG part to the
    iteration count the GPU needed on the same design.  With the Helmholtz
    filter, one filter solve (Jacobi-PCG to 1e-10, as the reference's Eigen_CG)
    is timed and counted once per filter call: 3 per iteration plus one per OC
    bisection step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    nx, ny, nz = cfg["dims"]
    pr = O.cantilever(nx, ny, nz)
    if cfg["field"] is None:
        x0 = np.full(pr.nelx, 0.3)
    else:
        x0, _ = O.gaussian_field(nx, ny, nz, cfg["field"])
    F = O.Filter(pr, cfg["rmin"])
    t = time.perf_counter()
    xf = F.density_filter(x0, False)
    t_filter = time.perf_counter() - t
    t = time.perf_counter()
    pr.set_density(xf)
    t_asm = time.perf_counter() - t
    t = time.perf_counter()
    u, it, _ = pr.scaling_cg(tol=1e-30, maxit=n_cg - 1)
    t_cg = (time.perf_counter() - t) / it
    t = time.perf_counter()
    nv = pr.node_values(u)
    df, dg = pr.sensitivity(nv)
    t_sens = time.perf_counter() - t
    t = time.perf_counter()
    dcf = F.density_filter(df, True)
    dvf = F.density_filter(dg, True)
    t_filter += time.perf_counter() - t
    t = time.perf_counter()
    _, oc_steps, _ = F.oc_update(x0, dcf, xf.sum(), dvf)
    t_oc = time.perf_counter() - t
    helm_note = ""
    if helm:
        H = O.HelmFilter((nx, ny, nz), cfg["rmin"], tol=1e-10)
        t = time.perf_counter()
        H.density_filter(x0)
        t_h = time.perf_counter() - t
        t_filter = 3 * t_h
        t_oc += oc_steps * t_h
        helm_note = (f"; Helmholtz filter: one solve timed ({t_h:.2f}s, {H.iters} CG "
                     f"iterations), counted 3 + {oc_steps} (OC steps) times")
    per_iter = t_filter + t_asm + t_cg * pcg_iters_per_solve + t_sens + t_oc
    return {
        "value": 1.0 / per_iter,
        "unit": "SIMP iters/s",
        "cores": O.lib().orc_num_threads(),
        "kind": "port",
        "sample": (f"CPU FP64 restatement of ICARAT's algorithm (reference Eigen build "
                   f"unavailable offline) on {nx}x{ny}x{nz}: filter + CSR assembly "
                   f"({t_asm:.2f}s) + {it} timed Jacobi-PCG iterations ({t_cg * 1e3:.1f} ms each, "
                   f"scaled to the GPU's {pcg_iters_per_solve:.0f} iterations/solve) + "
                   f"sensitivity + 2 chain-rule filters + OC ({t_oc:.2f}s)" + helm_note),
    }


def run_batch(A, a):
    """The batched workload: 24 seeded problems (sizes 32^3..160x80x80, 1..4 load
    cases each); each problem runs its own design loop and the running solves
    share one captured graph per launch."""
    import numpy as np
    from icarat2_amd import workloads as W
    pre = METHODS[a.method][1]
    specs = W.batch_workload()
    ctxs = [W.build_context(A, s) for s in specs]
    x0s = [W.initial_design(s) for s in specs]
    for c, s, x in zip(ctxs, specs, x0s):
        c.set_density(c.density_filter(x))
    b = A.Batch(ctxs)
    if a.budget is not None:
        b.set_budget(a.budget)
    log("[bench] autotune")
    cfg, cfg_ms = b.autotune(reps=3)

    def timed(precond):
        if a.warmup > 0:
            log(f"[bench] warmup {a.warmup} batched SIMP iteration(s), precond {precond}")
            b.simp_run(x0s, a.warmup, tol=a.tol, maxit=100000, precond=precond)
        log(f"[bench] timed {a.steps} batched SIMP iteration(s) of {len(specs)} problems")
        t0 = time.perf_counter()
        out = b.simp_run(x0s, a.steps, tol=a.tol, maxit=100000, precond=precond)
        return out, time.perf_counter() - t0

    (xs, hists, sts), dt = timed(pre)
    # the same problems one after the other (each alone on the GPU)
    log("[bench] sequential reference: each problem alone")
    t1 = time.perf_counter()
    for c, x in zip(ctxs, x0s):
        c.simp_run(x, a.steps, tol=a.tol, maxit=100000, precond=pre)
    dt_seq = time.perf_counter() - t1
    iters = [float(h[:, 2].sum()) / a.steps for h in hists]  # PCG iterations per design iteration
    out = {
        "metric": "SIMP iters/s + solve time to 1e-8 FP64 residual: learned/mixed vs plain FP64",
        "value": len(specs) * a.steps / dt,
        "unit": "SIMP iters/s (all problems)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded problem specs, uniform initial designs)",
        "config": {"workload": "c5: 24 independent problems (seeds 100..123), 32^3..160x80x80, "
                               "1-4 face load cases each, one design loop per problem, OC",
                   "method": a.method,
                   "elements": int(sum(c.nelx for c in ctxs)),
                   "load_cases": int(sum(len(s["loads"]) for s in specs)),
                   "autotune_config": cfg, "autotune_ms_per_launch": cfg_ms,
                   "admission_budget": a.budget if a.budget is not None else "default"},
        "sequential_value": len(specs) * a.steps / dt_seq,
        "batch_speedup_vs_sequential": dt_seq / dt,
        "max_true_relres": max(st.max_true_relres for st in sts),
        "pcg_iters_per_design_iter": iters,
        "graph_launches": b.info()[1],
    }
    if not a.no_cases:
        out["load_cases_together"] = cases_leg(A, specs, a.tol, pre)
    if pre != 0 and not a.no_plain:
        (xp, hp, sp), dtp = timed(0)
        out["plain_fp64"] = {
            "method": "fp64 (Jacobi-PCG), batched", "value": len(specs) * a.steps / dtp,
            "max_true_relres": max(st.max_true_relres for st in sp),
            "design_max_abs_diff": float(max(np.max(np.abs(x1 - x2)) for x1, x2 in zip(xs, xp)))}
        out["speedup_vs_plain_fp64"] = out["value"] / out["plain_fp64"]["value"]
    return out


def cases_leg(A, specs, tol, precond):
    """One FP64 solve (Jacobi-PCG or multigrid PCG, the bench's method) of
    every load case of each multi-case problem (initial design): all cases in
    one multi-column iteration (icarat_hip_solve_cases) against one case
    after another; iteration counts must agree."""
    from icarat2_amd import workloads as W
    mg = precond == A.PRECOND_MG
    rows, t_all, t_seq, t_auto = [], 0.0, 0.0, 0.0
    for s in specs:
        if len(s["loads"]) < 2:
            continue
        c = W.build_context(A, s)
        c.set_density(c.density_filter(W.initial_design(s)))
        kw = dict(tol=tol, maxit=100000, want_u=False, precond=precond)
        c.solve_cases(**kw)  # graphs built
        t0 = time.perf_counter()
        _, st = c.solve_cases(**kw)
        t1 = time.perf_counter()
        its = []
        for lc in range(c.load_case_count()):
            c.select_load_case(lc)
            _, s1 = c.solve(**kw)
            its.append(s1.iters)
        t2 = time.perf_counter()
        if [x.iters for x in st] != its:
            raise RuntimeError("load cases together: iteration counts differ from single solves")
        rows.append({"dims": "x".join(map(str, s["dims"])), "cases": len(its),
                     "together_ms": (t1 - t0) * 1e3, "one_after_another_ms": (t2 - t1) * 1e3})
        t_all += t1 - t0
        t_seq += t2 - t1
        # simp_run's automatic mode: multigrid together at every size,
        # Jacobi-PCG together up to 600,000 elements
        t_auto += (t1 - t0) if (mg or c.nelx <= 600000) else (t2 - t1)
    log(f"[bench] load cases together {t_all:.3f} s, one after another {t_seq:.3f} s")
    return {"method": ("fp64 multigrid PCG" if mg else "fp64 Jacobi-PCG")
                      + ", one solve of each multi-case problem's cases",
            "together_s": t_all, "one_after_another_s": t_seq,
            "speedup": t_seq / t_all if t_all > 0 else None,
            "automatic_s": t_auto, "automatic_speedup": t_seq / t_auto if t_auto > 0 else None,
            "problems": rows}


def precision_leg(A, cfg, method, tol, maxit, ref):
    """One solve of the first design (filtered initial field) with a precision
    variant, bounded by maxit: iterations, time, the verified residual, and the
    compliance against the value method's solve of the same design."""
    prec, pre = METHODS[method]
    c, x0 = make_problem(A, cfg)
    c.set_density(c.density_filter(x0))
    log(f"[bench] one solve, {method}, maxit {maxit}")
    try:
        u, st = c.solve(tol=tol, maxit=maxit, precision=prec, precond=pre, want_u=False,
                        allow_cap=True)
    except A.IcaratError as e:
        # e.g. a breakdown of the low-precision Jacobi-PCG on this contrast
        return {"reached": False, "error": str(e)}
    comp, _, _ = c.compliance_sensitivity()
    out = {"reached": bool(st.status == 0 and st.true_relres <= tol), "iterations": int(st.iters),
           "verifications": int(st.refine_steps), "solve_s": st.seconds,
           "ms_per_iteration": st.seconds / max(st.iters, 1) * 1e3,
           "true_relres": st.true_relres,
           "compliance_rel_diff_vs_value_method": abs(comp - ref) / abs(ref)}
    if pre == 2:
        out["fine_level_fp32"] = bool(c.mg_status().fine_fp32)
    del c
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--method", default="mg", choices=sorted(METHODS))
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 Jacobi-PCG run beside a faster method")
    ap.add_argument("--plain-maxit", type=int, default=20000,
                    help="iteration bound of the plain Jacobi-PCG solve of the first design")
    ap.add_argument("--no-precisions", action="store_true",
                    help="skip the one-solve precision variants")
    ap.add_argument("--tol", type=float, default=None,
                    help="verified FP64 relative residual (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--reps", type=int, default=20, help="launches timed for the roofline")
    ap.add_argument("--filter", default="cone", choices=["cone", "helm"],
                    help="density filter of the loop: cone (filter.hpp) or Helmholtz "
                         "(filter_helm.hpp)")
    ap.add_argument("--budget", type=int, default=None,
                    help="c5: admission budget in elements (0: all problems at once)")
    ap.add_argument("--no-cases", action="store_true",
                    help="c5: skip the load-cases-together leg")
    ap.add_argument("--roofline-only", action="store_true",
                    help="time only the dominant, vector and copy kernels on the first design "
                         "(for profiler passes); prints their figures, not the contract line")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU (slab "
              "decomposition over RCCL) path is not built", file=sys.stderr)
        sys.exit(2)
    if a.tol is None:
        a.tol = CONFIG_TOL[a.config]
    from icarat2_amd import _capi as A
    if not A.device_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    import numpy as np
    cfg = CONFIGS[a.config]
    if a.config == "c5":
        if a.method not in ("fp64", "mg"):
            print("bench.py: the batched workload runs FP64 Jacobi-PCG or multigrid PCG",
                  file=sys.stderr)
            sys.exit(2)
        print(json.dumps(run_batch(A, a)), flush=True)
        return
    helm = a.filter == "helm"
    dims_s = "x".join(map(str, cfg["dims"]))
    if a.roofline_only:
        c, x0 = make_problem(A, cfg, helm)
        c.set_density(c.density_filter(x0))
        kid = DOMINANT[a.method][0]
        res = {}
        hooks = [("dominant", kid), ("vector", 11 if METHODS[a.method][1] == 2 else 1),
                 ("copy", 7)]
        if helm:
            hooks.append(("helmholtz_filter", 9))
        for name, k in hooks:
            ms, fl, by = c.time_kernel(k, a.reps)
            res[name] = {"timing_hook": k, "avg_ms": ms, "flops": fl, "bytes": by}
            log(f"[bench] {name} kernel (hook {k}): {ms * 1e3:.2f} us")
        print(json.dumps({"workload": dims_s, "method": a.method, "kernels": res}), flush=True)
        return
    c, x0, x, hist, st, dt = run_gpu(A, cfg, a.method, a.steps, a.warmup, a.tol, helm)
    pcg_per_solve = float(hist[:, 2].mean())
    # the plain method (FP64 Jacobi-PCG) on the same workload and tolerance:
    # first one bounded solve of the first design; its SIMP loop only if that
    # solve reaches the tolerance
    plain = None
    jacobi_iters = pcg_per_solve
    plain_reached = True
    if a.method != "fp64" and not a.no_plain:
        cp, x0p = make_problem(A, cfg, helm)
        cp.set_density(cp.density_filter(x0p))
        log(f"[bench] plain FP64 Jacobi-PCG: one solve of the first design, maxit {a.plain_maxit}")
        _, s1 = cp.solve(tol=a.tol, maxit=a.plain_maxit, want_u=False, allow_cap=True)
        plain_reached = s1.status == 0 and s1.true_relres <= a.tol
        first = {"iterations": int(s1.iters), "solve_s": s1.seconds,
                 "ms_per_iteration": s1.seconds / max(s1.iters, 1) * 1e3,
                 "true_relres": s1.true_relres, "reached": bool(plain_reached)}
        del cp
        if plain_reached:
            cp, _, xp, hp, sp, dtp = run_gpu(A, cfg, "fp64", a.steps, a.warmup, a.tol, helm)
            jacobi_iters = float(hp[:, 2].mean())
            plain = {"method": "fp64 (Jacobi-PCG)", "value": a.steps / dtp,
                     "solve_s_per_iter": sp.solve_s / a.steps,
                     "pcg_iters_per_solve": jacobi_iters, "max_true_relres": sp.max_true_relres,
                     "design_max_abs_diff": float(np.max(np.abs(x - xp))),
                     "compliance_max_rel_diff": float(np.max(np.abs(hist[:, 0] - hp[:, 0]) /
                                                             np.abs(hp[:, 0]))),
                     "first_solve": first}
            del cp
        else:
            jacobi_iters = float(a.plain_maxit)
            plain = {"method": "fp64 (Jacobi-PCG)", "value": None,
                     "not_reached": f"the tolerance {a.tol:g} is not reached in "
                                    f"{s1.iters} iterations (true residual "
                                    f"{s1.true_relres:.3g}) on the first design",
                     "first_solve": first,
                     # a solve needs more than this: the multigrid loop's rate over
                     # it is a lower bound of the speedup
                     "speedup_lower_bound": (s1.seconds / (st.solve_s / a.steps))}
    # one solve of the first design per precision variant, same tolerance
    variants = None
    if not a.no_precisions and not helm and a.method in ("mg", "fp64"):
        c.set_density(c.density_filter(x0))
        _, s_ref = c.solve(tol=a.tol, maxit=100000, precision=METHODS[a.method][0],
                           precond=METHODS[a.method][1], want_u=False, allow_cap=True)
        ref_c, _, _ = c.compliance_sensitivity()
        variants = {"value_method_first_solve": {"iterations": int(s_ref.iters),
                                                 "verifications": int(s_ref.refine_steps),
                                                 "solve_s": s_ref.seconds,
                                                 "true_relres": s_ref.true_relres,
                                                 "reached": bool(s_ref.status == 0)}}
        bound = {"mg32": 60000, "fp32": 3000 if a.config == "c3" else 30000,
                 "bf16": 3000 if a.config == "c3" else 30000}
        for m in ("mg32", "fp32", "bf16"):
            variants[m] = precision_leg(A, cfg, m, a.tol, bound[m], ref_c)
    # roofline of the dominant kernel, timed with HIP events on the library's
    # stream (the stream the kernel is launched on)
    c.set_density(c.density_filter(x0))
    kid, kname, peak_key, tkey = DOMINANT[a.method]
    ms, flops, nbytes = c.time_kernel(kid, a.reps)
    t_s = ms * 1e-3
    peak_f = MI355X_PEAK[peak_key]
    compute_bound = flops / (peak_f * 1e12) >= nbytes / (MI355X_PEAK["hbm_gbs"] * 1e9)
    if compute_bound:
        achieved, peak, unit = flops / t_s / 1e12, peak_f, "TFLOP/s"
    else:
        achieved, peak, unit = nbytes / t_s / 1e9, MI355X_PEAK["hbm_gbs"], "GB/s"
    mg_like = METHODS[a.method][1] == 2
    vk = 11 if mg_like else 1
    ms_v, _, bytes_v = c.time_kernel(vk, a.reps)
    ms_c, _, bytes_c = c.time_kernel(7, a.reps)
    tol_note = ("" if a.tol <= 1e-8 else
                " -- the tightest verified FP64 tolerance this field reaches reliably; 1e-8 is "
                "below the FP64 representation floor of its solution")
    out = {
        "metric": "SIMP iters/s + solve time to 1e-8 FP64 residual: learned/mixed vs plain FP64",
        "value": a.steps / dt,
        "unit": "SIMP iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mg": "f64", "mg32": "f32+f64", "fp32": "f32+f64",
                  "learned": "f32+f64", "bf16": "bf16+f64"}[a.method],
        "data": "synthetic (seeded Gaussian-field densities, no external data)",
        "config": {"workload": f"{a.config}: {dims_s} voxel hexa8 cantilever SIMP "
                               f"({'Helmholtz ' if helm else ''}filter, solve to a {a.tol:g} "
                               f"verified FP64 relative residual{tol_note}, sensitivities, OC)",
                   "method": a.method, "tol": a.tol, "elements": int(c.nelx),
                   "free_dofs": int(c.numeq), "rmin": cfg["rmin"],
                   "filter": "helmholtz" if helm else "cone"},
        "solve_s_per_iter": st.solve_s / a.steps,
        "pcg_iters_per_solve": pcg_per_solve,
        "max_true_relres": st.max_true_relres,
        "train_s_timed": st.train_s,
        "fallbacks": int(st.fallbacks),
        "roofline": {"bound": BOUND_OF_PEAK[peak_key] if compute_bound else "hbm",
                     "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
                     "traffic": traffic_from_profiles(tkey, dims_s),
                     "kernel": kname, "avg_ms": ms, "algorithmic_flops": flops,
                     "compulsory_bytes": nbytes, "hbm_gbs_at_compulsory_bytes": nbytes / t_s / 1e9},
        "vector_kernel": {"kernel": ("k_fpcg_update_alpha<compensated x>" if mg_like
                                     else "k_pcg_update64f"),
                          "avg_ms": ms_v, "bytes": bytes_v, "gbs": bytes_v / (ms_v * 1e-3) / 1e9},
        "copy_kernel": {"kernel": "k_copy64", "avg_ms": ms_c, "bytes": bytes_c,
                        "gbs": bytes_c / (ms_c * 1e-3) / 1e9},
    }
    if mg_like:
        out["config"]["mg_levels"] = c.mg_info()
        ms_ = c.mg_status()
        out["config"]["mg_w_levels"] = int(ms_.w_levels)
        out["mg_status"] = {"factorizations": int(ms_.factorizations), "shifted": int(ms_.shifted),
                            "w_levels": int(ms_.w_levels), "w_fallbacks": int(ms_.w_fallbacks)}
        n_it = float(hist[:, 2].sum())
        if n_it > 0:
            t_it = st.solve_s / n_it
            b_it = mg_fine_bytes(float(c.numnp), float(c.nelx))
            out["iteration_roofline"] = {
                "scope": "fine-level compulsory bytes of one multigrid PCG iteration (25 lattice "
                         "vectors + young and mask per fine K.u; coarse levels -- twice per "
                         "visit in the W-cycle phase --, launch gaps and the per-solve set-up "
                         "in the time, not in the bytes)",
                "bytes": b_it, "s_per_iteration": t_it, "gbs": b_it / t_it / 1e9,
                "frac": b_it / t_it / 1e9 / MI355X_PEAK["hbm_gbs"]}
    # the same K.u (y = K x) on the VALU and on the matrix cores (FP64 MFMA
    # element contraction); the solver runs the VALU kernel
    ms_k8, fl_k8, _ = c.time_kernel(8, a.reps)
    ms_k12, fl_k12, _ = c.time_kernel(12, a.reps)
    out["k_apply_variants"] = {
        "fp64_valu": {"kernel": "k_stencil64z<false, SYM, false, false>", "avg_ms": ms_k8,
                      "tflops": fl_k8 / (ms_k8 * 1e-3) / 1e12},
        "fp64_mfma": {"kernel": "k_apply64_mfma (v_mfma_f64_4x4x4_4b)", "avg_ms": ms_k12,
                      "tflops": fl_k12 / (ms_k12 * 1e-3) / 1e12},
        "used": "fp64_valu" if ms_k8 <= ms_k12 else "fp64_mfma (not wired into the solver)"}
    if helm:
        ms_h, fl_h, by_h = c.time_kernel(9, a.reps)
        out["helm_filter"] = {"kernel": "Helmholtz filter pass (element loads, 6 FP64 GEMM "
                                        "transforms + spectral scaling, read-back)",
                              "avg_ms": ms_h, "flops": fl_h,
                              "tflops": fl_h / (ms_h * 1e-3) / 1e12,
                              "frac_fp64_peak": fl_h / (ms_h * 1e-3) / 1e12 /
                              MI355X_PEAK["fp64_tflops"]}
    if plain is not None:
        out["plain_fp64"] = plain
        out["speedup_vs_plain_fp64"] = (out["value"] / plain["value"]
                                        if plain.get("value") else None)
    if variants is not None:
        out["precision_variants"] = variants
    del c
    if not a.no_cpu:
        # the CPU runs the reference's Jacobi-PCG: scaled by Jacobi's iterations
        log("[bench] cpu baseline sample")
        cb = cpu_baseline(cfg, jacobi_iters, helm=helm)
        if not plain_reached:
            cb["sample"] += (f"; Jacobi-PCG does not reach {a.tol:g} in {a.plain_maxit} "
                             f"iterations, so this value (at {a.plain_maxit} iterations per "
                             f"solve) is an upper bound")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Newton-step benchmark of the MI355X Kadath/FUKA path.

Workload (the spinning-BH configuration on one GPU): XCTS black-hole system on a
Space_spheric with 5 solved domains (4 shells + compactified) at 17x13x12
collocation points, n = 58905 unknowns, dense Jacobian 27.8 GB in FP64; seeded
1e-3 perturbation of the flat guess; horizon rotation Ome = 0.05.

One "step" = one System_of_eqs::do_newton iteration: residual F(u) + batched-JVP
Jacobian fill + factorisation + solve (+ fp64 refinement) + field update, all
device-resident.  The K timed steps are the iterations of the Newton solve from
the perturbed guess (restarting from the guess once ||F||_inf < 1e-10).  The
system is assembled in domain-major order (block-tridiagonal J) and the FP64 LU
skips the structurally zero tile pairs.  The headline solver is FP64 (the one
that reaches 1e-10 here); FP32-LU + FP64 GMRES-IR runs beside it to the same
test, with time to solution and the converged fields compared.

Prints ONE JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RES = (17, 13, 12)
BOUNDS = [1.0, 2.0, 4.0, 8.0, 16.0]
N0, A, OME = 0.5, 1.0, 0.05
SEED = 20261202
PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # MI355X dense matrix TFLOP/s (vendor figures)
# HBM bytes per launch, from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE),
# FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0: the factors at which the same
# counters read a 256 MiB device-to-device copy (131093.6 KiB fetched, 262144.0 KiB
# written).
#   apply3d: over the residual and Jacobian fill of this workload (--fill-only, 6566
#     launches, 3.405e11 algorithmic bytes), traffic / algorithmic bytes = 1.0040; the
#     bench reports that ratio times its live algorithmic bytes per launch.
#   gemm_k4: mean over all launches of three FP64 Newton iterations of this workload,
#     measured on an earlier build of the LU (same GEMM kernel; 7580 launches per
#     iteration against 8256 now).
APPLY3D_TRAFFIC_RATIO = (2.0 * 83618276.7 + 166635597.5) * 1024.0 / 340514582928.0
GEMM_TRAFFIC_B = (2.0 * 18463.96 + 23028.09) * 1024.0
GEMM_TRAFFIC_NOTE = ("rocprofv3 --pmc of the same GEMM kernel on an earlier build of the LU (7580 launches per "
                     "iteration, mean over all); the pass over the current build crashed inside the profiler's "
                     "host code and was not repeated")


def cpu_baseline(budget_s=25.0):
    """Oracle (numpy/scipy restatement of the reference's FP64 Newton step) timed on
    the host on a bounded sample, extrapolated to one full step of the same workload."""
    import numpy as np
    import scipy.linalg
    from oracle import xcts as OX
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sp = OX.Space(RES, BOUNDS)
    sy = OX.XCTSSystem(sp, 1, len(BOUNDS), n0=N0, a=A, ome=OME)
    X0 = OX.initial_perturbation(sy, SEED)
    sy.set_fields(OX.initial_fields(sy, A, X0))
    lay, n = sy.unknown_layout()
    t0 = time.perf_counter()
    sy.residual()
    t_res = time.perf_counter() - t0
    # Jacobian columns: a sample from every (field, domain) block, batched like do_col_J
    cols = []
    for (_, _, off, cnt) in lay:
        cols += list(range(off, off + min(16, cnt)))
    t0 = time.perf_counter()
    sy.jacobian(cols=cols, chunk=16)
    t_cols = time.perf_counter() - t0 - t_res  # jacobian() re-evaluates F once
    t_fill = t_cols / len(cols) * n
    # LU (LAPACK dgetrf, the reference's Matrice::set_lu) on a sub-size matrix, scaled as n^3
    ns = 3000
    M = np.random.default_rng(1).standard_normal((ns, ns))
    t0 = time.perf_counter()
    lu = scipy.linalg.lu_factor(M, overwrite_a=True, check_finite=False)
    t_lu_s = time.perf_counter() - t0
    del lu
    t_lu = t_lu_s * (n / ns) ** 3
    step = t_res + t_fill + t_lu
    return {"value": step, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64: residual + {len(cols)} AD Jacobian columns (scaled to n={n}) + "
                      f"LAPACK dgetrf at n={ns} (scaled as n^3); fill {t_fill:.1f} s, LU {t_lu:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed Newton iterations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solver", default="f64", choices=["f64", "f32-ir", "f32-gmres", "bf16-gmres"])
    ap.add_argument("--compare", default="f32-gmres", choices=["none", "f64", "f32-ir", "f32-gmres", "bf16-gmres"],
                    help="solver run to ||F|| < 1e-10 beside the headline one")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--order", type=int, default=1, choices=[0, 1],
                    help="unknown/row order: 0 reference, 1 domain-major (block-tridiagonal J)")
    ap.add_argument("--max-newton", type=int, default=12)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy a 256 MiB device buffer once (reference kernel for HBM counters)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fill-only", action="store_true",
                    help="only the residual and the Jacobian fill once (the run behind the HBM counter passes of "
                         "apply3d): prints the launches and algorithmic bytes of apply3d and exits")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only, "
                                   "--gpus must be 1"}), flush=True)
        sys.exit(2)

    import ctypes
    import numpy as np
    import torch
    from kadath_frontera_amd.device import Context
    from kadath_frontera_amd.xcts import XctsSystem
    from kadath_frontera_amd._lib import lib

    if args.copy_calibration:
        src = torch.empty(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ctx = Context(0)
    sysd = XctsSystem(ctx, RES, BOUNDS, n0=N0, a=A, ome=OME, batch=args.batch, order=args.order)
    n = sysd.n
    X0 = sysd.perturbation(SEED)
    prec_of = {"f64": "f64", "f32-ir": "f32", "f32-gmres": "f32", "bf16-gmres": "bf16"}
    prec = 1e-10

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    if args.fill_only:
        sysd.set_guess(X0)
        lib.kad_kernel_timing(ctx.h, 1, 1)
        sysd.residual()
        sysd.jacobian()
        torch.cuda.synchronize()
        w, kms, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib.kad_kernel_timing_read(ctx.h, 1, ctypes.byref(w), ctypes.byref(kms), ctypes.byref(nl))
        lib.kad_kernel_timing(ctx.h, 1, 0)
        print(json.dumps({"fill_only": True, "n_unknowns": n, "apply3d_launches": nl.value,
                          "apply3d_algorithmic_bytes": w.value, "apply3d_ms": kms.value}), flush=True)
        return

    def newton_iter(solver):
        """One do_newton call; returns (converged_at_entry, |F| at entry, wall s, phases, ir stats)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conv, err, times, ir = sysd.do_newton(prec, solver)
        torch.cuda.synchronize()
        return conv, err, time.perf_counter() - t0, times, ir

    def solve(solver, timed_steps=None):
        """Newton from the perturbed guess until ||F||_inf < 1e-10.  Returns the per-iteration
        records, the time to solution and the converged coefficients."""
        sysd.set_guess(X0)
        recs, t_total, conv = [], 0.0, False
        for it in range(args.max_newton + 1):
            conv, err, dt, times, ir = newton_iter(solver)
            t_total += dt
            log(f"{solver} it {it} |F| {err:.3e} {dt:.3f} s phases {[round(float(x), 4) for x in times]} ir {ir}")
            if conv:
                break
            recs.append((err, dt, list(times), ir))
        fields = np.concatenate([sysd.coef(f, d).ravel() for f in range(5) for d in range(1, len(BOUNDS) + 1)])
        return recs, t_total, conv, err, fields

    # warmup: first iteration(s) from the guess (allocations, kernel loads)
    for _ in range(args.warmup):
        sysd.set_guess(X0)
        newton_iter(args.solver)
    # timed: K Newton iterations along the converging sequence (restarting from the
    # guess when it has converged); every timed call is a full iteration
    sysd.set_guess(X0)
    step_times, step_phases, step_ir = [], [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    while len(step_times) < args.steps:
        conv, err, dt, times, ir = newton_iter(args.solver)
        if conv:
            sysd.set_guess(X0)
            continue
        step_times.append(dt)
        step_phases.append(list(times))
        step_ir.append(ir)
    torch.cuda.synchronize()
    t_region = time.perf_counter() - t_start
    ms_step = 1e3 * float(np.mean(step_times))
    phases = np.mean(step_phases, axis=0)
    t_fact = float(np.mean([p[4] for p in step_phases]))
    lu_tflops = (2.0 / 3.0) * n ** 3 / t_fact / 1e12
    log(f"timed {len(step_times)} iterations: {ms_step:.1f} ms/iteration (region {t_region:.2f} s)")

    # live roofline of the two kernel classes (HIP events on the launch stream) over one
    # iteration from the guess: kind 0 MFMA GEMMs (flops), kind 1 fused 3-D spectral
    # passes of the residual / Jacobian programs (algorithmic HBM bytes)
    sysd.set_guess(X0)
    lib.kad_kernel_timing(ctx.h, 0, 1)
    lib.kad_kernel_timing(ctx.h, 1, 1)
    newton_iter(args.solver)
    meas = {}
    for kind in (0, 1):
        w, kms, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib.kad_kernel_timing_read(ctx.h, kind, ctypes.byref(w), ctypes.byref(kms), ctypes.byref(nl))
        lib.kad_kernel_timing(ctx.h, kind, 0)
        meas[kind] = (w.value, kms.value, nl.value)
    gw, gms, gn = meas[0]
    aw, ams, an = meas[1]
    roof_gemm = {"bound": "mfma", "achieved": gw / (gms / 1e3) / 1e12, "peak": PEAK[prec_of[args.solver]],
                 "unit": "TFLOP/s", "traffic": GEMM_TRAFFIC_B if args.solver == "f64" else None,
                 "traffic_note": GEMM_TRAFFIC_NOTE,
                 "kernel": "gemm_k4_kernel: LU trailing update + blocked-TRSM updates, executed tile pairs "
                           "(%d launches, %.3e flop, %.1f ms per iteration)" % (gn, gw, gms)}
    roof_gemm["frac"] = roof_gemm["achieved"] / roof_gemm["peak"]
    roof_apply = {"bound": "hbm", "achieved": aw / (ams / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                  "traffic": APPLY3D_TRAFFIC_RATIO * aw / max(an, 1), "algorithmic_bytes_per_launch": aw / max(an, 1),
                  "kernel": "apply3d_kernel: fused 3-D spectral passes of the residual and Jacobian programs "
                            "(%d launches, %.3e B, %.1f ms per iteration)" % (an, aw, ams)}
    roof_apply["frac"] = roof_apply["achieved"] / roof_apply["peak"]
    dominant, other = (roof_apply, roof_gemm) if ams > gms else (roof_gemm, roof_apply)

    # time to solution: headline solver and the comparison solver beside it
    runs = {args.solver: solve(args.solver)}
    if args.compare not in ("none", args.solver):
        runs[args.compare] = solve(args.compare)
    conv_info = {k: {"newton_iters": len(v[0]), "time_to_solution_s": v[1], "converged": v[2], "final_Finf": v[3],
                     "ms_per_iteration": 1e3 * float(np.mean([r[1] for r in v[0]])) if v[0] else None,
                     "factor_s_per_iteration": float(np.mean([r[2][4] for r in v[0]])) if v[0] else None,
                     "gmres_iters": [r[3][1] for r in v[0]] if k != "f64" else None}
                 for k, v in runs.items()}
    diff = None
    if len(runs) == 2:
        ref_key = "f64" if "f64" in runs else args.solver
        oth_key = [k for k in runs if k != ref_key][0]
        a, b = runs[oth_key][4], runs[ref_key][4]
        diff = {"solver": oth_key, "vs": ref_key,
                "max_rel": float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))}
    cmp_info = conv_info.get(args.compare) if args.compare in conv_info and args.compare != args.solver else None
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "Newton-step s + LU TFLOP/s: FP32/BF16+IR vs FP64 at ||F||inf<1e-10, 1 GPU",
        "value": ms_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32-ir": "f32-lu+f64-ir", "f32-gmres": "f32-lu+f64-gmres-ir",
                  "bf16-gmres": "bf16-lu+f64-gmres-ir"}[args.solver],
        "data": "synthetic (seeded 1e-3 perturbation of the flat guess)",
        "config": {"workload": "spinning-BH XCTS Newton iteration, 5 spherical domains 17x13x12, n=%d" % n,
                   "n_unknowns": n, "solver": args.solver, "jvp_batch": args.batch,
                   "order": ["reference", "domain-major"][args.order]},
        "lu_tflops_dense_equivalent": lu_tflops,
        "phases_s": {"residual": phases[0], "jacobian_fill": phases[1], "factor_solve": phases[2],
                     "update": phases[3], "factor": phases[4]},
        "compare": cmp_info,
        "convergence": {"solvers": conv_info, "fields_rel_diff": diff},
        "roofline": dominant,
        "roofline_other": other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

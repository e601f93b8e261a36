"""Newton-step benchmark of the MI355X Kadath/FUKA path.

Workload (the spinning-BH configuration on one GPU): XCTS black-hole system on a
Space_spheric with 5 solved domains (4 shells + compactified) at 17x13x12
collocation points, n = 58905 unknowns, dense Jacobian 27.8 GB in FP64; seeded
1e-3 perturbation of the flat guess; horizon rotation Ome = 0.05.

One "step" = one System_of_eqs::do_newton iteration from the perturbed guess:
residual F(u) + batched-JVP Jacobian fill + factorisation + solve (+ fp64
refinement) + field update, all device-resident.  The headline solver is
FP32-LU + FP64 GMRES-IR; the plain FP64-LU step is timed beside it, and both are
then iterated to ||F||_inf < 1e-10 and their converged fields compared.

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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solver", default="f32-gmres", choices=["f64", "f32-ir", "f32-gmres", "bf16-gmres"])
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(see DESIGN.md), --gpus must be 1"}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch
    from kadath_frontera_amd.device import Context
    from kadath_frontera_amd.xcts import XctsSystem
    from kadath_frontera_amd._lib import lib

    ctx = Context(0)
    sysd = XctsSystem(ctx, RES, BOUNDS, n0=N0, a=A, ome=OME, batch=args.batch)
    n = sysd.n
    X0 = sysd.perturbation(SEED)
    prec_of = {"f64": "f64", "f32-ir": "f32", "f32-gmres": "f32", "bf16-gmres": "bf16"}

    def one_step(solver, gemm_timing=False):
        sysd.set_guess(X0)
        torch.cuda.synchronize()
        if gemm_timing:
            lib.kad_gemm_timing(ctx.h, 1)
        t0 = time.perf_counter()
        conv, err, times, ir = sysd.do_newton(1e-10, solver)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(f"[bench] {solver} step {dt:.3f} s phases {[round(x, 4) for x in times]} ir {ir} |F| {err:.3e}",
              file=sys.stderr, flush=True)
        g = None
        if gemm_timing:
            import ctypes
            fl, ms, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
            lib.kad_gemm_timing_read(ctx.h, ctypes.byref(fl), ctypes.byref(ms), ctypes.byref(nl))
            lib.kad_gemm_timing(ctx.h, 0)
            g = (fl.value, ms.value, nl.value)
        return dt, times, ir, g

    for _ in range(args.warmup):
        one_step(args.solver)
    steps = [one_step(args.solver) for _ in range(args.steps)]
    ms_step = 1e3 * float(np.mean([s[0] for s in steps]))
    t_fact = float(np.mean([s[1][4] for s in steps]))
    lu_tflops = (2.0 / 3.0) * n ** 3 / t_fact / 1e12
    phases = np.mean([s[1] for s in steps], axis=0)
    # live GEMM roofline (HIP events on the launch stream) over one step
    _, _, _, g = one_step(args.solver, gemm_timing=True)
    gemm_flops, gemm_ms, gemm_n = g
    achieved = gemm_flops / (gemm_ms / 1e3) / 1e12
    peak = PEAK[prec_of[args.solver]]
    # plain FP64 step beside it
    f64 = None
    if args.solver != "f64":
        one_step("f64")
        d64 = [one_step("f64") for _ in range(max(1, min(2, args.steps)))]
        f64 = {"ms_per_step": 1e3 * float(np.mean([d[0] for d in d64])),
               "lu_tflops": (2.0 / 3.0) * n ** 3 / float(np.mean([d[1][4] for d in d64])) / 1e12}
    # convergence of both solvers to ||F||_inf < 1e-10 and field agreement
    conv_info = None
    if not args.no_converge:
        fields = {}
        iters = {}
        for solver in ([args.solver, "f64"] if args.solver != "f64" else ["f64"]):
            sysd.set_guess(X0)
            it, conv, err = 0, False, None
            while not conv and it < 12:
                conv, err, _, _ = sysd.do_newton(1e-10, solver)
                print(f"[bench] converge {solver} it {it} |F| {err:.3e}", file=sys.stderr, flush=True)
                it += 1
            iters[solver] = (it - 1, err, conv)
            fields[solver] = np.concatenate([sysd.coef(f, d).ravel() for f in range(5)
                                             for d in range(1, len(BOUNDS) + 1)])
        diff = None
        if len(fields) == 2:
            a, b = fields[args.solver], fields["f64"]
            diff = float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))
        conv_info = {"newton_iters": {k: v[0] for k, v in iters.items()},
                     "final_Finf": {k: v[1] for k, v in iters.items()},
                     "converged": all(v[2] for v in iters.values()),
                     "fields_rel_diff_vs_fp64": diff}
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
        "config": {"workload": "spinning-BH XCTS Newton step, 5 spherical domains 17x13x12, n=%d" % n,
                   "n_unknowns": n, "solver": args.solver, "jvp_batch": args.batch},
        "lu_tflops": lu_tflops,
        "phases_s": {"residual": phases[0], "jacobian_fill": phases[1], "factor_solve": phases[2],
                     "update": phases[3], "factor": phases[4]},
        "fp64_step": f64,
        "speedup_vs_fp64": (f64["ms_per_step"] / ms_step) if f64 else None,
        "convergence": conv_info,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "gemm trailing update (%d launches, %.3e flop, %.1f ms)" % (gemm_n, gemm_flops,
                                                                                        gemm_ms)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

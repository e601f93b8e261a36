"""Newton-step benchmark of the MI355X Kadath/FUKA path.

Workload (the spinning-BH configuration on one GPU): XCTS black-hole system on a
Space_spheric with 5 solved domains (4 shells + compactified) at 17x13x12
collocation points, n = 58905 unknowns, dense Jacobian 27.8 GB in FP64.  The horizon
rotation Ome = 0.1421 makes a hole of dimensionless spin chi = J_ADM / M_ADM^2 = 0.600
at the solution (the bench reports the value it reaches).  Guess: the converged hole
at Ome = 0.10 (reached from the flat guess through Ome = 0.05, before any timing)
plus the seeded 1e-3 perturbation of every coefficient.

One "step" = one System_of_eqs::do_newton iteration: residual F(u) + batched-JVP
Jacobian fill + factorisation + solve (+ fp64 refinement) + field update, all
device-resident.  The K timed steps are the iterations of the Newton solve from
the guess (restarting from the guess once ||F||_inf < 1e-10, with no factors kept).
The system is assembled in domain-major order.  The headline solver ("f64-dd")
factors every step's own Jacobian in FP64: the five domain blocks factored together
(one panel chain of 12032 columns instead of 58905), with partial pivoting inside each
or with the pivot orders of the previous step (static pivoting, refreshed after large
multipliers), the interface rows through a Woodbury capacitance matrix, then FP64
flexible GMRES on J to a relative residual of 1e-13; it reaches the same iterates as
the single FP64 LU of the whole J (the reference's pdgesv_, "f64", reported beside
it) to 1e-14.  The other legs run to the same test from the same guess with time to
solution and the converged fields compared: the domain factors of an earlier
Jacobian reused ("f64-dd-reuse"), the FP64 LU of an earlier Jacobian reused
("f64-reuse"), and FP32-LU / BF16-LU + FP64 GMRES-IR.

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
CFG1_RES, CFG1_BOUNDS = (9, 7, 4), [1.0, 2.0, 4.0]  # the CPU-runnable configuration
N0, A, OME = 0.5, 1.0, 0.1421  # chi = J_ADM / M_ADM^2 = 0.600 at the solution
OME_PATH = (0.05, 0.10)         # continuation from the flat guess to the guess's hole
SEED = 20261202
PEAK = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # MI355X dense matrix TFLOP/s (vendor figures)
# HBM bytes per launch, from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE),
# FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0: the factors at which the same
# counters read a 256 MiB device-to-device copy (131093.6 KiB fetched, 262144.0 KiB
# written).
#   apply3d: over the residual and Jacobian fill of this workload (--fill-only, 6566
#     launches, 3.736e11 algorithmic bytes with the element-wise instructions applied in
#     their store phase), traffic / algorithmic bytes = 1.0050; the bench reports that
#     ratio times its live algorithmic bytes per launch.
APPLY3D_TRAFFIC_RATIO = (2.0 * 99981830.5625 + 166654442.78125) * 1024.0 / 373567286352.0
# gemm_k4: the launches of one iteration span many shapes (64- and 128-tiles, the domain
# blocks' updates, the triangular solves' and the Woodbury products), and counter passes
# over the whole LU are not taken (the one tried ended in a segfault of the profiler's host
# counter path, with the persistent panel kernel spinning beside the GEMM), so the line's
# traffic stays null. The kernel is measured on its own at one domain block's trailing
# update instead (--gemm-only: C 12032 x 12032 -= A 12032 x 256 B, same calibration):
# 2 x 981412.0 KiB fetched + 1131008 KiB written = 3.168 GB per launch, 1.339 x the
# algorithmic 2.366 GB (C read and written once, the A and B panels re-fetched about 17
# times over eight L2s), at 23.4 flop/B, above the FP64 ridge: compute-bound.
# traffic_model_bytes is the operand traffic per launch of the executed tiles (each 128 x
# 128 C tile reads its 128 x k A and k x 128 B panels, reads and writes C), an upper bound
# on HBM bytes when nothing is reused through L2.
GEMM_TRAFFIC_NOTE = ("traffic is not measured over the LU's mixed GEMM launches; measured alone at a domain "
                     "block's trailing update (12032 x 12032 x 256, --gemm-only) gemm_k4 moves 3.168 GB per "
                     "launch = 1.339 x its algorithmic bytes (C once, A/B panels re-fetched), compute-bound at "
                     "23.4 flop/B; traffic_model_bytes = operand bytes per launch of the executed tiles with no "
                     "L2 reuse")


def GEMM_TRAFFIC_MODEL(flops, launches):
    # per executed tile pair (128 x 128 x k): 2*128*128*k flop; operands 8*(2*128*k + 2*128*128) bytes;
    # the launches' mean k is 256 in the trailing update (the panel width), so bytes/flop =
    # 8*(2*128*256 + 2*128*128) / (2*128*128*256)
    if launches == 0:
        return None
    per_flop = 8.0 * (2 * 128 * 256 + 2 * 128 * 128) / (2.0 * 128 * 128 * 256)
    return per_flop * flops / launches


def _host_cores():
    """(threads available to this process, physical cores from lscpu or None)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", avail))
    phys = None
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        phys = len(cores) or None
    except Exception:
        pass
    return min(threads, avail), phys


def cpu_baseline():
    """The oracle (numpy/scipy restatement of the reference's FP64 Newton step) on the
    host.  value: one Newton step of this workload extrapolated from a bounded sample
    (one residual, 16 AD columns per (field, domain) block, LAPACK dgetrf at n = 3000
    scaled as n^3).  cfg1_full_step: one complete Newton step (residual, every AD
    column, LU, update) of the 9x7x4 three-domain system, measured, not extrapolated,
    with all host threads and with one."""
    import numpy as np
    import scipy.linalg
    from threadpoolctl import threadpool_limits
    from oracle import xcts as OX
    threads, phys = _host_cores()
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
    # the small configuration in full (do_newton_seq: residual, every column, LU, update)
    full = {}
    for label, nthreads in (("all_threads", threads), ("one_thread", 1)):
        with threadpool_limits(limits=nthreads):
            s1 = OX.XCTSSystem(OX.Space(CFG1_RES, CFG1_BOUNDS), 1, len(CFG1_BOUNDS), n0=N0, a=A, ome=0.0)
            x1 = OX.initial_perturbation(s1, SEED)
            s1.set_fields(OX.initial_fields(s1, A, x1))
            t0 = time.perf_counter()
            OX.newton_step(s1)
            full[label] = time.perf_counter() - t0
    return {"value": step, "unit": "s", "cores": threads, "kind": "port",
            "physical_cores": phys,
            "sample": f"oracle FP64 on {threads} host threads: residual + {len(cols)} AD Jacobian columns "
                      f"(scaled to n={n}) + LAPACK dgetrf at n={ns} (scaled as n^3); fill {t_fill:.1f} s, "
                      f"LU {t_lu:.1f} s",
            "cfg1_full_step_s": {"n": s1.unknown_layout()[1], "threads": threads, **full}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed Newton iterations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--solver", default="f64-dd",
                    choices=["f64", "f32-ir", "f32-gmres", "bf16-gmres", "f64-reuse", "f64-dd", "f64-dd-reuse"])
    ap.add_argument("--compare", default="f64,f64-dd-reuse,f64-reuse,f32-gmres,bf16-gmres",
                    help="comma-separated solvers run to ||F|| < 1e-10 beside the headline one (or 'none')")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--fill-streams", type=int, default=3,
                    help="streams of the Jacobian fill (the tangent programs run concurrently, 1-4)")
    ap.add_argument("--order", type=int, default=1, choices=[0, 1],
                    help="unknown/row order: 0 reference, 1 domain-major (block-tridiagonal J)")
    ap.add_argument("--max-newton", type=int, default=12)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy a 256 MiB device buffer once (reference kernel for HBM counters)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fill-only", action="store_true",
                    help="only the residual and the Jacobian fill once (the run behind the HBM counter passes of "
                         "apply3d): prints the launches and algorithmic bytes of apply3d and exits")
    ap.add_argument("--gemm-only", action="store_true",
                    help="only four FP64 GEMM launches (one untimed) of a domain block's trailing-update shape, C (12032 x 12032) "
                         "-= A (12032 x 256) B (256 x 12032) (the run behind the HBM counter passes of gemm_k4): "
                         "prints their algorithmic bytes and exits")
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
    if args.gemm_only:
        mg, kg = 12032, 256
        Ag = torch.randn(kg, mg, dtype=torch.float64, device="cuda").T  # column-major views
        Bg = torch.randn(mg, kg, dtype=torch.float64, device="cuda").T
        Cg = torch.randn(mg, mg, dtype=torch.float64, device="cuda").T
        torch.cuda.synchronize()
        ctx.gemm("f64", 0, Ag, Bg, Cg)  # untimed: loads the code object
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.gemm("f64", 0, Ag, Bg, Cg)
        ctx.sync()
        dt = (time.perf_counter() - t0) / 3
        print(json.dumps({"gemm_only": True, "m": mg, "n": mg, "k": kg, "launches": 4, "timed": 3,
                          "ms_per_launch": 1e3 * dt, "tflops": 2.0 * mg * mg * kg / dt / 1e12,
                          "algorithmic_bytes_per_launch": 8.0 * (2 * mg * mg + 2 * mg * kg)}), flush=True)
        return
    sysd = XctsSystem(ctx, RES, BOUNDS, n0=N0, a=A, ome=OME, batch=args.batch, order=args.order)
    sysd.set_option(fill_streams=args.fill_streams)
    n = sysd.n
    X0 = sysd.perturbation(SEED)
    prec_of = {"f64": "f64", "f32-ir": "f32", "f32-gmres": "f32", "bf16-gmres": "bf16", "f64-reuse": "f64",
               "f64-dd": "f64", "f64-dd-reuse": "f64"}
    prec = 1e-10

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    if args.fill_only:
        # the residual and the Jacobian once, at the flat guess plus the perturbation: the
        # launches and their sizes do not depend on the field values, so no continuation
        sysd.set_guess(X0)
        sysd.set_option(fill_streams=1)  # each launch timed alone
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

    # the guess's hole: Newton continuation in Ome from the flat guess (untimed), with
    # the headline's fp64 direct solver when it is one (a profile then sees its kernels only)
    cont_solver = args.solver if args.solver in ("f64", "f64-dd") else "f64"
    sysd.set_guess(None)
    for om in OME_PATH:
        sysd.set_ome(om)
        for it in range(30):
            conv, err, _, _ = sysd.do_newton(prec, cont_solver)
            if conv:
                break
        if not conv:
            raise RuntimeError(f"continuation at Ome = {om} did not converge (|F| = {err:.3e})")
        log(f"guess hole at Ome = {om}: {it} iterations")
    guess = {(f, d): sysd.coef(f, d).copy() for f in range(5) for d in range(1, len(BOUNDS) + 1)}
    sysd.set_ome(OME)
    X0c = np.ascontiguousarray(X0, dtype=np.float64)

    def set_guess():
        for (f, d), cf in guess.items():
            sysd.set_coef(f, d, cf)
        rc = lib.kad_system_add_unknowns(sysd.h, X0c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise RuntimeError(f"kad_system_add_unknowns failed with status {rc}")

    def newton_iter(solver):
        """One do_newton call; returns (converged_at_entry, |F| at entry, wall s, phases, ir stats)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conv, err, times, ir = sysd.do_newton(prec, solver)
        torch.cuda.synchronize()
        return conv, err, time.perf_counter() - t0, times, ir

    def fields():
        return np.concatenate([sysd.coef(f, d).ravel() for f in range(5) for d in range(1, len(BOUNDS) + 1)])

    def solve(solver):
        """Newton from the perturbed guess until ||F||_inf < 1e-10 (the time to solution),
        then the fields and global quantities there, and again after one more step of the
        same solver ("polished": ||F|| at the floor of fp64 residual evaluation)."""
        set_guess()
        sysd.set_option(drop_factors=True)  # every leg starts without kept factors
        recs, t_total, conv = [], 0.0, False
        for it in range(args.max_newton + 1):
            conv, err, dt, times, ir = newton_iter(solver)
            t_total += dt
            log(f"{solver} it {it} |F| {err:.3e} {dt:.3f} s phases {[round(float(x), 4) for x in times]} ir {ir}")
            if conv:
                break
            recs.append((err, dt, list(times), ir))
        f_stop, q_stop = fields(), sysd.observables()
        dd = sysd.dd_info() if solver.startswith("f64-dd") else None
        _, err_p, _, _, _ = newton_iter_prec(solver, 0.0)
        f_pol, q_pol = fields(), sysd.observables()
        return dict(recs=recs, t=t_total, conv=conv, err=err, f_stop=f_stop, q_stop=q_stop, f_pol=f_pol,
                    q_pol=q_pol, dd=dd)

    def newton_iter_prec(solver, p):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conv, err, times, ir = sysd.do_newton(p, solver)
        torch.cuda.synchronize()
        return conv, err, time.perf_counter() - t0, times, ir

    # warmup: first iteration(s) from the guess (allocations, kernel loads)
    for _ in range(args.warmup):
        set_guess()
        sysd.set_option(drop_factors=True)
        newton_iter(args.solver)
    # timed: K Newton iterations along the converging sequence (restarting from the
    # guess when it has converged); every timed call is a full iteration
    set_guess()
    sysd.set_option(drop_factors=True)
    step_times, step_phases, step_ir = [], [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    while len(step_times) < args.steps:
        conv, err, dt, times, ir = newton_iter(args.solver)
        if conv:
            set_guess()
            sysd.set_option(drop_factors=True)  # a restart is a new solve: no factors kept
            continue
        step_times.append(dt)
        step_phases.append(list(times))
        step_ir.append(ir)
    torch.cuda.synchronize()
    t_region = time.perf_counter() - t_start
    ms_step = 1e3 * float(np.mean(step_times))
    phases = np.mean(step_phases, axis=0)
    t_fact = float(np.mean([p[4] for p in step_phases]))
    log(f"timed {len(step_times)} iterations: {ms_step:.1f} ms/iteration (region {t_region:.2f} s)")

    # live roofline of the two kernel classes (HIP events on the launch stream) over one
    # iteration from the guess: kind 0 MFMA GEMMs (flops), kind 1 fused 3-D spectral
    # passes of the residual / Jacobian programs (algorithmic HBM bytes).  The fill runs
    # on one stream here, so that each spectral pass is timed alone (on several streams
    # the launches overlap and each event pair spans its neighbours' work too).
    set_guess()
    sysd.set_option(drop_factors=True, fill_streams=1)
    lib.kad_kernel_timing(ctx.h, 0, 1)
    lib.kad_kernel_timing(ctx.h, 1, 1)
    newton_iter(args.solver)
    sysd.set_option(fill_streams=args.fill_streams)
    meas = {}
    for kind in (0, 1):
        w, kms, nl = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib.kad_kernel_timing_read(ctx.h, kind, ctypes.byref(w), ctypes.byref(kms), ctypes.byref(nl))
        lib.kad_kernel_timing(ctx.h, kind, 0)
        meas[kind] = (w.value, kms.value, nl.value)
    gw, gms, gn = meas[0]
    aw, ams, an = meas[1]
    # LU TFLOP/s: the flops its GEMMs executed (structurally zero tile pairs are skipped
    # and not counted) over the factorisation's wall time
    lu_tflops = gw / t_fact / 1e12 if t_fact > 0 else None
    roof_gemm = {"bound": "mfma", "achieved": gw / (gms / 1e3) / 1e12, "peak": PEAK[prec_of[args.solver]],
                 "unit": "TFLOP/s", "traffic": None, "traffic_model_bytes": GEMM_TRAFFIC_MODEL(gw, gn),
                 "traffic_note": GEMM_TRAFFIC_NOTE,
                 "work_per_launch_flop": gw / max(gn, 1),
                 "kernel": ("gemm_k4_kernel: batched trailing updates of the domain-block LU, blocked-TRSM "
                            "updates and the Woodbury products" if args.solver.startswith("f64-dd") else
                            "gemm_k4_kernel: LU trailing update + blocked-TRSM updates, executed tile pairs") +
                           " (%d launches, %.3e flop, %.1f ms per iteration)" % (gn, gw, gms)}
    roof_gemm["frac"] = roof_gemm["achieved"] / roof_gemm["peak"]
    roof_apply = {"bound": "hbm", "achieved": aw / (ams / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                  "traffic": APPLY3D_TRAFFIC_RATIO * aw / max(an, 1), "algorithmic_bytes_per_launch": aw / max(an, 1),
                  "kernel": "apply3d_kernel: fused 3-D spectral passes of the residual and Jacobian programs "
                            "(%d launches, %.3e B, %.1f ms per iteration, timed with the fill on one stream)"
                            % (an, aw, ams)}
    roof_apply["frac"] = roof_apply["achieved"] / roof_apply["peak"]
    dominant, other = (roof_apply, roof_gemm) if ams > gms else (roof_gemm, roof_apply)

    # time to solution: headline solver and the comparison solvers beside it
    legs = [args.solver] + [c for c in args.compare.split(",") if c and c != "none" and c != args.solver]
    runs = {k: solve(k) for k in legs}
    ref_key = "f64" if "f64" in runs else args.solver
    ref = runs[ref_key]
    sc = max(1.0, float(np.abs(ref["f_stop"]).max()))
    conv_info = {}
    for k, v in runs.items():
        recs = v["recs"]
        info = {"newton_iters": len(recs), "time_to_solution_s": v["t"], "converged": v["conv"], "final_Finf": v["err"],
                "ms_per_iteration": 1e3 * float(np.mean([r[1] for r in recs])) if recs else None,
                "factor_s_per_iteration": float(np.mean([r[2][4] for r in recs])) if recs else None}
        if k != "f64":
            info["gmres_inner_steps"] = [r[3][1] for r in recs]
            info["refinement_rel_resid"] = [r[3][2] for r in recs]
            info["fp64_fallback_steps"] = int(sum(r[3][3] for r in recs))
            info["factorizations"] = int(sum(r[3][4] for r in recs))
        if k.startswith("f64-dd"):
            info["domain_blocks"] = v["dd"]
        if k != ref_key:
            info["vs_" + ref_key] = {
                "fields_max_rel_at_stop": float(np.abs(v["f_stop"] - ref["f_stop"]).max() / sc),
                "fields_max_rel_polished": float(np.abs(v["f_pol"] - ref["f_pol"]).max() / sc),
                "M_adm_abs_at_stop": abs(v["q_stop"]["M_adm"] - ref["q_stop"]["M_adm"]),
                "J_adm_abs_at_stop": abs(v["q_stop"]["J_adm"] - ref["q_stop"]["J_adm"]),
                "M_adm_abs_polished": abs(v["q_pol"]["M_adm"] - ref["q_pol"]["M_adm"]),
                "J_adm_abs_polished": abs(v["q_pol"]["J_adm"] - ref["q_pol"]["J_adm"])}
        conv_info[k] = info
    conv_info[ref_key]["observables_at_stop"] = ref["q_stop"]
    chi = ref["q_stop"]["J_adm"] / ref["q_stop"]["M_adm"] ** 2
    # the CPU-runnable configuration on the GPU (one full Newton step), beside the CPU's
    s1 = XctsSystem(ctx, CFG1_RES, CFG1_BOUNDS, n0=N0, a=A, ome=0.0, batch=args.batch, order=args.order)
    s1.set_guess(s1.perturbation(SEED))
    s1.do_newton(prec, "f64")  # warm-up (allocations)
    s1.set_guess(s1.perturbation(SEED))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s1.do_newton(prec, "f64")
    torch.cuda.synchronize()
    gpu_cfg1 = time.perf_counter() - t0
    n_cfg1 = s1.n
    s1.close()
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
                  "bf16-gmres": "bf16-lu+f64-gmres-ir", "f64-reuse": "f64", "f64-dd": "f64",
                  "f64-dd-reuse": "f64"}[args.solver],
        "data": "synthetic (the converged Ome = 0.10 hole plus a seeded 1e-3 perturbation)",
        "config": {"workload": "spinning-BH XCTS Newton iteration, 5 spherical domains 17x13x12, n=%d" % n,
                   "n_unknowns": n, "solver": args.solver, "jvp_batch": args.batch, "fill_streams": args.fill_streams,
                   "order": ["reference", "domain-major"][args.order],
                   "horizon_rotation": OME, "chi_at_solution": chi},
        "lu_tflops": lu_tflops,
        "lu_tflops_note": "GEMM flops executed in one FP64 factorisation (structurally zero tile pairs skipped, "
                          "not counted; for f64-dd the domain blocks, D^-1 P and the capacitance matrix with its LU) "
                          "/ the factorisation's wall time",
        "phases_s": {"residual": phases[0], "jacobian_fill": phases[1], "factor_solve": phases[2],
                     "update": phases[3], "factor": phases[4]},
        "convergence": conv_info,
        "gpu_cfg1_newton_step_s": {"config": "9x7x4, 3 domains, n=%d, f64" % n_cfg1, "value": gpu_cfg1},
        "roofline": dominant,
        "roofline_other": other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

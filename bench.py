#!/usr/bin/env python3
"""Benchmark of the adaptive embedded RK N-body path on one MI355X.

Workload (the plain FP64 GPU configuration): N = 16384 Plummer sphere (seed 0,
scale radius 1, total mass 1, G = 1), softening 1e-2, DOP853, rtol = atol =
1e-10, 200 accepted steps (step-count-limited solve, tf = 100 unreachable).
A bench "step" is one complete solve with the initial state already resident
in HBM; value = N * accepted / seconds (particle-steps/s).

Also reported on the same JSON line: RHS evaluations per solve; the
roofline of the dominant kernel (the N-body RHS, FP64 VALU-bound: 20 FLOP per
directed pair interaction, N(N-1) per launch, timed with HIP events on the
library's stream); the bit-exact and mixed-precision RHS variants; and the
reference's own thread-pool CPU path (oracle/_ref, built from the reference
headers) timed on this host on a bounded sample (2 accepted steps at full N).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
FLOP_PER_INTERACTION = 20


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rhs_kernel_ms(exe, oh, xs, ks, ode, reps):
    for _ in range(2):
        ode(exe, xs, ks)
    exe.synchronize()
    exe.timer_start()
    for _ in range(reps):
        ode(exe, xs, ks)
    return exe.timer_stop() / reps


def cpu_baseline(x0, m, softening, tol, n, accepted):
    """The reference's ThreadPoolExecutor path on this host (oracle/_ref)."""
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if n == 16384 and os.path.exists(ref):
        L = C.CDLL(ref)
        dp = C.POINTER(C.c_double)
        L.ref_tp_solve_16384_dop853.argtypes = [C.c_int, dp, C.c_double, C.c_double,
                                                C.c_double, C.c_long, C.c_double, dp]
        L.ref_tp_solve_16384_dop853.restype = C.c_double
        xo = np.zeros(6 * n)
        secs = L.ref_tp_solve_16384_dop853(threads, x0.ctypes.data_as(dp), tol, softening,
                                           float(m[0]), accepted, 100.0, xo.ctypes.data_as(dp))
        kind = "reference"
        what = (f"reference RKEmbeddedParallel<DOP853> + ThreadPoolExecutor({threads}) + "
                f"NBodyODE (atomic pair RHS), N={n}, {accepted} accepted steps incl. the "
                "initial-step estimate, uniform masses applied by scaling the unit-mass RHS")
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        t = time.perf_counter()
        O.integrate("DOP853", O.RHS_NBODY, x0, 0.0, 100.0, tol, tol, softening=softening,
                    masses=m, max_accepted=accepted, store_states=False, threads=threads)
        secs = time.perf_counter() - t
        kind = "port"
        what = (f"oracle restatement (gather RHS, {threads} threads), N={n}, "
                f"{accepted} accepted steps incl. the initial-step estimate")
    return {"value": n * accepted / secs, "unit": "particle-steps/s", "cores": threads,
            "kind": kind, "sample": what, "seconds": secs}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves")
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--accepted", type=int, default=200)
    ap.add_argument("--variant", default="exact", choices=["fast", "exact", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--profile-rhs", action="store_true",
                    help="only run each RHS variant 3x and a 64 MiB calibration copy "
                         "(the workload of the committed counter profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): this build "
              "measures the single-GPU path; multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)

    import ode_prototyping_amd as oh

    if args.profile_rhs:
        x0, m = oh.scenarios.plummer(args.n, seed=0)
        exe = oh.HipExecutor(0)
        xs, ks = oh.HipState(exe, data=x0), oh.HipState(exe, 6 * args.n)
        for var in ("exact", "fast", "mixed"):
            o = oh.NBodyODE(1e-2, m, var)
            for _ in range(3):
                o(exe, xs, ks)
        a, b = oh.HipState(exe, 8 << 20), oh.HipState(exe, 8 << 20)
        for _ in range(3):
            b.assign(a)
        exe.synchronize()
        log("profile-rhs done")
        return

    n, steps, warmup = args.n, args.steps, args.warmup
    softening, tol, tableau = 1e-2, 1e-10, "DOP853"
    x0, m = oh.scenarios.plummer(n, seed=0)
    exe = oh.HipExecutor(0)
    x0_dev = oh.HipState(exe, data=x0)
    work = oh.HipState(exe, 6 * n)
    atol = oh.HipState(exe, data=np.full(6 * n, tol))
    rtol = oh.HipState(exe, data=np.full(6 * n, tol))
    ode = oh.NBodyODE(softening, m, args.variant)
    ode.masses_device(exe)

    def solve(o, **kw):
        return oh.integrate(tableau, x0_dev, 0.0, 100.0, atol, rtol, o, exe=exe,
                            max_accepted=args.accepted, x_out=work, fetch=False, **kw)

    log(f"bench: N={n} {tableau} variant={args.variant} warmup={warmup} steps={steps}")
    for w in range(warmup):
        r = solve(ode)
        log(f"warmup {w}: {r}")
    exe.synchronize()
    exe.timer_start()
    t0 = time.perf_counter()
    results = []
    for s in range(steps):
        results.append(solve(ode))
    exe.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = exe.timer_stop()
    r = results[-1]
    log(f"timed: {steps} solves in {wall:.3f} s (events {ev_ms:.1f} ms): {r}")
    accepted = r.accepted
    ms_per_step = 1000.0 * wall / steps
    value = n * accepted * steps / wall

    # roofline of the dominant kernel (the RHS), HIP events on the library stream
    xs = work
    ks = oh.HipState(exe, 6 * n)
    rhs_ms = rhs_kernel_ms(exe, oh, xs, ks, ode, 20)
    interactions = n * (n - 1)
    achieved = FLOP_PER_INTERACTION * interactions / (rhs_ms * 1e-3) / 1e12
    rhs_total_ms = r.rhs_evals * rhs_ms
    log(f"rhs kernel ({args.variant}): {rhs_ms:.4f} ms/launch, {achieved:.2f} TFLOP/s; "
        f"RHS share of solve {rhs_total_ms / ms_per_step:.3f}")

    # sustained FP64 peak of this device measured live (FMA probe)
    sus_tflops, sus_mhz = exe.probe_fp64()
    log(f"sustained FP64 FMA: {sus_tflops:.1f} TFLOP/s at {sus_mhz:.0f} MHz")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"nbody_rhs_n{n}_traffic.json")
    if os.path.exists(tfile):
        tk = json.load(open(tfile))["kernels"].get(args.variant)
        traffic = tk["hbm_bytes_per_launch"] if tk else None

    out = {
        "metric": "particle-steps/s at rtol 1e-10 + RHS evals/solve: chosen vs plain FP64, 1 GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 N-body solve: {n}-body Plummer sphere (seed 0, "
                               f"a=1, M=1), eps={softening}, {tableau}, rtol=atol={tol}, "
                               f"{args.accepted} accepted steps",
                   "n_particles": n, "tableau": tableau, "rtol": tol, "atol": tol,
                   "softening": softening, "accepted_steps": accepted,
                   "rhs_variant": args.variant, "parallelism": "single GPU"},
        "rhs_evals_per_solve": r.rhs_evals,
        "accepted": r.accepted,
        "rejected": r.rejected,
        "ms_per_step_events": ev_ms / steps,
        "roofline": {"bound": "valu", "pipe": "FP64 vector ALU (no dense contraction)",
                     "kernel": f"k_nbody_{args.variant}", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel_ms": rhs_ms,
                     "sustained_fp64_tflops": sus_tflops, "clock_mhz": sus_mhz,
                     "frac_of_sustained": achieved / sus_tflops,
                     "algorithmic": f"{FLOP_PER_INTERACTION} FLOP x N(N-1) = "
                                    f"{FLOP_PER_INTERACTION * interactions:.4g} FLOP/launch"},
    }

    if not args.no_extras:
        extras = {}
        for var in ("exact", "fast", "mixed"):
            if var == args.variant:
                continue
            o2 = oh.NBodyODE(softening, m, var)
            t = time.perf_counter()
            r2 = solve(o2)
            exe.synchronize()
            secs = time.perf_counter() - t
            kms = rhs_kernel_ms(exe, oh, xs, ks, o2, 5)
            extras[var] = {"particle_steps_per_s": n * r2.accepted / secs,
                           "ms_per_solve": 1000 * secs, "rhs_evals_per_solve": r2.rhs_evals,
                           "rhs_kernel_ms": kms,
                           "rhs_tflops": FLOP_PER_INTERACTION * interactions / (kms * 1e-3) / 1e12}
            log(f"{var}: {extras[var]}")
        # chosen path on this workload: mixed-precision RHS + k1 reuse after rejections
        o3 = oh.NBodyODE(softening, m, "mixed")
        t = time.perf_counter()
        r3 = solve(o3, fsal=True)
        exe.synchronize()
        secs = time.perf_counter() - t
        final_plain = oh.integrate(tableau, x0_dev, 0.0, 100.0, atol, rtol, ode, exe=exe,
                                   max_accepted=args.accepted).x_final
        final_mixed = oh.integrate(tableau, x0_dev, 0.0, 100.0, atol, rtol, o3, exe=exe,
                                   max_accepted=args.accepted, fsal=True).x_final
        rel = float(np.max(np.abs(final_mixed - final_plain)) / np.max(np.abs(final_plain)))
        extras["chosen_mixed_fsal"] = {"particle_steps_per_s": n * r3.accepted / secs,
                                       "rhs_evals_per_solve": r3.rhs_evals,
                                       "final_state_rel_err_vs_plain": rel}
        log(f"chosen: {extras['chosen_mixed_fsal']}")
        out["variants"] = extras

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(x0, m, softening, tol, n, 2)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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

Further blocks (each a separate workload, not the headline value):
* "ensemble": 2048 ragged systems (N log-uniform in [16, 4096], uniform-cube
  positions, N(0, 0.1^2) velocities, masses 1/N, eps 1e-3), DOPRI5, rtol =
  atol = 1e-10, t in [0, 0.1] (--ensemble-tf for the configured [0, 2]), every
  system with its own step size: plain FP64 (exact RHS, reference controller)
  against the chosen path (mixed-precision RHS with the close-encounter near
  field, learned controller, k1 reuse) and the variants between them, with the
  targets (particle-steps/s, RHS evals/solve, final state, energy drift) stated;
* "controller": 4096 held-out systems of the same distribution, DOP853,
  t in [0, 0.05], mixed-precision RHS: reference vs PI vs learned controller;
* "precision": N = 65536 colliding Plummer spheres, eps 1e-2, DOP853, 10
  accepted steps: mixed-precision against the FP64 path (final-state and
  energy error).
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
# FP64 VALU wave-instructions per 64 directed interactions of the FP64 RHS
# kernels (SQ_INSTS_VALU x 64 / N^2 at N=16384 with masses 1/N = 2^-14, which both kernels
# take as unit-mass terms and one scale per sum;
# profiles/nbody_rhs_n16384_valu_counters.txt) and how many of them are quarter-rate
# transcendentals (v_rsq_f64, v_rcp_f64).
VALU_PER_64_INTERACTIONS = {"exact": (38.38, 2), "fast": (17.18, 1)}
SIMDS = 1024  # 256 CUs x 4 SIMDs


def instruction_bound(variant, n, clock_mhz, kernel_ms):
    """Issue-time bound of an FP64 RHS launch: every 64-lane FP64 instruction
    occupies its SIMD for 4 cycles, a quarter-rate one for 16, at the clock the
    live FMA probe measured."""
    if variant not in VALU_PER_64_INTERACTIONS:
        return None
    total, quarter = VALU_PER_64_INTERACTIONS[variant]
    cycles = (total - quarter) * 4 + quarter * 16
    bound_ms = n * n / 64 / SIMDS * cycles / (clock_mhz * 1e3)
    return {"valu_per_64_interactions": total, "quarter_rate": quarter,
            "cycles_per_64_interactions": cycles, "bound_ms": bound_ms,
            "frac": bound_ms / kernel_ms,
            "source": "SQ_INSTS_VALU per interaction from the committed counter profile; "
                      "4 cycles per FP64 wave64 instruction, 16 for v_rsq_f64 / v_rcp_f64"}


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


def nbody_energy(oh, exe, x_dev_or_host, n, eps, m):
    """(KE + PE) of one system on the device."""
    x = x_dev_or_host if isinstance(x_dev_or_host, oh.HipState) else oh.HipState(
        exe, data=x_dev_or_host)
    md = oh.HipState(exe, data=m)
    out = np.zeros(2)
    exe.check(oh.lib().odehip_nbody_energy(exe.handle, x.ptr, n, eps, md.ptr,
                                           out.ctypes.data_as(C.POINTER(C.c_double))),
              "nbody_energy")
    return out.sum()


def ensemble_block(oh, exe, tf=0.1):
    """Batched ensemble of 2048 ragged systems, DOPRI5: the plain FP64 path against
    the chosen path and the variants between them."""
    from ode_prototyping_amd.ensemble import solve_ensemble, unpack
    from ode_prototyping_amd import train_controller as trc
    eps, tol, tab = 1e-3, 1e-10, "DOPRI5"
    sizes, x0, m = trc.workload_ensemble(2048, 0)
    x0d, md = oh.HipState(exe, data=x0), oh.HipState(exe, data=m)
    out_state = oh.HipState(exe, x0.size)
    net, info = trc.pretrained(exe, tab)
    inter = float(np.sum(sizes.astype(float) * (sizes - 1)))
    close = oh.MIXED_NEAR_CLOSE
    # name: (RHS variant, controller, network, near-field coefficient, k1 reuse)
    runs = {"plain_exact": ("exact", "reference", None, 0.0, False),
            "fast": ("fast", "reference", None, 0.0, False),
            "fast_fsal": ("fast", "reference", None, 0.0, True),
            "mixed": ("mixed", "reference", None, 0.0, False),
            "mixed_near": ("mixed", "reference", None, close, False),
            "chosen": ("mixed", "learned", net, close, True)}
    block = {"workload": f"{len(sizes)} systems, N log-uniform [16, 4096] (sum N "
                         f"{int(sizes.sum())}), uniform cube, masses 1/N, eps={eps}, {tab}, "
                         f"rtol=atol={tol}, t in [0, {tf}], per-system step size",
             "chosen_path": "mixed-precision RHS (FP32 pair terms, FP64 sums) with the "
                            f"close-encounter near field (pairs with r^2 < {close}/N in FP64), "
                            "learned step-size controller (PI fallback), first-same-as-last "
                            "k1 reuse",
             "interactions_per_rhs": inter, "learned_net": info}
    finals = {}
    for name, (var, ctrl, mlp, near, fsal) in runs.items():
        exe.near_field = near
        kw = dict(variant=var, exe=exe, controller=ctrl, mlp=mlp, fsal=fsal, x_out=out_state,
                  fetch=False)
        # untimed: the tile autotuner runs on the first solve of a variant
        solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, max_attempts=1, **kw)
        exe.synchronize()
        t = time.perf_counter()
        r = solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, **kw)
        exe.synchronize()
        secs = time.perf_counter() - t
        finals[name] = out_state.copy_to_span()
        peak = 2 * FP64_PEAK_TFLOPS if var == "mixed" else FP64_PEAK_TFLOPS
        ach = FLOP_PER_INTERACTION * inter / (r.rhs_ms * 1e-3) / 1e12 if r.rhs_ms else None
        block[name] = {"variant": var, "controller": ctrl, "near_field": near, "fsal": fsal,
                       "seconds": secs, "particle_steps_per_s": r.particle_steps / secs,
                       "rhs_evals_per_solve": r.total_rhs_evals / len(sizes),
                       "accepted": r.total_accepted, "rejected": r.total_rejected,
                       "rounds": r.rounds, "learned_steps": r.total_learned_steps,
                       "tile_config": r.tile_config, "rhs_ms": r.rhs_ms,
                       "rhs_tflops": ach, "rhs_peak_tflops": peak,
                       "rhs_frac": (ach / peak) if ach else None}
        log(f"ensemble {name}: {block[name]}")
    exe.near_field = oh.MIXED_NEAR_DEFAULT
    # every system ends at t = tf, so final states compare at equal t; the plain path
    # equals the FP64 oracle bit for bit (tests/test_gpu_configs.py)
    ref = unpack(finals["plain_exact"], sizes)
    ms = unpack(m, sizes, per_particle=1)
    x0s = unpack(x0, sizes)
    idx = np.linspace(0, len(sizes) - 1, 64).astype(int)  # energy on 64 systems
    e0 = {int(k): nbody_energy(oh, exe, x0s[k], int(sizes[k]), eps, ms[k]) for k in idx}
    for name in runs:
        got = unpack(finals[name], sizes)
        if name != "plain_exact":
            err = np.array([np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(got, ref)])
            block[name]["final_state_rel_err_vs_plain"] = {
                "median": float(np.median(err)), "p90": float(np.quantile(err, 0.9)),
                "max": float(np.max(err)), "frac_le_1e-9": float(np.mean(err <= 1e-9))}
        drift = [abs((nbody_energy(oh, exe, got[k], int(sizes[k]), eps, ms[k]) - e0[int(k)]) /
                     e0[int(k)]) for k in idx]
        block[name]["energy_rel_drift_64_systems"] = {"median": float(np.median(drift)),
                                                      "max": float(np.max(drift))}
    p = block["plain_exact"]

    def targets(c):
        fe = c["final_state_rel_err_vs_plain"]
        return {
            "particle_steps_per_s_vs_plain": c["particle_steps_per_s"] / p["particle_steps_per_s"],
            "rhs_evals_per_solve_vs_plain": c["rhs_evals_per_solve"] / p["rhs_evals_per_solve"],
            "final_state_max_rel_err": fe["max"], "final_state_median_rel_err": fe["median"],
            "energy_drift_median_chosen": c["energy_rel_drift_64_systems"]["median"],
            "energy_drift_median_plain": p["energy_rel_drift_64_systems"]["median"],
            "beats_plain_particle_steps": c["particle_steps_per_s"] > p["particle_steps_per_s"],
            "beats_plain_rhs_evals": c["rhs_evals_per_solve"] < p["rhs_evals_per_solve"],
            "final_state_le_1e-9": fe["max"] <= 1e-9,
            "energy_drift_no_worse": (c["energy_rel_drift_64_systems"]["median"] <=
                                      p["energy_rel_drift_64_systems"]["median"])}

    # the chosen path (mixed RHS + near field + learned controller + k1 reuse) and the fast
    # FP64 + k1 reuse path, against the same four targets
    block["targets"] = targets(block["chosen"])
    block["targets_fast_fsal"] = targets(block["fast_fsal"])
    log(f"ensemble targets: {block['targets']}; fast + k1 reuse: {block['targets_fast_fsal']}")
    # DOPRI5 dense output of every system at 20 common times (fast FP64 RHS)
    times = np.linspace(0.0, tf, 21)[1:]
    exe.synchronize()
    t = time.perf_counter()
    r = solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, variant="fast", exe=exe,
                       x_out=out_state, fetch=False, dense_times=times)
    exe.synchronize()
    secs = time.perf_counter() - t
    block["fast"]["dense_output"] = {"seconds": secs,
                                     "overhead": secs / block["fast"]["seconds"] - 1.0,
                                     "rows_per_system": int(np.min(r.n_dense)),
                                     "bytes_to_host": int(r.dense_states.nbytes)}
    log(f"ensemble fast + dense output: {block['fast']['dense_output']}")
    del r
    net.close()
    block["cpu_baseline"] = ensemble_cpu_baseline(sizes, x0, m, eps, tol, tab)
    log(f"ensemble cpu baseline: {block['cpu_baseline']}")
    return block


def near_sweep_block(oh, exe, tf=0.1, coefficients=(20.0, 100.0, 500.0, 2000.0)):
    """C4 to t = tf with the mixed RHS (reference controller, k1 reuse) at several
    near-field coefficients c (pairs with r^2 < c / N in FP64), against the plain FP64 path
    and fast FP64 + k1 reuse: time, RHS evaluations, final state vs plain, energy drift."""
    from ode_prototyping_amd.ensemble import solve_ensemble, unpack
    from ode_prototyping_amd import train_controller as trc
    eps, tol, tab = 1e-3, 1e-10, "DOPRI5"
    sizes, x0, m = trc.workload_ensemble(2048, 0)
    x0d, md = oh.HipState(exe, data=x0), oh.HipState(exe, data=m)
    out_state = oh.HipState(exe, x0.size)
    ms, x0s = unpack(m, sizes, per_particle=1), unpack(x0, sizes)
    idx = np.linspace(0, len(sizes) - 1, 64).astype(int)
    e0 = {int(k): nbody_energy(oh, exe, x0s[k], int(sizes[k]), eps, ms[k]) for k in idx}
    runs = [("plain_exact", "exact", 0.0, False), ("fast_fsal", "fast", 0.0, True)]
    runs += [(f"mixed_near_{c:g}", "mixed", c, True) for c in coefficients]
    block, ref = {"workload": f"2048-system ensemble, {tab}, t in [0, {tf}], reference "
                              "controller"}, None
    for name, var, near, fsal in runs:
        exe.near_field = near
        kw = dict(variant=var, exe=exe, fsal=fsal, x_out=out_state, fetch=False)
        solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, max_attempts=1, **kw)
        exe.synchronize()
        t = time.perf_counter()
        r = solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, **kw)
        exe.synchronize()
        secs = time.perf_counter() - t
        got = unpack(out_state.copy_to_span(), sizes)
        row = {"near_field": near, "seconds": secs,
               "rhs_evals_per_solve": r.total_rhs_evals / len(sizes), "rhs_ms": r.rhs_ms}
        if ref is None:
            ref = got
        else:
            err = np.array([np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(got, ref)])
            row["final_state_rel_err_vs_plain"] = {"median": float(np.median(err)),
                                                   "max": float(np.max(err)),
                                                   "frac_le_1e-9": float(np.mean(err <= 1e-9))}
        drift = [abs((nbody_energy(oh, exe, got[k], int(sizes[k]), eps, ms[k]) - e0[int(k)]) /
                     e0[int(k)]) for k in idx]
        row["energy_rel_drift_median"] = float(np.median(drift))
        block[name] = row
        log(f"near sweep {name}: {row}")
    exe.near_field = oh.MIXED_NEAR_DEFAULT
    return block


def ensemble_horizon_block(oh, exe, tf=2.0, chunk=0.125):
    """C4 over its configured horizon t in [0, tf] for the plain, fast FP64 + k1 reuse and
    chosen paths.  Each path runs in chunks of `chunk` (every chunk a fresh solve from
    the previous chunk's end state, logged as it ends), so no call is silent for minutes;
    a restart re-estimates every system's initial step.  Reports time, RHS evaluations
    and energy drift at tf over 64 systems (final states at tf are not compared: over
    this horizon close encounters make the systems chaotic)."""
    from ode_prototyping_amd.ensemble import solve_ensemble, unpack
    from ode_prototyping_amd import train_controller as trc
    eps, tol, tab = 1e-3, 1e-10, "DOPRI5"
    sizes, x0, m = trc.workload_ensemble(2048, 0)
    md = oh.HipState(exe, data=m)
    a, b = oh.HipState(exe, data=x0), oh.HipState(exe, x0.size)
    net, info = trc.pretrained(exe, tab)
    close = oh.MIXED_NEAR_CLOSE
    runs = {"plain_exact": ("exact", "reference", None, 0.0, False),
            "fast_fsal": ("fast", "reference", None, 0.0, True),
            "chosen": ("mixed", "learned", net, close, True)}
    ms = unpack(m, sizes, per_particle=1)
    x0s = unpack(x0, sizes)
    idx = np.linspace(0, len(sizes) - 1, 64).astype(int)
    e0 = {int(k): nbody_energy(oh, exe, x0s[k], int(sizes[k]), eps, ms[k]) for k in idx}
    n_chunks = int(round(tf / chunk))
    block = {"workload": f"{len(sizes)} systems as the ensemble block, t in [0, {tf}] in "
                         f"{n_chunks} chunks of {chunk}", "learned_net": info}
    for name, (var, ctrl, mlp, near, fsal) in runs.items():
        exe.near_field = near
        a.upload(x0)
        secs = evals = steps = 0.0
        for c in range(n_chunks):
            t0, t1 = c * chunk, (c + 1) * chunk
            exe.synchronize()
            t = time.perf_counter()
            r = solve_ensemble(tab, sizes, a, t0, t1, tol, tol, eps, md, variant=var, exe=exe,
                               controller=ctrl, mlp=mlp, fsal=fsal, x_out=b, fetch=False)
            exe.synchronize()
            secs += time.perf_counter() - t
            evals += r.total_rhs_evals
            steps += r.particle_steps
            a, b = b, a
            log(f"horizon {name}: chunk {c + 1}/{n_chunks} t={t1:g} {secs:.1f} s")
        got = unpack(a.copy_to_span(), sizes)
        drift = [abs((nbody_energy(oh, exe, got[k], int(sizes[k]), eps, ms[k]) - e0[int(k)]) /
                     e0[int(k)]) for k in idx]
        block[name] = {"seconds": secs, "particle_steps_per_s": steps / secs,
                       "rhs_evals_per_solve": evals / len(sizes),
                       "energy_rel_drift_64_systems": {"median": float(np.median(drift)),
                                                       "max": float(np.max(drift))}}
        log(f"horizon {name}: {block[name]}")
    exe.near_field = oh.MIXED_NEAR_DEFAULT
    net.close()
    return block


def ensemble_cpu_baseline(sizes, x0, m, eps, tol, tab, tf=0.005, count=12):
    """CPU baseline of the ensemble workload on a bounded sample: 12 systems spread over
    the size order, each integrated alone to tf.  The reference's own ThreadPoolExecutor
    path (oracle/_ref: RKEmbeddedParallel<BTDOPRI5> + NBodyODE, instantiated at these
    systems' sizes) where built; otherwise the oracle restatement."""
    from ode_prototyping_amd.ensemble import unpack
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    order = np.argsort(sizes, kind="stable")
    idx = order[np.linspace(0, len(sizes) - 1, count).astype(int)]
    xs, ms = unpack(x0, sizes), unpack(m, sizes, per_particle=1)
    ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    steps = 0.0
    if os.path.exists(ref) and tab == "DOPRI5":
        L = C.CDLL(ref)
        dp = C.POINTER(C.c_double)
        L.ref_solve_dopri5_n.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double,
                                         C.c_double, C.c_double, dp, C.POINTER(C.c_long)]
        L.ref_solve_dopri5_n.restype = C.c_double
        secs = 0.0
        for k in idx:
            n = int(sizes[k])
            xo = np.zeros(6 * n)
            saved = C.c_long()
            dt = L.ref_solve_dopri5_n(n, threads, np.ascontiguousarray(xs[k]).ctypes.data_as(dp),
                                      tol, eps, float(ms[k][0]), tf, xo.ctypes.data_as(dp),
                                      C.byref(saved))
            if dt < 0:
                raise RuntimeError(f"oracle/_ref has no DOPRI5 instance for N={n}")
            secs += dt
            steps += float(n) * (saved.value - 1)
        kind = "reference"
        what = (f"reference RKEmbeddedParallel<DOPRI5> + ThreadPoolExecutor({threads}) + "
                f"NBodyODE (atomic pair RHS), {count} systems spread over the size order "
                f"(N {int(sizes[idx].min())}..{int(sizes[idx].max())}), each solved alone for "
                f"t in [0, {tf}], uniform masses applied by scaling the unit-mass RHS")
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        t = time.perf_counter()
        for k in idx:
            r = O.integrate(tab, O.RHS_NBODY, xs[k], 0.0, tf, tol, tol, softening=eps,
                            masses=ms[k], store_states=False, threads=threads)
            steps += float(sizes[k]) * r.accepted
        secs = time.perf_counter() - t
        kind = "port"
        what = (f"oracle restatement ({threads} threads), {count} systems spread over "
                f"the size order (N {int(sizes[idx].min())}..{int(sizes[idx].max())}), "
                f"each solved alone for t in [0, {tf}]")
    return {"value": steps / secs, "unit": "particle-steps/s", "cores": threads,
            "kind": kind, "seconds": secs, "sample": what}


def controller_block(oh, exe):
    """Learned vs PI vs reference controller on 4096 held-out systems."""
    from ode_prototyping_amd.ensemble import solve_ensemble
    from ode_prototyping_amd import train_controller as trc
    tf, eps, tol, tab = 0.05, 1e-3, 1e-10, "DOP853"
    sizes, x0, m = trc.workload_ensemble(4096, 9000)
    x0d, md = oh.HipState(exe, data=x0), oh.HipState(exe, data=m)
    out_state = oh.HipState(exe, x0.size)
    net, info = trc.pretrained(exe, tab)
    net_fast, info_fast = trc.pretrained(exe, tab, "fast")
    block = {"workload": f"{len(sizes)} held-out systems (seed 9000), N log-uniform "
                         f"[16, 4096], eps={eps}, {tab}, rtol=atol={tol}, t in [0, {tf}], "
                         "mixed-precision and fast FP64 RHS", "learned_net": info,
             "learned_net_fast": info_fast}
    w2 = sizes.astype(float) ** 2
    # (key, RHS variant, near-field coefficient, controller, network): the mixed RHS with
    # the network trained on its traces, then with the close-encounter near field; in
    # FP64 (fast RHS) the mixed-trained network, and the one trained on fast-RHS traces
    # (its target error chosen on a separate validation ensemble)
    runs = [("reference", "mixed", 0.0, "reference", None), ("pi", "mixed", 0.0, "pi", None),
            ("learned", "mixed", 0.0, "learned", net),
            ("near_reference", "mixed", oh.MIXED_NEAR_CLOSE, "reference", None),
            ("near_learned", "mixed", oh.MIXED_NEAR_CLOSE, "learned", net),
            ("fast_reference", "fast", 0.0, "reference", None),
            ("fast_learned_mixed_net", "fast", 0.0, "learned", net),
            ("fast_learned", "fast", 0.0, "learned", net_fast)]
    for key, var, near, ctrl, mlp in runs:
        exe.near_field = near
        solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, variant=var, exe=exe,
                       controller=ctrl, mlp=mlp, max_attempts=1, x_out=out_state, fetch=False)
        exe.synchronize()
        t = time.perf_counter()
        r = solve_ensemble(tab, sizes, x0d, 0.0, tf, tol, tol, eps, md, variant=var,
                           exe=exe, controller=ctrl, mlp=mlp, x_out=out_state, fetch=False)
        exe.synchronize()
        secs = time.perf_counter() - t
        block[key] = {"variant": var, "near_field": near, "controller": ctrl, "seconds": secs,
                      "rhs_evals_per_solve": r.total_rhs_evals / len(sizes),
                      "pair_work_per_solve": float(np.sum(w2 * r.rhs_evals)) / len(sizes),
                      "accepted": r.total_accepted, "rejected": r.total_rejected,
                      "rounds": r.rounds, "learned_steps": r.total_learned_steps,
                      "particle_steps_per_s": r.particle_steps / secs}
        log(f"controller {key}: {block[key]}")
    exe.near_field = oh.MIXED_NEAR_DEFAULT

    def gain(learned, reference):
        lr, rf = block[learned], block[reference]
        return {"rhs_evals": 1.0 - lr["rhs_evals_per_solve"] / rf["rhs_evals_per_solve"],
                "pair_work": 1.0 - lr["pair_work_per_solve"] / rf["pair_work_per_solve"],
                "seconds": 1.0 - lr["seconds"] / rf["seconds"]}

    block["learned_gain"] = gain("learned", "reference")
    block["learned_gain_near_field"] = gain("near_learned", "near_reference")
    block["learned_gain_fast_fp64"] = gain("fast_learned", "fast_reference")
    block["learned_gain_fast_fp64_mixed_net"] = gain("fast_learned_mixed_net", "fast_reference")
    net.close()
    net_fast.close()
    return block


def outputs_block(oh, exe, x0_dev, atol, rtol, ode, n, eps, m, accepted, plain_secs, t_end):
    """The output paths and the leapfrog integrator on the headline workload:
    every accepted state kept (RawOutput), DOP853 dense output at 100 times,
    the .bin trajectory write, and fixed-step leapfrog on the same RHS."""
    import tempfile
    tab = "DOP853"
    block = {"workload": f"{n}-body Plummer sphere, eps={eps}, {tab}, rtol=atol=1e-10, "
                         f"{accepted} accepted steps (plain FP64 solve: {plain_secs:.3f} s)"}

    def timed(**kw):
        exe.synchronize()
        t = time.perf_counter()
        r = oh.integrate(tab, x0_dev, 0.0, 100.0, atol, rtol, ode, exe=exe,
                         max_accepted=accepted, **kw)
        exe.synchronize()
        return r, time.perf_counter() - t

    raw = oh.RawOutput()
    r, secs = timed(output=raw)
    block["raw_output"] = {"seconds": secs, "overhead": secs / plain_secs - 1.0,
                           "states": len(raw.times()),
                           "bytes_to_host": len(raw.times()) * 6 * n * 8}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "trajectory.bin")
        t = time.perf_counter()
        oh.nbody_io.write_output(path, raw, softening=eps)
        wsecs = time.perf_counter() - t
        size = os.path.getsize(path)
    block["trajectory_file"] = {"seconds": wsecs, "bytes": size, "MB_per_s": size / wsecs / 1e6}
    del raw
    times = np.linspace(0.0, t_end, 102)[1:-1]
    r, secs = timed(dense_times=times, fsal=True)
    block["dense_output"] = {"seconds": secs, "overhead": secs / plain_secs - 1.0,
                             "n_dense": r.n_dense, "dense_steps": r.dense_steps,
                             "dense_rhs_evals": r.dense_rhs_evals,
                             "rhs_evals": r.rhs_evals, "accepted": r.accepted}
    # leapfrog: 100 fixed steps of dt = 1e-3 on the same RHS kernel
    dt, steps = 1e-3, 100
    e0 = nbody_energy(oh, exe, x0_dev, n, eps, m)
    exe.synchronize()
    t = time.perf_counter()
    x_lf, lf = oh.leapfrog(x0_dev, 0.0, steps * dt - 0.5 * dt, dt, ode, exe=exe)
    exe.synchronize()
    secs = time.perf_counter() - t
    e1 = nbody_energy(oh, exe, x_lf, n, eps, m)
    block["leapfrog"] = {"dt": dt, "steps": lf["steps"], "t": lf["t"], "seconds": secs,
                         "particle_steps_per_s": n * lf["steps"] / secs,
                         "ms_per_step": 1000 * secs / lf["steps"],
                         "rhs_evals": lf["rhs_evals"], "energy_rel_err": abs((e1 - e0) / e0)}
    for k, v in block.items():
        if k != "workload":
            log(f"outputs {k}: {v}")
    return block


def c1_block(oh, exe, reps=5):
    """The reference's CPU-runnable configuration: N = 64 uniform cube (seed 0, masses 1/64),
    eps 1e-3, DOPRI5, rtol = atol = 1e-10, t in [0, 1].  The device solve against the
    reference's own single-thread and thread-pool paths (oracle/_ref) on this host, and
    the largest final-state difference between them (the device reduces the error norm in
    its canonical order and uses a correctly rounded pow, the reference's executors their
    own orders and glibc's pow)."""
    n, eps, tol, tf = 64, 1e-3, 1e-10, 1.0
    x0, m = oh.scenarios.uniform_cube(n, seed=0)
    ode = oh.NBodyODE(eps, m, "exact")
    x0_dev = oh.HipState(exe, data=x0)
    res = oh.integrate("DOPRI5", x0_dev, 0.0, tf, tol, tol, ode, exe=exe)  # untimed
    exe.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        res = oh.integrate("DOPRI5", x0_dev, 0.0, tf, tol, tol, ode, exe=exe)
    exe.synchronize()
    secs = (time.perf_counter() - t) / reps
    block = {"workload": f"N={n} uniform cube (seed 0, masses 1/{n}), eps={eps}, DOPRI5, "
                         f"rtol=atol={tol}, t in [0, {tf}]",
             "gpu": {"seconds": secs, "accepted": res.accepted, "rejected": res.rejected,
                     "rhs_evals": res.rhs_evals, "particle_steps_per_s": n * res.accepted / secs}}
    log(f"c1 gpu: {block['gpu']}")
    ref = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if os.path.exists(ref):
        L = C.CDLL(ref)
        dp = C.POINTER(C.c_double)
        L.ref_solve_64_dopri5.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, dp, C.POINTER(C.c_long)]
        L.ref_solve_64_dopri5.restype = C.c_double
        threads = max(2, min(16, len(os.sched_getaffinity(0))))
        for key, th in (("cpu_reference_single_thread", 1), ("cpu_reference_thread_pool", threads)):
            xo = np.zeros(6 * n)
            saved = C.c_long()
            sec = L.ref_solve_64_dopri5(th, np.ascontiguousarray(x0).ctypes.data_as(dp), tol, eps,
                                        float(m[0]), tf, xo.ctypes.data_as(dp), C.byref(saved))
            acc = saved.value - 1
            block[key] = {"threads": th, "seconds": sec, "accepted": acc,
                          "particle_steps_per_s": n * acc / sec,
                          "final_state_rel_diff_vs_gpu": float(
                              np.max(np.abs(xo - res.x_final)) / np.max(np.abs(xo)))}
            log(f"c1 {key}: {block[key]}")
    x0_dev.free()
    return block


def precision_block(oh, exe):
    """Mixed precision on 65536 colliding Plummer bodies, DOP853."""
    n, eps, tol, tab, acc = 65536, 1e-2, 1e-10, "DOP853", 10
    x0, m = oh.scenarios.colliding_plummer(n, seed=0)
    x0d = oh.HipState(exe, data=x0)
    atol = oh.HipState(exe, data=np.full(6 * n, tol))
    e0 = nbody_energy(oh, exe, x0d, n, eps, m)
    block = {"workload": f"{n} bodies: two Plummer spheres (a=1, M=1 total) at +-2 on x "
                         f"closing at 0.3, eps={eps}, {tab}, rtol=atol={tol}, {acc} accepted "
                         "steps"}
    finals = {}
    for name in ("exact", "mixed", "mixed_near_field"):
        var = "mixed" if name == "mixed_near_field" else name
        exe.near_field = (oh.MIXED_NEAR_CLOSE if name == "mixed_near_field"
                          else oh.MIXED_NEAR_DEFAULT)
        o = oh.NBodyODE(eps, m, var)
        t = time.perf_counter()
        r = oh.integrate(tab, x0d, 0.0, 100.0, atol, atol, o, exe=exe, max_accepted=acc)
        secs = time.perf_counter() - t
        finals[name] = r.x_final
        e1 = nbody_energy(oh, exe, r.x_final, n, eps, m)
        block[name] = {"seconds": secs, "particle_steps_per_s": n * r.accepted / secs,
                       "accepted": r.accepted, "rejected": r.rejected, "t": r.t,
                       "rhs_evals": r.rhs_evals, "energy_rel_err": abs((e1 - e0) / e0),
                       "near_field": exe.near_field if var == "mixed" else None}
        log(f"precision {name}: {block[name]}")
    exe.near_field = oh.MIXED_NEAR_DEFAULT
    # the exact path equals the FP64 oracle bit for bit; compare at the same time
    o = oh.NBodyODE(eps, m, "exact")
    for name in ("mixed", "mixed_near_field"):
        ref = oh.integrate(tab, x0d, 0.0, block[name]["t"], atol, atol, o, exe=exe).x_final
        block[name]["final_state_rel_err_vs_fp64_same_t"] = float(
            np.max(np.abs(finals[name] - ref)) / np.max(np.abs(ref)))
    return block


def near_field_block(oh, exe):
    """Mixed RHS through close encounters: all pairs FP32 vs the FP64 near field vs FP64."""
    tf, eps, tol, tab = 0.1, 1e-3, 1e-10, "DOPRI5"
    block = {"workload": f"single uniform-cube systems (seed = N, masses 1/N), eps={eps}, {tab}, "
                         f"rtol=atol={tol}, t in [0, {tf}]; final state vs the exact FP64 solve",
             "near_field": oh.MIXED_NEAR_CLOSE}
    for n in (28, 52, 89, 126):
        x0, m = oh.scenarios.uniform_cube(n, seed=n)
        ref = oh.integrate(tab, x0, 0.0, tf, tol, tol, oh.NBodyODE(eps, m, "exact"), exe=exe)
        row = {"fp64_rhs_evals": ref.rhs_evals}
        for name, c in (("all_fp32", 0.0), ("near_field", oh.MIXED_NEAR_CLOSE)):
            exe.near_field = c
            r = oh.integrate(tab, x0, 0.0, tf, tol, tol, oh.NBodyODE(eps, m, "mixed"), exe=exe)
            row[name] = {"rhs_evals": r.rhs_evals, "final_state_rel_err": float(
                np.max(np.abs(r.x_final - ref.x_final)) / np.max(np.abs(ref.x_final)))}
        exe.near_field = oh.MIXED_NEAR_DEFAULT
        block[f"n{n}"] = row
        log(f"near field n={n}: {row}")
    return block


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
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-tf", type=float, default=0.1,
                    help="end time of the ensemble block (the configuration's horizon is 2)")
    ap.add_argument("--near-sweep", action="store_true",
                    help="only run the 2048-system ensemble (t <= 0.1) with the mixed RHS at "
                         "near-field coefficients 20 .. 2000, against plain and fast FP64")
    ap.add_argument("--ensemble-horizon", action="store_true",
                    help="only run the ensemble over t in [0, 2] (or --ensemble-tf if larger "
                         "than 0.1) in chunks: plain, fast FP64 + k1 reuse and chosen paths")
    ap.add_argument("--no-controller", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-outputs", action="store_true",
                    help="skip the output-path block (also skipped by --no-extras)")
    ap.add_argument("--profile-ensemble", action="store_true",
                    help="only run short ensemble solves (2048 systems, t in [0, 0.005]) in "
                         "each RHS variant (the workload of the committed ensemble profiles)")
    ap.add_argument("--profile-ensemble-rhs", action="store_true",
                    help="only run the ensemble right-hand side of the 2048-system input 8x "
                         "per variant (exact, fast, mixed, mixed with the near field; the last "
                         "5 timed), tile configuration 7, no graphs (the workload of the "
                         "committed ensemble RHS profile)")
    ap.add_argument("--profile-ensemble-tail", action="store_true",
                    help="only run the ensemble right-hand side of the 32 and of the 8 "
                         "largest systems of the 2048-system input (a solve's tail: few "
                         "running systems), exact and fast, one wave per tile "
                         "(configuration 7) against the shared tiles the tail rule picks "
                         "(4 and 8 waves per tile: configurations 15 and 17); 8 launches "
                         "each, the last 5 timed, no graphs")
    ap.add_argument("--profile-rhs", action="store_true",
                    help="only run each RHS variant 3x, a 64 MiB calibration copy and "
                         "the near-field mixed RHS 3x (the workload of the committed "
                         "counter profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): this benchmark "
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
        exe.near_field = oh.MIXED_NEAR_CLOSE  # the mixed kernel's near-field instance
        o = oh.NBodyODE(1e-2, m, "mixed")
        for _ in range(3):
            o(exe, xs, ks)
        exe.near_field = oh.MIXED_NEAR_DEFAULT
        for _ in range(3):
            b.assign(a)
        exe.synchronize()
        log("profile-rhs done")
        return

    if args.near_sweep:
        exe = oh.HipExecutor(0)
        print(json.dumps({"near_sweep": near_sweep_block(oh, exe)}), flush=True)
        return

    if args.ensemble_horizon:
        exe = oh.HipExecutor(0)
        print(json.dumps({"ensemble_horizon": ensemble_horizon_block(oh, exe, args.ensemble_tf
                                                                      if args.ensemble_tf > 0.1
                                                                      else 2.0)}), flush=True)
        return

    if args.profile_ensemble_rhs:
        from ode_prototyping_amd.ensemble import ensemble_rhs
        from ode_prototyping_amd import train_controller as trc
        sizes, x0, m = trc.workload_ensemble(2048, 0)
        exe = oh.HipExecutor(0)
        x0d, md = oh.HipState(exe, data=x0), oh.HipState(exe, data=m)
        kd = oh.HipState(exe, x0.size)
        for var, near in (("exact", 0.0), ("fast", 0.0), ("mixed", 0.0),
                          ("mixed", oh.MIXED_NEAR_CLOSE)):
            exe.near_field = near
            ms = [ensemble_rhs(exe, sizes, x0d, 1e-3, md, var, tile_config=7, dxdt=kd)[1]
                  for _ in range(8)][3:]  # three warm-up launches, then five
            log(f"profile-ensemble-rhs {var} near {near}: {min(ms):.3f} ms (HIP events)")
        exe.near_field = oh.MIXED_NEAR_DEFAULT
        return

    if args.profile_ensemble_tail:
        from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
        from ode_prototyping_amd import train_controller as trc
        sizes, x0, m = trc.workload_ensemble(2048, 0)
        order = np.argsort(-sizes, kind="stable")
        xs, ms = unpack(x0, sizes), unpack(m, sizes, per_particle=1)
        exe = oh.HipExecutor(0)
        for count, split_cfg in ((32, 15), (8, 17)):
            pick = order[:count]
            sub = sizes[pick]
            x0d = oh.HipState(exe, data=pack([xs[k] for k in pick]))
            md = oh.HipState(exe, data=pack([ms[k] for k in pick]))
            kd = oh.HipState(exe, x0d.size)
            for var in ("exact", "fast"):
                for cfg in (7, split_cfg):
                    t = [ensemble_rhs(exe, sub, x0d, 1e-3, md, var, tile_config=cfg,
                                      dxdt=kd)[1] for _ in range(8)][3:]
                    log(f"profile-ensemble-tail {count} largest (sum n {int(sub.sum())}) "
                        f"{var} config {cfg}: {min(t):.4f} ms min, "
                        f"{sum(t) / len(t):.4f} ms mean (HIP events)")
        return

    if args.profile_ensemble:
        from ode_prototyping_amd.ensemble import solve_ensemble
        from ode_prototyping_amd import train_controller as trc
        sizes, x0, m = trc.workload_ensemble(2048, 0)
        exe = oh.HipExecutor(0)
        x0d, md = oh.HipState(exe, data=x0), oh.HipState(exe, data=m)
        out_state = oh.HipState(exe, x0.size)
        for var in ("exact", "fast", "mixed"):
            for rep in range(2):
                r = solve_ensemble("DOPRI5", sizes, x0d, 0.0, 0.005, 1e-10, 1e-10, 1e-3, md,
                                   variant=var, exe=exe, x_out=out_state, fetch=False)
                exe.synchronize()
            log(f"profile-ensemble {var}: rounds {r.rounds} tile {r.tile_config} "
                f"rhs_ms {r.rhs_ms:.3f}")
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
    plain_final = work.copy_to_span()  # the plain path's state at r.t
    log(f"timed: {steps} solves in {wall:.3f} s (events {ev_ms:.1f} ms): {r}")
    accepted = r.accepted
    ms_per_step = 1000.0 * wall / steps
    value = n * accepted * steps / wall

    # roofline of the dominant kernel (the RHS).  Its time per launch inside the timed
    # region: the timed solves' HIP-event time (library stream) over their RHS launches --
    # an upper bound, since it includes the attempts' stage / final / controller kernels
    # (about 1.5 % of a solve in the kernel trace).  Beside it, 20 isolated back-to-back
    # launches, which long uninterrupted runs can catch in a clock transient.
    xs = work
    ks = oh.HipState(exe, 6 * n)
    rhs_ms = ev_ms / steps / r.rhs_evals
    iso_ms = rhs_kernel_ms(exe, oh, xs, ks, ode, 20)
    interactions = n * (n - 1)
    achieved = FLOP_PER_INTERACTION * interactions / (rhs_ms * 1e-3) / 1e12
    log(f"rhs kernel ({args.variant}): {rhs_ms:.4f} ms/launch in the timed solves "
        f"({iso_ms:.4f} ms isolated), {achieved:.2f} TFLOP/s")

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
                     "kernel_ms_source": "timed solves' event time / their RHS launches "
                                         "(upper bound: includes ~1.5 % element kernels)",
                     "kernel_ms_isolated": iso_ms,
                     "sustained_fp64_tflops": sus_tflops, "clock_mhz": sus_mhz,
                     "frac_of_sustained": achieved / sus_tflops,
                     "instruction_bound": instruction_bound(args.variant, n, sus_mhz, rhs_ms),
                     "algorithmic": f"{FLOP_PER_INTERACTION} FLOP x N(N-1) = "
                                    f"{FLOP_PER_INTERACTION * interactions:.4g} FLOP/launch"},
    }

    if not args.no_extras:
        extras = {}
        for name in ("exact", "fast", "mixed", "mixed_near_field"):
            var = "mixed" if name == "mixed_near_field" else name
            if name == args.variant:
                continue
            exe.near_field = (oh.MIXED_NEAR_CLOSE if name == "mixed_near_field"
                              else oh.MIXED_NEAR_DEFAULT)
            o2 = oh.NBodyODE(softening, m, var)
            t = time.perf_counter()
            r2 = solve(o2)
            exe.synchronize()
            secs = time.perf_counter() - t
            kms = rhs_kernel_ms(exe, oh, xs, ks, o2, 5)
            extras[name] = {"particle_steps_per_s": n * r2.accepted / secs,
                            "near_field": exe.near_field if var == "mixed" else None,
                            "ms_per_solve": 1000 * secs, "rhs_evals_per_solve": r2.rhs_evals,
                            "rhs_kernel_ms": kms,
                            "rhs_tflops": FLOP_PER_INTERACTION * interactions / (kms * 1e-3) / 1e12}
            log(f"{name}: {extras[name]}")
        exe.near_field = oh.MIXED_NEAR_DEFAULT
        # chosen path on this workload: mixed-precision RHS + k1 reuse after rejections
        o3 = oh.NBodyODE(softening, m, "mixed")
        t = time.perf_counter()
        r3 = solve(o3, fsal=True)
        exe.synchronize()
        secs = time.perf_counter() - t
        # accuracy at equal t: the chosen path integrated to the plain solve's end time
        # r.t (its last step clipped there), against the plain path's state at r.t
        at_t = oh.integrate(tableau, x0_dev, 0.0, r.t, atol, rtol, o3, exe=exe, fsal=True)
        rel = float(np.max(np.abs(at_t.x_final - plain_final)) / np.max(np.abs(plain_final)))
        extras["chosen_mixed_fsal"] = {"particle_steps_per_s": n * r3.accepted / secs,
                                       "rhs_evals_per_solve": r3.rhs_evals,
                                       "final_state_rel_err_vs_plain": rel,
                                       "compared_at_t": r.t}
        log(f"chosen: {extras['chosen_mixed_fsal']}")
        out["variants"] = extras
    if not args.no_extras and not args.no_outputs:
        out["outputs"] = outputs_block(oh, exe, x0_dev, atol, rtol, ode, n, softening, m,
                                       args.accepted, wall / steps, r.t)

    if not args.no_ensemble:
        out["ensemble"] = ensemble_block(oh, exe, args.ensemble_tf)
    if not args.no_controller:
        out["controller"] = controller_block(oh, exe)
    if not args.no_precision:
        out["precision"] = precision_block(oh, exe)
        out["near_field"] = near_field_block(oh, exe)
    if not args.no_extras:
        out["c1"] = c1_block(oh, exe)

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(x0, m, softening, tol, n, 2)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Train the learned step-size controller on the GPU and save its weights.

    python -m ode_prototyping_amd.train_controller --tableau DOP853 --out step_mlp_dop853.npz

Training data: ensembles of the batched-ensemble workload (N log-uniform in
[16, 4096], positions uniform in [-1, 1]^3, velocities N(0, 0.1^2), masses
1/N, softening 1e-3, rtol = atol = 1e-10) integrated from t = 0 towards 2
with the PI controller for a fixed number of attempts per system, one
ensemble per training seed, using the right-hand-side variant the
controller is meant for (default: mixed precision); every attempt after the second gives one
(features, target) pair (see controller.training_pairs).  The network is
trained by the library's FP32 forward/backward kernels with Adam.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import controller as ctl
from .core import HipExecutor
from .ensemble import pack, solve_ensemble
from . import scenarios

ORDERS = {"DOPRI5": 5, "DOP853": 6, "RKF78": 8, "DVERK": 6, "RKF45": 5, "HE21": 2}
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def workload_ensemble(count, seed, n_min=16, n_max=4096):
    """Sizes and packed initial state / masses of one ensemble (seeded)."""
    sizes = scenarios.log_uniform_sizes(count, n_min, n_max, seed=seed)
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = scenarios.uniform_cube(int(n), seed=seed * 100003 + k)
        xs.append(x)
        ms.append(m)
    return sizes, pack(xs), pack(ms)


def collect_traces(exe, tableau, seeds, systems, attempts, variant="mixed", log=None,
                   near_field=0.0):
    traces, sizes_all = [], []
    saved = exe.near_field
    exe.near_field = near_field
    for seed in seeds:
        sizes, x0, m = workload_ensemble(systems, seed)
        res = solve_ensemble(tableau, sizes, x0, 0.0, 2.0, 1e-10, 1e-10, 1e-3, m,
                             variant=variant, exe=exe, controller="pi",
                             max_attempts=attempts, trace_cap=attempts, fetch=False)
        traces += res.trace
        sizes_all += [int(n) for n in sizes]
        if log:
            log(f"seed {seed}: {res.total_attempts} attempts, {res.total_rejected} rejected")
    exe.near_field = saved
    return traces, sizes_all


def train(exe, tableau, seeds=range(1000, 1008), systems=256, attempts=400, tau=0.65,
          epochs=10, batch=4096, lr=2e-3, seed=1, variant="mixed", near_field=0.0, log=None):
    """Returns (StepMLP, info dict).  `variant` (and `near_field`, its FP64 near-pair
    coefficient for the mixed variant): the right-hand side the traces are integrated
    with -- the one the controller will run with (its error norms carry that variant's
    rounding noise)."""
    t0 = time.perf_counter()
    traces, sizes = collect_traces(exe, tableau, seeds, systems, attempts, variant=variant,
                                   log=log, near_field=near_field)
    t_traces = time.perf_counter() - t0
    k = ORDERS[tableau]
    x, t = ctl.training_pairs(traces, sizes, 2.0, k, tau=tau)
    mean, std = ctl.normalisation(x)
    net = ctl.StepMLP(exe, mean, std, seed=seed)
    t1 = time.perf_counter()
    losses = net.train(x, t, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    t_train = time.perf_counter() - t1
    info = {"tableau": tableau, "k": k, "tau": tau, "variant": variant,
            "near_field": near_field,
            "samples": int(x.shape[0]),
            "seeds": [int(s) for s in seeds], "systems_per_seed": systems,
            "attempts_per_system": attempts, "epochs": epochs, "batch": batch, "lr": lr,
            "epoch_loss": [float(v) for v in losses], "trace_seconds": t_traces,
            "train_seconds": t_train}
    return net, info


def save(net, info, path):
    np.savez(path, params=net.params(), mean=net.mean, std=net.std,
             info=np.array(json.dumps(info)))


def load(exe, path):
    """A StepMLP from a file written by save (plain arrays, no pickles)."""
    with np.load(path, allow_pickle=False) as z:
        net = ctl.StepMLP(exe, z["mean"], z["std"], params=z["params"])
        info = json.loads(str(z["info"]))
    return net, info


def pretrained(exe, tableau, variant="mixed"):
    """The network shipped with the package for this tableau, trained on traces of the
    given right-hand-side variant (step_mlp_<tableau>.npz: mixed precision;
    step_mlp_<tableau>_fast.npz: fast FP64, DOP853 only).  The error norms of the two
    carry different rounding noise, so each network is for its own variant."""
    suffix = "" if variant == "mixed" else f"_{variant}"
    path = os.path.join(DATA_DIR, f"step_mlp_{tableau.lower()}{suffix}.npz")
    if not os.path.exists(path):
        raise FileNotFoundError(f"no trained controller for {tableau}, {variant} RHS ({path})")
    return load(exe, path)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--tableau", default="DOP853", choices=sorted(ORDERS))
    ap.add_argument("--out", required=True)
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--attempts", type=int, default=400)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--tau", type=float, default=0.65)
    ap.add_argument("--variant", default="mixed", choices=["exact", "fast", "mixed"])
    ap.add_argument("--near-field", type=float, default=0.0,
                    help="FP64 near-pair coefficient of the mixed RHS for the traces")
    ap.add_argument("--epochs", type=int, default=10)
    args = ap.parse_args(argv)
    exe = HipExecutor(0)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    net, info = train(exe, args.tableau, seeds=range(1000, 1000 + args.seeds),
                      systems=args.systems, attempts=args.attempts, tau=args.tau,
                      epochs=args.epochs, variant=args.variant,
                      near_field=args.near_field, log=log)
    save(net, info, args.out)
    log(json.dumps(info))


if __name__ == "__main__":
    main()

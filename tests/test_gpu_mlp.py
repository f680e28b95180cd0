"""The learned controller's network kernels against the CPU oracle, and the
learned controller inside the solves.

* the FP32 training kernels' loss and gradient equal the oracle's (double)
  to FP32 accumulation accuracy;
* the BF16 matrix-core inference equals the oracle's bfloat16-emulated
  forward to the level of one bfloat16 rounding of an activation;
* training is deterministic and fits a teacher network;
* with the learned controller every accepted attempt still has an error norm
  <= 1, the PI fallback sizes the first attempts, and the single-system and
  ensemble solves agree bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _net_data(n=300, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.25, O.MLP_PARAMS).astype(np.float32)
    mean = rng.normal(0.0, 1.0, O.MLP_IN).astype(np.float32)
    std = rng.uniform(0.5, 2.0, O.MLP_IN).astype(np.float32)
    x = (mean + std * rng.normal(size=(n, O.MLP_IN))).astype(np.float32)
    t = rng.normal(0.0, 0.5, n).astype(np.float32)
    return p, mean, std, x, t


@pytest.mark.parametrize("n", [1, 63, 300, 5000])
def test_gradient_matches_oracle(exe, n):
    from ode_prototyping_amd.controller import StepMLP
    p, mean, std, x, t = _net_data(n, seed=n)
    net = StepMLP(exe, mean, std, params=p)
    loss, g = net.loss_grad(x, t)
    loss_o, g_o = O.mlp_loss_grad(p.astype(np.float64), mean, std, x, t)
    assert abs(loss - loss_o) <= 1e-5 * max(loss_o, 1e-6)
    scale = np.max(np.abs(g_o))
    assert np.max(np.abs(g - g_o)) <= 2e-5 * scale, np.max(np.abs(g - g_o)) / scale
    net.close()


def test_inference_matches_bf16_oracle(exe):
    from ode_prototyping_amd.controller import StepMLP
    p, mean, std, x, _ = _net_data(1000, seed=5)
    net = StepMLP(exe, mean, std, params=p)
    y = net.infer(x)
    yb = O.mlp_forward(p.astype(np.float64), mean, std, x, bf16=True)
    yf = O.mlp_forward(p.astype(np.float64), mean, std, x, bf16=False)
    d = np.abs(y - yb)
    assert np.median(d) < 1e-4 and np.max(d) < 2e-2, (np.median(d), np.max(d))
    # and bfloat16 inference tracks the float network
    assert np.max(np.abs(y - yf)) < 0.1 * max(1.0, np.max(np.abs(yf)))
    # rows are independent: a single row gives the same value
    y1 = net.infer(x[7:8])
    assert y1[0] == y[7]
    net.close()


def test_training_fits_teacher_and_is_deterministic(exe):
    from ode_prototyping_amd.controller import StepMLP
    p, mean, std, x, _ = _net_data(20000, seed=9)
    t = O.mlp_forward(p.astype(np.float64), mean, std, x).astype(np.float32)
    a = StepMLP(exe, mean, std, seed=1)
    la = a.train(x, t, epochs=8, batch=1024, lr=3e-3, seed=4)
    b = StepMLP(exe, mean, std, seed=1)
    lb = b.train(x, t, epochs=8, batch=1024, lr=3e-3, seed=4)
    assert np.array_equal(a.params(), b.params()) and np.array_equal(la, lb)
    assert la[-1] < 0.25 * la[0] and np.all(np.diff(la) < 0), la
    # the published inference copy is the trained network
    y = a.infer(x[:500])
    yb = O.mlp_forward(a.params().astype(np.float64), mean, std, x[:500], bf16=True)
    assert np.median(np.abs(y - yb)) < 1e-4
    a.close()
    b.close()


def _small_net(exe):
    """A network that proposes a constant ratio 1.3 (zero weights, b3)."""
    from ode_prototyping_amd.controller import StepMLP
    p = np.zeros(O.MLP_PARAMS, np.float32)
    p[-1] = np.log(1.3)
    return StepMLP(exe, np.zeros(10, np.float32), np.ones(10, np.float32), params=p)


def test_learned_controller_keeps_the_acceptance_test(exe):
    """A network that always proposes dt * 1.3: the error norm still decides
    acceptance, the PI controller sizes attempts 1-3 and after two
    rejections in a row, and the run reaches tf."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    net = _small_net(exe)
    sizes = [20, 90, 300]
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = oh.scenarios.uniform_cube(n, seed=40 + k)
        xs.append(x)
        ms.append(m)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.02, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe, controller="learned", mlp=net, trace_cap=8192)
    assert res.n_unfinished == 0 and np.all(res.t == 0.02)
    ratios = []
    for k in range(len(sizes)):
        tr = res.trace[k]
        assert np.all(tr[tr[:, 2] == 1, 0] <= 1.0) and np.all(tr[tr[:, 2] == 0, 0] > 1.0)
        ratios.append(tr[1:, 1] / tr[:-1, 1])
    # learned steps: dt grows by exactly 1.3 after an accepted attempt with
    # history (unless the clip to tf shortens the step)
    assert np.sum(np.isclose(np.concatenate(ratios), 1.3, rtol=1e-6)) >= 3
    assert res.total_learned_steps > 0
    assert res.total_learned_steps < res.total_attempts  # PI sized some attempts
    # single-system solve, same network: bit-identical to the ensemble member
    one = oh.integrate("DOPRI5", xs[1], 0.0, 0.02, 1e-10, 1e-10,
                       oh.NBodyODE(1e-3, ms[1], "exact"), exe=exe, controller="learned",
                       mlp=net)
    assert (one.accepted, one.rejected) == (res.accepted[1], res.rejected[1])
    assert np.array_equal(one.x_final, res.states()[1])
    net.close()


def test_trained_controller_gain(exe):
    """The shipped DOP853 network (trained on the GPU from PI-controller
    traces of the ensemble workload, mixed-precision RHS) against the
    method without it, on held-out systems: fewer RHS evaluations to reach
    the same time, and every step still accepted only at error norm <= 1."""
    from ode_prototyping_amd import train_controller as trc
    from ode_prototyping_amd.ensemble import solve_ensemble
    net, info = trc.pretrained(exe, "DOP853")
    assert info["tableau"] == "DOP853" and info["samples"] > 100000
    sizes, x0, m = trc.workload_ensemble(256, 9100)
    runs = {}
    for ctrl in ("reference", "pi", "learned"):
        runs[ctrl] = solve_ensemble("DOP853", sizes, x0, 0.0, 0.05, 1e-10, 1e-10, 1e-3, m,
                                    variant="mixed", exe=exe, controller=ctrl,
                                    mlp=net if ctrl == "learned" else None, trace_cap=1 << 14)
    for ctrl, r in runs.items():
        print(ctrl, r.total_rhs_evals, r.total_accepted, r.total_rejected, r.rounds)
        assert r.n_unfinished == 0 and np.all(r.t == 0.05)
        for tr in r.trace:
            assert np.all(tr[tr[:, 2] == 1, 0] <= 1.0)
    gain = 1.0 - runs["learned"].total_rhs_evals / runs["reference"].total_rhs_evals
    assert gain > 0.05, gain
    assert runs["learned"].total_rhs_evals < runs["pi"].total_rhs_evals
    net.close()


def test_trained_controller_gain_fp64(exe):
    """The shipped DOP853 network for the fast FP64 RHS (trained on its PI-controller
    traces, target error chosen on a validation ensemble) against the reference
    controller on held-out systems: fewer RHS evaluations, every accepted step at error
    norm <= 1, and final states within 1e-8 of the reference controller's."""
    from ode_prototyping_amd import train_controller as trc
    from ode_prototyping_amd.ensemble import solve_ensemble
    net, info = trc.pretrained(exe, "DOP853", "fast")
    assert info["variant"] == "fast" and info["samples"] > 100000
    assert info["tau_selection"]["validation_seed"] != 9100
    sizes, x0, m = trc.workload_ensemble(256, 9100)
    runs = {}
    for ctrl in ("reference", "learned"):
        runs[ctrl] = solve_ensemble("DOP853", sizes, x0, 0.0, 0.05, 1e-10, 1e-10, 1e-3, m,
                                    variant="fast", exe=exe, controller=ctrl,
                                    mlp=net if ctrl == "learned" else None, trace_cap=1 << 14)
    for ctrl, r in runs.items():
        print(ctrl, r.total_rhs_evals, r.total_accepted, r.total_rejected, r.rounds)
        # every system reaches t = 0.05 (a first step longer than the interval is not
        # clipped, as in the reference: such a system ends past it)
        assert r.n_unfinished == 0 and np.all(r.t >= 0.05)
        for tr in r.trace:
            assert np.all(tr[tr[:, 2] == 1, 0] <= 1.0)
    gain = 1.0 - runs["learned"].total_rhs_evals / runs["reference"].total_rhs_evals
    assert gain > 0.05, gain
    same_t = np.flatnonzero(runs["learned"].t == runs["reference"].t)
    assert same_t.size >= 250
    a, b = runs["learned"].states(), runs["reference"].states()
    for k in same_t:
        assert np.max(np.abs(a[k] - b[k])) <= 1e-8 * np.max(np.abs(b[k])), k
    net.close()

"""DOP853 dense output in the CPU oracle (CPU only).

Pins of the continuous extension (the reference's
StepperDopr853::prepare_dense / dense_out, NumericalRecipes/
StepperDopr853.hpp:200-289, with the DenseOutput policy, Output.hpp:63-147):

* the reference's two consistency goldens
  (NumericalRecipes/test/StepperDopr853_test.cpp:168-252 and :396-480): the
  x values bit for bit, the y values to the reference's own EXPECT_DOUBLE_EQ
  (4 ulp) -- 77 of the 80 values are bit-identical, three differ by one ulp
  (at x = 0.5 and x = 0.9999999999999999 of the Van der Pol run and x = 0.5 of
  the oscillator run; no evaluation order of s or of the polynomial tried
  reproduces all 80, so the goldens were most likely produced by a build that
  rounded differently) -- and the final state bit for bit.  The reference
  compares with ranges::zip, which stops at the shorter sequence: the runs
  produce 20 points (0 .. 1.9000000000000006); the golden lists' 21st entry
  (x = 2.0) is never reached, because the policy only writes points strictly
  before the end of a step;
* the scipy comparison goldens of the same file (:79-166) to their 1e-4;
* the mathematics: the extension reproduces the step's start exactly at s = 0
  and its end at s = 1, and its error inside a step falls as h^8 (a 7th-order
  interpolant), which a wrong coefficient breaks; on the oscillator it tracks
  cos/sin to the tolerance;
* the RK-form integrator's dense output agrees with its own accepted states
  at the accepted times.
"""
import numpy as np
import pytest

import oracle_lib as O

GOLD_X = [0.0, 0.1, 0.2, 0.30000000000000004, 0.4, 0.5, 0.6, 0.7, 0.7999999999999999,
          0.8999999999999999, 0.9999999999999999, 1.0999999999999999, 1.2, 1.3,
          1.4000000000000001, 1.5000000000000002, 1.6000000000000003, 1.7000000000000004,
          1.8000000000000005, 1.9000000000000006, 2.0]

# StepperDopr853_test.cpp:197-238 (Van der Pol, eps = 1e-3, tol 1e-3)
VDP_Y0 = [2.0, 1.9316122783182343, 1.8584787214893352, 1.7797321406395954, 1.6936478062588005,
          1.5973048438124426, 1.4853961797009534, 1.3445081885185826, 1.1034118596940761,
          -1.9595852385244072, -1.8883998283360151, -1.8120698785806237, -1.7291573373244966,
          -1.6373650886237239, -1.53268730742399, -1.4060320034274552, -1.2275084884890473,
          1.9872073270417665, 1.9177895398084071, 1.8436813288440266, 1.7636585542718883]
VDP_Y1 = [0.0, -0.7094442261444629, -0.72022548954393439, -0.80686507047834188,
          -0.93894845532670956, -0.99943292259922056, -1.2246400286700401, -1.6518197427494905,
          -4.6193173417203504, 0.69181635473357528, 0.77341556725241589, 0.83049596472069132,
          0.88152147969282724, 0.94175345678023936, 1.17655985351827, 1.393853095827446,
          2.3214362560484116, -0.66890691862254825, -0.70612818420425061, -0.76466904126707647,
          -0.8354674509696991]
# :425-466 (oscillator, tol 1e-10)
OSC_Y0 = [1.0, 0.99500416528781477, 0.98006657793938101, 0.95533648916707048,
          0.92106099380834583, 0.87758256188944672, 0.82533561509979825, 0.7648421870658686,
          0.69670670935715528, 0.62160996830538506, 0.54030230597309425, 0.45359612124685283,
          0.36235775449376306, 0.26749882869694513, 0.16996714287483936, 0.070737201691768836,
          -0.029199522276831924, -0.12884449426873257, -0.22720209466067331,
          -0.32328956683362298, -0.41614683651801287]
OSC_Y1 = [0.0, -0.099833416648732465, -0.19866933080879839, -0.29552020666300693,
          -0.38941834227306915, -0.47942553859720011, -0.56464247349047925,
          -0.64421768710214034, -0.71735609089146357, -0.78332690964802465,
          -0.84147098490570094, -0.89120735982110444, -0.93203908596047802,
          -0.96355818555607042, -0.98544972984832069, -0.99749498659822255,
          -0.99957360305615339, -0.9916648106167375, -0.97384763057852497,
          -0.94630008769357532, -0.90929742683477]
# :108-149 (scipy solve_ivp values for the Van der Pol run)
SCIPY_Y0 = [2.0, 1.9316122783182301, 1.858478721489275, 1.7797321406396824, 1.6936478062587523,
            1.5973048438123825, 1.48539617970096, 1.344508188518583, 1.1034118596992013,
            -1.9595852384343717, -1.8883998282486796, -1.8120698785117855, -1.7291573370452291,
            -1.6373650883017334, -1.5326873072225298, -1.406032003621479, -1.2275084897783375,
            1.9872064158183738, 1.9177885756335038, 1.843680289685086, 1.7636574228798523]
SCIPY_Y1 = [0.0, -0.7094442261327178, -0.7202254893938206, -0.8068650706654498,
            -0.9389484552331121, -0.9994329225021683, -1.224640028673981, -1.6518197427464374,
            -4.619317342577298, 0.6918163442868651, 0.7734155795544163, 0.8304960343582835,
            0.8815211399286773, 0.9417531200490238, 1.1765597740458502, 1.3938534632445547,
            2.321437090654019, -0.6689116264185984, -0.7061440632836788, -0.7646734056087162,
            -0.8354687358368221]


def ulp_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def run(problem):
    if problem == "vanderpol":
        return O.nr_dop853_dense(1, [2.0, 0.0], 0.0, 2.0, 1e-3, 1e-3, 0.01, 20, mu_eps=1e-3)
    return O.nr_dop853_dense(0, [1.0, 0.0], 0.0, 2.0, 1e-10, 1e-10, 0.01, 20)


@pytest.mark.parametrize("problem,g0,g1,max_inexact", [
    ("vanderpol", VDP_Y0, VDP_Y1, 2), ("oscillator", OSC_Y0, OSC_Y1, 1)])
def test_nr_dense_consistency_goldens(problem, g0, g1, max_inexact):
    st, xs, ys, yf = run(problem)
    assert st == 0
    n = len(xs)
    assert n == 20  # the golden lists' last point (x = 2.0) is never written
    assert np.array_equal(xs, np.array(GOLD_X[:n]))
    d = np.concatenate([ulp_diff(ys[:, 0], g0[:n]), ulp_diff(ys[:, 1], g1[:n])])
    assert np.max(d) <= 4.0  # EXPECT_DOUBLE_EQ
    assert np.count_nonzero(d) <= max_inexact and np.max(d) <= 1.0
    assert yf[0] == g0[-1] and yf[1] == g1[-1]


def test_nr_dense_matches_scipy_goldens():
    st, xs, ys, yf = run("vanderpol")
    n = len(xs)
    assert np.max(np.abs(ys[:, 0] - SCIPY_Y0[:n])) < 1e-4
    assert np.max(np.abs(ys[:, 1] - SCIPY_Y1[:n])) < 1e-4
    assert abs(yf[0] - SCIPY_Y0[-1]) < 1e-4 and abs(yf[1] - SCIPY_Y1[-1]) < 1e-4


def test_oscillator_dense_tracks_closed_form():
    st, xs, ys, yf = run("oscillator")
    assert np.max(np.abs(ys[:, 0] - np.cos(xs))) < 1e-8
    assert np.max(np.abs(ys[:, 1] + np.sin(xs))) < 1e-8


def test_extension_endpoints_and_order():
    """One RK-form DOP853 step of y' = y: s = 0 returns the start exactly,
    s = 1 the end; inside the step the error against exp falls by ~2^8 per
    halving of h (local error O(h^8) of a 7th-order interpolant)."""
    y = np.array([1.0, 2.0, -0.5])
    s = np.array([0.0, 0.3, 0.5, 0.8, 1.0])
    errs = []
    for h in (0.4, 0.2):
        yout, d = O.dop853_step_dense(O.RHS_EXPONENTIAL, y, h, s)
        assert np.array_equal(d[0], y)
        assert np.max(ulp_diff(d[-1], yout)) <= 1.0
        errs.append(np.max(np.abs(d[1:4] - np.exp(np.outer(s[1:4], h)) * y), axis=1))
    ratio = errs[0] / errs[1]
    assert np.all(ratio > 180) and np.all(ratio < 360), ratio


def test_extension_order_vanderpol():
    """The same order on a nonlinear problem (Van der Pol, eps = 1) against a
    tight reference solve."""
    y = np.array([2.0, 0.0])
    s = np.array([0.5, 0.8])
    errs = []
    for h in (0.1, 0.05):
        _, d = O.dop853_step_dense(O.RHS_VANDERPOL, y, h, s)
        ref = np.array([O.integrate("DOP853", O.RHS_VANDERPOL, y, 0.0, sk * h, 1e-14, 1e-14,
                                    store_states=False).x_final for sk in s])
        errs.append(np.max(np.abs(d - ref), axis=1))
    ratio = errs[0] / errs[1]
    assert np.all(ratio > 150) and np.all(ratio < 400), ratio


def test_rk_form_dense_output_agrees_with_accepted_states():
    """Dense times placed on the accepted step times are written by the next
    step at s ~ 0: they reproduce the accepted states to rounding; and a
    solve with dense output takes exactly the steps of one without it."""
    x0 = np.array([2.0, 0.0])
    plain = O.integrate("DOP853", O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-8, 1e-8)
    inner = plain.times[1:-1]
    r = O.integrate("DOP853", O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-8, 1e-8, dense_times=inner)
    assert np.array_equal(r.times, plain.times) and np.array_equal(r.x_final, plain.x_final)
    assert r.n_dense == len(inner) and r.dense_steps == len(inner)
    assert r.dense_rhs_evals == 4 * len(inner) and r.rhs_evals == plain.rhs_evals
    states = np.array([plain.state(k) for k in range(1, len(plain.times) - 1)])
    assert np.max(np.abs(r.dense_states - states)) < 1e-12 * np.max(np.abs(states))


def test_rk_form_dense_needs_dop853():
    with pytest.raises(RuntimeError):
        O.integrate("DOPRI5", O.RHS_VANDERPOL, [2.0, 0.0], 0.0, 1.0, 1e-8, 1e-8,
                    dense_times=[0.5])

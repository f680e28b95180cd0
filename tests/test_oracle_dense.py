"""DOP853 and DOPRI5 dense output in the CPU oracle (CPU only).

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

DOPRI5 (StepperDopr5::prepare_dense / dense_out, StepperDopr5.hpp:181-221,
no extra stages): the NR DOPRI5 restatement reproduces both raw-output
goldens bit for bit and the nonzero-start dense golden bit for bit; the two
other dense goldens within their EXPECT_DOUBLE_EQ (see the test); the
extension is 4th order (error ~h^5 inside a step).
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


def test_rk_form_dense_needs_dop853_or_dopri5():
    with pytest.raises(RuntimeError):
        O.integrate("RKF78", O.RHS_VANDERPOL, [2.0, 0.0], 0.0, 1.0, 1e-8, 1e-8,
                    dense_times=[0.5])


# ---------------------------------------------------------------- DOPRI5 --
# NumericalRecipes/test/SimpleOscillatorTest.cpp and VanDerPolTest.cpp
# (StepperDopr5 with RawOutput and DenseOutput).
D5_OSC_Y0 = [1.0, 0.995004165274381, 0.9800665778342761, 0.9553364891156162, 0.9210609939762899,
             0.8775825618689674, 0.8253356148906368, 0.7648421872376728, 0.6967067092917164,
             0.6216099682354068, 0.540302305843194, 0.45359612136859212, 0.36235775441965129,
             0.2674988286034845, 0.1699671428597921, 0.070737201637418756, -0.029199522309753027,
             -0.12884449432018022, -0.22720209468205022, -0.32328956688476052,
             -0.41614683651997625]
D5_OSC_Y1 = [0.0, -0.09983341667072441, -0.19866933081341775, -0.29552020666230877,
             -0.38941834234718053, -0.47942553861305215, -0.5646424733839779, -0.6442176872563701,
             -0.7173560909168458, -0.7833269096128057, -0.8414709847747586, -0.8912073600416666,
             -0.9320390859400902, -0.9635581853673072, -0.9854497299379401, -0.9974949865471938,
             -0.9995736029766767, -0.99166481038394072, -0.97384763080764247,
             -0.94630008760441731, -0.9092974267525671]
D5_VDP_Y0 = [2.0, 1.9316114668840196, 1.8584936578161697, 1.779738838649533, 1.6936304673029232,
             1.5973233455330451, 1.4853988650659193, 1.3445145377549466, 1.1027432107238968,
             -1.9607574926907294, -1.8896363707127886, -1.8134015921664299, -1.7306276596111723,
             -1.6390360829064694, -1.5345832043822263, -1.4085081362013523, -1.231625033046776,
             1.9878333640723733, 1.9184567236794956, 1.844394686471839, 1.7644320190605265]
D5_VDP_Y1 = [0.0, -0.70723585006735068, -0.75690138151347441, -0.82140897529267398,
             -0.90658463142947066, -1.0280880445966889, -1.2278827867432773, -1.6569111260367424,
             -4.4853706789511456, 0.68962769177543404, 0.73633857873973074, 0.79157138514815228,
             0.86880585055426407, 0.96881359325147665, 1.1314276223440602, 1.428107104729363,
             2.3390319866124534, -0.67412345520940009, -0.71620744619303756, -0.7682889892922099,
             -0.83427005245677999]


def run5(problem, intervals=0, x1=0.0):
    if problem == "vanderpol":
        return O.nr_dopr5(1, [2.0, 0.0], x1, 2.0, 1e-3, 1e-3, 0.01, intervals, mu_eps=1e-3)
    return O.nr_dopr5(0, [1.0, 0.0], x1, 2.0, 1e-10, 1e-10, 0.01, intervals)


@pytest.mark.parametrize("problem,count,x_mid,y_mid,y_last", [
    ("oscillator", 49, 1.0004288786936162, (0.53994136718681807, -0.84170263152593661),
     (-0.41614683651997603, -0.90929742675256664)),
    ("vanderpol", 1177, 0.98436957642198519, (-1.9010597230846302, 0.72447110662737324),
     (1.7644320190605265, -0.83427005245677999))])
def test_nr_dopr5_raw_goldens(problem, count, x_mid, y_mid, y_last):
    """SimpleOscillatorTest.cpp:40-63, VanDerPolTest.cpp:38-76: the step
    sequence of the NR DOPRI5 stepper, bit for bit."""
    st, xs, ys, yf = run5(problem)
    assert st == 0 and len(xs) == count
    m = count // 2
    assert xs[0] == 0.0 and xs[m] == x_mid and xs[-1] == 2.0
    assert tuple(ys[m]) == y_mid and tuple(ys[-1]) == y_last and tuple(yf) == y_last


@pytest.mark.parametrize("problem,g0,g1", [("oscillator", D5_OSC_Y0, D5_OSC_Y1),
                                           ("vanderpol", D5_VDP_Y0, D5_VDP_Y1)])
def test_nr_dopr5_dense_consistency_goldens(problem, g0, g1):
    """SimpleOscillatorTest.cpp:154-240, VanDerPolTest.cpp:167-250, to the
    goldens' EXPECT_DOUBLE_EQ: these runs carry a slightly different
    trajectory than the raw goldens of the same files (the oscillator's final
    state differs from the raw run's by 4 ulp there, while the restatement's
    equals the raw golden bit for bit), so they are met within 4 ulp, not
    exactly."""
    st, xs, ys, yf = run5(problem, 20)
    assert st == 0 and len(xs) == 20
    assert np.array_equal(xs, np.array(GOLD_X[:20]))
    d = np.concatenate([ulp_diff(ys[:, 0], g0[:20]), ulp_diff(ys[:, 1], g1[:20]),
                        ulp_diff(yf, [g0[-1], g1[-1]])])
    assert np.max(d) <= 4.0


def test_nr_dopr5_dense_nonzero_start_golden():
    """SimpleOscillatorTest.cpp:242-280 (start at x = 1, five intervals): every
    y value and the final state bit for bit; the last x is 2 - 2^-52 (the
    golden writes it as 2.0, within its EXPECT_DOUBLE_EQ)."""
    st, xs, ys, yf = run5("oscillator", 5, x1=1.0)
    gx = [1.0, 1.2, 1.3999999999999999, 1.5999999999999999, 1.7999999999999998, 2.0]
    g0 = [1.0, 0.9800665778342762, 0.92106099397628982, 0.82533561489063678, 0.6967067092917163,
          0.54030230584326977]
    g1 = [0.0, -0.19866933081341756, -0.38941834234718037, -0.56464247338397766,
          -0.71735609091684549, -0.8414709847748435]
    assert st == 0 and len(xs) == 6
    assert np.array_equal(xs[:5], gx[:5]) and xs[5] == np.nextafter(2.0, 0.0)
    assert np.array_equal(ys[:, 0], g0) and np.array_equal(ys[:, 1], g1)
    assert yf[0] == 0.54030230584326955 and yf[1] == -0.84147098477484361


def test_dopri5_extension_endpoints_and_order():
    """DOPRI5's continuous extension is 4th order: inside a step of y' = y its
    error falls by ~2^5 per halving of h; s = 0 and s = 1 reproduce the
    step's ends."""
    y = np.array([1.0, 2.0, -0.5])
    s = np.array([0.0, 0.3, 0.5, 0.8, 1.0])
    errs = []
    for h in (0.2, 0.1):
        yout, d = O.step_dense("DOPRI5", O.RHS_EXPONENTIAL, y, h, s)
        assert np.array_equal(d[0], y)
        assert np.max(ulp_diff(d[-1], yout)) <= 1.0
        errs.append(np.max(np.abs(d[1:4] - np.exp(np.outer(s[1:4], h)) * y), axis=1))
    ratio = errs[0] / errs[1]
    assert np.all(ratio > 25) and np.all(ratio < 40), ratio


def test_rk_form_dopri5_dense_output_agrees_with_accepted_states():
    x0 = np.array([2.0, 0.0])
    plain = O.integrate("DOPRI5", O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-8, 1e-8)
    inner = plain.times[1:-1]
    r = O.integrate("DOPRI5", O.RHS_VANDERPOL, x0, 0.0, 2.0, 1e-8, 1e-8, dense_times=inner)
    assert np.array_equal(r.times, plain.times) and np.array_equal(r.x_final, plain.x_final)
    assert r.n_dense == len(inner) and r.dense_rhs_evals == 0
    states = np.array([plain.state(k) for k in range(1, len(plain.times) - 1)])
    assert np.max(np.abs(r.dense_states - states)) < 1e-12 * np.max(np.abs(states))

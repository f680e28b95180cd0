"""Batched ensemble solve and right-hand side against the CPU oracle.

Every system of an ensemble must come out bit for bit as the oracle (HIP
reduction order, correctly rounded pow) integrates it alone: final state,
counters, per-attempt trace; and as the single-system device solve does.
The ensemble right-hand side in the exact variant must equal the oracle's
(the reference's single-threaded pair loop) bit for bit in every tile
configuration; the fast and mixed variants hold the documented tolerances.
Sizes cover 1, 2 and 3 bodies, sizes around the 64- and 128-target tile
edges and the partner-group edges of the kernels.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 16, 37, 63, 64, 65, 100, 127, 128, 129, 200, 257]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def make_ensemble(sizes, seed0=0):
    import ode_prototyping_amd as oh
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = oh.scenarios.uniform_cube(int(n), seed=seed0 + k)
        xs.append(x)
        ms.append(m)
    return xs, ms


@pytest.fixture(scope="module")
def ens():
    return make_ensemble(SIZES)


@pytest.mark.parametrize("cfg", list(range(18)))
@pytest.mark.parametrize("with_mass", [True, False])
def test_ensemble_rhs_exact_bitwise(exe, ens, cfg, with_mass):
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    xs, ms = ens
    masses = pack(ms) if with_mass else None
    d, ms_ = ensemble_rhs(exe, SIZES, pack(xs), 1e-3, masses, "exact", tile_config=cfg)
    for k, (x, part) in enumerate(zip(xs, unpack(d, SIZES))):
        ref = O.rhs(O.RHS_NBODY, x, 1e-3, ms[k] if with_mass else None)
        assert np.array_equal(bits(part), bits(ref)), (cfg, SIZES[k])
    assert ms_ > 0.0


@pytest.mark.parametrize("cfg", [-1, 1, 6, 11, 12, 15, 17])
def test_ensemble_rhs_exact_bitwise_many_systems(exe, cfg):
    """600 systems of 16 to 300 bodies (about 1700 tiles): more workgroups
    than the 256 CUs hold at once, XCD placement on and off, 128-target tiles
    for the systems of at least 256 bodies in configuration 11, a tile shared
    by 2, 4 and 8 waves in configurations 12, 15 and 17."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    rng = np.random.default_rng(600)
    sizes = [int(v) for v in rng.integers(16, 301, size=600)]
    xs, ms = make_ensemble(sizes, seed0=1000)
    d, _ = ensemble_rhs(exe, sizes, pack(xs), 1e-3, pack(ms), "exact", tile_config=cfg)
    for k, part in enumerate(unpack(d, sizes)):
        ref = O.rhs(O.RHS_NBODY, xs[k], 1e-3, ms[k])
        assert np.array_equal(bits(part), bits(ref)), (cfg, k, sizes[k])


@pytest.mark.parametrize("cfg", [7, 13, 15, 17])
def test_ensemble_rhs_exact_bitwise_large_shared_tiles(exe, cfg):
    """Systems of 513 to 4085 bodies -- the sizes whose tiles the tail rule shares
    among 8 waves in a solve -- bit for bit as the oracle, one wave per tile and 2, 4 or
    8 waves per tile (blocks of 8 to 32 partners, several blocks per tile, the own
    targets' block in the middle of the range)."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    sizes = [513, 1024, 2047, 4085]
    xs, ms = make_ensemble(sizes, seed0=700)
    d, _ = ensemble_rhs(exe, sizes, pack(xs), 1e-3, pack(ms), "exact", tile_config=cfg)
    for k, part in enumerate(unpack(d, sizes)):
        ref = O.rhs(O.RHS_NBODY, xs[k], 1e-3, ms[k], threads=8)
        assert np.array_equal(bits(part), bits(ref)), (cfg, sizes[k])


@pytest.mark.parametrize("cfg", [2, 12, 14, 16])
def test_ensemble_rhs_exact_zero_softening(exe, cfg):
    """eps = 0: the IEEE term path (no shared reciprocal) and the skipped
    diagonal (0/0 would be NaN), one wave per tile and shared tiles."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    sizes = [7, 70, 130]
    xs, ms = make_ensemble(sizes, seed0=50)
    d, _ = ensemble_rhs(exe, sizes, pack(xs), 0.0, pack(ms), "exact", tile_config=cfg)
    for k, part in enumerate(unpack(d, sizes)):
        ref = O.rhs(O.RHS_NBODY, xs[k], 0.0, ms[k])
        assert np.array_equal(bits(part), bits(ref)), (cfg, sizes[k])


@pytest.mark.parametrize("variant,tol", [("fast", 2e-14), ("mixed", 5e-6)])
@pytest.mark.parametrize("cfg", [0, 5, 10])
def test_ensemble_rhs_fast_mixed(exe, ens, variant, tol, cfg):
    """Relative error of every acceleration against the exact oracle, scaled
    by the sum of the absolute pair terms (the conditioning of the sum)."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    xs, ms = ens
    d, _ = ensemble_rhs(exe, SIZES, pack(xs), 1e-3, pack(ms), variant, tile_config=cfg)
    for k, part in enumerate(unpack(d, SIZES)):
        n = SIZES[k]
        ref = O.rhs(O.RHS_NBODY, xs[k], 1e-3, ms[k])
        assert np.array_equal(part[:3 * n], ref[:3 * n])  # velocities copied
        pos = xs[k][:3 * n].reshape(n, 3)
        r = pos[None, :, :] - pos[:, None, :]
        w = ms[k][None, :] / (np.sum(r * r, axis=2) + 1e-6) ** 1.5
        scale = np.sum(np.abs(r) * w[:, :, None], axis=1).ravel() + 1e-300
        err = np.abs(part[3 * n:] - ref[3 * n:]) / scale
        assert err.max() < tol, (variant, n, err.max())


def mass_forms(sizes, seed):
    """Per system, in turn: equal masses 1 / n, equal 0.37 / n (not a power
    of two), masses spread over [0.5, 2] / n, and 1 / n but for the last
    body (1.5 / n)."""
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        if k % 4 == 0:
            m = np.full(n, 1.0 / n)
        elif k % 4 == 1:
            m = np.full(n, 0.37 / n)
        elif k % 4 == 2:
            m = rng.uniform(0.5, 2.0, n) / n
        else:
            m = np.full(n, 1.0 / n)
            m[-1] = 1.5 / n
        out.append(m)
    return out


@pytest.mark.parametrize("variant,tol", [("fast", 2e-14), ("mixed", 5e-6)])
@pytest.mark.parametrize("cfg", [0, 7, 11, 12, 17])
def test_ensemble_rhs_fast_mixed_mass_forms(exe, variant, tol, cfg):
    """Systems with equal masses (unit-mass terms, one scale of each finished
    sum) beside systems with spread masses and with one differing mass
    (per-term products), in one-wave, 128-target and shared tiles: every
    acceleration within the variant's bar of the oracle, scaled as in
    test_ensemble_rhs_fast_mixed."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    sizes = SPLIT_SIZES
    xs, _ = make_ensemble(sizes, seed0=300)
    ms = mass_forms(sizes, 300)
    d, _ = ensemble_rhs(exe, sizes, pack(xs), 1e-3, pack(ms), variant, tile_config=cfg)
    for k, part in enumerate(unpack(d, sizes)):
        n = sizes[k]
        ref = O.rhs(O.RHS_NBODY, xs[k], 1e-3, ms[k])
        assert np.array_equal(part[:3 * n], ref[:3 * n])
        pos = xs[k][:3 * n].reshape(n, 3)
        r = pos[None, :, :] - pos[:, None, :]
        w = ms[k][None, :] / (np.sum(r * r, axis=2) + 1e-6) ** 1.5
        scale = np.sum(np.abs(r) * w[:, :, None], axis=1).ravel() + 1e-300
        err = np.abs(part[3 * n:] - ref[3 * n:]) / scale
        assert err.max() < tol, (variant, cfg, n, k % 4, err.max())


@pytest.mark.parametrize("variant,bar", [("fast", 1e-9), ("mixed", 1e-6)])
def test_ensemble_solve_fast_mixed_mass_forms(exe, variant, bar):
    """An ensemble solve records at its start whether each system's masses are
    all equal, and its right-hand sides take the equal-mass form from that
    record: systems of every mass form follow their exact oracle solve to
    the variant's accuracy (a misweighed body would move the state by about
    1e-4 of its size by t = 0.01)."""
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [40, 41, 42, 43, 300, 301, 302, 303]
    xs, _ = make_ensemble(sizes, seed0=400)
    ms = mass_forms(sizes, 400)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.01, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant=variant, exe=exe)
    assert res.n_unfinished == 0
    finals = res.states()
    for k in range(len(sizes)):
        ref = oracle_solve("DOPRI5", xs[k], 0.01, 1e-10, 1e-3, ms[k])
        err = np.max(np.abs(finals[k] - ref.x_final)) / np.max(np.abs(ref.x_final))
        assert err < bar, (variant, sizes[k], k % 4, err)


def oracle_solve(tableau, x, tf, tol, eps, m, **kw):
    return O.integrate(tableau, O.RHS_NBODY, x, 0.0, tf, tol, tol, softening=eps, masses=m,
                       reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT, store_states=False,
                       store_trace=True, threads=8, **kw)


def assert_system_matches(res, k, ref, x_final, trace=True):
    assert (res.accepted[k], res.rejected[k], res.rhs_evals[k]) == (
        ref.accepted, ref.rejected, ref.rhs_evals), k
    assert res.t[k] == ref.t and res.dt[k] == ref.dt and res.dt_initial[k] == ref.dt_initial
    assert np.array_equal(bits(x_final), bits(ref.x_final)), k
    if trace:
        tr = res.trace[k]
        assert np.array_equal(bits(tr[:, 0]), bits(ref.trace_err))
        assert np.array_equal(bits(tr[:, 1]), bits(ref.trace_dt))
        assert np.array_equal(tr[:, 2].astype(int), ref.trace_acc)


# tile configurations: auto; 64-target tiles; 128-target tiles from 1024 bodies
# (none here); 128-target tiles from 256 bodies (the 257-body system), which
# compaction splits into 64-target ones once systems finish
@pytest.mark.parametrize("tableau,tf,cfg", [("DOPRI5", 0.05, -1), ("DOP853", 0.03, 7),
                                            ("RKF45", 0.02, 3), ("DOPRI5", 0.05, 11)])
def test_ensemble_solve_matches_oracle(exe, ens, tableau, tf, cfg):
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    xs, ms = ens
    res = solve_ensemble(tableau, SIZES, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant="exact", exe=exe, tile_config=cfg, trace_cap=4096)
    assert res.n_unfinished == 0
    finals = res.states()
    for k, n in enumerate(SIZES):
        ref = oracle_solve(tableau, xs[k], tf, 1e-10, 1e-3, ms[k])
        assert_system_matches(res, k, ref, finals[k])
    assert res.total_accepted == res.accepted.sum()
    assert res.particle_steps == float(np.sum(np.asarray(SIZES) * res.accepted))
    assert res.rounds == res.attempts.max()


def test_ensemble_solve_many_systems_with_compaction(exe):
    """120 systems of 16 to 160 bodies: they finish at different rounds, so
    the running set is compacted (tables rebuilt, graph re-captured) several
    times; every system still matches its own oracle solve bit for bit."""
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    rng = np.random.default_rng(120)
    sizes = [int(v) for v in rng.integers(16, 161, size=120)]
    xs, ms = make_ensemble(sizes, seed0=2000)
    # two rounds per graph replay: the host sees the running count often
    # enough to compact at every halving
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.1, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant="exact", exe=exe, trace_cap=4096, rounds_per_launch=2)
    assert res.n_unfinished == 0
    # finishing times spread enough for the running set to halve repeatedly
    # (attempts range from about 5 to 300)
    a = np.sort(res.attempts)
    assert a[len(a) // 2] < a[-1] // 8 and a[3 * len(a) // 4] < a[-1] // 4
    finals = res.states()
    for k in range(len(sizes)):
        ref = oracle_solve("DOPRI5", xs[k], 0.1, 1e-10, 1e-3, ms[k])
        assert_system_matches(res, k, ref, finals[k])


def test_ensemble_dense_output_across_compaction(exe):
    """DOPRI5 dense rows of the 120-system compaction case at 25 common
    times: every system's rows equal its oracle solve's rows bit for bit,
    written before and after the running set is rebuilt."""
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    rng = np.random.default_rng(120)
    sizes = [int(v) for v in rng.integers(16, 161, size=120)]
    xs, ms = make_ensemble(sizes, seed0=2000)
    times = np.linspace(0.0, 0.1, 26)[1:-1]
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.1, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant="exact", exe=exe, rounds_per_launch=2, dense_times=times)
    assert res.n_unfinished == 0 and np.all(res.n_dense == len(times))
    for k in range(len(sizes)):
        ref = oracle_solve("DOPRI5", xs[k], 0.1, 1e-10, 1e-3, ms[k], dense_times=times)
        assert np.array_equal(bits(res.dense_rows(k)), bits(ref.dense_states)), k


def test_ensemble_matches_single_solves_fsal(exe, ens):
    """FSAL on, max_accepted budget: each system as odehip_rk_solve alone."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    xs, ms = ens
    res = solve_ensemble("DOPRI5", SIZES, pack(xs), 0.0, 1.0, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant="exact", exe=exe, fsal=True, max_accepted=25,
                         rounds_per_launch=3)
    finals = res.states()
    for k, n in enumerate(SIZES):
        one = oh.integrate("DOPRI5", xs[k], 0.0, 1.0, 1e-10, 1e-10,
                           oh.NBodyODE(1e-3, ms[k], "exact"), exe=exe, fsal=True,
                           max_accepted=25)
        assert (res.accepted[k], res.rejected[k], res.rhs_evals[k], res.status_names[k]) == (
            one.accepted, one.rejected, one.rhs_evals, one.status)
        assert res.t[k] == one.t and res.dt[k] == one.dt
        assert np.array_equal(bits(finals[k]), bits(one.x_final))
    assert set(res.status_names) <= {"max_accepted", "reached_end"}
    assert np.all(res.accepted <= 25)


@pytest.mark.parametrize("cfg", [0, 5, 10])
def test_ensemble_rhs_mixed_all_near_is_fp64(exe, ens, cfg):
    """A near field covering every pair turns the ensemble's mixed RHS into
    FP64 terms (the fast variant's bar)."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, unpack
    xs, ms = ens
    exe.near_field = 1e30
    try:
        d, _ = ensemble_rhs(exe, SIZES, pack(xs), 1e-3, pack(ms), "mixed", tile_config=cfg)
    finally:
        exe.near_field = oh.MIXED_NEAR_DEFAULT
    for k, part in enumerate(unpack(d, SIZES)):
        n = SIZES[k]
        ref = O.rhs(O.RHS_NBODY, xs[k], 1e-3, ms[k])
        pos = xs[k][:3 * n].reshape(n, 3)
        r = pos[None, :, :] - pos[:, None, :]
        w = ms[k][None, :] / (np.sum(r * r, axis=2) + 1e-6) ** 1.5
        scale = np.sum(np.abs(r) * w[:, :, None], axis=1).ravel() + 1e-300
        err = np.abs(part[3 * n:] - ref[3 * n:]) / scale
        assert err.max() < 2e-14, (n, err.max())


SPLIT_SIZES = [1, 3, 64, 127, 128, 129, 255, 256, 257, 300, 1023, 1024, 1025, 2000]


@pytest.mark.parametrize("variant", ["fast", "mixed", "mixed_near"])
def test_ensemble_rhs_split_tiles_bitwise_unsplit(exe, variant):
    """The fast and mixed sums run over partner chunks of 128 in chunk order, so one
    tile's chunks shared by 2, 4 or 8 waves of a workgroup (configurations 12-17, XCD
    placement off and on) give the same bits as one wave per tile, with 64- or
    128-target tiles (configurations 0-11).  Systems of 1 to 16 chunks, with equal masses,
    unit masses and the mixed mass forms of mass_forms, and (mixed_near) a near field
    covering every pair."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import (ENS_TILE_CONFIGS, ENS_TILE_UNSPLIT, ensemble_rhs,
                                              pack)
    xs, ms = make_ensemble(SPLIT_SIZES, seed0=90)
    var = "fast" if variant == "fast" else "mixed"
    if variant == "mixed_near":
        exe.near_field = 1e30
    try:
        for masses in (pack(ms), None, pack(mass_forms(SPLIT_SIZES, 91))):
            base, _ = ensemble_rhs(exe, SPLIT_SIZES, pack(xs), 1e-3, masses, var, tile_config=6)
            assert np.all(np.isfinite(base))
            for cfg in list(range(ENS_TILE_UNSPLIT)) + list(range(ENS_TILE_UNSPLIT,
                                                                  ENS_TILE_CONFIGS)):
                d, _ = ensemble_rhs(exe, SPLIT_SIZES, pack(xs), 1e-3, masses, var,
                                    tile_config=cfg)
                assert np.array_equal(bits(d), bits(base)), (variant, cfg, masses is None)
    finally:
        exe.near_field = oh.MIXED_NEAR_DEFAULT


@pytest.mark.parametrize("variant", ["exact", "fast", "mixed"])
def test_ensemble_solve_independent_of_split(exe, variant):
    """128 systems of 2048 bodies (4096 tiles: one wave per tile at the start; as
    systems finish, compaction moves the rest to tiles shared by several waves) give
    every system the same bits as its solve alone (eight waves a tile from the start):
    the split changes no result, whatever the other systems do.  (Exact: both also
    equal the oracle, which the smaller ensemble tests check.)"""
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [2048] * 128
    xs, ms = make_ensemble(sizes, seed0=400)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 2e-3, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant=variant, exe=exe, trace_cap=512)
    assert len(set(res.attempts.tolist())) > 1  # systems finish in different rounds
    finals = res.states()
    for k in (0, 1, 37, 64, 100, 127):
        one = solve_ensemble("DOPRI5", [2048], xs[k], 0.0, 2e-3, 1e-10, 1e-10, 1e-3, ms[k],
                             variant=variant, exe=exe, trace_cap=512)
        assert (one.accepted[0], one.rejected[0], one.rhs_evals[0]) == (
            res.accepted[k], res.rejected[k], res.rhs_evals[k]), k
        assert np.array_equal(bits(one.states()[0]), bits(finals[k])), k
        assert np.array_equal(bits(one.trace[0]), bits(res.trace[k])), k


@pytest.mark.parametrize("cfg", [0, 1, 7])
def test_ensemble_rhs_after_solve_and_solve_after_rhs(exe, ens, cfg):
    """A solve that never compacts (two rounds) and then a right-hand side of the same
    systems in the same tile configuration, and the reverse order: each call reads the
    tile list of its own configuration (the device tables are reused only when they sit
    at the same arena offsets), and both results equal the oracle bit for bit."""
    from ode_prototyping_amd.ensemble import ensemble_rhs, pack, solve_ensemble, unpack
    xs, ms = ens
    refs = [O.rhs(O.RHS_NBODY, x, 1e-3, m) for x, m in zip(xs, ms)]

    def rhs_ok():
        d, _ = ensemble_rhs(exe, SIZES, pack(xs), 1e-3, pack(ms), "exact", tile_config=cfg)
        for k, part in enumerate(unpack(d, SIZES)):
            assert np.array_equal(bits(part), bits(refs[k])), (cfg, SIZES[k])

    def solve_ok():
        res = solve_ensemble("DOPRI5", SIZES, pack(xs), 0.0, 1.0, 1e-10, 1e-10, 1e-3,
                             pack(ms), variant="exact", exe=exe, tile_config=cfg,
                             max_attempts=2, rounds_per_launch=2, trace_cap=8)
        finals = res.states()
        for k in range(len(SIZES)):
            ref = oracle_solve("DOPRI5", xs[k], 1.0, 1e-10, 1e-3, ms[k], max_attempts=2)
            assert np.array_equal(bits(finals[k]), bits(ref.x_final)), (cfg, SIZES[k])

    solve_ok()
    rhs_ok()
    solve_ok()
    rhs_ok()


def test_ensemble_mixed_near_field_through_close_encounters(exe):
    """The close-encounter systems of the single-system near-field test (uniform cubes of
    28 and 52 bodies, seed n, eps 1e-3, DOPRI5, t in [0, 0.1]) solved together as one
    ensemble with the mixed RHS: with the close-encounter near field every system takes
    the FP64 oracle's number of RHS evaluations and ends within 5e-9 of its final state;
    with every pair in FP32 at least one system needs more evaluations, every system ends
    >= 30x further off and at least one >= 100x."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [28, 52]
    xs, ms = [], []
    for n in sizes:
        x, m = oh.scenarios.uniform_cube(n, seed=n)
        xs.append(x)
        ms.append(m)
    refs = [oracle_solve("DOPRI5", xs[k], 0.1, 1e-10, 1e-3, ms[k]) for k in range(2)]
    runs = {}
    for c in (oh.MIXED_NEAR_CLOSE, 0.0):
        exe.near_field = c
        try:
            res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.1, 1e-10, 1e-10, 1e-3,
                                 pack(ms), variant="mixed", exe=exe)
        finally:
            exe.near_field = oh.MIXED_NEAR_DEFAULT
        assert np.all(res.t == 0.1)
        finals = res.states()
        runs[c] = [(np.max(np.abs(finals[k] - refs[k].x_final)) /
                    np.max(np.abs(refs[k].x_final)), int(res.rhs_evals[k])) for k in range(2)]
        print(f"near field {c}: {runs[c]} (oracle {[r.rhs_evals for r in refs]})")
    for k in range(2):
        assert runs[oh.MIXED_NEAR_CLOSE][k][0] <= 5e-9
        assert runs[oh.MIXED_NEAR_CLOSE][k][1] == refs[k].rhs_evals
        assert runs[0.0][k][0] >= 30 * runs[oh.MIXED_NEAR_CLOSE][k][0]
    assert any(runs[0.0][k][0] >= 100 * runs[oh.MIXED_NEAR_CLOSE][k][0] for k in range(2))
    assert any(runs[0.0][k][1] > refs[k].rhs_evals for k in range(2))


@pytest.mark.parametrize("variant", ["fast", "mixed"])
def test_ensemble_matches_single_solves_variants(exe, variant):
    """The fast and mixed right-hand sides of the ensemble sum partners in
    increasing j per target like the single-system kernels' waves, but the
    single-system kernels split j over 16 waves: compare at fixed time."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [64, 200, 333]
    xs, ms = make_ensemble(sizes, seed0=7)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.05, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant=variant, exe=exe)
    tol = 1e-11 if variant == "fast" else 1e-6
    for k, n in enumerate(sizes):
        ref = oracle_solve("DOPRI5", xs[k], 0.05, 1e-10, 1e-3, ms[k])
        assert res.t[k] == 0.05
        err = np.abs(res.states()[k] - ref.x_final).max() / np.abs(ref.x_final).max()
        assert err < tol, (variant, n, err)


def test_ensemble_shared_tiles_dense_output_and_k1_reuse(exe):
    """Systems of up to 1500 bodies, where a solve shares each tile among 8 waves from
    the start: the shared tiles' first-same-as-last k1 copy and DOP853 dense-output
    evaluations.  Exact: every system's dense rows, final state and counters bit for bit
    as its single-system solve.  Fast: with and without k1 reuse the same bits (the chunk
    order makes the copied k1 equal to a recomputed one), with fewer RHS evaluations."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [20, 90, 300, 1000, 1500]
    xs, ms = make_ensemble(sizes, seed0=80)
    tf = 0.01
    times = np.linspace(0.0, tf, 6)[1:-1]
    res = solve_ensemble("DOP853", sizes, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe, dense_times=times, fsal=True)
    for k in range(len(sizes)):
        one = oh.integrate("DOP853", xs[k], 0.0, tf, 1e-10, 1e-10,
                           oh.NBodyODE(1e-3, ms[k], "exact"), exe=exe, dense_times=times,
                           fsal=True)
        assert (res.accepted[k], res.rejected[k], res.rhs_evals[k]) == (
            one.accepted, one.rejected, one.rhs_evals), k
        assert np.array_equal(bits(res.dense_rows(k)), bits(one.dense_states)), k
        assert np.array_equal(bits(res.states()[k]), bits(one.x_final)), k
    runs = [solve_ensemble("DOP853", sizes, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                           variant="fast", exe=exe, dense_times=times, fsal=f)
            for f in (False, True)]
    assert np.array_equal(runs[0].accepted, runs[1].accepted)
    assert np.array_equal(runs[0].rejected, runs[1].rejected)
    assert np.all(runs[1].rhs_evals <= runs[0].rhs_evals)  # a one-step system reuses none
    assert runs[1].total_rhs_evals < runs[0].total_rhs_evals
    assert np.array_equal(bits(runs[0].x_final), bits(runs[1].x_final))
    assert np.array_equal(bits(runs[0].dense_states), bits(runs[1].dense_states))


def test_ensemble_pi_controller_matches_oracle(exe):
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [16, 65, 150]
    xs, ms = make_ensemble(sizes, seed0=21)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 0.05, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe, controller="pi", trace_cap=4096)
    finals = res.states()
    for k in range(len(sizes)):
        ref = oracle_solve("DOPRI5", xs[k], 0.05, 1e-10, 1e-3, ms[k],
                           controller=O.CONTROLLER_PI)
        assert_system_matches(res, k, ref, finals[k])


def test_single_solve_pi_controller_matches_oracle(exe):
    import ode_prototyping_amd as oh
    x, m = oh.scenarios.plummer(300, seed=3)
    res = oh.integrate("DOP853", x, 0.0, 0.5, 1e-10, 1e-10, oh.NBodyODE(1e-2, m, "exact"),
                       exe=exe, controller="pi", trace_cap=4096)
    ref = O.integrate("DOP853", O.RHS_NBODY, x, 0.0, 0.5, 1e-10, 1e-10, softening=1e-2,
                      masses=m, reduce=O.HIP_ORDER, pow_mode=O.POW_CORRECT,
                      store_states=False, store_trace=True, threads=8,
                      controller=O.CONTROLLER_PI)
    assert (res.accepted, res.rejected, res.rhs_evals) == (ref.accepted, ref.rejected,
                                                           ref.rhs_evals)
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))
    assert np.array_equal(bits(res.trace[:, 1]), bits(ref.trace_dt))


def test_ensemble_edge_cases(exe):
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    # one system; and t0 == tf (nothing to do: the state comes back as given)
    xs, ms = make_ensemble([40], seed0=99)
    res = solve_ensemble("DOPRI5", [40], pack(xs), 0.0, 0.01, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe)
    ref = oracle_solve("DOPRI5", xs[0], 0.01, 1e-10, 1e-3, ms[0])
    assert np.array_equal(bits(res.x_final), bits(ref.x_final))
    xs, ms = make_ensemble([5, 9], seed0=3)
    res = solve_ensemble("DOPRI5", [5, 9], pack(xs), 1.0, 1.0, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe)
    assert np.array_equal(bits(res.x_final), bits(pack(xs)))
    assert res.total_attempts == 0 and res.status_names == ["reached_end"] * 2


def test_ensemble_autotune_reports_all_candidates(exe):
    from ode_prototyping_amd.ensemble import (ENS_TILE_CONFIGS, ENS_TILE_UNSPLIT, pack,
                                              solve_ensemble)
    sizes = [300, 20, 1000, 64, 513]
    xs, ms = make_ensemble(sizes, seed0=11)
    res = solve_ensemble("DOPRI5", sizes, pack(xs), 0.0, 1e-3, 1e-10, 1e-10, 1e-3, pack(ms),
                         variant="fast", exe=exe)
    assert res.tune_ms.size == ENS_TILE_CONFIGS and np.all(res.tune_ms > 0.0)
    # the choice: within 3 % of the fastest candidate, the most target tiles
    # (128-target tiles for systems of at least 1024 / 256 bodies halve them;
    # the shared-tile configurations have 64-target tiles), then the fastest
    tpl2 = ([None, None, 1024, 1024, 256, 256] * 2 +
            [None] * (ENS_TILE_CONFIGS - ENS_TILE_UNSPLIT))
    tiles = [sum(-(-n // (128 if t and n >= t else 64)) for n in sizes) for t in tpl2]
    near = [c for c in range(len(tpl2)) if res.tune_ms[c] <= 1.03 * res.tune_ms.min()]
    most = max(tiles[c] for c in near)
    pick = min((c for c in near if tiles[c] == most), key=lambda c: (res.tune_ms[c], c))
    assert res.tile_config == pick
    assert res.rhs_ms == res.tune_ms[res.tile_config]


def test_ensemble_rejects_bad_arguments(exe):
    from ode_prototyping_amd.core import OdeHipError
    from ode_prototyping_amd.ensemble import ensemble_rhs, solve_ensemble
    with pytest.raises(OdeHipError):
        ensemble_rhs(exe, [4, 0], np.zeros(24), 1e-3)
    with pytest.raises(OdeHipError):
        ensemble_rhs(exe, [4], np.zeros(24), 1e-3, tile_config=18)
    with pytest.raises(ValueError):
        ensemble_rhs(exe, [4], np.zeros(23), 1e-3)
    with pytest.raises(OdeHipError):
        solve_ensemble("DOPRI5", [4], np.zeros(24), 0.0, 1.0, 1e-8, 1e-8, 1e-3,
                       controller="learned", exe=exe)


@pytest.mark.parametrize("tableau,fsal", [("DOPRI5", False), ("DOP853", False),
                                          ("DOP853", True)])
def test_ensemble_dense_output_matches_single_system(exe, tableau, fsal):
    """Dense output of every system at common times: each system's rows equal
    the single-system solve's dense rows bit for bit (systems of different
    sizes finish in different rounds, so the running-system lists are
    rebuilt during the solve), and so do its final state and counters; rows
    a system does not reach are NaN."""
    import ode_prototyping_amd as oh
    from ode_prototyping_amd.ensemble import pack, solve_ensemble
    sizes = [20, 90, 300, 33, 150, 7]
    xs, ms = [], []
    for k, n in enumerate(sizes):
        x, m = oh.scenarios.uniform_cube(n, seed=60 + k)
        xs.append(x)
        ms.append(m)
    tf = 0.02
    times = np.linspace(0.0, tf, 11)[1:]  # the last (tf) is never written
    res = solve_ensemble(tableau, sizes, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                         exe=exe, dense_times=times, fsal=fsal)
    assert res.n_unfinished == 0 and np.all(res.n_dense == 9)
    for k, n in enumerate(sizes):
        one = oh.integrate(tableau, xs[k], 0.0, tf, 1e-10, 1e-10,
                           oh.NBodyODE(1e-3, ms[k], "exact"), exe=exe, dense_times=times,
                           fsal=fsal)
        assert one.n_dense == 9
        assert np.array_equal(bits(res.dense_rows(k)), bits(one.dense_states))
        assert np.array_equal(bits(res.states()[k]), bits(one.x_final))
        assert (res.accepted[k], res.rejected[k], res.rhs_evals[k]) == (
            one.accepted, one.rejected, one.rhs_evals)
    assert np.all(np.isnan(res.dense_states[9]))
    # an accepted-step budget stops systems at different times
    b = solve_ensemble(tableau, sizes, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                       exe=exe, dense_times=times, max_accepted=6, fsal=fsal)
    for k in range(len(sizes)):
        nk = int(np.sum(times < b.t[k]))
        assert b.n_dense[k] == nk
        rows = b.dense_states[:, 6 * sum(sizes[:k]):6 * sum(sizes[:k + 1])]
        assert not np.any(np.isnan(rows[:nk])) and np.all(np.isnan(rows[nk:]))
        assert np.array_equal(bits(rows[:nk]), bits(res.dense_rows(k)[:nk]))
    with pytest.raises(oh.core.OdeHipError, match="DOPRI5 or DOP853"):
        solve_ensemble("RKF78", sizes, pack(xs), 0.0, tf, 1e-10, 1e-10, 1e-3, pack(ms),
                       exe=exe, dense_times=times)

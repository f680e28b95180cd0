"""N-body RHS and executor-level kernels on the GPU against the CPU oracle.

Bar: the exact variant is bit-identical to the oracle (itself bit-identical
to the reference's NBodyODE with SingleThreadedExecutor); the fast FP64
variant within 5e-14 of the largest acceleration component; the mixed
variant within 2e-6, and its near-field pairs to FP64 accuracy.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 63, 64, 65, 100, 257, 1000, 1024, 2049]


def ic(n, seed=0, masses=True):
    from ode_prototyping_amd import scenarios
    x, m = scenarios.uniform_cube(n, seed=seed)
    return x, (m if masses else None)


def gpu_rhs(exe, x, softening, masses, variant):
    import ode_prototyping_amd as oh
    xs = oh.HipState(exe, data=x)
    ks = oh.HipState(exe, x.size)
    oh.NBodyODE(softening, masses, variant)(exe, xs, ks)
    return ks.to_numpy()


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("masses", [False, True])
def test_exact_rhs_bit_identical(exe, n, masses):
    x, m = ic(n, seed=n, masses=masses)
    got = gpu_rhs(exe, x, 1e-3, m, "exact")
    want = O.rhs(O.RHS_NBODY, x, 1e-3, m, threads=8)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), \
        f"n={n}: max abs diff {np.max(np.abs(got - want))}"


@pytest.mark.parametrize("n", [4095, 8192, 8193, 8208, 8256, 12289, 16384, 20001, 24576])
def test_exact_rhs_bit_identical_large(exe, n):
    """Both tile shapes at the sizes where the dispatch switches (32-target
    tiles up to 8192, 64-target tiles of 48 partners above); above, with a
    partial last tile (8193 = 170 * 48 + 33, 12289 = 256 * 48 + 1, 16384 =
    341 * 48 + 16) and without one (8208 = 171 * 48, 8256 = 172 * 48), so
    the two-tile steady-state loop ends on either parity of full tiles; and
    20001 and 24576 bodies, where the 64-target grid would leave its last round
    of workgroups half empty and the 32-target kernel runs in 3 rounds."""
    x, m = ic(n, seed=n, masses=True)
    got = gpu_rhs(exe, x, 1e-3, m, "exact")
    want = O.rhs(O.RHS_NBODY, x, 1e-3, m, threads=8)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), \
        f"n={n}: max abs diff {np.max(np.abs(got - want))}"


def test_exact_rhs_unsoftened(exe):
    """eps = 0: the diagonal term must be skipped, not evaluated (0/0)."""
    x, _ = ic(300, seed=3, masses=False)
    got = gpu_rhs(exe, x, 0.0, None, "exact")
    want = O.rhs(O.RHS_NBODY, x, 0.0, None)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, want)


def test_exact_rhs_reference_spinning_cube(exe):
    """The reference's own 1024-body test state (SpinningCubeNBodyTest.hpp)."""
    x, soft, _, _ = O.spinning_cube_ic(1024)
    got = gpu_rhs(exe, x, soft, None, "exact")
    want = O.rhs(O.RHS_NBODY, x, soft, None, threads=8)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [2, 65, 1000, 4096])
@pytest.mark.parametrize("eps", [1e-2, 0.0])
def test_fast_rhs_close(exe, n, eps):
    x, m = ic(n, seed=7)
    got = gpu_rhs(exe, x, eps, m, "fast")
    want = O.rhs(O.RHS_NBODY, x, eps, m, threads=8)
    acc_g, acc_w = got[3 * n:], want[3 * n:]
    assert np.array_equal(got[:3 * n], want[:3 * n])
    scale = np.max(np.abs(acc_w))
    assert np.max(np.abs(acc_g - acc_w)) <= 5e-14 * scale


@pytest.mark.parametrize("n", [2, 65, 1000, 4096])
def test_mixed_rhs_close(exe, n):
    x, m = ic(n, seed=11)
    got = gpu_rhs(exe, x, 1e-2, m, "mixed")
    want = O.rhs(O.RHS_NBODY, x, 1e-2, m, threads=8)
    acc_g, acc_w = got[3 * n:], want[3 * n:]
    assert np.array_equal(got[:3 * n], want[:3 * n])
    scale = np.max(np.abs(acc_w))
    assert np.max(np.abs(acc_g - acc_w)) <= 2e-6 * scale


def mass_case(n, case, seed):
    """Masses of one mass form: all equal (0.37 / n, not a power of two),
    spread over [0.5, 2] / n, or equal 1 / n but for the first or the last
    body (1.5 / n)."""
    if case == "equal":
        return np.full(n, 0.37 / n)
    if case == "spread":
        return np.random.default_rng(seed).uniform(0.5, 2.0, n) / n
    m = np.full(n, 1.0 / n)
    m[0 if case == "first_differs" else -1] = 1.5 / n
    return m


@pytest.mark.parametrize("variant,bar", [("fast", 5e-14), ("mixed", 2e-6), ("mixed_near", 2e-6)])
@pytest.mark.parametrize("case", ["equal", "spread", "first_differs", "last_differs"])
@pytest.mark.parametrize("n", [1000, 4096])
def test_fast_mixed_rhs_mass_forms(exe, variant, bar, case, n):
    """Equal masses take unit-mass terms and one scale of each finished sum,
    any other masses the per-term products.  Every form meets the variant's
    bar against the oracle's per-term products; a body whose mass differs
    from the rest, first or last, must keep the workgroups off the equal-mass
    form (which would misweigh its terms by a third: about 1e-3 of the
    largest acceleration here, far above the mixed bar)."""
    import ode_prototyping_amd as oh
    x, _ = ic(n, seed=19)
    m = mass_case(n, case, 19)
    if variant == "mixed_near":
        exe.near_field = oh.MIXED_NEAR_CLOSE
    try:
        got = gpu_rhs(exe, x, 1e-2, m, "fast" if variant == "fast" else "mixed")
    finally:
        exe.near_field = oh.MIXED_NEAR_DEFAULT
    want = O.rhs(O.RHS_NBODY, x, 1e-2, m, threads=8)
    assert np.array_equal(got[:3 * n], want[:3 * n])
    scale = np.max(np.abs(want[3 * n:]))
    err = np.max(np.abs(got[3 * n:] - want[3 * n:]))
    assert err <= bar * scale, (variant, case, n, err / scale)


@pytest.mark.parametrize("n", [16384, 16385, 20001])
@pytest.mark.parametrize("variant,bar", [("fast", 5e-14), ("mixed", 2e-6)])
@pytest.mark.parametrize("case", ["equal", "spread"])
def test_fast_mixed_rhs_either_side_of_one_workgroup_per_cu(exe, n, variant, bar, case):
    """The fast kernel's partner loop follows the grid: with at most one 64-target
    workgroup per CU (n <= 16384 on 256 CUs) it prefetches sets of partners, with more it
    loads 8 partners per step and waits; both add the terms in increasing j and meet the
    same bar against the oracle, with equal and with per-term masses."""
    x, _ = ic(n, seed=23)
    m = mass_case(n, case, 23)
    got = gpu_rhs(exe, x, 1e-2, m, variant)
    want = O.rhs(O.RHS_NBODY, x, 1e-2, m, threads=8)
    assert np.array_equal(got[:3 * n], want[:3 * n])
    scale = np.max(np.abs(want[3 * n:]))
    err = np.max(np.abs(got[3 * n:] - want[3 * n:]))
    assert err <= bar * scale, (variant, case, n, err / scale)


@pytest.mark.parametrize("n", [2, 65, 1000])
@pytest.mark.parametrize("eps", [1e-2, 0.0])
def test_mixed_rhs_all_near_is_fp64(exe, n, eps):
    """With a near field that covers every pair, the mixed RHS computes every
    term in FP64: the fast variant's bar (5e-14), not FP32's."""
    import ode_prototyping_amd as oh
    x, m = ic(n, seed=13)
    exe.near_field = 1e30
    try:
        got = gpu_rhs(exe, x, eps, m, "mixed")
    finally:
        exe.near_field = oh.MIXED_NEAR_DEFAULT
    want = O.rhs(O.RHS_NBODY, x, eps, m, threads=8)
    scale = np.max(np.abs(want[3 * n:]))
    assert np.max(np.abs(got[3 * n:] - want[3 * n:])) <= 5e-14 * scale


def test_mixed_rhs_close_pair_in_fp64(exe):
    """A close pair (distance 1e-3, softening 1e-4) among 500 bodies: its two
    accelerations are dominated by the pair term, which the close-encounter
    near field (20) computes in FP64 (rel. error <= 1e-10); the default
    all-FP32 form (near field 0) only meets the FP32 bar."""
    import ode_prototyping_amd as oh
    n = 500
    x, m = ic(n, seed=17)
    x[3:6] = x[0:3] + np.array([6e-4, -5e-4, 6e-4])  # |d| ~ 1e-3
    want = O.rhs(O.RHS_NBODY, x, 1e-4, m, threads=8)
    assert exe.near_field == oh.MIXED_NEAR_DEFAULT == 0.0
    exe.near_field = oh.MIXED_NEAR_CLOSE
    try:
        got = gpu_rhs(exe, x, 1e-4, m, "mixed")
    finally:
        exe.near_field = oh.MIXED_NEAR_DEFAULT
    pair = want[3 * n:3 * n + 6]
    rel = np.max(np.abs(got[3 * n:3 * n + 6] - pair)) / np.max(np.abs(pair))
    assert rel <= 1e-10, rel
    got32 = gpu_rhs(exe, x, 1e-4, m, "mixed")
    rel32 = np.max(np.abs(got32[3 * n:3 * n + 6] - pair)) / np.max(np.abs(pair))
    assert rel32 <= 2e-6
    scale = np.max(np.abs(want[3 * n:]))
    assert np.max(np.abs(got[3 * n:] - want[3 * n:])) <= 2e-6 * scale


def test_near_field_setter_rejects_bad_values(exe):
    import ode_prototyping_amd as oh
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(oh.OdeHipError):
            exe.near_field = bad
    assert exe.near_field == oh.MIXED_NEAR_DEFAULT


def test_nbody_acc_matches_rhs(exe):
    """NBodyODE::acc (positions -> accelerations) equals the RHS's acc half."""
    import ode_prototyping_amd as oh
    n = 300
    x, m = ic(n, seed=5)
    ode = oh.NBodyODE(1e-3, m, "exact")
    xs = oh.HipState(exe, data=x)
    acc = oh.HipState(exe, 3 * n)
    ode.acc(exe, xs, acc)
    want = O.rhs(O.RHS_NBODY, x, 1e-3, m)[3 * n:]
    assert np.array_equal(acc.to_numpy(), want)


def test_simple_odes(exe):
    import ode_prototyping_amd as oh
    x = np.array([2.0, 0.5])
    xs, ks = oh.HipState(exe, data=x), oh.HipState(exe, 2)
    oh.VanDerPolODE()(exe, xs, ks)
    assert np.array_equal(ks.to_numpy(), O.rhs(O.RHS_VANDERPOL, x))
    x = np.arange(10.0) * 1.5
    xs, ks = oh.HipState(exe, data=x), oh.HipState(exe, 10)
    oh.ExponentialODE()(exe, xs, ks)
    assert np.array_equal(ks.to_numpy(), x)


def test_state_semantics(exe):
    """HipState mirrors CudaState (CudaStateTest.cu:18-92): zero default,
    host round trip, deep copy, independence, fill."""
    import ode_prototyping_amd as oh
    s = oh.HipState(exe, 10)
    assert np.array_equal(s.to_numpy(), np.zeros(10))
    data = np.arange(10.0)
    a = oh.HipState(exe, data=data)
    assert np.array_equal(a.to_numpy(), data)
    b = a.copy()
    exe.fill(a, 3.0)
    assert np.array_equal(b.to_numpy(), data)
    assert np.array_equal(a.to_numpy(), np.full(10, 3.0))


def test_rk_norm_canonical_order(exe):
    """rk_norm follows the documented group order bit for bit, for sizes
    around the group boundaries; with n <= 2 it equals the sequential sum."""
    import ode_prototyping_amd as oh
    rng = np.random.default_rng(0)
    for n in [1, 2, 3, 64, 255, 256, 257, 1000, 98304]:
        v = rng.normal(size=n) * 1e-9
        sc = 1e-10 + 1e-10 * np.abs(rng.normal(size=n))
        got = exe.rk_norm(oh.HipState(exe, data=v), oh.HipState(exe, data=sc))
        assert got == O.rk_norm(v, sc, O.HIP_ORDER), n
        if n <= 2:
            assert got == O.rk_norm(v, sc)


@pytest.mark.parametrize("tableau", ["HE21", "RKF45", "DOPRI5", "DVERK", "RKF78", "DOP853"])
def test_stage_update_target_kernels(exe, tableau):
    """rk_stage / update_state_and_error / update_error_target kernels equal
    the reference expressions (RKEmbeddedParallel.hpp:91-130) evaluated in
    the same order in IEEE double on the host."""
    import ode_prototyping_amd as oh
    T = O.tableau(tableau)
    S, n = T["S"], 777
    rng = np.random.default_rng(1)
    ks_h = rng.normal(size=S * n)
    x0_h = rng.normal(size=n)
    dt = 0.0123
    ks, x0 = oh.HipState(exe, data=ks_h), oh.HipState(exe, data=x0_h)
    temp = oh.HipState(exe, n)
    K = ks_h.reshape(S, n)
    for stage in range(1, S):
        exe.rk_stage(tableau, stage, dt, temp, ks, x0)
        tmp = np.zeros(n)
        for j in range(stage):
            tmp = tmp + T["a"][stage - 1][j] * K[j]
        assert np.array_equal(temp.to_numpy(), x0_h + tmp * dt), stage
    x, err = oh.HipState(exe, n), oh.HipState(exe, n)
    exe.rk_update_state_and_error(tableau, dt, x, err, ks, x0)
    xs, es = np.zeros(n), np.zeros(n)
    db = T["b"] - T["bt"]
    for j in range(S):
        xs = xs + T["b"][j] * K[j]
        es = es + db[j] * K[j]
    assert np.array_equal(x.to_numpy(), x0_h + xs * dt)
    assert np.array_equal(err.to_numpy(), es * dt)
    atol = np.full(n, 1e-9)
    rtol = np.full(n, 1e-8)
    sc = oh.HipState(exe, n)
    exe.rk_update_error_target(sc, oh.HipState(exe, data=atol), oh.HipState(exe, data=rtol),
                               x, x0)
    xh = x.to_numpy()
    want = atol + rtol * np.where(np.abs(xh) < np.abs(x0_h), np.abs(x0_h), np.abs(xh))
    assert np.array_equal(sc.to_numpy(), want)


def test_energy_and_momentum(exe):
    import ode_prototyping_amd as oh
    from ode_prototyping_amd import _lib
    import ctypes as C
    n = 1500
    x, m = ic(n, seed=9)
    xs, ms = oh.HipState(exe, data=x), oh.HipState(exe, data=m)
    out = np.zeros(3)
    exe.check(_lib.lib().odehip_nbody_energy(exe.handle, xs.ptr, n, 1e-3, ms.ptr,
                                             out.ctypes.data_as(C.POINTER(C.c_double))), "e")
    ke, pe = O.energy(x, n, 1e-3, m)
    assert abs(out[0] - ke) <= 1e-13 * abs(ke)
    assert abs(out[1] - pe) <= 1e-13 * abs(pe)
    exe.check(_lib.lib().odehip_nbody_momentum(exe.handle, xs.ptr, n, ms.ptr,
                                               out.ctypes.data_as(C.POINTER(C.c_double))), "p")
    p = O.momentum(x, n, m)
    assert np.all(np.abs(out - p) <= 1e-17)


@pytest.mark.parametrize("n", [200, 9000])
@pytest.mark.parametrize("eps", [1e-3, 1e-70, 0.0])
def test_exact_rhs_extreme_inputs_bit_identical(exe, eps, n):
    """Inputs that leave the exact kernel's shared-reciprocal range: zero
    displacement components (coplanar particles), separations below 2^-300
    and above 2^100, and softenings outside [2^-100, 2^50].  Every term must
    still match the IEEE expression bit for bit (NaN where the reference
    also produces one)."""
    rng = np.random.default_rng(42)
    pos = rng.uniform(-1, 1, (n, 3))
    pos[:40, 2] = 0.0                              # coplanar block: dz == 0
    pos[40:60] = pos[40] + rng.uniform(-1, 1, (20, 3)) * 1e-95   # tiny separations
    pos[60:80] *= 1e35                             # huge separations
    pos[80:90, 0] = pos[0, 0]                      # shared x coordinate
    x = np.concatenate([pos.ravel(), rng.normal(size=3 * n)])
    m = rng.uniform(0.5, 1.5, n) / n
    got = gpu_rhs(exe, x, eps, m, "exact")
    want = O.rhs(O.RHS_NBODY, x, eps, m, threads=8)
    finite = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), finite)
    assert np.array_equal(got[finite].view(np.uint64), want[finite].view(np.uint64))


@pytest.mark.parametrize("case", ["pow2", "pow2_first_only", "pow2_last_differs", "not_pow2",
                                  "pow2_out_of_range", "pow2_extreme_inputs"])
@pytest.mark.parametrize("n", [12289, 4097, 4095])
def test_exact_rhs_uniform_pow2_masses_bit_identical(exe, case, n):
    """Masses that are all one power of two 2^k (|k| <= 300) take the unit-mass terms and
    one exact scale of each sum: still bit-identical to the oracle's per-term products.
    Arrays that only start with a power of two, that end with another power of two, a
    uniform mass that is not one, a power of two outside the range, and a uniform power
    of two with coordinates outside the shared-reciprocal range (IEEE path) take the
    per-term path, equally exact.  12289 bodies: the 64-target kernel, 4097: the 32-target
    one, both with a partial last tile; 4095: the 32-target kernel's instance without the
    power-of-two path (below 4096 bodies), per-term products for every case."""
    rng = np.random.default_rng(99)
    x, _ = ic(n, seed=5)
    m = np.full(n, 2.0 ** -14)
    if case == "pow2_first_only":
        m[1:] = rng.uniform(0.5, 1.5, n - 1) / n
    elif case == "pow2_last_differs":
        m[-1] = 2.0 ** -13
    elif case == "not_pow2":
        m[:] = 1.0 / 12289.0
    elif case == "pow2_out_of_range":
        m[:] = 2.0 ** -400
    elif case == "pow2_extreme_inputs":
        x[3] = 1e35  # a coordinate above 2^98: the launch takes the IEEE path
    got = gpu_rhs(exe, x, 1e-3, m, "exact")
    want = O.rhs(O.RHS_NBODY, x, 1e-3, m, threads=8)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), \
        f"{case}: max abs diff {np.max(np.abs(got - want))}"

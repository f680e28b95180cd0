"""Host logic of the ensemble solve, on the CPU: the production planner
(ode_prototyping_amd/csrc/ensemble_plan.hpp) through tests/cpp/libensemble_plan_check.so.

Covers the two ways the tile tables could go wrong on the device:
* the tile region's capacity must hold the list of *every* running subset under every
  tile configuration (compaction rebuilds the list for the running systems, and greedy
  XCD dealing of a subset can load one queue more than the full set did);
* a right-hand-side call and a solve on the same systems must not share a cached
  upload unless their tables sit at the same arena offsets.
"""
import ctypes
import os

import numpy as np
import pytest

from ode_prototyping_amd.scenarios import log_uniform_sizes

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "cpp", "libensemble_plan_check.so")

PLAN_FIELDS = ["sys", "groups", "tiles", "groups_act", "st", "n_active", "xb0", "xb1", "ks",
               "ytmp", "kfsal", "sc0", "partA", "partB", "trace", "dense_t", "dense_out",
               "total"]
TILE_BYTES = 8  # EnsTile: two ints


@pytest.fixture(scope="module")
def P():
    if not os.path.exists(SO):
        pytest.fail("run __graft_entry__.build() (it builds tests/cpp)")
    lib = ctypes.CDLL(SO)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.epc_tiles_capacity.restype = ctypes.c_long
    lib.epc_tiles_capacity.argtypes = [ctypes.c_int, ip]
    lib.epc_full_list_max.restype = ctypes.c_long
    lib.epc_full_list_max.argtypes = [ctypes.c_int, ip]
    lib.epc_make_tiles.restype = ctypes.c_long
    lib.epc_make_tiles.argtypes = [ctypes.c_int, ip, ctypes.c_int, ctypes.c_char_p, ip, ip,
                                   ctypes.c_long]
    lib.epc_check_tiles.argtypes = [ctypes.c_int, ip, ip, ip, ctypes.c_long, ctypes.c_long]
    lib.epc_plan.argtypes = [ctypes.c_int, ip, ctypes.c_int, ctypes.c_long, ctypes.c_int,
                             ctypes.c_long, ctypes.POINTER(ctypes.c_ulong)]
    lib.epc_table_key_hit.argtypes = [ctypes.c_int, ip] + [ctypes.c_int] * 5
    lib.epc_table_key_hit_shifted.argtypes = [ctypes.c_int, ip, ctypes.c_int, ctypes.c_long]
    lib.epc_config.argtypes = [ctypes.c_int, ip, ip, ip, ip]
    lib.epc_fill_waves.restype = ctypes.c_long
    lib.epc_tail_config.argtypes = [ctypes.c_int, ip, ctypes.c_char_p, ctypes.c_int,
                                    ctypes.c_int]
    return lib


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _sizes(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _config(P, c):
    """(waves per workgroup, 128-target threshold, xcd placement, waves per tile)"""
    w, t, x, sp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert P.epc_config(c, ctypes.byref(w), ctypes.byref(t), ctypes.byref(x),
                        ctypes.byref(sp)) == 0
    return w.value, t.value, x.value, sp.value


def _tiles(P, sizes, cfg, live=None):
    n = len(sizes)
    mask = None if live is None else bytes(np.asarray(live, dtype=np.uint8))
    count = P.epc_make_tiles(n, _ip(sizes), cfg, mask, None, None, 0)
    sys = np.zeros(max(count, 1), np.int32)
    i0 = np.zeros(max(count, 1), np.int32)
    assert P.epc_make_tiles(n, _ip(sizes), cfg, mask, _ip(sys), _ip(i0), count) == count
    return sys[:count], i0[:count]


def _check_coverage(P, sizes, cfg, live, sys, i0):
    """Every target of every running system in exactly one tile; nothing else."""
    wpb, tpl2_min_n, _, _ = _config(P, cfg)
    assert len(sys) % wpb == 0
    real = sys >= 0
    run = np.ones(len(sizes), bool) if live is None else np.asarray(live, bool)
    assert np.all(run[sys[real]]), "a tile of a finished system"
    covered = {}
    for s, a in zip(sys[real], i0[real]):
        step = 64 * (2 if sizes[s] >= tpl2_min_n else 1)
        assert a % step == 0 and 0 <= a < sizes[s]
        covered.setdefault(int(s), []).append(int(a))
    for s in np.nonzero(run)[0]:
        step = 64 * (2 if sizes[s] >= tpl2_min_n else 1)
        assert sorted(covered.get(int(s), [])) == list(range(0, int(sizes[s]), step))


def _subsets(rng, sizes, count):
    n = len(sizes)
    order = np.argsort(-sizes, kind="stable")
    yield np.ones(n, bool)
    for frac in (0.5, 0.25, 0.1):
        live = np.zeros(n, bool)
        live[order[:max(1, int(frac * n))]] = True  # the largest systems run longest
        yield live
        live = np.zeros(n, bool)
        live[order[-max(1, int(frac * n)):]] = True
        yield live
    for _ in range(count):
        live = rng.random(n) < rng.uniform(0.02, 0.9)
        if not live.any():
            live[rng.integers(n)] = True
        yield live


ENSEMBLES = {
    "c4_2048": lambda: log_uniform_sizes(2048, 16, 4096, seed=0),
    "lu32": lambda: log_uniform_sizes(32, 16, 4096, seed=0),
    "mixed34": lambda: np.rint(np.geomspace(16, 1943, 34)).astype(np.int64),
    "small": lambda: np.array([1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257]),
}


@pytest.mark.parametrize("name", sorted(ENSEMBLES))
def test_tile_capacity_holds_every_running_subset(P, name):
    sizes = _sizes(ENSEMBLES[name]())
    n = len(sizes)
    cap = P.epc_tiles_capacity(n, _ip(sizes))
    rng = np.random.default_rng(1234)
    trials = 12 if n > 1000 else 60
    for live in _subsets(rng, sizes, trials):
        for cfg in range(P.epc_n_configs()):
            # compaction may also switch to the configuration without 128-target tiles
            for c in {cfg, P.epc_without_tpl2(cfg)}:
                sys, i0 = _tiles(P, sizes, c, live)
                assert len(sys) <= cap, (name, c, int(live.sum()))
                _check_coverage(P, sizes, c, live, sys, i0)
                assert P.epc_check_tiles(n, _ip(sizes), _ip(sys), _ip(i0), len(sys), cap) == 0


def test_old_capacity_rule_overflows_on_subsets(P):
    """The rule the library used to size the region (longest full-ensemble list) is
    exceeded by running subsets: the CPU form of the out-of-bounds tile upload."""
    over = 0
    for name in ("lu32", "mixed34"):
        sizes = _sizes(ENSEMBLES[name]())
        n = len(sizes)
        old = P.epc_full_list_max(n, _ip(sizes))
        new = P.epc_tiles_capacity(n, _ip(sizes))
        assert new >= old
        rng = np.random.default_rng(7)
        for live in _subsets(rng, sizes, 200):
            for cfg in range(P.epc_n_configs()):
                sys, i0 = _tiles(P, sizes, cfg, live)
                if len(sys) > old:
                    over += 1
                    # the host check refuses such a list against the old capacity
                    assert P.epc_check_tiles(n, _ip(sizes), _ip(sys), _ip(i0), len(sys), old) == 1
    assert over > 0


def test_check_tiles_rejects_bad_entries(P):
    sizes = _sizes([100, 300])
    sys = np.array([0, 1, 1, 2], np.int32)
    i0 = np.array([0, 0, 64, 0], np.int32)
    assert P.epc_check_tiles(2, _ip(sizes), _ip(sys), _ip(i0), 4, 100) == 1  # system 2
    sys[3] = -1
    assert P.epc_check_tiles(2, _ip(sizes), _ip(sys), _ip(i0), 4, 100) == 0  # empty slot
    i0[1] = 300
    assert P.epc_check_tiles(2, _ip(sizes), _ip(sys), _ip(i0), 4, 100) == 1  # past n
    i0[1] = 0
    assert P.epc_check_tiles(2, _ip(sizes), _ip(sys), _ip(i0), 4, 3) == 1  # over capacity


def _plan(P, sizes, S, trace_cap, with_state, n_dense=0):
    out = (ctypes.c_ulong * 18)()
    assert P.epc_plan(len(sizes), _ip(sizes), S, trace_cap, int(with_state), n_dense, out) == 0
    return dict(zip(PLAN_FIELDS, list(out)))


@pytest.mark.parametrize("name", sorted(ENSEMBLES))
def test_rhs_and_solve_plans_share_table_offsets(P, name):
    sizes = _sizes(ENSEMBLES[name]())
    cap = P.epc_tiles_capacity(len(sizes), _ip(sizes))
    rhs = _plan(P, sizes, 0, 0, False)
    for S, trace, dense in ((7, 0, 0), (12, 50, 3), (2, 1, 0)):
        sol = _plan(P, sizes, S, trace, True, dense)
        for f in ("sys", "groups", "tiles"):
            assert rhs[f] == sol[f], f
        # regions in order, non-overlapping, the tile region sized for the capacity
        assert sol["tiles"] + cap * TILE_BYTES <= sol["groups_act"]
        assert sol["groups_act"] + 8 * (sol["groups"] - sol["sys"]) >= sol["groups_act"]
        assert sol["groups_act"] < sol["st"] <= sol["n_active"] < sol["xb0"] < sol["total"]
    assert rhs["tiles"] + cap * TILE_BYTES <= rhs["st"]
    assert rhs["groups_act"] == rhs["groups"]


@pytest.mark.parametrize("name", sorted(ENSEMBLES))
def test_table_cache_across_rhs_and_solve(P, name):
    sizes = _sizes(ENSEMBLES[name]())
    n = len(sizes)
    for cfg in (0, 1, 7):
        # solve -> rhs and rhs -> solve on one arena: the offsets agree, so the cached
        # upload is the right one
        assert P.epc_table_key_hit(n, _ip(sizes), cfg, 7, 1, 0, 0) == 1
        assert P.epc_table_key_hit(n, _ip(sizes), cfg, 0, 0, 12, 1) == 1
        # a key set with other table offsets never hits
        assert P.epc_table_key_hit_shifted(n, _ip(sizes), cfg, 256) == 0
        assert P.epc_table_key_hit_shifted(n, _ip(sizes), cfg, 0) == 1


def _chunk_waves(sizes, live, split):
    """Waves with partner chunks (128 partners each) when `split` waves share a
    64-target tile of every running system."""
    n = np.asarray(sizes, np.int64)[np.asarray(live, bool)]
    return int(np.sum(-(-n // 64) * np.minimum(split, -(-n // 128))))


def _tail_model(P, sizes, live, cfg, exact):
    """The decision table the library documents for tail_config, restated."""
    n = np.asarray(sizes, np.int64)[np.asarray(live, bool)]
    tiles = int(np.sum(-(-n // 64)))
    ratio = float(n.max()) * 1024 / float(np.sum(-(-n // 64) * n))
    wpb, tpl2, xcd, split = _config(P, cfg)
    unsplit = cfg if split == 1 else P.epc_split_config(xcd, 1)

    def useful(c):
        while c > 1 and _chunk_waves(sizes, live, c) <= _chunk_waves(sizes, live, c // 2):
            c //= 2
        return c

    if exact:
        if ratio > 1.2 and useful(8) == 8:
            return P.epc_split_config(xcd, 8)
        _, t2, _, _ = _config(P, unsplit)
        real = int(np.sum(-(-n // np.where(n >= t2, 128, 64))))
        return P.epc_without_tpl2(unsplit) if real < 4096 else unsplit
    if tiles >= 8192:
        return cfg
    c = useful(8 if ratio >= 1.0 else 4)
    return P.epc_split_config(xcd, c) if c > 1 else P.epc_without_tpl2(unsplit)


@pytest.mark.parametrize("name", sorted(ENSEMBLES))
def test_tail_config_rule(P, name):
    """The configuration compaction (and the start of a solve) moves the running
    systems to follows the documented decision table for every running subset and
    starting configuration, keeps the XCD placement, and never shares a tile among
    more waves than add working waves."""
    sizes = _sizes(ENSEMBLES[name]())
    n = len(sizes)
    assert P.epc_fill_waves() == 4096 and P.epc_chunk() == 128
    rng = np.random.default_rng(99)
    for live in _subsets(rng, sizes, 10 if n > 1000 else 40):
        mask = bytes(live.astype(np.uint8))
        for cfg in range(P.epc_n_configs()):
            xcd = _config(P, cfg)[2]
            for exact in (0, 1):
                got = P.epc_tail_config(n, _ip(sizes), mask, cfg, exact)
                assert got == _tail_model(P, sizes, live, cfg, exact), (name, cfg, exact)
                gw, gt, gx, gs = _config(P, got)
                assert gx == xcd
                if gs > 1:
                    assert gw == 1 and gt == 1 << 30
                    assert _chunk_waves(sizes, live, gs) > _chunk_waves(sizes, live, gs // 2)
                if exact:
                    assert gs in (1, 8)


def test_tail_config_cases(P):
    """Subsets of the 2048-system ensemble whose right-hand sides were timed for the
    table, starting from configuration 7 (one wave per tile, XCD placement)."""
    sizes = _sizes(ENSEMBLES["c4_2048"]())
    order = np.argsort(-sizes, kind="stable")
    cases = {  # subset: (exact, fast) waves per tile
        "top8": (order[:8], 8, 8), "top16": (order[:16], 1, 8),
        "top32": (order[:32], 1, 4), "top128": (order[:128], 1, 4),
        "every32": (order[::32], 8, 8), "every16": (order[::16], 8, 8),
        "mid64": (order[256:320], 1, 4), "mid16": (order[512:528], 8, 8)}
    for name, (pick, want_exact, want_fast) in cases.items():
        live = np.zeros(len(sizes), bool)
        live[pick] = True
        mask = bytes(live.astype(np.uint8))
        for exact, want in ((1, want_exact), (0, want_fast)):
            got = P.epc_tail_config(len(sizes), _ip(sizes), mask, 7, exact)
            assert _config(P, got)[3] == want, (name, exact)
    # the full ensemble keeps its tuned configuration
    for cfg in (1, 3, 7):
        for exact in (0, 1):
            assert P.epc_tail_config(len(sizes), _ip(sizes), None, cfg, exact) == cfg
    # one chunk per system: nothing to share
    small = _sizes([64] * 1000)
    assert P.epc_tail_config(len(small), _ip(small), None, 3, 0) == 1

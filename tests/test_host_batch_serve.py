"""Host-side logic of the batched launch and the server (no device)."""
import numpy as np
import pytest

from sde_mc_amd import _lib as L
from sde_mc_amd.batch import chunks_of, job_cost, plan_segments
from sde_mc_amd.serve import ChunkPool
from sde_mc_amd import workloads as W
from tests import jobs as J


def test_chunks_of():
    assert [chunks_of(n) for n in (1, 63, 64, 65, 1 << 22)] == [1, 1, 1, 2, 1 << 16]


def _expand(plan):
    """(slot, chunk) -> times covered, walking each group's tiles the way the
    persistent kernel does (tile t -> the last segment with tile_begin <= t)."""
    seen = {}
    for g in range(len(plan.groups)):
        s0, ns, tiles, _ = plan.hdr[g]
        segs = plan.segs[s0:s0 + ns]
        for t in range(tiles):
            s = segs[np.searchsorted(segs[:, 4], t, side='right') - 1]
            c0 = s[1] + (t - s[4]) * s[3]
            for c in range(c0, min(c0 + s[3], s[1] + s[2])):
                seen[(s[0], c)] = seen.get((s[0], c), 0) + 1
    return seen


def test_segments_cover_every_chunk_once_costliest_first():
    rng = np.random.default_rng(0)
    n = 50
    counts = rng.integers(0, 300, n)
    firsts = rng.integers(0, 1000, n)
    costs = rng.uniform(1, 100, n)
    for classes in (None, rng.integers(0, 9, n), np.full(n, 3)):
        for per in (1, 2, 4, 7):
            plan = plan_segments(np.arange(n), firsts, counts, costs, per, classes,
                                 min_tile_cost=50.0 * per)
            segs = plan.segs
            assert len(segs) == int((counts > 0).sum())
            assert len(plan.groups) <= L.MAX_GROUPS
            assert plan.hdr[:len(plan.groups), 1].sum() == len(segs)
            nt = (segs[:, 2] + segs[:, 3] - 1) // segs[:, 3]
            assert plan.tiles == nt.sum() == plan.hdr[:, 2].sum()
            for g, c in enumerate(plan.groups):
                s0, ns, tiles, _ = plan.hdr[g]
                gs = segs[s0:s0 + ns]
                # consecutive tile ranges from 0 within the group
                gnt = (gs[:, 2] + gs[:, 3] - 1) // gs[:, 3]
                assert gs[0, 4] == 0 and np.array_equal(gs[1:, 4], np.cumsum(gnt)[:-1])
                assert tiles == gnt.sum()
                # costliest first within the group
                assert np.all(np.diff(costs[gs[:, 0]]) <= 0)
                if classes is not None and c != L.CLASS_GENERIC:
                    assert np.all(classes[gs[:, 0]] == c)
            # expand tiles -> chunks: exactly the requested ranges
            want = {(i, firsts[i] + c): 1 for i in range(n) for c in range(counts[i])}
            assert _expand(plan) == want


def test_short_jobs_get_several_chunks_per_tile():
    from sde_mc_amd.batch import MIN_TILE_COST
    costs = np.array([MIN_TILE_COST * 2, MIN_TILE_COST / 8, MIN_TILE_COST / 7.5])
    plan = plan_segments([0, 1, 2], [0, 0, 0], [100, 100, 100], costs, 1)
    per = dict(zip(plan.segs[:, 0], plan.segs[:, 3]))
    assert per == {0: 1, 1: 8, 2: 8}
    plan = plan_segments([0, 1, 2], [0, 0, 0], [100, 100, 100], costs, 4)
    assert dict(zip(plan.segs[:, 0], plan.segs[:, 3])) == {0: 4, 1: 8, 2: 8}


def test_all_classes_fold_into_eight_groups():
    n = L.CLASS_GENERIC + 1  # every class, ten
    plan = plan_segments(np.arange(n), np.zeros(n), np.full(n, 4), np.arange(1.0, n + 1), 1,
                         np.arange(n))
    assert len(plan.groups) == L.MAX_GROUPS and L.CLASS_GENERIC in plan.groups
    # the cheapest dedicated classes (0 and 1, costs 1 and 2) join the generic group
    assert 0 not in plan.groups and 1 not in plan.groups
    g = plan.groups.index(L.CLASS_GENERIC)
    s0, ns, _, _ = plan.hdr[g]
    assert sorted(plan.segs[s0:s0 + ns, 0]) == [0, 1, L.CLASS_GENERIC]


def test_tile_cost_floor_follows_the_launch():
    from sde_mc_amd.batch import MAX_TILE_COST, MIN_TILE_COST, BatchTune, tile_cost_floor
    t = BatchTune(2, 16)  # 32 resident wavefronts per CU
    # the served mix (1.9e11 cost units) takes the largest tiles, the MLMC
    # level set (3e8) the smallest; in between, 64 tiles per resident wavefront
    assert tile_cost_floor(1.9e11, t, 256) == MAX_TILE_COST
    assert tile_cost_floor(3.0e8, t, 256) == MIN_TILE_COST
    mid = 64 * 256 * 32 * 3.0e4
    assert tile_cost_floor(mid, t, 256) == pytest.approx(3.0e4)


def test_job_cost_orders_levy_above_merton():
    m = W.request_job(W.Request(0, 'merton', 1.0, 1.0, 200, 1 << 16, 1))
    l = W.request_job(W.Request(1, 'levy', 1.0, 1.0, 200, 1 << 16, 2))
    assert job_cost(l) > 2 * job_cost(m)
    c = J.job_for(J.merton_sde(), 1.0, 200, 1000, coarse_factor=2)
    assert job_cost(c) > job_cost(J.job_for(J.merton_sde(), 1.0, 200, 1000))


def test_chunk_pool_never_exceeds_capacity():
    rng = np.random.default_rng(1)
    pool = ChunkPool(1000)
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.45:
            s, n = live.pop(rng.integers(len(live)))
            pool.release(s, n)
        else:
            n = int(rng.integers(1, 200))
            s = pool.alloc(n)
            if s is not None:
                assert 0 <= s and s + n <= 1000
                for (s2, n2) in live:
                    assert s + n <= s2 or s2 + n2 <= s  # no overlap
                live.append((s, n))
        assert pool.used == sum(n for _, n in live) <= pool.capacity
    for s, n in live:
        pool.release(s, n)
    assert pool.free == [(0, 1000)] and pool.used == 0


def test_request_mix_distribution():
    reqs = W.request_mix(2000, seed=3, rate=200)
    paths = np.array([r.paths for r in reqs])
    steps = np.array([r.steps for r in reqs])
    assert paths.min() >= 2 ** 14 and paths.max() <= 2 ** 22
    assert steps.min() >= 50 and steps.max() <= 800
    assert set(r.maturity for r in reqs) == {0.25, 0.5, 1.0, 2.0}
    assert 0.4 < np.mean([r.model == 'merton' for r in reqs]) < 0.6
    gaps = np.diff([r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 200) < 0.1 / 200 * 3
    # log-uniform: the median path count is the geometric midpoint 2^18
    assert 2 ** 17.5 < np.median(paths) < 2 ** 18.5


def test_request_jobs_validate_on_host():
    import ctypes
    lib = L.load()
    for r in W.request_mix(20, seed=4):
        job = W.request_job(r)
        buf = ctypes.create_string_buffer(lib.sdemc_sizeof_prepared_job())
        base = (ctypes.c_int64 * 1)(0)
        cs = ctypes.c_int32(-1)
        assert lib.sdemc_prepare_jobs(ctypes.byref(job), 1, base, buf, ctypes.byref(cs)) == L.OK
        # 1-D Merton, or 2-D Levy with per-step counts or the arrival clock;
        # single level, uniform grid
        assert cs.value in (0, 4, L.CLASS_INDEP2_CLOCK)


def test_job_classes():
    from sde_mc_amd.batch import job_class
    m1 = J.job_for(J.merton_sde(), 1.0, 10, 100)
    m1a = J.job_for(J.merton_sde(), 1.0, 10, 100, grid='adapted')
    m1c = J.job_for(J.merton_sde(), 1.0, 10, 100, coarse_factor=2)
    l2 = J.job_for(J.levy_rainbow_sde(), 1.0, 50, 100)  # lambda h ~ 0.5: per-step counts
    l2ac = J.job_for(J.levy_rainbow_sde(), 1.0, 10, 100, grid='adapted', coarse_factor=2)
    m3 = J.job_for(J.merton_sde(dim=3), 1.0, 10, 100)
    l2k = J.job_for(J.levy_rainbow_sde(), 1.0, 800, 100)  # lambda h ~ 0.03: arrival clock
    assert [job_class(j) for j in (m1, m1a, m1c, l2, l2ac, m3, l2k)] == [0, 2, 1, 4, 7, 9, 8]
    bad = J.job_for(J.merton_sde(), 1.0, 10, 100)
    bad.steps = 0
    with pytest.raises(L.SdemcError):
        job_class(bad)


def test_autotune_prefers_in_turn_groups_within_the_margin(monkeypatch):
    from types import SimpleNamespace
    from sde_mc_amd import autotune as A
    from sde_mc_amd.batch import BatchTune

    class NoRunner:
        def __init__(self, **kw):
            pass

        def close(self):
            pass

    side, turn, other = BatchTune(4, 4, 2, 1, 1), BatchTune(2, 16, 1, 0, 1), BatchTune(8, 2, 1, 1, 1)
    sweep, confirm = {}, {}
    monkeypatch.setattr(A, 'BatchRunner', NoRunner)
    monkeypatch.setattr(A, 'job_class', lambda j: 0)
    monkeypatch.setattr(A, '_device_key', lambda: 'host')
    monkeypatch.setattr(A, 'time_tune', lambda r, jobs, t, reps=2: sweep[t.key()])
    monkeypatch.setattr(A, 'time_tunes_interleaved',
                        lambda r, jobs, keys, rounds=3: [(k, confirm[k]) for k in keys])
    jobs = [SimpleNamespace(n_paths=1 << 12)]
    cands = [side, turn, other]

    def pick(times, confirm_times=None):
        sweep.clear()
        sweep.update({t.key(): ms for t, ms in zip(cands, times)})
        kw = {}
        if confirm_times is not None:
            confirm.clear()
            for t, ms in zip(cands, confirm_times):
                confirm[t.key()] = ms
                flip = t.key()[:3] + (1 - t.key()[3],) + t.key()[4:]
                confirm.setdefault(flip, 99.0)
            kw = dict(confirm_jobs=jobs, confirm_top=3)
        return A.autotune(jobs, candidates=cands, use_cache=False, **kw)[0].key()

    assert A.CONCURRENT_MARGIN == 0.02
    assert pick([10.0, 10.15, 11.0]) == turn.key()   # 1.5% behind: in turn kept
    assert pick([10.0, 10.3, 11.0]) == side.key()    # 3% behind: side by side kept
    assert pick([11.0, 12.0, 10.0]) == other.key()   # no in-turn shape close
    # the confirm stage's times decide, not the sweep's
    assert pick([10.0, 10.3, 11.0], [10.0, 10.1, 12.0]) == turn.key()
    assert pick([10.0, 10.1, 11.0], [10.0, 10.5, 12.0]) == side.key()


def test_server_keeps_the_tuned_step_unroll():
    # served rounds run the tuned shape with one-chunk tiles; the step unroll
    # (and every other field) is the tune's
    from sde_mc_amd.batch import BatchTune
    from sde_mc_amd.serve import server_tune
    t = server_tune(BatchTune(2, 16, 2, 0, 2))
    assert t.key() == (2, 16, 1, 0, 2)
    assert server_tune(BatchTune(4, 7, 1, 1, 1), tile_chunks=2).key() == (4, 7, 2, 1, 1)
    assert server_tune(None).key() == BatchTune().key()[:2] + (1,) + BatchTune().key()[3:]


@pytest.mark.parametrize("flags,status,text", [
    (L.FLAG_OVERFLOW, L.ERR_OVERFLOW, ''),
    (L.FLAG_JUMP_CSR, L.ERR_ARG, 'jump CSR'),
    (L.FLAG_CLASS, L.ERR_UNSUPPORTED, 'class'),
    (L.FLAG_TABLE, L.ERR_ARG, 'segment table'),
    (L.FLAG_TABLE | L.FLAG_OVERFLOW, L.ERR_ARG, 'segment table'),
    (float('nan'), L.ERR_ARG, 'never written'),
    (64.0, L.ERR_ARG, 'unknown'),
])
def test_statistics_flags_raise(flags, status, text):
    from sde_mc_amd import engine as E
    with pytest.raises(L.SdemcError) as e:
        E.check_flags(flags, 'job')
    assert e.value.status == status and text in str(e.value)
    E.check_flags(0.0, 'job')


def test_autotune_cache_key_follows_the_build(monkeypatch, tmp_path):
    # a cached pick is keyed by the ABI version, the built library and the
    # candidate list: a rebuild or a new candidate list retunes
    import os
    from sde_mc_amd import autotune as A
    from sde_mc_amd.batch import BatchTune
    lib = tmp_path / 'libsdemc.so'
    lib.write_bytes(b'x' * 10)
    monkeypatch.setattr(L, 'lib_path', lambda: str(lib))
    k1 = A._build_key(A.CANDIDATES)
    assert 'abi%d' % L.ABI_VERSION in k1
    os.utime(lib, (1, 12345))
    k2 = A._build_key(A.CANDIDATES)
    assert k2 != k1
    assert A._build_key(A.CANDIDATES[:5]) != k2
    assert A._build_key([BatchTune(2, 16, 1, 0, 1)]) != A._build_key([BatchTune(2, 16, 1, 0, 2)])


def test_plan_runs_a_lone_request_whole_and_splits_shared_work():
    # the round budget (cost_rate x round_s = 1000 cost units) splits work
    # only while something else needs the device: a request alone, with none
    # waiting for admission, takes all its remaining chunks in one round
    from collections import deque
    from sde_mc_amd.serve import PricingServer, _Active
    srv = object.__new__(PricingServer)
    srv.plain_single, srv.policy, srv.queue = True, 'srpt', deque()
    srv.cost_rate, srv.round_s, srv.fill_chunks = 1e6, 1e-3, 4

    def active(slot, nc, cost):
        a = _Active()
        a.slot, a.nc, a.next, a.cost, a.admitted = slot, nc, 0, cost, float(slot)
        return a

    big = active(0, 100, 50.0)  # 5000 units
    srv.active = {0: big}
    assert srv._plan() == ([0], [0], [100], [50.0], [0]) and big.next == 100
    # with a second request the shortest goes first and the budget cuts the rest
    big.next = 0
    small = active(1, 2, 50.0)
    srv.active = {0: big, 1: small}
    slots, firsts, counts, _, final = srv._plan()
    assert (slots, firsts, counts, final) == ([1, 0], [0, 0], [2, 18], [1]) and big.next == 18
    # alone again after a shared round: the rest in one round
    srv.active = {0: big}
    assert srv._plan() == ([0], [18], [82], [50.0], [0])
    # a request waiting for admission keeps the budget in force, and a round
    # below the device fill is topped up in the same order
    big.next = 0
    srv.queue.append(object())
    assert srv._plan() == ([0], [0], [20], [50.0], [])
    srv.fill_chunks = 30
    assert srv._plan() == ([0], [20], [30], [50.0], [])


def test_serving_loops_pause_the_cyclic_collector():
    # a full collection stops the host for tens of milliseconds; the serving
    # loops run with the cyclic collector off and restore it after, also
    # when the loop raises
    import gc
    from sde_mc_amd.serve import no_cyclic_gc
    assert gc.isenabled()
    with no_cyclic_gc():
        assert not gc.isenabled()
    assert gc.isenabled()
    with pytest.raises(RuntimeError):
        with no_cyclic_gc():
            raise RuntimeError('x')
    assert gc.isenabled()
    gc.disable()
    try:
        with no_cyclic_gc():
            pass
        assert not gc.isenabled()  # left as found
    finally:
        gc.enable()

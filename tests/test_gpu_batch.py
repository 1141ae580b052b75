"""The batched persistent launch: each job's statistics equal its plain launch's
bit for bit, whatever it is packed with, whatever the tile size, and on
every replay of the captured graph."""
import math

import numpy as np
import pytest
import torch

import sde_mc_amd as sm
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from sde_mc_amd.batch import BatchRunner, BatchTune
from tests import jobs as J

pytestmark = pytest.mark.gpu


def _mix(n=12, seed=3):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        P = int(2 ** rng.uniform(6, 16))
        N = int(2 ** rng.uniform(3, 8))
        grid = 'adapted' if i % 3 == 0 else 'uniform'
        if i % 2 == 0:
            sde, opt = J.merton_sde(), sm.EuroCall(float(rng.uniform(0.8, 1.2)))
        else:
            sde, opt = J.levy_rainbow_sde(), sm.Rainbow(1.0, log=True, discount=math.exp(0.02))
        M = 2 if (i % 4 == 1 and N % 2 == 0) else 0
        jobs.append(J.job_for(sde, float(rng.choice([0.25, 0.5, 1.0, 2.0])), N, P, grid=grid,
                              coarse_factor=M, option=opt, disc=0.98, seed=1000 + i))
    return jobs


def _plain(job):
    return E.simulate(job, E.require_device()).cpu().numpy()


def test_batch_equals_plain_bitwise():
    from sde_mc_amd.batch import job_class
    jobs = _mix()
    # the uniform single-level classes of both jump modes are in the mix:
    # 1-D clock, 2-D per-step counts, 2-D clock
    assert {0, 4, L.CLASS_INDEP2_CLOCK} <= {job_class(j) for j in jobs}
    plain = np.array([_plain(j) for j in jobs])
    r = BatchRunner(max_jobs=64)
    for tune in (BatchTune(4, 2, 1), BatchTune(1, 4, 4), BatchTune(8, 1, 2),
                 BatchTune(4, 7, 1, concurrent_groups=0), BatchTune(2, 14, 2),
                 BatchTune(4, 8, 1, step_unroll=2), BatchTune(8, 4, 2, 0, 2)):
        got = r.run(jobs, tune=tune)
        assert np.array_equal(got, plain), tune


def test_job_alone_equals_job_in_mix():
    jobs = _mix(10, seed=4)
    r = BatchRunner(max_jobs=64)
    mixed = r.run(jobs)
    for i, j in enumerate(jobs):
        alone = r.run([j])
        assert np.array_equal(alone[0], mixed[i])


def test_graph_replays_are_consistent():
    jobs = _mix(8, seed=5)
    other = _mix(6, seed=9)
    r = BatchRunner(max_jobs=64, use_graph=True)
    r2 = BatchRunner(max_jobs=64, use_graph=False)
    other_eager = r2.run(other)
    first = r.run(jobs)
    for i in range(4):
        torch.cuda.synchronize()
        again = r.run(jobs)
        assert np.array_equal(again, first), i
        # eager device work between replays: a plain launch, a torch
        # reduction, and another job set through the same captured plans
        assert np.array_equal(_plain(other[i]), other_eager[i])
        torch.arange(1 << 20, device=r.device, dtype=torch.float64).sum().item()
        assert np.array_equal(r.run(other), other_eager), i
    assert np.array_equal(r2.run(jobs), first)
    r3 = BatchRunner(max_jobs=64, use_graph=False, tune=BatchTune(concurrent_groups=0))
    assert np.array_equal(r3.run(jobs), first)


def test_jump_queue_overflow_and_long_windows_equal_plain():
    # the arrival clock's LDS queue at its edges, in the batched class kernels
    # against the plain ones: several jumps per lane and step (most fill
    # windows overflow the queue and draw in the step), single and coupled
    # levels, both jump rules; and a low rate whose window spans the horizon
    jobs = [J.job_for(J.merton_sde(rate=30.0), 1.0, 20, 3000, option=sm.EuroCall(1.0), seed=11),
            J.job_for(J.merton_sde(rate=30.0), 1.0, 40, 3000, coarse_factor=2,
                      option=sm.EuroCall(1.0), seed=12, exact_jumps=True),
            J.job_for(J.levy_rainbow_sde(f=(0.2, 0.2), eps=0.02), 1.0, 1000, 2000,
                      option=sm.Rainbow(1.0, log=True), seed=13),
            J.job_for(J.merton_sde(rate=0.05), 0.25, 700, 5000, option=sm.EuroCall(1.0), seed=14)]
    plain = np.array([_plain(j) for j in jobs])
    assert np.array_equal(BatchRunner(max_jobs=8).run(jobs), plain)
    assert np.all(plain[:, 3] == 0)


def test_generic_class_with_dedicated_classes():
    # a 3-D model runs in the generic kernel, next to a dedicated 1-D group
    corr = torch.tensor([[1.0, 0.2, 0.1], [0.2, 1.0, 0.3], [0.1, 0.3, 1.0]])
    jobs = [J.job_for(J.merton_sde(dim=3, corr=corr), 1.0, 32, 3000, option=sm.Rainbow(1.0), seed=7),
            J.job_for(J.merton_sde(), 1.0, 40, 5000, option=sm.EuroCall(1.0), seed=8)]
    plain = np.array([_plain(j) for j in jobs])
    assert np.array_equal(BatchRunner(max_jobs=8).run(jobs), plain)


def test_class_mismatch_is_flagged_not_silent():
    # a coupled job planned into the single-level kernel's group is refused by
    # the host check before upload and, staged anyway, by the kernel (NaN
    # partials, flag 4), which the host reports
    from sde_mc_amd import _lib as L
    from sde_mc_amd.batch import plan_segments
    job = J.job_for(J.merton_sde(), 1.0, 32, 500, coarse_factor=2, option=sm.EuroCall(1.0))
    r = BatchRunner(max_jobs=4)
    assert list(r.prepare([job], [0], [0])) == [1]
    r.upload_slots([0])
    plan = plan_segments([0], [0], [8], [1.0], 1, [0])
    buf = r.bufs[0]
    with pytest.raises(L.SdemcError):
        r.stage_round(buf, plan, [0])
    r.stage_round(buf, plan, [0], check=False)
    r.launch(buf, plan.groups)
    r.read_stats(buf, [0])
    buf.done.synchronize()
    st = buf.h_stats[:4].numpy()
    assert int(st[3]) & 4 and np.isnan(st[0])
    with pytest.raises(L.SdemcError):
        E.check_flags(st[3], 'job')


def test_set_chunk_flags_do_not_outlive_their_round():
    # the group kernels write a chunk's flag only when it is set and the
    # finalisation resets the flags it reads: a round whose job overflows its
    # adapted-grid cap is flagged, and the next job on the same partial slots
    # is clean and equals its plain launch
    from sde_mc_amd import _lib as L
    r = BatchRunner(max_jobs=4)
    bad = J.job_for(J.merton_sde(rate=30.0), 1.0, 20, 4000, grid='adapted', max_steps=22,
                    option=sm.EuroCall(1.0), seed=5)
    with pytest.raises(L.SdemcError, match='max_steps'):
        r.run([bad])
    good = J.job_for(J.merton_sde(), 1.0, 20, 4000, grid='adapted', option=sm.EuroCall(1.0),
                     seed=6)
    again = r.run([good])
    assert again[0, 3] == 0
    assert np.array_equal(again[0], np.array(_plain(good)))


def test_autotune_picks_a_candidate_and_it_prices_the_same():
    from sde_mc_amd.autotune import autotune
    jobs = _mix(6, seed=11)
    cands = [BatchTune(2, 4, 1), BatchTune(4, 8, 2)]
    tune, table = autotune(jobs, candidates=cands, use_cache=False)
    assert tune.key() in [c.key() for c in cands] and len(table) == 2
    plain = np.array([_plain(j) for j in jobs])
    assert np.array_equal(BatchRunner(max_jobs=16, tune=tune).run(jobs), plain)
    # confirming the fastest shapes on the real batch keeps one of them, its
    # class groups side by side or in turn
    cands3 = cands + [BatchTune(4, 4, 2, 1, 2)]
    tune2, _ = autotune(jobs[:3], candidates=cands3, use_cache=False, confirm_jobs=jobs,
                        confirm_top=2)
    keys = [c.key() for c in cands3]
    assert tune2.key() in keys + [k[:3] + (1 - k[3],) + k[4:] for k in keys]
    assert np.array_equal(BatchRunner(max_jobs=16, tune=tune2).run(jobs), plain)


def _misplanned_round(mutate):
    """Run three jobs once (slots 0-2 prepared), then launch a round whose
    segment table `mutate` spoiled, staged without the host check; returns
    the finalised statistics [3, 4] and whether the host check refused it."""
    from sde_mc_amd.batch import chunks_of, job_cost, plan_segments
    jobs = [J.job_for(J.merton_sde(), 1.0, 32, 5000, option=sm.EuroCall(1.0), seed=61),
            J.job_for(J.merton_sde(), 1.0, 16, 700, option=sm.EuroCall(1.0), seed=62),
            J.job_for(J.merton_sde(), 1.0, 24, 3000, option=sm.EuroCall(0.9), seed=63)]
    r = BatchRunner(max_jobs=8, use_graph=False)
    good = r.run(jobs)
    counts = np.array([chunks_of(j.n_paths) for j in jobs], dtype=np.int64)
    plan = plan_segments(np.arange(3), np.zeros(3, np.int64), counts,
                         np.array([job_cost(j) for j in jobs]), 1, r.slot_class[:3],
                         min_tile_cost=1.0)
    assert plan.groups == (0,) and plan.hdr[0, 1] == 3
    mutate(plan)
    buf = r.bufs[0]
    try:
        r.stage_round(buf, plan, range(3))
        refused = False
    except L.SdemcError:
        refused = True
    r.stage_round(buf, plan, range(3), check=False)
    print('misplanned round: segments %s header %s' % (plan.segs.tolist(), plan.hdr[0].tolist()),
          flush=True)
    r.launch(buf, plan.groups)
    torch.cuda.synchronize()
    return good, r.stats[:12].cpu().numpy().reshape(3, 4), refused, (r, jobs)


def _set(row, col, delta):
    def f(plan):
        plan.segs[row, col] += delta
    return f


@pytest.mark.parametrize("name,mutate", [
    ('unprepared slot', lambda p: p.segs.__setitem__((1, 0), 5)),
    ('chunks past the job', _set(0, 2, 1)),
    ('tile ranges overlap', _set(2, 4, -1)),
    ('tile count short', lambda p: p.hdr.__setitem__((0, 2), p.hdr[0, 2] - 1)),
])
def test_misplanned_round_is_flagged_not_stale(name, mutate):
    # a segment table that does not describe the job table is refused by the
    # host check before upload and, launched anyway, runs only in-bounds work
    # and flags every slot it finalises: no round returns the previous
    # round's statistics as its result
    good, st, refused, _ = _misplanned_round(mutate)
    assert refused, name
    assert np.all(good[:, 3] == 0)
    for i in range(3):
        assert int(st[i, 3]) & L.FLAG_TABLE, (name, i, st[i])
        with pytest.raises(L.SdemcError):
            E.check_flags(st[i, 3], 'slot %d' % i)


def test_wellplanned_round_after_a_misplanned_one_is_clean():
    # the fault word is per launch: the next good round carries no flag
    good, st, _, (r, jobs) = _misplanned_round(_set(0, 2, 1))
    assert int(st[0, 3]) & L.FLAG_TABLE
    again = r.run(jobs)
    assert np.all(again[:, 3] == 0) and np.array_equal(again, good)

"""Batched persistent launches: many jobs (uneven path and step counts, mixed
models, MLMC levels) in one kernel launch.

A :class:`BatchRunner` owns a device job table (prepared slots), a pool of
chunk partials, per-slot statistics and one or more round buffers.  A round
buffer holds a segment list -- (slot, first chunk, chunks, chunks per tile,
first tile) -- grouped by kernel class, and the mask of slots to finalise.
Each class group runs in its own persistent kernel, specialised to one
(model family, time grid, single level or coupled pair) exactly like the
plain launch's kernel for it, or the generic kernel for anything else; its
wavefronts take tile indices from the group's device counter until the
group's tiles are exhausted.  The launches of a round (counter reset, the
group kernels, finalisation) are captured once per (round buffer, class
groups) in a HIP graph and replayed; the host only uploads the small
segment list and, for new jobs, their slots.

A job's result is independent of what it is packed with or how it is cut:
the Philox streams are keyed by (seed, path) and the reduction order is
fixed per job (64-path chunks, then chunks in order), so a job alone and in
any mix gives the same bits.
"""
import ctypes
import itertools

import numpy as np
import torch

from . import _lib as L
from . import engine as E

NS = L.SEGMENT_INTS


def chunks_of(n_paths):
    return (int(n_paths) + L.CHUNK - 1) // L.CHUNK


def noise_width(model):
    """Normals drawn per step."""
    d = model.dim
    if model.noise in (L.NOISE_DIAG, L.NOISE_HESTON):
        return d
    return d + 1 if model.noise == L.NOISE_INDEP_COMMON else 2 * d


# Issue cost (VALU slots per wavefront step) of one normal and of a
# wavefront-wide jump draw, fitted to the measured per-model throughput of the
# plain kernels (1-D Merton at 500 steps and the 2-D Levy rainbow at 200);
# the uniform grid's arrival clock, for either jump law, at the time of the
# served mix's 2-D Levy clock-mode kernel (its cost per unit then equals the
# Merton and per-step-count kernels', so a round's cost predicts its time
# whatever its mix of classes).
SLOTS_PER_NORMAL = 60.0
SLOTS_PER_JUMP_STEP = {L.JUMP_LOGNORMAL: 280.0, L.JUMP_LEVY: 475.0}
SLOTS_PER_CLOCK_JUMP_STEP = 280.0
POIS_MIN_MU = 1.0 / 16.0  # the uniform grid's Levy jumps switch to per-step counts here


def job_cost(job):
    """Cost of one chunk (one wavefront of 64 paths) of a job in VALU issue
    slots: per step the normals drawn plus the jump work a wavefront sees (a
    jump draw whenever some lane jumps), x(1 + 1/M) for coupled pairs."""
    m = job.model
    mu = m.rate * job.T / job.steps
    wave_jumps = min(1.0, 64.0 * mu)
    per_jump = SLOTS_PER_JUMP_STEP.get(m.jump_law, 300.0)
    if job.grid == L.GRID_UNIFORM and not (m.jump_law == L.JUMP_LEVY and mu >= POIS_MIN_MU):
        per_jump = SLOTS_PER_CLOCK_JUMP_STEP
    c = job.steps * (SLOTS_PER_NORMAL * noise_width(m) + per_jump * wave_jumps)
    if job.coarse_factor:
        c *= 1.0 + 1.0 / job.coarse_factor
    return c


def job_class(job):
    """Kernel class of a job (sdemc_job_class): 0..7 are the dedicated kernels
    of the 1-D 'diag' (0..3) and 2-D 'indep' (4..7) families, 2 * grid +
    coupled within each, 8 the 2-D uniform single-level kernel for the
    arrival clock (4 takes per-step Poisson counts); 9 is the generic kernel
    that runs every job."""
    c = L.lib().sdemc_job_class(job)
    if c < 0:
        raise L.SdemcError(L.ERR_ARG, 'sdemc_job_class')
    return int(c)


class BatchTune:
    """Launch shape of the persistent kernels: wavefronts per workgroup (paths
    per workgroup = 64 x that), resident workgroups per CU, least chunks (64
    paths each) per tile, whether a round's class groups run concurrently,
    and the step unroll (4-step groups per loop iteration, 1 or 2) of the
    uniform single-level kernels."""

    def __init__(self, waves_per_block=4, blocks_per_cu=7, tile_chunks=1, concurrent_groups=1,
                 step_unroll=1):
        self.waves_per_block = int(waves_per_block)
        self.blocks_per_cu = int(blocks_per_cu)
        self.tile_chunks = int(tile_chunks)
        self.concurrent_groups = int(concurrent_groups)
        self.step_unroll = int(step_unroll)

    def c(self):
        t = L.BatchTune()
        t.waves_per_block = self.waves_per_block
        t.blocks_per_cu = self.blocks_per_cu
        t.concurrent_groups = self.concurrent_groups
        t.step_unroll = self.step_unroll
        return t

    def key(self):
        return (self.waves_per_block, self.blocks_per_cu, self.tile_chunks, self.concurrent_groups,
                self.step_unroll)

    def __repr__(self):
        return ('BatchTune(wpb=%d, bpc=%d, tile_chunks=%d, concurrent_groups=%d, step_unroll=%d)'
                % self.key())


class RoundPlan:
    """One round's segment table grouped by kernel class.

    segs   [n, 6] int32: (slot, first chunk, chunks, chunks per tile, first
           tile within its group, 0), each group's rows contiguous and
           costliest chunk first;
    hdr    [MAX_GROUPS, 4] int32: (first segment, segments, tiles, 0);
    groups the kernel class of each group, in launch order;
    tiles  tiles over all groups."""

    __slots__ = ('segs', 'hdr', 'groups', 'tiles')

    def __init__(self, segs, hdr, groups, tiles):
        self.segs, self.hdr, self.groups, self.tiles = segs, hdr, tuple(groups), int(tiles)


# Least work of one tile, in job_cost units (about 300 steps of a 1-D chunk):
# a wavefront takes a tile index from its group's counter with one atomic, so
# short jobs (e.g. MLMC level 0, 4 steps) get several chunks per tile.
MIN_TILE_COST = 2.0e4
MAX_TILE_COST = 5.0e4


def tile_cost_floor(total_cost, tune, cus):
    """Least work per tile for a launch of `total_cost` (job_cost units): a
    large launch can afford bigger tiles (fewer counter atomics and segment
    searches: the served mix runs 2.5% faster at 5e4 than at 2e4 with one
    chunk per tile), a small one needs many tiles per resident wavefront to
    fill the device (the MLMC level set, 3e8 in all, runs 20% slower at
    5e4); so the floor is the launch's work over 64 tiles per resident
    wavefront, kept within [MIN_TILE_COST, MAX_TILE_COST]."""
    waves = max(1, int(cus) * tune.waves_per_block * tune.blocks_per_cu)
    return float(min(MAX_TILE_COST, max(MIN_TILE_COST, total_cost / (64.0 * waves))))


def plan_segments(slots, firsts, counts, costs, tile_chunks, classes=None,
                  min_tile_cost=MIN_TILE_COST):
    """Group the segments by kernel class (generic when `classes` is None) and
    lay out consecutive tile ranges within each group; returns a RoundPlan.
    A segment's tiles hold max(tile_chunks, ceil(min_tile_cost / cost))
    chunks.  Groups launch one after another, the costliest group first; when
    more classes are present than a launch has groups, the cheapest dedicated
    groups fold into the generic one."""
    slots = np.asarray(slots, dtype=np.int64)
    firsts = np.asarray(firsts, dtype=np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    costs = np.asarray(costs, dtype=np.float64)
    if classes is None:
        cls = np.full(slots.size, L.CLASS_GENERIC, dtype=np.int64)
    else:
        cls = np.asarray(classes, dtype=np.int64).copy()
    keep = counts > 0
    slots, firsts, counts, costs, cls = slots[keep], firsts[keep], counts[keep], costs[keep], cls[keep]
    if np.any((cls < 0) | (cls > L.CLASS_GENERIC)):
        raise ValueError('kernel class out of range')
    work = {int(c): float((counts[cls == c] * costs[cls == c]).sum()) for c in np.unique(cls)}
    while len(work) > L.MAX_GROUPS:
        c = min((k for k in work if k != L.CLASS_GENERIC), key=lambda k: work[k])
        cls[cls == c] = L.CLASS_GENERIC
        work[L.CLASS_GENERIC] = work.get(L.CLASS_GENERIC, 0.0) + work.pop(c)
    groups = sorted(work, key=lambda k: (-work[k], k))
    rank = np.zeros(L.CLASS_GENERIC + 1, dtype=np.int64)
    for g, c in enumerate(groups):
        rank[c] = g
    order = np.lexsort((-costs, rank[cls])) if slots.size else np.zeros(0, np.int64)
    per = np.maximum(max(1, int(tile_chunks)),
                     np.ceil(float(min_tile_cost) / np.maximum(costs[order], 1e-30))).astype(np.int64)
    per = np.minimum(per, 1 << 20)
    nt = (counts[order] + per - 1) // per
    grp = rank[cls[order]]
    hdr = np.zeros((L.MAX_GROUPS, 4), dtype=np.int32)
    begin = np.zeros(order.size, dtype=np.int64)
    for g in range(len(groups)):
        idx = np.nonzero(grp == g)[0]
        t = nt[idx]
        begin[idx[1:]] = np.cumsum(t)[:-1]
        total = int(t.sum())
        if total >= 2 ** 31:
            raise ValueError('too many tiles in one launch')
        hdr[g] = (idx[0], idx.size, total, 0)
    seg = np.zeros((order.size, NS), dtype=np.int32)
    seg[:, 0] = slots[order]
    seg[:, 1] = firsts[order]
    seg[:, 2] = counts[order]
    seg[:, 3] = per
    seg[:, 4] = begin
    return RoundPlan(seg, hdr, groups, int(nt.sum()) if order.size else 0)


class RoundBuffer:
    """Device + pinned staging for one round's segment list, finalise mask and
    statistics read-back."""

    def __init__(self, device, max_jobs, max_segments):
        self.max_segments = int(max_segments)
        self.segs = torch.zeros(self.max_segments * NS, dtype=torch.int32, device=device)
        self.hdr = torch.zeros(4 * L.MAX_GROUPS, dtype=torch.int32, device=device)
        self.final = torch.zeros(max_jobs, dtype=torch.int32, device=device)
        self.h_segs = torch.zeros(self.max_segments * NS, dtype=torch.int32, pin_memory=True)
        self.h_hdr = torch.zeros(4 * L.MAX_GROUPS, dtype=torch.int32, pin_memory=True)
        self.h_final = torch.zeros(max_jobs, dtype=torch.int32, pin_memory=True)
        self.h_stats = torch.zeros(max_jobs * 4, dtype=torch.float64, pin_memory=True)
        self.done = None  # event recorded after the round's read-back


class BatchRunner:
    """Device-resident job table + partial pool + persistent launches."""

    def __init__(self, max_jobs=1024, max_segments=None, partial_chunks=1 << 18, tune=None,
                 use_graph=True, n_buffers=1):
        self.device = E.require_device()
        self.cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.lib = L.lib()
        self.kb = self.lib.sdemc_sizeof_prepared_job()
        self.max_jobs = int(max_jobs)
        self.max_segments = int(max_segments or max(64, 2 * self.max_jobs))
        self.tune = tune or BatchTune()
        self.use_graph = use_graph
        dev = self.device
        self.jobs = torch.zeros(self.max_jobs * self.kb, dtype=torch.uint8, device=dev)
        self.h_jobs = torch.zeros(self.max_jobs * self.kb, dtype=torch.uint8, pin_memory=True)
        self.stats = torch.zeros(self.max_jobs * 4, dtype=torch.float64, device=dev)
        # the groups' tile counters, then the launch's fault word
        self.counter = torch.zeros(L.MAX_GROUPS + 1, dtype=torch.int32, device=dev)
        self.slot_class = np.full(self.max_jobs, L.CLASS_GENERIC, dtype=np.int32)
        self.bufs = [RoundBuffer(dev, self.max_jobs, self.max_segments) for _ in range(n_buffers)]
        self._plans = {}
        self._alloc_partials(partial_chunks)

    def _alloc_partials(self, n):
        self.close()
        self.partial_chunks = int(n)
        self.partials = torch.zeros(self.partial_chunks * 2, dtype=torch.float64, device=self.device)
        self.cflags = torch.zeros(self.partial_chunks, dtype=torch.int32, device=self.device)

    def close(self):
        for p in getattr(self, '_plans', {}).values():
            self.lib.sdemc_batch_plan_destroy(p)
        self._plans = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def batch_dev(self, buf, groups):
        b = L.BatchDev()
        b.jobs_dev = self.jobs.data_ptr()
        b.segs_dev = buf.segs.data_ptr()
        b.seg_hdr_dev = buf.hdr.data_ptr()
        b.finalize_dev = buf.final.data_ptr()
        b.partials_dev = self.partials.data_ptr()
        b.chunk_flags_dev = self.cflags.data_ptr()
        b.stats_dev = self.stats.data_ptr()
        b.counter_dev = self.counter.data_ptr()
        b.partial_chunks = self.partial_chunks
        b.max_jobs = self.max_jobs
        b.max_segments = buf.max_segments
        if not 1 <= len(groups) <= L.MAX_GROUPS:
            raise ValueError('a launch runs 1..%d class groups' % L.MAX_GROUPS)
        b.n_groups = len(groups)
        for g, c in enumerate(groups):
            b.group_class[g] = int(c)
        return b

    # -- tables ----------------------------------------------------------
    def prepare(self, jobs, chunk_base, slots):
        """Validate jobs and stage their prepared bytes into slots (pinned host
        table); returns their kernel classes."""
        n = len(jobs)
        if n == 0:
            return np.zeros(0, dtype=np.int32)
        arr = (L.Job * n)(*jobs)
        base = (ctypes.c_int64 * n)(*[int(b) for b in chunk_base])
        buf = ctypes.create_string_buffer(n * self.kb)
        cls = (ctypes.c_int32 * n)()
        L.check(self.lib.sdemc_prepare_jobs(arr, n, base, buf, cls), 'sdemc_prepare_jobs')
        raw = np.frombuffer(buf.raw, dtype=np.uint8).reshape(n, self.kb)
        hj = self.h_jobs.numpy().reshape(self.max_jobs, self.kb)
        sl = np.asarray(slots, dtype=np.int64)
        hj[sl] = raw
        out = np.frombuffer(cls, dtype=np.int32).copy()
        self.slot_class[sl] = out
        return out

    def upload_slots(self, slots=None):
        """H2D copy of staged slots (all when None), ordered on the current stream."""
        if slots is None:
            self.jobs.copy_(self.h_jobs, non_blocking=True)
            return
        slots = sorted(int(s) for s in slots)
        if not slots:
            return
        # one copy of the span: the copies are ordered on the stream after the
        # rounds already queued, and slots inside the span that are not
        # re-staged get the bytes they already hold, so a round's host cost
        # does not grow with the number of its new requests
        a, b = slots[0] * self.kb, (slots[-1] + 1) * self.kb
        self.jobs[a:b].copy_(self.h_jobs[a:b], non_blocking=True)

    def check_tables(self, buf, groups):
        """sdemc_batch_check of the staged round against the host copy of the
        job table: raises SdemcError before a launch could run a mis-planned
        segment list."""
        groups = tuple(int(g) for g in groups) or (L.CLASS_GENERIC,)
        gc = (ctypes.c_int32 * len(groups))(*groups)
        L.check(self.lib.sdemc_batch_check(
            self.h_jobs.data_ptr(), self.max_jobs, buf.h_segs.data_ptr(), buf.max_segments,
            buf.h_hdr.data_ptr(), len(groups), gc, self.partial_chunks), 'sdemc_batch_check')

    def stage_round(self, buf, plan, finalize_slots, check=True):
        """Stage a round's segment list, header and finalise mask into `buf`
        (pinned host copy, then H2D on the current stream), after
        sdemc_batch_check of the tables; check=False skips the host check (the
        device then flags what the check would have refused)."""
        segs = plan.segs
        n = len(segs)
        if n > buf.max_segments:
            raise ValueError('%d segments exceed max_segments %d' % (n, buf.max_segments))
        if n:
            buf.h_segs.numpy()[:NS * n] = segs.reshape(-1)
        buf.h_hdr.numpy()[:] = plan.hdr.reshape(-1)
        if check:
            self.check_tables(buf, plan.groups)
        if n:
            buf.segs[:NS * n].copy_(buf.h_segs[:NS * n], non_blocking=True)
        buf.hdr.copy_(buf.h_hdr, non_blocking=True)
        hf = buf.h_final.numpy()
        hf[:] = 0
        fs = np.asarray(list(finalize_slots), dtype=np.int64)
        if fs.size:
            hf[fs] = 1
        buf.final.copy_(buf.h_final, non_blocking=True)

    # -- launch ----------------------------------------------------------
    def launch(self, buf, groups, tune=None):
        """Launch the round staged in `buf`: one persistent kernel per class
        group in `groups` (a RoundPlan's groups), then the finalisation."""
        tune = tune or self.tune
        groups = tuple(int(g) for g in groups) or (L.CLASS_GENERIC,)
        stream = E.stream_ptr()
        if self.use_graph:
            plan = self._graph(buf, groups, tune)
            L.check(self.lib.sdemc_batch_plan_launch(plan, stream), 'sdemc_batch_plan_launch')
        else:
            L.check(self.lib.sdemc_batch_launch(ctypes.byref(self.batch_dev(buf, groups)),
                                                ctypes.byref(tune.c()), stream),
                    'sdemc_batch_launch')

    def _graph(self, buf, groups, tune):
        """The captured graph of a round in `buf` with these class groups and
        launch shape, captured on first use."""
        key = (id(buf), groups, tune.key(), self.partials.data_ptr())
        plan = self._plans.get(key)
        if plan is None:
            h = ctypes.c_void_p()
            L.check(self.lib.sdemc_batch_plan_create(
                ctypes.byref(self.batch_dev(buf, groups)), ctypes.byref(tune.c()),
                ctypes.byref(h)), 'sdemc_batch_plan_create')
            plan = h
            self._plans[key] = plan
        return plan

    def warm(self, classes, tune=None):
        """Capture the graph of every round the given kernel classes can form
        (each ordered selection of them, as a round lists its groups, on
        every round buffer), so that no later round pays for a capture: a
        serving loop's start-up.  Returns the number of graphs held."""
        if self.use_graph:
            tune = tune or self.tune
            cls = sorted(set(int(c) for c in classes))
            for r in range(1, min(len(cls), L.MAX_GROUPS) + 1):
                for groups in itertools.permutations(cls, r):
                    for buf in self.bufs:
                        self._graph(buf, groups, tune)
        return len(self._plans)

    def read_stats(self, buf, slots):
        """Queue the D2H copy of the finalised slots' statistics into the
        round's pinned buffer and record its completion event."""
        slots = [int(s) for s in slots]
        if slots:
            # one copy of the span of the finalised slots (the others' rows
            # come along and are not read)
            a, b = 4 * min(slots), 4 * (max(slots) + 1)
            buf.h_stats[a:b].copy_(self.stats[a:b], non_blocking=True)
        buf.done = torch.cuda.Event(enable_timing=True)
        buf.done.record()

    def run(self, jobs, tune=None, sync=True, events=None):
        """Price `jobs` (list of L.Job) in one batched launch; returns [n, 4] fp64
        (sum, sum of squares, n, flags) per job."""
        tune = tune or self.tune
        n = len(jobs)
        if n == 0:
            return np.zeros((0, 4))
        if n > self.max_jobs:
            raise ValueError('%d jobs exceed max_jobs %d' % (n, self.max_jobs))
        counts = np.array([chunks_of(j.n_paths) for j in jobs], dtype=np.int64)
        base = np.zeros(n, dtype=np.int64)
        base[1:] = np.cumsum(counts)[:-1]
        total = int(counts.sum())
        if total > self.partial_chunks:
            self._alloc_partials(max(total, 2 * self.partial_chunks))
        slots = np.arange(n)
        cls = self.prepare(jobs, base, slots)
        self.upload_slots(range(n))
        costs = np.array([job_cost(j) for j in jobs])
        floor = tile_cost_floor(float(np.dot(counts, costs)), tune, self.cus)
        plan = plan_segments(slots, np.zeros(n, np.int64), counts, costs, tune.tile_chunks, cls,
                             min_tile_cost=floor)
        buf = self.bufs[0]
        self.stage_round(buf, plan, slots)
        if events is not None:
            events[0].record()
        self.launch(buf, plan.groups, tune)
        if events is not None:
            events[1].record()
        buf.h_stats[:4 * n].copy_(self.stats[:4 * n], non_blocking=True)
        if not sync:
            return None
        torch.cuda.current_stream().synchronize()
        st = buf.h_stats[:4 * n].numpy().reshape(n, 4).copy()
        for i in range(n):
            E.check_flags(st[i, 3], 'batch job %d' % i)
        return st

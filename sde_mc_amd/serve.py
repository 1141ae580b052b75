"""Continuous-batching pricing server on one GPU.

Requests (model, strike, maturity, steps, paths, seed) arrive over time.  A
request is admitted when a job slot is free and its chunk partials fit in
the HBM budget; its paths are then priced in rounds.  Each round is one
batched persistent launch (sde_mc_amd.batch) over chunks of the admitted
requests -- by default the requests with the least work left first, so
small requests finish within a round or two of their arrival while large
ones keep the device busy -- so requests that arrive while others are on the
GPU join the next round and their work shares the device with everything
else in flight.  Two rounds are kept in flight so the host plans round r+1
while round r runs.  A request's partial sums stay on the device until its
last chunk is done; it is then finalised in-kernel and its (sum, sum of
squares, n) is read back.

Results do not depend on the mix: each request's paths are keyed by its own
seed and path index, and its reduction order is fixed (see batch.py), so a
request priced alone and in any mix gives the same bits.
"""
import contextlib
import gc
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import engine as E
from .batch import BatchRunner, BatchTune, chunks_of, job_cost, plan_segments
from .workloads import request_job

PARTIAL_BYTES = 2 * 8 + 4  # one chunk: fp64 (sum, sum of squares) + int32 flag


@contextlib.contextmanager
def no_cyclic_gc():
    """Python's cyclic collector off for the block: a full collection over a
    large heap stops the host for tens of milliseconds (63 ms measured inside a
    2000 req/s stream, which then queued up to a 68 ms p99), and a loop that
    feeds the device round by round must not stop.  Reference counting still
    frees everything that is not in a cycle; cycles wait for the next
    collection after the block."""
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


class ChunkPool:
    """First-fit allocator of contiguous chunk-partial ranges (the HBM budget)."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]  # sorted (start, length)
        self.used = 0

    def alloc(self, n):
        for i, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, ln - n)
                self.used += n
                return s
        return None

    def release(self, start, n):
        self.used -= n
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((s, ln))
        self.free = merged


@dataclass
class Result:
    rid: int
    mean: float
    se: float
    n: int
    arrival: float
    admitted: float
    finished: float
    path_steps: int
    stats: tuple = field(default=None, repr=False)
    error: str = None

    @property
    def latency(self):
        return self.finished - self.arrival


class _Active:
    __slots__ = ('req', 'job', 'slot', 'base', 'nc', 'next', 'cost', 'admitted', 'in_table')


def server_tune(tune, tile_chunks=1):
    """The launch shape of served rounds: `tune` (default BatchTune()) with
    `tile_chunks` chunks per tile; every other field, the step unroll
    included, as tuned."""
    t = tune or BatchTune()
    return BatchTune(t.waves_per_block, t.blocks_per_cu, tile_chunks, t.concurrent_groups,
                     t.step_unroll)


class PricingServer:
    """Admission by slots and HBM budget, rounds of batched launches.

    policy 'srpt' (default) fills each round with the requests that have the
    least work left, each taking all its remaining chunks while the round's
    budget lasts; 'fair' gives the oldest request half the budget, the next
    half of the rest, and so on.  A round that holds a single request from its
    first chunk to its last runs that request's plain kernel instead (the
    same bits, less launch overhead), and a request that is alone -- the only
    one with work left and none waiting for admission -- takes all its
    remaining chunks in one round whatever the budget (its plain kernel if it
    has not started): split into rounds it would only add host gaps while
    nothing else needs the device."""

    def __init__(self, max_slots=256, hbm_budget_bytes=256 << 20, round_ms=4.0, tune=None,
                 grid=L.GRID_UNIFORM, job_builder=None, policy='srpt', plain_single=True,
                 tile_chunks=1, rate_cap=2.5e12):
        if policy not in ('srpt', 'fair'):
            raise ValueError("policy must be 'srpt' or 'fair'")
        self.policy = policy
        self.plain_single = plain_single
        self.pool = ChunkPool(hbm_budget_bytes // PARTIAL_BYTES)
        # rounds are a few ms long, so tiles stay one chunk (a long tile at the
        # end of a round leaves the rest of the device idle)
        t = server_tune(tune, tile_chunks)
        self.runner = BatchRunner(max_jobs=max_slots, max_segments=2 * max_slots,
                                  partial_chunks=self.pool.capacity, tune=t, n_buffers=2)
        self.max_slots = max_slots
        self.free_slots = list(range(max_slots))[::-1]
        self.queue = deque()
        self.active = {}
        self.round_s = round_ms * 1e-3
        self.cost_rate = 2.0e12     # cost units (VALU slots) per second, refined per round
        # upper bound of the refined rate: the served mix settles at 2.0-2.8e12
        # on MI355X, and an overestimate makes every round longer
        self.rate_cap = rate_cap
        # a round never holds fewer chunks than fill the device twice over
        # (unless less work is left): smaller rounds leave CUs idle
        t = self.runner.tune
        cus = torch.cuda.get_device_properties(self.runner.device).multi_processor_count
        self.fill_chunks = 2 * cus * t.waves_per_block * t.blocks_per_cu * max(1, t.tile_chunks)
        self.grid = grid
        self.job_builder = job_builder or (lambda r: request_job(r, grid))
        self.in_flight = deque()
        self._last_done = None
        self.results = {}
        self.peak_used = 0
        self.rounds = 0
        self.t0 = None

    def warm(self, classes):
        """Capture the round graphs of the given kernel classes up front (a
        round's first use of a graph otherwise pays its capture, 0.1-4 ms);
        returns the number of graphs held."""
        return self.runner.warm(classes)

    # -- admission -------------------------------------------------------------
    def submit(self, req):
        self.queue.append(req)

    def _now(self):
        return time.perf_counter() - self.t0

    def _admit(self):
        while self.queue and self.free_slots:
            req = self.queue[0]
            nc = chunks_of(req.paths)
            if nc > self.pool.capacity:
                self.queue.popleft()
                now = self._now()
                self.results[req.rid] = Result(req.rid, math.nan, math.nan, 0, req.arrival, now,
                                               now, 0, error='exceeds the HBM budget')
                continue
            base = self.pool.alloc(nc)
            if base is None:
                break  # wait for memory to free up (FIFO: no overtaking)
            self.queue.popleft()
            slot = self.free_slots.pop()
            a = _Active()
            a.req, a.slot, a.base, a.nc, a.next = req, slot, base, nc, 0
            a.job = self.job_builder(req)
            a.cost = job_cost(a.job)
            a.admitted = self._now()
            a.in_table = False
            self.active[slot] = a
        self.peak_used = max(self.peak_used, self.pool.used)

    def _to_table(self, slots):
        """Prepare and upload the job-table entries of requests entering their
        first batched round (a request that runs alone on its plain kernel
        never needs one)."""
        new = [self.active[s] for s in slots if not self.active[s].in_table]
        if new:
            self.runner.prepare([a.job for a in new], [a.base for a in new],
                                [a.slot for a in new])
            self.runner.upload_slots([a.slot for a in new])
            for a in new:
                a.in_table = True

    # -- rounds ------------------------------------------------------------------
    def _plan(self):
        """Chunks of each admitted request for the next round (see the class
        docstring for the policies).  A round below `fill_chunks` is topped
        up, in the same order, so it still fills the device.  A request whose
        last chunk is in the round is finalised in it."""
        todo = [a for a in self.active.values() if a.next < a.nc]
        if not todo:
            return None
        if self.plain_single and len(todo) == 1 and not self.queue:
            a = todo[0]
            first, a.next = a.next, a.nc
            return [a.slot], [first], [a.nc - first], [a.cost], [a.slot]
        budget = self.cost_rate * self.round_s
        take = {a.slot: 0 for a in todo}
        left = budget
        if self.policy == 'srpt':
            todo.sort(key=lambda a: ((a.nc - a.next) * a.cost, a.admitted))
            for a in todo:
                if left <= 0:
                    break
                rem = a.nc - a.next
                k = min(rem, max(1, int(left / a.cost)))
                take[a.slot] = k
                left -= k * a.cost
        else:
            todo.sort(key=lambda a: a.admitted)
            for _ in range(2):
                for a in todo:
                    rem = a.nc - a.next - take[a.slot]
                    if rem <= 0:
                        continue
                    if left <= 0 and take[a.slot] > 0:
                        continue
                    k = min(rem, max(1, int(max(0.5 * left, 0.0) / a.cost)))
                    take[a.slot] += k
                    left -= k * a.cost
                if left <= 0:
                    break
        short = self.fill_chunks - sum(take.values())
        for a in todo:
            if short <= 0:
                break
            k = min(a.nc - a.next - take[a.slot], short)
            if k > 0:
                take[a.slot] += k
                short -= k
        slots, firsts, counts, costs, final = [], [], [], [], []
        for a in todo:
            k = take[a.slot]
            if k <= 0:
                continue
            slots.append(a.slot)
            firsts.append(a.next)
            counts.append(k)
            costs.append(a.cost)
            a.next += k
            if a.next >= a.nc:
                final.append(a.slot)
        return slots, firsts, counts, costs, final

    def _launch(self, plan):
        slots, firsts, counts, costs, final = plan
        start = torch.cuda.Event(enable_timing=True)
        cost = float(np.dot(counts, costs))
        full = sum(counts) >= self.fill_chunks
        # rounds that fill the device refine the rate that sizes rounds
        rated = cost if full else 0.0
        if (self.plain_single and len(slots) == 1 and firsts[0] == 0
                and final == slots):
            # one whole request: its plain kernel gives the same bits
            a = self.active[slots[0]]
            if not self.in_flight and not full:
                # a small request on an idle device: wait on its read-back, as
                # the plain form does, which returns sooner than polling an
                # event (a round that fills the device stays timed: it refines
                # the rate that sizes rounds)
                s4 = E.simulate(a.job, self.runner.device).cpu().numpy()
                self._finish(slots[0], s4, self._now())
                self.rounds += 1
                return
            start.record()
            stats = E.simulate(a.job, self.runner.device)
            host = self._host_stats()
            host.copy_(stats, non_blocking=True)
            done = torch.cuda.Event(enable_timing=True)
            done.record()
            self.in_flight.append(('plain', (host, done, slots[0]), final, start, rated))
        else:
            self._to_table(slots)
            buf = next(b for b in self.runner.bufs
                       if b not in [f[1] for f in self.in_flight if f[0] == 'batch'])
            rp = plan_segments(slots, firsts, counts, costs, self.runner.tune.tile_chunks,
                               self.runner.slot_class[np.asarray(slots, dtype=np.int64)])
            self.runner.stage_round(buf, rp, final)
            start.record()
            self.runner.launch(buf, rp.groups)
            self.runner.read_stats(buf, final)
            self.in_flight.append(('batch', buf, final, start, rated))
        self.rounds += 1

    def _host_stats(self):
        """A pinned 4-double buffer for a plain round's read-back (a small ring:
        at most two rounds are in flight)."""
        ring = getattr(self, '_ring', None)
        if ring is None:
            ring = self._ring = [torch.zeros(4, dtype=torch.float64, pin_memory=True)
                                 for _ in range(4)]
            self._ring_i = 0
        self._ring_i = (self._ring_i + 1) % len(ring)
        return ring[self._ring_i]

    def _complete(self, block):
        """Retire finished rounds (the oldest first); block on it if asked."""
        while self.in_flight:
            kind, what, final, start, cost = self.in_flight[0]
            done = what[1] if kind == 'plain' else what.done
            if not block and not done.query():
                return
            done.synchronize()
            block = False
            self.in_flight.popleft()
            ms = start.elapsed_time(done)
            prev, self._last_done = self._last_done, done
            if prev is not None:
                # a round queued behind the previous one starts when that ends:
                # its time is its own, not the wait
                ms -= max(0.0, start.elapsed_time(prev))
            if cost > 0 and ms > 0.05:  # rounds that filled the device refine the rate
                # one round moves the estimate by at most a factor of two
                r = min(max(cost / (ms * 1e-3), 0.5 * self.cost_rate), 2.0 * self.cost_rate)
                self.cost_rate = min(0.7 * self.cost_rate + 0.3 * r, self.rate_cap)
            now = self._now()
            for s in final:
                if kind == 'plain':
                    s4 = what[0].numpy().copy()
                else:
                    s4 = what.h_stats.numpy()[4 * s:4 * s + 4].copy()
                self._finish(s, s4, now)

    def _finish(self, slot, s4, now):
        """Record the result of the request in `slot` from its (sum, sum of
        squares, n, flags) and free its slot and chunk range."""
        a = self.active.pop(slot)
        err = None
        try:
            E.check_flags(s4[3], 'request %d' % a.req.rid)
        except L.SdemcError as e:
            err = str(e)
        mean, se = E.mean_and_se(s4[0], s4[1], int(s4[2]))
        self.results[a.req.rid] = Result(a.req.rid, mean, se, int(s4[2]), a.req.arrival,
                                         a.admitted, now, a.req.path_steps, tuple(s4), err)
        self.pool.release(a.base, a.nc)
        self.free_slots.append(slot)

    # -- driver ------------------------------------------------------------------
    def serve(self, requests, realtime=True):
        """Serve `requests` (with .arrival in seconds from start when realtime)
        to completion, with the cyclic garbage collector paused; returns
        {rid: Result}."""
        with no_cyclic_gc():
            return self._serve(requests, realtime)

    def _serve(self, requests, realtime):
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = time.perf_counter()
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        n_total = len(pending)
        while len(self.results) < n_total:
            now = self._now()
            while pending and (not realtime or pending[0].arrival <= now):
                self.submit(pending.popleft())
            self._admit()
            self._complete(block=False)
            if len(self.in_flight) < len(self.runner.bufs):
                plan = self._plan()
                if plan is not None:
                    self._launch(plan)
                    continue
            if self.in_flight:
                self._complete(block=len(self.in_flight) == len(self.runner.bufs) or not pending)
            elif pending:
                wait = pending[0].arrival - self._now()
                if wait > 0:
                    time.sleep(min(wait, 0.002))
        return self.results


def serve_plain(requests, realtime=True, grid=L.GRID_UNIFORM):
    """The plain per-job path under the same arrivals: one launch per request in
    arrival order, each read back before the next (mc_simple's pattern), with
    the cyclic garbage collector paused as in PricingServer.serve."""
    with no_cyclic_gc():
        return _serve_plain(requests, realtime, grid)


def _serve_plain(requests, realtime, grid):
    dev = E.require_device()
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    t0 = time.perf_counter()
    out = {}
    while pending:
        now = time.perf_counter() - t0
        r = pending[0]
        if realtime and r.arrival > now:
            time.sleep(min(r.arrival - now, 0.002))
            continue
        pending.popleft()
        start = time.perf_counter() - t0
        job = request_job(r, grid)
        s, q, n = E.stats_to_host(E.simulate(job, dev))
        mean, se = E.mean_and_se(s, q, n)
        fin = time.perf_counter() - t0
        out[r.rid] = Result(r.rid, mean, se, n, r.arrival if realtime else 0.0, start, fin,
                            r.path_steps, (s, q, n, 0.0))
    return out


def latency_summary(results, span=None):
    lat = np.array([r.latency for r in results.values() if r.error is None])
    ps = sum(r.path_steps for r in results.values() if r.error is None)
    t_end = max(r.finished for r in results.values())
    t_beg = min(r.arrival for r in results.values())
    span = span or (t_end - t_beg)
    return {"requests": len(results), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "throughput_path_steps_s": ps / span, "span_s": span}

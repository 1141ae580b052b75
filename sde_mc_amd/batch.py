"""Batched persistent launches: many jobs (uneven path and step counts, mixed
models, MLMC levels) packed into one kernel launch.

A :class:`BatchRunner` owns a device job table, a tile list, the chunk
partials and the per-job statistics.  ``run`` prices a list of jobs in one
launch: every job's paths are cut into 64-path chunks, chunks are grouped
into tiles, tiles are ordered longest-first, and the persistent kernel's
wavefronts take tiles from a device counter until the list is exhausted.
The three launches of a round (counter reset, tile kernel, finalisation) are
captured once in a HIP graph and replayed.

A job's result is independent of what it is packed with: the Philox streams
are keyed by (seed, path) and the reduction order is fixed per job (64-path
chunks, then chunks in order), so a job alone and in any mix gives the same
bits.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L
from . import engine as E


def chunks_of(n_paths):
    return (int(n_paths) + L.CHUNK - 1) // L.CHUNK


def job_cost(job):
    """Relative cost of one chunk of a job (fine steps, x2 for coupled pairs, per normal)."""
    m = job.model
    k = m.dim if m.noise == L.NOISE_DIAG else (m.dim + 1 if m.noise == L.NOISE_INDEP_COMMON
                                               else 2 * m.dim)
    c = job.steps * (1.0 + 0.25 * (k - 1))
    if job.coarse_factor:
        c *= 1.0 + 1.0 / job.coarse_factor
    return c


class BatchTune:
    def __init__(self, waves_per_block=4, blocks_per_cu=2, tile_chunks=1):
        self.waves_per_block = int(waves_per_block)
        self.blocks_per_cu = int(blocks_per_cu)
        self.tile_chunks = int(tile_chunks)

    def c(self):
        t = L.BatchTune()
        t.waves_per_block = self.waves_per_block
        t.blocks_per_cu = self.blocks_per_cu
        return t

    def key(self):
        return (self.waves_per_block, self.blocks_per_cu, self.tile_chunks)

    def __repr__(self):
        return 'BatchTune(wpb=%d, bpc=%d, tile_chunks=%d)' % self.key()


class BatchRunner:
    """Device-resident job table + tile list + persistent launch (graph-replayed)."""

    def __init__(self, max_jobs=1024, max_tiles=1 << 18, partial_chunks=1 << 18, tune=None,
                 use_graph=True):
        self.device = E.require_device()
        self.lib = L.lib()
        self.kb = self.lib.sdemc_sizeof_prepared_job()
        self.max_jobs = int(max_jobs)
        self.max_tiles = int(max_tiles)
        self.tune = tune or BatchTune()
        self.use_graph = use_graph
        dev = self.device
        self.jobs = torch.zeros(self.max_jobs * self.kb, dtype=torch.uint8, device=dev)
        self.tiles = torch.zeros(self.max_tiles * 4, dtype=torch.int32, device=dev)
        self.n_tiles = torch.zeros(1, dtype=torch.int32, device=dev)
        self.finalize = torch.zeros(self.max_jobs, dtype=torch.int32, device=dev)
        self.stats = torch.zeros(self.max_jobs * 4, dtype=torch.float64, device=dev)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self._alloc_partials(partial_chunks)
        # pinned staging for the per-round tables
        self.h_jobs = torch.zeros(self.max_jobs * self.kb, dtype=torch.uint8, pin_memory=True)
        self.h_tiles = torch.zeros(self.max_tiles * 4, dtype=torch.int32, pin_memory=True)
        self.h_ntiles = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        self.h_final = torch.zeros(self.max_jobs, dtype=torch.int32, pin_memory=True)
        self.h_stats = torch.zeros(self.max_jobs * 4, dtype=torch.float64, pin_memory=True)
        self.class_set = 1
        self._plans = {}

    def _alloc_partials(self, n):
        self.partial_chunks = int(n)
        self.partials = torch.zeros(self.partial_chunks * 2, dtype=torch.float64, device=self.device)
        self.cflags = torch.zeros(self.partial_chunks, dtype=torch.int32, device=self.device)
        self._plans = {}

    def close(self):
        for p in self._plans.values():
            self.lib.sdemc_batch_plan_destroy(p)
        self._plans = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def batch_dev(self, class_set):
        b = L.BatchDev()
        b.jobs_dev = self.jobs.data_ptr()
        b.tiles_dev = self.tiles.data_ptr()
        b.n_tiles_dev = self.n_tiles.data_ptr()
        b.finalize_dev = self.finalize.data_ptr()
        b.partials_dev = self.partials.data_ptr()
        b.chunk_flags_dev = self.cflags.data_ptr()
        b.stats_dev = self.stats.data_ptr()
        b.counter_dev = self.counter.data_ptr()
        b.partial_chunks = self.partial_chunks
        b.max_jobs = self.max_jobs
        b.max_tiles = self.max_tiles
        b.class_set = int(class_set)
        return b

    # -- tables ----------------------------------------------------------
    def prepare(self, jobs, chunk_base, slots=None):
        """Validate and stage jobs into slots (host pinned buffer); returns class set."""
        n = len(jobs)
        arr = (L.Job * n)(*jobs)
        base = (ctypes.c_int64 * n)(*[int(b) for b in chunk_base])
        buf = ctypes.create_string_buffer(n * self.kb)
        cs = ctypes.c_int32(0)
        L.check(self.lib.sdemc_prepare_jobs(arr, n, base, buf, ctypes.byref(cs)),
                'sdemc_prepare_jobs')
        raw = np.frombuffer(buf.raw, dtype=np.uint8).reshape(n, self.kb)
        hj = self.h_jobs.numpy().reshape(self.max_jobs, self.kb)
        slots = list(range(n)) if slots is None else list(slots)
        for i, s in enumerate(slots):
            hj[s] = raw[i]
        return int(cs.value)

    def upload_jobs(self, slots=None):
        if slots is None:
            self.jobs.copy_(self.h_jobs, non_blocking=True)
        else:
            for s in slots:
                a, b = s * self.kb, (s + 1) * self.kb
                self.jobs[a:b].copy_(self.h_jobs[a:b], non_blocking=True)

    def set_tiles(self, tiles):
        """tiles: int array [n, 4] of (slot, first chunk, count, cost)."""
        n = len(tiles)
        if n > self.max_tiles:
            raise ValueError('%d tiles exceed max_tiles %d' % (n, self.max_tiles))
        ht = self.h_tiles.numpy()
        if n:
            ht[:4 * n] = np.asarray(tiles, dtype=np.int32).reshape(-1)
        self.h_ntiles[0] = n
        self.tiles[:4 * n].copy_(self.h_tiles[:4 * n], non_blocking=True)
        self.n_tiles.copy_(self.h_ntiles, non_blocking=True)

    def set_finalize(self, slots):
        hf = self.h_final.numpy()
        hf[:] = 0
        for s in slots:
            hf[s] = 1
        self.finalize.copy_(self.h_final, non_blocking=True)

    # -- launch ----------------------------------------------------------
    def launch(self, class_set=1, tune=None):
        tune = tune or self.tune
        stream = E.stream_ptr()
        if self.use_graph:
            key = (class_set, tune.waves_per_block, tune.blocks_per_cu, self.partials.data_ptr())
            plan = self._plans.get(key)
            if plan is None:
                h = ctypes.c_void_p()
                L.check(self.lib.sdemc_batch_plan_create(ctypes.byref(self.batch_dev(class_set)),
                                                         ctypes.byref(tune.c()), ctypes.byref(h)),
                        'sdemc_batch_plan_create')
                plan = h
                self._plans[key] = plan
            L.check(self.lib.sdemc_batch_plan_launch(plan, stream), 'sdemc_batch_plan_launch')
        else:
            L.check(self.lib.sdemc_batch_launch(ctypes.byref(self.batch_dev(class_set)),
                                                ctypes.byref(tune.c()), stream),
                    'sdemc_batch_launch')

    @staticmethod
    def plan_tiles(jobs, slots, chunk_counts, tile_chunks, first_chunks=None):
        """Tiles of `tile_chunks` chunks for each job, longest (costliest) first."""
        out = []
        for i, job in enumerate(jobs):
            nc = chunk_counts[i]
            f0 = 0 if first_chunks is None else first_chunks[i]
            cost = int(min(job_cost(job), 2 ** 30))
            for f in range(f0, f0 + nc, tile_chunks):
                out.append((slots[i], f, min(tile_chunks, f0 + nc - f), cost))
        out.sort(key=lambda t: -t[3])
        return out

    def run(self, jobs, tune=None, sync=True, events=None):
        """Price `jobs` (list of L.Job) in one batched launch; returns [n, 4] fp64
        (sum, sum of squares, n, flags) per job (host numpy when sync)."""
        tune = tune or self.tune
        n = len(jobs)
        if n == 0:
            return np.zeros((0, 4))
        if n > self.max_jobs:
            raise ValueError('%d jobs exceed max_jobs %d' % (n, self.max_jobs))
        counts = [chunks_of(j.n_paths) for j in jobs]
        base = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        total = int(sum(counts))
        if total > self.partial_chunks:
            self._alloc_partials(max(total, 2 * self.partial_chunks))
        cs = self.prepare(jobs, base)
        self.upload_jobs()
        tiles = self.plan_tiles(jobs, list(range(n)), counts, tune.tile_chunks)
        self.set_tiles(tiles)
        self.set_finalize(range(n))
        if events is not None:
            events[0].record()
        self.launch(cs, tune)
        if events is not None:
            events[1].record()
        self.h_stats[:4 * n].copy_(self.stats[:4 * n], non_blocking=True)
        if not sync:
            return None
        torch.cuda.current_stream().synchronize()
        st = self.h_stats[:4 * n].numpy().reshape(n, 4).copy()
        for i in range(n):
            E.check_flags(st[i, 3], 'batch job %d' % i)
        return st

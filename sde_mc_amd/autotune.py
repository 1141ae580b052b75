"""Autotuner of the batched persistent launch.

Sweeps the launch shape -- wavefronts per workgroup (so paths per workgroup
= 64 x wavefronts), resident workgroups per CU, chunks per tile and the step
unroll -- on a calibration batch, timing the tile kernel with device events, and keeps the
fastest shape per (device, set of kernel classes in the batch).  Results are cached in a JSON
file next to the library so later processes start tuned.
"""
import hashlib
import json
import os
import time

import numpy as np
import torch

from . import _lib as L
from .batch import BatchRunner, BatchTune, chunks_of, job_class

# A CU has 4 SIMDs and the tile kernels fit 7 or 8 wavefronts per SIMD, so
# the shapes swept are workgroups of 2, 4 or 8 wavefronts (128, 256 or 512
# paths) with 3..8 resident wavefronts per SIMD in all, 1 or 2 chunks per
# tile, a step unroll of 1 or 2 four-step groups per loop iteration, and the
# class groups side by side or one after another.
WAVES_PER_SIMD = (3, 4, 5, 6, 7, 8)
CANDIDATES = [BatchTune(w, 4 * s // w, t, c, u) for w in (2, 4, 8) for s in WAVES_PER_SIMD
              for t in (1, 2) for c in (1, 0) for u in (1, 2) if (4 * s) % w == 0]

# A shape whose class groups run side by side replays a graph with parallel
# branches, and its time has varied between processes (the same shape and
# batch: 53.5 ms in one, 59.2 ms in another), while groups run in turn form
# one chain and repeat within 0.3%.  The confirm stage keeps a side-by-side
# shape only when it beats the best in-turn shape by more than this margin.
CONCURRENT_MARGIN = 0.02


def cache_path():
    return os.environ.get('SDEMC_TUNE_CACHE',
                          os.path.join(os.path.dirname(L.lib_path()), 'autotune.json'))


def _device_key():
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return '%s/%d' % (getattr(p, 'gcnArchName', p.name), p.multi_processor_count)


def _build_key(candidates):
    """What a cached pick depends on besides the device and the classes: the
    library's ABI version, the built library itself (size and modification
    time, so a rebuild of the kernels retunes) and the candidate list."""
    try:
        st = os.stat(L.lib_path())
        stamp = '%d-%d' % (st.st_size, int(st.st_mtime))
    except OSError:
        stamp = 'nolib'
    cands = hashlib.sha1(repr([c.key() for c in candidates]).encode()).hexdigest()[:12]
    return 'abi%d/lib%s/cand%s' % (L.ABI_VERSION, stamp, cands)


def load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c):
    try:
        with open(cache_path(), 'w') as f:
            json.dump(c, f, indent=1, sort_keys=True)
    except OSError:
        pass


def time_tune(runner, jobs, tune, reps=2):
    """Mean tile-kernel time (ms) of `jobs` in one batched launch with `tune`."""
    runner.run(jobs, tune=tune)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for e in ev:
        runner.run(jobs, tune=tune, events=e)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in ev]))


def time_tunes_interleaved(runner, jobs, keys, rounds=3):
    """Median tile-kernel time (ms) of `jobs` under each shape in `keys`,
    as [(key, ms)].  Every shape runs once untimed, then the shapes are timed
    one after another in each of `rounds` rounds, so a drift of the device's
    clock over the sweep reaches every shape alike and one slow or fast
    launch does not decide between close shapes."""
    tunes = [BatchTune(*k) for k in keys]
    for t in tunes:
        runner.run(jobs, tune=t)
    ms = [[] for _ in tunes]
    for _ in range(rounds):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in tunes]
        for t, e in zip(tunes, ev):
            runner.run(jobs, tune=t, events=e)
        torch.cuda.synchronize()
        for m, (a, b) in zip(ms, ev):
            m.append(a.elapsed_time(b))
    return [(k, float(np.median(m))) for k, m in zip(keys, ms)]


def autotune(jobs, candidates=None, reps=2, use_cache=True, verbose=False, confirm_jobs=None,
             confirm_top=8):
    """Best BatchTune for batches like `jobs`; returns (tune, table of timings).

    With `confirm_jobs` (the real batch, when `jobs` is a cut-down copy of
    it), the `confirm_top` fastest shapes of the sweep are timed again on
    `confirm_jobs`, each also with its class groups run one after another
    instead of side by side, and the fastest there is kept: a copy with
    capped path counts ranks close shapes differently from the full-size
    batch.  Either way, the fastest in-turn shape replaces a side-by-side
    winner that beats it by no more than CONCURRENT_MARGIN."""
    classes = sorted(set(job_class(j) for j in jobs))
    key = '%s/classes%s/%s' % (_device_key(), '-'.join(str(c) for c in classes),
                                _build_key(candidates or CANDIDATES))
    cache = load_cache() if use_cache else {}
    if use_cache and key in cache:
        return BatchTune(*cache[key]['best']), cache[key]['table']
    counts = sum(chunks_of(j.n_paths) for j in jobs)
    runner = BatchRunner(max_jobs=max(8, len(jobs)), partial_chunks=counts)
    table = []
    t0 = time.time()
    for tune in candidates or CANDIDATES:
        ms = time_tune(runner, jobs, tune, reps)
        table.append((tune.key(), ms))
        if verbose:
            print('autotune %r: %.3f ms' % (tune, ms), flush=True)
    runner.close()
    timed = table
    if confirm_jobs is not None:
        ranked = [k for k, _ in sorted(table, key=lambda r: r[1])[:confirm_top]]
        # and each of them with the class groups one after another: several
        # groups of unequal work can finish sooner in turn, each on the whole
        # device, than side by side, where the longest one ends alone
        ranked += [f for f in (k[:3] + (1 - k[3],) + k[4:] for k in ranked) if f not in ranked]
        counts = sum(chunks_of(j.n_paths) for j in confirm_jobs)
        runner = BatchRunner(max_jobs=max(8, len(confirm_jobs)), partial_chunks=counts)
        confirmed = time_tunes_interleaved(runner, confirm_jobs, ranked)
        runner.close()
        if verbose:
            for k, ms in confirmed:
                print('autotune confirm %r: %.3f ms' % (k, ms), flush=True)
        timed = confirmed
    best = min(timed, key=lambda r: r[1])
    in_turn = [r for r in timed if r[0][3] == 0]
    if best[0][3] != 0 and in_turn:
        bt = min(in_turn, key=lambda r: r[1])
        if bt[1] <= best[1] * (1.0 + CONCURRENT_MARGIN):
            best = bt
    cache = load_cache()
    cache[key] = {'best': list(best[0]), 'table': table, 'seconds': time.time() - t0}
    if use_cache:
        save_cache(cache)
    return BatchTune(*best[0]), table

"""The C ABI library: it loads, exports every entry point include/sdemc.h
declares, its structures match the ctypes mirrors, and host-side validation
rejects bad jobs (no device needed for any of this)."""
import ctypes
import os
import re

import pytest

from sde_mc_amd import _lib as L
from tests import jobs as J

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'sdemc.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(sdemc_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.exported_symbols())


def test_abi_version_and_layout():
    lib = L.load()
    assert lib.sdemc_abi_version() == L.ABI_VERSION == 5
    L.check_layout()
    assert lib.sdemc_sizeof_prepared_job() % 8 == 0


def test_status_strings():
    L.load()
    assert L.status_string(L.OK) == 'ok'
    assert 'argument' in L.status_string(L.ERR_ARG)
    with pytest.raises(L.SdemcError):
        L.check(L.ERR_UNSUPPORTED, 'x')


def test_workspace_bytes():
    lib = L.load()
    assert lib.sdemc_workspace_bytes(1) == 256
    assert lib.sdemc_workspace_bytes(1 << 22) >= (1 << 16) * 20


def _prepare(job):
    lib = L.load()
    buf = ctypes.create_string_buffer(lib.sdemc_sizeof_prepared_job())
    base = (ctypes.c_int64 * 1)(0)
    cs = ctypes.c_int32(-1)
    st = lib.sdemc_prepare_jobs(ctypes.byref(job), 1, base, buf, ctypes.byref(cs))
    return st, cs.value


def test_prepare_accepts_workloads_and_reports_classes():
    st, cs = _prepare(J.job_for(J.merton_sde(), 1.0, 500, 1 << 22))
    assert st == L.OK and cs == 0
    st, cs = _prepare(J.job_for(J.levy_rainbow_sde(), 1.0, 200, 1 << 22, grid='adapted'))
    assert st == L.OK and cs == 6
    st, cs = _prepare(J.job_for(J.merton_sde(), 1.0, 200, 1 << 10, coarse_factor=2))
    assert st == L.OK and cs == 1
    st, cs = _prepare(J.job_for(J.merton_sde(dim=3), 1.0, 50, 100))
    assert st == L.OK and cs == L.CLASS_GENERIC


@pytest.mark.parametrize("mutate,status", [
    (lambda j: setattr(j, 'steps', 0), L.ERR_ARG),
    (lambda j: setattr(j, 'T', -1.0), L.ERR_ARG),
    (lambda j: setattr(j.model, 'dim', 5), L.ERR_UNSUPPORTED),
    (lambda j: setattr(j.model, 'noise', 7), L.ERR_UNSUPPORTED),
    (lambda j: setattr(j.payoff, 'kind', 42), L.ERR_UNSUPPORTED),
    (lambda j: setattr(j.model, 'rate', float('nan')), L.ERR_ARG),
    (lambda j: setattr(j.model, 'rate', 1e9), L.ERR_ARG),
    (lambda j: setattr(j, 'coarse_factor', 3), L.ERR_ARG),      # 500 % 3 != 0
    (lambda j: setattr(j, 'coarse_factor', 1), L.ERR_ARG),
    (lambda j: (setattr(j, 'grid', L.GRID_ADAPTED), setattr(j, 'max_steps', 0)), L.ERR_ARG),
    (lambda j: setattr(j, 'n_paths', 1 << 40), L.ERR_ARG),
])
def test_prepare_rejects_bad_jobs(mutate, status):
    job = J.job_for(J.merton_sde(), 1.0, 500, 1000)
    mutate(job)
    assert _prepare(job)[0] == status


def _general_job(**kw):
    import torch
    import sde_mc_amd as sm
    sde = sm.GeneralSde(0.0, 0.02, torch.tensor([[0.1, 0.2, 0.0], [0.0, 0.1, -0.1]]),
                        0.0, torch.ones(2), 2, 3)
    return J.job_for(sde, 1.0, 50, 1000, **kw)


def test_general_noise_validation():
    # 'general' runs on the uniform grid in the generic batched class; the
    # brown_dim, the adapted grid and non-finite matrix entries are refused
    st, cs = _prepare(_general_job())
    assert st == L.OK and cs == L.CLASS_GENERIC
    assert _prepare(_general_job(coarse_factor=2))[0] == L.OK
    job = _general_job()
    job.grid = L.GRID_ADAPTED
    assert _prepare(job)[0] == L.ERR_UNSUPPORTED
    for bd in (0, 5):
        job = _general_job()
        job.model.brown_dim = bd
        assert _prepare(job)[0] == L.ERR_UNSUPPORTED
    job = _general_job()
    job.model.bm1[1 * 4 + 2] = float('nan')
    assert _prepare(job)[0] == L.ERR_ARG
    job = _general_job()
    job.model.bm1[1 * 4 + 3] = float('nan')  # outside (dim, brown_dim): ignored
    assert _prepare(job)[0] == L.OK
    job = _general_job()
    job.model.chol[2 * 4 + 1] = float('inf')
    assert _prepare(job)[0] == L.ERR_ARG


def test_simulate_rejects_before_touching_device():
    # validation happens on the host: a bad job never reaches a launch
    lib = L.load()
    job = J.job_for(J.merton_sde(), 1.0, 500, 1000)
    job.steps = 0
    st = lib.sdemc_simulate(ctypes.byref(job), None, None, None, 0, 4, None)
    assert st == L.ERR_ARG
    job.steps = 10
    st = lib.sdemc_simulate(ctypes.byref(job), ctypes.c_void_p(8), None, ctypes.c_void_p(8), 0, 4,
                            None)
    assert st == L.ERR_WORKSPACE
    st = lib.sdemc_simulate(ctypes.byref(job), ctypes.c_void_p(8), None, ctypes.c_void_p(8),
                            1 << 20, 3, None)
    assert st == L.ERR_ARG  # waves_per_block must be 1, 2, 4 or 8


# ---- host check of a batched launch's tables (sdemc_batch_check) -----------
def _batch_tables():
    """A real plan of four jobs of three classes and the prepared slots."""
    import numpy as np
    from sde_mc_amd.batch import chunks_of, job_cost, plan_segments
    lib = L.load()
    jobs = [J.job_for(J.merton_sde(), 1.0, 100, 5000),
            J.job_for(J.merton_sde(), 1.0, 50, 300),
            J.job_for(J.levy_rainbow_sde(), 1.0, 64, 2000),
            J.job_for(J.merton_sde(), 1.0, 64, 1000, coarse_factor=2)]
    n = len(jobs)
    counts = np.array([chunks_of(j.n_paths) for j in jobs], dtype=np.int64)
    base = np.zeros(n, dtype=np.int64)
    base[1:] = np.cumsum(counts)[:-1]
    kb = lib.sdemc_sizeof_prepared_job()
    max_jobs = 8
    prepared = ctypes.create_string_buffer(max_jobs * kb)
    cls = (ctypes.c_int32 * n)()
    arr = (L.Job * n)(*jobs)
    assert lib.sdemc_prepare_jobs(arr, n, (ctypes.c_int64 * n)(*base.tolist()), prepared, cls) == L.OK
    plan = plan_segments(np.arange(n), np.zeros(n, np.int64), counts,
                         np.array([job_cost(j) for j in jobs]), 2, list(cls))
    return prepared, max_jobs, plan, int(counts.sum())


def _check(prepared, max_jobs, segs, hdr, groups, partial_chunks, max_segments=64):
    import numpy as np
    lib = L.load()
    s = np.zeros((max_segments, L.SEGMENT_INTS), dtype=np.int32)
    s[:len(segs)] = segs
    h = np.ascontiguousarray(hdr, dtype=np.int32)
    g = (ctypes.c_int32 * len(groups))(*groups)
    return lib.sdemc_batch_check(prepared, max_jobs, s.ctypes.data, max_segments, h.ctypes.data,
                                 len(groups), g, partial_chunks)


def test_batch_check_accepts_a_planned_round():
    prepared, mj, plan, chunks = _batch_tables()
    assert len(plan.groups) == 3
    assert _check(prepared, mj, plan.segs, plan.hdr, plan.groups, chunks) == L.OK
    # every job is also valid in the generic kernel's group
    from sde_mc_amd.batch import plan_segments
    import numpy as np
    segs = plan.segs.copy()
    n = len(segs)
    gen = plan_segments(segs[:, 0], segs[:, 1], segs[:, 2], np.ones(n), 1)
    assert _check(prepared, mj, gen.segs, gen.hdr, gen.groups, chunks) == L.OK


def _mutations():
    def seg(i, col, v):
        return lambda s, h, g, pc: (s.__setitem__((i, col), v), (s, h, g, pc))[1]
    return [
        ('unprepared slot', seg(0, 0, 5)),
        ('slot beyond the table', seg(0, 0, 8)),
        ('negative first chunk', seg(0, 1, -1)),
        ('chunks past the job', lambda s, h, g, pc: (s.__setitem__((0, 2), s[0, 2] + 1), (s, h, g, pc))[1]),
        ('zero chunks per tile', seg(0, 3, 0)),
        ('tile gap', lambda s, h, g, pc: (s.__setitem__((1, 4), s[1, 4] + 1), (s, h, g, pc))[1]),
        ('tile overlap', lambda s, h, g, pc: (s.__setitem__((1, 4), s[1, 4] - 1), (s, h, g, pc))[1]),
        ('header tile count', lambda s, h, g, pc: (h.__setitem__((0, 2), h[0, 2] + 1), (s, h, g, pc))[1]),
        ('header past the list', lambda s, h, g, pc: (h.__setitem__((0, 0), 63), (s, h, g, pc))[1]),
        ('wrong class', lambda s, h, g, pc: (s, h, tuple(reversed(g)), pc)),
        ('partial pool too small', lambda s, h, g, pc: (s, h, g, pc - 1)),
    ]


@pytest.mark.parametrize("name,mut", _mutations(), ids=[m[0] for m in _mutations()])
def test_batch_check_rejects_misplanned_tables(name, mut):
    prepared, mj, plan, chunks = _batch_tables()
    # the first group must hold at least two segments for the tile mutations
    assert plan.hdr[0, 1] >= 2
    s, h, g, pc = mut(plan.segs.copy(), plan.hdr.copy(), plan.groups, chunks)
    assert _check(prepared, mj, s, h, g, pc) == L.ERR_ARG, name


def test_noalloc_entry_point_refuses_before_touching_the_device():
    # sizes beyond the library's array and invalid jobs are refused on the host
    lib = L.load()
    job = J.job_for(J.merton_sde(), 1.0, 10, (1 << 24) + 1)
    assert lib.sdemc_simulate_noalloc(ctypes.byref(job), ctypes.c_void_p(8), None, 4, None) == \
        L.ERR_WORKSPACE
    job.steps = 0
    assert lib.sdemc_simulate_noalloc(ctypes.byref(job), ctypes.c_void_p(8), None, 4, None) == \
        L.ERR_ARG


def test_hbm_read_probe_refuses_bad_arguments():
    # null or misaligned source, a size that is not whole 16-byte words, no
    # workgroups: refused on the host, nothing launched
    lib = L.load()
    out = ctypes.c_void_p(64)
    assert lib.sdemc_hbm_read_probe_dev(None, 1024, out, 4, None) == L.ERR_ARG
    assert lib.sdemc_hbm_read_probe_dev(ctypes.c_void_p(16), 1000, out, 4, None) == L.ERR_ARG
    assert lib.sdemc_hbm_read_probe_dev(ctypes.c_void_p(8), 1024, out, 4, None) == L.ERR_ARG
    assert lib.sdemc_hbm_read_probe_dev(ctypes.c_void_p(16), 1024, out, 0, None) == L.ERR_ARG
    assert lib.sdemc_hbm_read_probe_dev(ctypes.c_void_p(16), 1024, None, 4, None) == L.ERR_ARG

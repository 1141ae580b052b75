"""The zero-allocation entry point sdemc_simulate_noalloc on the device: its
statistics and per-path outputs equal sdemc_simulate's (the entry point with
a caller workspace) bit for bit, on two calls in a row -- the second with a
smaller job, so the library's array still holds the first call's partials --
and jobs beyond the array are refused."""
import ctypes

import numpy as np
import pytest
import torch

import sde_mc_amd as sm
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from tests import jobs as J

pytestmark = pytest.mark.gpu


def _noalloc(job, dev, out=None):
    stats = torch.full((4,), float('nan'), dtype=torch.float64, device=dev)
    st = L.lib().sdemc_simulate_noalloc(ctypes.byref(job), ctypes.c_void_p(stats.data_ptr()),
                                        ctypes.byref(out) if out is not None else None, 4,
                                        E.stream_ptr())
    L.check(st, 'sdemc_simulate_noalloc')
    return stats.cpu().numpy()


def _with_outputs(fn, job, dev):
    P, d = int(job.n_paths), job.model.dim
    lv = 2 if job.coarse_factor else 1
    pay = torch.empty(P, dtype=torch.float32, device=dev)
    term = torch.empty((lv, d, P), dtype=torch.float32, device=dev)
    out = L.PathOut()
    out.payoff_dev, out.terminal_dev = pay.data_ptr(), term.data_ptr()
    stats = fn(job, dev, out)
    return stats, pay.cpu().numpy(), term.cpu().numpy()


def test_noalloc_equals_the_workspace_entry_point_on_two_calls():
    dev = E.require_device()
    big = J.job_for(J.levy_rainbow_sde(), 1.0, 64, (1 << 20) + 17,
                    option=sm.Rainbow(1.0, log=True, discount=1.02), disc=0.98, seed=71)
    small = J.job_for(J.merton_sde(), 1.0, 100, 5000 + 3, grid='adapted',
                      option=sm.EuroCall(1.0), disc=0.98, seed=72)
    for job in (big, small):
        print('noalloc job P=%d N=%d grid=%d' % (job.n_paths, job.steps, job.grid), flush=True)
        ref = _with_outputs(lambda j, d, o: E.simulate(j, d, out=o).cpu().numpy(), job, dev)
        got = _with_outputs(_noalloc, job, dev)
        assert got[0][2] == job.n_paths and got[0][3] == 0
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_noalloc_refuses_jobs_beyond_its_array():
    job = J.job_for(J.merton_sde(), 1.0, 10, (1 << 24) + 1)
    stats = torch.zeros(4, dtype=torch.float64, device=E.require_device())
    st = L.lib().sdemc_simulate_noalloc(ctypes.byref(job), ctypes.c_void_p(stats.data_ptr()), None,
                                        4, E.stream_ptr())
    assert st == L.ERR_WORKSPACE
    assert L.lib().sdemc_workspace_bytes(1 << 24) <= 5 << 20 < L.lib().sdemc_workspace_bytes((1 << 24) + 64)

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
    assert lib.sdemc_abi_version() == L.ABI_VERSION == 3
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

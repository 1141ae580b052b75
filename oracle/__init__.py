"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the sde_mc hot path.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
to check and time the HIP path; the sde_mc_amd package never imports it.

  oracle.closed_form  fp64 closed forms (Merton series, Lewis/Fourier Merton,
                      Black-Scholes, Stulz call on the max of two assets)
  oracle.refcpu       torch-CPU restatement of the reference solver loop,
                      op for op (the timed "reference CPU path")
  liboracle.so        C fp64 per-path restatement (sdemc_oracle.c), wrapped
                      by simulate / simulate_fixed / philox / normal below
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib_path():
    return os.path.join(_HERE, "lib", "liboracle.so")


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(lib_path()):
            from . import build as _b
            _b.build()
        lib = ctypes.CDLL(lib_path())
        lib.oracle_philox.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        lib.oracle_philox.restype = None
        lib.oracle_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        lib.oracle_normal.restype = ctypes.c_double
        lib.oracle_simulate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_simulate.restype = ctypes.c_int
        lib.oracle_simulate_fixed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_simulate_fixed.restype = ctypes.c_int
        lib.oracle_levy_icdf.argtypes = [ctypes.c_void_p, ctypes.c_float]
        lib.oracle_levy_icdf.restype = ctypes.c_double
        lib.oracle_ln_unit.argtypes = [ctypes.c_uint32]
        lib.oracle_ln_unit.restype = ctypes.c_float
        lib.oracle_poisson_thresholds.argtypes = [ctypes.c_double, ctypes.c_void_p,
                                                  ctypes.c_void_p]
        lib.oracle_poisson_thresholds.restype = ctypes.c_int
        lib.oracle_step_jumps.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32,
                                          ctypes.c_void_p]
        lib.oracle_step_jumps.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def philox(ctr, key):
    """Philox4x32-10 of one counter (4 uint32) under key (2 uint32)."""
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ctypes.c_uint32 * 4)()
    load().oracle_philox(c, k, o)
    return [int(v) for v in o]


def normal(seed, path, i):
    return load().oracle_normal(seed, path, i)


def _dims(job):
    d = job.model.dim
    return d, (2 if job.coarse_factor > 0 else 1)


def simulate(job, p0, n):
    """Paths p0..p0+n-1 of `job` (a ctypes sdemc_job) in fp64:
    (payoff [n], terminal [levels, dim, n], nsteps [n])."""
    d, lv = _dims(job)
    pay = np.empty(n, dtype=np.float64)
    term = np.empty((lv, d, n), dtype=np.float64)
    ns = np.empty(n, dtype=np.int32)
    load().oracle_simulate(ctypes.addressof(job), p0, n, _p(pay), _p(term), _p(ns))
    return pay, term, ns


def simulate_fixed(job, fixed, P):
    """Fixed-increment paths from the host arrays behind `fixed` (ctypes
    sdemc_fixed_in holding host pointers)."""
    d, lv = _dims(job)
    pay = np.empty(P, dtype=np.float64)
    term = np.empty((lv, d, P), dtype=np.float64)
    ns = np.empty(P, dtype=np.int32)
    load().oracle_simulate_fixed(ctypes.addressof(job), ctypes.addressof(fixed), P, _p(pay),
                                 _p(term), _p(ns))
    return pay, term, ns


def levy_icdf(job, y):
    return load().oracle_levy_icdf(ctypes.addressof(job), float(np.float32(y)))


def ln_unit(w):
    """fp32 ln((w | 1) 2^-32) as the kernels compute the arrival gaps."""
    return load().oracle_ln_unit(int(w) & 0xFFFFFFFF)


def poisson_thresholds(mu):
    """(thresholds T_1..T_16, reciprocal interval widths [16], counts above 16
    negligible) of the per-step Poisson count for mean mu."""
    T = (ctypes.c_uint32 * 16)()
    inv = (ctypes.c_double * 16)()
    ok = load().oracle_poisson_thresholds(float(mu), T, inv)
    return [int(v) for v in T], [float(v) for v in inv], bool(ok)


def step_jumps(job, path, n):
    """(count, first mark) of uniform-grid step n in per-step Poisson mode."""
    fm = ctypes.c_double()
    c = load().oracle_step_jumps(ctypes.addressof(job), int(path), int(n), ctypes.byref(fm))
    return c, fm.value

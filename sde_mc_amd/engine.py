"""Plumbing between the Python API and libsdemc: job descriptors, device
buffers (torch's caching allocator on the current HIP stream) and the calls
into the C ABI.  All arithmetic on paths happens in the HIP kernels.
"""
import ctypes
import math

import numpy as np
import torch
from scipy.stats import poisson

from . import _lib as L


def require_device():
    """The HIP device results are computed on; raises if there is none."""
    if not torch.cuda.is_available():
        raise L.HipUnavailableError('no HIP device: the sde_mc_amd kernels run on MI355X only')
    L.load()
    return torch.device('cuda', torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def reference_max_jumps(rate, T):
    """The reference's storage bound on jumps per path (solvers.py:133)."""
    if rate <= 0:
        return 5
    return max(int(T * poisson.ppf(1 - 1 / 1e9, rate)), 5)


def safe_max_jumps(rate, T):
    """A cap on jumps per path whose excess probability is below 1e-15."""
    if rate <= 0:
        return 8
    return int(poisson.ppf(1 - 1e-15, rate * T)) + 16


def to_float(v):
    if torch.is_tensor(v):
        return float(v.detach().double().cpu().reshape(-1)[0])
    return float(v)


def fill_payoff(job, option, disc):
    """Fill job.payoff from an Option; returns False when the payoff must be
    applied in torch on the terminal states instead."""
    p = job.payoff
    p.disc = float(disc)
    p.scale = 1.0
    p.log = 0
    p.strike = 0.0
    if option is not None and option.kernel_payoff(p):
        return True
    p.kind = L.PAYOFF_IDENTITY
    p.log = 0
    p.scale = 1.0
    return False


def make_job(model, T, steps, grid, n_paths, seed, path_offset, coarse_factor=0, max_steps=0,
             exact_jumps=False):
    job = L.Job()
    model.fill(job.model, exact_jumps=exact_jumps)
    job.T = float(T)
    job.steps = int(steps)
    job.grid = int(grid)
    job.coarse_factor = int(coarse_factor)
    job.max_steps = int(max_steps)
    job.n_paths = int(n_paths)
    job.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    job.path_offset = int(path_offset)
    job.payoff.kind = L.PAYOFF_IDENTITY
    job.payoff.scale = 1.0
    job.payoff.disc = 1.0
    return job


class Workspace:
    """Grow-only device scratch for chunk partials, reused across calls on one
    (device, stream): launches on one stream run in order, so they may share
    it, while launches on different streams may overlap and get their own.
    A buffer that grows is freed back to torch's caching allocator, which
    reuses it only in that stream's order."""

    def __init__(self):
        self.buf = {}

    def get(self, nbytes, device):
        key = (str(device), torch.cuda.current_stream(device).cuda_stream)
        b = self.buf.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            self.buf[key] = b
        return b


_WS = Workspace()


def simulate(job, device, out=None, wpb=4, stats=None):
    """Launch the plain RNG-mode kernel; returns the device stats tensor [4] fp64."""
    lib = L.lib()
    nbytes = lib.sdemc_workspace_bytes(job.n_paths)
    ws = _WS.get(nbytes, device)
    if stats is None:
        stats = torch.empty(4, dtype=torch.float64, device=device)
    po = ctypes.byref(out) if out is not None else None
    st = lib.sdemc_simulate(ctypes.byref(job), ptr(stats), po, ptr(ws), ws.numel(), int(wpb),
                            stream_ptr())
    L.check(st, 'sdemc_simulate')
    return stats


def simulate_fixed(job, fixed, device, out=None, stats=None):
    lib = L.lib()
    nbytes = lib.sdemc_workspace_bytes(job.n_paths)
    ws = _WS.get(nbytes, device)
    if stats is None:
        stats = torch.empty(4, dtype=torch.float64, device=device)
    po = ctypes.byref(out) if out is not None else None
    st = lib.sdemc_simulate_fixed(ctypes.byref(job), ctypes.byref(fixed), ptr(stats), po, ptr(ws),
                                  ws.numel(), stream_ptr())
    L.check(st, 'sdemc_simulate_fixed')
    return stats


def check_flags(flags, where):
    """Raise SdemcError for a statistics row whose flags word is not 0."""
    if flags != flags:  # NaN: no finalisation wrote this row
        raise L.SdemcError(L.ERR_ARG, where + ' (statistics never written)')
    f = int(flags)
    if f & L.FLAG_TABLE:
        raise L.SdemcError(L.ERR_ARG, where + ' (segment table does not match the job table)')
    if f & L.FLAG_OVERFLOW:
        raise L.SdemcError(L.ERR_OVERFLOW, where)
    if f & L.FLAG_JUMP_CSR:
        raise L.SdemcError(L.ERR_ARG, where + ' (jump CSR read outside jump_size)')
    if f & L.FLAG_CLASS:
        raise L.SdemcError(L.ERR_UNSUPPORTED, where + ' (job outside the launch class set)')
    if f:
        raise L.SdemcError(L.ERR_ARG, where + ' (unknown flags %d)' % f)


def stats_to_host(stats, where='simulate'):
    """(sum, sum of squares, n) as python floats, after the stream drains."""
    s = stats.double().cpu().numpy()
    check_flags(s[3], where)
    return float(s[0]), float(s[1]), int(round(s[2]))


def mean_and_se(s, q, n):
    """Mean and standard error from fp64 sums; the unbiased variance of
    helpers.py:51-68 divided by n (mc.py:119-120)."""
    mean = s / n
    var = (q - s * s / n) / (n - 1) if n > 1 else float('nan')
    var = max(var, 0.0)
    return mean, math.sqrt(var / n)

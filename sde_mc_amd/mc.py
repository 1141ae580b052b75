"""Plain Monte Carlo estimators (sde_mc/mc.py:13-123, 418-440).

``mc_simple`` keeps the reference's signature; its time loop, payoff and
fp64 sums run in one HIP launch per batch.  The reference's control-variate
estimators (mc.py:126-467) consume stored paths for neural networks and are
not part of this library.
"""
import time

import numpy as np
import torch

from . import engine as E
from .options import ConstantShortRate


class MCStatistics:
    """Monte Carlo statistics (mc.py:13-50): mean, standard error, wall time, N."""

    def __init__(self, sample_mean, sample_std, time_elapsed, num_trials, paths=None,
                 payoffs=None, normals=None):
        self.sample_mean = sample_mean.item() if torch.is_tensor(sample_mean) else float(sample_mean)
        self.sample_std = sample_std.item() if torch.is_tensor(sample_std) else float(sample_std)
        self.time_elapsed = time_elapsed
        self.num_trials = num_trials
        self.paths = paths
        self.payoffs = payoffs
        self.normals = normals

    def __str__(self):
        return 'Mean: {:.6f}  +/- {:.6f}    Time taken (s): {:.2f}    N: {:.2E}'.format(
            self.sample_mean, self.sample_std * 1.96, self.time_elapsed, self.num_trials)


def _disc(discounter, solver):
    return E.to_float(discounter(solver.time_interval))


def mc_simple(num_trials, sde_solver, payoff, discounter=None, bs=None, return_normals=False,
              payoff_time='terminal', max_path_bytes=1 << 28):
    """Monte Carlo estimate of E[discount(T) payoff(X_T)] (mc.py:53-123).

    Without ``bs``, per-path payoffs are returned and the paths are stored when
    they fit in ``max_path_bytes``.  With ``bs`` the work runs in launches of
    ``bs`` paths and only the fp64 sums come back.  The payoff is always taken
    at the terminal time: the reference's payoff_time='terminal' reads column
    num_steps of a jump-adapted path, which is an earlier state whenever a
    path jumped (mc.py:84-85); 'adapted' (mc.py:90-91) is what is computed.
    """
    if discounter is None:
        discounter = ConstantShortRate(r=0.0)
    num_trials = int(num_trials)
    disc = _disc(discounter, sde_solver)
    start = time.time()
    if not bs:
        d = sde_solver.sde.dim
        rows = sde_solver.num_steps + E.safe_max_jumps(sde_solver._rate(), sde_solver.time_interval)
        store = num_trials * (rows + 1) * d * 4 <= max_path_bytes
        out = normals = None
        if store:
            off = sde_solver.path_counter
            out, normals = sde_solver.solve(bs=num_trials, return_normals=return_normals)
            sde_solver.path_counter = off  # the payoff pass reuses the same paths
        payoffs, _ = sde_solver.simulate_payoffs(num_trials, payoff, disc)
        p64 = payoffs.double()
        s, q = float(p64.sum()), float((p64 * p64).sum())
        mn, se = E.mean_and_se(s, q, num_trials)
        tt = time.time() - start
        od = sde_solver._out_device()
        return MCStatistics(mn, se, tt, num_trials, out, payoffs.to(od), normals)
    bs = int(bs)
    remaining = num_trials
    s = q = 0.0
    while remaining:
        b = min(bs, remaining)
        remaining -= b
        bs_s, bs_q, _ = sde_solver.simulate_stats(b, payoff, disc)
        s += bs_s
        q += bs_q
    mn = s / num_trials
    sd = np.sqrt(max(q / num_trials - mn * mn, 0.0) * (num_trials / (num_trials - 1))) / np.sqrt(num_trials)
    return MCStatistics(mn, sd, time.time() - start, num_trials)


def find_num_trials(problem, eps, models=None, init_trials=1e5, bs=1e5):
    """Pilot run, then n = ceil((1.96 se / eps)^2 n0) (mc.py:418-427)."""
    if models is not None:
        raise NotImplementedError('control-variate models are not supported by sde_mc_amd')
    stats = mc_simple(init_trials, problem.solver, problem.payoff, problem.discounter, bs,
                      payoff_time='adapted')
    ratio = (stats.sample_std * 1.96 / eps) ** 2
    return int(np.ceil(ratio * init_trials))


def run_mc(problem, eps, bs=1e5, init_trials=1e5):
    """Price to a 95% half-width eps (mc.py:437-440)."""
    trials = find_num_trials(problem, eps, None, init_trials, bs)
    return mc_simple(trials, problem.solver, problem.payoff, problem.discounter, bs=bs,
                     payoff_time='adapted')

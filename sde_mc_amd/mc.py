"""Monte Carlo estimators (API of sde_mc/mc.py).

``mc_simple`` keeps the reference's signature; its time loop, payoff and
fp64 sums run in one HIP launch per batch.  The control-variate estimators
(mc_control_variates, mc_adaptive_cv, mc_apply_cvs, run_cv_mc) train the
networks with torch on paths stored by the HIP kernels, then apply them in
the fused control-variate kernel: the MLPs run on the matrix cores inside
the path loop and no path is stored (cv.py).  Networks, discounters or
payoffs the fused kernel cannot take (recurrent models, batch statistics,
non-constant rates, payoffs outside the kernel's set) go through the
reference's algorithm in torch on stored paths instead.
"""
import math
import time

import numpy as np
import torch
from torch.utils.data import DataLoader

from . import _lib as L
from . import cv as CV
from . import engine as E
from .helpers import ceil_mult, mc_estimates, remove_steps, sample_cov
from .nets import AdaptedPathData, NormalPathData
from .options import ConstantShortRate
from .varred import (EarlyStopping, apply_adapted_control_variates,
                     apply_diffusion_control_variate, train_adapted_control_variates,
                     train_diffusion_control_variate)


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


# ---------------------------------------------------------------------------
# Control variates
# ---------------------------------------------------------------------------
FUSED_MIN_LAUNCH = 1 << 20   # paths per fused launch at least (fills the device)


class _FusedCv:
    """One estimator's networks packed for the fused kernel."""

    def __init__(self, solver, payoff, discounter, f, g, cv_steps, rate, comp):
        self.solver, self.payoff = solver, payoff
        self.disc = _disc(discounter, solver)
        self.dev = E.require_device()
        sde = solver.sde
        d = sde.dim
        self.f = CV.DeviceNetwork(f, d + 1, sde.brown_dim, self.dev)
        self.g = CV.DeviceNetwork(g, d + 1, d, self.dev) if g is not None else None
        self.cv_steps, self.rate, self.comp = cv_steps, rate, comp

    def run(self, n, gamma=False):
        job, in_kernel = self.solver.make_job(n, option=self.payoff, discount=self.disc)
        assert in_kernel
        steps = self.cv_steps if self.cv_steps is not None else int(job.max_steps)
        stats, g = CV.cv_simulate(job, self.f, self.g, steps, self.rate, self.comp, self.dev,
                                  gamma=gamma)
        s, q, _ = E.stats_to_host(stats, 'control variates')
        return s, q, g


def _is_eval_packable(model):
    if model is None or CV.is_zero_network(model):
        return True
    if getattr(model, 'sequential', False):
        return False
    try:
        CV._affine_layers(model)
    except NotImplementedError:
        return False
    return True


def fused_control_variates(models, solver, payoff, discounter, tol=0):
    """The fused-kernel form of these networks, or None when the kernel
    cannot evaluate them (the torch path on stored paths is used then)."""
    if not isinstance(discounter, ConstantShortRate):
        return None
    probe = L.Payoff()
    if not payoff.kernel_payoff(probe):
        return None
    sde = solver.sde
    struct = sde.diffusion_struct
    if solver.has_jumps:
        if not isinstance(models, (list, tuple)) or len(models) != 2 or tol != 0:
            return None
        if solver.grid != 'adapted' or struct not in ('diag', 'indep'):
            return None
        f, g = models
        comp = E.to_float(torch.as_tensor(sde.jump_rate()) * torch.as_tensor(sde.jump_mean()))
        cv_steps = None       # every step of the adapted grid
    else:
        f, g, comp = (models[0] if isinstance(models, (list, tuple)) else models), None, 0.0
        if solver.grid != 'uniform':
            return None
        steps = solver.num_steps
        cv_steps = remove_steps(tol, steps, solver.time_interval) if tol != 0 else steps
        cv_steps = max(0, min(steps, cv_steps))
    if sde.dim + 2 > 8 or sde.brown_dim > CV.MAX_OUT:
        return None
    if not (_is_eval_packable(f) and _is_eval_packable(g)):
        return None
    try:
        return _FusedCv(solver, payoff, discounter, f, g, cv_steps, float(discounter.r), comp)
    except NotImplementedError:
        return None


def _store_paths(trials, solver, payoff, discounter):
    """solve() with increments plus the discounted payoffs of the same paths."""
    off = solver.path_counter
    paths, aux = solver.solve(bs=trials, return_normals=True)
    solver.path_counter = off
    payoffs, _ = solver.simulate_payoffs(trials, payoff, _disc(discounter, solver))
    return paths, aux, payoffs.to(solver._out_device())


def simulate_data(trials, solver, payoff, discounter, bs=1000, inference=False):
    """DataLoader of NormalPathData on HIP-simulated paths (mc.py:378-388)."""
    trials = int(trials)
    if inference:
        assert not trials % bs, 'Batch size should partition total trials evenly'
    paths, normals, payoffs = _store_paths(trials, solver, payoff, discounter)
    return DataLoader(NormalPathData(paths, payoffs, normals), batch_size=int(bs),
                      shuffle=not inference, drop_last=not inference)


def simulate_adapted_data(trials, solver, payoff, discounter, bs=1000, inference=False):
    """DataLoader of AdaptedPathData on HIP-simulated jump-adapted paths (mc.py:391-399)."""
    trials = int(trials)
    if getattr(solver, 'grid', 'adapted') != 'adapted':
        raise NotImplementedError('jump control variates need the jump-adapted grid '
                                  '(grid=\'adapted\'): the uniform grid keeps no pre-jump states')
    paths, aux, payoffs = _store_paths(trials, solver, payoff, discounter)
    normals, time_paths, left_paths, total, jump_paths = aux
    ds = AdaptedPathData(paths[:, :total + 1], payoffs, normals[:, :total],
                         left_paths[:, :total + 1], time_paths[:, :total + 1],
                         jump_paths[:, :total + 1], total)
    return DataLoader(ds, batch_size=int(bs), shuffle=not inference, drop_last=not inference)


def sim_train_control_variates(models, opt, solver, trials, payoff, discounter, sim_bs, bs,
                               epochs=10, print_losses=True, tol=0, early_stopping=None):
    """Simulate training paths and fit the networks (mc.py:402-411)."""
    if solver.has_jumps:
        dl = simulate_adapted_data(trials, solver, payoff, discounter, bs=bs)
        return train_adapted_control_variates(models, opt, dl, solver, discounter, epochs,
                                              print_losses, tol, early_stopping)
    dl = simulate_data(trials, solver, payoff, discounter, bs=bs)
    _, losses = train_diffusion_control_variate(models, opt, dl, solver, discounter, epochs,
                                                print_losses, tol, early_stopping)
    return losses


def mc_apply_cvs(models, solver, trials, payoff, discounter, sim_bs=1e5, bs=1000, tol=0):
    """Estimate with trained control variates applied (mc.py:187-232).

    Fused kernel when the networks allow it (launches of at least
    FUSED_MIN_LAUNCH paths; `bs` is unused there), else the reference's
    simulate-then-apply loop in torch with `sim_bs` paths per simulation and
    network batches of `bs`."""
    start = time.time()
    trials = int(trials)
    fused = fused_control_variates(models, solver, payoff, discounter, tol)
    s = q = 0.0
    remaining = trials
    chunk = int(sim_bs) if fused is None else max(int(sim_bs), FUSED_MIN_LAUNCH)
    while remaining > 0:
        b = min(chunk, remaining)
        remaining -= b
        if fused is not None:
            x, y, _ = fused.run(b)
        elif solver.has_jumps:
            dl = simulate_adapted_data(b, solver, payoff, discounter, bs=bs, inference=True)
            x, y = apply_adapted_control_variates(models, dl, solver, discounter, tol)
        else:
            dl = simulate_data(b, solver, payoff, discounter, bs=bs, inference=True)
            x, y = apply_diffusion_control_variate(models, dl, solver, discounter, tol)
        s += float(x)
        q += float(y)
    mn, var = mc_estimates(s, q, trials)
    sd = math.sqrt(max(var, 0.0)) / math.sqrt(trials)
    return MCStatistics(mn, sd, time.time() - start, trials)


def mc_control_variates(models, opt, solver, trials, steps, payoff, discounter, sim_bs=(1e5, 1e5),
                        bs=(1000, 1000), epochs=10, print_losses=True, tol=0,
                        early_stopping=None):
    """Neural control variates for a model without jumps (mc.py:126-184):
    train on trials[0] paths of steps[0] steps, estimate on trials[1] paths
    of steps[1] steps."""
    assert not solver.has_jumps
    if early_stopping is not None:
        early_stopping.batch_size = bs[1]
    solver.num_steps = steps[0]
    t0 = time.time()
    sim_train_control_variates(models, opt, solver, trials[0], payoff, discounter, sim_bs[0],
                               bs[0], epochs, print_losses, tol, early_stopping)
    train_time = time.time() - t0
    solver.num_steps = steps[1]
    stats = mc_apply_cvs(models, solver, trials[1], payoff, discounter, sim_bs[1], bs[1], tol)
    stats.time_elapsed += train_time
    return stats


def mc_adaptive_cv(models, opt, solver, trials, steps, payoff, discounter, sim_bs=(1e4, 1e4),
                   bs=(1000, 1000), epochs=10, print_losses=True, pre_trained=False, tol=0,
                   early_stopping=None):
    """Neural control variates [f, g] for a jump model on the adapted grid
    (mc.py:235-326)."""
    if early_stopping is not None:
        early_stopping.batch_size = bs[1]
    solver.num_steps = steps[0]
    t0 = time.time()
    if not pre_trained:
        dl = simulate_adapted_data(trials[0], solver, payoff, discounter, bs=bs[0])
        train_adapted_control_variates(models, opt, dl, solver, discounter, epochs, print_losses,
                                       tol, early_stopping=early_stopping)
    train_time = time.time() - t0
    solver.num_steps = steps[1]
    stats = mc_apply_cvs(models, solver, trials[1], payoff, discounter, sim_bs[1], bs[1], tol)
    stats.time_elapsed += train_time
    return stats


def mc_terminal_cv(num_trials, sde_solver, payoff, discounter=None, bs=None, return_normals=False):
    """The discounted terminal spot as control variate (mc.py:329-375):
    gamma = payoff - b (disc(T) S_T - S_0) with b = cov / var estimated on
    the whole sample, or on the first batch when `bs` is given."""
    if discounter is None:
        discounter = ConstantShortRate(r=0.0)
    num_trials = int(num_trials)
    disc = _disc(discounter, sde_solver)
    s0 = float(torch.as_tensor(sde_solver.sde.init_value).reshape(-1)[0])
    start = time.time()
    bs = int(bs) if bs else num_trials
    remaining = num_trials
    s = q = 0.0
    b = None
    while remaining:
        n = min(bs, remaining)
        remaining -= n
        pay, x = sde_solver.simulate_payoffs(n, payoff, disc)
        cvar = disc * x[:, 0] - s0
        if b is None:
            b = sample_cov(cvar, pay) / cvar.var()
        g = (pay - b * cvar).double()
        s += float(g.sum())
        q += float((g * g).sum())
    mn = s / num_trials
    sd = math.sqrt(max(q / num_trials - mn * mn, 0.0) * num_trials / (num_trials - 1)) / \
        math.sqrt(num_trials)
    return MCStatistics(mn, sd, time.time() - start, num_trials)


def sample_batch_cost(solver, option, discounter, models, trials, bs, nn_bs):
    """Seconds per network batch of mc_apply_cvs (mc.py:414-416)."""
    out = mc_apply_cvs(models, solver, trials, option, discounter, bs, nn_bs)
    return out.time_elapsed / (trials / nn_bs)


def find_num_trials(problem, eps, models=None, init_trials=1e5, bs=1e5):
    """Pilot run, then n = ceil((1.96 se / eps)^2 n0) (mc.py:418-427)."""
    if models is None:
        stats = mc_simple(init_trials, problem.solver, problem.payoff, problem.discounter, bs,
                          payoff_time='adapted')
    else:
        stats = mc_apply_cvs(models, problem.solver, init_trials, problem.payoff,
                             problem.discounter, bs)
    ratio = (stats.sample_std * 1.96 / eps) ** 2
    return int(np.ceil(ratio * init_trials))


def run_mc(problem, eps, bs=1e5, init_trials=1e5):
    """Price to a 95% half-width eps (mc.py:437-440)."""
    trials = find_num_trials(problem, eps, None, init_trials, bs)
    return mc_simple(trials, problem.solver, problem.payoff, problem.discounter, bs=bs,
                     payoff_time='adapted')


def find_num_trials_terminal_cv(problem, eps, init_trials, bs):
    """Trials for a 95% half-width eps with the terminal control variate (mc.py:430-434)."""
    stats = mc_terminal_cv(init_trials, problem.solver, problem.payoff, problem.discounter, bs)
    return int(np.ceil((stats.sample_std * 1.96 / eps) ** 2 * init_trials))


def run_mc_terminal_cv(problem, eps, bs=1e5, init_trials=1e5):
    """Price to a 95% half-width eps with the terminal control variate (mc.py:470-472)."""
    trials = find_num_trials_terminal_cv(problem, eps, init_trials, bs)
    return mc_terminal_cv(trials, problem.solver, problem.payoff, problem.discounter, bs=bs)


def run_cv_mc(problem, models, opt, eps, train_size, step_factor=30, sim_bs=1e5, train_bs=1e3,
              nn_bs=1e3, epochs=10, early_stopping=False, print_losses=True, init_trials=1e5):
    """Train control variates on a grid step_factor times coarser, size the
    sample for a 95% half-width eps, then estimate (mc.py:443-467).  Returns
    (MCStatistics, training seconds, estimation seconds)."""
    es = None
    if early_stopping:
        cost = sample_batch_cost(problem.solver, problem.payoff, problem.discounter, models,
                                 sim_bs, sim_bs, nn_bs)
        es = EarlyStopping(eps, 1.96, cost, 1)
        es.batch_size = nn_bs
    steps = problem.solver.num_steps
    problem.solver.num_steps = int(np.ceil(steps / step_factor))
    t0 = time.time()
    sim_train_control_variates(models, opt, problem.solver, train_size, problem.payoff,
                               problem.discounter, sim_bs, train_bs, epochs, print_losses, 0, es)
    train_time = time.time() - t0
    problem.solver.num_steps = steps
    trials = ceil_mult(find_num_trials(problem, eps, models, init_trials, sim_bs), nn_bs)
    stats = mc_apply_cvs(models, problem.solver, trials, problem.payoff, problem.discounter,
                         sim_bs, nn_bs)
    test_time = stats.time_elapsed
    stats.time_elapsed += train_time
    return stats, train_time, test_time

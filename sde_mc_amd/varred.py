"""Training and application of neural control variates on stored paths
(API of sde_mc/varred.py).

These loops run in torch on whatever device the DataLoader's tensors live
on; the paths themselves come from the HIP kernels (solver.solve).  The
estimators in mc.py do not use `apply_*` when the networks can be packed:
they evaluate the networks inside the path kernel (cv.py) and store no path.

The control variate of a trajectory is
    Brownian:    sum_n disc(t_n) <f(t_n, X_n), dW_n>
    jumps:       sum_n disc(t_n) <g(t_n, X_n-), 1> J_n
    compensator: - lambda E[J] sum_n disc(t_n) <g(t_n, X_n-), 1> (t_{n+1} - t_n)
and gamma = discounted payoff + control variate; the networks are trained
to minimise the batch variance of gamma.
"""
import time

import numpy as np
import torch

from .helpers import partition, remove_steps


class EarlyStopping:
    """Stop training when an epoch's loss gain no longer pays for its cost
    (varred.py:7-20): stop iff delta < alpha cost_epoch eps^2 batch_size /
    (cost_batch quantile^2)."""

    def __init__(self, eps, quantile, cost_batch, alpha):
        self.eps = eps
        self.quantile = quantile
        self.cost_batch = cost_batch
        self.alpha = alpha
        self.cost_epoch = None
        self.batch_size = None

    def threshold(self):
        return (self.alpha * self.cost_epoch * self.eps ** 2 * self.batch_size /
                (self.cost_batch * self.quantile ** 2))

    def stop(self, delta_gamma):
        return delta_gamma < self.threshold()


def integrate_cv(normals, f_out, discounts, diffusion_struct, tol=0, time_interval=None):
    """Per-trajectory sum_n disc(t_n) <f_n, dW_n> (varred.py:202-214); with
    tol > 0 only the steps before time_interval - tol count."""
    if tol != 0:
        assert time_interval is not None
        keep = remove_steps(tol, normals.shape[1], time_interval)
        normals, f_out, discounts = normals[:, :keep], f_out[:, :keep], discounts[:, :keep]
    prod = normals * f_out
    if diffusion_struct != 'diag':
        prod = prod.sum(-1)          # (bs, steps, dim, k) -> (bs, steps, dim)
    return (prod * discounts).sum(-1).sum(-1)


def _point_inputs(model, times, states):
    """Network inputs: (bs, steps, 1 + dim) for sequential models, else
    flattened to (bs * steps, 1 + dim)."""
    bs, steps, dim = states.shape
    t = times.expand(bs, steps, 1) if times.dim() == 3 else times
    x = torch.cat([t, states], dim=-1)
    return x if model.sequential else x.reshape(bs * steps, dim + 1)


def _diffusion_gammas(model, paths, normals, payoffs, time_points, discounts, solver, tol):
    bs, steps, dim = paths.shape
    times = time_points.view(1, steps, 1).to(paths.dtype)
    f_out = model(_point_inputs(model, times, paths)).reshape(normals.shape)
    cv = integrate_cv(normals, f_out, discounts, solver.sde.diffusion_struct, tol=tol,
                      time_interval=solver.time_interval)
    return payoffs + cv


def _grid(solver, device):
    tp = partition(solver.time_interval, solver.num_steps, ends='left', device=device)
    return tp


def _epoch_report(epoch, loss, trials):
    print('{}: Train loss: {:.5f}     95% confidence interval: {:.5f}'.format(
        epoch, loss, np.sqrt(loss) * 2 / np.sqrt(trials)))


def _maybe_stop(early_stopping, losses, epoch, cost_total):
    if early_stopping is None or epoch == 0:
        return False
    early_stopping.cost_epoch = cost_total / (epoch + 1)
    return early_stopping.stop(losses[epoch - 1] - losses[epoch])


def train_diffusion_control_variate(model, opt, dl, solver, discounter, epochs, print_losses=True,
                                    tol=0, early_stopping=None):
    """Fit f on a NormalPathData loader (varred.py:23-72); returns (seconds,
    per-epoch mean batch variance)."""
    trials = dl.dataset.paths.shape[0]
    time_points = _grid(solver, dl.dataset.paths.device)
    discounts = discounter(time_points).view(1, -1, 1)
    losses = []
    cost_total = 0.0
    start = time.time()
    for epoch in range(epochs):
        t0 = time.time()
        model.train()
        total = 0.0
        for (paths, normals), payoffs in dl:
            opt.zero_grad()
            loss = _diffusion_gammas(model, paths, normals, payoffs, time_points, discounts,
                                     solver, tol).var()
            total += loss.item()
            loss.backward()
            opt.step()
        losses.append(total / len(dl))
        if print_losses:
            _epoch_report(epoch, losses[epoch], trials)
        model.eval()
        cost_total += time.time() - t0
        if _maybe_stop(early_stopping, losses, epoch, cost_total):
            break
    return time.time() - start, losses


def apply_diffusion_control_variate(model, dl, solver, discounter, tol=0):
    """(sum, sum of squares) of gamma over a NormalPathData loader (varred.py:75-95)."""
    time_points = _grid(solver, dl.dataset.paths.device)
    discounts = discounter(time_points).view(1, -1, 1)
    s = q = 0
    with torch.inference_mode():
        for (paths, normals), payoffs in dl:
            g = _diffusion_gammas(model, paths, normals, payoffs, time_points, discounts, solver,
                                  tol)
            s += g.sum()
            q += (g * g).sum()
    return s, q


def _adapted_gammas(models, batch, payoffs, solver, discounter, tol):
    f, g = models
    paths, normals, left_paths, time_paths, jump_paths = batch
    bs, steps, dim = paths.shape
    dt = torch.diff(time_paths, dim=1)
    discounts = discounter(time_paths)
    f_out = f(_point_inputs(f, time_paths, paths)).reshape(normals.shape)
    brownian = integrate_cv(normals, f_out, discounts, solver.sde.diffusion_struct, tol=tol,
                            time_interval=solver.time_interval)
    g_out = g(_point_inputs(g, time_paths, left_paths)).reshape(bs, steps, dim)
    jumps = (g_out * discounts * jump_paths).sum(-1).sum(-1)
    rate_mean = solver.sde.jump_rate() * solver.sde.jump_mean()
    comp = -(rate_mean * g_out[:, :-1] * discounts[:, :-1] * dt).sum(-1).sum(-1)
    return payoffs + brownian + jumps + comp


def apply_adapted_control_variates(models, dl, solver, discounter, tol=0):
    """(sum, sum of squares) of gamma over an AdaptedPathData loader (varred.py:98-131)."""
    s = q = 0
    with torch.inference_mode():
        for batch, payoffs in dl:
            gm = _adapted_gammas(models, batch, payoffs, solver, discounter, tol)
            s += gm.sum()
            q += (gm * gm).sum()
    return s, q


def train_adapted_control_variates(models, opt, dl, solver, discounter, epochs=10,
                                   print_losses=True, tol=0, early_stopping=None):
    """Fit [f, g] on an AdaptedPathData loader (varred.py:134-199); returns the
    per-epoch mean batch variances."""
    trials = dl.dataset.paths.shape[0]
    losses = []
    cost_total = 0.0
    for epoch in range(epochs):
        t0 = time.time()
        for m in models:
            m.train()
        total = 0.0
        for batch, payoffs in dl:
            opt.zero_grad()
            loss = _adapted_gammas(models, batch, payoffs, solver, discounter, tol).var()
            total += loss.item()
            loss.backward()
            opt.step()
        losses.append(total / len(dl))
        if print_losses:
            _epoch_report(epoch, losses[epoch], trials)
        for m in models:
            m.eval()
        cost_total += time.time() - t0
        if _maybe_stop(early_stopping, losses, epoch, cost_total):
            break
    return losses

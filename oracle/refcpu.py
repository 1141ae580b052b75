"""TEST INFRASTRUCTURE ONLY: torch-CPU restatement of the reference's solver
loop, op for op, for affine models (Merton, GBM, the Levy models).

This is the "reference CPU path" that bench.py times on the host cores and
that tests use to cross-check the C oracle's event grid: the same fp32
tensor operations per step as JumpDiffusionSolver.solve
(sde_mc/solvers.py:164-226) -- per-step host loop, dense path storage,
a jump mark drawn for every path at every step (solvers.py:211) -- with the
model's coefficients written as affine functions (sde.py:368-375,
levy.py:74-80, 113-120) and the payoff of mc.py:87-100.

Randomness comes either from a torch.Generator (timing) or from fixed host
arrays (cross-checks): normals [S][K][P] standard, jump times and marks
[J][P].
"""
import math

import numpy as np
import torch


class AffineModel:
    """Coefficients a0 + a1 x, b0 + b1 x, c0 + c1 x, j0 + j1 x per component."""

    def __init__(self, x0, a0, a1, b0, b1, c0=None, c1=None, j0=None, j1=None, chol=None,
                 noise='diag', rate=0.0, lognormal=(0.0, 0.0), icdf=None):
        d = len(x0)
        z = [0.0] * d
        f = lambda v: torch.tensor(v if v is not None else z, dtype=torch.float32)
        self.dim = d
        self.x0 = f(x0)
        self.a0, self.a1, self.b0, self.b1 = f(a0), f(a1), f(b0), f(b1)
        self.c0, self.c1, self.j0, self.j1 = f(c0), f(c1), f(j0), f(j1)
        self.L = torch.tensor(np.eye(d) if chol is None else chol, dtype=torch.float32)
        self.noise = noise
        self.rate = float(rate)
        self.lognormal = lognormal
        self.icdf = icdf

    def sample_marks(self, bs, gen):
        if self.icdf is not None:
            u = torch.rand((bs, 1), generator=gen) + 5.960464477539063e-08 / 3
            return self.icdf(u)
        alpha, gamma = self.lognormal
        return (torch.randn((bs, 1), generator=gen) * gamma + alpha).exp() - 1


class TruncatedPowerIcdf:
    """Inverse CDF of levy.py:22-30 in torch (fp32 in, fp32 out), restated
    for the CPU loop: exponential tails beyond |z| = 1 (rate mu, mu = inf
    removes them) and c |z|^{-1-alpha} on eps < |z| < 1."""

    def __init__(self, cm, cp, mu, alpha, eps):
        im = 0.0 if math.isinf(mu) else 1.0 / mu
        self.cm, self.cp, self.im, self.a, self.eps = cm, cp, im, alpha, eps
        self.lda = (cm + cp) * (im + (eps ** (-alpha) - 1) / alpha)
        self.y1 = cm * im / self.lda
        self.y2 = (cm * im + cm * (eps ** (-alpha) - 1) / alpha) / self.lda
        self.y3 = 1 - cp * im / self.lda

    def __call__(self, y):
        a, lda, cm, cp, im = self.a, self.lda, self.cm, self.cp, self.im
        epa = self.eps ** (-a)
        x2 = -(a * (lda * y / cm - im) + 1) ** (-1 / a)
        x3 = ((-a / cp) * (lda * y - cm * im - cm * (epa - 1) / a) + epa) ** (-1 / a)
        if im > 0:
            x1 = torch.log(lda * y / (cm * im)) * im - 1
            x4 = 1 - im * torch.log(lda * (1 - y) / (cp * im))
        else:
            x1 = x4 = torch.zeros_like(y)
        return torch.where(y <= self.y1, x1, torch.where(y < self.y2, x2,
                                                         torch.where(y < self.y3, x3, x4)))


def merton_model(mu, sigma, lam, alpha, gamma, x0=1.0):
    kbar = math.exp(alpha + 0.5 * gamma * gamma) - 1
    return AffineModel([x0], [0.0], [mu - lam * kbar], [0.0], [sigma], j1=[1.0], rate=lam,
                       lognormal=(alpha, gamma))


def solve_adapted(model, T, N, bs, max_jumps, gen=None, fixed=None, exact_jumps=False):
    """Paths (bs, total+1, d) and total_steps of the jump-adapted Euler loop.

    fixed = (z [S][K][bs] float32, jump_times [J][bs], marks [J][bs]) replaces
    the draws; row `step - 1` of z drives step `step` for every path."""
    d = model.dim
    h = torch.tensor(T / N)
    x = model.x0.unsqueeze(0).repeat(bs, 1)
    t = torch.zeros((bs, 1))
    paths = torch.zeros((bs, N + max_jumps + 1, d))
    paths[:, 0] = x
    if fixed is None:
        jump_times = torch.empty((bs, max_jumps, 1)).exponential_(model.rate, generator=gen)
        jump_times = jump_times.cumsum(dim=1) if model.rate > 0 else torch.full_like(jump_times, math.inf)
    else:
        jump_times = torch.as_tensor(np.ascontiguousarray(fixed[1].T)).unsqueeze(-1)
        marks = torch.as_tensor(np.ascontiguousarray(fixed[2].T))
        zfix = torch.as_tensor(fixed[0])
    jump_idx = torch.zeros((bs, 1), dtype=torch.long)
    total = 0
    zero = torch.tensor(0.)
    while torch.any(t < T):
        total += 1
        nxt = jump_times[torch.arange(bs), jump_idx.squeeze(-1), :]
        h = torch.minimum(h, torch.maximum(T - t, zero))
        dt = torch.minimum(h, nxt - t)
        sq = torch.sqrt(dt.unsqueeze(-1))
        if fixed is None:
            z = torch.randn((bs, d, 1), generator=gen)
            zc = torch.randn((bs, 1, 1), generator=gen) if model.noise == 'common' else None
        else:
            row = zfix[total - 1].T  # (bs, K)
            z = row[:, :d].unsqueeze(-1)
            zc = row[:, d:d + 1].unsqueeze(-1) if model.noise == 'common' else None
        corr = torch.matmul(model.L, z * sq).squeeze(-1)
        old_x = x
        drift = model.a0 + model.a1 * x
        diff = (model.b0 + model.b1 * x) * corr
        if zc is not None:
            common = (zc * sq).squeeze(-1).repeat(1, d)
            diff = diff + (model.c0 + model.c1 * x) * common
        x = x + drift * dt + diff
        t += dt
        if fixed is None:
            mark = model.sample_marks(bs, gen).repeat(1, d)
        else:
            idx = jump_idx.squeeze(-1).clamp(max=marks.shape[1] - 1)
            mark = marks[torch.arange(bs), idx].unsqueeze(-1).repeat(1, d)
        hit = torch.isclose(nxt, t, atol=1e-12)
        cur = torch.where(hit, mark, torch.zeros_like(mark))
        y = x if exact_jumps else old_x
        x = x + (model.j0 + model.j1 * y) * cur
        paths[:, total] = x
        jump_idx = torch.where(hit, jump_idx + 1, jump_idx)
    return paths[:, :total + 1], total


def price(model, payoff, disc, T, N, n, max_jumps, gen):
    """mc_simple without batching (mc.py:87-100): mean and standard error of
    disc * payoff(x_T), the terminal column of each adapted path."""
    paths, total = solve_adapted(model, T, N, n, max_jumps, gen=gen)
    pay = payoff(paths[:, total]) * disc
    return float(pay.mean()), float(pay.std() / math.sqrt(n))

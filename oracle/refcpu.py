"""TEST INFRASTRUCTURE ONLY: torch-CPU restatement of the reference's solver
loops, op for op, for affine models (Merton, GBM, the Levy models).

This is the "reference CPU path" that bench.py times on the host cores and
that tests use to cross-check the C oracle's event grid: the same fp32
tensor operations per step as JumpDiffusionSolver.solve
(sde_mc/solvers.py:164-226) -- per-step host loop, dense path storage,
a jump mark drawn for every path at every step (solvers.py:211) -- with the
model's coefficients written as affine functions (sde.py:368-375,
levy.py:74-80, 113-120) and the payoff of mc.py:87-100; and the coupled
fine/coarse loop of JumpDiffusionSolver.multilevel_solve
(solvers.py:228-307) on fixed draws (multilevel_adapted).

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


def model_from_kernel(km):
    """AffineModel carrying the coefficients of a kernel model description
    (an sdemc_model: fields dim, noise, rate, x0, a0, a1, b0, b1, c0, c1, j0,
    j1, chol row-major with stride 4).  Noise 1 (the jump solver's 'indep':
    one common second driver, solvers.py:198-201) maps to 'common'."""
    d = int(km.dim)
    v = lambda f: [float(getattr(km, f)[i]) for i in range(d)]
    chol = [[float(km.chol[4 * i + j]) for j in range(d)] for i in range(d)]
    return AffineModel(v('x0'), v('a0'), v('a1'), v('b0'), v('b1'), v('c0'), v('c1'), v('j0'),
                       v('j1'), chol=chol, noise='common' if int(km.noise) == 1 else 'diag',
                       rate=float(km.rate))


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


def _affine_step(model, x, dt, corr, common):
    """EulerScheme.step (schemes.py:5-10) for the affine coefficients:
    x + a(x) dt + b(x) dW (+ c(x) dB, the common second driver)."""
    diff = (model.b0 + model.b1 * x) * corr
    if common is not None:
        diff = diff + (model.c0 + model.c1 * x) * common
    return x + (model.a0 + model.a1 * x) * dt + diff


def multilevel_adapted(model, T, fine, coarse, bs, fixed, exact_jumps=False, plant=None):
    """JumpDiffusionSolver.multilevel_solve (solvers.py:228-307) op for op in
    torch fp32, on fixed draws.

    fixed = (z [S][K][bs] standard normals, jump_times [J][bs], marks [J][bs]):
    fine sub-step number k (counted from 0 over the whole loop, the
    reference's total_steps_fine - 1) takes row k of z; K = dim, plus one
    common column for noise 'common'.  Returns (x_fine [bs, d], x_coarse
    [bs, d], iterations [bs]): the terminal states and, per path, the loop
    iterations (one coarse step each) it took while its fine clock was below
    T -- the reference keeps stepping finished paths with zero-length fine
    steps until every path is done, which leaves them unchanged.

    plant: None, or the name of a deliberate mistake for tests that show a
    comparison with this loop can fail ('coarse_post_state': the coarse jump
    coefficient at the post-step state; 'coarse_untruncated': the coarse
    step not cut at the jump time; 'sum_unscaled': the coarse increment sums
    the unscaled normals of the fine sub-steps)."""
    d = model.dim
    factor = int(fine / coarse)
    h_fine = torch.tensor(T / fine)
    h_coarse = factor * h_fine
    t_fine = torch.zeros((bs, 1))
    t_coarse = torch.zeros((bs, 1))
    x_fine = model.x0.unsqueeze(0).repeat(bs, 1)
    x_coarse = model.x0.unsqueeze(0).repeat(bs, 1)
    zfix = torch.as_tensor(fixed[0])
    jump_times = torch.as_tensor(np.ascontiguousarray(fixed[1].T)).unsqueeze(-1)
    marks = torch.as_tensor(np.ascontiguousarray(fixed[2].T))
    jump_idx = torch.zeros((bs, 1), dtype=torch.long)
    iters = torch.zeros(bs, dtype=torch.long)
    zero = torch.tensor(0.)
    total_fine = 0
    while torch.any(t_fine < T):
        iters += (t_fine.squeeze(-1) < T).long()
        run_sum = 0
        run_sum_common = 0
        nxt = jump_times[torch.arange(bs), jump_idx.clamp(max=jump_times.shape[1] - 1).squeeze(-1), :]
        for _ in range(factor):
            total_fine += 1
            h_fine = torch.minimum(h_fine, torch.maximum(T - t_fine, zero))
            dt_fine = torch.minimum(h_fine, nxt - t_fine)
            sq = torch.sqrt(dt_fine.unsqueeze(-1))
            row = zfix[total_fine - 1].T  # (bs, K)
            z = row[:, :d].unsqueeze(-1)
            corr = torch.matmul(model.L, z * sq).squeeze(-1)
            common = None
            if model.noise == 'common':
                common = (row[:, d:d + 1].unsqueeze(-1) * sq).squeeze(-1).repeat(1, d)
            old_x_fine = x_fine
            x_fine = _affine_step(model, x_fine, dt_fine, corr, common)
            t_fine = t_fine + dt_fine
            if plant == 'sum_unscaled':
                run_sum = run_sum + torch.matmul(model.L, z).squeeze(-1)
            else:
                run_sum = run_sum + corr
            if common is not None:
                run_sum_common = run_sum_common + common
        h_coarse = torch.minimum(h_coarse, torch.maximum(T - t_coarse, zero))
        if plant == 'coarse_untruncated':
            dt_coarse = h_coarse
        else:
            dt_coarse = torch.minimum(h_coarse, nxt - t_coarse)
        old_x_coarse = x_coarse
        x_coarse = _affine_step(model, x_coarse, dt_coarse, run_sum,
                                run_sum_common if model.noise == 'common' else None)
        t_coarse = t_coarse + dt_coarse
        hit = torch.isclose(nxt, t_fine, atol=1e-12)
        idx = jump_idx.squeeze(-1).clamp(max=marks.shape[1] - 1)
        mark = marks[torch.arange(bs), idx].unsqueeze(-1).repeat(1, d)
        cur = torch.where(hit, mark, torch.zeros_like(mark))
        yf = x_fine if exact_jumps else old_x_fine
        yc = x_coarse if (exact_jumps or plant == 'coarse_post_state') else old_x_coarse
        x_fine = x_fine + (model.j0 + model.j1 * yf) * cur
        x_coarse = x_coarse + (model.j0 + model.j1 * yc) * cur
        jump_idx = torch.where(hit, jump_idx + 1, jump_idx)
    return x_fine, x_coarse, iters


def general_uniform(x0, a0, a1, B0, B1, chol, T, z, coarse_factor=0):
    """DiffusionSolver.solve / multilevel_solve (solvers.py:68-119) for a
    'general' model in fp64 torch, on fixed standard normals z [N][K][bs]
    (K >= brown_dim; columns past brown_dim unused).

    The increments are corr_normals = L (z sqrt(h)) (solvers.py:51-56) and the
    step is the matrix branch of EulerScheme.step (schemes.py:12) as intended,
    x + a(x) h + torch.matmul(B(x), dW), with a(x) = a0 + a1 x and
    B(x) = B0 + x[..., None] B1 of shape (bs, dim, brown_dim).  With
    coarse_factor M >= 2 the coarse path takes steps of M h with the sum of
    its M fine increments (solvers.py:104-119).  Returns x_fine [bs, dim]
    (and x_coarse)."""
    f64 = torch.float64
    L = torch.as_tensor(np.asarray(chol), dtype=f64)
    nb = L.shape[0]
    B0 = torch.as_tensor(np.asarray(B0), dtype=f64)
    B1 = torch.as_tensor(np.asarray(B1), dtype=f64)
    a0 = torch.as_tensor(np.asarray(a0), dtype=f64)
    a1 = torch.as_tensor(np.asarray(a1), dtype=f64)
    z = torch.as_tensor(np.asarray(z), dtype=f64)
    N, bs = z.shape[0], z.shape[2]
    h = T / N

    def step(x, dt, dW):
        B = B0 + x.unsqueeze(-1) * B1
        return x + (a0 + a1 * x) * dt + torch.matmul(B, dW.unsqueeze(-1)).squeeze(-1)

    x = torch.as_tensor(np.asarray(x0), dtype=f64).unsqueeze(0).repeat(bs, 1)
    xc = x.clone()
    M = int(coarse_factor)
    run = torch.zeros((bs, nb), dtype=f64)
    for n in range(N):
        dW = torch.matmul(L, z[n, :nb] * math.sqrt(h)).T  # (bs, brown_dim)
        x = step(x, h, dW)
        if M:
            run = run + dW
            if (n + 1) % M == 0:
                xc = step(xc, M * h, run)
                run = torch.zeros_like(run)
    return (x, xc) if M else x

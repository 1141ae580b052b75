"""Shared builders for tests: model jobs for the workloads and host-generated
fixed increments (numpy, seeded)."""
import ctypes
import math

import numpy as np

import sde_mc_amd as sm
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E

# the 1-D Merton workload: S0=K=1, r=0.02, sigma=0.2, T=1, lambda=2,
# log-jump N(-0.1, 0.15^2)
MERTON = dict(mu=0.02, sigma=0.2, rate=2.0, alpha=-0.1, gamma=0.15)
MERTON_PRICE = 0.13205167640082938


def merton_sde(mu=0.02, sigma=0.2, rate=2.0, alpha=-0.1, gamma=0.15, s0=1.0, dim=1, corr=None):
    import torch
    return sm.Merton(mu, sigma, rate, alpha, gamma, torch.tensor([s0] * dim), dim,
                     corr_matrix=corr)


def levy_rainbow_sde(rate_scale=1.0, rho=0.5, sigma=(0.2, 0.3), f=(1.0, 1.0), alpha=0.8,
                     eps=0.05, mu=math.inf):
    """The 2-D exponential-Levy rainbow model: sigma=(0.2,0.3), Brownian rho,
    truncated measure |z|^{-1-alpha} on eps<|z|<1 (mu = inf)."""
    import torch
    corr = torch.tensor([[1.0, rho], [rho, 1.0]])
    chol = torch.linalg.cholesky(corr)
    levy = sm.ExampleLevy(1.0, 1.0, alpha, mu, 0.02, torch.tensor(list(sigma)),
                          torch.tensor(list(f)), chol, eps, 2)
    return sm.LevySde(levy, torch.tensor([0.0, 0.0]), corr_matrix=corr)


def job_for(sde, T, steps, n_paths, grid='uniform', seed=1234, path_offset=0, coarse_factor=0,
            option=None, disc=1.0, max_steps=None, exact_jumps=False):
    model = sde.kernel_model()
    g = L.GRID_ADAPTED if grid == 'adapted' else L.GRID_UNIFORM
    if max_steps is None:
        max_steps = steps + E.safe_max_jumps(model.rate, T)
    job = E.make_job(model, T, steps, g, n_paths, seed, path_offset, coarse_factor=coarse_factor,
                     max_steps=max_steps, exact_jumps=exact_jumps)
    E.fill_payoff(job, option, disc)
    return job


def n_normals(job):
    d, nz = job.model.dim, job.model.noise
    if nz == L.NOISE_GENERAL:
        return L.MAX_DIM
    if nz in (L.NOISE_DIAG, L.NOISE_HESTON):
        return d
    return d + 1 if nz == L.NOISE_INDEP_COMMON else 2 * d


def mark_sampler(job, rng):
    m = job.model
    if m.jump_law == L.JUMP_LEVY:
        from sde_mc_amd.levy import InverseCdf
        cm, cp, im, al, eps = m.jp[0], m.jp[1], m.jp[2], m.jp[3], m.jp[4]
        icdf = InverseCdf(cm, cp, math.inf if im == 0 else 1.0 / im, al, eps)
        import torch

        def draw(n):
            y = torch.tensor(rng.random(n), dtype=torch.float32) + np.float32(5.960464477539063e-08 / 3)
            return icdf(y.double()).numpy().astype(np.float32)
        return draw
    a, g = m.jp[0], m.jp[1]
    return lambda n: (np.exp(a + g * rng.standard_normal(n)) - 1.0).astype(np.float32)


class FixedInputs:
    """Host arrays of a fixed-increment job plus ctypes views for the oracle."""

    def __init__(self, arrays, n_rows, max_jumps=0):
        self.a = arrays
        self.n_rows = n_rows
        self.max_jumps = max_jumps

    def host_struct(self):
        s = L.FixedIn()
        for name, field in (('dw', 'dw_dev'), ('counts', 'counts_dev'), ('off', 'jump_off_dev'),
                            ('jt', 'jump_time_dev'), ('js', 'jump_size_dev')):
            arr = self.a.get(name)
            setattr(s, field, arr.ctypes.data if arr is not None and arr.size else None)
        js = self.a.get('js')
        s.n_marks = int(js.size) if js is not None else 0
        s.n_rows = self.n_rows
        s.max_jumps = self.max_jumps
        return s

    def device_struct(self, device):
        import torch
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device)
                  for k, v in self.a.items() if v is not None and v.size}
        s = L.FixedIn()
        for name, field in (('dw', 'dw_dev'), ('counts', 'counts_dev'), ('off', 'jump_off_dev'),
                            ('jt', 'jump_time_dev'), ('js', 'jump_size_dev')):
            t = self.t.get(name)
            setattr(s, field, t.data_ptr() if t is not None else None)
        js = self.a.get('js')
        s.n_marks = int(js.size) if js is not None else 0
        s.n_rows = self.n_rows
        s.max_jumps = self.max_jumps
        return s


def fixed_uniform(job, seed=0):
    """dw [N][K][P] scaled+correlated fp32, counts [N][P] Poisson(lambda h),
    marks in CSR order (path-major)."""
    rng = np.random.default_rng(seed)
    P, N, T = int(job.n_paths), int(job.steps), float(job.T)
    d = job.model.dim
    K = n_normals(job)
    h = T / N
    nb = job.model.brown_dim if job.model.noise == L.NOISE_GENERAL else d
    L_ = np.array([[job.model.chol[i * 4 + j] for j in range(nb)] for i in range(nb)])
    z = rng.standard_normal((N, K, P)) * math.sqrt(h)
    dw = np.zeros((N, K, P), dtype=np.float32)
    dw[:, :nb, :] = np.einsum('ij,njp->nip', L_, z[:, :nb, :])
    if job.model.noise == L.NOISE_INDEP_COMMON:
        dw[:, d, :] = z[:, d, :]
    elif job.model.noise == L.NOISE_INDEP_FULL:
        dw[:, d:, :] = np.einsum('ij,njp->nip', L_, z[:, d:, :])
    lam = job.model.rate
    counts = rng.poisson(lam * h, size=(N, P)).astype(np.uint8) if lam > 0 else np.zeros((N, P), np.uint8)
    per_path = counts.sum(axis=0).astype(np.int64)
    off = np.zeros(P + 1, dtype=np.int64)
    off[1:] = np.cumsum(per_path)
    js = mark_sampler(job, rng)(int(off[-1])) if off[-1] > 0 else np.zeros(0, np.float32)
    return FixedInputs(dict(dw=dw, counts=counts, off=off, js=js), N)


def fixed_adapted(job, seed=0, extra_rows=0):
    """z [S][K][P] standard normals, jump times [J][P] (fp32 of fp64 cumulative
    Exp(lambda) gaps, last one beyond T), marks [J][P]."""
    rng = np.random.default_rng(seed)
    P, T = int(job.n_paths), float(job.T)
    M = max(int(job.coarse_factor), 1)
    S = int(job.max_steps) * M + extra_rows
    K = n_normals(job)
    z = rng.standard_normal((S, K, P)).astype(np.float32)
    lam = job.model.rate
    J = max(E.safe_max_jumps(lam, T), 1) if lam > 0 else 1
    if lam > 0:
        gaps = rng.exponential(1.0 / lam, size=(J, P))
        jt = np.cumsum(gaps, axis=0)
        jt[-1, :] = np.maximum(jt[-1, :], 2 * T)
        jt = jt.astype(np.float32)
    else:
        jt = np.full((J, P), np.inf, dtype=np.float32)
    js = mark_sampler(job, rng)(J * P).reshape(J, P) if lam > 0 else np.zeros((J, P), np.float32)
    return FixedInputs(dict(dw=z, jt=jt, js=js), S, J)


def reference_kats():
    """The reference's own known-answer vectors (tests/golden/reference_kats.json)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), 'golden', 'reference_kats.json')) as f:
        return json.load(f)


def payoff_kat_cases():
    """(kind, strike, terminal state, expected payoff) for every row of the
    reference's option known-answer tests."""
    k = reference_kats()
    out = []
    for p in k['payoffs']:
        for x, want in zip(k[p['terminals']], p['expected']):
            out.append((p['kind'], p['strike'], x, want))
    return out


def frozen_state_job(kind, strike, x):
    """One path, one uniform step, every coefficient zero and no jumps: the
    terminal state is x exactly, so the job's payoff is the payoff of x."""
    job = L.Job()
    m = job.model
    m.dim = len(x)
    m.noise = L.NOISE_DIAG
    m.jump_law = L.JUMP_LOGNORMAL
    m.rate = 0.0
    for i, v in enumerate(x):
        m.x0[i] = float(v)
    for i in range(L.MAX_DIM):
        m.chol[i * L.MAX_DIM + i] = 1.0
    job.payoff.kind = getattr(L, 'PAYOFF_' + kind)
    job.payoff.strike = float(strike)
    job.payoff.scale = 1.0
    job.payoff.disc = 1.0
    job.T = 1.0
    job.steps = 1
    job.grid = L.GRID_UNIFORM
    job.n_paths = 1
    job.seed = 5
    return job

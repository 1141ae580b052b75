"""The pricing workloads the library is measured on, as job builders.

  merton_1d      1-D Merton call: S0=K=1, r=0.02, sigma=0.2, T=1, lambda=2,
                 log-jump N(-0.1, 0.15^2); closed form 0.13205167640082938
  levy_rainbow   2-D exponential-Levy call on the max: sigma=(0.2, 0.3),
                 Brownian rho=0.5, r=0.02, truncated measure |z|^{-1-alpha},
                 alpha=0.8, on 0.05<|z|<1 (c+ = c- = 1), jump scale f=0.2 per
                 asset (the reference preset's scale, problem.py:158),
                 small jumps replaced by a Gaussian of matched variance
  request_mix    served requests: model Merton or Levy rainbow, strike
                 U[0.8,1.2], maturity {0.25,0.5,1,2}, steps log-uniform
                 [50,800], paths log-uniform [2^14,2^22], own seed each
"""
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from . import engine as E
from .levy import ExampleLevy, LevySde
from .options import EuroCall, Rainbow
from .sde import Merton

R = 0.02


def merton_1d():
    return Merton(R, 0.2, 2.0, -0.1, 0.15, torch.tensor([1.0]), 1)


def levy_rainbow(f=0.2, rho=0.5):
    corr = torch.tensor([[1.0, rho], [rho, 1.0]])
    chol = torch.linalg.cholesky(corr)
    levy = ExampleLevy(1.0, 1.0, 0.8, math.inf, R, torch.tensor([0.2, 0.3]),
                       torch.tensor([f, f]), chol, 0.05, 2)
    return LevySde(levy, torch.tensor([0.0, 0.0]), corr_matrix=corr)


_MODELS = {}


def model_description(name):
    """KernelModel of a named workload model (cached: the Levy compensator is
    a quadrature)."""
    if name not in _MODELS:
        sde = merton_1d() if name == 'merton' else levy_rainbow()
        _MODELS[name] = sde.kernel_model()
    return _MODELS[name]


@dataclass
class Request:
    rid: int
    model: str      # 'merton' | 'levy'
    strike: float
    maturity: float
    steps: int
    paths: int
    seed: int
    arrival: float = 0.0

    @property
    def path_steps(self):
        return self.paths * self.steps


def request_mix(n, seed=0, rate=None, paths_range=(14, 22), steps_range=(50, 800)):
    """n requests drawn as described above; with `rate` (req/s) Poisson arrival
    times are attached."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(n):
        model = 'merton' if rng.random() < 0.5 else 'levy'
        K = float(rng.uniform(0.8, 1.2))
        T = float(rng.choice([0.25, 0.5, 1.0, 2.0]))
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        paths = int(round(2 ** rng.uniform(*paths_range)))
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(i, model, K, T, steps, paths, int(rng.integers(1, 2 ** 62)), t))
    return out


def request_job(req, grid=L.GRID_UNIFORM):
    """The pricing job of one request: discounted call (Merton) or call on the
    max of the forwards (Levy rainbow), Philox-keyed by the request's seed."""
    model = model_description(req.model)
    T = req.maturity
    job = E.make_job(model, T, req.steps, grid, req.paths, req.seed, 0,
                     max_steps=req.steps + E.safe_max_jumps(model.rate, T))
    disc = math.exp(-R * T)
    if req.model == 'merton':
        E.fill_payoff(job, EuroCall(req.strike), disc)
    else:
        E.fill_payoff(job, Rainbow(req.strike, log=True, discount=math.exp(R * T)), disc)
    return job


# Algorithmic fp32 FLOPs per path-step (add, mul, compare, convert = 1, FMA = 2,
# each transcendental 1), the count the fp32 roofline prices:
#   1-D Merton: a Box-Muller normal amortised 5, the increment's scale 1, the
#     Euler update 4, the per-step jump test 2 (12 in all);
#   2-D Levy rainbow: 3 normals (two correlated Brownian, one common small-jump
#     Gaussian) 15, the 2x2 correlation 3, their scales 3, two asset updates
#     with the second driver 12, the jump test 2 (35 in all);
# plus 10 per expected jump (the mark's draw and its application) and 10 per
# path for the payoff and the reduction, spread over the steps.
def fp32_flops_per_path_step(model, steps, T):
    m = model_description(model) if isinstance(model, str) else model
    base = 12.0 if m.dim == 1 and m.noise == L.NOISE_DIAG else 35.0
    return base + 10.0 * m.rate * T / steps + 10.0 / steps


# VALU lane-operations per path-step, counted from the algorithm (not from a
# profile): Philox4x32-10 is 10 rounds of 2 mul-lo + 2 mul-hi + 4 xor per
# 4 words (key schedule is wave-uniform, on the scalar unit) = 20 per word;
# Box-Muller is 11 per pair of normals (2 cvt, 2 scale, log, mul, sqrt, sin,
# cos, 2 mul); the Euler update of one component is 5 (two affine
# coefficients, scale of dW, two fused multiply-adds); the jump-step test is
# 2.  Jumps add one Philox block, an fp64 log (~30) and the mark per jump.
def valu_ops_per_path_step(model, steps, T, dim=None):
    m = model_description(model) if isinstance(model, str) else model
    d = m.dim
    K = d if m.noise == L.NOISE_DIAG else (d + 1 if m.noise == L.NOISE_INDEP_COMMON else 2 * d)
    per_normal = 20.0 + 5.5
    ops = K * per_normal + 5.0 * d + 2.0
    if d > 1:
        ops += d * (d + 1) / 2.0          # correlate
    if m.noise != L.NOISE_DIAG:
        ops += 2.0 * d                     # second driver
    jumps_per_step = m.rate * T / steps
    mark = 12.0 if m.jump_law == L.JUMP_LEVY else 8.0
    ops += jumps_per_step * (80.0 + 30.0 + mark + 2.0 * d)
    return ops

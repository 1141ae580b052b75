"""Estimator helpers and host-side precomputations (API of sde_mc/helpers.py).

Host functions only: grid points, the estimator arithmetic on (sum, sum of
squares, n), the correlation-matrix builder and the Levy drift compensator,
which is computed once per model.
"""
import math

import numpy as np
import torch
from scipy.integrate import quad


def partition(interval, steps, ends='right', device='cpu'):
    """Points interval * i / steps of the uniform grid; `ends` chooses which
    of the two endpoints are kept: 'right' (default), 'left', 'both' or
    'none' (helpers.py:6-33)."""
    if ends not in ('right', 'left', 'both', 'none'):
        raise AssertionError("ends must be 'right', 'left', 'both' or 'none'")
    first = 0 if ends in ('left', 'both') else 1
    last = steps if ends in ('right', 'both') else steps - 1
    idx = range(first, last + 1)
    return torch.tensor([interval * i / steps for i in idx], device=device)


def solve_quadratic(coefficients):
    """The larger real root of a x^2 + b x + c = 0, elementwise over batched
    coefficient tensors (a, b, c) (helpers.py:36-48)."""
    a, b, c = coefficients
    root_disc = torch.sqrt(b * b - 4 * a * c)
    both = torch.stack([(root_disc - b) / (2 * a), -(root_disc + b) / (2 * a)])
    return both.amax(dim=0)


def mc_estimates(run_sum, run_sum_squares, n):
    """(mean, unbiased variance) of n samples given their sum and sum of
    squares (helpers.py:51-68)."""
    mean = run_sum / n
    return mean, (run_sum_squares - run_sum * mean) / (n - 1)


def remove_steps(time_tol, steps, time_interval):
    """Index of the last grid step before time_interval - time_tol on a grid
    of `steps` equal steps (helpers.py:71-74)."""
    return int(math.floor(steps - time_tol * steps / time_interval))


def get_corr_matrix(rhos):
    """n x n correlation matrix whose strict upper triangle, row by row, is
    `rhos` (n (n - 1) / 2 entries); raises RuntimeError when it is not
    positive definite (helpers.py:77-100)."""
    vals = [float(r) for r in (rhos.tolist() if torch.is_tensor(rhos) else rhos)]
    m = len(vals)
    n = (1 + math.isqrt(1 + 8 * m)) // 2
    if n * (n - 1) // 2 != m:
        raise AssertionError('Length of correlation vector is not triangular')
    corr = torch.eye(n)
    k = 0
    for i in range(n):
        for j in range(i + 1, n):
            corr[i, j] = corr[j, i] = vals[k]
            k += 1
    if torch.linalg.cholesky_ex(corr).info.item() != 0:
        raise RuntimeError('Matrix is not positive semidefinite')
    return corr


def ceil_mult(x, n):
    """Least multiple of n not below x (helpers.py:103-113)."""
    return int(math.ceil(float(x) / n)) * int(n)


def get_jump_comp(c_plus, c_minus, alpha, mu, f):
    """Drift compensator of f z under the Levy measure
        nu(dz) = c_- |z|^{-1-alpha} dz on (-1, 0), c_+ z^{-1-alpha} dz on (0, 1),
                 c_+- e^{-mu (|z| - 1)} dz beyond |z| = 1 (tempered tails),
    i.e. int (e^{f z} - 1 - f z 1{|z| < 1}) nu(dz) (helpers.py:116-133).

    The body is a quadrature of (expm1(f z) - f z) |z|^{-1-alpha}; the tails
    have the closed form c_+ (e^f / (mu - f) - 1 / mu) + c_- (e^{-f} / (mu + f)
    - 1 / mu), finite for mu > |f|; mu = inf truncates the measure to |z| < 1.
    """
    def body(z):
        return (math.expm1(f * z) - f * z) * abs(z) ** (-1.0 - alpha)

    total = c_minus * quad(body, -1.0, 0.0)[0] + c_plus * quad(body, 0.0, 1.0)[0]
    if math.isfinite(mu):
        if mu <= abs(f):
            return math.inf
        total += c_plus * (math.exp(f) / (mu - f) - 1.0 / mu)
        total += c_minus * (math.exp(-f) / (mu + f) - 1.0 / mu)
    return total


def sample_cov(x, y):
    """Unbiased sample covariance of two 1-D tensors."""
    return torch.cov(torch.stack([x, y]))[0, 1]

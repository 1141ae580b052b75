"""Estimator helpers and host-side precomputations (sde_mc/helpers.py).

These are host functions: scalar maths on the estimator sums, the correlation
matrix builder and the Levy drift compensator (a one-off scipy quadrature per
model, sde_mc/helpers.py:116-133).
"""
import numpy as np
import torch
from scipy.integrate import quad


def partition(interval, steps, ends='right', device='cpu'):
    """Grid points interval * i / steps with the chosen endpoints (helpers.py:6-33)."""
    assert ends in ['right', 'left', 'both', 'none']
    right = steps + 1 if ends in ['right', 'both'] else steps
    left = 0 if ends in ['left', 'both'] else 1
    return torch.tensor([interval * i / steps for i in range(left, right)], device=device)


def solve_quadratic(coefficients):
    """Larger root of a x^2 + b x + c = 0, batched (helpers.py:36-48)."""
    a, b, c = coefficients
    disc = torch.sqrt(b * b - 4 * a * c)
    return torch.maximum((-b + disc) / (2 * a), (-b - disc) / (2 * a))


def mc_estimates(run_sum, run_sum_squares, n):
    """Sample mean and unbiased variance from (sum, sum of squares, n) (helpers.py:51-68)."""
    sample_mean = run_sum / n
    sample_var = (run_sum_squares - (run_sum * run_sum) / n) / (n - 1)
    return sample_mean, sample_var


def remove_steps(time_tol, steps, time_interval):
    """Index of the last step once time_tol is cut from the interval (helpers.py:71-74)."""
    return int(np.floor(steps - time_tol / (time_interval / steps)))


def get_corr_matrix(rhos):
    """Correlation matrix from its upper-triangle entries (helpers.py:77-100).

    Raises RuntimeError when the matrix is not positive definite.
    """
    rhos_tensor = torch.tensor(rhos, dtype=torch.float32)
    t_n = len(rhos)
    n = (1 + int(round(np.sqrt(1 + 8 * t_n)))) // 2
    if n * (n - 1) // 2 != t_n:
        raise AssertionError('Length of correlation vector is not triangular')
    corr = torch.eye(n)
    iu = torch.triu_indices(row=n, col=n, offset=1)
    corr[iu[0], iu[1]] = rhos_tensor
    corr = corr + torch.triu(corr, 1).t()
    try:
        torch.linalg.cholesky(corr)
    except Exception:
        raise RuntimeError('Matrix is not positive semidefinite')
    return corr


def ceil_mult(x, n):
    """Smallest multiple of n that is >= x (helpers.py:103-113)."""
    return int(np.ceil(float(x) / n) * n)


def get_jump_comp(c_plus, c_minus, alpha, mu, f):
    """Compensator int (e^{f z} - 1 - f z 1{|z|<1}) nu(dz) of the Levy measure
    nu(dz) = c_- |z|^{-1-alpha} (-1<z<0), c_+ z^{-1-alpha} (0<z<1) with
    exponential tails c_+- e^{-mu(|z|-1)} beyond |z| = 1 (helpers.py:116-133).
    mu = inf truncates the measure to |z| < 1.
    """
    def g(x):
        return np.exp(f * x) - 1 - f * x

    def q(x):
        return np.exp(f * x) - 1

    def inner(x):
        return np.abs(x) ** (-alpha - 1)

    total = quad(lambda x: c_minus * g(x) * inner(x), -1, 0)[0]
    total += quad(lambda x: c_plus * g(x) * inner(x), 0, 1)[0]
    if np.isfinite(mu):
        def tail(x):
            return np.exp(-mu * (np.abs(x) - 1))
        total += quad(lambda x: c_minus * q(x) * tail(x), -np.inf, -1)[0]
        total += quad(lambda x: c_plus * q(x) * tail(x), 1, np.inf)[0]
    return total


def sample_cov(x, y):
    """Unbiased sample covariance (helpers.py:136-137)."""
    return torch.sum((x - x.mean()) * (y - y.mean())) / (len(x) - 1)

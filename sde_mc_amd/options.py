"""Payoffs, discounting and closed-form prices (sde_mc/options.py).

Each Option is a torch functor on terminal states, exactly as in the
reference; the HIP kernels evaluate the recognised ones in-kernel from
:meth:`Option.kernel_payoff`.  Other Option subclasses are applied in torch to
the terminal states the kernel returns.
"""
from abc import ABC, abstractmethod
from math import factorial

import numpy as np
import torch
from scipy.integrate import quad
from scipy.stats import lognorm, norm

from . import _lib as L


def bs_binary_aon(spot, strike, expiry, r, sigma):
    """Black-Scholes asset-or-nothing price, as normalised in options.py:8-32."""
    upper = (np.log(spot / strike) + (r + 0.5 * sigma * sigma) * expiry) / (np.sqrt(expiry) * sigma)
    return quad(lambda z: np.exp(-0.5 * z * z), -np.inf, upper)[0] / np.sqrt(2 * np.pi)


def bs_call(spot, strike, expiry, r, sigma):
    """Black-Scholes European call (options.py:35-58)."""
    d1 = (np.log(spot / strike) + (r + sigma ** 2 / 2) * expiry) / (sigma * np.sqrt(expiry))
    d2 = d1 - sigma * np.sqrt(expiry)
    return spot * norm.cdf(d1) - strike * np.exp(-r * expiry) * norm.cdf(d2)


def merton_call(spot, strike, expiry, r, sigma, alpha, gamma, rate, terms=40):
    """Merton jump-diffusion call as a Poisson mixture of Black-Scholes prices
    (options.py:61-99): sum_k e^{-lambda' T} (lambda' T)^k / k! BS(r_k, sigma_k)
    with lambda' = lambda (1 + kbar), r_k = r - lambda kbar + k ln(1 + kbar) / T and
    sigma_k^2 = sigma^2 + k gamma^2 / T.  Uses math.factorial (numpy >= 2 has
    no np.math)."""
    beta = np.exp(alpha + 0.5 * gamma * gamma) - 1
    total = 0.0
    for k in range(terms):
        r_k = r - rate * beta + (k * np.log(beta + 1)) / expiry
        sigma_k = np.sqrt(sigma ** 2 + (k * gamma ** 2) / expiry)
        lt = (beta + 1) * rate * expiry
        total += np.exp(-lt) * lt ** k / factorial(k) * bs_call(spot, strike, expiry, r_k, sigma_k)
    return total


def bs_digital_call(spot, strike, expiry, r, sigma):
    """Black-Scholes cash-or-nothing digital (options.py:102-125)."""
    mn = np.log(spot) + (r - 0.5 * sigma * sigma) * expiry
    sd = sigma * np.sqrt(expiry)
    return (1 - lognorm.cdf(strike, s=sd, scale=np.exp(mn))) * np.exp(-r * expiry)


def bs_asian_call(spot, strike, expiry, r, sigma):
    """Geometric-average Asian call under Black-Scholes (options.py:128-153)."""
    sig_g = sigma / np.sqrt(3)
    b = 0.5 * (r - 0.5 * sig_g ** 2)
    d1 = (np.log(spot / strike) + (b + sig_g ** 2 / 2) * expiry) / (sig_g * np.sqrt(expiry))
    d2 = d1 - sig_g * np.sqrt(expiry)
    return spot * np.exp((b - r) * expiry) * norm.cdf(d1) - strike * np.exp(-r * expiry) * norm.cdf(d2)


def _f(v):
    return float(v.detach().double().cpu()) if torch.is_tensor(v) else float(v)


class Option(ABC):
    """Payoff of a transformed terminal value: payoff(discount * (exp(x) if log else x))
    (options.py:156-176)."""

    kernel_kind = None

    def __init__(self, log=False, discount=1):
        self.log = log
        self.discount = discount

    def transform(self, x):
        spot = torch.exp(x) if self.log else x
        return self.discount * spot

    @abstractmethod
    def payoff(self, x):
        pass

    def __call__(self, x):
        return self.payoff(self.transform(x))

    def kernel_payoff(self, p):
        """Fill a ctypes sdemc_payoff; returns False if there is no in-kernel form."""
        owner = next(c for c in type(self).__mro__ if 'payoff' in c.__dict__)
        if self.kernel_kind is None or owner not in _PAYOFF_IMPL:
            return False
        p.kind = self.kernel_kind
        p.log = 1 if self.log else 0
        p.strike = _f(getattr(self, 'strike', 0.0))
        p.scale = _f(self.discount)
        return True


class EuroCall(Option):
    """European call (options.py:179-199)."""
    kernel_kind = L.PAYOFF_EURO_CALL

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        x = x[:, 0]
        return torch.where(x > self.strike, x - self.strike,
                           torch.tensor(0., dtype=x.dtype, device=x.device))


class EuroPut(Option):
    """European put (options.py:202-222)."""
    kernel_kind = L.PAYOFF_EURO_PUT

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        x = x[:, 0]
        return torch.where(x < self.strike, self.strike - x,
                           torch.tensor(0., dtype=x.dtype, device=x.device))


class BinaryAoN(Option):
    """Binary asset-or-nothing (options.py:225-244)."""
    kernel_kind = L.PAYOFF_BINARY_AON

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        x = x[:, 0]
        return torch.where(x >= self.strike, x, torch.tensor(0, dtype=x.dtype, device=x.device))


class Basket(Option):
    """Arithmetic or geometric basket call (options.py:247-259)."""

    def __init__(self, strike, average_type='arithmetic', log=False, discount=1):
        assert average_type in ['arithmetic', 'geometric']
        super().__init__(log, discount)
        self.strike = strike
        self.average_type = average_type
        self.kernel_kind = (L.PAYOFF_BASKET_ARITH if average_type == 'arithmetic'
                            else L.PAYOFF_BASKET_GEO)

    def payoff(self, x):
        spot = x.mean(1) if self.average_type == 'arithmetic' else torch.exp(torch.log(x).mean(1))
        return torch.where(spot > self.strike, spot - self.strike,
                           torch.tensor(0., dtype=spot.dtype, device=spot.device))


class Rainbow(Option):
    """Call on the maximum (options.py:262-272)."""
    kernel_kind = L.PAYOFF_RAINBOW

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        spot = x.max(1).values
        return torch.where(spot > self.strike, spot - self.strike,
                           torch.tensor(0., dtype=spot.dtype, device=spot.device))


class Digital(Option):
    """Digital (cash-or-nothing) call (options.py:275-284)."""
    kernel_kind = L.PAYOFF_DIGITAL

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        x = x[:, 0]
        return torch.where(x > self.strike, torch.ones_like(x), torch.zeros_like(x))


class AsianCall(Option):
    """Asian call on the running integral in component 1 (options.py:287-298)."""

    def __init__(self, time_interval, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.time_interval = time_interval
        self.strike = strike

    def payoff(self, x):
        spot = torch.exp(x[:, 1] / self.time_interval) if self.log else x[:, 1] / self.time_interval
        return torch.where(spot > self.strike, spot - self.strike,
                           torch.tensor(0., dtype=spot.dtype, device=spot.device))


class HestonRainbow(Option):
    """Call on the max of the even components (options.py:301-311)."""

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        even = torch.tensor([i for i in range(len(x[0])) if not i % 2], device=x.device)
        spot = torch.index_select(x, 1, even).max(1).values
        return torch.where(spot > self.strike, spot - self.strike,
                           torch.tensor(0., dtype=spot.dtype, device=spot.device))


class BestOf(Option):
    """max(max_i S_i, K) (options.py:314-321)."""
    kernel_kind = L.PAYOFF_BEST_OF

    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike

    def payoff(self, x):
        m = torch.max(x, dim=1).values
        return torch.maximum(m, self.strike * torch.ones_like(m))


class ConstantShortRate:
    """Discount factor exp(-r t) (options.py:324-337)."""

    def __init__(self, r):
        self.r = r

    def __call__(self, t):
        if not torch.is_tensor(t):
            t = torch.tensor(t)
        return torch.exp(-t * self.r)


# payoff implementations the kernels know; a subclass overriding payoff()
# is applied in torch instead
_PAYOFF_IMPL = (EuroCall, EuroPut, BinaryAoN, Basket, Rainbow, Digital, BestOf)

"""Payoffs, discounting and closed-form prices (API of sde_mc/options.py).

An Option maps terminal states x (P, d) to payoffs (P,): the state is first
exponentiated when ``log`` is set and multiplied by ``discount``, then the
payoff is applied.  The HIP kernels evaluate the payoffs listed in
``_IN_KERNEL`` themselves (``Option.kernel_payoff`` fills the C descriptor);
any other Option, including a subclass that overrides ``payoff``, is applied
in torch to the terminal states the kernel returns.
"""
import math
from abc import ABC, abstractmethod

import numpy as np
import torch
from scipy.special import ndtr

from . import _lib as L


# ---- closed forms --------------------------------------------------------------
def _d1_d2(spot, strike, expiry, r, sigma):
    vol = sigma * math.sqrt(expiry)
    d1 = (math.log(spot / strike) + (r + 0.5 * sigma * sigma) * expiry) / vol
    return d1, d1 - vol


def bs_binary_aon(spot, strike, expiry, r, sigma):
    """N(d1) of Black-Scholes: the asset-or-nothing price per unit of spot, the
    normalisation of options.py:8-32."""
    return float(ndtr(_d1_d2(spot, strike, expiry, r, sigma)[0]))


def bs_call(spot, strike, expiry, r, sigma):
    """Black-Scholes European call (options.py:35-58)."""
    d1, d2 = _d1_d2(spot, strike, expiry, r, sigma)
    return float(spot * ndtr(d1) - strike * math.exp(-r * expiry) * ndtr(d2))


def merton_call(spot, strike, expiry, r, sigma, alpha, gamma, rate, terms=40):
    """Merton (1976) call with log-normal jumps log(1 + J) ~ N(alpha, gamma^2):
    conditional on n jumps the price is Black-Scholes with variance
    sigma^2 + n gamma^2 / T and rate r - lambda kbar + n log(1 + kbar) / T,
    kbar = E[J]; the n-th term is weighted by the Poisson(lambda (1 + kbar) T)
    probability of n (options.py:61-99, first `terms` terms)."""
    kbar = math.expm1(alpha + 0.5 * gamma * gamma)
    mass = rate * (1.0 + kbar) * expiry
    price = 0.0
    for n in range(terms):
        if mass > 0:
            weight = math.exp(n * math.log(mass) - mass - math.lgamma(n + 1))
        else:
            weight = 1.0 if n == 0 else 0.0
        vol_n = math.sqrt(sigma * sigma + n * gamma * gamma / expiry)
        r_n = r - rate * kbar + n * math.log1p(kbar) / expiry
        price += weight * bs_call(spot, strike, expiry, r_n, vol_n)
    return price


def bs_digital_call(spot, strike, expiry, r, sigma):
    """Cash-or-nothing call e^{-rT} P(S_T > K) under Black-Scholes
    (options.py:102-125)."""
    return float(math.exp(-r * expiry) * ndtr(_d1_d2(spot, strike, expiry, r, sigma)[1]))


def bs_asian_call(spot, strike, expiry, r, sigma):
    """Continuously monitored geometric-average Asian call under Black-Scholes
    (Kemna-Vorst): the average is log-normal with volatility sigma / sqrt(3)
    and cost of carry (r - sigma^2 / 6) / 2 (options.py:128-153)."""
    vol_g = sigma / math.sqrt(3.0)
    carry = 0.5 * (r - 0.5 * vol_g * vol_g)
    sd = vol_g * math.sqrt(expiry)
    d1 = (math.log(spot / strike) + (carry + 0.5 * vol_g * vol_g) * expiry) / sd
    fwd = spot * math.exp((carry - r) * expiry)
    return float(fwd * ndtr(d1) - strike * math.exp(-r * expiry) * ndtr(d1 - sd))


def _f(v):
    return float(v.detach().double().cpu()) if torch.is_tensor(v) else float(v)


def _floor_at_zero(v):
    return torch.clamp(v, min=0.0)


# ---- options ---------------------------------------------------------------------
class Option(ABC):
    """payoff(discount * (exp(x) if log else x)) on terminal states x (P, d)
    (options.py:156-176)."""

    kernel_kind = None

    def __init__(self, log=False, discount=1):
        self.log = log
        self.discount = discount

    def transform(self, x):
        return self.discount * (x.exp() if self.log else x)

    @abstractmethod
    def payoff(self, x):
        """Payoff of transformed states (P, d) -> (P,)."""

    def __call__(self, x):
        return self.payoff(self.transform(x))

    def kernel_payoff(self, p):
        """Fill a ctypes sdemc_payoff; False when the kernels cannot evaluate
        this payoff (it is then applied in torch)."""
        impl = next(c for c in type(self).__mro__ if 'payoff' in c.__dict__)
        if self.kernel_kind is None or impl not in _IN_KERNEL:
            return False
        p.kind = self.kernel_kind
        p.log = int(bool(self.log))
        p.strike = _f(getattr(self, 'strike', 0.0))
        p.scale = _f(self.discount)
        return True


class _Struck(Option):
    def __init__(self, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.strike = strike


class EuroCall(_Struck):
    """(S_0 - K)^+ (options.py:179-199)."""
    kernel_kind = L.PAYOFF_EURO_CALL

    def payoff(self, x):
        return _floor_at_zero(x[:, 0] - self.strike)


class EuroPut(_Struck):
    """(K - S_0)^+ (options.py:202-222)."""
    kernel_kind = L.PAYOFF_EURO_PUT

    def payoff(self, x):
        return _floor_at_zero(self.strike - x[:, 0])


class BinaryAoN(_Struck):
    """S_0 1{S_0 >= K} (options.py:225-244)."""
    kernel_kind = L.PAYOFF_BINARY_AON

    def payoff(self, x):
        s = x[:, 0]
        return s * (s >= self.strike).to(s.dtype)


class Basket(Option):
    """Call on the arithmetic or geometric mean of the components
    (options.py:247-259)."""

    def __init__(self, strike, average_type='arithmetic', log=False, discount=1):
        if average_type not in ('arithmetic', 'geometric'):
            raise AssertionError("average_type must be 'arithmetic' or 'geometric'")
        super().__init__(log, discount)
        self.strike = strike
        self.average_type = average_type
        self.kernel_kind = (L.PAYOFF_BASKET_ARITH if average_type == 'arithmetic'
                            else L.PAYOFF_BASKET_GEO)

    def payoff(self, x):
        if self.average_type == 'arithmetic':
            avg = x.mean(dim=1)
        else:
            avg = x.log().mean(dim=1).exp()
        return _floor_at_zero(avg - self.strike)


class Rainbow(_Struck):
    """Call on the maximum, (max_i S_i - K)^+ (options.py:262-272)."""
    kernel_kind = L.PAYOFF_RAINBOW

    def payoff(self, x):
        return _floor_at_zero(x.amax(dim=1) - self.strike)


class Digital(_Struck):
    """1{S_0 > K} (options.py:275-284)."""
    kernel_kind = L.PAYOFF_DIGITAL

    def payoff(self, x):
        s = x[:, 0]
        return (s > self.strike).to(s.dtype)


class AsianCall(Option):
    """Call on the time average A_T / T held in component 1, exponentiated
    when ``log`` (options.py:287-298)."""

    def __init__(self, time_interval, strike, log=False, discount=1):
        super().__init__(log, discount)
        self.time_interval = time_interval
        self.strike = strike

    kernel_kind = L.PAYOFF_ASIAN_CALL

    def payoff(self, x):
        avg = x[:, 1] / self.time_interval
        if self.log:
            avg = avg.exp()
        return _floor_at_zero(avg - self.strike)

    def kernel_payoff(self, p):
        if not super().kernel_payoff(p):
            return False
        p.param = _f(self.time_interval)
        return True


class HestonRainbow(_Struck):
    """Call on the maximum of the price components 0, 2, 4, ... of a
    price/variance state (options.py:301-311)."""
    kernel_kind = L.PAYOFF_RAINBOW_EVEN

    def payoff(self, x):
        return _floor_at_zero(x[:, 0::2].amax(dim=1) - self.strike)


class BestOf(_Struck):
    """max(max_i S_i, K) (options.py:314-321)."""
    kernel_kind = L.PAYOFF_BEST_OF

    def payoff(self, x):
        return torch.clamp(x.amax(dim=1), min=float(self.strike))


class ConstantShortRate:
    """Discount factor exp(-r t) (options.py:324-337)."""

    def __init__(self, r):
        self.r = r

    def __call__(self, t):
        return torch.exp(-self.r * torch.as_tensor(t))


# payoff implementations the kernels evaluate themselves
_IN_KERNEL = (EuroCall, EuroPut, BinaryAoN, Basket, Rainbow, Digital, BestOf, AsianCall,
              HestonRainbow)

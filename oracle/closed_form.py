"""TEST INFRASTRUCTURE ONLY: fp64 closed forms that pin the simulator.

Each function names the reference routine it restates (or the textbook
result it adds) and is itself pinned in tests/test_oracle_closed_form.py:
  bs_call            options.py:35-58   golden tests/test_options.py:12
  bs_binary_aon      options.py:8-32    golden tests/test_options.py:6
  merton_call        options.py:61-99   golden tests/test_options.py:18; also
                                        checked against merton_call_fourier
  merton_call_fourier  Gil-Pelaez inversion of the Merton characteristic
                       function (independent of the Poisson-series form)
  stulz_call_max     call on the max of two lognormal assets (Stulz 1982,
                     Johnson 1987), checked against call_max_quadrature
  call_max_quadrature  the same price by conditioning and 1-D quadrature
  heston_cf / heston_call  Heston (1993) call by Gil-Pelaez probabilities
                       (the reference has no Heston closed form); checked
                       against lewis_call on the same characteristic function,
                       the Black-Scholes limit and put-call parity
  lewis_call         Lewis (2001) single-integral call price from any
                     characteristic function; checked against bs_call
"""
import math

import numpy as np
from scipy.integrate import quad
from scipy.stats import norm


def bs_call(S, K, T, r, sigma):
    d1 = (math.log(S / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * math.sqrt(T))
    d2 = d1 - sigma * math.sqrt(T)
    return S * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)


def bs_binary_aon(S, K, T, r, sigma):
    """The reference's normalisation: N(d1) (options.py:29-32)."""
    d1 = (math.log(S / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * math.sqrt(T))
    return float(norm.cdf(d1))


def merton_call(S, K, T, r, sigma, alpha, gamma, lam, terms=50):
    """sum_k Pois(k; lam' T) BS(S, K, T, r_k, sigma_k) with lam' = lam (1 + kbar),
    r_k = r - lam kbar + k log(1 + kbar) / T, sigma_k^2 = sigma^2 + k gamma^2 / T."""
    kbar = math.exp(alpha + 0.5 * gamma * gamma) - 1.0
    lt = lam * (1.0 + kbar) * T
    total = 0.0
    for k in range(terms):
        w = math.exp(-lt + k * math.log(lt) - math.lgamma(k + 1)) if lt > 0 else (1.0 if k == 0 else 0.0)
        r_k = r - lam * kbar + k * math.log1p(kbar) / T
        s_k = math.sqrt(sigma * sigma + k * gamma * gamma / T)
        total += w * bs_call(S, K, T, r_k, s_k)
    return total


def merton_cf(u, S, T, r, sigma, alpha, gamma, lam):
    """E[exp(i u log S_T)] under the risk-neutral Merton dynamics."""
    kbar = math.exp(alpha + 0.5 * gamma * gamma) - 1.0
    mu = math.log(S) + (r - lam * kbar - 0.5 * sigma * sigma) * T
    return np.exp(1j * u * mu - 0.5 * sigma * sigma * u * u * T
                  + lam * T * (np.exp(1j * u * alpha - 0.5 * gamma * gamma * u * u) - 1.0))


def merton_call_fourier(S, K, T, r, sigma, alpha, gamma, lam):
    """C = S P1 - K e^{-rT} P2 with P_j = 1/2 + 1/pi int_0^inf Re[e^{-iu ln K} f_j(u) / (iu)] du,
    f_2 = phi, f_1(u) = phi(u - i) / phi(-i)."""
    lk = math.log(K)
    phi_mi = merton_cf(-1j, S, T, r, sigma, alpha, gamma, lam)

    def integrand(u, j):
        if j == 1:
            f = merton_cf(u - 1j, S, T, r, sigma, alpha, gamma, lam) / phi_mi
        else:
            f = merton_cf(u, S, T, r, sigma, alpha, gamma, lam)
        return (np.exp(-1j * u * lk) * f / (1j * u)).real

    p1 = 0.5 + quad(integrand, 1e-12, 200, args=(1,), limit=400, epsabs=1e-13)[0] / math.pi
    p2 = 0.5 + quad(integrand, 1e-12, 200, args=(2,), limit=400, epsabs=1e-13)[0] / math.pi
    return S * p1 - K * math.exp(-r * T) * p2


def bvn_cdf(a, b, rho):
    """P(X <= a, Y <= b) for standard normals with correlation rho (1-D quadrature)."""
    s = math.sqrt(1.0 - rho * rho)
    f = lambda x: norm.pdf(x) * norm.cdf((b - rho * x) / s)
    return quad(f, -np.inf, a, epsabs=1e-14, epsrel=1e-12, limit=200)[0]


def stulz_call_max(S1, S2, K, T, r, s1, s2, rho):
    """European call on max(S1_T, S2_T), two correlated GBMs with drift r."""
    sq = math.sqrt(T)
    sig = math.sqrt(s1 * s1 + s2 * s2 - 2 * rho * s1 * s2)
    d = (math.log(S1 / S2) + 0.5 * sig * sig * T) / (sig * sq)
    y1 = (math.log(S1 / K) + (r + 0.5 * s1 * s1) * T) / (s1 * sq)
    y2 = (math.log(S2 / K) + (r + 0.5 * s2 * s2) * T) / (s2 * sq)
    rho1 = (s1 - rho * s2) / sig
    rho2 = (s2 - rho * s1) / sig
    return (S1 * bvn_cdf(y1, d, rho1) + S2 * bvn_cdf(y2, -d + sig * sq, rho2)
            - K * math.exp(-r * T) * (1.0 - bvn_cdf(-y1 + s1 * sq, -y2 + s2 * sq, rho)))


def call_max_quadrature(S1, S2, K, T, r, s1, s2, rho):
    """The same price by conditioning on the first driver z1: given z1 the
    payoff is max(m, b) - K with m = max(a(z1), K) and b lognormal in z2, whose
    expectation is m N(c) + B e^{v^2/2} N(v - c); then 1-D quadrature in z1."""
    sq = math.sqrt(T)
    v = s2 * sq * math.sqrt(1 - rho * rho)

    def inner(z1):
        a = S1 * math.exp((r - 0.5 * s1 * s1) * T + s1 * sq * z1)
        B = S2 * math.exp((r - 0.5 * s2 * s2) * T + s2 * sq * rho * z1)
        m = max(a, K)
        c = math.log(m / B) / v
        return (m * norm.cdf(c) + B * math.exp(0.5 * v * v) * norm.cdf(v - c) - K) * norm.pdf(z1)

    z_kink = (math.log(K / S1) - (r - 0.5 * s1 * s1) * T) / (s1 * sq)
    lo = quad(inner, -40.0, z_kink, epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    hi = quad(inner, z_kink, 40.0, epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    return math.exp(-r * T) * (lo + hi)


# ---- Heston (1993) ---------------------------------------------------------
def heston_cf(u, S, T, r, v0, kappa, theta, xi, rho):
    """E[exp(i u log S_T)] under the Heston model, in the rotation-free form
    of Albrecher, Mayer, Schoutens and Tistaert (2007), which keeps the complex
    logarithm on its principal branch; u may be complex."""
    u = np.asarray(u, dtype=complex)
    iu = 1j * u
    beta = kappa - rho * xi * iu
    d = np.sqrt(beta * beta + xi * xi * (iu + u * u))
    g = (beta - d) / (beta + d)
    e = np.exp(-d * T)
    C = kappa * theta / (xi * xi) * ((beta - d) * T - 2.0 * np.log((1.0 - g * e) / (1.0 - g)))
    D = (beta - d) / (xi * xi) * (1.0 - e) / (1.0 - g * e)
    return np.exp(iu * (math.log(S) + r * T) + C + D * v0)


def heston_call(S, K, T, r, v0, kappa, theta, xi, rho):
    """Heston call by the two Gil-Pelaez probabilities:
    S P1 - K e^{-rT} P2, P_j = 1/2 + (1/pi) int_0^inf Re[e^{-iu log K} f_j(u) / (iu)] du
    with f_2 = phi, f_1(u) = phi(u - i) / phi(-i)."""
    lk = math.log(K)
    fwd = S * math.exp(r * T)

    def integrand(u, j):
        if j == 1:
            f = heston_cf(u - 1j, S, T, r, v0, kappa, theta, xi, rho) / fwd
        else:
            f = heston_cf(u, S, T, r, v0, kappa, theta, xi, rho)
        return (np.exp(-1j * u * lk) * f / (1j * u)).real

    p1 = 0.5 + quad(integrand, 1e-12, 200.0, args=(1,), limit=400)[0] / math.pi
    p2 = 0.5 + quad(integrand, 1e-12, 200.0, args=(2,), limit=400)[0] / math.pi
    return S * p1 - K * math.exp(-r * T) * p2


def lewis_call(S, K, T, r, cf):
    """Call price by Lewis' (2001) single integral from the characteristic
    function cf(u) = E[exp(i u log S_T)] of any model with forward S e^{rT}:
        C = S - (sqrt(K) e^{-rT} / pi) int_0^inf Re[e^{-iu log K} cf(u - i/2)] / (u^2 + 1/4) du.
    An integration path independent of heston_call's two probabilities."""
    lk = math.log(K)

    def integrand(u):
        z = cf(u - 0.5j)
        return (np.exp(-1j * u * lk) * z).real / (u * u + 0.25)

    val = quad(integrand, 0.0, 200.0, limit=400)[0]
    return S - math.sqrt(K) * math.exp(-r * T) / math.pi * val

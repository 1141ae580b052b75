"""The oracle's closed forms against the reference's golden values
(tests/test_options.py of sde_mc) and against independent computations."""
import math

import numpy as np
import pytest

from oracle import closed_form as cf


def test_bs_call_golden():
    # sde_mc tests/test_options.py:11-14
    assert math.isclose(cf.bs_call(1, 1, 3, 0.02, 0.2), 0.1646004265, rel_tol=1e-5)


def test_bs_binary_aon_golden():
    # sde_mc tests/test_options.py:5-8
    assert math.isclose(cf.bs_binary_aon(1, 1, 3, 0.02, 0.2), 0.63548275523, rel_tol=1e-9)


def test_merton_call_golden():
    # sde_mc tests/test_options.py:17-20: 0.36328189504657027
    assert math.isclose(cf.merton_call(1, 1, 3, 0.02, 0.3, -0.05, 0.3, 2), 0.36328189504657027,
                        rel_tol=1e-13)


@pytest.mark.parametrize("args", [
    (1, 1, 3, 0.02, 0.3, -0.05, 0.3, 2),
    (1, 1, 1, 0.02, 0.2, -0.1, 0.15, 2.0),
    (1, 0.8, 0.25, 0.02, 0.2, -0.1, 0.15, 2.0),
    (1, 1.2, 2.0, 0.05, 0.35, 0.05, 0.4, 0.5),
])
def test_merton_series_matches_fourier(args):
    # two unrelated derivations: Poisson mixture of BS prices vs Gil-Pelaez
    # inversion of the characteristic function
    assert abs(cf.merton_call(*args) - cf.merton_call_fourier(*args)) < 1e-9


def test_merton_config_value():
    # the Merton call of the 1-D workloads (S0=K=1, r=0.02, sigma=0.2, T=1,
    # lambda=2, log-jump N(-0.1, 0.15^2))
    assert math.isclose(cf.merton_call(1, 1, 1, 0.02, 0.2, -0.1, 0.15, 2.0),
                        0.13205167640082938, rel_tol=1e-13)


def test_merton_reduces_to_bs():
    assert abs(cf.merton_call(1, 1.1, 2, 0.03, 0.25, -0.1, 0.15, 0.0) -
               cf.bs_call(1, 1.1, 2, 0.03, 0.25)) < 1e-15


@pytest.mark.parametrize("args", [
    (1, 1, 1, 1, 0.02, 0.2, 0.3, 0.5),
    (1.1, 0.9, 1.05, 2, 0.03, 0.25, 0.15, -0.3),
    (1, 1, 0.8, 0.25, 0.02, 0.2, 0.3, 0.9),
])
def test_stulz_matches_quadrature(args):
    assert abs(cf.stulz_call_max(*args) - cf.call_max_quadrature(*args)) < 1e-10


def test_stulz_limits():
    # rho -> 1 with equal vols: max of two identical assets is one asset
    v = cf.call_max_quadrature(1, 1, 1, 1, 0.02, 0.2, 0.2, 0.999999)
    assert abs(v - cf.bs_call(1, 1, 1, 0.02, 0.2)) < 2e-4


def test_golden_fixture_closed_forms():
    from tests.jobs import reference_kats
    k = reference_kats()['closed_forms']
    for name, rel in (('bs_call', 1e-5), ('bs_binary_aon', 1e-9), ('merton_call', 1e-13)):
        assert math.isclose(getattr(cf, name)(*k[name]['args']), k[name]['expected'], rel_tol=rel)


def test_oracle_payoffs_match_reference_option_kats():
    # the reference's option known answers (tests/test_options.py:23-50),
    # through the oracle's whole path with the state frozen at the terminal
    import numpy as np
    import oracle
    from tests.jobs import frozen_state_job, payoff_kat_cases
    for kind, K, x, want in payoff_kat_cases():
        pay, term, _ = oracle.simulate(frozen_state_job(kind, K, x), 0, 1)
        assert np.allclose(term[0, :, 0], np.float32(x)), (kind, x)
        assert abs(pay[0] - want) < 1e-6, (kind, x, pay[0], want)


HESTON = (1.0, 1.0, 1.0, 0.02, 0.04, 1.5, 0.04, 0.3, -0.7)   # S K T r v0 kappa theta xi rho


def test_lewis_integral_matches_black_scholes():
    for S, K, T, r, s in ((1.0, 0.9, 1.0, 0.02, 0.25), (1.0, 1.3, 2.0, 0.05, 0.4)):
        bscf = lambda u: np.exp(1j * u * (math.log(S) + (r - 0.5 * s * s) * T) - 0.5 * u * u * s * s * T)
        assert abs(cf.lewis_call(S, K, T, r, bscf) - cf.bs_call(S, K, T, r, s)) < 1e-10


@pytest.mark.parametrize("K", [0.7, 1.0, 1.4])
def test_heston_two_integrations_agree(K):
    S, _, T, r, v0, ka, th, xi, rho = HESTON
    a = cf.heston_call(S, K, T, r, v0, ka, th, xi, rho)
    b = cf.lewis_call(S, K, T, r, lambda u: cf.heston_cf(u, S, T, r, v0, ka, th, xi, rho))
    assert abs(a - b) < 1e-10


def test_heston_limits_and_parity():
    S, K, T, r, v0, ka, th, xi, rho = HESTON
    # martingale: E[S_T] = S e^{rT}
    assert abs(cf.heston_cf(-1j, S, T, r, v0, ka, th, xi, rho) - S * math.exp(r * T)) < 1e-12
    # vanishing vol of vol with v0 = theta: Black-Scholes at sigma^2 = theta,
    # approached linearly in xi
    d3 = cf.heston_call(S, 0.9, T, r, th, ka, th, 1e-3, rho) - cf.bs_call(S, 0.9, T, r, math.sqrt(th))
    d4 = cf.heston_call(S, 0.9, T, r, th, ka, th, 1e-4, rho) - cf.bs_call(S, 0.9, T, r, math.sqrt(th))
    assert abs(d4) < 2e-6 and abs(d3 / d4 - 10.0) < 0.5
    # put-call parity through the closed-form put (K e^{-rT} - S + C)
    for K_ in (0.8, 1.2):
        c = cf.heston_call(S, K_, T, r, v0, ka, th, xi, rho)
        p_int = cf.lewis_call(S, K_, T, r, lambda u: cf.heston_cf(u, S, T, r, v0, ka, th, xi, rho)) \
            - S + K_ * math.exp(-r * T)
        assert p_int > 0 and abs(c - (p_int + S - K_ * math.exp(-r * T))) < 1e-10

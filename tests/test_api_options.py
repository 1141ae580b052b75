"""The package's Option classes and closed forms (the drop-in API of
sde_mc/options.py) against the reference's own known answers and, where
the reference holds none, against an independent computation."""
import math

import numpy as np
import pytest
import torch
from scipy.integrate import quad
from scipy.stats import norm

import sde_mc_amd as sm
from tests.jobs import reference_kats

_CLASSES = {'EURO_CALL': sm.EuroCall, 'BINARY_AON': sm.BinaryAoN, 'DIGITAL': sm.Digital,
            'BASKET_ARITH': sm.Basket, 'RAINBOW': sm.Rainbow, 'BEST_OF': sm.BestOf}


def test_option_payoffs_match_reference_kats():
    k = reference_kats()
    for p in k['payoffs']:
        x = torch.tensor(k[p['terminals']])
        got = _CLASSES[p['kind']](p['strike'])(x)
        assert torch.allclose(got, torch.tensor(p['expected'])), p['kind']


def test_closed_forms_match_reference_goldens():
    k = reference_kats()['closed_forms']
    assert math.isclose(sm.bs_call(*k['bs_call']['args']), k['bs_call']['expected'], rel_tol=1e-5)
    assert math.isclose(sm.bs_binary_aon(*k['bs_binary_aon']['args']),
                        k['bs_binary_aon']['expected'], rel_tol=1e-9)
    assert math.isclose(sm.merton_call(*k['merton_call']['args']), k['merton_call']['expected'],
                        rel_tol=1e-12)


def test_discounter_matches_reference():
    d = reference_kats()['discount']
    t = torch.tensor(d['times'])
    assert torch.allclose(sm.ConstantShortRate(d['rate'])(t), torch.exp(-d['rate'] * t))


@pytest.mark.parametrize("spot,strike,T,r,s", [(1, 1, 1, 0.02, 0.2), (1, 0.8, 2, 0.05, 0.3)])
def test_digital_and_asian_by_integration(spot, strike, T, r, s):
    # digital: e^{-rT} P(S_T > K); geometric Asian: the continuous geometric
    # average of GBM is log-normal with log-mean log S + (r - s^2/2) T / 2 and
    # log-variance s^2 T / 3 -- both integrated against the log-normal density
    m, v = math.log(spot) + (r - 0.5 * s * s) * T, s * s * T
    p_exceed = 1 - norm.cdf((math.log(strike) - m) / math.sqrt(v))
    assert math.isclose(sm.bs_digital_call(spot, strike, T, r, s), math.exp(-r * T) * p_exceed,
                        rel_tol=1e-10)
    mg, vg = math.log(spot) + 0.5 * (r - 0.5 * s * s) * T, s * s * T / 3
    f = lambda y: max(math.exp(y) - strike, 0.0) * norm.pdf(y, mg, math.sqrt(vg))
    want = math.exp(-r * T) * quad(f, math.log(strike), mg + 12 * math.sqrt(vg))[0]
    assert math.isclose(sm.bs_asian_call(spot, strike, T, r, s), want, rel_tol=1e-7)


def test_transform_log_and_discount():
    x = torch.tensor([[0.0], [math.log(2.0)]])
    c = sm.EuroCall(1.0, log=True, discount=1.5)
    assert torch.allclose(c(x), torch.tensor([0.5, 2.0]))
    assert sm.HestonRainbow(0.5)(torch.tensor([[1.0, 9.0, 2.0, 9.0]])).item() == 1.5
    g = sm.Basket(1.0, average_type='geometric')(torch.tensor([[4.0, 1.0]]))
    assert torch.allclose(g, torch.tensor([1.0]))


def test_uniform_grid_iterates_the_left_points_and_the_end():
    from sde_mc_amd import UniformGrid
    pts = list(UniformGrid(0.0, 3.0, 10))
    # stops once t > end - 1e-8: ten left points, the accumulated end excluded
    assert len(pts) == 10 and pts[0] == 0.0
    assert max(abs(p - 0.3 * k) for k, p in enumerate(pts)) < 1e-14
    with pytest.raises(AssertionError):
        UniformGrid(0.0, 1e-9, 1)


def test_jump_solver_torch_helpers():
    """The by-hand stepping helpers of JumpDiffusionSolver (solvers.py:140-148)."""
    import sde_mc_amd as sm
    p = sm.MertonEuroCall.default_params(10, 'cpu')
    s = p.solver
    torch.manual_seed(0)
    times = s.sample_jump_times((4000, 6, 1))
    assert times.shape == (4000, 6, 1) and bool((times[:, 1:] >= times[:, :-1]).all())
    gaps = torch.diff(times[..., 0], dim=1, prepend=torch.zeros(4000, 1))
    assert abs(float(gaps.mean()) - 1.0 / float(s.sde.jump_rate())) < 0.05
    j = s.sample_one_jump(5)
    assert j.shape == (5, s.sde.dim) and bool((j > -1).all())
    x = torch.ones(5, 1)
    assert torch.allclose(s.add_jumps(0.0, x, x, j), x + j)

"""GPU prices against closed forms (3-sigma plus a stated Euler-bias bound),
and the reference API (solve / mc_simple / mc_multilevel) on the device."""
import math

import numpy as np
import pytest
import torch

import sde_mc_amd as sm
from oracle import closed_form as cf
from sde_mc_amd import engine as E
from tests import jobs as J

pytestmark = pytest.mark.gpu


def _price(job):
    s, q, n = E.stats_to_host(E.simulate(job, E.require_device()))
    return E.mean_and_se(s, q, n)


@pytest.mark.parametrize("grid", ['uniform', 'adapted'])
def test_merton_2p22_x_500_matches_closed_form(grid):
    # the 1-D Merton call at 2^22 paths x 500 steps; Euler bias at 500 steps is
    # ~3e-6 (measured against an exact-scheme coupling), far below 3 se ~ 3e-4
    job = J.job_for(J.merton_sde(), 1.0, 500, 1 << 22, grid=grid, option=sm.EuroCall(1.0),
                    disc=math.exp(-0.02), seed=31)
    m, se = _price(job)
    print('merton %s: %.6f +- %.6f (closed form %.6f)' % (grid, m, se, J.MERTON_PRICE))
    assert abs(m - J.MERTON_PRICE) < 3 * se + 1e-5


@pytest.mark.parametrize("K,T", [(0.8, 1.0), (1.2, 1.0), (1.0, 0.25), (1.0, 2.0)])
def test_merton_strikes_and_maturities(K, T):
    job = J.job_for(J.merton_sde(), T, 200, 1 << 20, option=sm.EuroCall(K), disc=math.exp(-0.02 * T),
                    seed=32)
    m, se = _price(job)
    want = cf.merton_call(1, K, T, 0.02, 0.2, -0.1, 0.15, 2.0)
    assert abs(m - want) < 4 * se + 5e-5


def test_levy_rainbow_jump_free_matches_stulz():
    sde = J.levy_rainbow_sde(f=(0.0, 0.0))
    model = sde.kernel_model()
    model.rate = 0.0
    job = E.make_job(model, 1.0, 200, 0, 1 << 22, 41, 0)
    E.fill_payoff(job, sm.Rainbow(1.0, log=True, discount=math.exp(0.02)), math.exp(-0.02))
    m, se = _price(job)
    want = cf.stulz_call_max(1, 1, 1, 1, 0.02, 0.2, 0.3, 0.5)
    assert abs(m - want) < 3 * se + 1e-5


@pytest.mark.parametrize("grid", ['uniform', 'adapted'])
def test_levy_rainbow_martingale(grid):
    # E[e^{-rT} S_i(T)] = S_i(0) = 1 for each asset under the compensated drift
    sde = J.levy_rainbow_sde()
    for i in range(2):
        job = J.job_for(sde, 1.0, 200, 1 << 20, grid=grid, seed=50 + i)
        E.fill_payoff(job, None, 1.0)
        job.payoff.kind = sm._lib.PAYOFF_IDENTITY
        job.payoff.log = 1
        # IDENTITY reads component 0; swap the model's components for i = 1
        if i == 1:
            mm = job.model
            for f in ('x0', 'a0', 'a1', 'b0', 'b1', 'c0', 'c1', 'j0', 'j1'):
                arr = getattr(mm, f)
                arr[0], arr[1] = arr[1], arr[0]
        m, se = _price(job)
        assert abs(m - 1.0) < 4 * se + 2e-4, (i, m, se)


def test_solver_api_shapes_like_reference_tests():
    # sde_mc tests/test_solvers.py:5-26 and tests/test_mc.py:5-18
    gbm = sm.Gbm(0.02, 0.2, torch.tensor([1., 2.]), dim=2)
    paths, normals = sm.EulerSolver(gbm, 3, 10).solve(bs=4, return_normals=True)
    assert paths.shape == (4, 11, 2) and normals.shape == (4, 10, 2)
    assert not torch.isnan(paths).any()
    merton = sm.Merton(0.02, 0.3, 2, -0.05, 0.3, torch.tensor([1.]), dim=1)
    paths, (nrm, tp, lp, total, jp) = sm.JumpEulerSolver(merton, 3, 10).solve(8)
    assert paths.shape == (8, total + 1, 1)
    for t in (paths, nrm, tp, lp, jp):
        assert not torch.isnan(t).any()
    assert torch.allclose(tp[:, total, 0], torch.tensor(3.0))
    stats = sm.mc_simple(100, sm.EulerSolver(gbm, 3, 10), sm.EuroCall(1), sm.ConstantShortRate(0.02))
    assert stats.sample_mean >= 0 and stats.sample_std > 0
    assert stats.paths.shape == (100, 11, 2)
    stats = sm.mc_simple(100, sm.EulerSolver(sm.Gbm(0.02, 0.2, torch.tensor([1.]), dim=1), 3, 10),
                         sm.EuroCall(1), sm.ConstantShortRate(0.02), bs=17)
    assert stats.sample_mean >= 0 and stats.sample_std > 0


def test_mc_simple_prices_merton_problem():
    prob = sm.MertonEuroCall(0.02, 0.2, 2.0, -0.1, 0.15, 1.0, 1.0, 1.0, 200, 'cpu')
    stats = sm.mc_simple(1 << 20, prob.solver, prob.payoff, prob.discounter, bs=1 << 18)
    assert abs(stats.sample_mean - J.MERTON_PRICE) < 4 * stats.sample_std + 5e-5


def test_mc_multilevel_matches_closed_form():
    prob = sm.MertonEuroCall(0.02, 0.2, 2.0, -0.1, 0.15, 1.0, 1.0, 1.0, 4, 'cpu', grid='uniform')
    levels = [4 * 2 ** l for l in range(7)]
    trials = [int(2 ** (22 - 10 * l / 6)) // 64 * 64 for l in range(7)]
    a = sm.mc_multilevel(trials, levels, prob.solver, prob.payoff, prob.discounter)
    prob.solver.path_counter = 0
    b = sm.mc_multilevel(trials, levels, prob.solver, prob.payoff, prob.discounter, batched=False)
    # batched and per-level launches compute the same jobs
    assert a.sample_mean == b.sample_mean and a.sample_std == b.sample_std
    assert abs(a.sample_mean - J.MERTON_PRICE) < 4 * a.sample_std + 2e-4


def test_run_mc_reaches_the_half_width():
    # find_num_trials / run_mc (mc.py:418-440): a pilot sizes the run so that
    # the 95% half-width is eps; the priced call then sits within it (plus
    # the 200-step Euler bias, ~0.018 T h) of the closed form
    prob = sm.MertonEuroCall(0.02, 0.2, 2.0, -0.1, 0.15, 1.0, 1.0, 1.0, 200, 'cpu')
    n = sm.find_num_trials(prob, 1e-3, init_trials=1 << 17, bs=1 << 17)
    assert 1.4e5 < n < 1.9e5          # (1.96 * 0.207 / 1e-3)^2 ~ 1.65e5
    stats = sm.run_mc(prob, 1e-3, bs=1 << 18, init_trials=1 << 17)
    print('run_mc: %d trials, %.6f +- %.6f (half-width %.2e)'
          % (stats.num_trials, stats.sample_mean, stats.sample_std, 1.96 * stats.sample_std))
    assert 1.96 * stats.sample_std < 1.03e-3
    assert abs(stats.sample_mean - J.MERTON_PRICE) < 4 * stats.sample_std + 0.018 * 0.005


def test_optimal_trials_reach_the_mlmc_target():
    # get_optimal_trials (mlmc.py:77-97) from pilot variances of levels 4..64,
    # then mc_multilevel with those trials: the estimate's standard error
    # meets eps / 1.96 and the price is within it of the closed form (plus
    # the finest level's Euler bias, ~0.018 h = 2.8e-4 at 64 steps)
    prob = sm.MertonEuroCall(0.02, 0.2, 2.0, -0.1, 0.15, 1.0, 1.0, 1.0, 4, 'cpu', grid='uniform')
    levels = [4, 8, 16, 32, 64]
    eps = 2e-3
    trials = sm.get_optimal_trials(1 << 15, levels, eps, prob.solver, prob.payoff, prob.discounter)
    assert trials[0] > trials[1] > trials[-1] > 0
    st = sm.mc_multilevel(trials, levels, prob.solver, prob.payoff, prob.discounter)
    print('mlmc trials %s: %.6f +- %.6f' % (trials, st.sample_mean, st.sample_std))
    assert st.sample_std < 1.05 * eps / 1.96
    assert abs(st.sample_mean - J.MERTON_PRICE) < 4 * st.sample_std + 0.018 / 64


def _preset_job(prob, n, option, disc, offset=0):
    job, in_kernel = prob.solver.make_job(n, option=option, discount=disc, path_offset=offset)
    assert in_kernel or option is None
    return job


def test_levy_call_on_max_preset_on_the_device():
    # LevyCallOnMax.default_params(2, ...) (problem.py:140-166): alpha 0.5,
    # tempered tails mu = 2, eps = 0.001 (~124 jumps a year), T = 3, on the
    # jump-adapted grid.  Paths follow the oracle; each asset's discounted
    # forward is a martingale, which pins the preset's compensator; the call
    # on the max lies between the larger single-asset call and their sum.
    import oracle
    from tests.test_gpu_parity import TOL, _close, _run_rng
    prob = sm.LevyCallOnMax.default_params(2, 64, 'cpu')
    disc = float(prob.discounter(3.0))
    job = _preset_job(prob, 512, prob.payoff, disc)
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, 512)
    assert np.array_equal(ns, ons) and ns.min() > 64 + 200
    assert _close(term, oterm) < TOL
    assert abs(s[0] / 512 - opay.mean()) < 1e-5
    n = 1 << 19
    m_call, se_call = _price(_preset_job(prob, n, prob.payoff, disc, offset=1 << 30))
    singles = []
    for i in range(2):
        job = _preset_job(prob, n, None, 1.0, offset=(2 + i) << 30)
        job.payoff.kind = sm._lib.PAYOFF_IDENTITY
        job.payoff.log = 1
        if i == 1:
            mm = job.model
            for f in ('x0', 'a0', 'a1', 'b0', 'b1', 'c0', 'c1', 'j0', 'j1'):
                arr = getattr(mm, f)
                arr[0], arr[1] = arr[1], arr[0]
        m, se = _price(job)
        print('LevyCallOnMax asset %d: E[exp(X_T)] = %.6f +- %.6f' % (i, m, se))
        assert abs(m - 1.0) < 4 * se + 1e-4
        call = _preset_job(prob, n, sm.EuroCall(1.0, log=True, discount=math.exp(0.06)), disc,
                           offset=(4 + i) << 30)
        if i == 1:
            mm = call.model
            for f in ('x0', 'a0', 'a1', 'b0', 'b1', 'c0', 'c1', 'j0', 'j1'):
                arr = getattr(mm, f)
                arr[0], arr[1] = arr[1], arr[0]
        singles.append(_price(call))
    print('LevyCallOnMax: %.6f +- %.6f, single calls %s' % (m_call, se_call, singles))
    lo = max(c for c, _ in singles)
    hi = sum(c for c, _ in singles)
    tol = 4 * (se_call + max(e for _, e in singles))
    assert lo - tol < m_call < hi + tol

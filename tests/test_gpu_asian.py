"""AsianWrapper's running-integral state and the Asian / Heston-rainbow
payoffs on the device: paths against the fp64 oracle (fixed and Philox
inputs, single level, coupled pair, adapted grid), payoff known answers
through the kernels, in-kernel payoff == torch payoff on the same terminal
states, batched == plain bitwise."""
import math

import numpy as np
import pytest
import torch

import oracle
import sde_mc_amd as sm
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from sde_mc_amd.batch import BatchRunner
from tests import jobs as J
from tests.test_gpu_parity import TOL, _close, _run_fixed, _run_rng
from tests.test_oracle_paths import EXTRA_PAYOFFS, _torch_payoff

pytestmark = pytest.mark.gpu


def _job(P, N, base=None, grid='uniform', coupled=0, seed=51, T=1.0):
    base = base or sm.Gbm(0.05, 0.3, torch.tensor([1.0]), 1)
    return J.job_for(sm.AsianWrapper(base), T, N, P, grid=grid, coarse_factor=coupled,
                     option=sm.AsianCall(T, 1.0), disc=math.exp(-0.05 * T), seed=seed)


@pytest.mark.parametrize("base,coupled", [('gbm', 0), ('merton', 0), ('gbm', 4)])
def test_asian_fixed_increments_match_oracle(base, coupled):
    P, N = 4096 + 5, 64
    b = J.merton_sde(mu=0.05) if base == 'merton' else None
    job = _job(P, N, base=b, coupled=coupled)
    fx = J.fixed_uniform(job, seed=61)
    pay, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("grid", ['uniform', 'adapted'])
def test_asian_philox_paths_match_oracle(grid):
    P, N = 2000, 80
    job = _job(P, N, base=J.merton_sde(mu=0.05), grid=grid, seed=63)
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    if grid == 'adapted':
        assert np.array_equal(ns, ons)
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("kind,make,x", EXTRA_PAYOFFS)
def test_asian_and_heston_rainbow_payoffs_through_the_kernel(kind, make, x):
    opt = make(sm)
    job = J.frozen_state_job(kind, opt.strike, x)
    assert opt.kernel_payoff(job.payoff)
    pay, _, _, _ = _run_rng(job)
    assert pay[0] == pytest.approx(_torch_payoff(opt, x), rel=2e-6, abs=1e-6)


def test_in_kernel_asian_payoff_equals_torch_on_the_same_states():
    p_T, K = 1.0, 1.02
    solver = sm.EulerSolver(sm.AsianWrapper(sm.Gbm(0.05, 0.3, torch.tensor([1.0]), 1)), p_T, 50,
                            device='cuda')
    n = 1 << 16
    off = solver.path_counter
    pay, x = solver.simulate_payoffs(n, sm.AsianCall(p_T, K), 1.0)
    torch_pay = sm.AsianCall(p_T, K)(x)
    assert torch.allclose(pay, torch_pay.to(pay.dtype), rtol=1e-6, atol=1e-6)
    # and the discrete Euler expectation of the average
    h = 1.0 / 50
    want = h * sum((1 + 0.05 * h) ** k for k in range(50))
    a = x[:, 1].double()
    assert abs(float(a.mean()) - want) < 4 * float(a.std()) / math.sqrt(n)
    solver.path_counter = off
    stats = sm.mc_simple(n, solver, sm.AsianCall(p_T, K), sm.ConstantShortRate(0.05), bs=n)
    assert stats.sample_mean == pytest.approx(float(pay.double().mean()) * math.exp(-0.05),
                                              rel=1e-6)


def test_asian_batched_equals_plain():
    jobs = [_job(3000, 64, seed=71), _job(5000, 128, base=J.merton_sde(), coupled=2, seed=72),
            _job(2000, 40, base=J.merton_sde(), grid='adapted', seed=73)]
    dev = E.require_device()
    plain = np.array([E.simulate(j, dev).cpu().numpy() for j in jobs])
    r = BatchRunner(max_jobs=8)
    assert np.array_equal(r.run(jobs), plain)

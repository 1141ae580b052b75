"""Heston family on the device: paths against the fp64 oracle (fixed and
Philox inputs, single level and coupled pair), batched == plain bitwise in the
generic batch class, and the price against the Heston closed form after
removing the scheme's first-order bias."""
import math

import numpy as np
import pytest
import torch

import oracle
import sde_mc_amd as sm
from oracle import closed_form as cf
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from sde_mc_amd.batch import BatchRunner
from tests import jobs as J
from tests.test_gpu_parity import TOL, _close, _run_fixed, _run_rng

pytestmark = pytest.mark.gpu

PARAMS = dict(r=0.02, kappa=1.5, theta=0.04, xi=0.3, rho=-0.7)


def _job(N, P, coupled=0, seed=21, T=1.0, option=None):
    sde = sm.Heston(PARAMS['r'], PARAMS['kappa'], PARAMS['theta'], PARAMS['xi'], PARAMS['rho'],
                    torch.tensor([1.0, 0.04]))
    return J.job_for(sde, T, N, P, coarse_factor=coupled, option=option or sm.EuroCall(1.0),
                     disc=math.exp(-0.02 * T), seed=seed)


@pytest.mark.parametrize("coupled", [0, 2])
def test_heston_fixed_increments_match_oracle(coupled):
    P, N = 4096 + 3, 64
    job = _job(N, P, coupled)
    fx = J.fixed_uniform(job, seed=31)
    pay, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("coupled", [0, 4])
def test_heston_philox_paths_match_oracle(coupled):
    P, N = 1000, 96
    job = _job(N, P, coupled, seed=33)
    pay, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


def test_heston_batched_equals_plain_in_the_generic_class():
    jobs = [_job(64, 3000, seed=41), _job(128, 5000, coupled=2, seed=42),
            J.job_for(J.merton_sde(), 1.0, 40, 4000, option=sm.EuroCall(1.0), seed=43)]
    dev = E.require_device()
    plain = np.array([E.simulate(j, dev).cpu().numpy() for j in jobs])
    r = BatchRunner(max_jobs=8)
    assert np.array_equal(r.run(jobs), plain)
    assert list(r.slot_class[:2]) == [L.CLASS_GENERIC, L.CLASS_GENERIC]


def test_heston_price_extrapolates_to_the_closed_form():
    want = cf.heston_call(1.0, 1.0, 1.0, PARAMS['r'], 0.04, PARAMS['kappa'], PARAMS['theta'],
                          PARAMS['xi'], PARAMS['rho'])
    P = 1 << 22
    dev = E.require_device()
    out = {}
    for N in (128, 512):
        job = _job(N, P, seed=55)
        pay = torch.empty(P, dtype=torch.float32, device=dev)
        po = L.PathOut()
        po.payoff_dev = pay.data_ptr()
        E.stats_to_host(E.simulate(job, dev, out=po))
        out[N] = pay.double().cpu().numpy()
    d = (4 * out[512] - out[128]) / 3
    extrap, se = d.mean(), d.std() / math.sqrt(P)
    assert abs(extrap - want) < 4 * se + 5e-5, (extrap, want, se)


def test_heston_api_solve_and_mc():
    problem = sm.HestonEuroCall.default_params(50, 'cpu')
    paths, normals = problem.solver.solve(bs=256)
    assert paths.shape == (256, 51, 2) and normals.shape == (256, 50, 2)
    assert torch.all(paths[:, :, 1] >= 0)
    stats = sm.mc_simple(1 << 18, problem.solver, problem.payoff, problem.discounter, bs=1 << 18)
    want = cf.heston_call(1.0, 1.0, 3.0, 0.02, 0.15, 0.25, 0.5, 0.3, -0.3)
    assert abs(stats.sample_mean - want) < 4 * stats.sample_std + 5e-3

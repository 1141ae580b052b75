"""The 'general' diffusion family (a (dim, brown_dim) matrix B(x) = B0 +
diag(x) B1 on the correlated Brownian increments, schemes.py:12) on the
device: paths against the oracle (Philox and fixed increments, single and
coupled levels, with jumps), stored increments replayed bit for bit, the
batched launch equal to the per-level one, and prices against closed forms."""
import math

import numpy as np
import pytest
import torch

import oracle
import sde_mc_amd as sm
from oracle import closed_form as cf
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from tests import jobs as J
from tests.test_gpu_parity import TOL, _close, _run_fixed, _run_rng
from tests.test_oracle_general import _corr, _general

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M", [0, 2])
@pytest.mark.parametrize("dim,nb", [(1, 1), (2, 3), (3, 2), (4, 4), (2, 4)])
def test_general_rng_matches_oracle(dim, nb, M):
    sde = _general(dim, nb, seed=dim * 10 + nb)
    P = 1000
    job = J.job_for(sde, 1.0, 40, P, coarse_factor=M, seed=71 + M, option=sm.EuroCall(1.0))
    pay, term, ns, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("M", [0, 4])
def test_general_fixed_matches_oracle(M):
    sde = _general(3, 4, seed=9)
    P = 777
    job = J.job_for(sde, 0.5, 48, P, coarse_factor=M, option=sm.EuroCall(1.0))
    fx = J.fixed_uniform(job, seed=3)
    _, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


def _jump_model():
    # 2 assets on 3 factors with proportional lognormal jumps
    B1 = np.array([[0.2, 0.05, -0.1], [0.0, 0.25, 0.1]])
    return sm.KernelModel(2, L.NOISE_GENERAL, [1.0, 1.1], a1=[0.01, -0.02], j1=[1.0, 0.5],
                          rate=3.0, jump_params=[-0.1, 0.15], corr=_corr(3), brown_dim=3,
                          bm1=B1)


@pytest.mark.parametrize("M", [0, 2])
def test_general_with_jumps_matches_oracle(M):
    P = 1500
    job = E.make_job(_jump_model(), 1.0, 60, L.GRID_UNIFORM, P, 17, 0, coarse_factor=M)
    E.fill_payoff(job, sm.Basket(1.0), 1.0)
    _, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5
    fx = J.fixed_uniform(job, seed=4)
    assert fx.a['js'].size > P  # jumps present
    _, tf, _, _ = _run_fixed(job, fx)
    _, otf, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(tf, otf) < TOL


def test_general_stored_increments_replay_the_paths():
    # solve(return_normals=True) returns the correlated increments (bs, N,
    # brown_dim); fed back as fixed rows they give the same paths bit for bit
    sde = _general(3, 2, seed=12)
    N, bs = 30, 300
    solver = sm.EulerSolver(sde, 1.0, N)
    paths, normals = solver.solve(bs, return_normals=True)
    assert paths.shape == (bs, N + 1, 3) and normals.shape == (bs, N, 2)
    dw = np.zeros((N, L.MAX_DIM, bs), dtype=np.float32)
    dw[:, :2, :] = normals.permute(1, 2, 0).numpy()
    job = J.job_for(sde, 1.0, N, bs, seed=solver.seed)
    fx = J.FixedInputs(dict(dw=dw, counts=np.zeros((N, bs), np.uint8),
                            off=np.zeros(bs + 1, np.int64), js=np.zeros(0, np.float32)), N)
    _, term, _, _ = _run_fixed(job, fx)
    assert np.array_equal(term[0].T, paths[:, N].numpy())
    # and the stored paths follow the oracle
    _, oterm, _ = oracle.simulate(job, 0, bs)
    assert _close(paths[:, N].numpy().T[None], oterm) < TOL


def test_general_batched_levels_equal_per_level():
    sde = sm.MultiFactorGbm(torch.tensor([0.02, 0.02]), torch.tensor([[0.2, 0.1, 0.0],
                                                                      [0.0, 0.15, 0.2]]),
                            torch.tensor([1.0, 1.0]), 2, 3, corr_matrix=torch.tensor(_corr(3)))
    solver = sm.EulerSolver(sde, 1.0, 4)
    levels = [4, 8, 16, 32]
    trials = [1 << 18, 1 << 16, 1 << 14, 1 << 12]
    disc = sm.ConstantShortRate(0.02)
    a = sm.mc_multilevel(trials, levels, solver, sm.EuroCall(1.0), disc)
    solver.path_counter = 0
    b = sm.mc_multilevel(trials, levels, solver, sm.EuroCall(1.0), disc, batched=False)
    assert a.sample_mean == b.sample_mean and a.sample_std == b.sample_std


def test_multi_factor_gbm_prices_at_2p22():
    # asset 0 of a 3-asset, 3-factor GBM is a GBM of volatility
    # sqrt((s C s^T)_00); its discounted call at r = mu_0 against
    # Black-Scholes (Euler bias at 500 steps < 1e-4), and every asset's mean
    # against the exact Euler mean S_i(0) (1 + mu_i h)^N
    mu = np.array([0.03, 0.01, 0.05])
    sig = np.array([[0.2, 0.1, -0.05], [0.05, 0.3, 0.0], [0.0, -0.1, 0.25]])
    C = _corr(3)
    T, N, P = 1.0, 500, 1 << 22
    sde = sm.MultiFactorGbm(torch.tensor(mu), torch.tensor(sig), torch.tensor([1.0, 1.0, 1.0]),
                            3, 3, corr_matrix=torch.tensor(C))
    vol = math.sqrt((sig @ C @ sig.T)[0, 0])
    job = J.job_for(sde, T, N, P, seed=2024, option=sm.EuroCall(1.0), disc=math.exp(-mu[0] * T))
    s, q, n = E.stats_to_host(E.simulate(job, E.require_device()))
    m, se = E.mean_and_se(s, q, n)
    want = cf.bs_call(1.0, 1.0, T, mu[0], vol)
    print('multi-factor GBM call %.6f +- %.6f, Black-Scholes %.6f' % (m, se, want))
    assert abs(m - want) < 4 * se + 1e-4
    for i in range(3):
        job = J.job_for(sde, T, N, P, seed=3000 + i)
        job.payoff.kind = L.PAYOFF_IDENTITY
        if i:
            mm = job.model
            for a in (mm.x0, mm.a0, mm.a1):
                a[0], a[i] = a[i], a[0]
            for j in range(3):
                mm.bm0[j], mm.bm0[i * 4 + j] = mm.bm0[i * 4 + j], mm.bm0[j]
                mm.bm1[j], mm.bm1[i * 4 + j] = mm.bm1[i * 4 + j], mm.bm1[j]
        s, q, n = E.stats_to_host(E.simulate(job, E.require_device()))
        m, se = E.mean_and_se(s, q, n)
        want = (1 + mu[i] * T / N) ** N
        assert abs(m - want) < 4 * se + 1e-6, (i, m, se, want)

"""Pins of the oracle's 'general' diffusion family (a (dim, brown_dim) matrix
B(x) = B0 + diag(x) B1 applied to the correlated Brownian increments, the
matrix branch of schemes.py:12): the 'diag' family it reduces to, the torch
matrix-product restatement of the reference loop, the Gaussian law of the
constant-coefficient case and the exact Euler moments of a multi-factor GBM.
Non-symmetric matrices throughout, so a transposed operand fails."""
import math

import numpy as np
import pytest
import torch

import oracle
import sde_mc_amd as sm
from oracle import refcpu
from sde_mc_amd import _lib as L
from tests import jobs as J


def _corr(nb, rho=0.4):
    c = np.full((nb, nb), rho)
    np.fill_diagonal(c, 1.0)
    c[0, -1] = c[-1, 0] = -0.3 if nb > 2 else c[0, -1]
    return c


def _general(dim, nb, seed, prop=True, const=True, a1=0.03):
    rng = np.random.default_rng(seed)
    B0 = rng.uniform(-0.3, 0.3, (dim, nb)) if const else np.zeros((dim, nb))
    B1 = rng.uniform(-0.3, 0.3, (dim, nb)) if prop else np.zeros((dim, nb))
    x0 = rng.uniform(0.8, 1.2, dim)
    a0 = rng.uniform(-0.05, 0.05, dim)
    return sm.GeneralSde(torch.tensor(a0), a1, torch.tensor(B0), torch.tensor(B1),
                         torch.tensor(x0), dim, nb, corr_matrix=torch.tensor(_corr(nb)))


def _z(job, P):
    """The oracle's normals of paths 0..P-1: z [N][4][P]."""
    N = int(job.steps)
    z = np.empty((N, L.MAX_DIM, P))
    for p in range(P):
        for n in range(N):
            for k in range(L.MAX_DIM):
                z[n, k, p] = oracle.normal(job.seed, job.path_offset + p, n * L.MAX_DIM + k)
    return z


@pytest.mark.parametrize("dim,nb", [(1, 1), (2, 3), (3, 2), (4, 4), (2, 4)])
def test_general_matches_torch_matrix_product(dim, nb):
    sde = _general(dim, nb, seed=dim * 10 + nb)
    job = J.job_for(sde, 1.0, 12, 48, option=sm.EuroCall(1.0))
    km = sde.kernel_model()
    _, term, ns = oracle.simulate(job, 0, 48)
    want = refcpu.general_uniform(km.x0, km.a0, km.a1, km.bm0, km.bm1, km.cholesky(), 1.0,
                                  _z(job, 48))
    assert np.all(ns == 12)
    assert np.abs(term[0].T - want.numpy()).max() < 1e-12


@pytest.mark.parametrize("M", [2, 4])
def test_general_coupled_matches_torch_matrix_product(M):
    sde = _general(3, 2, seed=M)
    job = J.job_for(sde, 1.0, 16, 32, coarse_factor=M, option=sm.EuroCall(1.0))
    km = sde.kernel_model()
    _, term, _ = oracle.simulate(job, 0, 32)
    xf, xc = refcpu.general_uniform(km.x0, km.a0, km.a1, km.bm0, km.bm1, km.cholesky(), 1.0,
                                    _z(job, 32), coarse_factor=M)
    assert np.abs(term[0].T - xf.numpy()).max() < 1e-12
    assert np.abs(term[1].T - xc.numpy()).max() < 1e-12


def test_general_with_diagonal_matrices_is_the_diag_family():
    # 4 components on 4 factors: 'general' draws 4 normals per step like
    # 'diag' at dim 4, and with B0 = diag(b0), B1 = diag(b1) adds only exact
    # zeros to the 'diag' step, so the paths agree bit for bit (jumps too)
    corr = torch.tensor(_corr(4))
    b0 = np.array([0.05, 0.0, 0.02, 0.0])
    b1 = np.array([0.2, 0.3, 0.1, 0.25])
    diag = sm.KernelModel(4, L.NOISE_DIAG, [1.0, 0.9, 1.1, 1.2], a1=[0.01] * 4, b0=list(b0),
                          b1=list(b1), j1=[1.0] * 4, rate=1.5, jump_params=[-0.1, 0.15],
                          corr=corr.numpy())
    gen = sm.KernelModel(4, L.NOISE_GENERAL, [1.0, 0.9, 1.1, 1.2], a1=[0.01] * 4,
                         j1=[1.0] * 4, rate=1.5, jump_params=[-0.1, 0.15], corr=corr.numpy(),
                         brown_dim=4, bm0=np.diag(b0), bm1=np.diag(b1))
    for M in (0, 2):
        jobs = []
        for km in (diag, gen):
            from sde_mc_amd import engine as E
            job = E.make_job(km, 1.0, 20, L.GRID_UNIFORM, 256, 77, 0, coarse_factor=M)
            E.fill_payoff(job, sm.Basket(1.0), 1.0)
            jobs.append(job)
        a = oracle.simulate(jobs[0], 0, 256)
        b = oracle.simulate(jobs[1], 0, 256)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_general_constant_coefficients_are_gaussian():
    # B1 = 0, a1 = 0: the Euler scheme is exact, X_T ~ N(x0 + a0 T,
    # B0 C B0^T T) with C the factor correlation
    dim, nb, T, P = 3, 2, 2.0, 1 << 15
    rng = np.random.default_rng(3)
    B0 = np.array([[0.3, -0.1], [0.05, 0.4], [-0.2, 0.1]])
    a0 = np.array([0.1, -0.05, 0.02])
    C = np.array([[1.0, 0.6], [0.6, 1.0]])
    sde = sm.GeneralSde(torch.tensor(a0), 0.0, torch.tensor(B0), 0.0, torch.tensor([1.0, 2.0, 0.5]),
                        dim, nb, corr_matrix=torch.tensor(C))
    job = J.job_for(sde, T, 5, P, seed=int(rng.integers(1 << 30)), option=sm.EuroCall(1.0))
    _, term, _ = oracle.simulate(job, 0, P)
    x = term[0].T
    cov = B0 @ C @ B0.T * T
    mean = np.array([1.0, 2.0, 0.5]) + a0 * T
    se = np.sqrt(np.diag(cov) / P)
    assert np.all(np.abs(x.mean(axis=0) - mean) < 4 * se)
    emp = np.cov(x.T)
    # sd of a sample covariance entry: sqrt((s_ii s_jj + s_ij^2) / P)
    sd = np.sqrt((np.outer(np.diag(cov), np.diag(cov)) + cov ** 2) / P)
    assert np.all(np.abs(emp - cov) < 5 * sd)
    # the transposed product B0^T C B0 is not even the right shape; the
    # un-correlated B0 B0^T differs by far more than the tolerance
    assert np.abs(B0 @ B0.T * T - cov).max() > 20 * sd.max()


def test_multi_factor_gbm_exact_euler_moments():
    # dS_i = mu_i S_i dt + S_i sum_j sigma_ij dW_j: each Euler factor
    # 1 + mu_i h + (sigma dW)_i is independent across steps, so
    #   E[S_i(T)] = S_i(0) (1 + mu_i h)^N,
    #   E[S_i S_k] = S_i S_k (1 + (mu_i + mu_k) h + mu_i mu_k h^2 + h (s C s^T)_ik)^N
    mu = np.array([0.05, -0.02, 0.1])
    sig = np.array([[0.2, 0.0, 0.1], [0.05, 0.3, -0.1], [0.0, -0.15, 0.25]])
    C = _corr(3)
    s0 = np.array([1.0, 0.8, 1.3])
    T, N, P = 1.0, 10, 1 << 16
    sde = sm.MultiFactorGbm(torch.tensor(mu), torch.tensor(sig), torch.tensor(s0), 3, 3,
                            corr_matrix=torch.tensor(C))
    job = J.job_for(sde, T, N, P, seed=99, option=sm.EuroCall(1.0))
    _, term, _ = oracle.simulate(job, 0, P)
    S = term[0].T
    h = T / N
    m1 = s0 * (1 + mu * h) ** N
    Q = sig @ C @ sig.T
    m2 = np.outer(s0, s0) * (1 + (mu[:, None] + mu[None, :]) * h + np.outer(mu, mu) * h * h
                             + h * Q) ** N
    se1 = np.sqrt(np.diag(m2) - m1 ** 2) / math.sqrt(P)
    assert np.all(np.abs(S.mean(axis=0) - m1) < 4 * se1)
    emp2 = (S[:, :, None] * S[:, None, :]).mean(axis=0)
    se2 = np.sqrt(((S[:, :, None] * S[:, None, :]) ** 2).mean(axis=0) - emp2 ** 2) / math.sqrt(P)
    assert np.all(np.abs(emp2 - m2) < 4 * se2)
    # a transposed loading matrix changes the cross moments beyond the tolerance
    Qt = sig.T @ C @ sig
    m2t = np.outer(s0, s0) * (1 + (mu[:, None] + mu[None, :]) * h + np.outer(mu, mu) * h * h
                              + h * Qt) ** N
    assert np.any(np.abs(m2t - m2) > 8 * se2)


def test_general_fixed_rows_are_the_correlated_increments():
    # fixed-increment rows hold dW = L z sqrt(h) (4 rows per step): fed the
    # increments of the generator's own normals, the fixed path is the RNG path
    sde = _general(2, 3, seed=5)
    P, N = 40, 9
    job = J.job_for(sde, 1.0, N, P, option=sm.EuroCall(1.0))
    km = sde.kernel_model()
    z = _z(job, P)
    Lc = np.zeros((4, 4))
    Lc[:3, :3] = km.cholesky()
    dw = np.einsum('ij,njp->nip', Lc, z * math.sqrt(1.0 / N)).astype(np.float32)
    fx = J.FixedInputs(dict(dw=dw, counts=np.zeros((N, P), np.uint8),
                            off=np.zeros(P + 1, np.int64), js=np.zeros(0, np.float32)), N)
    _, tf, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    _, tr, _ = oracle.simulate(job, 0, P)
    assert np.abs(tf - tr).max() < 1e-6


def test_general_model_description():
    sde = _general(3, 2, seed=1)
    km = sde.kernel_model()
    m = km.fill(L.Job().model)
    assert m.noise == L.NOISE_GENERAL and m.brown_dim == 2
    assert m.bm0[1 * 4 + 1] == pytest.approx(float(sde.B0[1, 1]))
    assert m.bm1[2 * 4 + 0] == pytest.approx(float(sde.B1[2, 0]))
    # the reference's torch step (the intended matmul of schemes.py:12)
    x = torch.rand(5, 3)
    b = sde.diffusion(0.0, x)
    assert b.shape == (5, 3, 2)
    assert torch.allclose(b[:, 1, :], sde.B0[1] + x[:, 1:2] * sde.B1[1])
    with pytest.raises(ValueError):
        sm.GeneralSde(0.0, 0.0, torch.zeros(3, 3), 0.0, torch.ones(3), 3, 2)

"""Pins of the C oracle's path restatement, each against something the
oracle's own code does not contain: known-answer vectors, closed forms,
product formulas, hand-built event grids, the torch-fp32 restatement of the
reference loop, statistical laws."""
import math

import numpy as np
import pytest
import torch
from scipy import stats
from scipy.integrate import quad

import oracle
from oracle import closed_form as cf
from oracle import refcpu
from sde_mc_amd import _lib as L
from sde_mc_amd import EuroCall, Rainbow
from tests import jobs as J


# ---- generator -----------------------------------------------------------
KAT = [  # Random123 known-answer vectors for philox4x32-10
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert oracle.philox(ctr, key) == want


def test_normals_are_standard_normal():
    z = np.array([oracle.normal(7, p, i) for p in range(400) for i in range(100)])
    assert abs(z.mean()) < 4 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 4 * math.sqrt(2 / z.size)
    assert stats.kstest(z, 'norm').pvalue > 1e-3
    # distinct streams per path and per index
    assert oracle.normal(7, 0, 0) != oracle.normal(7, 1, 0) != oracle.normal(8, 0, 0)


# ---- deterministic path pins ----------------------------------------------
def _gbm_job(P, N, grid='uniform', mu=0.05, sigma=0.3, T=1.0):
    import sde_mc_amd as sm
    sde = sm.Gbm(mu, sigma, torch.tensor([1.0]), 1)
    return J.job_for(sde, T, N, P, grid=grid)


def test_zero_vol_gives_compound_drift():
    job = _gbm_job(8, 50, sigma=0.0)
    fx = J.fixed_uniform(job, seed=1)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), 8)
    assert np.allclose(term[0, 0], (1 + 0.05 / 50) ** 50, rtol=1e-14)


def test_gbm_fixed_matches_product_formula():
    P, N = 16, 40
    job = _gbm_job(P, N)
    fx = J.fixed_uniform(job, seed=2)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    dw = fx.a['dw'][:, 0, :].astype(np.float64)
    want = np.prod(1 + 0.05 / N + 0.3 * dw, axis=0)
    assert np.allclose(term[0, 0], want, rtol=1e-13)


def test_single_forced_jump_uniform():
    # one jump of mark 0.5 in step 3: x_4 = x_3 (1 + mu h + sigma dW_3 + J)
    P, N = 4, 10
    job = J.job_for(J.merton_sde(), 1.0, N, P)
    fx = J.fixed_uniform(job, seed=3)
    fx.a['counts'][:] = 0
    fx.a['counts'][3, :] = 1
    fx.a['off'] = np.arange(P + 1, dtype=np.int64)
    fx.a['js'] = np.full(P, 0.5, dtype=np.float32)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    a1 = job.model.a1[0]
    dw = fx.a['dw'][:, 0, :].astype(np.float64)
    f = 1 + a1 / N + 0.2 * dw
    f[3] += 0.5
    assert np.allclose(term[0, 0], np.prod(f, axis=0), rtol=1e-13)


def test_exact_jump_rule_uses_post_step_state():
    # exact_jumps: the jump coefficient is taken after the diffusion step
    # (solvers.py:214-215): x_4 = x_3 (1 + mu h + sigma dW_3)(1 + J)
    P, N = 4, 10
    job = J.job_for(J.merton_sde(), 1.0, N, P, exact_jumps=True)
    fx = J.fixed_uniform(job, seed=3)
    fx.a['counts'][:] = 0
    fx.a['counts'][3, :] = 1
    fx.a['off'] = np.arange(P + 1, dtype=np.int64)
    fx.a['js'] = np.full(P, 0.5, dtype=np.float32)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    a1 = job.model.a1[0]
    dw = fx.a['dw'][:, 0, :].astype(np.float64)
    f = 1 + a1 / N + 0.2 * dw
    f[3] *= 1.5
    assert np.allclose(term[0, 0], np.prod(f, axis=0), rtol=1e-13)


def test_double_noise_product_formula():
    # 'indep' diffusion with two independent drivers (DoubleGbm, sde.py:222-255):
    # x_N = x_0 prod(1 + mu h + s1 dW_n + s2 dB_n), dW and dB the two dw rows
    import sde_mc_amd as sm
    P, N = 12, 30
    job = J.job_for(sm.DoubleGbm(0.03, 0.2, 0.4, torch.tensor([1.0]), 1), 1.0, N, P)
    fx = J.fixed_uniform(job, seed=8)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    dw = fx.a['dw'].astype(np.float64)
    want = np.prod(1 + 0.03 / N + 0.2 * dw[:, 0, :] + 0.4 * dw[:, 1, :], axis=0)
    assert np.allclose(term[0, 0], want, rtol=1e-13)


def test_put_call_parity_and_basket_invariants():
    # call(x) - put(x) = x - K with the call pinned by the reference's known
    # answers; geometric basket <= arithmetic basket, equal on equal components
    from tests.jobs import frozen_state_job
    for x in (0.3, 0.9, 1.0, 1.7):
        c = oracle.simulate(frozen_state_job('EURO_CALL', 1.0, [x]), 0, 1)[0][0]
        p = oracle.simulate(frozen_state_job('EURO_PUT', 1.0, [x]), 0, 1)[0][0]
        assert abs((c - p) - (x - 1.0)) < 1e-12
    for x in ([1.5, 1.5], [2.0, 0.7], [0.9, 1.4, 1.1]):
        g = oracle.simulate(frozen_state_job('BASKET_GEO', 0.5, x), 0, 1)[0][0]
        a = oracle.simulate(frozen_state_job('BASKET_ARITH', 0.5, x), 0, 1)[0][0]
        assert g <= a + 1e-12
        if len(set(x)) == 1:
            assert abs(g - a) < 1e-12
        else:
            assert abs(g - (np.exp(np.mean(np.log(x))) - 0.5)) < 1e-12


def test_adapted_grid_restarts_at_jump():
    # h = 0.25, one jump at t = 0.3: steps end at .25, .3 (jump), .55, .8, 1.0
    P = 3
    job = J.job_for(J.merton_sde(sigma=0.0), 1.0, 4, P, grid='adapted')
    fx = J.fixed_adapted(job, seed=4)
    fx.a['jt'][:] = 5.0
    fx.a['jt'][0, :] = np.float32(0.3)
    fx.a['js'][0, :] = 0.25
    _, term, ns = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert (ns == 5).all()
    a1 = job.model.a1[0]
    f32 = np.float32
    t1 = f32(0.25)
    t2 = f32(0.3)                      # 0.25 + (0.3 - 0.25) is exact
    t3 = f32(t2 + f32(0.25))
    t4 = f32(t3 + f32(0.25))
    dts = [t1, f32(t2 - t1), f32(0.25), f32(0.25), f32(f32(1.0) - t4)]
    x = 1.0
    for k, dt in enumerate(dts):
        xo = x
        x = x + a1 * x * float(dt)
        if k == 1:                     # the jump lands at the end of step 2
            x += xo * 0.25
    assert np.allclose(term[0, 0], x, rtol=1e-7)


def test_adapted_matches_torch_fp32_restatement():
    # the C oracle's event grid equals the op-for-op torch-fp32 loop of the
    # reference on the same fixed draws: same step counts, states within fp32
    P, N = 64, 20
    job = J.job_for(J.merton_sde(rate=6.0), 1.0, N, P, grid='adapted')
    fx = J.fixed_adapted(job, seed=5)
    _, term, ns = oracle.simulate_fixed(job, fx.host_struct(), P)
    m = refcpu.merton_model(0.02, 0.2, 6.0, -0.1, 0.15)
    paths, total = refcpu.solve_adapted(m, 1.0, N, P, int(job.max_steps) - N,
                                        fixed=(fx.a['dw'], fx.a['jt'], fx.a['js']))
    assert total == ns.max()
    assert np.allclose(paths[:, total, 0].numpy(), term[0, 0], rtol=2e-5)


def _coupled_adapted_case(sde, N, M, P, seed, exact_jumps=False):
    job = J.job_for(sde, 1.0, N, P, grid='adapted', coarse_factor=M, exact_jumps=exact_jumps)
    fx = J.fixed_adapted(job, seed=seed)
    _, term, ns = oracle.simulate_fixed(job, fx.host_struct(), P)
    m = refcpu.model_from_kernel(job.model)
    fixed = (fx.a['dw'], fx.a['jt'], fx.a['js'])
    return job, term, ns, m, fixed


COUPLED_ADAPTED = [
    ('merton_old_state', lambda: J.merton_sde(rate=6.0), 24, 2, False),
    ('merton_exact', lambda: J.merton_sde(rate=6.0), 24, 2, True),
    ('merton_m4', lambda: J.merton_sde(rate=9.0), 32, 4, False),
    ('levy_indep', lambda: J.levy_rainbow_sde(eps=0.2), 16, 2, False),
    ('levy_indep_exact', lambda: J.levy_rainbow_sde(f=(0.2, 0.2)), 16, 2, True),
]


@pytest.mark.parametrize("name,mk,N,M,exact", COUPLED_ADAPTED, ids=[c[0] for c in COUPLED_ADAPTED])
def test_adapted_coupled_matches_torch_fp32_restatement(name, mk, N, M, exact):
    # the C oracle's coupled jump-adapted pair equals the op-for-op torch-fp32
    # restatement of multilevel_solve (solvers.py:228-307) on the same fixed
    # draws: identical iteration counts per path, fine and coarse states
    # within fp32
    P = 64
    job, term, ns, m, fixed = _coupled_adapted_case(mk(), N, M, P, seed=21, exact_jumps=exact)
    xf, xc, iters = refcpu.multilevel_adapted(m, 1.0, N, N // M, P, fixed, exact_jumps=exact)
    assert np.array_equal(iters.numpy(), ns)
    # jumps happened inside the window, several paths took event steps
    assert (ns > N // M).sum() >= P // 4
    for i in range(job.model.dim):
        assert np.allclose(xf[:, i].numpy(), term[0, i], rtol=2e-5, atol=2e-6), (name, 'fine', i)
        assert np.allclose(xc[:, i].numpy(), term[1, i], rtol=2e-5, atol=2e-6), (name, 'coarse', i)


@pytest.mark.parametrize("plant", ['coarse_post_state', 'coarse_untruncated', 'sum_unscaled'])
def test_adapted_coupled_restatement_detects_planted_mistakes(plant):
    # the comparison above is sharp enough to see each plausible mistake of a
    # coupled jump-adapted step: a loop with the mistake planted disagrees
    # with the oracle on the coarse path
    P, N, M = 64, 24, 2
    job, term, ns, m, fixed = _coupled_adapted_case(J.merton_sde(rate=6.0), N, M, P, seed=21)
    xf, xc, _ = refcpu.multilevel_adapted(m, 1.0, N, N // M, P, fixed, plant=plant)
    assert np.allclose(xf[:, 0].numpy(), term[0, 0], rtol=2e-5)
    assert not np.allclose(xc[:, 0].numpy(), term[1, 0], rtol=2e-5, atol=2e-6)


def test_oracle_mlmc_telescopes_adapted():
    # on the jump-adapted grid too, E[P_fine - P_coarse] from the coupled
    # pair equals the difference of the two single-level adapted prices, and
    # the correction's variance is far below a level's
    n = 1 << 14
    sde = J.merton_sde(rate=4.0)
    opt, disc = EuroCall(1.0), math.exp(-0.02)
    c = J.job_for(sde, 1.0, 8, n, grid='adapted', coarse_factor=2, option=opt, disc=disc)
    d, _, _ = oracle.simulate(c, 0, n)
    f = J.job_for(sde, 1.0, 8, n, grid='adapted', option=opt, disc=disc, seed=11)
    g = J.job_for(sde, 1.0, 4, n, grid='adapted', option=opt, disc=disc, seed=12)
    pf, _, _ = oracle.simulate(f, 0, n)
    pg, _, _ = oracle.simulate(g, 0, n)
    diff = pf.mean() - pg.mean()
    se = math.sqrt(pf.var() / n + pg.var() / n + d.var() / n)
    assert abs(d.mean() - diff) < 4 * se
    assert d.var() < 0.1 * pf.var()


def test_coupled_coarse_equals_single_on_summed_increments():
    # the coarse path of a coupled pair is a single-level run on the coarse
    # grid whose increments are pairwise sums and whose jumps are the fine ones
    P, N, M = 32, 24, 2
    sde = J.merton_sde(rate=5.0)
    cj = J.job_for(sde, 1.0, N, P, coarse_factor=M)
    fx = J.fixed_uniform(cj, seed=6)
    _, tc, _ = oracle.simulate_fixed(cj, fx.host_struct(), P)
    sj = J.job_for(sde, 1.0, N // M, P)
    dw = fx.a['dw'].astype(np.float64)
    sfx = J.FixedInputs(dict(dw=(dw[0::2] + dw[1::2]).astype(np.float32),
                             counts=(fx.a['counts'][0::2] + fx.a['counts'][1::2]).astype(np.uint8),
                             off=fx.a['off'], js=fx.a['js']), N // M)
    _, ts, _ = oracle.simulate_fixed(sj, sfx.host_struct(), P)
    # the summed fp32 increments differ from the fp64 sums by one rounding
    assert np.allclose(tc[1, 0], ts[0, 0], rtol=1e-6)
    _, tf, _ = oracle.simulate_fixed(J.job_for(sde, 1.0, N, P), fx.host_struct(), P)
    assert np.array_equal(tc[0, 0], tf[0, 0])


# ---- Levy inverse CDF ----------------------------------------------------
def _levy_cdf(z, cm, cp, mu, alpha, eps, lda):
    im = 0.0 if math.isinf(mu) else 1.0 / mu

    def dens(x):
        ax = abs(x)
        c = cm if x < 0 else cp
        if ax < eps:
            return 0.0
        if ax <= 1:
            return c * ax ** (-1 - alpha) / lda
        return c * math.exp(-mu * (ax - 1)) / lda if im > 0 else 0.0
    lo = -1.0 if im == 0 else -60.0
    pts = [p for p in (-1.0, -eps, eps, 1.0) if lo < p < z]
    return quad(dens, lo, z, points=pts or None, limit=200, epsabs=1e-13)[0]


@pytest.mark.parametrize("mu", [2.0, math.inf])
def test_levy_icdf_inverts_density(mu):
    import sde_mc_amd as sm
    icdf = sm.InverseCdf(1.0, 1.0, mu, 0.8, 0.05)
    sde = sm.LevySde(sm.ExpExampleLevy(1.0, 1.0, 0.8, mu, 0.02, 0.2, 1.0, 0.05, 1),
                     torch.tensor([1.0]))
    job = J.job_for(sde, 1.0, 10, 1)
    zs = [-0.9, -0.3, -0.06, 0.06, 0.3, 0.9] + ([-1.5, 1.7] if not math.isinf(mu) else [])
    for z in zs:
        y = _levy_cdf(z, 1.0, 1.0, mu, 0.8, 0.05, icdf.lda)
        assert abs(oracle.levy_icdf(job, y) - z) < 2e-5 * max(1, abs(z)), (z, y)


# ---- statistical pins (whole RNG-mode path) ------------------------------
def test_oracle_rng_merton_price_uniform():
    n = 1 << 16
    job = J.job_for(J.merton_sde(), 1.0, 50, n, option=EuroCall(1.0), disc=math.exp(-0.02))
    pay, _, _ = oracle.simulate(job, 0, n)
    se = pay.std() / math.sqrt(n)
    assert abs(pay.mean() - J.MERTON_PRICE) < 4 * se + 2e-4


def test_oracle_rng_merton_price_adapted():
    n = 1 << 16
    job = J.job_for(J.merton_sde(), 1.0, 50, n, grid='adapted', option=EuroCall(1.0),
                    disc=math.exp(-0.02))
    pay, _, ns = oracle.simulate(job, 0, n)
    se = pay.std() / math.sqrt(n)
    assert abs(pay.mean() - J.MERTON_PRICE) < 4 * se + 2e-4
    # a path with k jumps before T takes between N + 1 and N + k grid steps
    # (k + 1 segments, each rounded up to whole steps), about N + (k + 1) / 2
    assert ns.min() >= 50 and 50 + 1.2 < ns.mean() < 50 + 2.0


def test_oracle_rng_levy_martingale():
    # exp-Levy log price with compensated drift: E[exp(X_T)] = 1 per asset
    n = 1 << 14
    sde = J.levy_rainbow_sde()
    job = J.job_for(sde, 1.0, 50, n)
    _, term, _ = oracle.simulate(job, 0, n)
    for i in range(2):
        e = np.exp(term[0, i])
        assert abs(e.mean() - 1.0) < 4 * e.std() / math.sqrt(n) + 1e-3


def test_oracle_rng_levy_jump_free_matches_stulz():
    # rate 0 and f = 0: the log-price Euler scheme is exact, so the call on the
    # max equals the two-asset closed form
    n = 1 << 15
    sde = J.levy_rainbow_sde(f=(0.0, 0.0))
    model = sde.kernel_model()
    model.rate = 0.0
    from sde_mc_amd import engine as E
    job = E.make_job(model, 1.0, 20, L.GRID_UNIFORM, n, 99, 0)
    E.fill_payoff(job, Rainbow(1.0, log=True, discount=math.exp(0.02)), math.exp(-0.02))
    pay, _, _ = oracle.simulate(job, 0, n)
    want = cf.stulz_call_max(1, 1, 1, 1, 0.02, 0.2, 0.3, 0.5)
    assert abs(pay.mean() - want) < 4 * pay.std() / math.sqrt(n)


def test_oracle_mlmc_telescopes():
    # E[P_fine - P_coarse] from the coupled pair equals the difference of the
    # two single-level prices
    n = 1 << 15
    sde = J.merton_sde()
    opt, disc = EuroCall(1.0), math.exp(-0.02)
    c = J.job_for(sde, 1.0, 16, n, coarse_factor=2, option=opt, disc=disc)
    d, _, _ = oracle.simulate(c, 0, n)
    f = J.job_for(sde, 1.0, 16, n, option=opt, disc=disc, seed=11)
    g = J.job_for(sde, 1.0, 8, n, option=opt, disc=disc, seed=12)
    pf, _, _ = oracle.simulate(f, 0, n)
    pg, _, _ = oracle.simulate(g, 0, n)
    diff = pf.mean() - pg.mean()
    se = math.sqrt(pf.var() / n + pg.var() / n)
    assert abs(d.mean() - diff) < 4 * se
    # the coupled correction has far smaller variance than either level
    assert d.var() < 0.05 * pf.var()


# ---- jump draws ------------------------------------------------------------
def test_ln_unit_accuracy():
    rng = np.random.default_rng(3)
    for w in list(rng.integers(0, 2 ** 32, 20000)) + [0, 1, 2 ** 31, 2 ** 32 - 1]:
        exact = math.log((int(w) | 1) * 2.0 ** -32)
        # 24 significant bits of w are used: |error| <= 2^-24 relative in u
        assert abs(oracle.ln_unit(w) - exact) <= 2e-7 * abs(exact) + 7e-8


def test_poisson_thresholds_match_scipy():
    for mu in (1e-4, 0.004, 0.125, 0.5, 1.0, 3.0):
        T, inv, ok = oracle.poisson_thresholds(mu)
        # counts above 16 must be negligible for the per-step mode to apply
        assert ok == (stats.poisson.sf(16, mu) * 2 ** 32 < 0.5)
        for k in range(1, 17):
            assert abs(T[k - 1] - stats.poisson.sf(k - 1, mu) * 2 ** 32) <= 0.51
        widths = [T[k] - (T[k + 1] if k < 15 else 0) for k in range(16)]
        assert inv == [1.0 / w if w else 0.0 for w in widths]


def test_levy_jump_mode_switches_at_one_sixteenth_jump_per_step():
    # per-step counts from lambda h = 1/16 on, the arrival clock below
    sde = J.levy_rainbow_sde()  # lambda ~ 24.97
    lam = J.job_for(sde, 1.0, 10, 1).model.rate
    n_at = int(lam * 16)  # lambda / n_at just above 1/16
    assert oracle.step_jumps(J.job_for(sde, 1.0, n_at, 1), 0, 0)[0] >= 0
    assert oracle.step_jumps(J.job_for(sde, 1.0, n_at + 1, 1), 0, 0)[0] == -1
    assert oracle.step_jumps(J.job_for(sde, 1.0, 200, 1, grid='adapted'), 0, 0)[0] == -1


def test_per_step_counts_are_poisson_and_marks_follow_the_levy_law():
    # the Levy rainbow at 50 steps per year: lambda h = 24.97 / 50 ~ 0.5
    sde = J.levy_rainbow_sde()
    job = J.job_for(sde, 1.0, 50, 1, seed=77)
    mu = job.model.rate / 50
    counts, firsts, firsts_multi = [], [], []
    for p in range(4000):
        for n in range(0, 50, 5):
            c, fm = oracle.step_jumps(job, p, n)
            counts.append(c)
            if c:
                firsts.append(fm)
            if c >= 2:
                firsts_multi.append(fm)
    counts = np.array(counts)
    kmax = 5
    obs = np.array([(counts == k).sum() for k in range(kmax)] + [(counts >= kmax).sum()])
    pmf = np.array([stats.poisson.pmf(k, mu) for k in range(kmax)] + [stats.poisson.sf(kmax - 1, mu)])
    chi2 = ((obs - pmf * counts.size) ** 2 / (pmf * counts.size)).sum()
    assert chi2 < stats.chi2.ppf(0.999, kmax)
    # the residual mark: the truncated |z|^-1.8 law on 0.05 < |z| < 1
    lda = job.model.jp[5]
    cdf = lambda z: np.array([_levy_cdf(v, 1.0, 1.0, math.inf, 0.8, 0.05, lda) for v in z])
    assert stats.kstest(np.array(firsts), cdf).pvalue > 1e-3
    # ... independently of the count: the first mark of multi-jump steps too
    assert len(firsts_multi) > 500
    assert stats.kstest(np.array(firsts_multi), cdf).pvalue > 1e-3


def _heston_job(N, P, xi=0.3, rho=-0.7, seed=21, T=1.0):
    import sde_mc_amd as sm
    sde = sm.Heston(0.02, 1.5, 0.04, xi, rho, torch.tensor([1.0, 0.04]))
    return J.job_for(sde, T, N, P, option=sm.EuroCall(1.0), disc=math.exp(-0.02 * T), seed=seed)


def test_heston_zero_vol_of_vol_product_formula():
    # xi = 0: the implicit variance equation gives v' = (v + kappa theta h) / (1 + kappa h)
    # (implicit Euler of dv = kappa (theta - v) dt), and S_{n+1} = S_n (1 + r h + sqrt(v_n) dW_n)
    P, N = 6, 20
    job = _heston_job(N, P, xi=0.3)
    job.model.jp[2] = 0.0          # the oracle takes any xi; the kernels require xi > 0
    job.model.x0[1] = 0.09
    fx = J.fixed_uniform(job, seed=5)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    h = 1.0 / N
    v, S = 0.09, np.ones(P)
    dw = fx.a['dw'].astype(np.float64)
    for n in range(N):
        S = S * (1 + 0.02 * h + math.sqrt(v) * dw[n, 0])
        v = (v + 1.5 * 0.04 * h) / (1 + 1.5 * h)
    assert np.allclose(term[0, 0], S, rtol=1e-13) and np.allclose(term[0, 1], v, rtol=1e-13)


def test_heston_oracle_price_extrapolates_to_the_closed_form():
    # first-order weak error: (4 p(4N) - p(N)) / 3 removes the h term; common
    # random numbers (same seed) keep the two runs' noise correlated
    from oracle import closed_form as cf
    want = cf.heston_call(1.0, 1.0, 1.0, 0.02, 0.04, 1.5, 0.04, 0.3, -0.7)
    P = 1 << 18
    p32, _, _ = oracle.simulate(_heston_job(32, P), 0, P)
    p128, _, _ = oracle.simulate(_heston_job(128, P), 0, P)
    extrap = (4 * p128.mean() - p32.mean()) / 3
    se = np.std(4 * p128 - p32) / 3 / math.sqrt(P)
    assert abs(extrap - want) < 4 * se + 1e-4, (extrap, want, se)
    assert p32.mean() < p128.mean() < want + 4 * se  # the bias shrinks with h


def _asian_job(P, N, sigma=0.3, mu=0.05, T=1.0, grid='uniform', base=None, seed=11):
    import sde_mc_amd as sm
    base = base or sm.Gbm(mu, sigma, torch.tensor([1.0]), 1)
    return J.job_for(sm.AsianWrapper(base), T, N, P, grid=grid,
                     option=sm.AsianCall(T, 1.0), seed=seed)


def test_asian_integral_is_the_left_riemann_sum():
    """A_N = sum_n S_n h with S_n the pre-step price (AsianWrapper's drift
    x[:, 0] at the start of the step, sde.py:389-390)."""
    P, N = 16, 40
    job = _asian_job(P, N)
    fx = J.fixed_uniform(job, seed=4)
    _, term, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    dw = fx.a['dw'][:, 0, :].astype(np.float64)
    S = np.vstack([np.ones((1, P)), np.cumprod(1 + 0.05 / N + 0.3 * dw, axis=0)])
    assert np.allclose(term[0, 0], S[-1], rtol=1e-13)
    assert np.allclose(term[0, 1], S[:-1].sum(axis=0) / N, rtol=1e-13)
    # the post-step price would give a different sum
    assert not np.allclose(term[0, 1], S[1:].sum(axis=0) / N, rtol=1e-6)


def test_asian_mean_matches_the_euler_expectation():
    """E[A_N] = S_0 h sum_{n<N} (1 + mu h)^n exactly for the Euler scheme."""
    P, N, mu = 1 << 15, 50, 0.05
    job = _asian_job(P, N, mu=mu)
    pay, term, _ = oracle.simulate(job, 0, P)
    h = 1.0 / N
    want = h * sum((1 + mu * h) ** n for n in range(N))
    a = term[0, 1]
    assert abs(a.mean() - want) < 4 * a.std() / math.sqrt(P)
    # the payoff is the call on A / T (options.py:295-298)
    assert np.allclose(pay, np.maximum(a / 1.0 - 1.0, 0.0), atol=1e-12)


def test_asian_over_merton_on_the_adapted_grid():
    """E[A_T] = S_0 (e^{mu T} - 1) / mu for the compensated Merton price; the
    adapted Euler grid is within its O(h) bias of it."""
    P, N, T = 1 << 14, 100, 1.0
    job = _asian_job(P, N, T=T, grid='adapted', base=J.merton_sde(mu=0.05))
    _, term, _ = oracle.simulate(job, 0, P)
    want = (math.exp(0.05 * T) - 1) / 0.05
    a = term[0, 1]
    assert abs(a.mean() - want) < 4 * a.std() / math.sqrt(P) + 2e-3


def _torch_payoff(option, x):
    return float(option(torch.tensor([x], dtype=torch.float64)))


EXTRA_PAYOFFS = [
    ('ASIAN_CALL', lambda sm: sm.AsianCall(2.0, 1.0), (1.1, 2.4)),
    ('ASIAN_CALL', lambda sm: sm.AsianCall(2.0, 1.3), (1.1, 2.4)),
    ('ASIAN_CALL', lambda sm: sm.AsianCall(2.0, 1.0, log=True, discount=0.9), (0.1, -0.2)),
    ('RAINBOW_EVEN', lambda sm: sm.HestonRainbow(1.0), (1.2, 0.04)),
    ('RAINBOW_EVEN', lambda sm: sm.HestonRainbow(1.1), (1.2, 0.04, 1.5, 0.09)),
    ('RAINBOW_EVEN', lambda sm: sm.HestonRainbow(1.6), (1.2, 0.04, 1.5, 0.09)),
]


@pytest.mark.parametrize("kind,make,x", EXTRA_PAYOFFS)
def test_oracle_asian_and_heston_rainbow_payoffs(kind, make, x):
    """The oracle's payoff of a frozen state against the torch option (the
    reference's composition: Option.transform, then payoff)."""
    import sde_mc_amd as sm
    opt = make(sm)
    job = J.frozen_state_job(kind, opt.strike, x)
    assert opt.kernel_payoff(job.payoff)
    pay, _, _ = oracle.simulate(job, 0, 1)
    assert pay[0] == pytest.approx(_torch_payoff(opt, x), rel=1e-12, abs=1e-14)

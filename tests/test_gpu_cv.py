"""Fused neural control variates on the device.

* zero networks: gamma equals the plain kernel's payoffs bit for bit (the
  control-variate kernels walk the same paths);
* random networks: gamma against a reference evaluation of the same bf16
  network (tests/cv_emul.py, fp64 sums) on the paths, increments, jump marks
  and grid times the plain kernel stores for the same job;
* trained networks: the control variate has mean zero and cuts the
  variance, and the estimators run end to end (fused and torch routes).
"""
import ctypes
import math

import numpy as np
import pytest
import torch

import sde_mc_amd as sm
from sde_mc_amd import _lib as L
from sde_mc_amd import cv as CV
from sde_mc_amd import engine as E
from tests import cv_emul
from tests import jobs as J

pytestmark = pytest.mark.gpu


def _dev():
    return E.require_device()


def _diag(job):
    return job.model.noise in (L.NOISE_DIAG, L.NOISE_HESTON)


def _store(job):
    """Rows of a plain run of `job` stored on the device, as numpy."""
    dev = _dev()
    P, d = int(job.n_paths), job.model.dim
    adapted = job.grid == L.GRID_ADAPTED
    rows = job.max_steps if adapted else job.steps
    kn = d if _diag(job) else 2 * d
    t = {'paths': torch.empty((rows + 1, d, P), dtype=torch.float32, device=dev),
         'normals': torch.zeros((rows, kn, P), dtype=torch.float32, device=dev),
         'jumps': torch.zeros((rows + 1, d, P), dtype=torch.float32, device=dev),
         'nsteps': torch.empty(P, dtype=torch.int32, device=dev),
         'payoff': torch.empty(P, dtype=torch.float32, device=dev)}
    out = L.PathOut()
    out.paths_dev, out.normals_dev = t['paths'].data_ptr(), t['normals'].data_ptr()
    out.jumps_dev, out.nsteps_dev = t['jumps'].data_ptr(), t['nsteps'].data_ptr()
    out.payoff_dev = t['payoff'].data_ptr()
    if adapted:
        t['times'] = torch.empty((rows + 1, P), dtype=torch.float32, device=dev)
        t['left'] = torch.empty((rows + 1, d, P), dtype=torch.float32, device=dev)
        out.times_dev, out.left_dev = t['times'].data_ptr(), t['left'].data_ptr()
    stats = E.simulate(job, dev, out=out)
    E.stats_to_host(stats)
    return {k: v.cpu().numpy() for k, v in t.items()}


def _net(n_in, layers, n_out, seed):
    torch.manual_seed(seed)
    return sm.Mlp(n_in, layers, n_out, batch_norm=False, batch_norm_init=False).eval()


def _features(t, X, d):
    F = np.zeros((X.shape[1], 8))
    F[:, 0] = t
    F[:, 1:1 + d] = X.T
    F[:, 1 + d] = 1.0
    return F


def _cv_run(job, f, g, cv_steps, rate, comp):
    dev = _dev()
    d = job.model.dim
    brown = d if _diag(job) else 2 * d
    fn = CV.DeviceNetwork(f, d + 1, brown, dev)
    gn = CV.DeviceNetwork(g, d + 1, d, dev) if g is not None else None
    print('cv job P=%d N=%d grid=%d noise=%d dim=%d f=%d g=%d' % (
        job.n_paths, job.steps, job.grid, job.model.noise, d, fn.hidden,
        gn.hidden if gn else -1), flush=True)
    stats, gam = CV.cv_simulate(job, fn, gn, cv_steps, rate, comp, dev, gamma=True)
    return stats.cpu().numpy(), gam.cpu().numpy()


def _uniform_reference(st, job, f, cv_steps, rate):
    d = job.model.dim
    fm = CV.padded_matrices(f, d + 1, d if _diag(job) else 2 * d)
    h = np.float32(job.T / job.steps)
    P = int(job.n_paths)
    cv = np.zeros(P)
    size = np.zeros(P)
    for n in range(cv_steps):
        t = float(np.float32(n) * h)
        fo = cv_emul.matrix_forward(fm, _features(t, st['paths'][n], d))
        dW = st['normals'][n].astype(np.float64)
        terms = fo[:, :dW.shape[0]] * dW.T
        disc = math.exp(-rate * t)
        cv += disc * terms.sum(1)
        size += disc * np.abs(terms).sum(1)
    return st['payoff'] + cv, size


def _adapted_reference(st, job, f, g, rate, comp):
    d = job.model.dim
    brown = d if _diag(job) else 2 * d
    fm = CV.padded_matrices(f, d + 1, brown)
    gm = CV.padded_matrices(g, d + 1, d)
    P = int(job.n_paths)
    ns = st['nsteps']
    rows = st['normals'].shape[0]
    bcv, jcv, ccv, size = np.zeros(P), np.zeros(P), np.zeros(P), np.zeros(P)
    for n in range(int(ns.max()) + 1):
        t = st['times'][n].astype(np.float64)
        disc = np.exp(-rate * t)
        go = cv_emul.matrix_forward(gm, _features(t, st['left'][n], d))
        gs = go[:, :d].sum(1)
        J = st['jumps'][n][0].astype(np.float64)
        jt = np.where(n <= ns, disc * gs * J, 0.0)
        jcv += jt
        size += np.abs(jt)
        if n < rows and n < ns.max():
            live = n < ns
            fo = cv_emul.matrix_forward(fm, _features(t, st['paths'][n], d))
            dW = st['normals'][n].astype(np.float64)
            bt = np.where(live, disc * (fo[:, :brown] * dW.T).sum(1), 0.0)
            dt = st['times'][n + 1].astype(np.float64) - t
            ct = np.where(live, disc * gs * dt, 0.0)
            bcv += bt
            ccv += ct
            size += np.abs(bt) + abs(comp) * np.abs(ct)
    return st['payoff'] + bcv + jcv - comp * ccv, size


def _compare(gamma, ref, size, what):
    err = np.abs(gamma.astype(np.float64) - ref)
    rel = err / (size + np.abs(ref) + 1e-6)
    print('%s: max err %.3e  max rel %.3e  mean rel %.3e' % (what, err.max(), rel.max(),
                                                             rel.mean()), flush=True)
    # bf16 roundings flip now and then under a different fp32 summation order
    assert rel.max() < 5e-3
    assert rel.mean() < 2e-4


UNIFORM_CASES = {
    'gbm1': lambda: J.job_for(sm.Gbm(0.02, 0.3, torch.tensor([1.0]), 1), 3.0, 40, 3000,
                              option=sm.EuroCall(1.0), disc=math.exp(-0.06), seed=5),
    'gbm3': lambda: J.job_for(sm.Gbm(0.02, 0.3, torch.tensor([1.0] * 3), 3,
                                     sm.get_corr_matrix([0.7, 0.2, -0.3])), 3.0, 24, 2000,
                              option=sm.Rainbow(1.0), disc=math.exp(-0.06), seed=6),
    'heston': lambda: J.job_for(sm.Heston(0.02, 1.5, 0.04, 0.3, -0.7, torch.tensor([1.0, 0.04])),
                                1.0, 32, 2500, option=sm.EuroCall(1.0), disc=math.exp(-0.02),
                                seed=7),
    'double2': lambda: J.job_for(sm.DoubleGbm(0.02, 0.2, 0.1, torch.tensor([1.0, 1.0]), 2),
                                 1.0, 20, 2000, option=sm.Basket(1.0), seed=8),
}

ADAPTED_CASES = {
    'merton1': lambda: J.job_for(J.merton_sde(), 1.0, 30, 3000, grid='adapted',
                                 option=sm.EuroCall(1.0), disc=math.exp(-0.02), seed=9),
    'levy2': lambda: J.job_for(J.levy_rainbow_sde(), 1.0, 20, 2000, grid='adapted',
                               option=sm.Rainbow(1.0, log=True), disc=math.exp(-0.02), seed=10),
}


@pytest.mark.parametrize("case", sorted(UNIFORM_CASES) + sorted(ADAPTED_CASES))
def test_zero_networks_give_plain_payoffs_exactly(case):
    job = (UNIFORM_CASES.get(case) or ADAPTED_CASES[case])()
    st = _store(job)
    s, gam = _cv_run(job, None, None, 0, 0.02, 0.3)
    assert np.array_equal(gam, st['payoff'])
    assert s[0] == pytest.approx(float(st['payoff'].astype(np.float64).sum()), rel=1e-12)
    assert s[2] == job.n_paths


@pytest.mark.parametrize("case,layers", [('gbm1', [50, 50, 50]), ('gbm3', [30, 30]),
                                         ('heston', [63]), ('double2', [40, 20, 40, 20])])
def test_uniform_network_matches_reference_evaluation(case, layers):
    job = UNIFORM_CASES[case]()
    d = job.model.dim
    f = _net(d + 1, layers, d if _diag(job) else 2 * d, seed=len(layers) + d)
    st = _store(job)
    cv_steps = job.steps - 3
    s, gam = _cv_run(job, f, None, cv_steps, 0.02, 0.0)
    ref, size = _uniform_reference(st, job, f, cv_steps, 0.02)
    _compare(gam, ref, size, case)
    assert s[0] == pytest.approx(float(gam.astype(np.float64).sum()), rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("case", sorted(ADAPTED_CASES))
def test_adapted_networks_match_reference_evaluation(case):
    job = ADAPTED_CASES[case]()
    d = job.model.dim
    f = _net(d + 1, [50, 50, 50], d if _diag(job) else 2 * d, seed=11)
    g = _net(d + 1, [40, 40], d, seed=12)
    comp = 0.37
    st = _store(job)
    s, gam = _cv_run(job, f, g, int(job.max_steps), 0.02, comp)
    ref, size = _adapted_reference(st, job, f, g, 0.02, comp)
    _compare(gam, ref, size, case)
    assert (st['jumps'] != 0).any()


def test_fused_adapted_agrees_with_the_torch_loop():
    """apply_adapted_control_variates (fp32 torch networks on stored paths)
    against the fused kernel on paths other than the batch's longest, whose
    final row the torch loop leaves out."""
    p = sm.MertonEuroCall.default_params(20, 'cuda')
    f, g = sm.get_mlps(p, 2, 20, 'cuda')
    f.eval()
    g.eval()
    solver = p.solver
    off = solver.path_counter
    n = 2048
    dl = sm.simulate_adapted_data(n, solver, p.payoff, p.discounter, bs=n, inference=True)
    (paths, normals, left, times, jumps), pay = next(iter(dl))
    gt = sm.varred._adapted_gammas([f, g], (paths, normals, left, times, jumps), pay, solver,
                                   p.discounter, 0).detach().cpu().numpy()
    solver.path_counter = off
    fused = sm.mc.fused_control_variates([f, g], solver, p.payoff, p.discounter)
    assert fused is not None
    _, _, gam = fused.run(n, gamma=True)
    gam = gam.cpu().numpy()
    total = dl.dataset.total_steps
    ts = times.cpu().numpy()[:, :, 0]
    # paths that reach T before the batch's last row are complete in both
    done = ts[:, -1] >= np.float32(solver.time_interval)
    assert done.sum() > n // 2
    err = np.abs(gam[done] - gt[done])
    print('fused vs torch loop: max %.3e mean %.3e (total steps %d)' % (err.max(), err.mean(),
                                                                          total), flush=True)
    assert err.mean() < 2e-3 and err.max() < 5e-2


def test_trained_control_variate_is_unbiased_and_reduces_variance():
    p = sm.BlackScholesEuroCall.default_params(50, 'cuda')
    f = sm.get_mlps(p, 2, 30, 'cuda')
    opt = sm.get_opt(f)
    sm.sim_train_control_variates(f, opt, p.solver, 20000, p.payoff, p.discounter, 20000, 500,
                                  epochs=8, print_losses=False)
    fused = sm.mc.fused_control_variates(f, p.solver, p.payoff, p.discounter)
    assert fused is not None
    n = 1 << 20
    off = p.solver.path_counter
    _, _, gam = fused.run(n, gamma=True)
    p.solver.path_counter = off
    pay, _ = p.solver.simulate_payoffs(n, p.payoff, float(p.discounter(p.solver.time_interval)))
    gam = gam.double()
    pay = pay.double()
    cv = gam - pay
    se = float(cv.std()) / math.sqrt(n)
    print('mean cv %.3e se %.3e  var ratio %.4f' % (float(cv.mean()), se,
                                                   float(gam.var() / pay.var())), flush=True)
    assert abs(float(cv.mean())) < 4 * se
    assert float(gam.var()) < 0.25 * float(pay.var())


def test_estimators_end_to_end():
    p = sm.BlackScholesEuroCall.default_params(20, 'cuda')
    f = sm.get_mlps(p, 1, 10, 'cuda')
    opt = sm.get_opt(f)
    stats = sm.mc_control_variates(f, opt, p.solver, (4000, 1 << 18), (10, 20), p.payoff,
                                   p.discounter, bs=(500, 1000), epochs=2, print_losses=False)
    assert math.isfinite(stats.sample_mean) and stats.sample_std > 0
    exact = sm.bs_call(1.0, 1.0, 3.0, 0.02, 0.3)
    assert abs(stats.sample_mean - exact) < 5 * stats.sample_std + 0.01
    # ZeroFunction through the fused path equals mc_simple on the same paths
    off = p.solver.path_counter
    z = sm.mc_apply_cvs(sm.ZeroFunction(1), p.solver, 1 << 16, p.payoff, p.discounter)
    p.solver.path_counter = off
    plain = sm.mc_simple(1 << 16, p.solver, p.payoff, p.discounter, bs=1 << 16)
    assert z.sample_mean == pytest.approx(plain.sample_mean, rel=1e-12)
    # a recurrent model takes the torch route on stored paths
    lstm = sm.Lstm(2, 8, 1, device='cuda')
    r = sm.mc_apply_cvs(lstm, p.solver, 2000, p.payoff, p.discounter, sim_bs=2000, bs=1000)
    assert math.isfinite(r.sample_mean)
    # jump model: [f, g] trained on the adapted grid, applied fused
    q = sm.MertonEuroCall.default_params(20, 'cuda')
    fg = sm.get_mlps(q, 1, 10, 'cuda')
    st = sm.mc_adaptive_cv(fg, sm.get_opt(fg), q.solver, (4000, 1 << 18), (10, 20), q.payoff,
                           q.discounter, bs=(500, 1000), epochs=2, print_losses=False)
    assert math.isfinite(st.sample_mean) and st.sample_std > 0
    tc = sm.mc_terminal_cv(1 << 16, p.solver, p.payoff, p.discounter, bs=1 << 15)
    assert math.isfinite(tc.sample_mean)


@pytest.mark.parametrize("P", [0, 1, 63, 65])
def test_cv_ragged_and_empty_jobs(P):
    job = J.job_for(J.merton_sde(), 1.0, 12, P, grid='adapted', option=sm.EuroCall(1.0), seed=21)
    f = _net(2, [20, 20], 1, seed=1)
    g = _net(2, [20], 1, seed=2)
    s, gam = _cv_run(job, f, g, int(job.max_steps), 0.02, 0.1)
    assert s[2] == P and gam.shape == (P,)
    if P:
        st = _store(job)
        ref, size = _adapted_reference(st, job, f, g, 0.02, 0.1)
        assert np.abs(gam - ref).max() <= 5e-3 * (size + np.abs(ref) + 1e-6).max()
        assert s[0] == pytest.approx(float(gam.astype(np.float64).sum()), rel=1e-9, abs=1e-12)
    else:
        assert s[0] == 0.0 and s[1] == 0.0


def test_cv_steps_zero_leaves_the_payoff():
    job = UNIFORM_CASES['gbm1']()
    st = _store(job)
    _, gam = _cv_run(job, _net(2, [30], 1, seed=3), None, 0, 0.02, 0.0)
    assert np.array_equal(gam, st['payoff'])


def test_tol_truncation_matches_the_torch_loop():
    """mc_apply_cvs' tol (integrate_cv's remove_steps cut) in the fused
    kernel against the torch loop on the same stored paths."""
    p = sm.BlackScholesEuroCall.default_params(30, 'cuda')
    f = sm.get_mlps(p, 2, 20, 'cuda')
    f.eval()
    tol = 0.35
    n = 2048
    off = p.solver.path_counter
    dl = sm.simulate_data(n, p.solver, p.payoff, p.discounter, bs=n, inference=True)
    (paths, normals), pay = next(iter(dl))
    tp = sm.varred._grid(p.solver, paths.device)
    disc = p.discounter(tp).view(1, -1, 1)
    gt = sm.varred._diffusion_gammas(f, paths, normals, pay, tp, disc, p.solver,
                                     tol).detach().cpu().numpy()
    p.solver.path_counter = off
    fused = sm.mc.fused_control_variates(f, p.solver, p.payoff, p.discounter, tol=tol)
    assert fused.cv_steps == sm.remove_steps(tol, 30, 3.0) == 26
    _, _, gam = fused.run(n, gamma=True)
    err = np.abs(gam.cpu().numpy() - gt)
    print('tol=%.2f fused vs torch: max %.3e mean %.3e' % (tol, err.max(), err.mean()), flush=True)
    assert err.mean() < 2e-3 and err.max() < 5e-2


def test_jump_cv_on_the_uniform_grid_is_refused():
    q = sm.MertonEuroCall(0.02, 0.2, 1, -0.05, 0.3, 1, 1, 3, 20, 'cuda', grid='uniform')
    fg = sm.get_mlps(q, 1, 10, 'cuda')
    assert sm.mc.fused_control_variates(fg, q.solver, q.payoff, q.discounter) is None
    with pytest.raises(NotImplementedError):
        sm.mc_apply_cvs(fg, q.solver, 1000, q.payoff, q.discounter, sim_bs=1000, bs=1000)


def test_cv_entry_point_refuses_bad_arguments():
    """sdemc_cv_simulate validates on the host and launches nothing on error."""
    dev = _dev()
    lib = L.lib()
    f = CV.DeviceNetwork(_net(2, [16], 1, seed=5), 2, 1, dev)
    job = UNIFORM_CASES['gbm1']()
    stats = torch.zeros(4, dtype=torch.float64, device=dev)
    ws = torch.empty(lib.sdemc_workspace_bytes(job.n_paths), dtype=torch.uint8, device=dev)

    def call(job, spec):
        return lib.sdemc_cv_simulate(ctypes.byref(job), ctypes.byref(spec), E.ptr(stats), None,
                                     E.ptr(ws), ws.numel(), E.stream_ptr())

    def spec(**kw):
        s = CV.CvSpec()
        s.f = f.c()
        s.cv_steps = job.steps
        s.rate = 0.02
        for k, v in kw.items():
            setattr(s, k, v)
        return s
    assert call(job, spec()) == L.OK
    bad = spec()
    bad.f.hidden_layers = 5
    assert call(job, bad) == L.ERR_ARG
    bad = spec()
    bad.f.n_out = 2
    assert call(job, bad) == L.ERR_ARG
    bad = spec()
    bad.f.weights_dev = f.words.data_ptr() + 2         # misaligned
    assert call(job, bad) == L.ERR_ARG
    assert call(job, spec(cv_steps=job.steps + 1)) == L.ERR_ARG
    assert call(job, spec(g=f.c())) == L.ERR_ARG       # jump network on the uniform grid
    dbl = J.job_for(sm.DoubleGbm(0.02, 0.2, 0.1, torch.tensor([1.0]), 1), 1.0, 10, 100,
                    grid='adapted', option=sm.EuroCall(1.0))
    f2 = CV.DeviceNetwork(_net(2, [16], 2, seed=6), 2, 2, dev)
    s2 = spec()
    s2.f = f2.c()
    assert call(dbl, s2) == L.ERR_UNSUPPORTED          # 'indep' full noise on the adapted grid
    mer = J.job_for(J.merton_sde(), 1.0, 10, 100, option=sm.EuroCall(1.0))
    assert call(mer, spec()) == L.ERR_UNSUPPORTED      # jumps on the uniform grid

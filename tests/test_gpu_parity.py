"""GPU parity: the HIP kernels (through the C ABI) against the fp64 C oracle on
the same inputs, path by path.

Tolerances (fp32 stepping against fp64):
  * terminal states: |x32 - x64| <= 2e-5 (1 + |x64|) per path, the fp32
    rounding of ~10^3 Euler steps plus the hardware transcendentals of
    Box-Muller (v_log/v_sqrt/v_sin/v_cos, ~1e-6 relative per normal);
  * adapted-grid step counts: identical (the clock is fp32 on both sides);
  * prices from the fp64 sums: the oracle mean within 1e-5 absolute.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

import oracle
from sde_mc_amd import _lib as L
from sde_mc_amd import engine as E
from sde_mc_amd import EuroCall, Rainbow
from tests import jobs as J

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _dev():
    return E.require_device()


def _run_rng(job):
    dev = _dev()
    P = int(job.n_paths)
    lv = 2 if job.coarse_factor else 1
    d = job.model.dim
    pay = torch.empty(P, dtype=torch.float32, device=dev)
    term = torch.empty((lv, d, P), dtype=torch.float32, device=dev)
    ns = torch.empty(P, dtype=torch.int32, device=dev)
    out = L.PathOut()
    out.payoff_dev, out.terminal_dev, out.nsteps_dev = pay.data_ptr(), term.data_ptr(), ns.data_ptr()
    print('rng job P=%d N=%d grid=%d M=%d noise=%d' % (P, job.steps, job.grid, job.coarse_factor,
                                                      job.model.noise), flush=True)
    stats = E.simulate(job, dev, out=out)
    s = stats.cpu().numpy()
    return pay.cpu().numpy(), term.cpu().numpy(), ns.cpu().numpy(), s


def _run_fixed(job, fx):
    dev = _dev()
    P = int(job.n_paths)
    lv = 2 if job.coarse_factor else 1
    d = job.model.dim
    pay = torch.empty(P, dtype=torch.float32, device=dev)
    term = torch.empty((lv, d, P), dtype=torch.float32, device=dev)
    ns = torch.empty(P, dtype=torch.int32, device=dev)
    out = L.PathOut()
    out.payoff_dev, out.terminal_dev, out.nsteps_dev = pay.data_ptr(), term.data_ptr(), ns.data_ptr()
    fin = fx.device_struct(dev)
    print('fixed job P=%d N=%d grid=%d M=%d' % (P, job.steps, job.grid, job.coarse_factor), flush=True)
    stats = E.simulate_fixed(job, fin, dev, out=out)
    s = stats.cpu().numpy()
    return pay.cpu().numpy(), term.cpu().numpy(), ns.cpu().numpy(), s


def _close(a32, b64, tol=TOL):
    err = np.abs(a32.astype(np.float64) - b64) / (1 + np.abs(b64))
    return float(err.max())


# ---- generator ------------------------------------------------------------
def test_device_philox_known_answers():
    dev = _dev()
    from tests.test_oracle_paths import KAT
    for ctr, key, want in KAT:
        c = torch.tensor([v - (1 << 32) if v >= 1 << 31 else v for v in ctr], dtype=torch.int32,
                         device=dev)
        o = torch.empty(4, dtype=torch.int32, device=dev)
        L.check(L.lib().sdemc_philox_dev(ctypes.c_void_p(c.data_ptr()), key[0], key[1],
                                         ctypes.c_void_p(o.data_ptr()), 1, E.stream_ptr()), 'philox')
        got = [int(v) & 0xFFFFFFFF for v in o.cpu().tolist()]
        assert got == want


def test_hbm_read_probe_reads_every_word():
    # the XOR of the probe's outputs is the XOR of every 32-bit word of the
    # source, at a size whose last grid-stride pass is ragged
    dev = _dev()
    words = np.random.default_rng(5).integers(0, 1 << 32, size=(1 << 18) + 4 * 37,
                                              dtype=np.uint32)
    src = torch.from_numpy(words.view(np.int32)).to(dev)
    blocks = 7
    out = torch.empty(blocks * 256, dtype=torch.int32, device=dev)
    L.check(L.lib().sdemc_hbm_read_probe_dev(ctypes.c_void_p(src.data_ptr()), words.nbytes,
                                             ctypes.c_void_p(out.data_ptr()), blocks,
                                             E.stream_ptr()), 'hbm read probe')
    got = np.bitwise_xor.reduce(out.cpu().numpy().view(np.uint32))
    assert got == np.bitwise_xor.reduce(words)


def test_device_normals_match_oracle():
    dev = _dev()
    P, n = 256, 64
    out = torch.empty((n, P), dtype=torch.float32, device=dev)
    L.check(L.lib().sdemc_normals_dev(12345, 1000, P, n, ctypes.c_void_p(out.data_ptr()),
                                      E.stream_ptr()), 'normals')
    z = out.cpu().numpy().astype(np.float64)
    ref = np.array([[oracle.normal(12345, 1000 + p, i) for p in range(P)] for i in range(n)])
    assert np.abs(z - ref).max() < 5e-6 * (1 + np.abs(ref)).max()


# ---- fixed-increment parity -------------------------------------------------
def test_fixed_uniform_merton_2p16_x_500():
    # 2^16 paths x 500 steps of host N(0, dt) increments, Poisson counts and
    # lognormal marks: every path against the fp64 oracle
    P, N = 1 << 16, 500
    job = J.job_for(J.merton_sde(), 1.0, N, P, option=EuroCall(1.0), disc=math.exp(-0.02))
    fx = J.fixed_uniform(job, seed=10)
    pay, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term[0, 0], oterm[0, 0]) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5
    assert s[2] == P and s[3] == 0


@pytest.mark.parametrize("grid,coupled", [('uniform', 2), ('adapted', 0), ('adapted', 2)])
def test_fixed_merton_variants(grid, coupled):
    P, N = 4096 + 37, 64
    job = J.job_for(J.merton_sde(rate=4.0), 1.0, N, P, grid=grid, coarse_factor=coupled,
                    option=EuroCall(1.0), disc=math.exp(-0.02))
    fx = J.fixed_adapted(job, seed=11) if grid == 'adapted' else J.fixed_uniform(job, seed=11)
    pay, term, ns, s = _run_fixed(job, fx)
    opay, oterm, ons = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    if grid == 'adapted':
        assert np.array_equal(ns, ons)
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("grid", ['uniform', 'adapted'])
def test_fixed_levy_rainbow(grid):
    P, N = 2048 + 5, 100
    job = J.job_for(J.levy_rainbow_sde(), 1.0, N, P, grid=grid,
                    option=Rainbow(1.0, log=True, discount=math.exp(0.02)), disc=math.exp(-0.02))
    fx = J.fixed_adapted(job, seed=12) if grid == 'adapted' else J.fixed_uniform(job, seed=12)
    pay, term, ns, s = _run_fixed(job, fx)
    opay, oterm, ons = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    if grid == 'adapted':
        assert np.array_equal(ns, ons)


# ---- RNG-mode parity (same Philox counters on both sides) ---------------------
CASES = [
    ('merton', 'uniform', 0), ('merton', 'adapted', 0), ('merton', 'uniform', 2),
    ('merton', 'adapted', 2), ('levy', 'uniform', 0), ('levy', 'adapted', 0),
    ('levy', 'uniform', 2), ('merton3', 'uniform', 0),
    # Levy law at lambda h ~ 0.042 < 1/16: the uniform grid takes the arrival
    # clock instead of per-step counts
    ('levy_clock', 'uniform', 0), ('levy_clock', 'uniform', 2),
]


@pytest.mark.parametrize("model,grid,coupled", CASES)
def test_rng_paths_match_oracle(model, grid, coupled):
    P, N = 1000, 96
    if model == 'merton':
        sde, opt = J.merton_sde(rate=3.0), EuroCall(1.0)
    elif model == 'merton3':
        corr = torch.tensor([[1.0, 0.3, 0.1], [0.3, 1.0, -0.2], [0.1, -0.2, 1.0]])
        sde, opt = J.merton_sde(rate=3.0, dim=3, corr=corr), Rainbow(1.0)
    elif model == 'levy_clock':
        sde, opt = J.levy_rainbow_sde(eps=0.3), Rainbow(1.0, log=True, discount=math.exp(0.02))
    else:
        sde, opt = J.levy_rainbow_sde(), Rainbow(1.0, log=True, discount=math.exp(0.02))
    job = J.job_for(sde, 1.0, N, P, grid=grid, coarse_factor=coupled, option=opt,
                    disc=math.exp(-0.02), seed=777, path_offset=12345)
    if model.startswith('levy') and grid == 'uniform':
        # the jump mode the case is meant to cover
        assert (oracle.step_jumps(job, 0, 0)[0] < 0) == (model == 'levy_clock')
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    if grid == 'adapted':
        assert np.array_equal(ns, ons)
    assert abs(s[0] / P - opay.mean()) < 1e-5
    assert s[2] == P


# the Levy rainbow exactly as it is benchmarked and served (jump scale f = 0.2
# per asset, sde_mc problem.py:158; eps = 0.05, so lambda ~ 25 per year): per
# step counts at 200 steps (lambda h ~ 0.12), the arrival clock at 800 (~0.03)
BENCH_LEVY = [('uniform', 200, 0), ('uniform', 800, 0), ('uniform', 200, 2),
              ('adapted', 200, 0), ('adapted', 200, 2)]


@pytest.mark.parametrize("grid,N,coupled", BENCH_LEVY)
def test_rng_paths_of_the_benchmarked_levy_model(grid, N, coupled):
    from sde_mc_amd import workloads as W
    P = 1000
    sde = W.levy_rainbow()
    job = J.job_for(sde, 1.0, N, P, grid=grid, coarse_factor=coupled,
                    option=Rainbow(1.0, log=True, discount=math.exp(0.02)),
                    disc=math.exp(-0.02), seed=4242, path_offset=77)
    assert abs(job.model.j1[0]) < 1e-12 and abs(job.model.j0[0] - 0.2) < 1e-7
    if grid == 'uniform' and not coupled:
        assert (oracle.step_jumps(job, 0, 0)[0] < 0) == (N == 800)
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    if grid == 'adapted':
        assert np.array_equal(ns, ons)
    assert abs(s[0] / P - opay.mean()) < 1e-5
    assert s[2] == P and s[3] == 0


def test_fixed_benchmarked_levy_model():
    from sde_mc_amd import workloads as W
    P, N = 4096 + 3, 200
    job = J.job_for(W.levy_rainbow(), 1.0, N, P,
                    option=Rainbow(1.0, log=True, discount=math.exp(0.02)), disc=math.exp(-0.02))
    fx = J.fixed_uniform(job, seed=13)
    pay, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


def _fixed_struct_shifted(fx, dev, shift_dw, shift_counts):
    """Device inputs of fx with dw placed `shift_dw` floats and counts
    `shift_counts` bytes past the start of an allocation: off the 16-byte /
    4-byte alignment the four-paths-per-lane kernel needs."""
    fin = fx.device_struct(dev)
    keep = []
    if shift_dw:
        dw = fx.t['dw'].reshape(-1)
        buf = torch.empty(dw.numel() + shift_dw, dtype=torch.float32, device=dev)
        buf[shift_dw:] = dw
        fin.dw_dev = buf.data_ptr() + 4 * shift_dw
        keep.append(buf)
    if shift_counts:
        c = fx.t['counts'].reshape(-1)
        buf = torch.empty(c.numel() + shift_counts, dtype=torch.uint8, device=dev)
        buf[shift_counts:] = c
        fin.counts_dev = buf.data_ptr() + shift_counts
        keep.append(buf)
    return fin, keep


@pytest.mark.parametrize("model,d", [('merton', 1), ('merton', 3), ('levy', 2)])
def test_fixed_four_paths_per_lane_equals_one_path_per_lane(model, d):
    # P % 4 == 0 with aligned rows runs the four-paths-per-lane kernel
    # (16-byte row loads); the same inputs placed off that alignment run the
    # one-path-per-lane kernel: every path, payoff and the job's sums agree
    # bit for bit (ragged last wavefront, jumps on)
    dev = _dev()
    P, N = 4 * (1024 + 3 * 64 + 5), 67
    if model == 'merton':
        sde, opt = J.merton_sde(rate=4.0, dim=d), EuroCall(1.0)
    else:
        sde, opt = J.levy_rainbow_sde(), Rainbow(1.0, log=True, discount=math.exp(0.02))
    job = J.job_for(sde, 1.0, N, P, option=opt, disc=math.exp(-0.02))
    fx = J.fixed_uniform(job, seed=21)
    assert int(fx.a['counts'].sum()) > 0
    res = []
    for shift in ((0, 0), (1, 0), (0, 1)):
        print('fixed job P=%d d=%d shift=%s' % (P, d, shift), flush=True)
        fin, keep = _fixed_struct_shifted(fx, dev, *shift)
        pay = torch.empty(P, dtype=torch.float32, device=dev)
        term = torch.empty((1, d, P), dtype=torch.float32, device=dev)
        ns = torch.empty(P, dtype=torch.int32, device=dev)
        out = L.PathOut()
        out.payoff_dev, out.terminal_dev, out.nsteps_dev = (pay.data_ptr(), term.data_ptr(),
                                                            ns.data_ptr())
        stats = E.simulate_fixed(job, fin, dev, out=out)
        res.append((pay.cpu().numpy(), term.cpu().numpy(), ns.cpu().numpy(),
                    stats.cpu().numpy()))
    assert res[0][3][3] == 0 and res[0][3][2] == P
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(a, b)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(res[0][1], oterm) < TOL


def test_fixed_jump_reads_outside_the_marks_raise():
    # counts that claim more marks than jump_size holds: the kernel refuses
    # the reads past n_marks (in bounds, mark 0) and flags the job, which the
    # host reports as an argument error instead of a price
    P, N = 256, 20
    job = J.job_for(J.merton_sde(rate=4.0), 1.0, N, P, option=EuroCall(1.0))
    fx = J.fixed_uniform(job, seed=14)
    assert fx.a['js'].size > 8
    fx.a['js'] = fx.a['js'][:8].copy()
    print('fixed job with 8 marks for %d jumps' % int(fx.a['counts'].sum()), flush=True)
    stats = E.simulate_fixed(job, fx.device_struct(_dev()), _dev())
    with pytest.raises(L.SdemcError) as e:
        E.stats_to_host(stats, 'short jump_size')
    assert e.value.status == L.ERR_ARG and 'jump CSR' in str(e.value)


@pytest.mark.parametrize("coupled", [0, 2])
def test_adapted_low_rate_long_refill_windows(coupled):
    """lambda h = 0.0015: the adapted grid refills the second jump slots only
    every 128 iterations (64 coarse iterations of a coupled pair), so lanes
    with two jumps inside one window draw the next in the step; the paths and
    step counts still follow the oracle."""
    P, N = 1500, 1000
    job = J.job_for(J.merton_sde(rate=1.5), 1.0, N, P, grid='adapted', coarse_factor=coupled,
                    option=EuroCall(1.0), disc=math.exp(-0.02), seed=929)
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert np.array_equal(ns, ons)
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("eps,N,coupled", [(0.05, 101, 0), (0.05, 102, 3), (0.3, 101, 0),
                                           (0.3, 102, 3)])
def test_levy_partial_last_jump_window(eps, N, coupled):
    """Step counts that end inside a jump window (per-step counts draw the
    jumps of 8 steps at a time, the clock fills its queue per window): the
    paths still follow the oracle, single level and coupled."""
    P = 1000
    job = J.job_for(J.levy_rainbow_sde(eps=eps), 1.0, N, P, coarse_factor=coupled,
                    option=Rainbow(1.0, log=True), seed=919)
    assert (oracle.step_jumps(job, 0, 0)[0] < 0) == (eps == 0.3)
    pay, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


@pytest.mark.parametrize("eps", [0.05, 0.3])
def test_levy_stored_paths_in_both_jump_modes(eps):
    """The path-materialising kernel in either uniform-grid jump mode
    (per-step counts at lambda h ~ 0.26, the arrival clock at ~0.042): its
    last stored row is the terminal state of the kernel that stores nothing,
    bit for bit, its first the initial state, and the steps with a stored
    jump sum occur at the Poisson rate 1 - exp(-lambda h)."""
    P, N = 777, 96
    job = J.job_for(J.levy_rainbow_sde(eps=eps), 1.0, N, P, option=Rainbow(1.0, log=True),
                    seed=43)
    assert (oracle.step_jumps(job, 0, 0)[0] < 0) == (eps == 0.3)
    _, term, _, _ = _run_rng(job)
    dev = _dev()
    paths = torch.full((N + 1, 2, P), float('nan'), dtype=torch.float32, device=dev)
    jumps = torch.zeros((N + 1, 2, P), dtype=torch.float32, device=dev)
    out = L.PathOut()
    out.paths_dev, out.jumps_dev = paths.data_ptr(), jumps.data_ptr()
    print('stored levy P=%d N=%d eps=%g' % (P, N, eps), flush=True)
    E.simulate(job, dev, out=out)
    p, jp = paths.cpu().numpy(), jumps.cpu().numpy()
    assert np.array_equal(p[N], term[0])
    assert np.all(p[0] == 0.0) and np.all(np.isfinite(p))
    # the jump sum is stored for both components
    assert np.array_equal(jp[1:, 0], jp[1:, 1])
    q = 1.0 - math.exp(-job.model.rate / N)
    hits = int((jp[1:, 0] != 0.0).sum())
    n = N * P
    assert abs(hits - n * q) < 5 * math.sqrt(n * q * (1 - q))


@pytest.mark.parametrize("P", [1, 63, 64, 65])
def test_rng_ragged_path_counts(P):
    job = J.job_for(J.merton_sde(), 1.0, 10, P, option=EuroCall(1.0), seed=5)
    pay, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert s[2] == P


def test_empty_job_gives_zero_stats():
    job = J.job_for(J.merton_sde(), 1.0, 10, 1, option=EuroCall(1.0))
    job.n_paths = 0
    s = E.simulate(job, _dev()).cpu().numpy()
    assert list(s) == [0.0, 0.0, 0.0, 0.0]


def test_batch_size_does_not_change_results():
    # paths are keyed by (seed, path index): two launches of 500 equal one of 1000
    job = J.job_for(J.merton_sde(), 1.0, 50, 1000, option=EuroCall(1.0), seed=9)
    a, ta, _, _ = _run_rng(job)
    j1 = J.job_for(J.merton_sde(), 1.0, 50, 500, option=EuroCall(1.0), seed=9, path_offset=0)
    j2 = J.job_for(J.merton_sde(), 1.0, 50, 500, option=EuroCall(1.0), seed=9, path_offset=500)
    b1, t1, _, _ = _run_rng(j1)
    b2, t2, _, _ = _run_rng(j2)
    assert np.array_equal(a, np.concatenate([b1, b2]))


def test_kernel_payoffs_match_reference_option_kats():
    # the reference's option known answers (tests/test_options.py:23-50)
    # through the device path, state frozen at the terminal
    from tests.jobs import frozen_state_job, payoff_kat_cases
    dev = E.require_device()
    for kind, K, x, want in payoff_kat_cases():
        s, q, n = E.stats_to_host(E.simulate(frozen_state_job(kind, K, x), dev))
        assert n == 1 and abs(s - want) < 1e-6, (kind, x, s, want)


@pytest.mark.parametrize("grid", ['uniform', 'adapted'])
def test_exact_jumps_match_oracle(grid):
    # post-step jump coefficient (solvers.py:214-215), fixed inputs and RNG mode
    P, N = 2048 + 3, 50
    job = J.job_for(J.merton_sde(rate=5.0), 1.0, N, P, grid=grid, option=EuroCall(1.0),
                    exact_jumps=True, seed=41)
    fx = J.fixed_adapted(job, seed=13) if grid == 'adapted' else J.fixed_uniform(job, seed=13)
    pay, term, ns, s = _run_fixed(job, fx)
    opay, oterm, ons = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    pay, term, ns, s = _run_rng(job)
    opay, oterm, ons = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL


def test_double_noise_matches_oracle():
    # two independent correlated drivers ('indep' diffusion, DoubleGbm)
    import sde_mc_amd as sm
    corr = torch.tensor([[1.0, 0.4], [0.4, 1.0]])
    sde = sm.DoubleGbm(0.03, 0.2, 0.3, torch.tensor([1.0, 1.1]), 2, corr_matrix=corr)
    P, N = 1500, 40
    job = J.job_for(sde, 1.0, N, P, option=sm.Basket(1.0), seed=43)
    fx = J.fixed_uniform(job, seed=14)
    pay, term, _, s = _run_fixed(job, fx)
    opay, oterm, _ = oracle.simulate_fixed(job, fx.host_struct(), P)
    assert _close(term, oterm) < TOL
    pay, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5


def test_put_and_geometric_basket_payoffs_match_oracle():
    from tests.jobs import frozen_state_job
    dev = E.require_device()
    for kind, K, x in (('EURO_PUT', 1.0, [0.4]), ('EURO_PUT', 1.0, [1.3]),
                       ('BASKET_GEO', 0.5, [2.0, 0.7]), ('BASKET_GEO', 0.5, [0.9, 1.4, 1.1])):
        job = frozen_state_job(kind, K, x)
        s, q, n = E.stats_to_host(E.simulate(job, dev))
        want = oracle.simulate(job, 0, 1)[0][0]
        assert n == 1 and abs(s - want) < 1e-6 * (1 + abs(want)), (kind, x, s, want)


def test_plain_launches_on_two_streams_keep_their_own_scratch():
    """Chunk partials live in a per-(device, stream) workspace: two jobs launched
    on two streams at once give their serial results."""
    dev = _dev()
    jobs = [J.job_for(J.merton_sde(), 1.0, 200, (1 << 18) + 7 * k, option=EuroCall(1.0),
                      seed=70 + k) for k in range(2)]
    serial = [E.simulate(j, dev).cpu().numpy() for j in jobs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    out = []
    for s, j in zip(streams, jobs):
        with torch.cuda.stream(s):
            out.append(E.simulate(j, dev))
    torch.cuda.synchronize()
    for o, r in zip(out, serial):
        assert np.array_equal(o.cpu().numpy(), r)


@pytest.mark.parametrize("model,exact", [('merton', False), ('merton', True), ('levy', False),
                                         ('levy', True)])
def test_rng_paths_match_oracle_at_high_jump_rates(model, exact):
    """Several jumps per lane and step: the kernels' jump queues overflow on
    most windows and fall back to in-step draws; the paths still follow the
    fp64 oracle on the same Philox counters."""
    P = 2000
    if model == 'merton':
        sde, opt, N = J.merton_sde(rate=30.0), EuroCall(1.0), 20      # lambda h = 1.5
    else:
        sde, opt, N = J.levy_rainbow_sde(eps=0.01), Rainbow(1.0, log=True), 50  # lambda h ~ 1.9
    job = J.job_for(sde, 1.0, N, P, option=opt, seed=881, exact_jumps=exact)
    pay, term, _, s = _run_rng(job)
    opay, oterm, _ = oracle.simulate(job, 0, P)
    assert _close(term, oterm) < TOL
    assert abs(s[0] / P - opay.mean()) < 1e-5 * max(1.0, abs(opay.mean()))

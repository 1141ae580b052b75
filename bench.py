"""Benchmark of the served / batched pricing path on one MI355X.

Workload: the served request mix (Merton 1-D calls and 2-D exponential-Levy
calls on the max; strike U[0.8,1.2], maturity {0.25,0.5,1,2}, steps
log-uniform [50,800], paths log-uniform [2^14,2^22]), R requests, synthetic.
One "step" prices the whole mix once in a batched persistent launch (job
table + tile list in HBM, graph-replayed) and reads every request's
(sum, sum of squares, n) back.  The plain form prices the same requests
one launch per request with a synchronising read-back, the per-job kernel of
sde_mc_amd.mc_simple.  Both produce bit-identical results per request, so
they are compared at equal confidence-interval half-widths.

Prints one JSON line (see the fields at the end of main()).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path-steps/s batched vs plain at equal CI; p50/p99 pricing latency"
PEAK_VALU_OPS = 78.6e12                   # fp32 vector FMA lane-ops/s (157.3 TFLOP/s / 2)
PEAK_FP32 = 157.3e12                      # MI355X fp32 vector peak, FLOP/s


def cpu_baseline(reqs, budget_s=6.0, cap_paths=1 << 12):
    """The reference's solver loop restated op for op in torch on the host
    cores (oracle/refcpu.py), over the mix's requests in order with each
    request's path count capped; path-steps/s of the sample (a second CPU
    figure beside the planned 2^16 x 100 Merton baseline, c0_baseline)."""
    import torch
    from oracle import refcpu
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    gen = torch.Generator().manual_seed(0)
    done_ps, t_total, n_req = 0, 0.0, 0
    for r in reqs:
        m = W.model_description(r.model)
        if r.model == 'merton':
            model = refcpu.merton_model(0.02, 0.2, 2.0, -0.1, 0.15)
        else:
            jp = list(m.jump_params)
            icdf = refcpu.TruncatedPowerIcdf(jp[0], jp[1], math.inf if jp[2] == 0 else 1 / jp[2],
                                             jp[3], jp[4])
            model = refcpu.AffineModel(m.x0, m.a0, m.a1, m.b0, m.b1, m.c0, m.c1, m.j0, m.j1,
                                       chol=m.cholesky(), noise='common', rate=m.rate, icdf=icdf)
        P = min(r.paths, cap_paths)
        t0 = time.perf_counter()
        refcpu.solve_adapted(model, r.maturity, r.steps, P,
                             E.safe_max_jumps(m.rate, r.maturity), gen=gen)
        t_total += time.perf_counter() - t0
        done_ps += P * r.steps
        n_req += 1
        if t_total > budget_s:
            break
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            cpu = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": done_ps / t_total, "unit": "path-steps/s", "cores": torch.get_num_threads(),
            "cpu": cpu, "kind": "port",
            "sample": "%d requests of the mix in order, paths capped at %d, reference solver "
                      "loop op for op in torch fp32 (%.1f s)" % (n_req, cap_paths, t_total)}


MERTON_CLOSED_FORM = 0.13205167640082938  # S0=K=1, r=0.02, sigma=0.2, lambda=2, N(-0.1, 0.15^2), T=1


def _median(v):
    v = sorted(v)
    n = len(v)
    return v[n // 2] if n % 2 else 0.5 * (v[n // 2 - 1] + v[n // 2])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            return next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown"


def c0_baseline(dev, runs=5, P=1 << 16, N=100, T=1.0):
    """The plain method on the CPU, as planned for the 1-D Merton call at 2^16
    paths x 100 steps, seed 0: the reference's jump-adapted solver loop with
    path storage and mc_simple's unbatched payoff (solvers.py:164-226,
    mc.py:87-100) restated op for op in torch fp32 (oracle/refcpu.py), on the
    host cores this process uses; 1 warm-up, median of `runs`.  Beside it the
    GPU's plain kernel on the same workload (same model, grid, paths and
    steps; its own Philox stream), median of `runs` event-timed launches."""
    import torch
    from scipy.stats import poisson
    from oracle import refcpu
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.options import EuroCall
    lam = 2.0
    max_jumps = max(int(T * poisson.ppf(1 - 1 / 1e9, lam)), 5)  # solvers.py:133
    model = refcpu.merton_model(0.02, 0.2, lam, -0.1, 0.15)
    disc = math.exp(-0.02 * T)
    payoff = lambda x: torch.clamp(x[:, 0] - 1.0, min=0.0)
    gen = torch.Generator().manual_seed(0)
    refcpu.price(model, payoff, disc, T, N, P, max_jumps, gen)          # warm-up
    times, prices = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        mean, se = refcpu.price(model, payoff, disc, T, N, P, max_jumps, gen)
        times.append(time.perf_counter() - t0)
        prices.append((mean, se))
    t = _median(times)
    mean, se = prices[times.index(t)] if t in times else prices[0]
    cpu = {"value": P * N / t, "unit": "path-steps/s", "cores": torch.get_num_threads(),
           "affinity_cpus": len(os.sched_getaffinity(0)), "cpu": _cpu_model(), "kind": "port",
           "sample": "the 1-D Merton call (S0=K=1, r=0.02, sigma=0.2, lambda=2, log-jump "
                     "N(-0.1, 0.15^2), T=1) at %d paths x %d steps, seed 0: the reference's "
                     "jump-adapted loop op for op in torch fp32 with path storage, 1 warm-up, "
                     "median of %d" % (P, N, runs),
           "runs_s": times, "price": mean, "half_width": 1.96 * se,
           "error_vs_closed_form": mean - MERTON_CLOSED_FORM}
    gpu = {}
    for grid, code in (("adapted", L.GRID_ADAPTED), ("uniform", L.GRID_UNIFORM)):
        job = E.make_job(W.model_description('merton'), T, N, code, P, 0, 0,
                         max_steps=N + E.safe_max_jumps(lam, T))
        E.fill_payoff(job, EuroCall(1.0), disc)
        stats = E.simulate(job, dev)
        torch.cuda.synchronize()
        ms = _median([_time_events(lambda: E.simulate(job, dev, stats=stats), 1)
                      for _ in range(runs)])
        s, q, n = E.stats_to_host(stats, 'C0 on the GPU')
        m, e = E.mean_and_se(s, q, n)
        gpu[grid] = {"path_steps_s": P * N / (ms * 1e-3), "kernel_ms": ms, "price": m,
                     "half_width": 1.96 * e, "error_vs_closed_form": m - MERTON_CLOSED_FORM}
    cpu["gpu_plain_same_workload"] = gpu
    cpu["gpu_over_cpu_adapted"] = gpu["adapted"]["path_steps_s"] / cpu["value"]
    return cpu


def _sample_job(job, n):
    """The first n paths of `job` (same seed and path offsets, so the same
    Philox streams as those paths in the full job)."""
    import ctypes
    from sde_mc_amd import _lib as L
    j = L.Job.from_buffer_copy(job)
    j.n_paths = min(int(n), int(job.n_paths))
    return j


def _fp64_error(dev, jobs, n):
    """fp32 kernel against the fp64 oracle on the first n paths of each job:
    max and mean |x32 - x64| / (1 + |x64|) over terminal states, and the
    difference of the sample means of the discounted payoffs."""
    import numpy as np
    import torch
    import oracle
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    rel_max, rel_sum, cnt, dprice = 0.0, 0.0, 0, []
    for job in jobs:
        j = _sample_job(job, n)
        P = int(j.n_paths)
        lv = 2 if j.coarse_factor else 1
        d = j.model.dim
        term = torch.empty((lv, d, P), dtype=torch.float32, device=dev)
        pay = torch.empty(P, dtype=torch.float32, device=dev)
        out = L.PathOut()
        out.terminal_dev, out.payoff_dev = term.data_ptr(), pay.data_ptr()
        E.stats_to_host(E.simulate(j, dev, out=out), 'fp64 error sample')
        opay, oterm, _ = oracle.simulate(j, 0, P)
        t32 = term.cpu().numpy().astype(np.float64)
        err = np.abs(t32 - oterm) / (1.0 + np.abs(oterm))
        rel_max = max(rel_max, float(err.max()))
        rel_sum += float(err.sum())
        cnt += err.size
        dprice.append(abs(float(pay.cpu().double().mean()) - float(opay.mean())))
    return {"jobs": len(jobs), "paths_per_job": int(n), "max_rel_state_err": rel_max,
            "mean_rel_state_err": rel_sum / max(cnt, 1), "max_price_diff": max(dprice),
            "mean_price_diff": float(np.mean(dprice))}


def fp64_error_statement(dev, reqs, mlmc_jobs=None, n=256):
    """The error of each result against fp64, stated per leg: the served
    mix (every request), the MLMC levels and the fixed-increment mode,
    recomputed on a sample of paths by the fp64 oracle (the checker; outside
    every timed region)."""
    from sde_mc_amd import workloads as W
    out = {"served_mix": _fp64_error(dev, [W.request_job(r) for r in reqs], n)}
    if mlmc_jobs:
        out["mlmc"] = _fp64_error(dev, mlmc_jobs, n)
    out["fixed_increment"] = _fixed_fp64_error(dev)
    out["note"] = ("states: max/mean over the sample of |x32 - x64| / (1 + |x64|); prices: "
                   "|mean of the fp32 kernel's payoffs - mean of the fp64 oracle's| on the "
                   "same paths; the job sums themselves are fp64")
    return out


def _fixed_fp64_error(dev, P=4096, N=500, seed=1):
    """Fixed-increment mode against the fp64 oracle on the same host inputs."""
    import numpy as np
    import torch
    import oracle
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.options import EuroCall
    rng = np.random.default_rng(seed)
    h = 1.0 / N
    dw = (rng.standard_normal((N, 1, P)) * math.sqrt(h)).astype(np.float32)
    counts = rng.poisson(2.0 * h, size=(N, P)).astype(np.uint8)
    off = np.zeros(P + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts.sum(axis=0))
    js = (np.exp(-0.1 + 0.15 * rng.standard_normal(max(int(off[-1]), 1))) - 1).astype(np.float32)
    job = E.make_job(W.model_description('merton'), 1.0, N, L.GRID_UNIFORM, P, 0, 0)
    E.fill_payoff(job, EuroCall(1.0), math.exp(-0.02))
    arrays = dict(dw=dw, counts=counts, off=off, js=js)

    def fixed_in(ptrs):
        f = L.FixedIn()
        f.dw_dev, f.counts_dev, f.jump_off_dev = ptrs['dw'], ptrs['counts'], ptrs['off']
        f.jump_time_dev, f.jump_size_dev = None, ptrs['js']
        f.n_marks, f.n_rows, f.max_jumps = int(js.size), N, 0
        return f
    dev_t = {k: torch.from_numpy(v).to(dev) for k, v in arrays.items()}
    term = torch.empty((1, 1, P), dtype=torch.float32, device=dev)
    pay = torch.empty(P, dtype=torch.float32, device=dev)
    out = L.PathOut()
    out.terminal_dev, out.payoff_dev = term.data_ptr(), pay.data_ptr()
    E.stats_to_host(E.simulate_fixed(job, fixed_in({k: t.data_ptr() for k, t in dev_t.items()}),
                                     dev, out=out), 'fixed fp64 error sample')
    opay, oterm, _ = oracle.simulate_fixed(job, fixed_in({k: v.ctypes.data for k, v in arrays.items()}), P)
    err = np.abs(term.cpu().numpy().astype(np.float64) - oterm) / (1.0 + np.abs(oterm))
    return {"jobs": 1, "paths_per_job": P, "max_rel_state_err": float(err.max()),
            "mean_rel_state_err": float(err.mean()),
            "max_price_diff": abs(float(pay.cpu().double().mean()) - float(opay.mean())),
            "mean_price_diff": abs(float(pay.cpu().double().mean()) - float(opay.mean()))}


def valu_issue_ceiling(dev, iters=8192, reps=3):
    """Measured VALU issue rates: back-to-back instructions of one kind from 8
    resident wavefronts per SIMD on every CU (sdemc_valu_probe_dev), wavefront
    instructions per second, best of `reps` launches each: 0 (v_fma_f32),
    1 (v_xor_b32 + v_add_u32) and 2 (v_mad_u64_u32 + v_xor_b32) check the
    issue ceiling of one wavefront instruction per CU-cycle that the path
    kernels' VALU counts are priced at; 3 (v_exp_f32 + 2 v_mul_f32) prices
    the transcendentals; 4-6 time a Philox block and the normals of one."""
    import ctypes
    import torch
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    blocks = cus * 8
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    lib = L.lib()
    clock_khz, cu, ws = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    lib.sdemc_device_info(torch.cuda.current_device(), ctypes.byref(cu), ctypes.byref(clock_khz),
                          ctypes.byref(ws))
    per_iter = {0: 32, 1: 64, 2: 64, 3: 96}   # VALU instructions per loop iteration
    names = {0: "v_fma_f32", 1: "v_xor_b32 + v_add_u32", 2: "v_mad_u64_u32 + v_xor_b32",
             3: "v_exp_f32 + 2 v_mul_f32", 4: "philox4x32_10 block",
             5: "philox4x32_10 block + 3 box_muller pairs (6 normals)",
             6: "2 independent philox4x32_10 blocks"}
    rec = {"clock_mhz": clock_khz.value / 1e3, "cus": cus,
           "probe": "sdemc_valu_probe_dev, %d workgroups of 4 wavefronts, %d iterations" % (blocks, iters)}
    waves = blocks * 4
    for kind in (0, 1, 2, 3, 4, 5, 6):
        it = iters if kind < 4 else iters // 4

        def run():
            L.check(lib.sdemc_valu_probe_dev(ctypes.c_void_p(out.data_ptr()), blocks, it, kind,
                                             E.stream_ptr()), 'sdemc_valu_probe_dev')
        run()
        torch.cuda.synchronize()
        ms = min(_time_events(run, 1) for _ in range(reps))
        cycles = ms * 1e-3 * clock_khz.value * 1e3
        if kind < 4:
            rate = waves * per_iter[kind] * it / (ms * 1e-3)
            rec[names[kind]] = {"wave_instr_per_s": rate, "kernel_ms": ms,
                                "wave_instr_per_cu_cycle": rate / (cus * clock_khz.value * 1e3)}
        else:
            # SIMD cycles one wavefront's iteration occupies (4 SIMDs per CU)
            rec[names[kind]] = {"kernel_ms": ms,
                                "simd_cycles_per_wave_iter": cycles * cus * 4 / (waves * it)}
    rec["wave_instr_per_s"] = rec[names[0]]["wave_instr_per_s"]
    return rec


def hbm_read_ceiling(dev, gib=4, reps=5):
    """Measured streaming-read rate of the device: sdemc_hbm_read_probe_dev
    over a `gib` GiB buffer (16 times the last-level cache), 16-byte loads at
    a grid stride, 4 or 8 workgroups of 256 threads per CU, best of `reps`
    launches each.  The fixed-increment kernel's bandwidth is stated against
    this as well as against the 8 TB/s datasheet figure."""
    import ctypes
    import torch
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    nbytes = gib << 30
    src = torch.ones(nbytes, dtype=torch.uint8, device=dev)
    out = torch.empty(cus * 8 * 256, dtype=torch.int32, device=dev)
    lib = L.lib()
    best = None
    for bpc in (4, 8):
        def run():
            L.check(lib.sdemc_hbm_read_probe_dev(ctypes.c_void_p(src.data_ptr()), nbytes,
                                                 ctypes.c_void_p(out.data_ptr()), cus * bpc,
                                                 E.stream_ptr()), 'sdemc_hbm_read_probe_dev')
        run()
        torch.cuda.synchronize()
        ms = min(_time_events(run, 1) for _ in range(reps))
        if best is None or ms < best[0]:
            best = (ms, bpc)
    del src, out
    torch.cuda.empty_cache()
    return {"gbs": nbytes / (best[0] * 1e-3) / 1e9, "bytes": nbytes, "kernel_ms": best[0],
            "blocks_per_cu": best[1],
            "probe": "sdemc_hbm_read_probe_dev: 16-byte loads at a grid stride, XOR-folded"}


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE,
# each scaled by the counter's reading on a 1 GiB device copy (fetch x2.0,
# write x1.0); profiles/hbm_traffic_served_mix_and_fixed_merton.md holds the
# runs.  Served mix round (k_batch group kernels, 128 requests, seed 0),
# keyed by the launch shape's chunks per tile: nearly all of it is the chunk
# partials' writes, which scatter more with one chunk per tile; and the
# fixed-increment kernel (2^21 x 500).
TRAFFIC_MIX_ROUND = {1: 9.504e7, 2: 7.103e7}  # tile_chunks -> HBM bytes of one mix round
TRAFFIC_FIXED = 5.711e9      # HBM bytes of one fixed-increment launch (2^21 x 500)
# VALU instructions (wavefront-level, SQ_INSTS_VALU summed over the group
# kernels) of one round of the served mix (128 requests, seed 0) and the fp32
# transcendentals among them (SQ_INSTS_VALU_TRANS_F32), from rocprofv3 --pmc;
# profiles/k_batch_served_mix_128req_sq_counters.md.  A wave64 VALU
# instruction holds its SIMD for 4 cycles, one per CU-cycle over the 4 SIMDs;
# the transcendentals overlap the instructions after them, so the issue
# figure counts the others.
VALU_INSTR_MIX_ROUND = 2.678e10
VALU_TRANS_MIX_ROUND = 2.031e9


def _time_events(fn, reps):
    """Mean device time (ms) of fn() over reps, bracketed by events on the
    current stream (the stream the library launches on)."""
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / reps


def fixed_increment_leg(dev, reps, P=1 << 21, N=500, seed=0):
    """1-D Merton call on the uniform grid fed from HBM (the parity mode at a
    size well past the 256 MB last-level cache): dw [N][1][P] fp32 N(0, h),
    counts [N][P] uint8 Poisson(lambda h), log-jump marks exp(N(-0.1, 0.15^2))-1
    in path-major CSR order.  Inputs are drawn on the device (synthetic)."""
    import torch
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.options import EuroCall
    T = 1.0
    h = T / N
    g = torch.Generator(device=dev).manual_seed(seed)
    dw = torch.randn((N, 1, P), device=dev, generator=g).mul_(math.sqrt(h))
    counts = torch.poisson(torch.full((N, P), 2.0 * h, device=dev), generator=g).to(torch.uint8)
    per_path = counts.sum(dim=0, dtype=torch.int64)
    off = torch.zeros(P + 1, dtype=torch.int64, device=dev)
    off[1:] = torch.cumsum(per_path, 0)
    n_marks = int(off[-1].item())
    js = torch.randn(max(n_marks, 1), device=dev, generator=g).mul_(0.15).add_(-0.1).exp_().sub_(1.0)
    fin = L.FixedIn()
    fin.dw_dev, fin.counts_dev, fin.jump_off_dev = dw.data_ptr(), counts.data_ptr(), off.data_ptr()
    fin.jump_time_dev, fin.jump_size_dev = None, js.data_ptr()
    fin.n_marks, fin.n_rows, fin.max_jumps = n_marks, N, 0
    job = E.make_job(W.model_description('merton'), T, N, L.GRID_UNIFORM, P, 0, 0)
    E.fill_payoff(job, EuroCall(1.0), math.exp(-0.02 * T))
    stats = torch.empty(4, dtype=torch.float64, device=dev)
    E.simulate_fixed(job, fin, dev, stats=stats)
    torch.cuda.synchronize()
    ms = _time_events(lambda: E.simulate_fixed(job, fin, dev, stats=stats), reps)
    s, q, n = E.stats_to_host(stats, 'fixed-increment bench')
    mean, se = E.mean_and_se(s, q, n)
    nbytes = 4 * N * P + N * P + 8 * P + 4 * n_marks + 16 * ((P + 63) // 64)
    return {"workload": "Merton 1-D call, fixed increments from HBM, %d paths x %d steps" % (P, N),
            "path_steps_s": P * N / (ms * 1e-3), "kernel_ms": ms,
            "bytes_per_launch": nbytes, "bytes_per_path_step": nbytes / (P * N),
            "roofline": {"bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": 8000.0,
                         "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / 8e12,
                         "traffic": TRAFFIC_FIXED if P == 1 << 21 and N == 500 else None,
                         "kernel": "k_fixed_uniform4<1, merton> + k_finalize_one"},
            "price": mean, "se": se, "closed_form": MERTON_CLOSED_FORM}


def equal_ci_leg(reqs, tune, dev, eps=1e-3, pilot=1 << 13, reps=3):
    """Every request of the mix priced to a 95% half-width <= eps: a pilot run
    per request (own Philox stream) sets the path count n = 1.2 (1.96 sd / eps)^2,
    the run_mc rule with a 20% margin for the pilot's error in sd, capped to
    [2^10, 2^22]; the same jobs are then priced
    batched and plain (bit-identical results, so equal confidence intervals).
    Timed twice: the main runs alone, and the whole run_mc workflow (pilots,
    then main runs; mc.py:418-440), whose pilots are 128 small jobs."""
    import numpy as np
    import torch
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.batch import BatchRunner, chunks_of
    pil = [W.request_job(W.Request(r.rid, r.model, r.strike, r.maturity, r.steps, pilot,
                                   r.seed ^ 0x5DEECE66D)) for r in reqs]
    prunner = BatchRunner(max_jobs=len(reqs), partial_chunks=sum(chunks_of(pilot) for _ in reqs),
                          tune=tune)
    st = prunner.run(pil)
    pref = st.copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = prunner.run(pil)
    tpb = (time.perf_counter() - t0) / reps
    for j in pil[:4]:  # warm-up of the plain launches
        E.simulate(j, dev).cpu()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        pp = [E.simulate(j, dev).cpu().numpy() for j in pil]
    tpp = (time.perf_counter() - t0) / reps
    assert np.array_equal(np.array(pp), pref), "plain and batched pilots differ"
    st = pref
    sd = np.sqrt(np.maximum((st[:, 1] - st[:, 0] ** 2 / st[:, 2]) / (st[:, 2] - 1), 0.0))
    n = np.clip(np.ceil(1.2 * (1.96 * sd / eps) ** 2), 1 << 10, 1 << 22).astype(np.int64)
    sized = [W.Request(r.rid, r.model, r.strike, r.maturity, r.steps, int(k), r.seed)
             for r, k in zip(reqs, n)]
    jobs = [W.request_job(r) for r in sized]
    ps = sum(r.path_steps for r in sized)
    runner = BatchRunner(max_jobs=len(jobs), partial_chunks=int(sum(chunks_of(k) for k in n)),
                         tune=tune)
    ref = runner.run(jobs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        st = runner.run(jobs)
    tb = (time.perf_counter() - t0) / reps
    assert np.array_equal(st, ref)
    for j in jobs[:4]:
        E.stats_to_host(E.simulate(j, dev))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        pst = [E.simulate(j, dev).cpu().numpy() for j in jobs]
    tp = (time.perf_counter() - t0) / reps
    assert np.array_equal(np.array(pst), st), "plain and batched results differ"
    hw = []
    for row in st:
        m, se = E.mean_and_se(row[0], row[1], int(row[2]))
        hw.append(1.96 * se)
    return {"eps": eps, "requests": len(jobs), "paths_min": int(n.min()), "paths_max": int(n.max()),
            "path_steps": ps, "batched_path_steps_s": ps / tb, "plain_path_steps_s": ps / tp,
            "speedup_vs_plain": tp / tb,
            "pilot_ms": {"batched": tpb * 1e3, "plain": tpp * 1e3},
            "workflow_speedup_vs_plain": (tp + tpp) / (tb + tpb),
            "max_half_width": float(max(hw)),
            "requests_over_eps": int(sum(h > eps for h in hw)),
            "note": "requests capped at 2^22 paths may stay above eps"}


def mlmc_leg(dev, reps=21):
    """MLMC on the 1-D Merton call: levels l=0..6, refinement 2, base 4 steps,
    P_l = 2^(22 - 10 l / 6) paths, coupled fine/coarse pairs for l >= 1, all
    levels in one batched launch (launch shape autotuned on the level set,
    a batch of other classes than the served mix) vs one plain launch per
    level.  Path-steps count fine steps only.  Returns (record, level jobs)."""
    import numpy as np
    import torch
    from sde_mc_amd import _lib as L
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.batch import BatchRunner, chunks_of
    from sde_mc_amd.options import EuroCall
    model = W.model_description('merton')
    T = 1.0
    jobs, ps = [], 0
    for l in range(7):
        P = int(round(2 ** (22 - 10 * l / 6)))
        N = 4 * 2 ** l
        job = E.make_job(model, T, N, L.GRID_UNIFORM, P, 7000 + l, 0, coarse_factor=2 if l else 0)
        E.fill_payoff(job, EuroCall(1.0), math.exp(-0.02 * T))
        jobs.append(job)
        ps += P * N
    from sde_mc_amd.autotune import autotune
    # the level set is a 0.5 ms batch: the sweep's fastest shapes are timed
    # again in interleaved rounds (and each with its class groups in turn),
    # so one noisy launch does not pick the shape
    tune, _ = autotune(jobs, confirm_jobs=jobs)
    runner = BatchRunner(max_jobs=8, partial_chunks=sum(chunks_of(j.n_paths) for j in jobs),
                         tune=tune)
    ref = runner.run(jobs)
    torch.cuda.synchronize()
    from sde_mc_amd.serve import no_cyclic_gc
    # a pass is 0.5 ms: the median of `reps` passes, each timed on its own,
    # so one stalled pass does not set the figure
    tbs, tps = [], []
    with no_cyclic_gc():
        for _ in range(reps):
            t0 = time.perf_counter()
            st = runner.run(jobs)
            tbs.append(time.perf_counter() - t0)
    tb = float(np.median(tbs))
    assert np.array_equal(st, ref)
    for j in jobs:  # warm-up of the plain launches (first-use costs of each kernel)
        E.simulate(j, dev).cpu()
    torch.cuda.synchronize()
    with no_cyclic_gc():
        for _ in range(reps):
            t0 = time.perf_counter()
            pst = [E.simulate(j, dev).cpu().numpy() for j in jobs]
            tps.append(time.perf_counter() - t0)
    tp = float(np.median(tps))
    assert np.array_equal(np.array(pst), st), "plain and batched MLMC levels differ"
    est, var = 0.0, 0.0
    for row in st:
        m, se = E.mean_and_se(row[0], row[1], int(row[2]))
        est += m
        var += se * se
    return {"levels": 7, "paths": [int(j.n_paths) for j in jobs], "path_steps": ps,
            "batched_path_steps_s": ps / tb, "plain_path_steps_s": ps / tp,
            "speedup_vs_plain": tp / tb, "batched_ms": tb * 1e3, "plain_ms": tp * 1e3,
            "timing": "median of %d passes of each form, host wall clock per pass "
                      "(launches, read-back and host work)" % reps,
            "tune": dict(zip(("waves_per_block", "blocks_per_cu", "tile_chunks",
                              "concurrent_groups", "step_unroll"), tune.key())),
            "estimate": est, "rmse_se": math.sqrt(var), "closed_form": MERTON_CLOSED_FORM,
            "error": est - MERTON_CLOSED_FORM}, jobs


def _mlp_flops(layers):
    """Useful flops of one evaluation of an MLP with these widths (biases included)."""
    return sum(2 * a * b + b for a, b in zip(layers[:-1], layers[1:]))


def cv_leg(dev, reps=3, P=1 << 20, N=1000, P_torch=32000):
    """Neural control variates on the reference's Merton experiment
    (examples/merton_1d_european/merton_cv_experiment.py): Merton 1-D call,
    T=3, N steps on the jump-adapted grid, f and g = MLP 2-50-50-50-1
    (random weights; the cost does not depend on them), P paths per fused
    launch.  Compared with the reference's route on the same device: paths
    stored by the HIP kernel, then apply_adapted_control_variates in torch
    (fp32 GEMMs) in network batches of 2000, P_torch paths."""
    import torch
    import sde_mc_amd as sm
    from sde_mc_amd import cv as CV
    from sde_mc_amd import engine as E
    from sde_mc_amd import mc as MC
    from sde_mc_amd import _lib as L
    p = sm.MertonEuroCall.default_params(N, 'cuda')
    torch.manual_seed(0)
    nets = [sm.Mlp(2, [50, 50, 50], 1, batch_norm=False, batch_norm_init=False,
                   device='cuda').eval() for _ in range(2)]
    fused = MC.fused_control_variates(nets, p.solver, p.payoff, p.discounter)
    disc = float(p.discounter(p.solver.time_interval))
    job, _ = p.solver.make_job(P, option=p.payoff, discount=disc)
    ns = torch.empty(P, dtype=torch.int32, device=dev)
    out = L.PathOut()
    out.nsteps_dev = ns.data_ptr()
    plain_stats = E.simulate(job, dev, out=out)
    plain_ms = _time_events(lambda: E.simulate(job, dev, stats=plain_stats), reps)
    path_steps = int(ns.to(torch.int64).sum())
    steps_max = job.max_steps

    def run():
        return CV.cv_simulate(job, fused.f, fused.g, steps_max, fused.rate, fused.comp, dev)[0]
    st = run()
    torch.cuda.synchronize()
    ms = _time_events(run, reps)
    s, q, n = E.stats_to_host(st, 'cv bench')
    mean, se = E.mean_and_se(s, q, n)
    flops = path_steps * 2 * _mlp_flops([2, 50, 50, 50, 1])
    mfma_instr = 16 * 3 - 4                      # 32x32x16 products per wave per evaluation
    waves = (P + 63) // 64
    # executed products: every row of the wave's longest path, both networks
    # (rows ~ path steps here: the steps per path vary by a few jumps)
    exec_flops = 2 * mfma_instr * 32 * 32 * 16 * 2 * (path_steps / P) * waves
    # reference route: store, then torch
    p2 = sm.MertonEuroCall.default_params(N, 'cuda')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dl = MC.simulate_adapted_data(P_torch, p2.solver, p2.payoff, p2.discounter, bs=2000,
                                  inference=True)
    x, y = sm.apply_adapted_control_variates(nets, dl, p2.solver, p2.discounter)
    float(x)
    torch.cuda.synchronize()
    dt_torch = time.perf_counter() - t0
    torch_ps = P_torch * (path_steps / P) / dt_torch
    fused_ps = path_steps / (ms * 1e-3)
    # diffusion form: Black-Scholes call, f only, uniform grid
    b = sm.BlackScholesEuroCall.default_params(N, 'cuda')
    fb = MC.fused_control_variates(nets[0], b.solver, b.payoff, b.discounter)
    jb, _ = b.solver.make_job(P, option=b.payoff, discount=float(b.discounter(3.0)))
    bplain = E.simulate(jb, dev)
    bplain_ms = _time_events(lambda: E.simulate(jb, dev, stats=bplain), reps)

    def runb():
        return CV.cv_simulate(jb, fb.f, None, N, fb.rate, 0.0, dev)[0]
    runb()
    bms = _time_events(runb, reps)
    bflops = P * N * _mlp_flops([2, 50, 50, 50, 1])
    diffusion = {"workload": "Black-Scholes call T=3, %d steps, uniform grid, f MLP 2-50-50-50-1, "
                             "%d paths" % (N, P),
                 "path_steps_s": P * N / (bms * 1e-3), "kernel_ms": bms,
                 "plain_kernel_ms": bplain_ms, "cv_overhead_vs_plain": bms / bplain_ms,
                 "useful_tflops": bflops / (bms * 1e-3) / 1e12,
                 "executed_tflops": mfma_instr * 32 * 32 * 16 * 2 * N * waves / (bms * 1e-3) / 1e12,
                 "kernel": "k_cv_uniform<1,merton>"}
    return {"workload": "Merton 1-D call T=3, %d steps, jump-adapted grid, f and g MLP "
                        "2-50-50-50-1 (bf16 MFMA, fp32 accumulation), %d paths" % (N, P),
            "path_steps_s": fused_ps, "kernel_ms": ms, "path_steps": path_steps,
            "plain_kernel_ms": plain_ms, "cv_overhead_vs_plain": ms / plain_ms,
            "roofline": {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12,
                         "peak": 2500.0, "unit": "TFLOP/s (bf16, useful flops)",
                         "frac": flops / (ms * 1e-3) / 2.5e15,
                         "executed_tflops": exec_flops / (ms * 1e-3) / 1e12,
                         "kernel": "k_cv_adapted<1,merton>"},
            "torch_route_path_steps_s": torch_ps,
            "speedup_vs_torch_route": fused_ps / torch_ps,
            "price": mean, "se": se, "diffusion": diffusion}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--requests", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the fixed-increment, equal-CI and MLMC legs")
    ap.add_argument("--tune", default=None,
                    help="W,B,T[,G[,U]] launch shape (waves/block, blocks/CU, chunks/tile, "
                         "concurrent groups, step unroll) instead of autotuning")
    ap.add_argument("--fixed-only", action="store_true",
                    help="run only the fixed-increment leg and print its record")
    ap.add_argument("--cv-only", action="store_true",
                    help="run only the control-variate leg and print its record")
    ap.add_argument("--probe-copy-mib", type=int, default=0,
                    help="first run a device copy of this many MiB (calibrates HBM counters)")
    ap.add_argument("--valu-probe-only", action="store_true",
                    help="run only the VALU issue-rate probe and print its record")
    ap.add_argument("--c0-only", action="store_true",
                    help="run only the CPU baseline on the 2^16 x 100 Merton call (and the GPU "
                         "plain kernel beside it) and print its record")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X per process and "
              "has no multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from sde_mc_amd import engine as E
    from sde_mc_amd import workloads as W
    from sde_mc_amd.batch import BatchRunner, chunks_of

    dev = E.require_device()
    if args.probe_copy_mib:
        src = torch.ones(args.probe_copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.fixed_only:
        fixed = fixed_increment_leg(dev, args.steps)
        fixed["read_ceiling"] = ceil = hbm_read_ceiling(dev)
        fixed["roofline"]["frac_of_read_ceiling"] = fixed["roofline"]["achieved"] / ceil["gbs"]
        print(json.dumps({"fixed_increment": fixed}))
        return
    if args.cv_only:
        print(json.dumps({"control_variates": cv_leg(dev, args.steps)}))
        return
    if args.valu_probe_only:
        print(json.dumps({"valu_issue_ceiling": valu_issue_ceiling(dev)}))
        return
    if args.c0_only:
        print(json.dumps({"cpu_baseline": c0_baseline(dev)}))
        return
    reqs = W.request_mix(args.requests, seed=args.seed)
    jobs = [W.request_job(r) for r in reqs]
    total_ps = sum(r.path_steps for r in reqs)
    ops = sum(r.path_steps * W.valu_ops_per_path_step(r.model, r.steps, r.maturity) for r in reqs)
    flops = sum(r.path_steps * W.fp32_flops_per_path_step(r.model, r.steps, r.maturity)
                for r in reqs)
    chunks = sum(chunks_of(r.paths) for r in reqs)
    # autotune the persistent launch shape on a calibration copy of the mix
    # (same requests, path counts capped at 2^18)
    from sde_mc_amd.autotune import autotune
    from sde_mc_amd.batch import BatchTune
    if args.tune:
        tune = BatchTune(*[int(v) for v in args.tune.split(',')])
    else:
        calib = [W.request_job(W.Request(r.rid, r.model, r.strike, r.maturity, r.steps,
                                         min(r.paths, 1 << 18), r.seed)) for r in reqs]
        tune, _ = autotune(calib, confirm_jobs=jobs)
    runner = BatchRunner(max_jobs=len(jobs), partial_chunks=chunks, tune=tune)

    ref = None
    for _ in range(args.warmup):
        ref = runner.run(jobs)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    from sde_mc_amd.serve import no_cyclic_gc
    with no_cyclic_gc():  # both timed loops run without collector pauses
        t0 = time.perf_counter()
        for k in range(args.steps):
            st = runner.run(jobs, events=ev[k])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    batched = total_ps * args.steps / dt
    assert ref is None or np.array_equal(st, ref)

    plain = None
    if not args.no_plain:
        for j in jobs[:4]:
            E.stats_to_host(E.simulate(j, dev))
        torch.cuda.synchronize()
        with no_cyclic_gc():
            t0 = time.perf_counter()
            pst = []
            for j in jobs:
                s = E.simulate(j, dev)
                pst.append(s.cpu().numpy())
            dtp = time.perf_counter() - t0
        plain = total_ps / dtp
        assert np.array_equal(np.array(pst), st), "plain and batched results differ"

    latency = None
    if not args.no_serve:
        from sde_mc_amd.batch import job_class
        from sde_mc_amd.serve import PricingServer, latency_summary, serve_plain
        latency = {}
        # 50, 200, 1000 req/s: the workload's rates; 1500, 2000 and 2500 req/s
        # offer 3.0e11, 4.4e11 and 5.0e11 path-steps/s: up to and past what
        # the plain per-request form (kernel, read-back and host work per
        # request) finishes, and up to what the batched rounds can
        for rate in (50, 200, 1000, 1500, 2000, 2500):
            n_req = args.requests * (4 if rate >= 1500 else 1)   # long enough to queue up
            arr = W.request_mix(n_req, seed=args.seed + rate, rate=rate)
            srv = PricingServer(max_slots=256, hbm_budget_bytes=512 << 20, tune=tune)
            # server start-up: the round graphs of the stream's kernel classes
            srv.warm({job_class(W.request_job(r)) for r in arr})
            res = srv.serve(arr)
            served = latency_summary(res)
            served["rounds"] = srv.rounds
            served["peak_partial_bytes"] = srv.peak_used * 20
            arr2 = W.request_mix(n_req, seed=args.seed + rate, rate=rate)
            plain_res = serve_plain(arr2)
            same = all(np.array_equal(np.array(res[i].stats[:3]), np.array(plain_res[i].stats[:3]))
                       for i in res)
            latency[str(rate)] = {"served": served, "plain": latency_summary(plain_res),
                                  "results_identical": bool(same)}

    extra = {}
    mlmc_jobs = None
    if not args.no_extra:
        extra["fixed_increment"] = fixed = fixed_increment_leg(dev, args.steps)
        ceil = hbm_read_ceiling(dev)
        fixed["read_ceiling"] = ceil
        fixed["roofline"]["frac_of_read_ceiling"] = fixed["roofline"]["achieved"] / ceil["gbs"]
        extra["equal_ci"] = equal_ci_leg(reqs, tune, dev)
        extra["mlmc"], mlmc_jobs = mlmc_leg(dev)
        extra["control_variates"] = cv_leg(dev)

    ceiling = valu_issue_ceiling(dev)
    cpu = None
    if not args.no_cpu:
        cpu = c0_baseline(dev)
        cpu["mix_sample"] = cpu_baseline(reqs)
        extra["fp64_error"] = fp64_error_statement(dev, reqs, mlmc_jobs)

    kern_s = kern_ms * 1e-3
    achieved = flops / kern_s
    mix_default = (args.requests, args.seed) == (128, 0)
    valu_instr = (VALU_INSTR_MIX_ROUND - VALU_TRANS_MIX_ROUND) if mix_default else None
    issue_peak = ceiling["cus"] * ceiling["clock_mhz"] * 1e6  # wave instr / s
    out = {
        "metric": METRIC,
        "value": batched,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "served request mix, %d requests (Merton 1-D / Levy rainbow 2-D), "
                               "%.3g path-steps per pass" % (len(reqs), total_ps),
                   "stepping": "fp32", "sums": "fp64", "grid": "uniform",
                   "launch": "batched persistent kernel, graph-replayed",
                   "tune": {"waves_per_block": tune.waves_per_block,
                            "blocks_per_cu": tune.blocks_per_cu,
                            "tile_chunks": tune.tile_chunks,
                            "concurrent_groups": tune.concurrent_groups,
                            "step_unroll": tune.step_unroll}},
        "plain_value": plain,
        "speedup_vs_plain": (batched / plain) if plain else None,
        "latency": latency,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s (fp32 vector, algorithmic: 12 per 1-D Merton path-step, "
                             "35 per 2-D Levy path-step, + 10 per expected jump)",
                     "frac": achieved / PEAK_FP32,
                     "traffic": TRAFFIC_MIX_ROUND.get(tune.tile_chunks) if mix_default else None,
                     "kernel": "k_batch (persistent tile kernel)", "kernel_ms": kern_ms,
                     "valu_issue": {
                         "instr_per_round": valu_instr,
                         "achieved_wave_instr_per_s": valu_instr / kern_s if valu_instr else None,
                         "ceiling_wave_instr_per_s": issue_peak,
                         "frac": valu_instr / kern_s / issue_peak if valu_instr else None,
                         "source": "non-transcendental VALU instructions of a round "
                                   "(SQ_INSTS_VALU - SQ_INSTS_VALU_TRANS_F32, rocprofv3 "
                                   "--pmc) over the live round time, against one wavefront "
                                   "instruction per CU-cycle at the device clock (the "
                                   "integer probes of valu_issue_ceiling reach 0.99-1.01)"},
                     "lane_ops": {"achieved": ops / kern_s / 1e12, "peak": PEAK_VALU_OPS / 1e12,
                                  "unit": "Top/s (VALU lane-ops counted from the algorithm, "
                                          "Philox 20 per word)",
                                  "frac": ops / kern_s / PEAK_VALU_OPS}},
        "valu_issue_ceiling": ceiling,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

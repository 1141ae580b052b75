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
PEAK_FP32 = 157.3e12


def cpu_baseline(reqs, budget_s=12.0, cap_paths=1 << 12):
    """The reference's solver loop restated op for op in torch on the host
    cores (oracle/refcpu.py), over the mix's requests in order with each
    request's path count capped; path-steps/s of the sample."""
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
    return {"value": done_ps / t_total, "unit": "path-steps/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": "%d requests of the mix in order, paths capped at %d, reference solver "
                      "loop op for op in torch fp32 (%.1f s)" % (n_req, cap_paths, t_total)}


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
    reqs = W.request_mix(args.requests, seed=args.seed)
    jobs = [W.request_job(r) for r in reqs]
    total_ps = sum(r.path_steps for r in reqs)
    ops = sum(r.path_steps * W.valu_ops_per_path_step(r.model, r.steps, r.maturity) for r in reqs)
    chunks = sum(chunks_of(r.paths) for r in reqs)
    # autotune the persistent launch shape on a calibration copy of the mix
    # (same requests, path counts capped at 2^16)
    from sde_mc_amd.autotune import autotune
    calib = [W.request_job(W.Request(r.rid, r.model, r.strike, r.maturity, r.steps,
                                     min(r.paths, 1 << 16), r.seed)) for r in reqs]
    tune, _ = autotune(calib)
    runner = BatchRunner(max_jobs=len(jobs), partial_chunks=chunks, tune=tune)

    for _ in range(args.warmup):
        ref = runner.run(jobs)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = runner.run(jobs, events=ev[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    batched = total_ps * args.steps / dt
    assert np.array_equal(st, ref)

    plain = None
    if not args.no_plain:
        for j in jobs[:4]:
            E.stats_to_host(E.simulate(j, dev))
        torch.cuda.synchronize()
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
        from sde_mc_amd.serve import PricingServer, latency_summary, serve_plain
        latency = {}
        for rate in (50, 200, 1000):
            arr = W.request_mix(args.requests, seed=args.seed + rate, rate=rate)
            srv = PricingServer(max_slots=256, hbm_budget_bytes=512 << 20, tune=tune)
            res = srv.serve(arr)
            served = latency_summary(res)
            served["rounds"] = srv.rounds
            served["peak_partial_bytes"] = srv.peak_used * 20
            arr2 = W.request_mix(args.requests, seed=args.seed + rate, rate=rate)
            plain_res = serve_plain(arr2)
            same = all(np.array_equal(np.array(res[i].stats[:3]), np.array(plain_res[i].stats[:3]))
                       for i in res)
            latency[str(rate)] = {"served": served, "plain": latency_summary(plain_res),
                                  "results_identical": bool(same)}

    cpu = None if args.no_cpu else cpu_baseline(reqs)

    achieved = ops / (kern_ms * 1e-3)
    price_err = []
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
                            "tile_chunks": tune.tile_chunks}},
        "plain_value": plain,
        "speedup_vs_plain": (batched / plain) if plain else None,
        "latency": latency,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_VALU_OPS / 1e12,
                     "unit": "Top/s (VALU lane-ops, algorithmic)",
                     "frac": achieved / PEAK_VALU_OPS, "traffic": None,
                     "kernel": "k_batch (persistent tile kernel)", "kernel_ms": kern_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

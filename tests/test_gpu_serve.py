"""The continuous-batching server on the device: every request's result is the
same bits alone and in the mix, the HBM budget bounds admission, and a
Poisson-arrival stream completes with latencies reported."""
import math

import numpy as np
import pytest

from oracle import closed_form as cf
from sde_mc_amd import workloads as W
from sde_mc_amd.serve import PricingServer, latency_summary, serve_plain

pytestmark = pytest.mark.gpu


def _small_mix(n, seed, rate=None):
    # the served distributions with path counts scaled down so the test is short
    return W.request_mix(n, seed=seed, rate=rate, paths_range=(10, 17), steps_range=(20, 200))


@pytest.mark.parametrize("policy,plain_single", [('srpt', True), ('srpt', False),
                                                  ('fair', True)])
def test_served_results_equal_alone_and_plain(policy, plain_single):
    reqs = _small_mix(24, seed=7)
    mix = PricingServer(max_slots=32, hbm_budget_bytes=64 << 20, policy=policy,
                        plain_single=plain_single).serve(reqs, realtime=False)
    for r in reqs[:6]:
        alone = PricingServer(max_slots=4, hbm_budget_bytes=8 << 20, policy=policy,
                              plain_single=plain_single).serve([r], realtime=False)
        assert np.array_equal(np.array(alone[r.rid].stats[:3]), np.array(mix[r.rid].stats[:3]))
    plain = serve_plain(_small_mix(24, seed=7), realtime=False)
    for r in reqs:
        assert np.array_equal(np.array(mix[r.rid].stats[:3]), np.array(plain[r.rid].stats[:3]))


def test_budget_limits_admission():
    reqs = _small_mix(40, seed=8)
    budget = 2 << 20
    srv = PricingServer(max_slots=8, hbm_budget_bytes=budget)
    res = srv.serve(reqs, realtime=False)
    assert len(res) == 40
    assert srv.peak_used * 20 <= budget
    assert all(r.error is None or 'budget' in r.error for r in res.values())
    big = W.Request(999, 'merton', 1.0, 1.0, 50, 1 << 22, 5)
    res = PricingServer(max_slots=4, hbm_budget_bytes=1 << 20).serve([big], realtime=False)
    assert res[999].error is not None


def test_poisson_stream_completes_with_latencies():
    reqs = _small_mix(60, seed=9, rate=400)
    srv = PricingServer(max_slots=64, hbm_budget_bytes=64 << 20)
    res = srv.serve(reqs)
    s = latency_summary(res)
    assert s["requests"] == 60 and s["p99_ms"] >= s["p50_ms"] > 0
    assert srv.rounds >= 2
    # Merton requests price near the closed form
    for r in reqs:
        if r.model == 'merton' and r.paths >= 2 ** 15:
            want = cf.merton_call(1, r.strike, r.maturity, 0.02, 0.2, -0.1, 0.15, 2.0)
            assert abs(res[r.rid].mean - want) < 5 * res[r.rid].se + 2e-3

"""The product path has no CPU fallback: without a HIP device every pricing
entry of the drop-in API raises, the package never imports the oracle, and
bench.py refuses GPU counts other than one."""
import ast
import os
import subprocess
import sys

import pytest
import torch

import sde_mc_amd as sm
from sde_mc_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _no_device():
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")


def _merton():
    return sm.Merton(0.02, 0.3, 2, -0.05, 0.3, torch.tensor([1.]), dim=1)


def test_solve_raises_without_a_device():
    _no_device()
    with pytest.raises(L.HipUnavailableError):
        sm.JumpEulerSolver(_merton(), 1.0, 10).solve(8)
    with pytest.raises(L.HipUnavailableError):
        sm.EulerSolver(sm.Gbm(0.02, 0.2, torch.tensor([1.]), dim=1), 1.0, 10).solve(8)


def test_estimators_raise_without_a_device():
    _no_device()
    prob = sm.MertonEuroCall(0.02, 0.2, 2.0, -0.1, 0.15, 1.0, 1.0, 1.0, 20, 'cpu')
    with pytest.raises(L.HipUnavailableError):
        sm.mc_simple(100, prob.solver, prob.payoff, prob.discounter)
    with pytest.raises(L.HipUnavailableError):
        sm.mc_simple(100, prob.solver, prob.payoff, prob.discounter, bs=64)
    with pytest.raises(L.HipUnavailableError):
        sm.mc_multilevel([100, 50], [4, 8], prob.solver, prob.payoff, prob.discounter)


def test_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, 'sde_mc_amd')
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if not f.endswith('.py'):
                continue
            tree = ast.parse(open(os.path.join(dirpath, f)).read())
            for node in ast.walk(tree):
                if isinstance(node, ast.Import):
                    names = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom):
                    names = [node.module or '']
                else:
                    continue
                assert not any(n == 'oracle' or n.startswith('oracle.') for n in names), f


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and r.stdout == ''

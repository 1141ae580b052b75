"""Host-side checks of the control-variate path: network packing against the
kernel's lane data flow, BatchNorm folding, the fused-path eligibility
rules and the torch reference loops (no GPU)."""
import numpy as np
import pytest
import torch
import torch.nn as nn

from sde_mc_amd import cv as CV
from sde_mc_amd import mc as MC
from sde_mc_amd.nets import Gru, Lstm, Mlp, ZeroFunction
from sde_mc_amd.options import EuroCall
from sde_mc_amd.varred import EarlyStopping, integrate_cv
from tests import cv_emul


def _features(rng, dim, n=64):
    F = np.zeros((n, 8))
    F[:, 0] = rng.uniform(0, 3, n)
    F[:, 1:1 + dim] = rng.uniform(0.2, 2.5, (n, dim))
    F[:, 1 + dim] = 1.0
    return F


@pytest.mark.parametrize("layers,n_out,dim", [([50], 1, 1), ([50, 50, 50], 1, 1),
                                              ([20, 63], 4, 2), ([63, 63, 63, 63], 8, 4),
                                              ([7, 9, 11], 3, 3)])
def test_packing_matches_lane_data_flow(layers, n_out, dim):
    torch.manual_seed(len(layers) + n_out)
    m = Mlp(dim + 1, layers, n_out, batch_norm=False, batch_norm_init=False)
    words, hidden = CV.pack_network(m, dim + 1, n_out)
    assert hidden == len(layers)
    assert words.size == CV.BLOCK * CV.blocks(hidden)
    mats = CV.padded_matrices(m, dim + 1, n_out)
    F = _features(np.random.default_rng(7), dim)
    lane = cv_emul.lane_forward(words, hidden, F)
    mat = cv_emul.matrix_forward(mats, F)
    scale = np.abs(mat).max() + 1e-6
    # same bf16 roundings; only the fp64 summation order differs
    assert np.abs(lane - mat).max() <= 1e-9 * scale + 1e-12
    # and the bf16 network is close to the fp32 torch network
    ref = m(torch.tensor(F[:, :dim + 1], dtype=torch.float32)).detach().numpy()
    assert np.abs(lane[:, :n_out] - ref).max() <= 0.05 * np.abs(ref).max() + 1e-3
    assert np.all(lane[:, n_out:] == 0.0)


def test_wrong_unit_order_is_detected():
    """A packer that ignored the accumulator row order would fail the check above."""
    torch.manual_seed(3)
    m = Mlp(2, [40, 40], 1, batch_norm=False, batch_norm_init=False)
    mats = CV.padded_matrices(m, 2, 1)
    F = _features(np.random.default_rng(1), 1)
    good = cv_emul.lane_forward(CV.pack_matrices(mats), 2, F)
    saved = CV.unit_order
    try:
        CV.unit_order = lambda c, h, q: 16 * c + 8 * h + q
        bad = cv_emul.lane_forward(CV.pack_matrices(mats), 2, F)
    finally:
        CV.unit_order = saved
    assert np.abs(good - bad).max() > 1e-3


def test_bias_rides_on_constant_unit():
    """A network that is only biases gives its output bias everywhere."""
    m = Mlp(2, [5, 5], 1, batch_norm=False, batch_norm_init=False)
    with torch.no_grad():
        for mod in m.modules():
            if isinstance(mod, nn.Linear):
                mod.weight.zero_()
                mod.bias.fill_(0.25)
        m.net[-1].weight.fill_(1.0)
    # hidden 1: relu(0.25) = 0.25 per unit; hidden 2: relu(0.25) each; out = 5 * 0.25 + 0.25
    F = _features(np.random.default_rng(0), 1)
    out = cv_emul.lane_forward(*CV.pack_network(m, 2, 1), F)
    assert np.allclose(out[:, 0], 1.5)


def test_batchnorm_folding():
    torch.manual_seed(0)
    m = Mlp(3, [30, 30], 2, batch_norm=True, batch_norm_init=True)
    for mod in m.modules():
        if isinstance(mod, nn.BatchNorm1d):
            with torch.no_grad():
                mod.running_mean.uniform_(-1, 1)
                mod.running_var.uniform_(0.5, 2)
                mod.weight.uniform_(0.5, 1.5)
                mod.bias.uniform_(-0.5, 0.5)
    m.eval()
    mats = CV.padded_matrices(m, 3, 2)
    x = torch.randn(16, 3, dtype=torch.float64)
    a = np.concatenate([x.numpy(), np.ones((16, 1)), np.zeros((16, 4))], axis=1)
    for j, M in enumerate(mats):
        a = a @ M.T
        if j < len(mats) - 1:
            a = np.maximum(a, 0.0)
    assert np.allclose(a[:, :2], m.double()(x).detach().numpy(), atol=1e-9)
    m.train()
    with pytest.raises(NotImplementedError):
        CV.padded_matrices(m, 3, 2)


def test_unsupported_networks():
    with pytest.raises(NotImplementedError):
        CV.padded_matrices(Mlp(2, [64], 1, batch_norm=False, batch_norm_init=False), 2, 1)
    with pytest.raises(NotImplementedError):
        CV.padded_matrices(Mlp(2, [8] * 5, 1, batch_norm=False, batch_norm_init=False), 2, 1)
    with pytest.raises(NotImplementedError):
        CV.padded_matrices(Mlp(2, [8], 1, activation=nn.Tanh, batch_norm=False,
                               batch_norm_init=False), 2, 1)
    with pytest.raises(NotImplementedError):
        CV.padded_matrices(Mlp(2, [8], 1, final_activation=nn.ReLU, batch_norm=False,
                               batch_norm_init=False), 2, 1)
    assert not MC._is_eval_packable(Lstm(2, 8, 1))
    assert not MC._is_eval_packable(Gru(2, 8, 1))
    assert MC._is_eval_packable(ZeroFunction(1))
    assert MC._is_eval_packable(Mlp(2, [8], 1, batch_norm=False, batch_norm_init=False))


class _SquaredCall(EuroCall):
    """A payoff the kernels do not know."""

    def payoff(self, x):
        return torch.clamp(x[:, 0] - self.strike, min=0.0) ** 2


def test_fused_eligibility_rules():
    from sde_mc_amd.problem import BlackScholesEuroCall, MertonEuroCall
    p = BlackScholesEuroCall.default_params(10, 'cpu')
    f = Mlp(2, [8], 1, batch_norm=False, batch_norm_init=False)
    # a discounter other than a constant short rate, or a payoff outside the
    # kernel's set, routes to the torch loops
    assert MC.fused_control_variates(f, p.solver, p.payoff, lambda t: torch.exp(-0.02 * t)) is None
    assert MC.fused_control_variates(f, p.solver, _SquaredCall(1.0), p.discounter) is None
    assert MC.fused_control_variates(Lstm(2, 4, 1), p.solver, p.payoff, p.discounter) is None
    q = MertonEuroCall.default_params(10, 'cpu')
    g = Mlp(2, [8], 1, batch_norm=False, batch_norm_init=False)
    assert MC.fused_control_variates(f, q.solver, q.payoff, q.discounter) is None  # needs [f, g]
    assert MC.fused_control_variates([f, g], q.solver, q.payoff, q.discounter, tol=0.1) is None


def test_integrate_cv_structures():
    """'diag' sums f dW over dims; 'indep' pairs the two noise columns."""
    torch.manual_seed(0)
    n = torch.randn(3, 5, 2)
    f = torch.randn(3, 5, 2)
    d = torch.exp(-0.1 * torch.arange(5.0)).view(1, 5, 1)
    out = integrate_cv(n, f, d, 'diag')
    ref = torch.tensor([sum(float(n[b, s, i] * f[b, s, i] * d[0, s, 0]) for s in range(5)
                            for i in range(2)) for b in range(3)])
    assert torch.allclose(out, ref, atol=1e-5)
    n4 = torch.randn(3, 5, 2, 2)
    f4 = torch.randn(3, 5, 2, 2)
    out4 = integrate_cv(n4, f4, d, 'indep')
    ref4 = ((n4 * f4).sum(-1).sum(-1) * d[..., 0]).sum(-1)
    assert torch.allclose(out4, ref4, atol=1e-5)
    # tol drops the steps after T - tol: remove_steps(1.0, 5, 5.0) = 4
    cut = integrate_cv(n, f, d, 'diag', tol=1.0, time_interval=5.0)
    assert torch.allclose(cut, integrate_cv(n[:, :4], f[:, :4], d[:, :4], 'diag'))


def test_early_stopping_threshold():
    es = EarlyStopping(eps=0.01, quantile=1.96, cost_batch=0.5, alpha=1.0)
    es.cost_epoch, es.batch_size = 2.0, 1000
    thr = 1.0 * 2.0 * 0.01 ** 2 * 1000 / (0.5 * 1.96 ** 2)
    assert es.threshold() == pytest.approx(thr)
    assert es.stop(thr * 0.99) and not es.stop(thr * 1.01)

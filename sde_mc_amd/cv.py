"""Neural control variates on the device: a torch ReLU MLP f(t, x) packed into
the bf16 MFMA fragments of sdemc_cv_simulate (include/sdemc.h), and the fused
run that simulates the paths, evaluates f at every step and accumulates

    gamma = disc * payoff(X_N) + sum_{n < cv_steps} exp(-r t_n) <f(t_n, X_n), dW_n>

(apply_diffusion_control_variate, sde_mc/varred.py:77-95, with integrate_cv's
'diag' sum, varred.py:202-214).

Supported networks: Linear / ReLU stacks with 1..3 hidden layers of at most
63 units, input (t, x) and output of the model's dimension, optionally with
BatchNorm1d layers in eval mode (folded into the neighbouring Linear).
Anything else raises NotImplementedError.
"""
import ctypes

import numpy as np
import torch
import torch.nn as nn

from . import _lib as L
from . import engine as E

HIDDEN_PAD = 64
MAX_HIDDEN = 4
MAX_OUT = 8


class CvMlp(ctypes.Structure):
    _fields_ = [("weights_dev", ctypes.c_void_p), ("hidden_layers", ctypes.c_int32),
                ("n_out", ctypes.c_int32)]


class CvSpec(ctypes.Structure):
    _fields_ = [("f", CvMlp), ("g", CvMlp), ("cv_steps", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("rate", ctypes.c_double),
                ("jump_comp", ctypes.c_double)]


def _affine_layers(model):
    """[(W, b)] of the network's Linear layers (fp64), eval-mode BatchNorm1d
    folded in; checks the layer pattern Linear (ReLU Linear)*."""
    mods = [m for m in model.modules() if not isinstance(m, nn.Sequential) and m is not model
            and len(list(m.children())) == 0]
    layers, pending_in, relus = [], None, []
    for m in mods:
        if isinstance(m, nn.BatchNorm1d):
            if m.training:
                raise NotImplementedError('BatchNorm1d must be in eval mode to be folded')
            s = (m.weight.detach().double() if m.affine else torch.ones(m.num_features, dtype=torch.float64))
            s = s / torch.sqrt(m.running_var.double() + m.eps)
            c = (m.bias.detach().double() if m.affine else 0.0) - s * m.running_mean.double()
            if layers and len(relus) == len(layers) - 1:
                W, b = layers[-1]                  # after a Linear: scale its outputs
                layers[-1] = (s[:, None] * W, s * b + c)
            else:                                   # on the inputs: fold into the next Linear
                pending_in = (s, c)
        elif isinstance(m, nn.Linear):
            W = m.weight.detach().double()
            b = m.bias.detach().double() if m.bias is not None else torch.zeros(W.shape[0], dtype=torch.float64)
            if pending_in is not None:
                s, c = pending_in
                b = b + W @ c
                W = W * s[None, :]
                pending_in = None
            if len(relus) != len(layers):
                raise NotImplementedError('Linear layers must be separated by ReLU')
            layers.append((W, b))
        elif isinstance(m, nn.ReLU):
            if len(relus) != len(layers) - 1:
                raise NotImplementedError('one ReLU after each hidden Linear')
            relus.append(m)
        else:
            raise NotImplementedError('%s has no device form' % type(m).__name__)
    if len(layers) < 2 or len(relus) != len(layers) - 1:
        raise NotImplementedError('need Linear (ReLU Linear)+ with no final activation')
    return [(W.cpu().numpy(), b.cpu().numpy()) for W, b in layers]


def _bf16_bits(a):
    """Round-to-nearest-even fp32 -> bf16 bit patterns (uint16)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def padded_matrices(model, n_in, n_out):
    """The network as padded transposed matrices Wt[o][i] (fp64): the first
    takes 8 inputs (the n_in features, 1, 0...), each hidden layer has 64
    units of which unit `width` is the constant 1 carrying the next layer's
    bias, the output has 32 rows of which the first n_out are used."""
    layers = _affine_layers(model)
    hidden = len(layers) - 1
    if hidden > MAX_HIDDEN:
        raise NotImplementedError('at most %d hidden layers' % MAX_HIDDEN)
    if layers[0][0].shape[1] != n_in or layers[-1][0].shape[0] != n_out:
        raise NotImplementedError('network must map R^%d to R^%d' % (n_in, n_out))
    if n_in + 1 > 8 or n_out > MAX_OUT:
        raise NotImplementedError('at most 7 inputs and %d outputs' % MAX_OUT)
    mats = []
    in_width = n_in
    for j, (W, b) in enumerate(layers):
        out_w = W.shape[0]
        last = j == hidden
        if not last and out_w > HIDDEN_PAD - 1:
            raise NotImplementedError('hidden layers of at most %d units' % (HIDDEN_PAD - 1))
        M = np.zeros((32 if last else HIDDEN_PAD, 8 if j == 0 else HIDDEN_PAD))
        M[:out_w, :in_width] = W
        M[:out_w, in_width] = b            # column `in_width` is the constant unit
        if not last:
            M[out_w, in_width] = 1.0       # keeps the constant for the next layer
        mats.append(M)
        in_width = out_w
    return mats


def unit_order(c, h, q):
    """Input unit at K position (h, q) of 16-unit chunk c: the order in which
    a 32x32 accumulator tile's rows sit in the lanes."""
    return 16 * c + 8 * (q // 4) + 4 * h + q % 4


BLOCK = 512   # bf16 words of one operand fragment (64 lanes x 8)


def blocks(hidden):
    return 8 * hidden - 2


def pack_matrices(mats):
    """bf16 words (uint16) of padded matrices [first (64x8), hidden (64x64)...,
    output (32x64)] in the fragment order of sdemc_cv_mlp."""
    hidden = len(mats) - 1
    lane = np.arange(64)
    row = lane % 32
    half = lane // 32
    q = np.arange(8)
    u = np.stack([unit_order(0, half, qq) for qq in q], axis=1)   # [64, 8] for chunk 0
    out = np.zeros((blocks(hidden), 64, 8))
    for m in range(2):
        out[m] = mats[0][32 * m + row][:, q]                       # both halves alike
    for k in range(1, hidden):
        for m in range(2):
            for c in range(4):
                out[2 + 8 * (k - 1) + 4 * m + c] = mats[k][(32 * m + row)[:, None], u + 16 * c]
    for c in range(4):
        out[2 + 8 * (hidden - 1) + c] = mats[-1][row[:, None], u + 16 * c]
    return _bf16_bits(out.astype(np.float32)).reshape(-1)


def pack_network(model, n_in, n_out):
    """bf16 words (uint16) and hidden-layer count of sdemc_cv_mlp."""
    mats = padded_matrices(model, n_in, n_out)
    return pack_matrices(mats), len(mats) - 1


def is_zero_network(model):
    return isinstance(model, ZeroNetwork) or getattr(model, 'is_zero_function', False)


class ZeroNetwork:
    """Marker of the zero function (no work in the kernel)."""


class DeviceNetwork:
    """A packed network resident on the device (None model: the zero function)."""

    def __init__(self, model, n_in, n_out, device):
        self.n_out = n_out
        if model is None or is_zero_network(model):
            self.hidden, self.words = 0, None
            return
        arr, self.hidden = pack_network(model, n_in, n_out)
        self.words = torch.from_numpy(arr.view(np.int16).copy()).to(device)

    def c(self):
        m = CvMlp()
        m.weights_dev = self.words.data_ptr() if self.words is not None else None
        m.hidden_layers = self.hidden
        m.n_out = self.n_out
        return m


def cv_simulate(job, f_net, g_net, cv_steps, rate, jump_comp, device, gamma=False):
    """Fused control-variate run of one job: device stats [4] fp64 (sum, sum
    of squares, n, flags) and, with gamma=True, the per-path values [P] fp32."""
    lib = L.lib()
    spec = CvSpec()
    spec.f = f_net.c()
    spec.g = g_net.c() if g_net is not None else CvMlp()
    spec.cv_steps = int(cv_steps)
    spec.rate = float(rate)
    spec.jump_comp = float(jump_comp)
    ws = E._WS.get(lib.sdemc_workspace_bytes(job.n_paths), device)
    stats = torch.empty(4, dtype=torch.float64, device=device)
    g = torch.empty(int(job.n_paths), dtype=torch.float32, device=device) if gamma else None
    st = lib.sdemc_cv_simulate(ctypes.byref(job), ctypes.byref(spec), E.ptr(stats), E.ptr(g),
                               E.ptr(ws), ws.numel(), E.stream_ptr())
    L.check(st, 'sdemc_cv_simulate')
    return stats, g

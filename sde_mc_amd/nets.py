"""Control-variate networks and path datasets (API of sde_mc/nets.py).

The networks are ordinary torch modules: they are trained with torch
autograd on the device (varred.py), and at inference the MLPs are packed
into bf16 matrix-core operands and evaluated inside the path kernel
(sde_mc_amd/cv.py).  Recurrent models (Lstm, Gru) read whole paths and are
evaluated by torch on stored paths.
"""
import torch
import torch.nn as nn
import torch.optim as optim
from torch.utils.data import Dataset


class ControlVariate(nn.Module):
    """Base of the control-variate models (nets.py:7-20): `sequential` models
    take a whole path (bs, steps, 1 + dim), the others single points
    (bs, 1 + dim)."""

    def __init__(self, sequential, device):
        super().__init__()
        self.sequential = sequential
        self.device = device


class ZeroFunction(ControlVariate):
    """f = 0 (nets.py:23-36): the plain estimator through the control-variate
    code path."""

    is_zero_function = True

    def __init__(self, output_dim):
        super().__init__(sequential=False, device=None)
        self.output_dim = output_dim

    def forward(self, x):
        return x.new_zeros((x.shape[0], self.output_dim))


class Mlp(ControlVariate):
    """Feed-forward network (nets.py:39-93): optional BatchNorm1d on the
    inputs, then per hidden layer Linear, optional BatchNorm1d and the
    activation, then a Linear output and an optional final activation.  The
    layers sit in `self.net` in that order, so state dicts carry over."""

    def __init__(self, input_size, layer_sizes, output_size, activation=nn.ReLU,
                 final_activation=None, batch_norm=True, batch_norm_init=True, device='cpu'):
        assert len(layer_sizes) > 0, "At least one hidden layer required."
        super().__init__(sequential=False, device=device)
        self.num_layers = len(layer_sizes)
        mods = [nn.BatchNorm1d(input_size, device=device)] if batch_norm_init else []
        width = input_size
        for size in layer_sizes:
            mods.append(nn.Linear(width, size, device=device))
            if batch_norm:
                mods.append(nn.BatchNorm1d(size, device=device))
            mods.append(activation())
            width = size
        mods.append(nn.Linear(width, output_size, device=device))
        if final_activation is not None:
            mods.append(final_activation())
        self.net = nn.Sequential(*mods)

    def forward(self, x):
        return self.net(x)


class _Recurrent(ControlVariate):
    """Recurrent model over a whole path with a linear read-out per step."""

    cell = None

    def __init__(self, in_dim, hidden_dim, out_dim, device='cpu'):
        super().__init__(sequential=True, device=device)
        self.in_dim, self.hidden_dim, self.out_dim = in_dim, hidden_dim, out_dim
        self.rnn = self.cell(in_dim, hidden_dim, batch_first=True).to(device)
        self.lin = nn.Linear(hidden_dim, out_dim).to(device)

    def init_hidden(self, bs):
        return torch.zeros((1, bs, self.hidden_dim), device=self.device)

    def forward(self, x):
        seq, _ = self.rnn(x, self.init_hidden(x.shape[0]))
        return self.lin(seq)


class Lstm(_Recurrent):
    """LSTM control variate (nets.py:96-128)."""

    cell = nn.LSTM

    def init_hidden(self, bs):
        z = super().init_hidden(bs)
        return z, z.clone()


class Gru(_Recurrent):
    """GRU control variate (nets.py:131-162)."""

    cell = nn.GRU


class NormalPathData(Dataset):
    """((path[:-1], increments), payoff) per trajectory (nets.py:165-175)."""

    def __init__(self, paths, payoffs, normals):
        self.paths = paths[:, :-1]
        self.payoffs = payoffs
        self.normals = normals

    def __len__(self):
        return len(self.payoffs)

    def __getitem__(self, idx):
        return (self.paths[idx], self.normals[idx]), self.payoffs[idx]


class NormalJumpsPathData(Dataset):
    """As NormalPathData with the jump marks (nets.py:178-189)."""

    def __init__(self, paths, payoffs, normals, jumps):
        self.paths = paths[:, :-1]
        self.payoffs = payoffs
        self.normals = normals
        self.jumps = jumps

    def __len__(self):
        return len(self.payoffs)

    def __getitem__(self, idx):
        return (self.paths[idx], self.normals[idx], self.jumps[idx]), self.payoffs[idx]


class AdaptedPathData(Dataset):
    """Rows 0..total_steps-1 of jump-adapted trajectories: states, increments,
    pre-jump states, grid times and jump marks (nets.py:192-207)."""

    def __init__(self, paths, payoffs, normals, left_paths, time_paths, jump_paths, total_steps):
        self.paths = paths[:, :-1]
        self.payoffs = payoffs
        self.normals = normals
        self.left_paths = left_paths[:, :-1]
        self.time_paths = time_paths[:, :-1]
        self.jump_paths = jump_paths[:, :-1]
        self.total_steps = total_steps

    def __len__(self):
        return len(self.payoffs)

    def __getitem__(self, idx):
        return ((self.paths[idx], self.normals[idx], self.left_paths[idx], self.time_paths[idx],
                 self.jump_paths[idx]), self.payoffs[idx])


def get_mlps(problem, num_layers, hidden_size, device):
    """Brownian network f: (t, x) -> R^brown_dim and, for jump models, the
    jump network g: (t, x) -> R^dim, each with num_layers hidden layers of
    hidden_size + dim + 1 units and no batch norm after the layers
    (nets.py:210-219)."""
    d = problem.dim() + 1
    layers = [hidden_size + d] * num_layers
    f = Mlp(d, layers, problem.solver.sde.brown_dim, batch_norm=False, device=device)
    if problem.solver.has_jumps:
        return [f, Mlp(d, layers, d - 1, batch_norm=False, device=device)]
    return f


def get_opt(models):
    """Adam over the parameters of one model or a [f, g] pair (nets.py:222-229)."""
    if isinstance(models, list):
        return optim.Adam(list(models[0].parameters()) + list(models[1].parameters()))
    return optim.Adam(models.parameters())

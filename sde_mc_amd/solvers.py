"""SDE solvers (sde_mc/solvers.py) whose time loop runs in the HIP kernels.

``solve`` and ``multilevel_solve`` keep the reference's signatures and return
layouts; the whole path (every step, the jumps and, in the estimators, the
payoff and the fp64 sums) runs in one kernel launch, one path per lane.
Random numbers come from a counter-based Philox stream keyed by the solver's
seed, so the paths a call produces depend only on (seed, path index); each
call takes the next unused range of path indices, the way successive calls of
the reference draw on from ``torch.manual_seed(seed)`` (solvers.py:37).

Results are computed on the current HIP device and returned on
``self.device`` ('cpu' by default, as in the reference).  There is no CPU
path: without a HIP device every simulation raises HipUnavailableError.
"""
import math
from abc import ABC

import numpy as np
import torch

from . import _lib as L
from . import engine as E
from .schemes import EulerScheme, HestonScheme


class SdeSolver(ABC):
    """Common state of the solvers (solvers.py:9-56)."""

    default_grid = 'uniform'

    def __init__(self, sde, time_interval, num_steps, device='cpu', seed=1):
        self.sde = sde
        self.time_interval = time_interval
        self.num_steps = num_steps
        self.device = device
        self.has_jumps = bool(torch.as_tensor(self.sde.jump_rate()).any())
        corr = torch.as_tensor(self.sde.corr_matrix)
        self.lower_cholesky = torch.linalg.cholesky(corr) if len(corr) > 1 else torch.ones(1, 1)
        torch.manual_seed(seed)  # the reference seeds torch here (solvers.py:37)
        self.seed = int(seed)
        self.path_counter = 0
        self.waves_per_block = 4
        self.exact_jumps = False
        self.grid = self.default_grid

    # -- reference helpers ------------------------------------------------
    def sample_corr_normals(self, size, h, corr=True):
        """N(0, h) increments of shape `size` (trailing axis of length 1), mixed
        by the Cholesky factor when `corr` (solvers.py:51-56); host torch RNG."""
        z = math.sqrt(float(h)) * torch.randn(size)
        mixed = self.lower_cholesky @ z if corr else z
        return mixed.squeeze(-1)

    def init_storage(self, bs, steps, low_storage=False):
        """Path storage (bs, steps + 1, dim) for callers that step by hand."""
        return torch.zeros(bs, steps + 1, self.sde.dim)

    def step(self, t, x, h, corr_normals):
        """One scheme step in torch (the scheme mixins provide it, schemes.py)."""
        raise NotImplementedError

    def solve(self, bs=1, return_normals=False):
        raise NotImplementedError

    # -- kernel plumbing --------------------------------------------------
    def _grid_code(self):
        if self.grid not in ('uniform', 'adapted'):
            raise ValueError("grid must be 'uniform' or 'adapted'")
        return L.GRID_ADAPTED if self.grid == 'adapted' else L.GRID_UNIFORM

    def _rate(self):
        return float(torch.as_tensor(self.sde.jump_rate()).double().sum())

    def _take_paths(self, n):
        off = self.path_counter
        self.path_counter += int(n)
        return off

    def _max_steps(self, steps):
        return int(steps) + E.safe_max_jumps(self._rate(), self.time_interval)

    def make_job(self, n, steps=None, coarse_factor=0, option=None, discount=1.0,
                 path_offset=None):
        """Job descriptor for n paths of this solver (and, optionally, a payoff)."""
        steps = self.num_steps if steps is None else steps
        model = self.sde.kernel_model()
        off = self._take_paths(n) if path_offset is None else path_offset
        job = E.make_job(model, self.time_interval, steps, self._grid_code(), n, self.seed, off,
                         coarse_factor=coarse_factor, max_steps=self._max_steps(steps),
                         exact_jumps=self.exact_jumps)
        in_kernel = E.fill_payoff(job, option, discount)
        return job, in_kernel

    def _out_device(self):
        return torch.device(self.device) if not isinstance(self.device, torch.device) else self.device

    def simulate_stats(self, n, option, discount, steps=None, coarse_factor=0):
        """Fused run: n paths, payoff and fp64 (sum, sum of squares, n).

        With an Option the kernels cannot evaluate, terminal states come back
        and the payoff is applied in torch on the device."""
        dev = E.require_device()
        job, in_kernel = self.make_job(n, steps, coarse_factor, option, discount)
        if in_kernel:
            stats = E.simulate(job, dev, wpb=self.waves_per_block)
            return E.stats_to_host(stats)
        d = self.sde.dim
        lv = 2 if coarse_factor else 1
        term = torch.empty((lv, d, n), dtype=torch.float32, device=dev)
        out = L.PathOut()
        out.terminal_dev = term.data_ptr()
        stats = E.simulate(job, dev, out=out, wpb=self.waves_per_block)
        E.stats_to_host(stats)
        pay = option(term[0].t()) * discount
        if coarse_factor:
            pay = pay - option(term[1].t()) * discount
        pay = pay.double()
        return float(pay.sum()), float((pay * pay).sum()), int(n)

    def simulate_payoffs(self, n, option, discount):
        """Per-path discounted payoffs (fp32, device) and the terminal states (n, dim)."""
        dev = E.require_device()
        job, in_kernel = self.make_job(n, None, 0, option, discount)
        d = self.sde.dim
        term = torch.empty((d, n), dtype=torch.float32, device=dev)
        pay = torch.empty(n, dtype=torch.float32, device=dev)
        out = L.PathOut()
        out.terminal_dev = term.data_ptr()
        out.payoff_dev = pay.data_ptr()
        stats = E.simulate(job, dev, out=out, wpb=self.waves_per_block)
        E.stats_to_host(stats)
        x = term.t()
        if not in_kernel:
            pay = option(x) * discount
        return pay, x

    def _normal_rows(self):
        """Rows per step of the stored increments: dim ('diag'), the padded
        Brownian dimension ('general'), else (dW, dB) pairs."""
        kind, d = self.sde.diffusion_struct, self.sde.dim
        return d if kind == 'diag' else (L.MAX_DIM if kind == 'general' else 2 * d)

    def _normals_out(self, nrm):
        """Stored increments [S][K][bs] in the reference's layout: (bs, S, dim)
        for 'diag', (bs, S, brown_dim) for 'general', (bs, S, dim, 2) else."""
        kind, d = self.sde.diffusion_struct, self.sde.dim
        if kind == 'diag':
            return nrm.permute(2, 0, 1)
        if kind == 'general':
            return nrm[:, :self.sde.brown_dim].permute(2, 0, 1)
        return nrm.reshape(nrm.shape[0], d, 2, nrm.shape[-1]).permute(3, 0, 1, 2)

    def _materialise(self, bs, low_storage):
        """One stored solve: paths [S+1][d][P] and the aux arrays, on device."""
        dev = E.require_device()
        d = self.sde.dim
        job, _ = self.make_job(bs)
        adapted = job.grid == L.GRID_ADAPTED
        rows = job.max_steps if adapted else job.steps
        kn = self._normal_rows()
        paths = torch.empty((rows + 1, d, bs), dtype=torch.float32, device=dev)
        nsteps = torch.empty(bs, dtype=torch.int32, device=dev)
        out = L.PathOut()
        out.paths_dev = paths.data_ptr()
        out.nsteps_dev = nsteps.data_ptr()
        aux = {}
        if not low_storage:
            aux['normals'] = torch.zeros((rows, kn, bs), dtype=torch.float32, device=dev)
            aux['jumps'] = torch.zeros((rows + 1, d, bs), dtype=torch.float32, device=dev)
            out.normals_dev = aux['normals'].data_ptr()
            out.jumps_dev = aux['jumps'].data_ptr()
            if adapted:
                aux['times'] = torch.empty((rows + 1, bs), dtype=torch.float32, device=dev)
                aux['left'] = torch.empty((rows + 1, d, bs), dtype=torch.float32, device=dev)
                out.times_dev = aux['times'].data_ptr()
                out.left_dev = aux['left'].data_ptr()
        stats = E.simulate(job, dev, out=out, wpb=self.waves_per_block)
        E.stats_to_host(stats, 'solve')
        return job, rows, paths, nsteps, aux


class DiffusionSolver(SdeSolver):
    """Fixed-grid solver for SDEs without jumps (solvers.py:59-119)."""

    def solve(self, bs=1, return_normals=False):
        """paths (bs, N+1, dim) and the scaled correlated increments (solvers.py:68-88)."""
        bs = int(bs)
        job, rows, paths, nsteps, aux = self._materialise(bs, low_storage=False)
        od = self._out_device()
        p = paths.permute(2, 0, 1).to(od)
        return p, self._normals_out(aux['normals']).to(od)

    def multilevel_solve(self, bs, levels, return_normals=False):
        """Coupled fine/coarse pair (solvers.py:90-119).  Returns (bs, 2, dim) tensors
        holding the initial and terminal state of each level; the coarse increments
        are the sums of the fine ones."""
        return _multilevel(self, bs, levels)


class JumpDiffusionSolver(SdeSolver):
    """Jump-adapted solver (solvers.py:130-307).

    grid='adapted' (default) is the reference's event grid: each step ends at
    min(t + h, next jump time, T) on an fp32 clock and a jump is detected with
    torch.isclose's tolerances.  grid='uniform' uses N steps of T/N with the
    jumps of each step applied at its end.
    """

    default_grid = 'adapted'

    def __init__(self, sde, time_interval, num_steps, device='cpu', seed=1, exact_jumps=False,
                 grid='adapted'):
        super().__init__(sde, time_interval, num_steps, device, seed)
        self.max_jumps = E.reference_max_jumps(self._rate(), time_interval)
        self.exact_jumps = exact_jumps
        self.grid = grid

    # -- the reference's torch helpers for stepping by hand (solvers.py:140-148)
    def add_jumps(self, t, old_x, x, jumps):
        """x plus the jump term c(t, old_x, J)."""
        return x + self.sde.jumps(t, old_x, jumps)

    def sample_jump_times(self, size):
        """Arrival times: cumulative Exp(lambda) gaps along axis 1, torch RNG."""
        rate = float(torch.as_tensor(self.sde.jump_rate()).sum())
        return torch.empty(size).exponential_(rate).cumsum(dim=1)

    def sample_one_jump(self, size):
        """One mark per path, repeated over the dimensions."""
        return self.sde.sample_jumps([size, 1], 'cpu').repeat(1, self.sde.dim)

    def solve(self, bs=1, return_normals=False, low_storage=False):
        """paths[:, :total_steps + 1] and (normals, time_paths, left_paths, total_steps,
        jump_paths) as in solvers.py:164-226 (aux is None under low_storage)."""
        bs = int(bs)
        job, rows, paths, nsteps, aux = self._materialise(bs, low_storage)
        od = self._out_device()
        d = self.sde.dim
        T = float(self.time_interval)
        if job.grid == L.GRID_ADAPTED:
            total = int(nsteps.max()) if bs > 0 else 0
        else:
            total = job.steps
        out_paths = paths[:total + 1].permute(2, 0, 1).to(od)
        if low_storage:
            return out_paths, (None, None, None, total, None)
        if job.grid == L.GRID_ADAPTED:
            times = aux['times']
            times[total + 1:] = T
            left = aux['left']
            left[total + 1:] = 0.0
            time_paths = times.t().unsqueeze(-1)
            left_paths = left.permute(2, 0, 1)
        else:
            # uniform grid: times n * h; pre-jump states are not kept
            grid_t = torch.arange(job.steps + 1, dtype=torch.float32, device=paths.device)
            time_paths = (grid_t * float(np.float32(T / job.steps))).expand(bs, -1).unsqueeze(-1)
            left_paths = None
        jumps = aux['jumps']
        jumps[total + 1:] = 0.0
        normals = self._normals_out(aux['normals'])
        left_out = left_paths.to(od) if left_paths is not None else None
        return out_paths, (normals.to(od), time_paths.to(od), left_out, total,
                           jumps.permute(2, 0, 1).to(od))

    def multilevel_solve(self, bs, levels, return_normals=False):
        """Coupled fine/coarse pair sharing increments and jumps (solvers.py:228-307).
        Returns (bs, 2, dim) tensors holding the initial and terminal state of
        each level."""
        return _multilevel(self, bs, levels)


def _multilevel(solver, bs, levels):
    bs = int(bs)
    fine, coarse = levels
    factor = int(fine / coarse)
    dev = E.require_device()
    d = solver.sde.dim
    job, _ = solver.make_job(bs, steps=fine, coarse_factor=factor)
    term = torch.empty((2, d, bs), dtype=torch.float32, device=dev)
    out = L.PathOut()
    out.terminal_dev = term.data_ptr()
    stats = E.simulate(job, dev, out=out, wpb=solver.waves_per_block)
    E.stats_to_host(stats, 'multilevel_solve')
    x0 = torch.tensor(solver.sde.kernel_model().x0, dtype=torch.float32, device=dev)
    od = solver._out_device()
    pf = torch.stack([x0.expand(bs, d), term[0].t()], dim=1).to(od)
    pc = torch.stack([x0.expand(bs, d), term[1].t()], dim=1).to(od)
    return (pf, pc), None


class EulerSolver(EulerScheme, DiffusionSolver):
    """Euler scheme on the fixed grid (solvers.py:122-123)."""


class HestonSolver(HestonScheme, DiffusionSolver):
    """Heston scheme on the fixed grid (solvers.py:126-127)."""


class JumpEulerSolver(EulerScheme, JumpDiffusionSolver):
    """Euler scheme on the jump-adapted grid (solvers.py:310-311)."""


class Grid(ABC):
    """Iterator over the time points of [start, end] (solvers.py:314-326)."""

    def __init__(self, start, end):
        self.start = start
        self.end = end
        self.time_interval = end - start
        self.t = start

    def __iter__(self):
        return self

    def __next__(self):
        raise NotImplementedError


class UniformGrid(Grid):
    """start, start + h, ... up to end (within 1e-8), h = (end - start) / num_steps
    (solvers.py:329-341)."""

    def __init__(self, start, end, num_steps):
        super().__init__(start, end)
        self.num_steps = num_steps
        self.h = (end - start) / num_steps
        assert self.h > 1e-8

    def __next__(self):
        if self.t > self.end - 1e-8:
            raise StopIteration
        t = self.t
        self.t += self.h
        return t

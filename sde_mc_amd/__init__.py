"""sde_mc_amd: the Monte Carlo path-integration hot path of sde_mc
(piers-hinds/sde_mc v0.4) on AMD MI355X.

The public names mirror sde_mc's (``from sde_mc_amd import *`` stands in for
``from sde_mc import *`` on the pricing path).  Solvers, payoffs and the
(sum, sum of squares) reduction run in hand-written gfx950 HIP kernels
reached through the C ABI of ``lib/libsdemc.so`` (include/sdemc.h).
"""
from .version import __version__
from .helpers import (partition, solve_quadratic, mc_estimates, remove_steps, get_corr_matrix,
                      ceil_mult, get_jump_comp, sample_cov)
from .sde import (Sde, DiffusionSde, Gbm, LogGbm, DoubleGbm, GeneralSde, MultiFactorGbm, Heston,
                  LogNormalJumpsSde, Merton, AsianWrapper, KernelModel)
from .levy import UNIFORM_TOL, InverseCdf, Levy, LevySde, ExampleLevy, ExpExampleLevy, Levy2d
from .schemes import EulerScheme, HestonScheme
from .options import (bs_binary_aon, bs_call, merton_call, bs_digital_call, bs_asian_call, Option,
                      EuroCall, EuroPut, BinaryAoN, Basket, Rainbow, Digital, AsianCall,
                      HestonRainbow, BestOf, ConstantShortRate)
from .solvers import (SdeSolver, DiffusionSolver, JumpDiffusionSolver, EulerSolver,
                      HestonSolver, JumpEulerSolver, Grid, UniformGrid)
from .mc import (MCStatistics, mc_simple, find_num_trials, run_mc, mc_control_variates,
                 mc_apply_cvs, mc_adaptive_cv, mc_terminal_cv, simulate_data,
                 simulate_adapted_data, sim_train_control_variates, sample_batch_cost,
                 find_num_trials_terminal_cv, run_mc_terminal_cv, run_cv_mc)
from .nets import (ControlVariate, ZeroFunction, Mlp, Lstm, Gru, NormalPathData,
                   NormalJumpsPathData, AdaptedPathData, get_mlps, get_opt)
from .varred import (EarlyStopping, train_diffusion_control_variate,
                     apply_diffusion_control_variate, apply_adapted_control_variates,
                     train_adapted_control_variates, integrate_cv)
from .mlmc import mc_multilevel, get_optimal_trials, mlmc_bs_from_trials
from .problem import (Problem, BlackScholesEuroCall, BlackScholesRainbow, HestonEuroCall,
                      MertonEuroCall,
                      LevyRainbow, LevyRainbowMLMC, LevyCall, LevyBestOf, LevyCallOnMax)
from ._lib import SdemcError, HipUnavailableError

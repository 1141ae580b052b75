// sdemc_kernels.h -- __global__ kernels of the plain (one job per launch)
// path and their host-side launchers, templated on the model family and
// dimension.  Each family is instantiated in its own translation unit.
#pragma once

#include "sdemc_core.h"

namespace sdemc {

// Plain RNG-mode kernel: one wavefront per 64-path chunk.  Paths, payoff and
// the chunk's fp64 partial sums in one pass; per-path outputs are optional.
// POIS: the uniform single-level path's jump mode (the job's pois_mode), so
// each kernel holds the registers of one mode only.
template <int D, int FAM, int GRID, bool COUPLED, bool STORE, bool POIS = false>
__global__ __launch_bounds__(512) void k_rng(const KJob job, double* __restrict__ partials,
                                             int32_t* __restrict__ cflags,
                                             const sdemc_path_out out, int64_t nchunks) {
  const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (chunk >= nchunks) return;  // whole wavefront
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  const KPayoff& pf = job.pay;
  PathCtx c;
  c.path = job.path_offset + (uint64_t)p;
  c.p = p;
  c.P = P;
  c.k0 = (uint32_t)job.seed;
  c.k1 = (uint32_t)(job.seed >> 32);
  StoreCtx st;
  st.paths = out.paths_dev;
  st.times = out.times_dev;
  st.left = out.left_dev;
  st.jumps = out.jumps_dev;
  st.normals = out.normals_dev;
  st.P = P;
  st.rows = GRID == SDEMC_GRID_ADAPTED ? job.max_steps : job.steps;
  float v = 0.0f;
  int ovf = 0;
  int32_t nsteps = job.steps;
  if (!COUPLED) {
    float x[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      path_uniform_rng_mode<D, FAM, STORE, STORE ? 1 : 2, POIS>(job, c, valid, x, st);
    } else {
      AdaptRngSrc<D, FAM> src;
      src.init(job.m, c);
      nsteps = path_adapted<D, FAM, STORE>(job, src, valid, p, x, ovf, st);
    }
    if (valid) {
      v = payoff_raw<D>(pf, x) * pf.disc;
      if (out.terminal_dev) {
#pragma unroll
        for (int i = 0; i < D; ++i) out.terminal_dev[(int64_t)i * P + p] = x[i];
      }
    }
  } else {
    float xf[D], xc[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      path_uniform_rng_coupled<D, FAM>(job, c, xf, xc);
    } else {
      AdaptRngSrc<D, FAM> src;
      src.init(job.m, c);
      nsteps = path_adapted_coupled<D, FAM>(job, src, valid, xf, xc, ovf);
    }
    if (valid) {
      v = pf.disc * (payoff_raw<D>(pf, xf) - payoff_raw<D>(pf, xc));
      if (out.terminal_dev) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          out.terminal_dev[(int64_t)i * P + p] = xf[i];
          out.terminal_dev[(int64_t)(D + i) * P + p] = xc[i];
        }
      }
    }
  }
  if (valid) {
    if (out.payoff_dev) out.payoff_dev[p] = v;
    if (out.nsteps_dev) out.nsteps_dev[p] = nsteps;
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store(dv, dv * dv, ovf, partials + 2 * chunk, cflags + chunk);
}

// Fixed increments, uniform grid, single level: every thread of a workgroup
// runs the same steps (threads past the last path repeat path P - 1 and
// discard it, none exits early) so the workgroup can keep its wavefronts in
// step with barriers (path_uniform_fixed<..., true>).
template <int D, int FAM>
__global__ __launch_bounds__(256) void k_fixed_uniform(const KJob job, const sdemc_fixed_in in,
                                                       double* __restrict__ partials,
                                                       int32_t* __restrict__ cflags,
                                                       const sdemc_path_out out,
                                                       int64_t nchunks) {
  const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  const int64_t pp = valid ? p : P - 1;
  const KPayoff& pf = job.pay;
  int bad = 0;
  float x[D];
  path_uniform_fixed<D, FAM, true>(job, in, pp, P, x, bad);
  if (chunk >= nchunks) return;  // whole wavefront, after the last barrier
  float v = 0.0f;
  if (valid) {
    v = payoff_raw<D>(pf, x) * pf.disc;
    if (out.terminal_dev) {
#pragma unroll
      for (int i = 0; i < D; ++i) out.terminal_dev[(int64_t)i * P + p] = x[i];
    }
    if (out.payoff_dev) out.payoff_dev[p] = v;
    if (out.nsteps_dev) out.nsteps_dev[p] = job.steps;
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store(dv, dv * dv, (valid && bad) ? 2 : 0, partials + 2 * chunk, cflags + chunk);
}

// Plain fixed-increment kernel: the same stepping fed from host arrays.
// Lanes past the last path read nothing.
template <int D, int FAM, int GRID, bool COUPLED>
__global__ __launch_bounds__(256) void k_fixed(const KJob job, const sdemc_fixed_in in,
                                               double* __restrict__ partials,
                                               int32_t* __restrict__ cflags,
                                               const sdemc_path_out out, int64_t nchunks) {
  const int64_t chunk = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (chunk >= nchunks) return;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  const KPayoff& pf = job.pay;
  StoreCtx st{};
  float v = 0.0f;
  int ovf = 0, bad = 0;
  int32_t nsteps = job.steps;
  if (!COUPLED) {
    float x[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      if (valid) path_uniform_fixed<D, FAM>(job, in, p, P, x, bad);
    } else {
      AdaptFixedSrc<D, FAM> src;
      src.init(in, p, P);
      nsteps = path_adapted<D, FAM, false>(job, src, valid, p, x, ovf, st);
    }
    if (valid) {
      v = payoff_raw<D>(pf, x) * pf.disc;
      if (out.terminal_dev) {
#pragma unroll
        for (int i = 0; i < D; ++i) out.terminal_dev[(int64_t)i * P + p] = x[i];
      }
    }
  } else {
    float xf[D], xc[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      if (valid) {
        UniFixedSrc<D, FAM> src;
        src.init(in, p, P);
        path_uniform_coupled<D, FAM>(job, src, xf, xc);
        bad = src.bad;
      }
    } else {
      AdaptFixedSrc<D, FAM> src;
      src.init(in, p, P);
      nsteps = path_adapted_coupled<D, FAM>(job, src, valid, xf, xc, ovf);
    }
    if (valid) {
      v = pf.disc * (payoff_raw<D>(pf, xf) - payoff_raw<D>(pf, xc));
      if (out.terminal_dev) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          out.terminal_dev[(int64_t)i * P + p] = xf[i];
          out.terminal_dev[(int64_t)(D + i) * P + p] = xc[i];
        }
      }
    }
  }
  if (valid) {
    if (out.payoff_dev) out.payoff_dev[p] = v;
    if (out.nsteps_dev) out.nsteps_dev[p] = nsteps;
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store(dv, dv * dv, ovf | (bad ? 2 : 0), partials + 2 * chunk, cflags + chunk);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
struct LaunchArgs {
  double* partials;
  int32_t* cflags;
  sdemc_path_out out;
  bool store;
  int wpb;
  hipStream_t stream;
};

template <int D, int FAM>
int launch_rng_d(const sdemc_job& sj, const LaunchArgs& a) {
  KJob job;
  make_kjob(sj, job);
  const int64_t nchunks = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
  const int threads = WAVE * a.wpb;
  const int64_t blocks = (nchunks + a.wpb - 1) / a.wpb;
  if (blocks <= 0 || blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  const bool coupled = job.M > 0;
  const bool adapted = job.grid == SDEMC_GRID_ADAPTED;
  const dim3 g((unsigned)blocks), b((unsigned)threads);
  // the per-wavefront jump lists of the uniform grid's per-step Poisson mode
  const size_t lds = (size_t)a.wpb * sizeof(PoisLds);
#define SDEMC_L(G, C, S)                                                                     \
  hipLaunchKernelGGL((k_rng<D, FAM, G, C, S>), g, b, lds, a.stream, job, a.partials, a.cflags,    \
                     a.out, nchunks)
#define SDEMC_LU(S)                                                                           \
  if (job.pois_mode)                                                                          \
    hipLaunchKernelGGL((k_rng<D, FAM, SDEMC_GRID_UNIFORM, false, S, true>), g, b, lds, a.stream, \
                       job, a.partials, a.cflags, a.out, nchunks);                            \
  else                                                                                        \
    SDEMC_L(SDEMC_GRID_UNIFORM, false, S)
  if constexpr (FAM == FAM_DOUBLE || FAM == FAM_HESTON || FAM == FAM_GENERAL) {
    if (adapted) return SDEMC_ERR_UNSUPPORTED;
    if (coupled) SDEMC_L(SDEMC_GRID_UNIFORM, true, false);
    else if (a.store) { SDEMC_LU(true); }
    else { SDEMC_LU(false); }
  } else {
    if (!coupled) {
      if (!adapted) {
        if (a.store) { SDEMC_LU(true); }
        else { SDEMC_LU(false); }
      } else {
        if (a.store) SDEMC_L(SDEMC_GRID_ADAPTED, false, true);
        else SDEMC_L(SDEMC_GRID_ADAPTED, false, false);
      }
    } else {
      if (a.store) return SDEMC_ERR_UNSUPPORTED;
      if (!adapted) SDEMC_L(SDEMC_GRID_UNIFORM, true, false);
      else SDEMC_L(SDEMC_GRID_ADAPTED, true, false);
    }
  }
#undef SDEMC_L
#undef SDEMC_LU
  return hipGetLastError() == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

template <int D, int FAM>
int launch_fixed_d(const sdemc_job& sj, const sdemc_fixed_in& in, const LaunchArgs& a) {
  KJob job;
  make_kjob(sj, job);
  const int64_t nchunks = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
  const int wpb = 4;
  const int64_t blocks = (nchunks + wpb - 1) / wpb;
  if (blocks <= 0 || blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  const bool coupled = job.M > 0;
  const bool adapted = job.grid == SDEMC_GRID_ADAPTED;
  const dim3 g((unsigned)blocks), b((unsigned)(WAVE * wpb));
#define SDEMC_L(G, C)                                                                        \
  hipLaunchKernelGGL((k_fixed<D, FAM, G, C>), g, b, 0, a.stream, job, in, a.partials, a.cflags, \
                     a.out, nchunks)
  if constexpr (FAM == FAM_DOUBLE || FAM == FAM_HESTON || FAM == FAM_GENERAL) {
    if (adapted) return SDEMC_ERR_UNSUPPORTED;
    if (coupled) SDEMC_L(SDEMC_GRID_UNIFORM, true);
    else SDEMC_L(SDEMC_GRID_UNIFORM, false);
  } else {
    if (!coupled) {
      if (!adapted)
        hipLaunchKernelGGL((k_fixed_uniform<D, FAM>), g, b, 0, a.stream, job, in, a.partials,
                           a.cflags, a.out, nchunks);
      else SDEMC_L(SDEMC_GRID_ADAPTED, false);
    } else {
      if (!adapted) SDEMC_L(SDEMC_GRID_UNIFORM, true);
      else SDEMC_L(SDEMC_GRID_ADAPTED, true);
    }
  }
#undef SDEMC_L
  return hipGetLastError() == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

template <int FAM>
int launch_rng_fam(const sdemc_job& job, const LaunchArgs& a) {
  switch (job.model.dim) {
    case 1: return launch_rng_d<1, FAM>(job, a);
    case 2: return launch_rng_d<2, FAM>(job, a);
    case 3: return launch_rng_d<3, FAM>(job, a);
    case 4: return launch_rng_d<4, FAM>(job, a);
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

template <int FAM>
int launch_fixed_fam(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  switch (job.model.dim) {
    case 1: return launch_fixed_d<1, FAM>(job, in, a);
    case 2: return launch_fixed_d<2, FAM>(job, in, a);
    case 3: return launch_fixed_d<3, FAM>(job, in, a);
    case 4: return launch_fixed_d<4, FAM>(job, in, a);
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

// Per-family entry points (one translation unit each).
int launch_rng_merton(const sdemc_job& job, const LaunchArgs& a);
int launch_rng_levy(const sdemc_job& job, const LaunchArgs& a);
int launch_rng_double(const sdemc_job& job, const LaunchArgs& a);
int launch_fixed_merton(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a);
int launch_fixed_levy(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a);
int launch_fixed_double(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a);
int launch_rng_heston(const sdemc_job& job, const LaunchArgs& a);
int launch_fixed_heston(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a);
int launch_rng_general(const sdemc_job& job, const LaunchArgs& a);
int launch_fixed_general(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a);

}  // namespace sdemc

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

// Fixed increments, uniform grid, single level, four paths per lane: lane l
// of a wavefront steps paths 4l .. 4l + 3 of the wavefront's 256, so a row
// of increments is one 16-byte load per lane and a row of counts one 4-byte
// load, whole 128-byte lines per wavefront, one step's rows in flight (more
// rows in flight per wavefront read slower here).  It needs P % 4 == 0,
// increments 16-byte and counts 4-byte aligned (then every row is), which
// the launcher checks.  Each path takes path_uniform_fixed's steps in its
// order; the wavefront's four 64-path chunks (lanes 16k .. 16k + 15) reduce
// in chunk_store's butterfly order: its lane distances 32, 16, 8, 4 are
// lanes 8, 4, 2, 1 apart here and 2, 1 the elements within a lane.  Results
// are k_fixed_uniform's bit for bit.
template <int D, int FAM>
__global__ __launch_bounds__(256) void k_fixed_uniform4(const KJob job, const sdemc_fixed_in in,
                                                        double* __restrict__ partials,
                                                        int32_t* __restrict__ cflags,
                                                        const sdemc_path_out out,
                                                        int64_t nchunks) {
  static_assert(FAM == FAM_MERTON || FAM == FAM_LEVY, "rows of dW and one common dB at most");
  constexpr int KF = NKF<D, FAM>::K;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (wave * 4 >= nchunks) return;  // whole wavefront
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p0 = wave * (4 * WAVE) + 4 * lane;
  const bool valid = p0 < P;               // all four paths or none (P % 4 == 0)
  const int64_t pb = valid ? p0 : P - 4;   // lanes past the end repeat the last four
  const KModel& f = job.m;
  const KPayoff& pf = job.pay;
  const int32_t N = job.steps;
  float x[4][D];
  int64_t off[4];
  int bad[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
#pragma unroll
    for (int i = 0; i < D; ++i) x[e][i] = f.x0[i];
    off[e] = in.jump_off_dev ? in.jump_off_dev[pb + e] : 0;
    bad[e] = 0;
  }
  const float* dwp = in.dw_dev + pb;
  const uint8_t* cntp = in.counts_dev ? in.counts_dev + pb : nullptr;
  for (int32_t n = 0; n < N; ++n) {
    float4 r[KF];
#pragma unroll
    for (int q = 0; q < KF; ++q)
      r[q] = *reinterpret_cast<const float4*>(dwp + ((int64_t)n * KF + q) * P);
    const uint32_t c4 = cntp ? *reinterpret_cast<const uint32_t*>(cntp + (int64_t)n * P) : 0u;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float dW[D], dB[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        dW[i] = e == 0 ? r[i].x : (e == 1 ? r[i].y : (e == 2 ? r[i].z : r[i].w));
        dB[i] = 0.0f;
        if (FAM == FAM_LEVY) {  // the common row, as fixed_row reads it
          const float4& c = r[KF - 1];
          dB[i] = e == 0 ? c.x : (e == 1 ? c.y : (e == 2 ? c.z : c.w));
        }
      }
      fixed_step<D, FAM>(job, in, x[e], dW, dB, (int)((c4 >> (8 * e)) & 0xFFu), off[e], bad[e]);
    }
  }
  double S[4], Q[4];
  int flag = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float v = 0.0f;
    if (valid) {
      const int64_t p = p0 + e;
      v = payoff_raw<D>(pf, x[e]) * pf.disc;
      if (out.terminal_dev) {
#pragma unroll
        for (int i = 0; i < D; ++i) out.terminal_dev[(int64_t)i * P + p] = x[e][i];
      }
      if (out.payoff_dev) out.payoff_dev[p] = v;
      if (out.nsteps_dev) out.nsteps_dev[p] = job.steps;
      if (bad[e]) flag = SDEMC_FLAG_JUMP_CSR;
    }
    S[e] = valid ? (double)v : 0.0;
    Q[e] = S[e] * S[e];
  }
  // chunk_store's levels 32, 16, 8, 4 (lanes of the chunk) ...
#pragma unroll
  for (int d = 8; d > 0; d >>= 1) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      S[e] += __shfl_xor(S[e], d);
      Q[e] += __shfl_xor(Q[e], d);
    }
  }
  // ... then 2 and 1 (elements of the lane)
  const double s = (S[0] + S[2]) + (S[1] + S[3]);
  const double q = (Q[0] + Q[2]) + (Q[1] + Q[3]);
  const int k = lane >> 4;
  const uint64_t fl = __ballot(flag != 0);
  const int64_t chunk = wave * 4 + k;
  if ((lane & 15) == 0 && chunk < nchunks) {
    partials[2 * chunk] = s;
    partials[2 * chunk + 1] = q;
    cflags[chunk] = ((fl >> (16 * k)) & 0xFFFFull) != 0 ? SDEMC_FLAG_JUMP_CSR : 0;
  }
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

// Whether the four-paths-per-lane kernel can read the rows: P % 4 == 0,
// increments 16-byte and counts 4-byte aligned.
inline bool fixed_rows_vector_aligned(const sdemc_fixed_in& in, uint64_t P) {
  return P % 4 == 0 && ((uintptr_t)in.dw_dev & 15u) == 0 &&
         ((uintptr_t)in.counts_dev & 3u) == 0;
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
      if (!adapted && fixed_rows_vector_aligned(in, job.n_paths)) {
        const int64_t waves4 = (nchunks + 3) / 4;
        hipLaunchKernelGGL((k_fixed_uniform4<D, FAM>), dim3((unsigned)((waves4 + wpb - 1) / wpb)),
                           b, 0, a.stream, job, in, a.partials, a.cflags, a.out, nchunks);
      } else if (!adapted)
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

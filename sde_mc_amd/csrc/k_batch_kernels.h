// k_batch_kernels.h -- device side of the batched persistent launch: the
// chunk body of each job class, the persistent tile kernel and its launch.
// The class kernels are instantiated in separate translation units
// (k_batch_*.hip) so they compile in parallel; k_batch.hip holds the host
// side and names them through the extern declarations below.
#pragma once

#include "sdemc_kernels.h"

namespace sdemc {

// One 64-path chunk of one job, RNG mode, no per-path outputs.  The result
// equals the plain kernel's chunk: same Philox streams, same butterfly.
// MODE: the uniform single-level path's jump mode, 0 the arrival clock, 1
// per-step Poisson counts, -1 whichever the job has.
template <int D, int FAM, int GRID, bool COUPLED, int U = 1, int MODE = -1>
__device__ __forceinline__ void chunk_body(const KJob& job, int64_t chunk, double* part,
                                           int32_t* cf) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  PathCtx c;
  c.path = job.path_offset + (uint64_t)p;
  c.p = p;
  c.P = P;
  c.k0 = (uint32_t)job.seed;
  c.k1 = (uint32_t)(job.seed >> 32);
  const StoreCtx st{};
  float v = 0.0f;
  int ovf = 0;
  if (!COUPLED) {
    float x[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      if constexpr (MODE < 0)
        path_uniform_rng<D, FAM, false, U>(job, c, valid, x, st);
      else
        path_uniform_rng_mode<D, FAM, false, U, MODE == 1>(job, c, valid, x, st);
    } else {
      AdaptRngSrc<D, FAM> src;
      src.init(job.m, c);
      path_adapted<D, FAM, false>(job, src, valid, p, x, ovf, st);
    }
    if (valid) v = payoff_raw<D>(job.pay, x) * job.pay.disc;
  } else {
    float xf[D], xc[D];
    if (GRID == SDEMC_GRID_UNIFORM) {
      path_uniform_rng_coupled<D, FAM>(job, c, xf, xc);
    } else {
      AdaptRngSrc<D, FAM> src;
      src.init(job.m, c);
      path_adapted_coupled<D, FAM>(job, src, valid, xf, xc, ovf);
    }
    if (valid) v = job.pay.disc * (payoff_raw<D>(job.pay, xf) - payoff_raw<D>(job.pay, xc));
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store<true>(dv, dv * dv, ovf, part, cf);
}

// Class of a job: dedicated kernels for the 1-D 'diag' and 2-D 'indep'
// families (grid and level type folded in), else the generic kernel.  The
// dedicated uniform single-level kernels each hold one jump mode, so none
// carries the other mode's registers: 1-D arrival clock (class 0), 2-D
// per-step Poisson counts (class 4) and 2-D arrival clock (class 8); 1-D
// per-step count jobs run in the generic kernel.
inline int32_t class_of(int noise, int dim, int grid, bool coupled, bool pois) {
  int32_t fam = -1;
  if (noise == SDEMC_NOISE_DIAG && dim == 1) fam = 0;
  if (noise == SDEMC_NOISE_INDEP_COMMON && dim == 2) fam = 1;
  if (fam < 0) return SDEMC_CLASS_GENERIC;
  if (grid == SDEMC_GRID_UNIFORM && !coupled) {
    if (fam == 0) return pois ? SDEMC_CLASS_GENERIC : SDEMC_CLASS_DIAG1;
    return pois ? SDEMC_CLASS_INDEP2 : SDEMC_CLASS_INDEP2_CLOCK;
  }
  return fam * 4 + grid * 2 + (coupled ? 1 : 0);
}

inline int32_t job_class_of(const sdemc_job& j) {
  return class_of(j.model.noise, j.model.dim, j.grid, j.coarse_factor > 0, job_pois_mode(j));
}

// The class of a prepared slot, decoded from its dispatch key
// ((noise * 4 + dim - 1) * 2 + grid) * 2 + coupled and its jump mode.
inline int32_t kjob_class_of(const KJob& k) {
  const int key = k.class_key;
  return class_of(key >> 4, ((key >> 2) & 3) + 1, (key >> 1) & 1, (key & 1) != 0,
                  k.pois_mode != 0);
}

// The generic class calls each case's chunk body as a function of its own,
// once per 64-path chunk: the cases are compiled one by one instead of all
// inlined into one kernel body (about 1.5 minutes of compile time instead of
// 15), and the call costs nothing next to a chunk's path work.
template <int D, int FAM, int GRID, bool COUPLED>
__device__ __attribute__((noinline)) void chunk_call(const KJob& job, int64_t chunk, double* part,
                                                     int32_t* cf) {
  chunk_body<D, FAM, GRID, COUPLED>(job, chunk, part, cf);
}

#define SDEMC_CASE(NOISE, D, FAM, GRID, C)                              \
  case ((NOISE * 4 + (D - 1)) * 2 + GRID) * 2 + C:                      \
    chunk_call<D, FAM, GRID, (C != 0)>(job, chunk, part, cf);        \
    return true;

#define SDEMC_CASES_DG(NOISE, D, FAM, GRID) \
  SDEMC_CASE(NOISE, D, FAM, GRID, 0)        \
  SDEMC_CASE(NOISE, D, FAM, GRID, 1)

// One chunk of a job in kernel class CLS: the dedicated classes run exactly
// one (family, grid, level type) -- the code of the plain kernel for it --
// and refuse anything else; the generic class runs every supported job.
template <int CLS, int U>
__device__ __forceinline__ bool run_chunk(const KJob& job, int64_t chunk, double* part,
                                          int32_t* cf) {
  if (CLS < SDEMC_CLASS_GENERIC) {
    // class 8 is class 4's family, grid and level type in the other mode
    constexpr int BASE = CLS == SDEMC_CLASS_INDEP2_CLOCK ? SDEMC_CLASS_INDEP2 : CLS;
    constexpr int FAMI = BASE / 4, GRID = (BASE / 2) % 2;
    constexpr bool COUPLED = (BASE % 2) != 0;
    constexpr int NOISE = FAMI == 0 ? SDEMC_NOISE_DIAG : SDEMC_NOISE_INDEP_COMMON;
    constexpr int D = FAMI == 0 ? 1 : 2;
    constexpr int FAM = FAMI == 0 ? FAM_MERTON : FAM_LEVY;
    if (job.class_key != ((NOISE * 4 + (D - 1)) * 2 + GRID) * 2 + (COUPLED ? 1 : 0)) return false;
    constexpr int MODE = (GRID == SDEMC_GRID_UNIFORM && !COUPLED)
                             ? (CLS == SDEMC_CLASS_INDEP2 ? 1 : 0)
                             : -1;
    if (MODE >= 0 && job.pois_mode != MODE) return false;
    chunk_body<D, FAM, GRID, COUPLED, U, MODE>(job, chunk, part, cf);
    return true;
  } else {
    switch (job.class_key) {
      SDEMC_CASES_DG(0, 1, FAM_MERTON, 0)
      SDEMC_CASES_DG(0, 1, FAM_MERTON, 1)
      SDEMC_CASES_DG(0, 2, FAM_MERTON, 0)
      SDEMC_CASES_DG(0, 2, FAM_MERTON, 1)
      SDEMC_CASES_DG(0, 3, FAM_MERTON, 0)
      SDEMC_CASES_DG(0, 3, FAM_MERTON, 1)
      SDEMC_CASES_DG(0, 4, FAM_MERTON, 0)
      SDEMC_CASES_DG(0, 4, FAM_MERTON, 1)
      SDEMC_CASES_DG(1, 1, FAM_LEVY, 0)
      SDEMC_CASES_DG(1, 1, FAM_LEVY, 1)
      SDEMC_CASES_DG(1, 2, FAM_LEVY, 0)
      SDEMC_CASES_DG(1, 2, FAM_LEVY, 1)
      SDEMC_CASES_DG(1, 3, FAM_LEVY, 0)
      SDEMC_CASES_DG(1, 3, FAM_LEVY, 1)
      SDEMC_CASES_DG(1, 4, FAM_LEVY, 0)
      SDEMC_CASES_DG(1, 4, FAM_LEVY, 1)
      SDEMC_CASES_DG(2, 1, FAM_DOUBLE, 0)
      SDEMC_CASES_DG(2, 2, FAM_DOUBLE, 0)
      SDEMC_CASES_DG(2, 3, FAM_DOUBLE, 0)
      SDEMC_CASES_DG(2, 4, FAM_DOUBLE, 0)
      SDEMC_CASES_DG(3, 2, FAM_HESTON, 0)
      SDEMC_CASES_DG(4, 1, FAM_GENERAL, 0)
      SDEMC_CASES_DG(4, 2, FAM_GENERAL, 0)
      SDEMC_CASES_DG(4, 3, FAM_GENERAL, 0)
      SDEMC_CASES_DG(4, 4, FAM_GENERAL, 0)
      default: return false;
    }
  }
}

#undef SDEMC_CASE
#undef SDEMC_CASES_DG

// A launch whose segment table does not describe the prepared job table
// (a segment naming an unprepared slot, chunks outside its job or the
// partial pool, tiles outside their segment, a header outside the segment
// list) computes what it can and sets the launch's fault word; every slot
// finalised by that launch then carries flag SDEMC_FLAG_TABLE, so a
// mis-planned round never returns stale statistics as a result.
__device__ __forceinline__ void raise_fault(int32_t* fault, int lane) {
  if (lane == 0) *fault = 1;
}

// Persistent tile kernel of one class group.  hdr = {first segment,
// segments, tiles}; tile indices come from this group's counter.  U is the
// step unroll of the uniform single-level path (classes 0, 4 and 8).
// Wavefronts per SIMD: the dedicated class kernels are held to 64 VGPRs so
// that 8 fit (the served mix's per-step count kernel went from 66 VGPRs, 7
// wavefronts, to 64 with 20 B of spills per lane, and the round from 41.4 to
// 39.7 ms; the coupled and adapted classes move by -7% .. +2%); the generic
// class, whose cases' register needs go up to ~250 VGPRs for the 4-D
// two-driver family, to 128 (4 wavefronts; a few of its cases spill a
// little) instead of running at 2 per SIMD.
template <int CLS, int U>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(
    CLS == SDEMC_CLASS_GENERIC ? 4 : 8))) void k_batch(const KJob* __restrict__ jobs,
                                               const sdemc_segment* __restrict__ segs,
                                               const int32_t* __restrict__ hdr,
                                               double* __restrict__ partials,
                                               int32_t* __restrict__ cflags,
                                               int32_t* __restrict__ counter,
                                               int32_t* __restrict__ fault,
                                               int64_t partial_chunks, int32_t max_jobs,
                                               int32_t max_segments) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int32_t s0 = hdr[0];
  int32_t nseg = hdr[1];
  const int32_t ntiles = hdr[2];
  if (ntiles <= 0) return;  // nothing to run in this group
  if (s0 < 0 || s0 >= max_segments || nseg <= 0) {
    raise_fault(fault, lane);
    return;
  }
  if (nseg > max_segments - s0) {
    raise_fault(fault, lane);
    nseg = max_segments - s0;
  }
  int32_t seg = s0;
  const int32_t last = s0 + nseg - 1;
  for (;;) {
    int32_t t = 0;
    if (lane == 0) t = atomicAdd(counter, 1);
    t = __shfl(t, 0);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= ntiles) break;
    // segment owning tile t: the last one whose tile_begin <= t; a wave's
    // tile indices only grow, so search [seg, last]
    int32_t lo = seg, hi = last;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (segs[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    seg = lo;
    const sdemc_segment sg = segs[seg];
    if (sg.job < 0 || sg.job >= max_jobs || sg.per_tile < 1 || sg.count < 1 || sg.first < 0 ||
        t < sg.tile_begin) {
      raise_fault(fault, lane);
      continue;
    }
    const KJob& job = jobs[sg.job];
    if (job.magic != KJOB_MAGIC) {
      raise_fault(fault, lane);
      continue;
    }
    const int64_t nc = (int64_t)((job.n_paths + WAVE - 1) / WAVE);
    const int64_t seg_end = (int64_t)sg.first + sg.count;
    const int64_t c0 = (int64_t)sg.first + (int64_t)(t - sg.tile_begin) * sg.per_tile;
    int64_t c1 = c0 + sg.per_tile;
    // the segment's last tile (the one before the next segment's first)
    // must reach its last chunk, or chunks of it would run in no tile
    const int32_t next_begin = seg < last ? segs[seg + 1].tile_begin : ntiles;
    const bool short_range = t == next_begin - 1 && c1 < seg_end;
    if (c1 > seg_end) c1 = seg_end;
    // a tile past its segment's chunks, or chunks past the job's
    if (short_range || c0 >= seg_end || seg_end > nc || job.chunk_base < 0 ||
        job.chunk_base + seg_end > partial_chunks) {
      raise_fault(fault, lane);
      if (c0 >= seg_end || seg_end > nc || job.chunk_base < 0 ||
          job.chunk_base + seg_end > partial_chunks)
        continue;
    }
    for (int64_t chunk = c0; chunk < c1; ++chunk) {
      const int64_t slot = job.chunk_base + chunk;
      if (!run_chunk<CLS, U>(job, chunk, partials + 2 * slot, cflags + slot) && lane == 0) {
        // a job outside this kernel's class: poison its partial and flag it
        partials[2 * slot] = __builtin_nan("");
        partials[2 * slot + 1] = __builtin_nan("");
        cflags[slot] = SDEMC_FLAG_CLASS;
      }
    }
  }
}

// Launch of group g of a round in kernel class CLS (one instantiation per
// class and unroll, spread over the k_batch_*.hip translation units).
template <int CLS, int U>
void launch_group(const sdemc_batch_dev& b, int g, unsigned grid, int wpb, hipStream_t s) {
  hipLaunchKernelGGL((k_batch<CLS, U>), dim3(grid), dim3(WAVE * wpb),
                     (size_t)wpb * sizeof(PoisLds), s, (const KJob*)b.jobs_dev,
                     b.segs_dev, b.seg_hdr_dev + 4 * g, b.partials_dev, b.chunk_flags_dev,
                     b.counter_dev + g, b.counter_dev + SDEMC_MAX_GROUPS, b.partial_chunks,
                     b.max_jobs, b.max_segments);
}

#define SDEMC_BATCH_INSTANCES(X) \
  X(0, 1) X(0, 2) X(1, 1) X(2, 1) X(3, 1) X(4, 1) X(4, 2) X(5, 1) X(6, 1) X(7, 1) \
  X(SDEMC_CLASS_INDEP2_CLOCK, 1) X(SDEMC_CLASS_INDEP2_CLOCK, 2) X(SDEMC_CLASS_GENERIC, 1)

#define SDEMC_LAUNCH_GROUP_SIG(CLS, U)                                                    \
  void launch_group<CLS, U>(const sdemc_batch_dev& b, int g, unsigned grid, int wpb, \
                            hipStream_t s);

}  // namespace sdemc

// sdemc_api.hip -- extern "C" entry points of libsdemc: host-side validation,
// dispatch to the family launchers, the fixed-order finalisation kernels and
// the generator test kernels.  Nothing here throws across the ABI.
#include <math.h>
#include <string.h>

#include <mutex>

#include "sdemc_kernels.h"
#include "cv_mlp.h"

namespace sdemc {

static thread_local int g_last_hip_error = 0;

static int hip_status(hipError_t e) {
  if (e == hipSuccess) return SDEMC_OK;
  g_last_hip_error = (int)e;
  return SDEMC_ERR_LAUNCH;
}

static bool finite(double v) { return isfinite(v) != 0; }

// Chunk-partial workspace of sdemc_simulate_noalloc: a file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (on the host the name is only a
// stand-in); it uses the device address hipGetSymbolAddress returns, looked
// up once per device on its first call and kept in scratch_dev[].
static __device__ __align__(256) unsigned char scratch[SDEMC_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static void* scratch_dev[kMaxDevices];
static std::mutex scratch_mu;

static int scratch_address(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e);
  if (dev < 0 || dev >= kMaxDevices) return SDEMC_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e);
    if (!p) return SDEMC_ERR_DEVICE;
    scratch_dev[dev] = p;
  }
  *ptr = scratch_dev[dev];
  return SDEMC_OK;
}

// Host-side checks of one job; everything the kernels assume is enforced
// here (dimensions, grid sizes, jump law constants, payoff kind).
int validate_job(const sdemc_job* job) {
  if (!job) return SDEMC_ERR_ARG;
  const sdemc_model& m = job->model;
  if (m.dim < 1 || m.dim > SDEMC_MAX_DIM) return SDEMC_ERR_UNSUPPORTED;
  if (m.noise < SDEMC_NOISE_DIAG || m.noise > SDEMC_NOISE_GENERAL) return SDEMC_ERR_UNSUPPORTED;
  if (m.jump_law != SDEMC_JUMP_LOGNORMAL && m.jump_law != SDEMC_JUMP_LEVY)
    return SDEMC_ERR_UNSUPPORTED;
  if (m.exact_jumps != 0 && m.exact_jumps != 1) return SDEMC_ERR_ARG;
  if (!finite(m.rate) || m.rate < 0.0) return SDEMC_ERR_ARG;
  for (int i = 0; i < m.dim; ++i) {
    if (!finite(m.x0[i]) || !finite(m.a0[i]) || !finite(m.a1[i]) || !finite(m.b0[i]) ||
        !finite(m.b1[i]) || !finite(m.c0[i]) || !finite(m.c1[i]) || !finite(m.j0[i]) ||
        !finite(m.j1[i]))
      return SDEMC_ERR_ARG;
    for (int j = 0; j < m.dim; ++j)
      if (m.noise != SDEMC_NOISE_GENERAL && !finite(m.chol[i * SDEMC_MAX_DIM + j]))
        return SDEMC_ERR_ARG;
    if (!finite(m.ax[i])) return SDEMC_ERR_ARG;
    if (m.ax[i] != 0.0 && (i == 0 || m.noise != SDEMC_NOISE_DIAG)) return SDEMC_ERR_UNSUPPORTED;
  }
  if (m.rate > 0.0) {
    if (m.jump_law == SDEMC_JUMP_LOGNORMAL) {
      if (!finite(m.jp[0]) || !finite(m.jp[1]) || m.jp[1] < 0.0) return SDEMC_ERR_ARG;
    } else {
      const double cm = m.jp[0], cp = m.jp[1], imu = m.jp[2], al = m.jp[3], eps = m.jp[4];
      const double lda = m.jp[5];
      if (!(cm > 0.0) || !(cp > 0.0) || !finite(cm) || !finite(cp)) return SDEMC_ERR_ARG;
      if (!finite(imu) || imu < 0.0) return SDEMC_ERR_ARG;
      if (!(al > 0.0 && al < 2.0) || !(eps > 0.0 && eps < 1.0)) return SDEMC_ERR_ARG;
      if (!(lda > 0.0) || !finite(lda)) return SDEMC_ERR_ARG;
      for (int k = 6; k <= 10; ++k)
        if (!finite(m.jp[k])) return SDEMC_ERR_ARG;
      if (!(m.jp[6] <= m.jp[7] && m.jp[7] <= m.jp[8])) return SDEMC_ERR_ARG;
    }
  }
  const sdemc_payoff& p = job->payoff;
  if (p.kind < SDEMC_PAYOFF_EURO_CALL || p.kind > SDEMC_PAYOFF_RAINBOW_EVEN)
    return SDEMC_ERR_UNSUPPORTED;
  if (p.kind == SDEMC_PAYOFF_ASIAN_CALL && (!finite(p.param) || !(p.param > 0.0)))
    return SDEMC_ERR_ARG;
  if (p.log != 0 && p.log != 1) return SDEMC_ERR_ARG;
  if (!finite(p.strike) || !finite(p.scale) || !finite(p.disc)) return SDEMC_ERR_ARG;
  if (!finite(job->T) || !(job->T > 0.0)) return SDEMC_ERR_ARG;
  if (job->steps < 1) return SDEMC_ERR_ARG;
  if (job->grid != SDEMC_GRID_UNIFORM && job->grid != SDEMC_GRID_ADAPTED) return SDEMC_ERR_ARG;
  if (job->coarse_factor != 0) {
    if (job->coarse_factor < 2) return SDEMC_ERR_ARG;
    if (job->steps % job->coarse_factor != 0) return SDEMC_ERR_ARG;
  }
  if (job->grid == SDEMC_GRID_ADAPTED && job->max_steps < 1) return SDEMC_ERR_ARG;
  if (job->n_paths > (1ull << 37)) return SDEMC_ERR_ARG;
  // bounded expected jumps per path keeps every per-step jump loop short
  if (m.rate * job->T > 1.0e5) return SDEMC_ERR_ARG;
  if ((m.noise == SDEMC_NOISE_INDEP_FULL || m.noise == SDEMC_NOISE_GENERAL) &&
      job->grid == SDEMC_GRID_ADAPTED)
    return SDEMC_ERR_UNSUPPORTED;
  if (m.noise == SDEMC_NOISE_GENERAL) {
    // the (dim, brown_dim) matrices and the brown_dim x brown_dim Cholesky
    // factor; entries outside them are ignored
    if (m.brown_dim < 1 || m.brown_dim > SDEMC_MAX_DIM) return SDEMC_ERR_UNSUPPORTED;
    for (int i = 0; i < m.brown_dim; ++i)
      for (int j = 0; j < m.brown_dim; ++j)
        if (!finite(m.chol[i * SDEMC_MAX_DIM + j])) return SDEMC_ERR_ARG;
    for (int i = 0; i < m.dim; ++i)
      for (int j = 0; j < m.brown_dim; ++j)
        if (!finite(m.bm0[i * SDEMC_MAX_DIM + j]) || !finite(m.bm1[i * SDEMC_MAX_DIM + j]))
          return SDEMC_ERR_ARG;
  }
  if (m.noise == SDEMC_NOISE_HESTON) {
    // (S, v) with no jumps on the uniform grid; the Feller condition keeps the
    // implicit variance equation's constant term positive (sde.py:260)
    if (m.dim != 2 || job->grid != SDEMC_GRID_UNIFORM) return SDEMC_ERR_UNSUPPORTED;
    if (m.rate != 0.0) return SDEMC_ERR_UNSUPPORTED;
    const double kappa = m.jp[0], theta = m.jp[1], xi = m.jp[2];
    if (!(kappa > 0.0) || !(theta > 0.0) || !(xi > 0.0) || !finite(kappa) || !finite(theta) ||
        !finite(xi))
      return SDEMC_ERR_ARG;
    if (!(2.0 * kappa * theta > xi * xi)) return SDEMC_ERR_ARG;
    if (!(m.x0[1] >= 0.0)) return SDEMC_ERR_ARG;
  }
  return SDEMC_OK;
}

__host__ __device__ int64_t n_chunks(uint64_t n_paths) { return (int64_t)((n_paths + WAVE - 1) / WAVE); }

size_t workspace_bytes(uint64_t n_paths) {
  const size_t nc = (size_t)n_chunks(n_paths);
  size_t b = nc * 2 * sizeof(double) + nc * sizeof(int32_t);
  return (b + 255) & ~(size_t)255;
}

__global__ __launch_bounds__(256) void k_finalize_one(const double* partials,
                                                      const int32_t* cflags, int64_t nc,
                                                      uint64_t n, double* stats) {
  finalize_block(partials, cflags, 0, nc, n, stats);
}

__global__ __launch_bounds__(256) void k_finalize(const double* partials, const int32_t* cflags,
                                                  const int64_t* chunk_base,
                                                  const uint64_t* n_paths, const int32_t* mask,
                                                  int32_t n_jobs, double* stats) {
  const int j = blockIdx.x;
  if (j >= n_jobs) return;
  if (mask && mask[j] == 0) return;
  const uint64_t n = n_paths[j];
  finalize_block(partials, cflags, chunk_base[j], n_chunks(n), n, stats + 4 * (int64_t)j);
}

__global__ void k_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const U4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

__global__ void k_normals(uint32_t k0, uint32_t k1, uint64_t path_offset, int64_t n_paths,
                          int32_t n_per_path, float* out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_paths) return;
  NormalStream ns;
  ns.init(k0, k1, path_offset + (uint64_t)p);
  for (int32_t i = 0; i < n_per_path; ++i) out[(int64_t)i * n_paths + p] = ns.get((uint32_t)i);
}

// VALU issue-rate probes: 8 independent chains per lane, 4 links of each per
// loop iteration (the loop counter lives on the scalar unit), so the
// wavefronts of a full device issue back-to-back instructions of one kind:
//   KIND 0  v_fma_f32                    x = x a + b
//   KIND 1  32-bit integer ALU           x = (x ^ a) + b
//   KIND 2  32 x 32 -> 64-bit multiply   (hi, lo) of x * a, x = hi ^ lo
//   KIND 3  fp32 transcendental + FMA    x = exp2(x a) b
//   KIND 4  one Philox4x32-10 block per iteration (the kernels' generator)
//   KIND 5  a Philox block and three Box-Muller pairs: the 6 normals of one
//           block as the path kernels draw them
//   KIND 6  two independent Philox blocks per iteration (the same work as
//           two iterations of KIND 4 with twice the instruction-level
//           parallelism)
// KIND 0 measures the ceiling the path kernels' VALU instruction counts are
// priced at; the others the costs of Philox (1, 2, 4) and Box-Muller (3, 5).
// HBM streaming-read probe: each thread XOR-folds 16-byte words taken at a
// grid stride and writes one word; over a buffer far larger than the
// last-level cache this measures the read rate that the fixed-increment
// kernels are held against.
__global__ __launch_bounds__(256) void k_hbm_read_probe(const uint4* __restrict__ src, int64_t n,
                                                        uint32_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0u;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint4 v = src[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int32_t iters,
                                                    float a, float b) {
  if constexpr (KIND == 4 || KIND == 5 || KIND == 6) {
    const uint32_t path = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0u;
    float zacc = 0.0f;
    for (int32_t i = 0; i < iters; ++i) {
      const U4 w = philox4x32_10((uint32_t)i, STREAM_BM, path, 0u, __float_as_uint(a), 7u);
      if constexpr (KIND == 6) {
        const U4 v = philox4x32_10((uint32_t)i, STREAM_JUMP, path, 0u, __float_as_uint(a), 7u);
        acc ^= (w.x ^ w.y ^ w.z ^ w.w) + (v.x ^ v.y ^ v.z ^ v.w);
      } else if constexpr (KIND == 4) {
        acc ^= w.x ^ w.y ^ w.z ^ w.w;
      } else {
        float z[NORMALS_PER_BLOCK];
        normals6(w, z);
        zacc += ((z[0] + z[1]) + (z[2] + z[3])) + (z[4] + z[5]);
      }
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] =
        KIND == 5 ? zacc : __uint_as_float(acc);
  } else if constexpr (KIND == 0 || KIND == 3) {
    float x[8];
    const float s = (float)threadIdx.x * 1e-3f;
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = s + (float)j;
    for (int32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if constexpr (KIND == 0) x[j] = __fmaf_rn(x[j], a, b);
          else x[j] = __builtin_amdgcn_exp2f(x[j] * a) * b;
        }
      }
    }
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) acc += x[j];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
  } else {
    uint32_t x[8];
    const uint32_t ua = __float_as_uint(a), ub = __float_as_uint(b);
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = threadIdx.x * 2654435761u + (uint32_t)j;
    for (int32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          if constexpr (KIND == 1) {
            x[j] = (x[j] ^ ua) + ub;
          } else {
            const uint64_t m = (uint64_t)x[j] * (uint64_t)ua;
            x[j] = (uint32_t)(m >> 32) ^ (uint32_t)m;
          }
        }
      }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) acc ^= x[j];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = __uint_as_float(acc);
  }
}

}  // namespace sdemc

using namespace sdemc;

extern "C" {

int sdemc_abi_version(void) { return SDEMC_ABI_VERSION; }

const char* sdemc_status_string(int status) {
  switch (status) {
    case SDEMC_OK: return "ok";
    case SDEMC_ERR_ARG: return "invalid argument";
    case SDEMC_ERR_UNSUPPORTED: return "unsupported model, payoff or grid";
    case SDEMC_ERR_LAUNCH: return "HIP launch or runtime error";
    case SDEMC_ERR_WORKSPACE: return "workspace too small";
    case SDEMC_ERR_OVERFLOW: return "a path exceeded max_steps";
    case SDEMC_ERR_DEVICE: return "no usable device";
    default: return "unknown status";
  }
}

int sdemc_last_hip_error(void) { return g_last_hip_error; }
size_t sdemc_sizeof_job(void) { return sizeof(sdemc_job); }
size_t sdemc_sizeof_model(void) { return sizeof(sdemc_model); }
size_t sdemc_sizeof_path_out(void) { return sizeof(sdemc_path_out); }
size_t sdemc_sizeof_fixed_in(void) { return sizeof(sdemc_fixed_in); }

size_t sdemc_workspace_bytes(uint64_t n_paths) { return workspace_bytes(n_paths); }

int sdemc_device_info(int device, int* cu_count, int* clock_khz, int* wave_size) {
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    g_last_hip_error = (int)e;
    return SDEMC_ERR_DEVICE;
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  if (wave_size) *wave_size = prop.warpSize;
  return SDEMC_OK;
}

static int dispatch_rng(const sdemc_job& job, const LaunchArgs& a) {
  switch (job.model.noise) {
    case SDEMC_NOISE_DIAG: return launch_rng_merton(job, a);
    case SDEMC_NOISE_INDEP_COMMON: return launch_rng_levy(job, a);
    case SDEMC_NOISE_INDEP_FULL: return launch_rng_double(job, a);
    case SDEMC_NOISE_HESTON: return launch_rng_heston(job, a);
    case SDEMC_NOISE_GENERAL: return launch_rng_general(job, a);
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

static int dispatch_fixed(const sdemc_job& job, const sdemc_fixed_in& in, const LaunchArgs& a) {
  switch (job.model.noise) {
    case SDEMC_NOISE_DIAG: return launch_fixed_merton(job, in, a);
    case SDEMC_NOISE_INDEP_COMMON: return launch_fixed_levy(job, in, a);
    case SDEMC_NOISE_INDEP_FULL: return launch_fixed_double(job, in, a);
    case SDEMC_NOISE_HESTON: return launch_fixed_heston(job, in, a);
    case SDEMC_NOISE_GENERAL: return launch_fixed_general(job, in, a);
    default: return SDEMC_ERR_UNSUPPORTED;
  }
}

static int finalize_plain(const sdemc_job& job, double* partials, int32_t* cflags,
                          double* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_one, dim3(1), dim3(256), 0, s, partials, cflags,
                     n_chunks(job.n_paths), job.n_paths, stats);
  return hip_status(hipGetLastError());
}

int sdemc_simulate(const sdemc_job* job, double* stats_dev, const sdemc_path_out* out,
                   void* workspace_dev, size_t workspace_bytes_, int waves_per_block,
                   void* stream) {
  int st = validate_job(job);
  if (st != SDEMC_OK) return st;
  if (!stats_dev || !workspace_dev) return SDEMC_ERR_ARG;
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4 &&
      waves_per_block != 8)
    return SDEMC_ERR_ARG;
  if (workspace_bytes_ < workspace_bytes(job->n_paths)) return SDEMC_ERR_WORKSPACE;
  const int64_t nc = n_chunks(job->n_paths);
  LaunchArgs a;
  a.partials = (double*)workspace_dev;
  a.cflags = (int32_t*)((char*)workspace_dev + (size_t)nc * 2 * sizeof(double));
  memset(&a.out, 0, sizeof(a.out));
  if (out) a.out = *out;
  a.store = a.out.paths_dev != nullptr;
  a.wpb = waves_per_block;
  a.stream = (hipStream_t)stream;
  if (a.store && job->coarse_factor > 0) return SDEMC_ERR_UNSUPPORTED;
  if (nc > 0) {
    st = dispatch_rng(*job, a);
    if (st != SDEMC_OK) return st;
  }
  return finalize_plain(*job, a.partials, a.cflags, stats_dev, a.stream);
}

int sdemc_simulate_noalloc(const sdemc_job* job, double* stats_dev, const sdemc_path_out* out,
                           int waves_per_block, void* stream) {
  int st = validate_job(job);
  if (st != SDEMC_OK) return st;
  const size_t need = workspace_bytes(job->n_paths);
  if (need > (size_t)SDEMC_NOALLOC_SCRATCH_BYTES) return SDEMC_ERR_WORKSPACE;
  void* ws = nullptr;
  st = scratch_address(&ws);
  if (st != SDEMC_OK) return st;
  // every chunk's partial and flag is written before the finalisation reads
  // them, so nothing left by an earlier call needs clearing
  return sdemc_simulate(job, stats_dev, out, ws, need, waves_per_block, stream);
}

size_t sdemc_cv_mlp_words(int32_t hidden_layers) {
  if (hidden_layers < 1 || hidden_layers > SDEMC_CV_MAX_HIDDEN) return 0;
  return (size_t)cv_words(hidden_layers);
}

static int check_mlp(const sdemc_cv_mlp& n, int n_out) {
  if (n.hidden_layers == 0) return SDEMC_OK;  // the zero network
  if (n.hidden_layers < 0 || n.hidden_layers > SDEMC_CV_MAX_HIDDEN) return SDEMC_ERR_ARG;
  if (!n.weights_dev || ((uintptr_t)n.weights_dev & 15u) != 0) return SDEMC_ERR_ARG;
  if (n.n_out != n_out) return SDEMC_ERR_ARG;
  return SDEMC_OK;
}

int sdemc_cv_simulate(const sdemc_job* job, const sdemc_cv_spec* cv, double* stats_dev,
                      float* gamma_dev, void* workspace_dev, size_t workspace_bytes_,
                      void* stream) {
  int st = validate_job(job);
  if (st != SDEMC_OK) return st;
  if (!cv || !stats_dev || !workspace_dev) return SDEMC_ERR_ARG;
  const sdemc_model& m = job->model;
  if (job->coarse_factor != 0 || m.noise == SDEMC_NOISE_GENERAL) return SDEMC_ERR_UNSUPPORTED;
  const bool adapted = job->grid == SDEMC_GRID_ADAPTED;
  if (adapted) {
    if (m.noise != SDEMC_NOISE_DIAG && m.noise != SDEMC_NOISE_INDEP_COMMON)
      return SDEMC_ERR_UNSUPPORTED;
  } else {
    if (m.rate != 0.0 || m.noise == SDEMC_NOISE_INDEP_COMMON) return SDEMC_ERR_UNSUPPORTED;
  }
  const bool diag = m.noise == SDEMC_NOISE_DIAG || m.noise == SDEMC_NOISE_HESTON;
  const int brown = diag ? m.dim : 2 * m.dim;
  if (brown > SDEMC_CV_MAX_OUT || m.dim + 2 > 8) return SDEMC_ERR_UNSUPPORTED;
  if ((st = check_mlp(cv->f, brown)) != SDEMC_OK) return st;
  if ((st = check_mlp(cv->g, m.dim)) != SDEMC_OK) return st;
  if (cv->g.hidden_layers != 0 && !adapted) return SDEMC_ERR_ARG;
  const int32_t limit = adapted ? job->max_steps : job->steps;
  if (cv->cv_steps < 0 || cv->cv_steps > limit) return SDEMC_ERR_ARG;
  if (!isfinite(cv->rate) || !isfinite(cv->jump_comp)) return SDEMC_ERR_ARG;
  if (workspace_bytes_ < workspace_bytes(job->n_paths)) return SDEMC_ERR_WORKSPACE;
  const int64_t nc = n_chunks(job->n_paths);
  CvLaunch a;
  a.fw = cv->f.weights_dev;
  a.f_hidden = cv->f.hidden_layers;
  a.gw = cv->g.weights_dev;
  a.g_hidden = cv->g.hidden_layers;
  a.rate = (float)cv->rate;
  a.comp = (float)cv->jump_comp;
  a.cv_steps = cv->cv_steps;
  a.partials = (double*)workspace_dev;
  a.cflags = (int32_t*)((char*)workspace_dev + (size_t)nc * 2 * sizeof(double));
  a.gamma = gamma_dev;
  a.stream = (hipStream_t)stream;
  if (nc > 0) {
    st = adapted ? launch_cv_adapted(*job, a) : launch_cv_uniform(*job, a);
    if (st != SDEMC_OK) return st;
  }
  return finalize_plain(*job, a.partials, a.cflags, stats_dev, a.stream);
}

int sdemc_simulate_fixed(const sdemc_job* job, const sdemc_fixed_in* in, double* stats_dev,
                         const sdemc_path_out* out, void* workspace_dev,
                         size_t workspace_bytes_, void* stream) {
  int st = validate_job(job);
  if (st != SDEMC_OK) return st;
  if (!in || !stats_dev || !workspace_dev || !in->dw_dev) return SDEMC_ERR_ARG;
  if (workspace_bytes_ < workspace_bytes(job->n_paths)) return SDEMC_ERR_WORKSPACE;
  const int M = job->coarse_factor;
  if (job->grid == SDEMC_GRID_UNIFORM) {
    if (in->n_rows < job->steps) return SDEMC_ERR_ARG;
    if ((in->counts_dev == nullptr) != (in->jump_off_dev == nullptr)) return SDEMC_ERR_ARG;
    if (in->n_marks < 0) return SDEMC_ERR_ARG;
    if (in->n_marks > 0 && !in->jump_size_dev) return SDEMC_ERR_ARG;
  } else {
    const int64_t need = (int64_t)job->max_steps * (M > 0 ? M : 1);
    if ((int64_t)in->n_rows < need) return SDEMC_ERR_ARG;
    if (in->max_jumps < 0) return SDEMC_ERR_ARG;
    if (in->max_jumps > 0 && (!in->jump_time_dev || !in->jump_size_dev)) return SDEMC_ERR_ARG;
  }
  const int64_t nc = n_chunks(job->n_paths);
  LaunchArgs a;
  a.partials = (double*)workspace_dev;
  a.cflags = (int32_t*)((char*)workspace_dev + (size_t)nc * 2 * sizeof(double));
  memset(&a.out, 0, sizeof(a.out));
  if (out) {
    a.out.payoff_dev = out->payoff_dev;
    a.out.terminal_dev = out->terminal_dev;
    a.out.nsteps_dev = out->nsteps_dev;
  }
  a.store = false;
  a.wpb = 4;
  a.stream = (hipStream_t)stream;
  if (nc > 0) {
    st = dispatch_fixed(*job, *in, a);
    if (st != SDEMC_OK) return st;
  }
  return finalize_plain(*job, a.partials, a.cflags, stats_dev, a.stream);
}

int sdemc_reduce_finalize(const double* partials_dev, const int32_t* chunk_flags_dev,
                          const int64_t* chunk_base_dev, const uint64_t* n_paths_dev,
                          int32_t n_jobs, double* stats_dev, void* stream) {
  if (n_jobs < 0) return SDEMC_ERR_ARG;
  if (n_jobs == 0) return SDEMC_OK;
  if (!partials_dev || !chunk_flags_dev || !chunk_base_dev || !n_paths_dev || !stats_dev)
    return SDEMC_ERR_ARG;
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)n_jobs), dim3(256), 0, (hipStream_t)stream,
                     partials_dev, chunk_flags_dev, chunk_base_dev, n_paths_dev,
                     (const int32_t*)nullptr, n_jobs, stats_dev);
  return hip_status(hipGetLastError());
}

int sdemc_philox_dev(const uint32_t* ctr_dev, uint32_t key0, uint32_t key1, uint32_t* out_dev,
                     int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!ctr_dev || !out_dev))) return SDEMC_ERR_ARG;
  if (n == 0) return SDEMC_OK;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  hipLaunchKernelGGL(k_philox, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, ctr_dev,
                     key0, key1, out_dev, n);
  return hip_status(hipGetLastError());
}

int sdemc_hbm_read_probe_dev(const void* src_dev, int64_t bytes, uint32_t* out_dev,
                             int32_t blocks, void* stream) {
  if (!src_dev || !out_dev || bytes < 16 || (bytes & 15) != 0 ||
      ((uintptr_t)src_dev & 15) != 0 || blocks < 1)
    return SDEMC_ERR_ARG;
  hipLaunchKernelGGL(k_hbm_read_probe, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const uint4*)src_dev, bytes / 16, out_dev);
  return hip_status(hipGetLastError());
}

int sdemc_valu_probe_dev(float* out_dev, int32_t blocks, int32_t iters, int32_t kind,
                         void* stream) {
  if (!out_dev || blocks < 1 || iters < 1 || kind < 0 || kind > 6) return SDEMC_ERR_ARG;
  const dim3 g((unsigned)blocks), b(256);
  const hipStream_t s = (hipStream_t)stream;
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_valu_probe<0>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    case 1: hipLaunchKernelGGL(k_valu_probe<1>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    case 2: hipLaunchKernelGGL(k_valu_probe<2>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    case 4: hipLaunchKernelGGL(k_valu_probe<4>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    case 5: hipLaunchKernelGGL(k_valu_probe<5>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    case 6: hipLaunchKernelGGL(k_valu_probe<6>, g, b, 0, s, out_dev, iters, 0.99999f, 1e-6f); break;
    default: hipLaunchKernelGGL(k_valu_probe<3>, g, b, 0, s, out_dev, iters, 0.001f, 0.999f); break;
  }
  return hip_status(hipGetLastError());
}

int sdemc_normals_dev(uint64_t seed, uint64_t path_offset, int64_t n_paths, int32_t n_per_path,
                      float* out_dev, void* stream) {
  if (n_paths < 0 || n_per_path < 0) return SDEMC_ERR_ARG;
  if (n_paths == 0 || n_per_path == 0) return SDEMC_OK;
  if (!out_dev) return SDEMC_ERR_ARG;
  const int64_t blocks = (n_paths + 255) / 256;
  if (blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  hipLaunchKernelGGL(k_normals, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (uint32_t)seed, (uint32_t)(seed >> 32), path_offset, n_paths, n_per_path,
                     out_dev);
  return hip_status(hipGetLastError());
}

}  // extern "C"

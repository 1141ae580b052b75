// sdemc_core.h -- device code of the gfx950 path-integration kernels.
//
// One Monte Carlo path per lane of a 64-wide wavefront.  The path state, the
// Philox counters and the pending jump live in registers for the whole
// horizon; nothing per step touches HBM in RNG mode.  A wavefront's 64 paths
// form one reduction chunk whose fp64 (sum, sum of squares) is reduced by a
// fixed butterfly of lane shuffles, so a job's result does not depend on which
// launch, workgroup or tile computed it.
//
// Reference semantics restated here (file:line in piers-hinds/sde_mc):
//   Euler step                      sde_mc/schemes.py:5-10
//   correlated increments L z sqrt(dt)   sde_mc/solvers.py:51-56, 195-201
//   jump-adapted grid and clock     sde_mc/solvers.py:182-225
//   coupled fine/coarse levels      sde_mc/solvers.py:228-307
//   lognormal marks                 sde_mc/sde.py:325-326
//   Levy marks (inverse CDF)        sde_mc/levy.py:7-30, 85-87
//   payoffs and discounting         sde_mc/options.py:167-176, 179-321, 334-337
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/sdemc.h"

// No implicit multiply-add fusion anywhere below: every fused operation is
// written as __fmaf_rn, so a path's arithmetic is the same instruction
// sequence in every kernel this code is inlined into (plain, fixed-increment
// and batched launches give the same bits for the same job).
#pragma clang fp contract(off)

namespace sdemc {

enum { FAM_MERTON = 0, FAM_LEVY = 1, FAM_DOUBLE = 2, FAM_HESTON = 3, FAM_GENERAL = 4 };

// 'general' noise draws SDEMC_MAX_DIM normals per step whatever its Brownian
// dimension: the Cholesky factor and the diffusion matrices are zero-padded,
// so the unused normals contribute exact zeros.
constexpr int GENERAL_K = SDEMC_MAX_DIM;

// Families whose increments are D correlated normals and no second driver.
template <int FAM>
struct DiagNoise {
  static constexpr bool value = FAM == FAM_MERTON || FAM == FAM_HESTON;
};
enum { STREAM_BM = 0, STREAM_JUMP = 1, STREAM_POIS = 2, STREAM_MARK = 3 };

constexpr int WAVE = 64;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies, key schedule by
// the Weyl constants.  Counter = (block, stream, path_lo, path_hi), key = seed.
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

// a ^ b ^ c for a wave-uniform c: one gfx950 3-input bitwise op (truth
// table 0x96 = odd parity) instead of two xors.
__host__ __device__ __forceinline__ uint32_t xor3_key(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t d;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return a ^ b ^ c;
#endif
}

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                      uint32_t c3, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
  // the key is wave-uniform (one seed per job): recompute its schedule per
  // block on the scalar unit rather than holding 20 round keys in scalar
  // registers for the whole kernel
  k0 = __builtin_amdgcn_readfirstlane(k0);
  k1 = __builtin_amdgcn_readfirstlane(k1);
  asm volatile("" : "+s"(k0), "+s"(k1));
#endif
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = xor3_key(hi1, c1, k0);
    c1 = lo1;
    c2 = xor3_key(hi0, c3, k1);
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// 1 + (w >> 9) 2^-23 in [1, 2), assembled as float bits by one
// v_alignbit_b32: ({0x7F, w} >> 9) = 0x3F800000 | (w >> 9).
__device__ __forceinline__ float one_plus_u23(uint32_t w) {
  return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, w, 9u));
}
// Uniform in (0,1): odd multiples of 2^-24, exact in fp32 (and in fp64):
// ((w >> 8) | 1) 2^-24 = (1 + (w >> 9) 2^-23) - (1 - 2^-24), the
// subtraction exact (Sterbenz).
__device__ __forceinline__ float u01_open(uint32_t w) {
  return one_plus_u23(w) - 0x1.fffffep-1f;
}
// Uniform in [0,1): multiples of 2^-24, exact in fp32.
__device__ __forceinline__ float u01_half(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

// Box-Muller on the hardware transcendental units: v_log_f32 is log2,
// v_sin_f32 / v_cos_f32 take the angle in revolutions and are periodic in
// it, so an angle a in [0, 1) is passed as the float 1 + a and needs no range
// reduction.  z0 = r cos(2 pi a), z1 = r sin(2 pi a), r = sqrt(-2 ln u).
__device__ __forceinline__ void box_muller_ua(float u, float one_plus_a, float& z0, float& z1) {
  const float r = __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u));
  z0 = r * __builtin_amdgcn_cosf(one_plus_a);
  z1 = r * __builtin_amdgcn_sinf(one_plus_a);
}

// Six standard normals from one Philox block (w0, w1, w2, w3), three
// Box-Muller pairs: pair p takes its radius uniform from word p,
// u = ((w_p >> 8) | 1) 2^-24, and a 19-bit angle a = ((w_p mod 2^9) 2^10 +
// bits [10p, 10p + 10) of w3) 2^-19 from the 9 low bits of word p, which the
// radius does not use, and 10 bits of word 3 (126 of the block's 128 bits).
// z[2p] = r_p cos(2 pi a_p), z[2p + 1] = r_p sin(2 pi a_p).
constexpr int NORMALS_PER_BLOCK = 6;
__device__ __forceinline__ void normals6(const U4& w, float* z) {
  const uint32_t wp[3] = {w.x, w.y, w.z};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    // 1 + a as float bits: 0x3F800000 | (w_p mod 2^9) << 14 | (w3 field) << 4
    const uint32_t lo = (((w.w >> (10 * p)) & 0x3FFu) << 4) | 0x3F800000u;
    const uint32_t bits = ((wp[p] << 14) & 0x007FC000u) | lo;
    box_muller_ua(u01_open(wp[p]), __uint_as_float(bits), z[2 * p], z[2 * p + 1]);
  }
}

// Normals of block b of a path's Brownian stream.
__device__ __forceinline__ void normal_block(uint32_t b, uint32_t plo, uint32_t phi, uint32_t k0,
                                             uint32_t k1, float* z) {
  normals6(philox4x32_10(b, STREAM_BM, plo, phi, k0, k1), z);
}

// Steps (or adapted-grid rows) whose K normals each are drawn together as
// whole blocks: 4 when 4K normals fill whole blocks (K = 3), else 6.
__host__ __device__ constexpr int group_steps(int K) { return (4 * K) % NORMALS_PER_BLOCK == 0 ? 4 : 6; }

__device__ __forceinline__ float fast_exp(float v) {
  return __builtin_amdgcn_exp2f(v * 1.44269504088896341f);
}
__device__ __forceinline__ float fast_pow(float b, float e) {  // b > 0
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(b));
}
__device__ __forceinline__ float fast_ln(float v) {
  return 0.693147180559945309f * __builtin_amdgcn_logf(v);
}

// ln u for u = (w | 1) 2^-32 in (0,1), in fp32 from correctly rounded
// operations only (division, fma, multiply, add), so a CPU reproduces it bit
// for bit: u = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
// s = (m-1)/(m+1), by its odd series to s^9 (|s| < 0.1716, error < 1e-9).
__device__ __forceinline__ float ln_unit(uint32_t w) {
#pragma clang fp contract(off)
  const uint32_t v = w | 1u;
  const int z = __clz((int)v);
  const uint32_t top = v << z;
  float m = (float)(top >> 8) * 0x1p-23f;
  int e = -(z + 1);
  const bool hi = m > 1.41421356f;
  m = hi ? m * 0.5f : m;
  e = hi ? e + 1 : e;
  const float sn = m - 1.0f;
  const float sd = m + 1.0f;
  const float sq = __fdiv_rn(sn, sd);
  const float s2 = sq * sq;
  float p = __fmaf_rn(s2, 0.222222222f, 0.285714286f);
  p = __fmaf_rn(s2, p, 0.4f);
  p = __fmaf_rn(s2, p, 0.666666667f);
  p = __fmaf_rn(s2, p, 2.0f);
  const float lnm = sq * p;
  const float et = (float)e * 0.693147181f;
  return lnm + et;
}

// Normal number i of one path (i = step * K + k): normal i mod 6 of Philox
// block i / 6 on the Brownian stream (normals6).  The last block is cached,
// so sequential reads cost one Philox block per 6 normals.
struct NormalStream {
  uint32_t k0, k1, plo, phi;
  uint32_t blk;
  float z[NORMALS_PER_BLOCK];
  __device__ __forceinline__ void init(uint32_t key0, uint32_t key1, uint64_t path) {
    k0 = key0;
    k1 = key1;
    plo = (uint32_t)path;
    phi = (uint32_t)(path >> 32);
    blk = 0xffffffffu;
  }
  __device__ __forceinline__ float get(uint32_t i) {
    const uint32_t b = i / (uint32_t)NORMALS_PER_BLOCK;
    if (b != blk) {
      blk = b;
      normal_block(b, plo, phi, k0, k1, z);
    }
    const uint32_t s = i - b * (uint32_t)NORMALS_PER_BLOCK;
    float v = z[0];
#pragma unroll
    for (int q = 1; q < NORMALS_PER_BLOCK; ++q) v = s == (uint32_t)q ? z[q] : v;
    return v;
  }
};

// ---------------------------------------------------------------------------
// Model constants as the kernels read them: fp32 coefficients and fp64 jump
// clock constants, converted once on the host so every lane reads them as
// wave-uniform scalars from the kernel arguments (or the job table).
// ---------------------------------------------------------------------------
struct KModel {
  float x0[SDEMC_MAX_DIM], a0[SDEMC_MAX_DIM], a1[SDEMC_MAX_DIM];
  float b0[SDEMC_MAX_DIM], b1[SDEMC_MAX_DIM], c0[SDEMC_MAX_DIM], c1[SDEMC_MAX_DIM];
  float j0[SDEMC_MAX_DIM], j1[SDEMC_MAX_DIM];
  float L[SDEMC_MAX_DIM][SDEMC_MAX_DIM];
  float jm, js;  // lognormal: mean and sd of the log-jump
  float lv[11];  // Levy inverse-CDF constants, see levy_icdf
  int32_t exact, law, dim, noise;
  double rate;
  float inv_rate;  // 1 / lambda (fp32 jump clock)
  float hv_kappa, hv_drift, hv_xi;  // Heston variance: kappa, kappa theta - xi^2/2, xi
  float ax[SDEMC_MAX_DIM];          // drift of component i gains ax[i] x_0 ('diag')
  // 'general' noise: B(x)_ij = bm0[i][j] + bm1[i][j] x_i, zero outside
  // (dim, brown_dim); L is then the brown_dim x brown_dim factor, zero-padded
  float bm0[SDEMC_MAX_DIM][SDEMC_MAX_DIM], bm1[SDEMC_MAX_DIM][SDEMC_MAX_DIM];
};

struct KPayoff {
  int32_t kind, log;
  float strike, scale, disc, param;
};

// Everything one job needs in the kernels, plus derived grid constants.
struct KJob {
  KModel m;
  KPayoff pay;
  double T;
  float n_over_T;   // fp32(steps / T): jump arrival time -> step index
  float h;          // fp32(T / steps): fine step (reference: torch.tensor(T / N))
  float sqh;        // sqrt(h)
  float hc;         // coupled uniform grid: fp32(T * M / steps)
  float sqhc;
  int32_t steps, grid, M, max_steps;
  uint64_t n_paths, seed, path_offset;
  int64_t chunk_base;  // first chunk partial of this job (batched launches)
  uint32_t magic;      // KJOB_MAGIC once prepared on the host
  int32_t class_key;   // batched-kernel dispatch key, see class_key_of
  // per-step Poisson counts (uniform grid, Levy marks): step n has at least
  // k+1 jumps iff its uniform word w < pois_T[k]; the first mark is the
  // offset of w within its count's interval (see poisson_count)
  int32_t pois_mode;
  int32_t clock_win;  // steps per jump-queue fill window (uniform grid), per jump-slot refill (adapted)
  uint32_t pois_T[16];
  double pois_inv[16];  // reciprocal width of count c's interval at [c - 1]
  // Uniform-grid step form (uniform_step): UFORM_GENERIC, or the step's
  // constants folded on the host for coefficients that are all constant
  // (UFORM_CONST: x += uA + uW z + uC z_D) or all proportional to the state
  // (UFORM_PROP: x += x (uA + uW z + uC z_D)), with uA = a h, uW = diag(b)
  // L sqrt(h) (lower triangular), uC = c sqrt(h).
  int32_t uform;
  int32_t additive;  // every jump coefficient constant (j1 = 0): x += j0 J
  float uA[SDEMC_MAX_DIM], uC[SDEMC_MAX_DIM];
  float uW[SDEMC_MAX_DIM][SDEMC_MAX_DIM];
};

enum { UFORM_GENERIC = 0, UFORM_CONST = 1, UFORM_PROP = 2 };

// Thresholds of the per-step jump count: pois_T[k] = round(P(N >= k+1) 2^32)
// for N ~ Poisson(mu), mu = rate * T / steps; the tail sums run from the top
// so small tails keep full relative precision.  inv[c-1] = 1 / (T[c-1] - T[c])
// (T[16] = 0), 0 for an empty interval.  Returns whether counts above 16 are
// negligible (P(N > 16) < 2^-33).
inline bool poisson_thresholds(double mu, uint32_t* T16, double* inv) {
#pragma clang fp contract(off)
  double p[64];
  p[0] = exp(-mu);
  for (int i = 1; i < 64; ++i) p[i] = p[i - 1] * mu / (double)i;
  double tail[65];
  tail[64] = 0.0;
  for (int i = 63; i >= 0; --i) tail[i] = tail[i + 1] + p[i];
  for (int k = 1; k <= 16; ++k) {
    const double v = tail[k] * 4294967296.0;
    T16[k - 1] = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)(v + 0.5);
  }
  for (int k = 0; k < 16; ++k) {
    const uint32_t width = T16[k] - (k < 15 ? T16[k + 1] : 0u);
    inv[k] = width ? 1.0 / (double)width : 0.0;
  }
  return tail[17] * 4294967296.0 < 0.5;
}

// Per-step Poisson counts pay one count word per step however rare the
// jumps; the arrival clock pays a Philox block per jump and its fill passes
// per window.  Levy-law jobs on the uniform grid take the counts from this
// many expected jumps per step on (measured crossover near 0.07 on the 2-D
// rainbow), the clock below it.
constexpr double POIS_MIN_MU = 0.0625;

constexpr uint32_t KJOB_MAGIC = 0x5DE3C0DEu;

// Dispatch key of a prepared job: ((noise * 4 + dim - 1) * 2 + grid) * 2 + coupled.
inline int32_t class_key_of(int noise, int dim, int grid, bool coupled) {
  return ((noise * 4 + (dim - 1)) * 2 + grid) * 2 + (coupled ? 1 : 0);
}

inline void make_kjob(const sdemc_job& j, KJob& k) {
  const sdemc_model& m = j.model;
  KModel& f = k.m;
  for (int i = 0; i < SDEMC_MAX_DIM; ++i) {
    f.x0[i] = (float)m.x0[i];
    f.a0[i] = (float)m.a0[i];
    f.a1[i] = (float)m.a1[i];
    f.b0[i] = (float)m.b0[i];
    f.b1[i] = (float)m.b1[i];
    f.c0[i] = (float)m.c0[i];
    f.c1[i] = (float)m.c1[i];
    f.j0[i] = (float)m.j0[i];
    f.j1[i] = (float)m.j1[i];
    for (int q = 0; q < SDEMC_MAX_DIM; ++q) f.L[i][q] = (float)m.chol[i * SDEMC_MAX_DIM + q];
    f.ax[i] = (float)m.ax[i];
  }
  const bool general = m.noise == SDEMC_NOISE_GENERAL;
  for (int i = 0; i < SDEMC_MAX_DIM; ++i) {
    for (int q = 0; q < SDEMC_MAX_DIM; ++q) {
      const bool in_b = general && i < m.dim && q < m.brown_dim;
      f.bm0[i][q] = in_b ? (float)m.bm0[i * SDEMC_MAX_DIM + q] : 0.0f;
      f.bm1[i][q] = in_b ? (float)m.bm1[i * SDEMC_MAX_DIM + q] : 0.0f;
      if (general && (i >= m.brown_dim || q >= m.brown_dim)) f.L[i][q] = 0.0f;
    }
  }
  f.exact = m.exact_jumps;
  f.law = m.jump_law;
  f.dim = m.dim;
  f.noise = m.noise;
  f.jm = (float)m.jp[0];
  f.js = (float)m.jp[1];
  for (int q = 0; q < 11; ++q) f.lv[q] = 0.0f;
  if (m.jump_law == SDEMC_JUMP_LEVY) {
    // jp = {cm, cp, 1/mu, alpha, eps, lambda, y1, y2, y3, eps^-alpha, -1/alpha}
    const double cm = m.jp[0], cp = m.jp[1], imu = m.jp[2], al = m.jp[3];
    const double lda = m.jp[5], epa = m.jp[9];
    f.lv[0] = (float)m.jp[6];
    f.lv[1] = (float)m.jp[7];
    f.lv[2] = (float)m.jp[8];
    f.lv[3] = (float)(imu > 0.0 ? lda / (cm * imu) : 0.0);  // mu*lda/cm
    f.lv[4] = (float)imu;                                   // 1/mu
    f.lv[5] = (float)(al * lda / cm);                       // alpha*lda/cm
    f.lv[6] = (float)(1.0 - al * imu);                      // 1 - alpha/mu
    f.lv[7] = (float)(-al / cp * lda);                      // -alpha*lda/cp
    f.lv[8] = (float)((-al / cp) * (-cm * imu - cm * (epa - 1.0) / al) + epa);
    f.lv[9] = (float)(imu > 0.0 ? lda / (cp * imu) : 0.0);  // mu*lda/cp
    f.lv[10] = (float)m.jp[10];                             // -1/alpha
  }
  f.rate = m.rate;
  f.inv_rate = m.rate > 0.0 ? (float)(1.0 / m.rate) : 0.0f;
  f.hv_kappa = f.hv_drift = f.hv_xi = 0.0f;
  if (m.noise == SDEMC_NOISE_HESTON) {
    f.hv_kappa = (float)m.jp[0];
    f.hv_drift = (float)(m.jp[0] * m.jp[1] - 0.5 * m.jp[2] * m.jp[2]);
    f.hv_xi = (float)m.jp[2];
  }
  k.pay.kind = j.payoff.kind;
  k.pay.log = j.payoff.log;
  k.pay.strike = (float)j.payoff.strike;
  k.pay.scale = (float)j.payoff.scale;
  k.pay.disc = (float)j.payoff.disc;
  k.pay.param = (float)j.payoff.param;
  k.T = j.T;
  k.steps = j.steps;
  k.grid = j.grid;
  k.M = j.coarse_factor;
  k.max_steps = j.max_steps;
  k.n_over_T = (float)((double)j.steps / j.T);
  k.h = (float)(j.T / (double)j.steps);
  k.sqh = sqrtf(k.h);
  if (j.coarse_factor > 0) {
    k.hc = (float)(j.T * (double)j.coarse_factor / (double)j.steps);
    k.sqhc = sqrtf(k.hc);
  } else {
    k.hc = k.h;
    k.sqhc = k.sqh;
  }
  k.n_paths = j.n_paths;
  k.seed = j.seed;
  k.path_offset = j.path_offset;
  k.chunk_base = 0;
  k.magic = KJOB_MAGIC;
  k.class_key = class_key_of(m.noise, m.dim, j.grid, j.coarse_factor > 0);
  k.pois_mode = 0;
  // A window's fill passes run when any lane of the wavefront jumps in it, so
  // longer windows mean fewer passes per jump (the wavefront's passes follow
  // the lane with the most jumps), until a jump beyond the queue (drawn in
  // the step) becomes common.  Uniform grid, eight-slot queue in LDS: about
  // four jumps per path per window measured best on the served mix (round
  // 42.7 / 41.9 / 41.6 / 41.5 ms at one / two / three / four).  Adapted grid:
  // its two jump slots are refilled every clock_win iterations, about half a
  // jump per path apart.  The step loop rounds the window up to its own
  // stride.
  {
    const double per_step = m.rate * j.T / (double)j.steps;
    const bool uniform = j.grid == SDEMC_GRID_UNIFORM;
    const double per_window = uniform ? 4.0 : 0.5;
    int32_t w = uniform ? 512 : 128;
    while (w > 4 && per_step * (double)w > per_window) w >>= 1;
    k.clock_win = w;
  }
  for (int q = 0; q < 16; ++q) {
    k.pois_T[q] = 0;
    k.pois_inv[q] = 0.0;
  }
  if (m.jump_law == SDEMC_JUMP_LEVY && j.grid == SDEMC_GRID_UNIFORM) {
    const double mu = m.rate * j.T / (double)j.steps;
    if (mu >= POIS_MIN_MU) k.pois_mode = poisson_thresholds(mu, k.pois_T, k.pois_inv) ? 1 : 0;
  }
  // Folded uniform-grid step constants ('diag' and the jump solver's 'indep'
  // noise, no cross drift): every coefficient constant or every one
  // proportional to the state.
  k.uform = UFORM_GENERIC;
  k.additive = 1;
  for (int i = 0; i < m.dim; ++i)
    if (m.j1[i] != 0.0) k.additive = 0;
  for (int i = 0; i < SDEMC_MAX_DIM; ++i) {
    k.uA[i] = k.uC[i] = 0.0f;
    for (int q = 0; q < SDEMC_MAX_DIM; ++q) k.uW[i][q] = 0.0f;
  }
  if (m.noise == SDEMC_NOISE_DIAG || m.noise == SDEMC_NOISE_INDEP_COMMON) {
    bool cst = true, prop = true;
    for (int i = 0; i < m.dim; ++i) {
      if (m.ax[i] != 0.0) cst = prop = false;
      if (m.a1[i] != 0.0 || m.b1[i] != 0.0 || m.c1[i] != 0.0) cst = false;
      if (m.a0[i] != 0.0 || m.b0[i] != 0.0 || m.c0[i] != 0.0) prop = false;
    }
    if (cst || prop) {
      k.uform = cst ? UFORM_CONST : UFORM_PROP;
      const double h = (double)k.h, sq = sqrt(h);
      for (int i = 0; i < m.dim; ++i) {
        const double a = cst ? m.a0[i] : m.a1[i];
        const double b = cst ? m.b0[i] : m.b1[i];
        const double c = cst ? m.c0[i] : m.c1[i];
        k.uA[i] = (float)(a * h);
        k.uC[i] = m.noise == SDEMC_NOISE_INDEP_COMMON ? (float)(c * sq) : 0.0f;
        for (int q = 0; q <= i; ++q) k.uW[i][q] = (float)(b * m.chol[i * SDEMC_MAX_DIM + q] * sq);
      }
    }
  }
}

// Whether a job's uniform-grid jumps come from per-step Poisson counts (the
// pois_mode make_kjob sets) rather than the arrival clock.
inline bool job_pois_mode(const sdemc_job& j) {
  if (j.model.jump_law != SDEMC_JUMP_LEVY || j.grid != SDEMC_GRID_UNIFORM) return false;
  const double mu = j.model.rate * j.T / (double)j.steps;
  if (!(mu >= POIS_MIN_MU)) return false;
  uint32_t T[16];
  double inv[16];
  return poisson_thresholds(mu, T, inv);
}

// Inverse CDF of the Levy jump law of levy.py:22-30 on y in (0,1):
//   y <= y1 : 1 - ... tail branch  x1 = ln(mu*lda*y/cm)/mu - 1
//   y <  y2 : x2 = -(alpha*(lda*y/cm - 1/mu) + 1)^(-1/alpha)
//   y <  y3 : x3 = ((-alpha/cp)*(lda*y - cm/mu - cm*(eps^-a - 1)/a) + eps^-a)^(-1/a)
//   else    : x4 = 1 - ln(mu*lda*(1-y)/cp)/mu
// Branch selection compares the fp32 uniform with fp32 thresholds, as the
// reference's torch.where on fp32 tensors does.
// All four branches share one logarithm (fast_pow is exp2 of e log2 b and
// fast_ln is ln 2 log2 v), so the function is straight-line code: the lanes
// of a wavefront that land in different branches cost one evaluation, and
// every value is the one the branch would compute.
__device__ __forceinline__ float levy_icdf(const KModel& f, float y) {
  const bool t1 = y <= f.lv[0];
  const bool t2 = !t1 && y < f.lv[1];
  const bool t3 = !t1 && !t2 && y < f.lv[2];
  const float arg = t1 ? f.lv[3] * y
                       : (t2 ? __fmaf_rn(f.lv[5], y, f.lv[6])
                             : (t3 ? __fmaf_rn(f.lv[7], y, f.lv[8]) : f.lv[9] * (1.0f - y)));
  const float lg = __builtin_amdgcn_logf(arg);
  const float pw = __builtin_amdgcn_exp2f(f.lv[10] * lg);
  const float ln = 0.693147180559945309f * lg;
  return t1 ? __fmaf_rn(ln, f.lv[4], -1.0f)
            : (t2 ? -pw : (t3 ? pw : __fmaf_rn(-f.lv[4], ln, 1.0f)));
}

// Jump mark from two Philox words.
__device__ __forceinline__ float jump_mark(const KModel& f, uint32_t m1, uint32_t m2) {
  if (f.law == SDEMC_JUMP_LEVY) {
    // torch.rand + UNIFORM_TOL / 3 (levy.py:7, 86), in fp32
    const float y = u01_half(m1) + 0x1.555556p-26f;
    return levy_icdf(f, y);
  }
  const float u = u01_open(m1);
  const float a = one_plus_u23(m2);  // the angle (m2 >> 9) 2^-23 in revolutions, plus 1
  const float z = __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u)) *
                  __builtin_amdgcn_cosf(a);
  return fast_exp(__fmaf_rn(f.js, z, f.jm)) - 1.0f;
}

// ---------------------------------------------------------------------------
// Payoff on the terminal state (options.py:167-176): spot = scale * (exp(x)
// if log else x), payoff(spot) * disc.
// ---------------------------------------------------------------------------
// Undiscounted payoff; the caller applies the discount factor, as
// mc.py:93 and mlmc.py:64-65 multiply after the payoff.
template <int D>
__device__ __forceinline__ float payoff_raw(const KPayoff& p, const float (&x)[D]) {
  float s[D];
#pragma unroll
  for (int i = 0; i < D; ++i) s[i] = p.scale * (p.log ? fast_exp(x[i]) : x[i]);
  const float K = p.strike;
  float v;
  switch (p.kind) {
    case SDEMC_PAYOFF_EURO_CALL: v = s[0] > K ? s[0] - K : 0.0f; break;
    case SDEMC_PAYOFF_EURO_PUT: v = s[0] < K ? K - s[0] : 0.0f; break;
    case SDEMC_PAYOFF_BINARY_AON: v = s[0] >= K ? s[0] : 0.0f; break;
    case SDEMC_PAYOFF_DIGITAL: v = s[0] > K ? 1.0f : 0.0f; break;
    case SDEMC_PAYOFF_BASKET_ARITH: {
      float m = 0.0f;
#pragma unroll
      for (int i = 0; i < D; ++i) m += s[i];
      m = m / (float)D;
      v = m > K ? m - K : 0.0f;
      break;
    }
    case SDEMC_PAYOFF_BASKET_GEO: {
      float m = 0.0f;
#pragma unroll
      for (int i = 0; i < D; ++i) m += fast_ln(s[i]);
      m = fast_exp(m / (float)D);
      v = m > K ? m - K : 0.0f;
      break;
    }
    case SDEMC_PAYOFF_RAINBOW: {
      float m = s[0];
#pragma unroll
      for (int i = 1; i < D; ++i) m = fmaxf(m, s[i]);
      v = m > K ? m - K : 0.0f;
      break;
    }
    case SDEMC_PAYOFF_BEST_OF: {
      float m = s[0];
#pragma unroll
      for (int i = 1; i < D; ++i) m = fmaxf(m, s[i]);
      v = fmaxf(m, K);
      break;
    }
    case SDEMC_PAYOFF_ASIAN_CALL: {
      // the running integral in component 1 over the averaging time, after
      // the transform above (AsianCall.payoff on Option.transform's output)
      float a = __fdiv_rn(s[D > 1 ? 1 : 0], p.param);
      if (p.log) a = fast_exp(a);
      v = a > K ? a - K : 0.0f;
      break;
    }
    case SDEMC_PAYOFF_RAINBOW_EVEN: {
      float m = s[0];
#pragma unroll
      for (int i = 2; i < D; i += 2) m = fmaxf(m, s[i]);
      v = m > K ? m - K : 0.0f;
      break;
    }
    default: v = s[0]; break;
  }
  return v;
}

// ---------------------------------------------------------------------------
// Euler step (schemes.py:5-10) for one component and the jump update.
// ---------------------------------------------------------------------------
template <int D, int FAM>
__device__ __forceinline__ void euler(const KModel& f, float (&x)[D], float dt,
                                      const float (&dW)[D], const float (&dB)[D]) {
  if constexpr (FAM == FAM_HESTON) {
    // x = (S, v): S by Euler with the pre-step sqrt(v); sqrt of the new v is
    // the positive root of (1 + kappa dt) y^2 - xi dW_1 y - (v + d dt) = 0,
    // y = (B + sqrt(B^2 + 4 A C)) / (2 A) (schemes.py:16-21, helpers.py:36-48)
    static_assert(D == 2, "Heston state is (S, v)");
    const float S = x[0], v = x[1];
    const float a = f.a1[0] * S;
    const float b = __fsqrt_rn(v) * S;
    x[0] = __fmaf_rn(b, dW[0], __fmaf_rn(a, dt, S));
    const float A = __fmaf_rn(f.hv_kappa, dt, 1.0f);
    const float C = __fmaf_rn(f.hv_drift, dt, v);
    const float B = f.hv_xi * dW[1];
    const float disc = __fmaf_rn(B, B, 4.0f * A * C);
    const float y = __fdiv_rn(B + __fsqrt_rn(disc), 2.0f * A);
    x[1] = y * y;
    return;
  }
  if constexpr (FAM == FAM_GENERAL) {
    // dW = bm0 w, dB = bm1 w (general_apply), so row i of B(x) w is
    // dW_i + x_i dB_i with x_i the pre-step state
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const float xi = x[i];
      const float a = __fmaf_rn(f.a1[i], xi, f.a0[i]);
      const float t = __fmaf_rn(a, dt, xi);
      x[i] = t + __fmaf_rn(xi, dB[i], dW[i]);
    }
    return;
  }
  const float s0 = x[0];  // pre-step x_0 for the cross drift ax[i] x_0
#pragma unroll
  for (int i = 0; i < D; ++i) {
    float a = __fmaf_rn(f.a1[i], x[i], f.a0[i]);
    if (FAM == FAM_MERTON && i > 0) a = __fmaf_rn(f.ax[i], s0, a);
    const float b = __fmaf_rn(f.b1[i], x[i], f.b0[i]);
    const float t = __fmaf_rn(a, dt, x[i]);
    if (FAM == FAM_MERTON) {
      x[i] = __fmaf_rn(b, dW[i], t);
    } else {
      const float c = __fmaf_rn(f.c1[i], x[i], f.c0[i]);
      x[i] = t + __fmaf_rn(c, dB[i], b * dW[i]);
    }
  }
}

// x += g(y) * J with g(y) = j0 + j1 * y (sde.py:374-375, levy.py:119-120).
template <int D>
__device__ __forceinline__ void add_jump(const KModel& f, float (&x)[D], const float (&y)[D],
                                         float J) {
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = __fmaf_rn(__fmaf_rn(f.j1[i], y[i], f.j0[i]), J, x[i]);
}

// Number of normals drawn per step.
template <int D, int FAM>
struct NK {
  static constexpr int K = FAM == FAM_GENERAL
                               ? GENERAL_K
                               : (DiagNoise<FAM>::value ? D : (FAM == FAM_LEVY ? D + 1 : 2 * D));
};

// 'general' noise: the correlated increments w = L (z sqrt(dt)) of the
// Brownian motion (solvers.py:51-56), all GENERAL_K of them (zero beyond
// brown_dim), ...
__device__ __forceinline__ void general_w(const KModel& f, const float* z, float sq,
                                          float (&w)[GENERAL_K]) {
  float zs[GENERAL_K];
#pragma unroll
  for (int q = 0; q < GENERAL_K; ++q) zs[q] = z[q] * sq;
#pragma unroll
  for (int i = 0; i < GENERAL_K; ++i) {
    float acc = f.L[i][0] * zs[0];
#pragma unroll
    for (int q = 1; q <= i; ++q) acc = __fmaf_rn(f.L[i][q], zs[q], acc);
    w[i] = acc;
  }
}

// ... and the two products the Euler step needs: e = bm0 w and g = bm1 w,
// with which B(x) w = e + diag(x) g.
template <int D>
__device__ __forceinline__ void general_apply(const KModel& f, const float (&w)[GENERAL_K],
                                              float (&e)[D], float (&g)[D]) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    float a = f.bm0[i][0] * w[0], b = f.bm1[i][0] * w[0];
#pragma unroll
    for (int q = 1; q < GENERAL_K; ++q) {
      a = __fmaf_rn(f.bm0[i][q], w[q], a);
      b = __fmaf_rn(f.bm1[i][q], w[q], b);
    }
    e[i] = a;
    g[i] = b;
  }
}

// Correlated, scaled increments from K standard normals:
//   dW = L (z[0:D] sqrt(dt)),  dB = z[D] sqrt(dt) (common) or L (z[D:2D] sqrt(dt)).
template <int D, int FAM>
__device__ __forceinline__ void correlate(const KModel& f, const float* z, float sq,
                                          float (&dW)[D], float (&dB)[D]) {
  if constexpr (FAM == FAM_GENERAL) {
    float w[GENERAL_K];
    general_w(f, z, sq, w);
    general_apply<D>(f, w, dW, dB);
    return;
  }
  float zs[D];
#pragma unroll
  for (int j = 0; j < D; ++j) zs[j] = z[j] * sq;
  if (D == 1) {
    dW[0] = zs[0];
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float acc = f.L[i][0] * zs[0];
#pragma unroll
      for (int j = 1; j <= i; ++j) acc = __fmaf_rn(f.L[i][j], zs[j], acc);
      dW[i] = acc;
    }
  }
  if (FAM == FAM_LEVY) {
    const float c = z[D] * sq;
#pragma unroll
    for (int i = 0; i < D; ++i) dB[i] = c;
  } else if (FAM == FAM_DOUBLE) {
    float zb[D];
#pragma unroll
    for (int j = 0; j < D; ++j) zb[j] = z[D + j] * sq;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float acc = f.L[i][0] * zb[0];
#pragma unroll
      for (int j = 1; j <= i; ++j) acc = __fmaf_rn(f.L[i][j], zb[j], acc);
      dB[i] = acc;
    }
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) dB[i] = 0.0f;
  }
}

// One Euler step of a uniform grid of step h = j.h from the step's K standard
// normals z (the diffusion part; the caller adds the jumps).  Jobs with folded
// constants (KJob::uform) take 3-5 operations per component instead of the
// generic correlate + euler: x_i + uA_i + sum_k uW_ik z_k + uC_i z_D, or
// x_i + x_i (uA_i + sum_k uW_ik z_k + uC_i z_D) -- the same Euler step
// (schemes.py:5-10) with the products of the constant factors taken once on
// the host.  Every kernel that steps a uniform grid from Philox normals
// calls this, so their paths stay bit-identical.
template <int D, int FAM>
__device__ __forceinline__ void uniform_step(const KJob& j, float (&x)[D], const float* z) {
  const KModel& f = j.m;
  if constexpr (FAM == FAM_MERTON || FAM == FAM_LEVY) {
    if (j.uform == UFORM_CONST) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        float t = x[i] + j.uA[i];
        if (FAM == FAM_LEVY) t = __fmaf_rn(j.uC[i], z[D], t);
#pragma unroll
        for (int q = 0; q <= i; ++q) t = __fmaf_rn(j.uW[i][q], z[q], t);
        x[i] = t;
      }
      return;
    }
    if (j.uform == UFORM_PROP) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        float g = __fmaf_rn(j.uW[i][0], z[0], j.uA[i]);
#pragma unroll
        for (int q = 1; q <= i; ++q) g = __fmaf_rn(j.uW[i][q], z[q], g);
        if (FAM == FAM_LEVY) g = __fmaf_rn(j.uC[i], z[D], g);
        x[i] = __fmaf_rn(x[i], g, x[i]);
      }
      return;
    }
  }
  float dW[D], dB[D];
  correlate<D, FAM>(f, z, j.sqh, dW, dB);
  euler<D, FAM>(f, x, j.h, dW, dB);
}

// ---------------------------------------------------------------------------
// Jump arrivals drawn in the kernel.  Jump j (j = 0, 1, ...) takes Philox
// block j of the jump stream: word x gives the exponential gap to it,
// words y, z its mark.  Arrival times accumulate in fp64.
// ---------------------------------------------------------------------------
struct JumpClock {
  float tau;    // arrival time of jump j (fp32, accumulated gap by gap)
  uint32_t j;   // index of the pending jump
  uint32_t m1, m2;
  __device__ __forceinline__ void draw(uint32_t k0, uint32_t k1, uint32_t plo, uint32_t phi,
                                       float inv_rate) {
#pragma clang fp contract(off)
    const U4 w = philox4x32_10(j, STREAM_JUMP, plo, phi, k0, k1);
    const float gap = -ln_unit(w.x) * inv_rate;
    tau = tau + gap;
    m1 = w.y;
    m2 = w.z;
  }
  __device__ __forceinline__ void init(uint32_t k0, uint32_t k1, uint32_t plo, uint32_t phi,
                                       double rate, float inv_rate) {
    j = 0;
    tau = 0.0f;
    if (rate > 0.0) {
      draw(k0, k1, plo, phi, inv_rate);
    } else {
      tau = __builtin_inff();
      m1 = m2 = 0;
    }
  }
};

// Step index (grid of n steps over [0,T]) in which time tau falls.
__device__ __forceinline__ int32_t step_of(float tau, float n_over_T, int32_t n) {
#pragma clang fp contract(off)
  const float q = tau * n_over_T;
  return q < (float)n ? (int32_t)q : 0x7fffffff;
}

// Count of uniform-grid step n in per-step Poisson mode: the count is c iff
// pois_T[c] <= w < pois_T[c-1] (pois_T[16] = 0), and given c the offset of
// w in that interval is uniform and independent of c, so it yields the
// first mark's uniform: y = (w - lo + 1/2) / (hi - lo), fp64, one rounding.
// Count c of a step with w < T[0]: T[c] <= w < T[c-1]; y is the first mark's
// uniform, the offset of w inside that interval times the interval's
// reciprocal width.  Counts 1 and 2 use values the compiler keeps in
// registers; larger counts (rare for every wavefront) scan the rest of the
// table in a loop kept rolled, so its entries are read from memory when
// needed instead of occupying scalar registers for the whole kernel.
__device__ __forceinline__ int poisson_count(const KJob& j, uint32_t w, float& y) {
  int c;
  uint32_t lo;
  double inv;
  if (w >= j.pois_T[1]) {
    c = 1;
    lo = j.pois_T[1];
    inv = j.pois_inv[0];
  } else if (w >= j.pois_T[2]) {
    c = 2;
    lo = j.pois_T[2];
    inv = j.pois_inv[1];
  } else {
    c = 3;
#pragma nounroll
    while (c < 16 && w < j.pois_T[c]) ++c;
    lo = c < 16 ? j.pois_T[c] : 0u;
    inv = j.pois_inv[c - 1];
  }
  const double r = ((double)(w - lo) + 0.5) * inv;
  y = fminf((float)r, 0x1.fffffep-1f);
  return c;
}

// Uniform of mark k >= 1 of step n: word (k-1) mod 4 of block n of the mark
// stream (stream index 3 + 16 floor((k-1)/4)).
__device__ __forceinline__ float poisson_mark_uniform(uint32_t k0, uint32_t k1, uint32_t plo,
                                                      uint32_t phi, int32_t n, int k) {
  const uint32_t q = (uint32_t)(k - 1);
  const U4 mw = philox4x32_10((uint32_t)n, STREAM_MARK + 16u * (q >> 2), plo, phi, k0, k1);
  const uint32_t sel = q & 3u;
  return u01_open(sel == 0 ? mw.x : (sel == 1 ? mw.y : (sel == 2 ? mw.z : mw.w)));
}

// Jumps of uniform-grid step n in per-step Poisson mode, given the step's
// count word w (w < pois_T[0]): marks are added with coefficients at xo
// (pre-step state) or at the running state (exact jumps).  Returns the sum
// of the marks.
template <int D>
__device__ __forceinline__ float poisson_jumps(const KJob& j, uint32_t k0, uint32_t k1,
                                               uint32_t plo, uint32_t phi, uint32_t w,
                                               int32_t n, float (&x)[D],
                                               const float (&xo)[D], int k_first = 0) {
  const KModel& f = j.m;
  float y;
  const int c = poisson_count(j, w, y);
  float jsum = 0.0f;
  for (int k = k_first; k < c; ++k) {
    if (k > 0) y = poisson_mark_uniform(k0, k1, plo, phi, n, k);
    const float J = levy_icdf(f, y);
    if (f.exact) {
      float yy[D];
#pragma unroll
      for (int i = 0; i < D; ++i) yy[i] = x[i];
      add_jump(f, x, yy, J);
    } else {
      add_jump(f, x, xo, J);
    }
    jsum += J;
  }
  return jsum;
}

// ---------------------------------------------------------------------------
// Per-step Poisson jumps of a window of eight uniform-grid steps, drawn by
// the whole wavefront from a list in LDS.  Each lane posts its jumping steps
// (count word below pois_T[0]); the lanes then take the list's items in
// turn, so a pass of the draw code serves up to 64 jumping steps whichever
// lanes they belong to, instead of one step per lane.  An item's first mark
// comes from its count word, marks 2, 3, ... from the mark stream; the owner
// reads them back at its step and applies them in the per-step form's order,
// so results are those of drawing in the step loop.  Eight steps per window
// (two 4-word count blocks) fill the passes better than four and halve the
// fixed cost of the list per step.
// ---------------------------------------------------------------------------
constexpr int POIS_WIN = 8;         // steps per window at most (8 for 4-step groups, else 6)
constexpr int POIS_XCAP = 128;      // marks 2, 3, ... held per window
constexpr uint32_t POIS_NOX = 0xFFu;  // item's further marks drawn by its owner

struct PoisLds {
  uint32_t v[POIS_WIN * WAVE];   // item: count word in, first mark (float bits) out
  uint16_t own[POIS_WIN * WAVE];  // owning lane * POIS_WIN + step within the window
  uint8_t cnt[POIS_WIN * WAVE];  // item's jump count
  uint8_t xb[POIS_WIN * WAVE];   // slot of the item's second mark in x, or POIS_NOX
  float x[POIS_XCAP];
};

// Arrival-clock jump queue of a wavefront (uniform grid): slot q of lane l
// holds the step and the mark of the lane's q-th queued jump of the current
// fill window, k = -1 past the last one.  It takes the space of PoisLds (a
// kernel runs one jump mode); lane-contiguous, so every access is
// conflict-free and each lane reads back only what it wrote itself.
constexpr int CLOCK_JQ = 8;
struct ClockLds {
  int32_t k[CLOCK_JQ][WAVE];
  float J[CLOCK_JQ][WAVE];
};
static_assert(sizeof(ClockLds) <= sizeof(PoisLds), "the clock queue shares the jump-list space");

// One PoisLds per wavefront of the workgroup, in the launch's dynamic LDS
// (the launchers size it as waves per block * sizeof(PoisLds)).
extern __shared__ PoisLds pois_lds_dyn[];

__device__ __forceinline__ PoisLds& pois_lds() { return pois_lds_dyn[threadIdx.x / WAVE]; }

// LDS written by some lanes of the wavefront, then read by others: the
// wavefront's LDS operations complete in issue order, so only the compiler
// has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive prefix sum over the active lanes of v < 2^BITS, and the total.
template <int BITS>
__device__ __forceinline__ uint32_t wave_prefix(uint32_t v, uint32_t& total) {
  uint32_t pre = 0u, tot = 0u;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint64_t m = __ballot((v >> b) & 1u);
    pre += lanes_below(m) << b;
    tot += (uint32_t)__builtin_popcountll(m) << b;
  }
  total = tot;
  return pre;
}

// Draws the jumps of window n0 (steps n0 .. n0 + 7) for the whole wavefront.
// jmask: the lane's jumping steps, w: its count words.  Returns the lane's
// first item; the item of step s is that plus the number of jumping steps
// before s.  Must be called by every lane of the wavefront together.
template <int WIN>
__device__ __forceinline__ uint32_t pois_draw_window(const KJob& j, uint32_t k0, uint32_t k1,
                                                     uint32_t plo, uint32_t phi, int32_t n0,
                                                     const uint32_t (&w)[WIN],
                                                     uint32_t jmask, PoisLds& L) {
  static_assert(WIN <= POIS_WIN, "window larger than the item list");
  uint32_t total;
  const uint32_t first = wave_prefix<4>((uint32_t)__builtin_popcount(jmask), total);
  if (total == 0u) return first;
  const uint32_t lane = __lane_id();
  wave_lds_sync();  // the previous window's items have been read
  {
    uint32_t it = first;
#pragma unroll
    for (int s = 0; s < WIN; ++s) {
      if ((jmask >> s) & 1u) {
        L.v[it] = w[s];
        L.own[it] = (uint16_t)(lane * (uint32_t)WIN + (uint32_t)s);
        ++it;
      }
    }
  }
  wave_lds_sync();
  const uint64_t act = __ballot(1);
  const uint32_t rank = lanes_below(act);
  const uint32_t nact = (uint32_t)__builtin_popcountll(act);
  uint32_t xused = 0u;
  for (uint32_t b = 0u; b < total; b += nact) {
    const uint32_t i = b + rank;
    const bool have = i < total;
    const uint32_t ii = have ? i : 0u;
    const uint32_t wi = L.v[ii];
    const uint32_t ow = L.own[ii];
    int cnt = 0;
    float J = 0.0f;
    if (have) {
      float y;
      cnt = poisson_count(j, wi, y);
      J = levy_icdf(j.m, y);
    }
    const uint32_t e = have ? (uint32_t)(cnt - 1) : 0u;  // further marks, < 16
    uint32_t xb = POIS_NOX;
    if (j.additive) {
      // constant jump coefficients: the step adds j0 times the sum of its
      // marks, so the item's lane draws the further marks itself and hands
      // over one total (count 1)
      if (__any(e != 0u)) {
        const uint32_t olo = __shfl(plo, (int)(ow / (uint32_t)WIN));
        const uint32_t ohi = __shfl(phi, (int)(ow / (uint32_t)WIN));
        const uint32_t n = (uint32_t)n0 + ow % (uint32_t)WIN;
        U4 mw = {0, 0, 0, 0};
        for (uint32_t q = 0u; q < e; ++q) {
          if ((q & 3u) == 0u) mw = philox4x32_10(n, STREAM_MARK + 16u * (q >> 2), olo, ohi, k0, k1);
          const uint32_t sel = q & 3u;
          const uint32_t u = sel == 0 ? mw.x : (sel == 1 ? mw.y : (sel == 2 ? mw.z : mw.w));
          J += levy_icdf(j.m, u01_open(u));
        }
      }
      if (have) {
        L.v[ii] = __float_as_uint(J);
        L.cnt[ii] = (uint8_t)1;
        L.xb[ii] = (uint8_t)POIS_NOX;
      }
      continue;
    }
    if (__any(e != 0u)) {
      uint32_t etot;
      const uint32_t xo = xused + wave_prefix<4>(e, etot);
      xused += etot;
      const uint32_t olo = __shfl(plo, (int)(ow / (uint32_t)WIN));
      const uint32_t ohi = __shfl(phi, (int)(ow / (uint32_t)WIN));
      if (e != 0u && xo + e <= (uint32_t)POIS_XCAP) {
        xb = xo;
        const uint32_t n = (uint32_t)n0 + ow % (uint32_t)WIN;
        U4 mw = {0, 0, 0, 0};
        for (uint32_t q = 0u; q < e; ++q) {
          // mark q + 2 of the step: word q mod 4 of mark-stream block q / 4
          if ((q & 3u) == 0u) mw = philox4x32_10(n, STREAM_MARK + 16u * (q >> 2), olo, ohi, k0, k1);
          const uint32_t sel = q & 3u;
          const uint32_t u = sel == 0 ? mw.x : (sel == 1 ? mw.y : (sel == 2 ? mw.z : mw.w));
          L.x[xo + q] = levy_icdf(j.m, u01_open(u));
        }
      }
    }
    if (have) {
      L.v[ii] = __float_as_uint(J);
      L.cnt[ii] = (uint8_t)cnt;
      L.xb[ii] = (uint8_t)xb;
    }
  }
  wave_lds_sync();
  return first;
}

// ---------------------------------------------------------------------------
// Chunk reduction: fp64 butterfly over the 64 lanes, lane 0 stores.
// ---------------------------------------------------------------------------
// SPARSE_FLAG: the flag is written only when set (the batched launch, whose
// flag slots start at zero and are cleared again by the finalisation that
// reads them), so a chunk's store is its 16-byte partial alone.
template <bool SPARSE_FLAG = false>
__device__ __forceinline__ void chunk_store(double s, double q, int flag, double* partial2,
                                            int32_t* cflag) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    q += __shfl_xor(q, off);
  }
  // flags are bits (SDEMC_FLAG_OVERFLOW, SDEMC_FLAG_JUMP_CSR): OR over the lanes
  const int any_flag = (__any(flag & SDEMC_FLAG_OVERFLOW) ? SDEMC_FLAG_OVERFLOW : 0) |
                       (__any(flag & SDEMC_FLAG_JUMP_CSR) ? SDEMC_FLAG_JUMP_CSR : 0);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    partial2[0] = s;
    partial2[1] = q;
    if (!SPARSE_FLAG || any_flag != 0) cflag[0] = any_flag;
  }
}

// Fixed-order reduction of chunk partials into {sum, sumsq, n, flags}: thread
// t takes chunks t, t+256, ...; then a 256-leaf LDS tree.  With `clear` (the
// batched launch's flag slots, the same array as cflags) every set flag read
// is reset to zero for the slot's next use.
__device__ __forceinline__ void finalize_block(const double* partials, const int32_t* cflags, int64_t base,
                               int64_t nc, uint64_t n, double* st, int extra_flags = 0,
                               int32_t* clear = nullptr) {
  __shared__ double ss[256];
  __shared__ double qs[256];
  __shared__ int fs[256];
  const int t = threadIdx.x;
  double s = 0.0, q = 0.0;
  int fl = 0;
  int64_t c = t;
  // eight chunks' loads issued before their adds, which keep chunk order:
  // the loop waits on memory once per eight chunks instead of once per chunk
  // (a 2^21-path job's finalisation: 47 us one by one, 14 us by eight, no
  // faster by sixteen)
  constexpr int U = 8;
  for (; c + (U - 1) * 256 < nc; c += U * 256) {
    double ps[U], pq[U];
    int32_t pf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ps[u] = partials[2 * (base + c + u * 256)];
      pq[u] = partials[2 * (base + c + u * 256) + 1];
      pf[u] = cflags[base + c + u * 256];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      s += ps[u];
      q += pq[u];
      fl |= pf[u];
      if (clear && pf[u] != 0) clear[base + c + u * 256] = 0;
    }
  }
  for (; c < nc; c += 256) {
    s += partials[2 * (base + c)];
    q += partials[2 * (base + c) + 1];
    const int32_t cf = cflags[base + c];
    fl |= cf;
    if (clear && cf != 0) clear[base + c] = 0;
  }
  ss[t] = s;
  qs[t] = q;
  fs[t] = fl;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      ss[t] += ss[t + w];
      qs[t] += qs[t + w];
      fs[t] |= fs[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    st[0] = ss[0];
    st[1] = qs[0];
    st[2] = (double)n;
    st[3] = (double)(fs[0] | extra_flags);
  }
}

// ---------------------------------------------------------------------------
// Sources of randomness.  RNG: Philox in registers.  Fixed: host arrays.
// ---------------------------------------------------------------------------
struct PathCtx {
  uint64_t path;    // global Philox path index
  int64_t p;        // index within the job
  int64_t P;        // paths in the job
  uint32_t k0, k1;  // key
};

// Per-path optional stores (plain kernel only).
struct StoreCtx {
  float* paths;
  float* times;
  float* left;
  float* jumps;
  float* normals;
  int64_t P;
  int32_t rows;  // S
};

// Coarse half of a coupled uniform pair (an MLMC level): the sums of the M
// fine increments and the fine jumps of the coarse step -- their marks
// summed, and under the post-step rule composed into one affine map per
// component -- in the order the fine steps take them.
template <int D>
struct CoarseAcc {
  float sW[D], sB[D], A[D], Bc[D];
  float Jsum;
  __device__ __forceinline__ void reset() {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      sW[i] = 0.0f;
      sB[i] = 0.0f;
      A[i] = 1.0f;
      Bc[i] = 0.0f;
    }
    Jsum = 0.0f;
  }
  __device__ __forceinline__ void increment(const float (&dW)[D], const float (&dB)[D]) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      sW[i] += dW[i];
      sB[i] += dB[i];
    }
  }
  __device__ __forceinline__ void jump(const KModel& f, float J) {
    if (f.exact) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        const float g = __fmaf_rn(f.j1[i], J, 1.0f);
        A[i] = A[i] * g;
        Bc[i] = __fmaf_rn(Bc[i], g, f.j0[i] * J);
      }
    }
    Jsum += J;
  }
  // the coarse step of size hc of the coarse state xc[0 .. D-1]
  template <int FAM>
  __device__ __forceinline__ void step(const KModel& f, float hc, float* xc) {
    float y[D], yo[D];
#pragma unroll
    for (int i = 0; i < D; ++i) y[i] = yo[i] = xc[i];
    euler<D, FAM>(f, y, hc, sW, sB);
    if (f.exact) {
#pragma unroll
      for (int i = 0; i < D; ++i) y[i] = __fmaf_rn(A[i], y[i], Bc[i]);
    } else if (Jsum != 0.0f) {
      add_jump(f, y, yo, Jsum);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) xc[i] = y[i];
  }
};

// ===========================================================================
// Uniform grid, single level: N steps of h = fp32(T/N).  The G K normals of
// a group of G steps (group_steps: 4 or 6) come from G K / 6 Philox blocks
// generated together, so every index into them is a compile-time constant.
// ===========================================================================
// U: G-step groups per iteration of the step loop (code shape only; the
// draws and the arithmetic are the same for every U).
// POIS: the job's jump mode (per-step Poisson counts or the arrival clock),
// a template parameter so each form keeps only its own state in registers;
// path_uniform_rng below dispatches on the job.
// COUPLED: x is the fine path of an MLMC pair and xc (then non-null) its
// coarse path, stepped every M = j.M fine steps by CoarseAcc.
template <int D, int FAM, bool STORE, int U, bool POIS, bool COUPLED = false>
__device__ void path_uniform_rng_mode(const KJob& j, const PathCtx& c, bool lane_valid,
                                      float (&x)[D], const StoreCtx& st,
                                      float* xc = nullptr) {
  constexpr int K = NK<D, FAM>::K;
  constexpr int G = group_steps(K);         // steps per group (whole normal blocks)
  constexpr int WIN = G == 4 ? 8 : 6;       // steps per Poisson window, a multiple of G
  const KModel& f = j.m;
  const int32_t N = j.steps;
  const float sq = j.sqh;
  const uint32_t plo = (uint32_t)c.path, phi = (uint32_t)(c.path >> 32);
  constexpr bool pois = POIS;
  JumpClock jc;
  int32_t kj = 0x7fffffff;
  if (!pois) {
    jc.init(c.k0, c.k1, plo, phi, f.rate, f.inv_rate);
    kj = step_of(jc.tau, j.n_over_T, N);
  }
  // Jump draws off the step loop.  Arrival-clock mode: the marks of the
  // coming steps are drawn ahead into a queue of CLOCK_JQ (step, mark) slots
  // per lane in LDS, in the order the per-step form applies them, by passes
  // in which every lane of the wavefront takes its next jump at once, so the
  // draw code runs about max-over-lanes times per window of clock_win steps
  // (sized by the job's jump rate, rounded up to whole strides of G U)
  // instead of once per step in which any lane jumps; a step compares its
  // index with the head's step, kept in a register, and a jump reads its
  // mark and the next head from LDS (no register queue to shift); jumps
  // beyond the queue are taken inside the step loop by code compiled into
  // the windows where some lane needs it.  Per-step Poisson mode: the
  // jumping steps of each window of WIN steps are drawn by the whole
  // wavefront from a list in LDS (pois_draw_window), and each step reads its
  // marks back.  Marks, order and arithmetic are those of the per-step form,
  // so results are unchanged.
  const bool clock = !pois && f.rate > 0.0;  // job-uniform
  PoisLds* const pl = pois ? &pois_lds() : nullptr;
  ClockLds* const cq = clock ? reinterpret_cast<ClockLds*>(&pois_lds()) : nullptr;
  const uint32_t lane = pois ? 0u : __lane_id();
  int32_t qh = 0;   // the lane's queue head
  int32_t hk = -1;  // step of the head's jump, -1 when the queue is empty
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = f.x0[i];
  if (STORE && lane_valid) {
#pragma unroll
    for (int i = 0; i < D; ++i) st.paths[(int64_t)i * st.P + c.p] = x[i];
  }
  CoarseAcc<D> acc;
  int32_t cm = 0;  // fine steps taken in the current coarse step
  if constexpr (COUPLED) {
#pragma unroll
    for (int i = 0; i < D; ++i) xc[i] = f.x0[i];
    acc.reset();
  }
  uint32_t wjm = 0u;     // per-step Poisson mode: the lane's jumping steps of the window
  uint32_t wfirst = 0u;  // and its first item in the window's jump list
  auto group = [&](int32_t n0, auto spill) {
    uint32_t jm = 0u;     // the lane's jumping steps of the group
    uint32_t first = 0u;  // its first item in the window's jump list
    if (pois) {
      if ((n0 % WIN) == 0) {
        // a new window: the count words of its steps (step n: word n mod 4
        // of block n / 4 of the count stream)
        uint32_t w[WIN];
        const uint32_t T0 = j.pois_T[0];
        wjm = 0u;
        // blocks a window touches: two (8 steps from word 0, or 6 from word 0 or 2)
        static_assert(WIN == 8 || WIN == 6, "count-word windows of 8 or 6 steps");
        constexpr int NB = 2;
        uint32_t cw[4 * NB];
#pragma unroll
        for (int g = 0; g < NB; ++g) {
          U4 pw = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
          const int32_t nb = ((n0 >> 2) + g) * 4;  // first step of the block
          if (nb < N)
            pw = philox4x32_10((uint32_t)(n0 >> 2) + (uint32_t)g, STREAM_POIS, plo, phi, c.k0, c.k1);
          cw[4 * g + 0] = pw.x;
          cw[4 * g + 1] = pw.y;
          cw[4 * g + 2] = pw.z;
          cw[4 * g + 3] = pw.w;
        }
        if constexpr (WIN % 4 == 0) {
#pragma unroll
          for (int q = 0; q < WIN; ++q) w[q] = cw[q];
        } else {
          // a 6-step window starts at word 0 or 2 of its first block
          const bool odd = (n0 & 3) != 0;
#pragma unroll
          for (int q = 0; q < WIN; ++q) w[q] = odd ? cw[q + 2] : cw[q];
        }
#pragma unroll
        for (int q = 0; q < WIN; ++q) wjm |= (w[q] < T0 ? 1u : 0u) << q;
        const int32_t inside = N - n0;
        if (inside < WIN) wjm &= (1u << inside) - 1u;
        wfirst = pois_draw_window<WIN>(j, c.k0, c.k1, plo, phi, n0, w, wjm, *pl);
      }
      const uint32_t sh = (uint32_t)(n0 % WIN);
      jm = (wjm >> sh) & ((1u << G) - 1u);
      first = wfirst + (uint32_t)__builtin_popcount(wjm & ((1u << sh) - 1u));
    }
    // the G K normals of the group's steps: G K / 6 whole blocks
    float z[G * K];
    const uint32_t b0 = (uint32_t)n0 * (uint32_t)K / (uint32_t)NORMALS_PER_BLOCK;
#pragma unroll
    for (int b = 0; b < G * K / NORMALS_PER_BLOCK; ++b)
      normal_block(b0 + b, plo, phi, c.k0, c.k1, &z[NORMALS_PER_BLOCK * b]);
#pragma unroll
    for (int s = 0; s < G; ++s) {
      const int32_t n = n0 + s;
      if (n < N) {
        float dW[D], dB[D], xo[D];
        // the increments themselves only for the coarse path and stored normals
        if (COUPLED || (STORE && st.normals)) correlate<D, FAM>(f, &z[s * K], sq, dW, dB);
#pragma unroll
        for (int i = 0; i < D; ++i) xo[i] = x[i];
        uniform_step<D, FAM>(j, x, &z[s * K]);
        float jsum = 0.0f;
        if (!pois && n == hk) {
          do {
            const float J = cq->J[qh][lane];
            if (j.additive) {  // fma(j1, y, j0) = j0 exactly: add_jump's bits, one operation
#pragma unroll
              for (int i = 0; i < D; ++i) x[i] = __fmaf_rn(f.j0[i], J, x[i]);
            } else if (f.exact) {
              float y[D];
#pragma unroll
              for (int i = 0; i < D; ++i) y[i] = x[i];
              add_jump(f, x, y, J);
            } else {
              add_jump(f, x, xo, J);
            }
            jsum += J;
            if constexpr (COUPLED) acc.jump(f, J);
            ++qh;
            hk = qh < CLOCK_JQ ? cq->k[qh][lane] : -1;
          } while (n == hk);
        }
        if (pois && j.additive && ((jm >> s) & 1u)) {
          // constant jump coefficients: one total per jumping step
          const uint32_t it = first + (uint32_t)__builtin_popcount(jm & ((1u << s) - 1u));
          const float J = __uint_as_float(pl->v[it]);
#pragma unroll
          for (int i = 0; i < D; ++i) x[i] = __fmaf_rn(f.j0[i], J, x[i]);
          jsum += J;
          if constexpr (COUPLED) acc.jump(f, J);
        } else if (pois && ((jm >> s) & 1u)) {
          const uint32_t it = first + (uint32_t)__builtin_popcount(jm & ((1u << s) - 1u));
          const PoisLds& L = *pl;
          const int cnt = L.cnt[it];
          const uint32_t xb = L.xb[it];
          float J = __uint_as_float(L.v[it]);
          for (int k = 0;;) {
            if (f.exact) {
              float y[D];
#pragma unroll
              for (int i = 0; i < D; ++i) y[i] = x[i];
              add_jump(f, x, y, J);
            } else {
              add_jump(f, x, xo, J);
            }
            jsum += J;
            if constexpr (COUPLED) acc.jump(f, J);
            if (++k >= cnt) break;
            J = xb != POIS_NOX ? L.x[xb + (uint32_t)(k - 1)]
                               : levy_icdf(f, poisson_mark_uniform(c.k0, c.k1, plo, phi, n, k));
          }
        }
        if (decltype(spill)::value && n == kj) {
          do {
            const float J = jump_mark(f, jc.m1, jc.m2);
            if (f.exact) {
              float y[D];
#pragma unroll
              for (int i = 0; i < D; ++i) y[i] = x[i];
              add_jump(f, x, y, J);
            } else {
              add_jump(f, x, xo, J);
            }
            jsum += J;
            if constexpr (COUPLED) acc.jump(f, J);
            ++jc.j;
            jc.draw(c.k0, c.k1, plo, phi, f.inv_rate);
            kj = step_of(jc.tau, j.n_over_T, N);
          } while (kj == n);
        }
        if constexpr (COUPLED) {
          acc.increment(dW, dB);
          if (++cm == j.M) {
            acc.template step<FAM>(f, j.hc, xc);
            acc.reset();
            cm = 0;
          }
        }
        if (STORE && lane_valid) {
          const int64_t r = (int64_t)(n + 1) * D;
#pragma unroll
          for (int i = 0; i < D; ++i) st.paths[(r + i) * st.P + c.p] = x[i];
          if (st.jumps) {
#pragma unroll
            for (int i = 0; i < D; ++i) st.jumps[(r + i) * st.P + c.p] = jsum;
          }
          if (st.normals) {
            const int64_t rn = (int64_t)n * D * (DiagNoise<FAM>::value ? 1 : 2);
            if constexpr (FAM == FAM_GENERAL) {
              // the correlated increments w, as fixed-increment rows take them
              float w[GENERAL_K];
              general_w(f, &z[s * K], sq, w);
              const int64_t rg = (int64_t)n * GENERAL_K;
#pragma unroll
              for (int q = 0; q < GENERAL_K; ++q) st.normals[(rg + q) * st.P + c.p] = w[q];
            } else if (DiagNoise<FAM>::value) {
#pragma unroll
              for (int i = 0; i < D; ++i) st.normals[(rn + i) * st.P + c.p] = dW[i];
            } else {
#pragma unroll
              for (int i = 0; i < D; ++i) {
                st.normals[(rn + 2 * i) * st.P + c.p] = dW[i];
                st.normals[(rn + 2 * i + 1) * st.P + c.p] = dB[i];
              }
            }
          }
        }
      }
    }
  };
  // arrival-clock fill windows: a whole number of step-loop strides
  const int32_t win = ((j.clock_win + G * U - 1) / (G * U)) * (G * U);
  int32_t wend = 0;
  bool spill = false;
  for (int32_t n0 = 0; n0 < N; n0 += G * U) {
    if (clock && n0 >= wend) {
      // the previous window's queue is used up: its jumps all fell before n0
      wend = n0 + win;
      int32_t tail = 0;
      for (int q = 0; q < CLOCK_JQ; ++q) {
        if (!__any(kj < wend)) break;
        if (kj < wend) {
          cq->k[q][lane] = kj;
          cq->J[q][lane] = jump_mark(f, jc.m1, jc.m2);
          ++jc.j;
          jc.draw(c.k0, c.k1, plo, phi, f.inv_rate);
          kj = step_of(jc.tau, j.n_over_T, N);
          tail = q + 1;
        }
      }
      if (tail < CLOCK_JQ) cq->k[tail][lane] = -1;
      qh = 0;
      hk = cq->k[0][lane];
      spill = __any(kj < wend);
    }
    if (spill) {
      group(n0, std::true_type{});
#pragma unroll
      for (int u = 1; u < U; ++u)
        if (n0 + G * u < N) group(n0 + G * u, std::true_type{});
    } else {
      group(n0, std::false_type{});
#pragma unroll
      for (int u = 1; u < U; ++u)
        if (n0 + G * u < N) group(n0 + G * u, std::false_type{});
    }
  }
}

template <int D, int FAM, bool STORE, int U = 1>
__device__ __forceinline__ void path_uniform_rng(const KJob& j, const PathCtx& c, bool lane_valid,
                                                 float (&x)[D], const StoreCtx& st) {
  if (j.pois_mode != 0)
    path_uniform_rng_mode<D, FAM, STORE, U, true>(j, c, lane_valid, x, st);
  else
    path_uniform_rng_mode<D, FAM, STORE, U, false>(j, c, lane_valid, x, st);
}

// Coupled fine/coarse pair on the uniform grid (an MLMC level), Philox
// inputs: the single-level step loop with the coarse half riding along, so
// it shares the 4-step normal blocks, the jump queue and the per-wavefront
// jump lists; the draws, their order and the arithmetic are those of
// path_uniform_coupled.
template <int D, int FAM, int U = 1>
__device__ __forceinline__ void path_uniform_rng_coupled(const KJob& j, const PathCtx& c,
                                                         float (&xf)[D], float (&xc)[D]) {
  const StoreCtx st{};
  if (j.pois_mode != 0)
    path_uniform_rng_mode<D, FAM, false, U, true, true>(j, c, true, xf, st, xc);
  else
    path_uniform_rng_mode<D, FAM, false, U, false, true>(j, c, true, xf, st, xc);
}

// Row layout of fixed increments per step: D ('diag'), D + 1 ('indep'
// common: the last row is the shared increment), 2D ('indep' full).
template <int D, int FAM>
struct NKF {
  static constexpr int K = NK<D, FAM>::K;
};

// One step's row of fixed increments: dW (and dB) as the kernels' Euler step
// takes them; 'general' rows hold the GENERAL_K correlated increments w.
template <int D, int FAM>
__device__ __forceinline__ void fixed_row(const KModel& f, const float* row, int64_t P,
                                          float (&dW)[D], float (&dB)[D]) {
  if constexpr (FAM == FAM_GENERAL) {
    float w[GENERAL_K];
#pragma unroll
    for (int q = 0; q < GENERAL_K; ++q) w[q] = row[(int64_t)q * P];
    general_apply<D>(f, w, dW, dB);
    return;
  }
#pragma unroll
  for (int i = 0; i < D; ++i) dW[i] = row[(int64_t)i * P];
  if (FAM == FAM_LEVY) {
    const float cc = row[(int64_t)D * P];
#pragma unroll
    for (int i = 0; i < D; ++i) dB[i] = cc;
  } else if (FAM == FAM_DOUBLE) {
#pragma unroll
    for (int i = 0; i < D; ++i) dB[i] = row[(int64_t)(D + i) * P];
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) dB[i] = 0.0f;
  }
}

// One uniform-grid step fed from fixed inputs: the Euler update with the
// step's increments, then its `cnt` jumps with marks read in CSR order.
template <int D, int FAM>
__device__ __forceinline__ void fixed_step(const KJob& j, const sdemc_fixed_in& in, float (&x)[D],
                                           const float (&dW)[D], const float (&dB)[D], int cnt,
                                           int64_t& off, int& bad) {
  const KModel& f = j.m;
  float xo[D];
#pragma unroll
  for (int i = 0; i < D; ++i) xo[i] = x[i];
  euler<D, FAM>(f, x, j.h, dW, dB);
  for (int k = 0; k < cnt; ++k) {
    const bool ok = off >= 0 && off < in.n_marks;
    const float J = ok ? in.jump_size_dev[off] : 0.0f;
    bad |= !ok;
    ++off;
    if (f.exact) {
      float y[D];
#pragma unroll
      for (int i = 0; i < D; ++i) y[i] = x[i];
      add_jump(f, x, y, J);
    } else {
      add_jump(f, x, xo, J);
    }
  }
}

// Fixed increments on the uniform grid: dw [N][K][P] already scaled and
// correlated, counts [N][P], marks in CSR order per path.  The inputs of G
// consecutive steps are loaded before the G dependent updates run, so a
// wavefront keeps G rows of loads in flight instead of one.  With SYNC the
// workgroup meets at a barrier after every 4G steps (every thread of it must
// call this with the same N): its wavefronts then read neighbouring halves
// of the same 128-byte count lines within a few steps of each other, while
// the lines are still in L2, instead of fetching them again after drifting
// apart.
template <int D, int FAM, bool SYNC = false>
__device__ void path_uniform_fixed(const KJob& j, const sdemc_fixed_in& in, int64_t p, int64_t P,
                                   float (&x)[D], int& bad) {
  constexpr int KF = NKF<D, FAM>::K;
  constexpr int G = D <= 2 ? 8 : 4;
  const KModel& f = j.m;
  const int32_t N = j.steps;
  int64_t off = in.jump_off_dev ? in.jump_off_dev[p] : 0;
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = f.x0[i];
  const float* dwp = in.dw_dev + p;
  const uint8_t* cntp = in.counts_dev ? in.counts_dev + p : nullptr;
  int32_t n = 0;
  for (; n + G <= N; n += G) {
    float dW[G][D], dB[G][D];
    int cnt[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      fixed_row<D, FAM>(f, dwp + (int64_t)(n + g) * KF * P, P, dW[g], dB[g]);
      cnt[g] = cntp ? (int)cntp[(int64_t)(n + g) * P] : 0;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) fixed_step<D, FAM>(j, in, x, dW[g], dB[g], cnt[g], off, bad);
    if (SYNC && ((n / G) & 3) == 3) __syncthreads();
  }
  for (; n < N; ++n) {
    float dW[D], dB[D];
    fixed_row<D, FAM>(f, dwp + (int64_t)n * KF * P, P, dW, dB);
    const int cnt = cntp ? (int)cntp[(int64_t)n * P] : 0;
    fixed_step<D, FAM>(j, in, x, dW, dB, cnt, off, bad);
  }
}

// ===========================================================================
// Uniform grid, coupled fine/coarse pair (MLMC level).  Fine steps of
// h = fp32(T/Nf), coarse steps of hc = fp32(M T/Nf) whose increment is the
// sum of its M fine ones; a jump in fine step n belongs to coarse step n / M.
// Coarse jumps under the post-step rule compose into one affine map per
// component, so any number of them in a coarse step is exact.  This form reads
// fixed increments (UniFixedSrc); Philox inputs go through
// path_uniform_rng_coupled, which takes the same steps in the same order.
// ===========================================================================
template <int D, int FAM, class SRC>
__device__ void path_uniform_coupled(const KJob& j, SRC& src, float (&xf)[D], float (&xc)[D]) {
  const KModel& f = j.m;
  const int32_t M = j.M;
  const int32_t Nc = j.steps / M;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    xf[i] = f.x0[i];
    xc[i] = f.x0[i];
  }
  for (int32_t m = 0; m < Nc; ++m) {
    float sW[D], sB[D], A[D], Bc[D];
    float Jsum = 0.0f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      sW[i] = 0.0f;
      sB[i] = 0.0f;
      A[i] = 1.0f;
      Bc[i] = 0.0f;
    }
    for (int32_t s = 0; s < M; ++s) {
      const int32_t n = m * M + s;
      float dW[D], dB[D], xo[D];
      src.incr(j, n, dW, dB);
#pragma unroll
      for (int i = 0; i < D; ++i) {
        sW[i] += dW[i];
        sB[i] += dB[i];
        xo[i] = xf[i];
      }
      euler<D, FAM>(f, xf, j.h, dW, dB);
      while (src.jump_pending(j, n)) {
        const float J = src.take_jump(j);
        if (f.exact) {
          float y[D];
#pragma unroll
          for (int i = 0; i < D; ++i) y[i] = xf[i];
          add_jump(f, xf, y, J);
#pragma unroll
          for (int i = 0; i < D; ++i) {
            const float g = __fmaf_rn(f.j1[i], J, 1.0f);
            A[i] = A[i] * g;
            Bc[i] = __fmaf_rn(Bc[i], g, f.j0[i] * J);
          }
        } else {
          add_jump(f, xf, xo, J);
        }
        Jsum += J;
      }
    }
    float xco[D];
#pragma unroll
    for (int i = 0; i < D; ++i) xco[i] = xc[i];
    euler<D, FAM>(f, xc, j.hc, sW, sB);
    if (f.exact) {
#pragma unroll
      for (int i = 0; i < D; ++i) xc[i] = __fmaf_rn(A[i], xc[i], Bc[i]);
    } else if (Jsum != 0.0f) {
      add_jump(f, xc, xco, Jsum);
    }
  }
}

// Fixed source for the uniform coupled path (fine-grid arrays).
template <int D, int FAM>
struct UniFixedSrc {
  const sdemc_fixed_in* in;
  int64_t p, P, off;
  int32_t cur_n, left;
  int bad;
  __device__ __forceinline__ void init(const sdemc_fixed_in& fin, int64_t pp, int64_t PP) {
    in = &fin;
    p = pp;
    P = PP;
    off = fin.jump_off_dev ? fin.jump_off_dev[pp] : 0;
    cur_n = -1;
    left = 0;
    bad = 0;
  }
  __device__ __forceinline__ void incr(const KJob& j, int32_t n, float (&dW)[D], float (&dB)[D]) {
    constexpr int KF = NKF<D, FAM>::K;
    fixed_row<D, FAM>(j.m, in->dw_dev + (int64_t)n * KF * P + p, P, dW, dB);
    cur_n = n;
    left = in->counts_dev ? (int)in->counts_dev[(int64_t)n * P + p] : 0;
  }
  __device__ __forceinline__ bool jump_pending(const KJob&, int32_t n) const {
    return n == cur_n && left > 0;
  }
  __device__ __forceinline__ float take_jump(const KJob&) {
    --left;
    const bool ok = off >= 0 && off < in->n_marks;
    const float J = ok ? in->jump_size_dev[off] : 0.0f;
    bad |= !ok;
    ++off;
    return J;
  }
};

// ===========================================================================
// Jump-adapted grid (solvers.py:164-226), fp32 clock.  Every iteration is
// one step for every unfinished path, so the normal index (step * K + k) is
// the same across the wavefront while its lanes are active.
// ===========================================================================
__device__ __forceinline__ bool isclose32(float tau, float t) {
#pragma clang fp contract(off)
  // torch.isclose(tau, t, atol=1e-12) with the default rtol=1e-5, in fp32
  const float err = fabsf(tau - t);
  const float allowed = 1e-12f + fabsf(1e-5f * t);
  return tau == t || err <= allowed;
}

// Philox source of the adapted grid.  The pending jump (tau1, J1) and the one
// after it (tau2, J2) are held drawn, marks included, so a jump costs the
// lane only a shift; refill(), called by the whole wavefront every few
// steps, draws the replacements of all its lanes in one pass.  A lane that
// jumps again before a refill draws in place.  Each jump j takes the same
// Philox block and the arrival times accumulate in the same fp32 order as a
// jump-by-jump clock, so the grid and the marks do not depend on when the
// draws happen.
template <int D, int FAM>
struct AdaptRngSrc {
  NormalStream ns;
  JumpClock jc;  // the last jump drawn
  uint32_t plo, phi, k0, k1;
  float tau1, J1, tau2, J2;
  bool have2;
  __device__ __forceinline__ void draw_next(const KModel& f, float& tau, float& J) {
    ++jc.j;
    jc.draw(k0, k1, plo, phi, f.inv_rate);
    tau = jc.tau;
    J = jump_mark(f, jc.m1, jc.m2);
  }
  __device__ __forceinline__ void init(const KModel& f, const PathCtx& c) {
    k0 = c.k0;
    k1 = c.k1;
    plo = (uint32_t)c.path;
    phi = (uint32_t)(c.path >> 32);
    ns.init(k0, k1, c.path);
    jc.init(k0, k1, plo, phi, f.rate, f.inv_rate);
    tau1 = jc.tau;
    J1 = f.rate > 0.0 ? jump_mark(f, jc.m1, jc.m2) : 0.0f;
    tau2 = __builtin_inff();
    J2 = 0.0f;
    have2 = false;
  }
  __device__ __forceinline__ float z(uint32_t row, int k) {
    constexpr int K = NK<D, FAM>::K;
    return ns.get(row * K + k);
  }
  // normals of rows row0 .. row0 + G - 1 (G = group_steps(K), row0 a
  // multiple of G): G K / 6 whole blocks, zq[s K + k] = z(row0 + s, k)
  static constexpr int G = group_steps(NK<D, FAM>::K);
  __device__ __forceinline__ void rows(uint32_t row0, bool, float (&zq)[G * NK<D, FAM>::K]) {
    constexpr int K = NK<D, FAM>::K;
    const uint32_t b0 = row0 * (uint32_t)K / (uint32_t)NORMALS_PER_BLOCK;
#pragma unroll
    for (int b = 0; b < G * K / NORMALS_PER_BLOCK; ++b)
      normal_block(b0 + b, plo, phi, k0, k1, &zq[NORMALS_PER_BLOCK * b]);
  }
  __device__ __forceinline__ float next_time() const { return tau1; }
  __device__ __forceinline__ float mark(const KModel&) const { return J1; }
  __device__ __forceinline__ void advance(const KModel& f) {
    if (have2) {
      tau1 = tau2;
      J1 = J2;
      have2 = false;
    } else {
      draw_next(f, tau1, J1);
    }
  }
  // whole wavefront: draw the second slot of every live lane that lacks one
  // and whose pending jump falls before T (later ones are never used)
  __device__ __forceinline__ void refill(const KModel& f, bool lane_live, float T32) {
    if (f.rate > 0.0) {
      const bool need = lane_live && !have2 && tau1 < T32;
      if (__any(need)) {
        if (need) {
          draw_next(f, tau2, J2);
          have2 = true;
        }
      }
    }
  }
};

template <int D, int FAM>
struct AdaptFixedSrc {
  const sdemc_fixed_in* in;
  int64_t p, P;
  int32_t jn;
  __device__ __forceinline__ void init(const sdemc_fixed_in& fin, int64_t pp, int64_t PP) {
    in = &fin;
    p = pp;
    P = PP;
    jn = 0;
  }
  __device__ __forceinline__ float z(uint32_t row, int k) {
    constexpr int K = NK<D, FAM>::K;
    if ((int32_t)row >= in->n_rows) return 0.0f;
    return in->dw_dev[((int64_t)row * K + k) * P + p];
  }
  // rows row0 .. row0 + G - 1 of a live lane (a lane past the last path, or
  // one whose path has ended, reads nothing)
  static constexpr int G = group_steps(NK<D, FAM>::K);
  __device__ __forceinline__ void rows(uint32_t row0, bool live, float (&zq)[G * NK<D, FAM>::K]) {
    constexpr int K = NK<D, FAM>::K;
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
      for (int k = 0; k < K; ++k) zq[q * K + k] = live ? z(row0 + (uint32_t)q, k) : 0.0f;
  }
  __device__ __forceinline__ float next_time() const {
    return jn < in->max_jumps ? in->jump_time_dev[(int64_t)jn * P + p] : __builtin_inff();
  }
  __device__ __forceinline__ float mark(const KModel&) const {
    return jn < in->max_jumps ? in->jump_size_dev[(int64_t)jn * P + p] : 0.0f;
  }
  __device__ __forceinline__ void advance(const KModel&) { ++jn; }
  __device__ __forceinline__ void refill(const KModel&, bool, float) {}
};

// Single level on the adapted grid.  Returns the number of steps taken;
// `overflow` is set when max_steps was reached before T.
template <int D, int FAM, bool STORE, class SRC>
__device__ int32_t path_adapted(const KJob& j, SRC& src, bool lane_valid, int64_t p,
                                float (&x)[D], int& overflow, const StoreCtx& st) {
  constexpr int K = NK<D, FAM>::K;
  const KModel& f = j.m;
  const float T32 = (float)j.T;
  const int32_t max_steps = j.max_steps;
  float h = j.h;
  float t = 0.0f;
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = f.x0[i];
  float tau = lane_valid ? src.next_time() : 0.0f;
  bool active = lane_valid && (t < T32);
  int32_t steps = 0;
  overflow = 0;
  if (STORE && lane_valid) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      st.paths[(int64_t)i * st.P + p] = x[i];
      if (st.left) st.left[(int64_t)i * st.P + p] = x[i];
    }
    if (st.times) st.times[p] = 0.0f;
  }
  // A live lane takes one step per iteration, so at the start of iteration i
  // every live lane is at row i: the normals of rows i .. i + G - 1 are drawn
  // for the wavefront together every G iterations (G = group_steps(K), whole
  // blocks), in the stream's layout (row r, normal k is index r K + k).  The
  // second jump slots are refilled in the first group at or after every
  // multiple of clock_win iterations (a power of two from 4: about half a
  // jump per path between refills), a jump beyond them being drawn in the
  // step.
  constexpr int G = group_steps(K);
  const uint32_t refill_mask = (uint32_t)j.clock_win - 1u;
  for (uint32_t row0 = 0; __any(active); row0 += G) {
    if ((row0 & refill_mask) < (uint32_t)G) src.refill(f, active, T32);
    float zq[G * K];
    src.rows(row0, active, zq);
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (active && steps >= max_steps) {
        overflow = 1;
        active = false;
      }
      if (active) {
        float dt, tn;
        {
#pragma clang fp contract(off)
          h = fminf(h, fmaxf(T32 - t, 0.0f));
          dt = fminf(h, tau - t);
        }
        const float sq = __builtin_amdgcn_sqrtf(dt);
        float zz[K];
#pragma unroll
        for (int k = 0; k < K; ++k) zz[k] = zq[q * K + k];
        float dW[D], dB[D], xo[D];
        correlate<D, FAM>(f, zz, sq, dW, dB);
#pragma unroll
        for (int i = 0; i < D; ++i) xo[i] = x[i];
        euler<D, FAM>(f, x, dt, dW, dB);
        {
#pragma clang fp contract(off)
          tn = t + dt;
        }
        t = tn;
        ++steps;
        if (STORE) {
          const int64_t r = (int64_t)steps * D;
          if (st.left) {
#pragma unroll
            for (int i = 0; i < D; ++i) st.left[(r + i) * st.P + p] = x[i];
          }
          if (st.times) st.times[(int64_t)steps * st.P + p] = t;
          if (st.normals) {
            const int64_t rn = (int64_t)(steps - 1) * D * (DiagNoise<FAM>::value ? 1 : 2);
            if (DiagNoise<FAM>::value) {
#pragma unroll
              for (int i = 0; i < D; ++i) st.normals[(rn + i) * st.P + p] = dW[i];
            } else {
#pragma unroll
              for (int i = 0; i < D; ++i) {
                st.normals[(rn + 2 * i) * st.P + p] = dW[i];
                st.normals[(rn + 2 * i + 1) * st.P + p] = dB[i];
              }
            }
          }
        }
        float Jap = 0.0f;
        if (isclose32(tau, t)) {
          const float J = src.mark(f);
          if (f.exact) {
            float y[D];
#pragma unroll
            for (int i = 0; i < D; ++i) y[i] = x[i];
            add_jump(f, x, y, J);
          } else {
            add_jump(f, x, xo, J);
          }
          Jap = J;
          src.advance(f);
          tau = src.next_time();
        }
        if (STORE) {
          const int64_t r = (int64_t)steps * D;
#pragma unroll
          for (int i = 0; i < D; ++i) {
            st.paths[(r + i) * st.P + p] = x[i];
            if (st.jumps) st.jumps[(r + i) * st.P + p] = Jap;
          }
        }
        active = t < T32;
      }
    }
  }
  if (STORE && lane_valid) {
    // rows past this path's end repeat its final state (a finished path
    // takes zero-length steps while others run, solvers.py:182-191)
    for (int32_t r = steps + 1; r <= st.rows; ++r) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        st.paths[((int64_t)r * D + i) * st.P + p] = x[i];
        if (st.left) st.left[((int64_t)r * D + i) * st.P + p] = x[i];
        if (st.jumps) st.jumps[((int64_t)r * D + i) * st.P + p] = 0.0f;
      }
      if (st.times) st.times[(int64_t)r * st.P + p] = t;
    }
  }
  return steps;
}

// Coupled pair on the adapted grid (solvers.py:228-307): per iteration M fine
// sub-steps toward the pending jump time, then one coarse step driven by the
// sum of the fine increments; a jump that falls now hits both levels, with
// the fine coefficient taken before the last fine sub-step.
template <int D, int FAM, class SRC>
__device__ int32_t path_adapted_coupled(const KJob& j, SRC& src, bool lane_valid,
                                        float (&xf)[D], float (&xc)[D], int& overflow) {
  constexpr int K = NK<D, FAM>::K;
  const KModel& f = j.m;
  const int32_t M = j.M;
  const int32_t max_steps = j.max_steps;
  const float T32 = (float)j.T;
  float hf = j.h;
  float hc;
  {
#pragma clang fp contract(off)
    hc = (float)M * hf;
  }
  float tf = 0.0f, tc = 0.0f;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    xf[i] = f.x0[i];
    xc[i] = f.x0[i];
  }
  float tau = lane_valid ? src.next_time() : 0.0f;
  bool active = lane_valid && (tf < T32);
  int32_t it = 0;
  overflow = 0;
  int32_t iter = 0;
  int32_t refill_mask = j.clock_win - 1;  // a power of two minus one
  while (refill_mask > 0 && (refill_mask + 1) * M > j.clock_win) refill_mask >>= 1;
  while (__any(active)) {
    // second jump slots for the wavefront together, every clock_win / M
    // iterations (an iteration is M fine steps; the jumps keep their Philox
    // blocks, so the marks are those drawn one by one)
    if ((iter++ & refill_mask) == 0) src.refill(f, active, T32);
    if (active && it >= max_steps) {
      overflow = 1;
      active = false;
    }
    if (active) {
      float sW[D], sB[D], xfo[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        sW[i] = 0.0f;
        sB[i] = 0.0f;
        xfo[i] = xf[i];
      }
      for (int32_t s = 0; s < M; ++s) {
        float dt;
        {
#pragma clang fp contract(off)
          hf = fminf(hf, fmaxf(T32 - tf, 0.0f));
          dt = fminf(hf, tau - tf);
        }
        const float sq = __builtin_amdgcn_sqrtf(dt);
        float zz[K];
#pragma unroll
        for (int k = 0; k < K; ++k) zz[k] = src.z((uint32_t)(it * M + s), k);
        float dW[D], dB[D];
        correlate<D, FAM>(f, zz, sq, dW, dB);
#pragma unroll
        for (int i = 0; i < D; ++i) xfo[i] = xf[i];
        euler<D, FAM>(f, xf, dt, dW, dB);
        {
#pragma clang fp contract(off)
          tf = tf + dt;
#pragma unroll
          for (int i = 0; i < D; ++i) {
            sW[i] = sW[i] + dW[i];
            sB[i] = sB[i] + dB[i];
          }
        }
      }
      float dtc;
      {
#pragma clang fp contract(off)
        hc = fminf(hc, fmaxf(T32 - tc, 0.0f));
        dtc = fminf(hc, tau - tc);
      }
      float xco[D];
#pragma unroll
      for (int i = 0; i < D; ++i) xco[i] = xc[i];
      euler<D, FAM>(f, xc, dtc, sW, sB);
      {
#pragma clang fp contract(off)
        tc = tc + dtc;
      }
      if (isclose32(tau, tf)) {
        const float J = src.mark(f);
        if (f.exact) {
          float y[D];
#pragma unroll
          for (int i = 0; i < D; ++i) y[i] = xf[i];
          add_jump(f, xf, y, J);
#pragma unroll
          for (int i = 0; i < D; ++i) y[i] = xc[i];
          add_jump(f, xc, y, J);
        } else {
          add_jump(f, xf, xfo, J);
          add_jump(f, xc, xco, J);
        }
        src.advance(f);
        tau = src.next_time();
      }
      ++it;
      active = tf < T32;
    }
  }
  return it;
}

}  // namespace sdemc

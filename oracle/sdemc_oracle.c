/*
 * sdemc_oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU fp64 restatement of the
 * sde_mc hot path, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  Nothing in the sde_mc_amd package links,
 * loads or calls it.
 *
 * It restates, path by path and in double precision:
 *   JumpDiffusionSolver.solve      sde_mc/solvers.py:164-226  (adapted grid)
 *   JumpDiffusionSolver.multilevel sde_mc/solvers.py:228-307  (adapted, coupled)
 *   DiffusionSolver.solve/multilevel sde_mc/solvers.py:68-119 (uniform grid)
 *   EulerScheme.step               sde_mc/schemes.py:5-10
 *   sample_corr_normals            sde_mc/solvers.py:51-56, 195-201
 *   LogNormalJumpsSde.sample_jumps sde_mc/sde.py:325-326
 *   InverseCdf                     sde_mc/levy.py:10-30
 *   Option.__call__ + discount     sde_mc/options.py:167-176, 179-321
 * plus the uniform grid with per-step jump counts that the kernels also run,
 * and the random-number layout the kernels use (Philox4x32-10 of Salmon et
 * al., SC'11, Box-Muller, exponential arrival gaps), so paths drawn in the
 * kernels can be recomputed here from the same counters.
 *
 * What stays in fp32 on purpose, because it decides *which* grid a path
 * takes and the reference computes it in fp32 tensors: the adapted-grid
 * clock t, h and dt (solvers.py:168,190-191,204), torch.isclose
 * (solvers.py:212), the jump times (fp32 tensor, solvers.py:144), and the
 * Levy uniform and its branch thresholds (levy.py:28-30, 86).  Everything
 * else (state, increments, marks, payoff) is fp64.
 *
 * Build: see oracle/build.py (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sdemc.h"

#define D_MAX SDEMC_MAX_DIM
#define ORACLE_API __attribute__((visibility("default")))

/* ---------------- Philox4x32-10 --------------------------------------- */
ORACLE_API void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static void philox_block(uint32_t blk, uint32_t stream, uint64_t path, uint64_t seed,
                         uint32_t out[4]) {
  uint32_t c[4] = {blk, stream, (uint32_t)path, (uint32_t)(path >> 32)};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  oracle_philox(c, k, out);
}

static double u_open(uint32_t w) { return (double)((w >> 8) | 1u) * 0x1p-24; }
/* Box-Muller angle in revolutions: the top 23 bits of the word */
static double u_angle(uint32_t w) { return (double)(w >> 9) * 0x1p-23; }

/* ln u for u = (w | 1) 2^-32, fp32: u = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s) with s = (m-1)/(m+1) by the odd series to s^9.  Every
 * operation is a single correctly rounded fp32 operation (this file is built
 * with -ffp-contract=off), which is what makes the exponential arrival gaps
 * of the kernels reproducible here bit for bit. */
ORACLE_API float oracle_ln_unit(uint32_t w) {
  uint32_t v = w | 1u;
  int z = __builtin_clz(v);
  uint32_t top = v << z;
  float m = (float)(top >> 8) * 0x1p-23f;
  int e = -(z + 1);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float sn = m - 1.0f, sd = m + 1.0f;
  float sq = sn / sd;
  float s2 = sq * sq;
  float p = fmaf(s2, 0.222222222f, 0.285714286f);
  p = fmaf(s2, p, 0.4f);
  p = fmaf(s2, p, 0.666666667f);
  p = fmaf(s2, p, 2.0f);
  float lnm = sq * p;
  float et = (float)e * 0.693147181f;
  return lnm + et;
}

/* P(N >= k) 2^32, k = 1..16, of N ~ Poisson(mu), rounded to integers; the
 * tail sums run from the top.  inv[c-1] = 1 / (T_c - T_{c+1}) (T_17 = 0) is
 * the reciprocal width of count c's interval, 0 for an empty interval.
 * Returns whether P(N > 16) < 2^-33. */
static int poisson_thresholds(double mu, uint32_t T16[16], double inv[16]) {
  double p[64], tail[65];
  p[0] = exp(-mu);
  for (int i = 1; i < 64; ++i) p[i] = p[i - 1] * mu / (double)i;
  tail[64] = 0.0;
  for (int i = 63; i >= 0; --i) tail[i] = tail[i + 1] + p[i];
  for (int k = 1; k <= 16; ++k) {
    double v = tail[k] * 4294967296.0;
    T16[k - 1] = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)(v + 0.5);
  }
  for (int k = 0; k < 16; ++k) {
    uint32_t width = T16[k] - (k < 15 ? T16[k + 1] : 0u);
    inv[k] = width ? 1.0 / (double)width : 0.0;
  }
  return tail[17] * 4294967296.0 < 0.5;
}

ORACLE_API int oracle_poisson_thresholds(double mu, uint32_t* T16, double* inv) {
  return poisson_thresholds(mu, T16, inv);
}

/* normal i of a path: normal i%6 of Brownian block i/6, six per block from three
 * Box-Muller pairs: pair p takes its radius uniform ((w_p >> 8) | 1) 2^-24 from
 * word p and its angle a = ((w_p mod 2^9) 2^10 + bits [10p, 10p + 10) of w_3) 2^-19;
 * z_2p = r cos 2 pi a, z_2p+1 = r sin 2 pi a */
ORACLE_API double oracle_normal(uint64_t seed, uint64_t path, uint32_t i) {
  uint32_t w[4];
  philox_block(i / 6u, 0, path, seed, w);
  uint32_t s = i % 6u, p = s >> 1;
  uint32_t ang = ((w[p] & 0x1FFu) << 10) | ((w[3] >> (10u * p)) & 0x3FFu);
  double r = sqrt(-2.0 * log(u_open(w[p])));
  double th = 2.0 * M_PI * (double)ang * 0x1p-19;
  return (s & 1u) ? r * sin(th) : r * cos(th);
}

/* ---------------- model ------------------------------------------------ */
typedef struct {
  const sdemc_job* job;
  int D, K, noise;
  int R; /* entries of dW per step: D, or 4 for 'general' (the Brownian increments) */
  double L[D_MAX][D_MAX];
  float lv_y1, lv_y2, lv_y3;
  float inv_rate, n_over_T;
  int pois;
  uint32_t pT[16];
  double pinv[16];
} model_t;

static void model_init(model_t* m, const sdemc_job* job) {
  m->job = job;
  m->D = job->model.dim;
  m->noise = job->model.noise;
  m->K = (m->noise == SDEMC_NOISE_DIAG || m->noise == SDEMC_NOISE_HESTON)
             ? m->D
             : (m->noise == SDEMC_NOISE_INDEP_COMMON ? m->D + 1 : 2 * m->D);
  m->R = m->D;
  /* 'general': 4 normals per step whatever brown_dim, the factor zero-padded */
  int nb = D_MAX;
  if (m->noise == SDEMC_NOISE_GENERAL) {
    m->K = m->R = D_MAX;
    nb = job->model.brown_dim;
  }
  for (int i = 0; i < D_MAX; ++i)
    for (int j = 0; j < D_MAX; ++j)
      m->L[i][j] = (i < nb && j < nb) ? job->model.chol[i * D_MAX + j] : 0.0;
  m->lv_y1 = (float)job->model.jp[6];
  m->lv_y2 = (float)job->model.jp[7];
  m->lv_y3 = (float)job->model.jp[8];
  m->inv_rate = job->model.rate > 0.0 ? (float)(1.0 / job->model.rate) : 0.0f;
  m->n_over_T = (float)((double)job->steps / job->T);
  m->pois = 0;
  memset(m->pT, 0, sizeof(m->pT));
  memset(m->pinv, 0, sizeof(m->pinv));
  /* per-step Poisson counts: Levy marks on the uniform grid from 1/16
     expected jumps per step on, the arrival clock below */
  if (job->model.jump_law == SDEMC_JUMP_LEVY && job->grid == SDEMC_GRID_UNIFORM) {
    double mu = job->model.rate * job->T / (double)job->steps;
    if (mu >= 0.0625) m->pois = poisson_thresholds(mu, m->pT, m->pinv);
  }
}

/* InverseCdf.__call__ (levy.py:22-30) in fp64, branch chosen on fp32 y */
static double levy_icdf(const model_t* m, float y) {
  const double* jp = m->job->model.jp;
  double cm = jp[0], cp = jp[1], imu = jp[2], al = jp[3], eps = jp[4], lda = jp[5];
  double epa = pow(eps, -al), yd = (double)y;
  if (y <= m->lv_y1) return log(lda * yd / (cm * imu)) * imu - 1.0;
  if (y < m->lv_y2) return -pow(al * (lda * yd / cm - imu) + 1.0, -1.0 / al);
  if (y < m->lv_y3) return pow((-al / cp) * (lda * yd - cm * imu - cm * (epa - 1.0) / al) + epa, -1.0 / al);
  return 1.0 - imu * log(lda * (1.0 - yd) / (cp * imu));
}

ORACLE_API double oracle_levy_icdf(const sdemc_job* job, float y) {
  model_t m;
  model_init(&m, job);
  return levy_icdf(&m, y);
}

/* mark of jump j from words (m1, m2) */
static double mark_from_words(const model_t* m, uint32_t m1, uint32_t m2) {
  const sdemc_model* md = &m->job->model;
  if (md->jump_law == SDEMC_JUMP_LEVY) {
    /* torch.rand + UNIFORM_TOL / 3 in fp32 (levy.py:86) */
    volatile float base = (float)(m1 >> 8) * 0x1p-24f;
    volatile float y = base + 0x1.555556p-26f;
    return levy_icdf(m, y);
  }
  double z = sqrt(-2.0 * log(u_open(m1))) * cos(2.0 * M_PI * u_angle(m2));
  return exp(md->jp[0] + md->jp[1] * z) - 1.0;
}

/* Euler step of schemes.py:5-10 in fp64; dW, dB already scaled/correlated */
static void euler(const model_t* m, double* x, double dt, const double* dW, const double* dB) {
  const sdemc_model* md = &m->job->model;
  if (m->noise == SDEMC_NOISE_HESTON) {
    /* HestonScheme.step (schemes.py:16-21): S by Euler with sqrt of the
     * pre-step v; the new v is the square of the larger root of
     * a y^2 + b y + c = 0, a = -1 - kappa dt, b = xi dW_1,
     * c = v + kappa theta dt - xi^2 dt / 2 (sde.py:274-278, helpers.py:36-48) */
    double S = x[0], v = x[1];
    double kappa = md->jp[0], theta = md->jp[1], xi = md->jp[2];
    double a = -1.0 - kappa * dt, b = xi * dW[1], c = v + kappa * theta * dt - 0.5 * dt * xi * xi;
    double sd = sqrt(b * b - 4.0 * a * c);
    double r1 = (-b + sd) / (2.0 * a), r2 = (-b - sd) / (2.0 * a);
    double y = r1 > r2 ? r1 : r2;
    x[0] = S + md->a1[0] * S * dt + sqrt(v) * S * dW[0];
    x[1] = y * y;
    return;
  }
  if (m->noise == SDEMC_NOISE_GENERAL) {
    /* x + a(x) dt + B(x) dW with the (dim, brown_dim) matrix
     * B(x)_ij = bm0_ij + bm1_ij x_i: the intended torch.matmul of
     * schemes.py:12 on the correlated increments */
    double xo[D_MAX];
    memcpy(xo, x, sizeof(double) * m->D);
    for (int i = 0; i < m->D; ++i) {
      double diff = 0.0;
      for (int j = 0; j < md->brown_dim; ++j)
        diff += (md->bm0[i * D_MAX + j] + md->bm1[i * D_MAX + j] * xo[i]) * dW[j];
      x[i] = xo[i] + (md->a0[i] + md->a1[i] * xo[i]) * dt + diff;
    }
    return;
  }
  const double s0 = x[0]; /* pre-step x_0: AsianWrapper's integrand (sde.py:389-390) */
  for (int i = 0; i < m->D; ++i) {
    double a = md->a0[i] + md->a1[i] * x[i] + (i > 0 ? md->ax[i] * s0 : 0.0);
    double b = md->b0[i] + md->b1[i] * x[i];
    double diff = b * dW[i];
    if (m->noise != SDEMC_NOISE_DIAG) diff += (md->c0[i] + md->c1[i] * x[i]) * dB[i];
    x[i] = x[i] + a * dt + diff;
  }
}

static void add_jump(const model_t* m, double* x, const double* y, double J) {
  const sdemc_model* md = &m->job->model;
  for (int i = 0; i < m->D; ++i) x[i] += (md->j0[i] + md->j1[i] * y[i]) * J;
}

/* sample_corr_normals: (z * sqrt(dt)) then L @ (solvers.py:51-56); the
 * 'indep' second column is one common scalar (solvers.py:198-201) or, for
 * the diffusion solver, a second correlated vector (solvers.py:80-81) */
static void correlate(const model_t* m, const double* z, double sq, double* dW, double* dB) {
  int D = m->D;
  if (m->noise == SDEMC_NOISE_GENERAL) {
    for (int i = 0; i < D_MAX; ++i) {
      double acc = 0.0;
      for (int j = 0; j <= i; ++j) acc += m->L[i][j] * (z[j] * sq);
      dW[i] = acc;
      dB[i] = 0.0;
    }
    return;
  }
  for (int i = 0; i < D; ++i) {
    double acc = 0.0;
    for (int j = 0; j <= i; ++j) acc += m->L[i][j] * (z[j] * sq);
    dW[i] = acc;
    dB[i] = 0.0;
  }
  if (m->noise == SDEMC_NOISE_INDEP_COMMON) {
    for (int i = 0; i < D; ++i) dB[i] = z[D] * sq;
  } else if (m->noise == SDEMC_NOISE_INDEP_FULL) {
    for (int i = 0; i < D; ++i) {
      double acc = 0.0;
      for (int j = 0; j <= i; ++j) acc += m->L[i][j] * (z[D + j] * sq);
      dB[i] = acc;
    }
  }
}

/* payoff (options.py:167-176, 179-321), undiscounted */
static double payoff_raw(const sdemc_payoff* p, int D, const double* x) {
  double s[D_MAX];
  for (int i = 0; i < D; ++i) s[i] = p->scale * (p->log ? exp(x[i]) : x[i]);
  double K = p->strike, v, mx;
  switch (p->kind) {
    case SDEMC_PAYOFF_EURO_CALL: return s[0] > K ? s[0] - K : 0.0;
    case SDEMC_PAYOFF_EURO_PUT: return s[0] < K ? K - s[0] : 0.0;
    case SDEMC_PAYOFF_BINARY_AON: return s[0] >= K ? s[0] : 0.0;
    case SDEMC_PAYOFF_DIGITAL: return s[0] > K ? 1.0 : 0.0;
    case SDEMC_PAYOFF_BASKET_ARITH:
      v = 0.0;
      for (int i = 0; i < D; ++i) v += s[i];
      v /= D;
      return v > K ? v - K : 0.0;
    case SDEMC_PAYOFF_BASKET_GEO:
      v = 0.0;
      for (int i = 0; i < D; ++i) v += log(s[i]);
      v = exp(v / D);
      return v > K ? v - K : 0.0;
    case SDEMC_PAYOFF_RAINBOW:
      mx = s[0];
      for (int i = 1; i < D; ++i) mx = s[i] > mx ? s[i] : mx;
      return mx > K ? mx - K : 0.0;
    case SDEMC_PAYOFF_BEST_OF:
      mx = s[0];
      for (int i = 1; i < D; ++i) mx = s[i] > mx ? s[i] : mx;
      return mx > K ? mx : K;
    case SDEMC_PAYOFF_ASIAN_CALL: /* options.py:295-298 on the transformed state */
      v = s[D > 1 ? 1 : 0] / p->param;
      if (p->log) v = exp(v);
      return v > K ? v - K : 0.0;
    case SDEMC_PAYOFF_RAINBOW_EVEN: /* options.py:305-311 */
      mx = s[0];
      for (int i = 2; i < D; i += 2) mx = s[i] > mx ? s[i] : mx;
      return mx > K ? mx - K : 0.0;
    default: return s[0];
  }
}

/* ---------------- random sources -------------------------------------- */
typedef struct {
  int fixed;
  uint64_t seed, path; /* RNG mode */
  int64_t p, P;        /* fixed mode */
  const float* dw;
  const uint8_t* counts;
  const int64_t* jump_off;
  const float* jump_time;
  const float* jump_size;
  int64_t n_marks, off;
  int32_t n_rows, max_jumps;
  /* in-kernel jump clock: fp32 arrival time accumulated gap by gap */
  float tau;
  uint32_t j, m1, m2;
  float inv_rate;
  int bad;
} src_t;

static void clock_draw(src_t* s) {
  uint32_t w[4];
  philox_block(s->j, 1, s->path, s->seed, w);
  float gap = -oracle_ln_unit(w[0]) * s->inv_rate;
  s->tau = s->tau + gap;
  s->m1 = w[1];
  s->m2 = w[2];
}

static void clock_init(src_t* s, const sdemc_job* job) {
  s->j = 0;
  s->tau = 0.0f;
  s->inv_rate = job->model.rate > 0.0 ? (float)(1.0 / job->model.rate) : 0.0f;
  if (job->model.rate > 0.0)
    clock_draw(s);
  else
    s->tau = INFINITY;
}

/* step index (grid of n steps on [0,T]) of time tau */
static int32_t step_of(float tau, float n_over_T, int32_t n) {
  float q = tau * n_over_T;
  return q < (float)n ? (int32_t)q : 0x7fffffff;
}

/* marks of step n under per-step Poisson counts: count c = #{k: w < T_k};
 * mark 0 from the offset of w in its count interval, ((w - T_{c+1}) + 1/2)
 * times the interval's reciprocal width (an fp64 product), mark k >= 1 from
 * word (k-1) % 4 of mark-stream block (n, 3 + 16 ((k-1) / 4)); returns the
 * number of marks */
static int poisson_marks(const model_t* m, src_t* s, int32_t n, double* marks, int cap) {
  uint32_t pw[4];
  philox_block((uint32_t)n >> 2, 2, s->path, s->seed, pw);
  uint32_t w = pw[n & 3];
  if (!(w < m->pT[0])) return 0;
  /* count c: T_{c+1} <= w < T_c; the offset within that interval is the
   * first mark's uniform, independent of c */
  int c = 1;
  while (c < 16 && w < m->pT[c]) ++c;
  uint32_t lo = c < 16 ? m->pT[c] : 0u;
  double r = ((double)(w - lo) + 0.5) * m->pinv[c - 1];
  float y = (float)r;
  if (y > 0x1.fffffep-1f) y = 0x1.fffffep-1f;
  for (int k = 0; k < c; ++k) {
    if (k > 0) {
      uint32_t q = (uint32_t)(k - 1), mw[4];
      philox_block((uint32_t)n, 3u + 16u * (q >> 2), s->path, s->seed, mw);
      y = (float)u_open(mw[q & 3u]);
    }
    if (k < cap) marks[k] = levy_icdf(m, y);
  }
  return c;
}

/* fixed-increment row n (uniform grid): D rows, then D+1-th (common) or D more */
static void fixed_row(const model_t* m, src_t* s, int32_t n, double* dW, double* dB) {
  int D = m->D, K = m->K;
  const float* row = s->dw + (int64_t)n * K * s->P + s->p;
  if (m->noise == SDEMC_NOISE_GENERAL) {
    for (int i = 0; i < D_MAX; ++i) {
      dW[i] = row[(int64_t)i * s->P];
      dB[i] = 0.0;
    }
    return;
  }
  for (int i = 0; i < D; ++i) dW[i] = row[(int64_t)i * s->P];
  for (int i = 0; i < D; ++i) dB[i] = 0.0;
  if (m->noise == SDEMC_NOISE_INDEP_COMMON)
    for (int i = 0; i < D; ++i) dB[i] = row[(int64_t)D * s->P];
  else if (m->noise == SDEMC_NOISE_INDEP_FULL)
    for (int i = 0; i < D; ++i) dB[i] = row[(int64_t)(D + i) * s->P];
}

static double fixed_mark(src_t* s) {
  int ok = s->off >= 0 && s->off < s->n_marks;
  double J = ok ? (double)s->jump_size[s->off] : 0.0;
  if (!ok) s->bad = 1;
  s->off++;
  return J;
}

/* jump count and first mark of step n of path `path` in per-step Poisson mode */
ORACLE_API int oracle_step_jumps(const sdemc_job* job, uint64_t path, int32_t n, double* first_mark) {
  model_t m;
  model_init(&m, job);
  if (!m.pois) return -1;
  src_t s;
  memset(&s, 0, sizeof(s));
  s.seed = job->seed;
  s.path = path;
  double marks[16];
  int c = poisson_marks(&m, &s, n, marks, 16);
  if (first_mark) *first_mark = c > 0 ? marks[0] : 0.0;
  return c;
}

/* ---------------- uniform grid ----------------------------------------- */
/* Single level: N steps of h = T/N; the jumps of step n (those with arrival
 * in [nh, (n+1)h)) are added at its end with the coefficient at the
 * pre-step state (or the post-step state with exact_jumps). */
static void uniform_single(const model_t* m, src_t* s, double* x) {
  const sdemc_job* job = m->job;
  int D = m->D, K = m->K, N = job->steps;
  double h = job->T / N, sq = sqrt(h);
  int32_t kj = 0;
  if (!s->fixed && !m->pois) {
    clock_init(s, job);
    kj = step_of(s->tau, m->n_over_T, N);
  }
  for (int i = 0; i < D; ++i) x[i] = job->model.x0[i];
  for (int32_t n = 0; n < N; ++n) {
    double dW[D_MAX], dB[D_MAX], xo[D_MAX], z[2 * D_MAX];
    if (s->fixed) {
      fixed_row(m, s, n, dW, dB);
    } else {
      for (int k = 0; k < K; ++k) z[k] = oracle_normal(s->seed, s->path, (uint32_t)(n * K + k));
      correlate(m, z, sq, dW, dB);
    }
    memcpy(xo, x, sizeof(double) * D);
    euler(m, x, h, dW, dB);
    if (s->fixed) {
      int c = s->counts ? s->counts[(int64_t)n * s->P + s->p] : 0;
      for (int k = 0; k < c; ++k) {
        double J = fixed_mark(s);
        add_jump(m, x, job->model.exact_jumps ? x : xo, J);
      }
    } else if (m->pois) {
      double marks[16];
      int c = poisson_marks(m, s, n, marks, 16);
      for (int k = 0; k < c; ++k) add_jump(m, x, job->model.exact_jumps ? x : xo, marks[k]);
    } else {
      while (kj == n) {
        double J = mark_from_words(m, s->m1, s->m2);
        double y[D_MAX];
        memcpy(y, job->model.exact_jumps ? x : xo, sizeof(double) * D);
        add_jump(m, x, y, J);
        s->j++;
        clock_draw(s);
        kj = step_of(s->tau, m->n_over_T, N);
      }
    }
  }
}

/* Coupled pair on the uniform grid: M fine steps per coarse step; the coarse
 * increment is the sum of the fine ones; fine step n's jumps belong to
 * coarse step n / M and are applied to the coarse path one by one after its
 * diffusion step. */
static void uniform_coupled(const model_t* m, src_t* s, double* xf, double* xc) {
  const sdemc_job* job = m->job;
  int D = m->D, K = m->K, N = job->steps, M = job->coarse_factor;
  double hf = job->T / N, hc = job->T * M / N, sq = sqrt(hf);
  int32_t kj = 0;
  if (!s->fixed && !m->pois) {
    clock_init(s, job);
    kj = step_of(s->tau, m->n_over_T, N);
  }
  for (int i = 0; i < D; ++i) xf[i] = xc[i] = job->model.x0[i];
  double* marks = NULL;
  int cap = 0;
  for (int32_t mc = 0; mc < N / M; ++mc) {
    double sW[D_MAX] = {0}, sB[D_MAX] = {0};
    int nm = 0;
    for (int q = 0; q < M; ++q) {
      int32_t n = mc * M + q;
      double dW[D_MAX], dB[D_MAX], xo[D_MAX], z[2 * D_MAX];
      if (s->fixed) {
        fixed_row(m, s, n, dW, dB);
      } else {
        for (int k = 0; k < K; ++k) z[k] = oracle_normal(s->seed, s->path, (uint32_t)(n * K + k));
        correlate(m, z, sq, dW, dB);
      }
      for (int i = 0; i < m->R; ++i) {
        sW[i] += dW[i];
        sB[i] += dB[i];
      }
      memcpy(xo, xf, sizeof(double) * D);
      euler(m, xf, hf, dW, dB);
      int c;
      double pm[16];
      int pk = 0;
      if (s->fixed) c = s->counts ? s->counts[(int64_t)n * s->P + s->p] : 0;
      else if (m->pois) c = poisson_marks(m, s, n, pm, 16);
      else c = 0;
      for (;;) {
        double J;
        if (s->fixed) {
          if (c-- <= 0) break;
          J = fixed_mark(s);
        } else if (m->pois) {
          if (pk >= c) break;
          J = pm[pk++];
        } else {
          if (kj != n) break;
          J = mark_from_words(m, s->m1, s->m2);
          s->j++;
          clock_draw(s);
          kj = step_of(s->tau, m->n_over_T, N);
        }
        add_jump(m, xf, job->model.exact_jumps ? xf : xo, J);
        if (nm == cap) {
          cap = cap ? 2 * cap : 16;
          marks = (double*)realloc(marks, sizeof(double) * cap);
        }
        marks[nm++] = J;
      }
    }
    double xco[D_MAX];
    memcpy(xco, xc, sizeof(double) * D);
    euler(m, xc, hc, sW, sB);
    for (int k = 0; k < nm; ++k) add_jump(m, xc, job->model.exact_jumps ? xc : xco, marks[k]);
  }
  free(marks);
}

/* ---------------- adapted grid (solvers.py:164-226) -------------------- */
static int isclose32(float tau, float t) {
  volatile float err = fabsf(tau - t);
  volatile float rt = 1e-5f * t;
  volatile float allowed = 1e-12f + fabsf(rt);
  return tau == t || err <= allowed;
}

static float next_time(src_t* s) {
  if (s->fixed) return s->j < (uint32_t)s->max_jumps ? s->jump_time[(int64_t)s->j * s->P + s->p] : INFINITY;
  return s->tau;
}

static double take_mark(const model_t* m, src_t* s) {
  double J;
  if (s->fixed) {
    J = s->j < (uint32_t)s->max_jumps ? (double)s->jump_size[(int64_t)s->j * s->P + s->p] : 0.0;
    s->j++;
  } else {
    J = mark_from_words(m, s->m1, s->m2);
    s->j++;
    clock_draw(s);
  }
  return J;
}

static double adapted_z(const model_t* m, src_t* s, uint32_t row, int k) {
  if (s->fixed) {
    if ((int32_t)row >= s->n_rows) return 0.0;
    return s->dw[((int64_t)row * m->K + k) * s->P + s->p];
  }
  return oracle_normal(s->seed, s->path, row * (uint32_t)m->K + (uint32_t)k);
}

static int32_t adapted_single(const model_t* m, src_t* s, double* x, int* overflow) {
  const sdemc_job* job = m->job;
  int D = m->D, K = m->K;
  volatile float T32 = (float)job->T;
  volatile float h = (float)(job->T / (double)job->steps);
  volatile float t = 0.0f;
  if (!s->fixed) clock_init(s, job);
  else s->j = 0;
  for (int i = 0; i < D; ++i) x[i] = job->model.x0[i];
  float tau = next_time(s);
  int32_t steps = 0;
  *overflow = 0;
  while (t < T32) {
    if (steps >= job->max_steps) {
      *overflow = 1;
      break;
    }
    volatile float rem = T32 - t;
    h = fminf(h, fmaxf(rem, 0.0f));
    volatile float gap = tau - t;
    volatile float dt = fminf(h, gap);
    double sq = sqrt((double)dt);
    double z[2 * D_MAX], dW[D_MAX], dB[D_MAX], xo[D_MAX];
    for (int k = 0; k < K; ++k) z[k] = adapted_z(m, s, (uint32_t)steps, k);
    correlate(m, z, sq, dW, dB);
    memcpy(xo, x, sizeof(double) * D);
    euler(m, x, (double)dt, dW, dB);
    t = t + dt;
    steps++;
    if (isclose32(tau, t)) {
      double J = take_mark(m, s);
      add_jump(m, x, job->model.exact_jumps ? x : xo, J);
      tau = next_time(s);
    }
  }
  return steps;
}

/* solvers.py:228-307 */
static int32_t adapted_coupled(const model_t* m, src_t* s, double* xf, double* xc, int* overflow) {
  const sdemc_job* job = m->job;
  int D = m->D, K = m->K, M = job->coarse_factor;
  volatile float T32 = (float)job->T;
  volatile float hf = (float)(job->T / (double)job->steps);
  volatile float hc = (float)M * hf;
  volatile float tf = 0.0f, tc = 0.0f;
  if (!s->fixed) clock_init(s, job);
  else s->j = 0;
  for (int i = 0; i < D; ++i) xf[i] = xc[i] = job->model.x0[i];
  float tau = next_time(s);
  int32_t it = 0;
  *overflow = 0;
  while (tf < T32) {
    if (it >= job->max_steps) {
      *overflow = 1;
      break;
    }
    double sW[D_MAX] = {0}, sB[D_MAX] = {0}, xfo[D_MAX];
    memcpy(xfo, xf, sizeof(double) * D);
    for (int q = 0; q < M; ++q) {
      volatile float rem = T32 - tf;
      hf = fminf(hf, fmaxf(rem, 0.0f));
      volatile float gap = tau - tf;
      volatile float dt = fminf(hf, gap);
      double sq = sqrt((double)dt);
      double z[2 * D_MAX], dW[D_MAX], dB[D_MAX];
      for (int k = 0; k < K; ++k) z[k] = adapted_z(m, s, (uint32_t)(it * M + q), k);
      correlate(m, z, sq, dW, dB);
      memcpy(xfo, xf, sizeof(double) * D);
      euler(m, xf, (double)dt, dW, dB);
      tf = tf + dt;
      for (int i = 0; i < D; ++i) {
        sW[i] += dW[i];
        sB[i] += dB[i];
      }
    }
    volatile float remc = T32 - tc;
    hc = fminf(hc, fmaxf(remc, 0.0f));
    volatile float gapc = tau - tc;
    volatile float dtc = fminf(hc, gapc);
    double xco[D_MAX];
    memcpy(xco, xc, sizeof(double) * D);
    euler(m, xc, (double)dtc, sW, sB);
    tc = tc + dtc;
    if (isclose32(tau, tf)) {
      double J = take_mark(m, s);
      add_jump(m, xf, job->model.exact_jumps ? xf : xfo, J);
      add_jump(m, xc, job->model.exact_jumps ? xc : xco, J);
      tau = next_time(s);
    }
    it++;
  }
  return it;
}

/* ---------------- entry points ---------------------------------------- */
static void run_path(const model_t* m, src_t* s, double* payoff, double* term, int32_t* nsteps,
                     int64_t idx, int64_t n) {
  const sdemc_job* job = m->job;
  int D = m->D, ovf = 0;
  int32_t ns = job->steps;
  double x[D_MAX], xc[D_MAX];
  int coupled = job->coarse_factor > 0;
  if (job->grid == SDEMC_GRID_UNIFORM) {
    if (coupled) uniform_coupled(m, s, x, xc);
    else uniform_single(m, s, x);
  } else {
    if (coupled) ns = adapted_coupled(m, s, x, xc, &ovf);
    else ns = adapted_single(m, s, x, &ovf);
  }
  double v = payoff_raw(&job->payoff, D, x);
  if (coupled) v -= payoff_raw(&job->payoff, D, xc);
  v *= job->payoff.disc;
  if (ovf || s->bad) v = NAN;
  if (payoff) payoff[idx] = v;
  if (term) {
    for (int i = 0; i < D; ++i) term[(int64_t)i * n + idx] = x[i];
    if (coupled)
      for (int i = 0; i < D; ++i) term[(int64_t)(D + i) * n + idx] = xc[i];
  }
  if (nsteps) nsteps[idx] = ns;
}

/* RNG mode: paths p0 .. p0+n-1 of the job, drawn from the kernels' counters. */
ORACLE_API int oracle_simulate(const sdemc_job* job, int64_t p0, int64_t n, double* payoff,
                               double* term, int32_t* nsteps) {
  model_t m;
  model_init(&m, job);
  for (int64_t q = 0; q < n; ++q) {
    src_t s;
    memset(&s, 0, sizeof(s));
    s.seed = job->seed;
    s.path = job->path_offset + (uint64_t)(p0 + q);
    run_path(&m, &s, payoff, term, nsteps, q, n);
  }
  return 0;
}

/* Fixed-increment mode: the same arrays sdemc_simulate_fixed reads. */
ORACLE_API int oracle_simulate_fixed(const sdemc_job* job, const sdemc_fixed_in* in, int64_t P,
                                     double* payoff, double* term, int32_t* nsteps) {
  model_t m;
  model_init(&m, job);
  for (int64_t p = 0; p < P; ++p) {
    src_t s;
    memset(&s, 0, sizeof(s));
    s.fixed = 1;
    s.p = p;
    s.P = P;
    s.dw = in->dw_dev;
    s.counts = in->counts_dev;
    s.jump_off = in->jump_off_dev;
    s.jump_time = in->jump_time_dev;
    s.jump_size = in->jump_size_dev;
    s.n_marks = in->n_marks;
    s.n_rows = in->n_rows;
    s.max_jumps = in->max_jumps;
    s.off = in->jump_off_dev ? in->jump_off_dev[p] : 0;
    run_path(&m, &s, payoff, term, nsteps, p, P);
  }
  return 0;
}

ORACLE_API int oracle_abi_version(void) { return SDEMC_ABI_VERSION; }

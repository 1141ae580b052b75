/*
 * sdemc.h -- C ABI of the MI355X (gfx950) Monte Carlo path-integration library.
 *
 * The library prices functionals of jump-diffusion and Levy-driven SDEs by
 * Euler-Maruyama stepping, one path per GPU lane, with the path state held in
 * registers and the estimator sums reduced in fp64.  It replaces the hot path
 * of piers-hinds/sde_mc:
 *
 *   JumpDiffusionSolver.solve            sde_mc/solvers.py:164-226
 *   JumpDiffusionSolver.multilevel_solve sde_mc/solvers.py:228-307
 *   DiffusionSolver.solve / multilevel   sde_mc/solvers.py:68-119
 *   EulerScheme.step                     sde_mc/schemes.py:5-13
 *   Option.__call__ + discounter         sde_mc/options.py:167-176, 324-337
 *   the (sum, sum of squares) reduction  sde_mc/mc.py:105-120, sde_mc/mlmc.py:45-69
 *
 * Every entry point returns an int status (SDEMC_OK == 0) and never throws.
 * Arguments are validated on the host before anything is launched.  Device
 * buffers are owned by the caller; the library allocates no device memory
 * (sdemc_batch_plan_create allocates HIP graph objects, which
 * sdemc_batch_plan_destroy releases).
 * All pointers named *_dev are device pointers; `stream` is a hipStream_t
 * (NULL = the default stream).
 */
#ifndef SDEMC_H
#define SDEMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDEMC_ABI_VERSION 5
#define SDEMC_MAX_DIM 4
#define SDEMC_CHUNK 64 /* paths per reduction chunk (one wavefront) */

/* ---- status codes ------------------------------------------------------ */
enum sdemc_status {
  SDEMC_OK = 0,
  SDEMC_ERR_ARG = 1,         /* invalid argument (checked on the host)          */
  SDEMC_ERR_UNSUPPORTED = 2, /* no kernel for this model / payoff / grid         */
  SDEMC_ERR_LAUNCH = 3,      /* a HIP call failed (see sdemc_last_hip_error)     */
  SDEMC_ERR_WORKSPACE = 4,   /* workspace too small                              */
  SDEMC_ERR_OVERFLOW = 5,    /* a path needed more grid steps than max_steps     */
  SDEMC_ERR_DEVICE = 6       /* no usable device                                 */
};

/* ---- model description --------------------------------------------------
 * Coefficients are affine in the state, per component i < dim:
 *   drift      a_i(x) = a0[i] + a1[i] * x_i
 *   diffusion  b_i(x) = b0[i] + b1[i] * x_i   (driven by the correlated BM W)
 *   second     c_i(x) = c0[i] + c1[i] * x_i   (driven by B, see `noise`)
 *   jump       g_i(x) = j0[i] + j1[i] * x_i   (multiplies the jump mark J)
 * One Euler step over dt with increments dW (= chol * z * sqrt(dt)) is
 *   x_i <- x_i + a_i(x) dt + b_i(x) dW_i + c_i(x) dB_i,
 * and a jump with mark J adds g_i(y) J with y the state before the step
 * (exact_jumps == 0, solvers.py:216-217) or after it (exact_jumps == 1,
 * solvers.py:214-215).  All jumps are common to every component
 * (solvers.py:146-148).
 */
enum sdemc_noise {
  SDEMC_NOISE_DIAG = 0,         /* 'diag':  dB = 0                             sde.py:24-25   */
  SDEMC_NOISE_INDEP_COMMON = 1, /* 'indep' in the jump solver: dB_i = one scalar
                                   N(0,dt) shared by all components
                                   (small-jump Gaussian, solvers.py:198-201)          */
  SDEMC_NOISE_INDEP_FULL = 2,   /* 'indep' in the diffusion solver: B is a second
                                   correlated BM independent of W (solvers.py:80-81)   */
  SDEMC_NOISE_HESTON = 3,       /* Heston (sde.py:258-279, schemes.py:16-21): dim 2,
                                   x = (S, v), correlated 'diag' increments dW;
                                   S += a1[0] S dt + sqrt(v) S dW_0 and sqrt(v_new)
                                   is the positive root of the drift-implicit
                                   equation (1 + kappa dt) y^2 - xi dW_1 y
                                   - (v + (kappa theta - xi^2/2) dt) = 0
                                   (helpers.py:36-48); jp = {kappa, theta, xi};
                                   no jumps, uniform grid, 2 kappa theta > xi^2  */
  SDEMC_NOISE_GENERAL = 4       /* 'general' (sde.py:31-32, schemes.py:12): a
                                   (dim, brown_dim) diffusion matrix applied to
                                   the correlated brown_dim-dimensional BM,
                                   B(x)_ij = bm0[i*4+j] + bm1[i*4+j] * x_i:
                                   x_i <- x_i + a_i(x) dt + sum_j B(x)_ij dW_j,
                                   dW = chol * z * sqrt(dt) with chol
                                   brown_dim x brown_dim; b0, b1, c0, c1
                                   unused; uniform grid                      */
};

enum sdemc_jump_law {
  SDEMC_JUMP_LOGNORMAL = 0, /* J = exp(jp[0] + jp[1] * Z) - 1            sde.py:325-326 */
  SDEMC_JUMP_LEVY = 1       /* J = InverseCdf(U + 2^-24/3)               levy.py:10-30, 85-87
                               jp = {c_minus, c_plus, 1/mu, alpha, eps, lambda,
                                     y1, y2, y3, eps^-alpha, -1/alpha}          */
};

typedef struct sdemc_model {
  int32_t dim;         /* 1..SDEMC_MAX_DIM                                  */
  int32_t noise;       /* enum sdemc_noise                                  */
  int32_t jump_law;    /* enum sdemc_jump_law                               */
  int32_t exact_jumps; /* 0: jump coefficient at pre-step state, 1: post    */
  double rate;         /* jump intensity lambda (0 = no jumps)              */
  double x0[SDEMC_MAX_DIM];
  double a0[SDEMC_MAX_DIM], a1[SDEMC_MAX_DIM];
  double b0[SDEMC_MAX_DIM], b1[SDEMC_MAX_DIM];
  double c0[SDEMC_MAX_DIM], c1[SDEMC_MAX_DIM];
  double j0[SDEMC_MAX_DIM], j1[SDEMC_MAX_DIM];
  double chol[SDEMC_MAX_DIM * SDEMC_MAX_DIM]; /* lower Cholesky factor, row-major */
  double jp[12];                              /* jump-law parameters            */
  double ax[SDEMC_MAX_DIM]; /* drift of component i gains ax[i] * x_0 ('diag' noise,
                               ax[0] = 0): AsianWrapper's running integral
                               dA = S dt (sde.py:378-406)                      */
  int32_t brown_dim;        /* 'general' noise: Brownian dimension 1..4        */
  int32_t reserved;
  double bm0[SDEMC_MAX_DIM * SDEMC_MAX_DIM]; /* 'general' noise: B(x) = bm0 + diag(x) bm1, */
  double bm1[SDEMC_MAX_DIM * SDEMC_MAX_DIM]; /* row-major (dim, brown_dim) in stride 4     */
} sdemc_model;

/* ---- payoff (options.py:156-321) ---------------------------------------- */
enum sdemc_payoff_kind {
  SDEMC_PAYOFF_EURO_CALL = 0,  /* options.py:179-199 */
  SDEMC_PAYOFF_EURO_PUT = 1,   /* options.py:202-222 */
  SDEMC_PAYOFF_BINARY_AON = 2, /* options.py:225-244 */
  SDEMC_PAYOFF_DIGITAL = 3,    /* options.py:275-284 */
  SDEMC_PAYOFF_BASKET_ARITH = 4, /* options.py:247-259 */
  SDEMC_PAYOFF_BASKET_GEO = 5,   /* options.py:247-259 */
  SDEMC_PAYOFF_RAINBOW = 6,    /* call on max, options.py:262-272 */
  SDEMC_PAYOFF_BEST_OF = 7,    /* max(max_i S_i, K), options.py:314-321 */
  SDEMC_PAYOFF_IDENTITY = 8,   /* x_0 itself (martingale checks)   */
  SDEMC_PAYOFF_ASIAN_CALL = 9, /* call on s_1 / param, param = the averaging time
                                  (exponentiated when log), options.py:287-298 */
  SDEMC_PAYOFF_RAINBOW_EVEN = 10 /* call on max of s_0, s_2, ..., options.py:301-311 */
};

typedef struct sdemc_payoff {
  int32_t kind;    /* enum sdemc_payoff_kind                                   */
  int32_t log;     /* Option.log: exponentiate the state first (options.py:168) */
  double strike;
  double scale;    /* Option.discount: spot multiplier (options.py:169)         */
  double disc;     /* discounter(T) applied to the payoff (mc.py:93)            */
  double param;    /* kind-specific: ASIAN_CALL's averaging time                */
} sdemc_payoff;

/* ---- one pricing job ------------------------------------------------------ */
enum sdemc_grid {
  SDEMC_GRID_UNIFORM = 0, /* N steps of T/N, per-step Poisson jump counts        */
  SDEMC_GRID_ADAPTED = 1  /* jump-adapted grid of solvers.py:182-225, fp32 clock  */
};

typedef struct sdemc_job {
  sdemc_model model;
  sdemc_payoff payoff;
  double T;              /* time_interval                                         */
  int32_t steps;         /* num_steps (the fine level when coarse_factor > 0)       */
  int32_t grid;          /* enum sdemc_grid                                       */
  int32_t coarse_factor; /* 0: single level; M >= 2: coupled fine/coarse pair with
                            steps / M coarse steps (solvers.py:228-307)            */
  int32_t max_steps;     /* adapted grid: cap on grid steps per path (the reference's
                            storage bound N + max_jumps, solvers.py:133,170)       */
  uint64_t n_paths;
  uint64_t seed;         /* Philox key                                            */
  uint64_t path_offset;  /* index of the first path in the Philox counter space   */
} sdemc_job;

/* ---- optional per-path outputs (plain entry points only) -----------------
 * Layouts are structure-of-arrays with the path index fastest, so one
 * wavefront's 64 lanes touch 256 contiguous bytes per row.
 */
typedef struct sdemc_path_out {
  float* payoff_dev;    /* [P] discounted payoff (coupled: fine minus coarse)        */
  float* terminal_dev;  /* [dim][P] state at T (coupled: [2][dim][P], fine first)    */
  int32_t* nsteps_dev;  /* [P] grid steps taken (adapted grid)                      */
  float* paths_dev;     /* [S+1][dim][P] states after each step, S = steps (uniform)
                           or max_steps (adapted); rows past a path's end repeat
                           its terminal state                                       */
  float* times_dev;     /* adapted only, [S+1][P] grid times                         */
  float* left_dev;      /* adapted only, [S+1][dim][P] pre-jump states              */
  float* jumps_dev;     /* adapted only, [S+1][dim][P] applied jump marks           */
  float* normals_dev;   /* [S][K][P] increments used (scaled and correlated);
                           K = dim ('diag'), 2*dim ('indep': column-major pairs),
                           4 ('general': dW_0..3, zero beyond brown_dim)            */
} sdemc_path_out;

/* ---- host-supplied random inputs (fixed-increment mode) ------------------
 * Uniform grid:  dw     [N][K][P] increments, already scaled by sqrt(dt) and
 *                       correlated (the reference's `normals`, solvers.py:206);
 *                       K = dim ('diag'), dim + 1 ('indep' common: the last row
 *                       is the shared small-jump increment), 2*dim ('indep' full),
 *                       4 ('general': dW_0..3; rows beyond brown_dim meet zero
 *                       matrix columns)
 *                counts [N][P] jumps arriving in step n
 *                jump_off [P+1] CSR offsets into jump_size (int64)
 *                jump_size marks J, path-major in arrival order
 * Adapted grid:  dw     [S][K][P] standard normals z (scaled in the kernel)
 *                jump_time [J][P] cumulative arrival times (fp32)
 *                jump_size [J][P] marks J
 * Coupled jobs use the fine grid's arrays.
 */
typedef struct sdemc_fixed_in {
  const float* dw_dev;
  const uint8_t* counts_dev;
  const int64_t* jump_off_dev;
  const float* jump_time_dev;
  const float* jump_size_dev;
  int64_t n_marks;   /* uniform: length of jump_size (reads outside it are
                        refused in the kernel and reported as SDEMC_ERR_ARG) */
  int32_t n_rows;    /* rows in dw (N or S; coupled adapted: S * M)          */
  int32_t max_jumps; /* adapted: J                                            */
} sdemc_fixed_in;

/* ---- statistics ------------------------------------------------------------
 * stats_dev receives 4 doubles per job: {sum, sum of squares, n, flags}
 * where flags is an OR of enum sdemc_flag bits (0: a valid result).  The sums
 * are taken in fp64 over fixed 64-path chunks and then over chunks in a fixed
 * order, so a job's result depends only on the job, never on the launch it
 * ran in.
 */
enum sdemc_flag {
  SDEMC_FLAG_OVERFLOW = 1, /* a path needed more grid steps than max_steps          */
  SDEMC_FLAG_JUMP_CSR = 2, /* fixed mode: a jump mark read outside jump_size         */
  SDEMC_FLAG_CLASS = 4,    /* batched: a job outside its group's kernel class        */
  SDEMC_FLAG_TABLE = 8     /* batched: the launch's segment table does not describe
                              the prepared job table (see sdemc_batch_check); the
                              sums are not a result                                 */
};

/* Library identity. */
int sdemc_abi_version(void);
const char* sdemc_status_string(int status);
int sdemc_last_hip_error(void);
size_t sdemc_sizeof_job(void);
size_t sdemc_sizeof_model(void);
size_t sdemc_sizeof_path_out(void);
size_t sdemc_sizeof_fixed_in(void);

/* Workspace (bytes) one plain launch needs: fp64 partials for every chunk. */
size_t sdemc_workspace_bytes(uint64_t n_paths); /* 2 doubles + 1 int32 per chunk */

/* Device properties: CU count, clock (kHz), wavefront size. */
int sdemc_device_info(int device, int* cu_count, int* clock_khz, int* wave_size);

/* Plain path, in-kernel Philox (RNG mode).  One launch for the paths, one for
 * the fixed-order finalisation.  Replaces JumpDiffusionSolver.solve +
 * payoff + the mc_simple sums (solvers.py:164-226, mc.py:87-117); with
 * coarse_factor > 0 it replaces multilevel_solve + the MLMC level sums
 * (solvers.py:228-307, mlmc.py:61-69).  `out` may be NULL. */
int sdemc_simulate(const sdemc_job* job, double* stats_dev, const sdemc_path_out* out,
                   void* workspace_dev, size_t workspace_bytes, int waves_per_block,
                   void* stream);

/* The same plain path with no workspace argument and no allocation of any
 * kind: the chunk partials live in a fixed device array of the library
 * (SDEMC_NOALLOC_SCRATCH_BYTES, one per GPU, held until the process ends),
 * so callers that cannot allocate -- code under HIP graph capture, serving
 * loops that forbid allocation -- can price.  Jobs needing more workspace
 * than the array holds (sdemc_workspace_bytes(n_paths) >
 * SDEMC_NOALLOC_SCRATCH_BYTES, i.e. above SDEMC_NOALLOC_MAX_PATHS paths) get
 * SDEMC_ERR_WORKSPACE.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other host threads) race.  No
 * counterpart in the reference; its result equals sdemc_simulate's. */
#define SDEMC_NOALLOC_MAX_PATHS (1ull << 24)
#define SDEMC_NOALLOC_SCRATCH_BYTES (5u << 20) /* >= 20 B x 2^18 chunks, 256 B aligned */
int sdemc_simulate_noalloc(const sdemc_job* job, double* stats_dev, const sdemc_path_out* out,
                           int waves_per_block, void* stream);

/* Plain path fed with host-generated increments, jump counts/times and marks
 * instead of in-kernel draws (parity mode; the HBM-bound variant). */
int sdemc_simulate_fixed(const sdemc_job* job, const sdemc_fixed_in* in, double* stats_dev,
                         const sdemc_path_out* out, void* workspace_dev,
                         size_t workspace_bytes, void* stream);

/* Reduce chunk partials {sum, sumsq} of n_jobs jobs into stats (4 doubles each).
 * chunk_base[j] is the first partial of job j, n_paths[j] its path count;
 * chunk_flags holds one overflow flag per chunk; all arrays live on the device.  Replaces mc_estimates' inputs
 * (helpers.py:51-68) with fp64 sums in a fixed order. */
int sdemc_reduce_finalize(const double* partials_dev, const int32_t* chunk_flags_dev,
                          const int64_t* chunk_base_dev, const uint64_t* n_paths_dev,
                          int32_t n_jobs, double* stats_dev, void* stream);

/* ---- neural control variates (mc.py:126-326, varred.py:77-131, 202-214) -
 * ReLU MLPs evaluated with bf16 matrix-core (MFMA) products and fp32
 * accumulation at every step of every path give the control-variate
 * estimators of apply_diffusion_control_variate (uniform grid, no jumps) and
 * apply_adapted_control_variates (jump-adapted grid):
 *   gamma = disc(T) payoff(X_T)
 *         + sum_{n < cv_steps} disc(t_n) <f(t_n, X_n), dW_n>          (Brownian)
 *         + sum_n disc(t_n) <g(t_n, X_n-), 1> J_n                     (jumps)
 *         - jump_comp sum_n disc(t_n) <g(t_n, X_n-), 1> (t_{n+1} - t_n)
 * with disc(t) = exp(-rate t), X_n the state at grid time t_n (after any
 * jump there), X_n- the state before it, J_n the jump mark applied at t_n
 * (0 at most times) and dW_n the step's correlated increments ('diag': dim
 * components; 'indep': dim pairs (dW_i, dB_i) matched with f outputs
 * 2i, 2i + 1).  The control variate has mean zero for any f and g (up to the
 * time discretisation of the compensator), so rounding the networks to bf16
 * costs variance reduction, never bias.  Paths are those of sdemc_simulate
 * for the same job (same Philox counters): f = g = 0 gives its payoffs
 * exactly.  Every row of every path is included; the reference truncates a
 * batch at its longest path's final row (varred.py:113-118 on
 * nets.py:192-207).
 *
 * Jobs: single level; uniform grid without jumps ('diag' dim 1..4, 'indep'
 * Double-GBM dim 1..4, Heston) or adapted grid with lognormal or Levy jumps
 * ('diag' / 'indep', dim 1..4); 'indep' needs 2 dim <= 8 f outputs.
 *
 * A network arrives as bf16 MFMA operand words built by the host
 * (sde_mc_amd/cv.py, pack_network): hidden layers of at most 63 units padded
 * to 64, the first padded unit a constant 1 carrying the next layer's bias;
 * inputs (t, x_1..x_dim, 1) padded to 8 and fed as bf16 high and low parts
 * (the first layer's weights are shared by both); outputs padded to 32.
 * The words form blocks of 512, word 512 b + 8 l + q of block b holding for
 * lane l = 0..63 and q = 0..7 the transposed weight W^T[o][i] (output unit
 * o, input unit i):
 *   first layer, blocks b = m (m = 0, 1):        W^T[32 m + l % 32][q]
 *   hidden layer k = 1..L-1, b = 2 + 8 (k - 1) + 4 m + c (m = 0, 1; c = 0..3):
 *                                                W^T[32 m + l % 32][u(c, l / 32, q)]
 *   output layer, b = 2 + 8 (L - 1) + c (c = 0..3): W^T[l % 32][u(c, l / 32, q)]
 * with u(c, h, q) = 16 c + 8 (q / 4) + 4 h + q % 4, the order in which a
 * 32x32 accumulator tile's rows sit in the lanes: 8 L - 2 blocks in all.
 */
#define SDEMC_CV_MAX_HIDDEN 4
#define SDEMC_CV_MAX_OUT 8
typedef struct sdemc_cv_mlp {
  const uint16_t* weights_dev; /* sdemc_cv_mlp_words(hidden_layers) bf16 words */
  int32_t hidden_layers;       /* 1..SDEMC_CV_MAX_HIDDEN; 0 = the zero network */
  int32_t n_out;               /* f: brown_dim, g: dim                         */
} sdemc_cv_mlp;

typedef struct sdemc_cv_spec {
  sdemc_cv_mlp f;    /* Brownian integrand f(t, X_t)                         */
  sdemc_cv_mlp g;    /* jump integrand g(t, X_t-); adapted grid only          */
  int32_t cv_steps;  /* Brownian terms of steps n < cv_steps (remove_steps)   */
  int32_t reserved;
  double rate;       /* discount exp(-rate t) of the integrands               */
  double jump_comp;  /* jump_rate * jump_mean of the compensator              */
} sdemc_cv_spec;

/* bf16 words of a packed network with this many hidden layers. */
size_t sdemc_cv_mlp_words(int32_t hidden_layers);

/* Control-variate run: stats_dev gets {sum, sum of squares, n, flags} of
 * gamma; gamma_dev (optional, [P] fp32) the per-path values. */
int sdemc_cv_simulate(const sdemc_job* job, const sdemc_cv_spec* cv, double* stats_dev,
                      float* gamma_dev, void* workspace_dev, size_t workspace_bytes,
                      void* stream);

/* ---- batched persistent launch ------------------------------------------
 * A job table of max_jobs slots lives on the device in the library's own
 * prepared layout (sdemc_sizeof_prepared_job() bytes per slot), filled by
 * the caller with bytes from sdemc_prepare_jobs.  A launch runs a list of
 * segments: segment s covers chunks [first, first + count) of slot `job`,
 * cut into tiles of `per_tile` chunks; segment s owns tiles
 * [tile_begin, tile_begin + ceil(count / per_tile)), the segments' tile
 * ranges are consecutive from 0, and the list order is the order in which
 * tiles are handed out.  Persistent wavefronts take tile indices from a
 * device counter until all tiles are taken, so jobs of any size and any mix
 * of models, grids and MLMC levels share one launch.  Partials of slot j
 * live at partials[chunk_base_j + c].  A launch whose segments name an
 * unprepared slot, chunks outside a slot or the partial pool, or tile ranges
 * that do not cover their chunks runs what it can and sets SDEMC_FLAG_TABLE in
 * every slot it finalises; sdemc_batch_check finds the same faults on the
 * host before a launch.
 */
typedef struct sdemc_segment {
  int32_t job;        /* slot index                                     */
  int32_t first;      /* first chunk                                    */
  int32_t count;      /* chunks in the segment                          */
  int32_t per_tile;   /* chunks per tile (>= 1)                         */
  int32_t tile_begin; /* index of the segment's first tile              */
  int32_t reserved;
} sdemc_segment;

/* Bytes of one prepared job slot. */
size_t sdemc_sizeof_prepared_job(void);

/* Kernel classes of the batched launch: a dedicated persistent kernel for
 * each of the 1-D 'diag' (Merton, GBM) and 2-D 'indep' (Levy) families on
 * either grid, single-level or coupled, and one generic kernel that runs
 * every supported job.  On the uniform grid at a single level each
 * dedicated kernel holds one jump mode: the 2-D family has a kernel for
 * per-step Poisson counts (class 4: Levy law at 1/16 or more expected jumps
 * per step) and one for the arrival clock (class 8: every other job); the
 * 1-D kernel (class 0) takes arrival-clock jobs and leaves 1-D per-step
 * count jobs to the generic kernel. */
#define SDEMC_MAX_GROUPS 8
enum sdemc_class {
  SDEMC_CLASS_DIAG1 = 0,  /* + 2 * grid + coupled: classes 0..3 */
  SDEMC_CLASS_INDEP2 = 4, /* + 2 * grid + coupled: classes 4..7 */
  SDEMC_CLASS_INDEP2_CLOCK = 8, /* 2-D 'indep', uniform grid, single level, arrival clock */
  SDEMC_CLASS_GENERIC = 9
};

/* The dedicated class of a job, SDEMC_CLASS_GENERIC, or -1 for an invalid job. */
int sdemc_job_class(const sdemc_job* job);

/* Validate n host jobs and write their prepared slots (host memory,
 * n * sdemc_sizeof_prepared_job() bytes).  chunk_base[i] is the first chunk
 * partial of job i in the batch partials array.  classes (nullable)
 * receives each job's class. */
int sdemc_prepare_jobs(const sdemc_job* jobs, int32_t n, const int64_t* chunk_base,
                       void* prepared_out, int32_t* classes);

/* A launch runs n_groups persistent kernels in stream order, one per class
 * group; group g's segments are segs[hdr[g][0] .. hdr[g][0] + hdr[g][1])
 * with hdr[g][2] tiles, taken from counter[g].  A segment whose job is not
 * of the group's class is refused in the kernel (NaN partials, chunk
 * flag 4) instead of being computed by the wrong code. */
typedef struct sdemc_batch_dev {
  const void* jobs_dev;          /* [max_jobs] prepared slots                   */
  const sdemc_segment* segs_dev; /* [max_segments], each group's contiguous     */
  const int32_t* seg_hdr_dev;    /* [SDEMC_MAX_GROUPS][4] {first segment,
                                    segments, tiles, 0}, read on the device     */
  const int32_t* finalize_dev;   /* [max_jobs] nonzero: finalise slot this launch */
  double* partials_dev;          /* [partial_chunks][2]                         */
  int32_t* chunk_flags_dev;      /* [partial_chunks] flags: zero when first
                                    handed to the library; the group kernels
                                    write only set flags and the finalisation
                                    resets those it reads                      */
  double* stats_dev;             /* [max_jobs][4]                               */
  int32_t* counter_dev;          /* [SDEMC_MAX_GROUPS + 1]: the groups' tile
                                    counters, then the launch's fault word; all
                                    zeroed by the launch                        */
  int64_t partial_chunks;        /* capacity of partials / chunk_flags          */
  int32_t max_jobs;
  int32_t max_segments;
  int32_t n_groups;              /* 1..SDEMC_MAX_GROUPS                         */
  int32_t group_class[SDEMC_MAX_GROUPS]; /* enum sdemc_class of each group      */
  int32_t reserved;
} sdemc_batch_dev;

typedef struct sdemc_batch_tune {
  int32_t waves_per_block;   /* 1, 2, 4 or 8                                  */
  int32_t blocks_per_cu;     /* persistent workgroups per CU                  */
  int32_t concurrent_groups; /* 0: group kernels run one after another on the
                                stream; 1: forked onto side streams (default
                                priority) and joined back, so a group's tail
                                overlaps the next group                        */
  int32_t step_unroll;       /* 4-step groups per loop iteration of the uniform
                                single-level path: 0 or 1 (one), 2 (two)       */
} sdemc_batch_tune;

/* Host-side check of a launch's tables before they are uploaded: the
 * prepared slots (host copy, max_jobs * sdemc_sizeof_prepared_job() bytes),
 * the segment list, the header [SDEMC_MAX_GROUPS][4] and the group classes.
 * Per group: the header names segments inside the list; their tile ranges
 * run consecutively from 0, each of ceil(count / per_tile) tiles, and add up
 * to the header's tile count; each segment names a prepared slot of the
 * group's class (any class in the generic group), chunks inside that slot's
 * job and partial slots inside the pool.  SDEMC_OK or SDEMC_ERR_ARG.
 * No counterpart in the reference (it has no batched launch); it replaces
 * the planner's trust in its own tables with a check. */
int sdemc_batch_check(const void* prepared_host, int32_t max_jobs, const sdemc_segment* segs,
                      int32_t max_segments, const int32_t* hdr, int32_t n_groups,
                      const int32_t* group_class, int64_t partial_chunks);

/* One batched launch: zero the tile counters, run the group kernels,
 * finalise the flagged slots.  Everything is ordered after earlier work on
 * `stream`, and later work on `stream` waits for all of it. */
int sdemc_batch_launch(const sdemc_batch_dev* b, const sdemc_batch_tune* tune, void* stream);

/* The same launches captured once into a HIP graph and replayed by
 * sdemc_batch_plan_launch; the graph reads the segment header and the
 * tables from the device, so one plan serves every launch that uses the
 * same buffers. */
typedef struct sdemc_batch_plan sdemc_batch_plan;
int sdemc_batch_plan_create(const sdemc_batch_dev* b, const sdemc_batch_tune* tune,
                            sdemc_batch_plan** plan);
int sdemc_batch_plan_launch(sdemc_batch_plan* plan, void* stream);
int sdemc_batch_plan_destroy(sdemc_batch_plan* plan);

/* Philox4x32-10 block (Salmon et al., SC'11) evaluated on the device for
 * n counters: out[4i..4i+3] = philox(ctr_i, key).  ctr is [n][4] uint32.
 * Exposed for known-answer tests of the in-kernel generator. */
int sdemc_philox_dev(const uint32_t* ctr_dev, uint32_t key0, uint32_t key1, uint32_t* out_dev,
                     int64_t n, void* stream);

/* HBM streaming-read probe (no counterpart in the reference; it measures the
 * read rate the fixed-increment kernels are held against): `blocks`
 * workgroups of 256 threads read `bytes` (a multiple of 16, src 16-byte
 * aligned) from src_dev as 16-byte words at a grid stride, each thread
 * XOR-folding its words into one uint32 of out_dev[blocks * 256] (so the XOR
 * of out_dev is the XOR of every 32-bit word of the source). */
int sdemc_hbm_read_probe_dev(const void* src_dev, int64_t bytes, uint32_t* out_dev,
                             int32_t blocks, void* stream);

/* VALU issue-rate probes (no counterpart in the reference; they measure the
 * ceilings the path kernels' roofline is priced at): `blocks` workgroups of
 * 256 threads each run `iters` iterations of 32 independent operations of one
 * kind and write one float to out_dev[blocks * 256].  kind 0: fp32 FMA
 * (v_fma_f32; a launch issues blocks * 4 * 32 * iters wavefront FMAs);
 * 1: 32-bit integer xor + add; 2: 32 x 32 -> 64-bit multiply; 3: fp32
 * exp2 and two multiplies; 4: one Philox4x32-10 block per iteration (not 32
 * operations); 5: a Philox block and the three Box-Muller pairs of its six
 * normals per iteration; 6: two independent Philox blocks per iteration. */
int sdemc_valu_probe_dev(float* out_dev, int32_t blocks, int32_t iters, int32_t kind,
                         void* stream);

/* Standard normals exactly as the stepping kernels draw them: normal i of
 * path p (i = step * K + k) for p in [0, n_paths), i in [0, n_per_path).
 * out is [n_per_path][n_paths]. */
int sdemc_normals_dev(uint64_t seed, uint64_t path_offset, int64_t n_paths, int32_t n_per_path,
                      float* out_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SDEMC_H */

// cv_mlp.h -- neural control variates on the matrix cores: every step of
// every path evaluates ReLU MLPs f(t, x) (and g(t, x-) for jump models) with
// bf16 v_mfma_f32_32x32x16_bf16 products and fp32 accumulation, and
// accumulates the discounted integrals of the control variate next to the
// path (varred.py:77-131, 202-214; mc.py:126-326).
//
// One wavefront carries 64 paths.  A layer is computed transposed,
// Z^T = W^T A^T, in 32x32 tiles: output units on the tile rows, paths on the
// columns, so path tile 0 holds paths 0..31 and tile 1 paths 32..63 of the
// wavefront.  In the instruction lane l supplies row l % 32 of A and column
// l % 32 of B, each with the 8 K positions of its half h = l / 32, and
// receives C rows 8 (r / 4) + 4 h + r % 4 of column l % 32 in accumulator r.
// Accumulators 8s..8s+7 of output tile m, read as the next layer's B
// operand, are input chunk c = 2m + s with K position (h, q) holding unit
// u(c, h, q) = 16 c + 8 (q / 4) + 4 h + q % 4; the host orders the weights
// the same way, so layers chain without data exchange between lanes, and
// only the inputs and outputs cross between lanes l and l ^ 32.
//
// The first layer takes the 8 features (t, x_1..x_dim, 1, 0..) as bf16
// high parts in half 0 and low parts (feature - high) in half 1, both against
// the same weights, so the inputs keep about 16 significant bits.
#pragma once

#include "sdemc_core.h"

namespace sdemc {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

constexpr int CV_BLOCK_WORDS = 512;  // one operand fragment: 64 lanes x 8 bf16

__host__ __device__ constexpr int cv_blocks(int hidden) { return 8 * hidden - 2; }
__host__ __device__ constexpr int cv_words(int hidden) {
  return hidden > 0 ? CV_BLOCK_WORDS * cv_blocks(hidden) : 0;
}

// Packed weights of the launch, f then g, staged once per block.
extern __shared__ u32x4_t cv_lds[];

__device__ __forceinline__ bf16x8_t cv_frag(int base_blk, int blk, int lane) {
  return __builtin_bit_cast(bf16x8_t, cv_lds[(base_blk + blk) * WAVE + lane]);
}

__device__ __forceinline__ f32x16_t mfma16(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t pack2(float a, float b) {
  const bf16x2_t v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}

// ReLU of accumulators 8s..8s+7 as a bf16 operand: round to bf16 pairs, then
// a packed signed 16-bit max with 0 (a negative bf16 is a negative int16,
// and rounding keeps the sign, so this is relu before rounding, bit for bit).
__device__ __forceinline__ uint32_t relu_pair(uint32_t v) {
  uint32_t r;
  asm("v_pk_max_i16 %0, %1, 0" : "=v"(r) : "v"(v));
  return r;
}

template <int S>
__device__ __forceinline__ bf16x8_t relu_chunk(const f32x16_t& c) {
  u32x4_t u;
#pragma unroll
  for (int i = 0; i < 4; ++i) u[i] = relu_pair(pack2(c[8 * S + 2 * i], c[8 * S + 2 * i + 1]));
  return __builtin_bit_cast(bf16x8_t, u);
}

// First-layer operands of the two path tiles from this lane's 8 features.
__device__ __forceinline__ void cv_inputs(const float (&F)[8], int lane, bf16x8_t& in0,
                                          bf16x8_t& in1) {
  float lo[8];
  u32x4_t hi_w, lo_w;
#pragma unroll
  for (int i = 0; i < 8; ++i) lo[i] = F[i] - (float)(__bf16)F[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    hi_w[i] = pack2(F[2 * i], F[2 * i + 1]);
    lo_w[i] = pack2(lo[2 * i], lo[2 * i + 1]);
  }
  const bool low = lane < 32;
  u32x4_t send, recv;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    send[i] = low ? lo_w[i] : hi_w[i];
    recv[i] = (uint32_t)__shfl_xor((int)send[i], 32);
  }
  // tile 0 (paths 0..31): half 0 high parts, half 1 low parts of path l % 32
  in0 = __builtin_bit_cast(bf16x8_t, low ? hi_w : recv);
  // tile 1 (paths 32..63)
  in1 = __builtin_bit_cast(bf16x8_t, low ? recv : lo_w);
}

// One further hidden layer (blocks blk .. blk + 7) from activations a to z,
// both path tiles.
__device__ __forceinline__ void cv_layer(int blk, int lane, const bf16x8_t (&a0)[4],
                                         const bf16x8_t (&a1)[4], bf16x8_t (&z0)[4],
                                         bf16x8_t (&z1)[4]) {
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    f32x16_t c0 = {}, c1 = {};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const bf16x8_t w = cv_frag(blk, 4 * m + c, lane);
      c0 = mfma16(w, a0[c], c0);
      c1 = mfma16(w, a1[c], c1);
    }
    z0[2 * m] = relu_chunk<0>(c0);
    z0[2 * m + 1] = relu_chunk<1>(c0);
    z1[2 * m] = relu_chunk<0>(c1);
    z1[2 * m + 1] = relu_chunk<1>(c1);
  }
}

// Network with `hidden` hidden layers whose blocks start at block `base`:
// out[q], q < 8, is output unit q for this lane's path.
__device__ __forceinline__ void cv_mlp(int base, int hidden, int lane, const float (&F)[8],
                                       float (&out)[8]) {
  bf16x8_t in0, in1;
  cv_inputs(F, lane, in0, in1);
  bf16x8_t b0[4], b1[4];
  {
    const f32x16_t zero = {};
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const bf16x8_t a = cv_frag(base, m, lane);
      const f32x16_t c0 = mfma16(a, in0, zero);
      const f32x16_t c1 = mfma16(a, in1, zero);
      b0[2 * m] = relu_chunk<0>(c0);
      b0[2 * m + 1] = relu_chunk<1>(c0);
      b1[2 * m] = relu_chunk<0>(c1);
      b1[2 * m + 1] = relu_chunk<1>(c1);
    }
  }
  // further hidden layers in pairs, the activations alternating between b
  // and n so no copy is needed; an odd one out copies back once
  int blk = base + 2;
  int k = 1;
  for (; k + 1 < hidden; k += 2) {
    bf16x8_t n0[4], n1[4];
    cv_layer(blk, lane, b0, b1, n0, n1);
    cv_layer(blk + 8, lane, n0, n1, b0, b1);
    blk += 16;
  }
  if (k < hidden) {
    bf16x8_t n0[4], n1[4];
    cv_layer(blk, lane, b0, b1, n0, n1);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      b0[c] = n0[c];
      b1[c] = n1[c];
    }
    blk += 8;
  }
  f32x16_t o0 = {}, o1 = {};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bf16x8_t a = cv_frag(blk, c, lane);
    o0 = mfma16(a, b0[c], o0);
    o1 = mfma16(a, b1[c], o1);
  }
  // rows 0..3 of a column sit in lane l % 32, rows 4..7 in lane 32 + l % 32
  const bool low = lane < 32;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float x1 = __shfl_xor(o1[q], 32);
    const float x0 = __shfl_xor(o0[q], 32);
    out[q] = low ? o0[q] : x1;
    out[4 + q] = low ? x0 : o1[q];
  }
}

__device__ __forceinline__ void cv_stage(const uint16_t* fw, int fwords, const uint16_t* gw,
                                         int gwords) {
  const u32x4_t* f4 = (const u32x4_t*)fw;
  const u32x4_t* g4 = (const u32x4_t*)gw;
  const int nf = fwords >> 3, ng = gwords >> 3;
  for (int i = threadIdx.x; i < nf + ng; i += blockDim.x)
    cv_lds[i] = i < nf ? f4[i] : g4[i - nf];
  __syncthreads();
}

template <int D, int FAM>
__device__ __forceinline__ void cv_features(float t, const float (&x)[D], float (&F)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) F[i] = 0.0f;
  F[0] = t;
#pragma unroll
  for (int i = 0; i < D; ++i) F[1 + i] = x[i];
  F[1 + D] = 1.0f;
}

// <f, dW> of one step: 'diag' and Heston pair output i with dW_i; 'indep'
// pairs outputs 2i, 2i+1 with (dW_i, dB_i) (varred.py:210-214).
template <int D, int FAM>
__device__ __forceinline__ float cv_dot(const float (&fo)[8], const float (&dW)[D],
                                        const float (&dB)[D]) {
  float s = 0.0f;
  if constexpr (DiagNoise<FAM>::value) {
#pragma unroll
    for (int i = 0; i < D; ++i) s = __fmaf_rn(fo[i], dW[i], s);
  } else {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      s = __fmaf_rn(fo[2 * i], dW[i], s);
      s = __fmaf_rn(fo[2 * i + 1], dB[i], s);
    }
  }
  return s;
}

// Uniform grid without jumps (apply_diffusion_control_variate): steps as in
// path_uniform_rng, t_n = n h.
template <int D, int FAM>
__global__ __launch_bounds__(256) void k_cv_uniform(const KJob job, const uint16_t* fw,
                                                    int32_t f_hidden, float rate,
                                                    int32_t cv_steps,
                                                    double* __restrict__ partials,
                                                    int32_t* __restrict__ cflags,
                                                    float* __restrict__ gamma,
                                                    int64_t nchunks) {
  cv_stage(fw, cv_words(f_hidden), nullptr, 0);
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t chunk = (int64_t)blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x / WAVE);
  if (chunk >= nchunks) return;  // whole wavefront
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  const KModel& f = job.m;
  constexpr int K = NK<D, FAM>::K;
  NormalStream ns;
  ns.init((uint32_t)job.seed, (uint32_t)(job.seed >> 32), job.path_offset + (uint64_t)p);
  float x[D];
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = f.x0[i];
  const float h = job.h, sq = job.sqh;
  const int32_t n_cv = f_hidden > 0 ? cv_steps : 0;
  float cv = 0.0f;
  for (int32_t n = 0; n < job.steps; ++n) {
    float z[K], dW[D], dB[D];
#pragma unroll
    for (int k = 0; k < K; ++k) z[k] = ns.get((uint32_t)n * K + k);
    correlate<D, FAM>(f, z, sq, dW, dB);
    if (n < n_cv) {  // wave-uniform: every lane takes part in the products
      const float t = (float)n * h;
      float F[8], fo[8];
      cv_features<D, FAM>(t, x, F);
      cv_mlp(0, f_hidden, lane, F, fo);
      cv = __fmaf_rn(fast_exp(-rate * t), cv_dot<D, FAM>(fo, dW, dB), cv);
    }
    uniform_step<D, FAM>(job, x, z);
  }
  float v = 0.0f;
  if (valid) {
    v = payoff_raw<D>(job.pay, x) * job.pay.disc + cv;
    if (gamma) gamma[p] = v;
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store(dv, dv * dv, 0, partials + 2 * chunk, cflags + chunk);
}

// Jump-adapted grid (apply_adapted_control_variates): the steps of
// path_adapted, row by row.  Row n of a path is its state at grid time t_n:
// the jump term of the mark J_n applied there uses g(t_n, X_n-); while
// t_n < T the row also takes the step to t_{n+1} with its Brownian term
// f(t_n, X_n) . dW_n and compensator g(t_n, X_n-) (t_{n+1} - t_n).  Rows run
// in lock-step across the wavefront so the products see all 64 lanes;
// finished lanes ride along with their results masked.
template <int D, int FAM>
__global__ __launch_bounds__(256) void k_cv_adapted(const KJob job, const uint16_t* fw,
                                                    int32_t f_hidden, const uint16_t* gw,
                                                    int32_t g_hidden, float rate, float comp,
                                                    int32_t cv_steps,
                                                    double* __restrict__ partials,
                                                    int32_t* __restrict__ cflags,
                                                    float* __restrict__ gamma,
                                                    int64_t nchunks) {
  cv_stage(fw, cv_words(f_hidden), gw, cv_words(g_hidden));
  const int g_base = cv_words(f_hidden) / CV_BLOCK_WORDS;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t chunk = (int64_t)blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x / WAVE);
  if (chunk >= nchunks) return;  // whole wavefront
  const int64_t P = (int64_t)job.n_paths;
  const int64_t p = chunk * WAVE + lane;
  const bool valid = p < P;
  const KModel& f = job.m;
  constexpr int K = NK<D, FAM>::K;
  PathCtx c;
  c.path = job.path_offset + (uint64_t)p;
  c.p = p;
  c.P = P;
  c.k0 = (uint32_t)job.seed;
  c.k1 = (uint32_t)(job.seed >> 32);
  AdaptRngSrc<D, FAM> src;
  src.init(f, c);
  const float T32 = (float)job.T;
  const int32_t max_steps = job.max_steps;
  float h = job.h;
  float t = 0.0f;
  float x[D], xl[D];
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = xl[i] = f.x0[i];
  float tau = valid ? src.next_time() : 0.0f;
  bool active = valid && (t < T32);
  bool row = valid;  // this lane has row `steps` still to account for
  float Jn = 0.0f;   // mark applied at this row's time
  int32_t steps = 0;
  int ovf = 0;
  float bcv = 0.0f, jcv = 0.0f, ccv = 0.0f;
  int32_t iter = 0;
  while (__any(row)) {
    if ((iter++ & (job.clock_win - 1)) == 0) src.refill(f, active, T32);  // see path_adapted
    if (active && steps >= max_steps) {
      ovf = 1;
      active = false;
    }
    const bool use_f = f_hidden > 0 && __any(active && steps < cv_steps);
    const bool use_g = g_hidden > 0 && __any(row && (active || Jn != 0.0f));
    float fo[8], go[8];
    if (use_g) {
      float F[8];
      cv_features<D, FAM>(t, xl, F);
      cv_mlp(g_base, g_hidden, lane, F, go);
    }
    if (use_f) {
      float F[8];
      cv_features<D, FAM>(t, x, F);
      cv_mlp(0, f_hidden, lane, F, fo);
    }
    if (row) {
      const float disc = fast_exp(-rate * t);
      float gs = 0.0f;
      if (use_g) {
#pragma unroll
        for (int i = 0; i < D; ++i) gs += go[i];
        jcv = __fmaf_rn(disc * gs, Jn, jcv);
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
        for (int k = 0; k < K; ++k) zz[k] = src.z((uint32_t)steps, k);
        float dW[D], dB[D], xo[D];
        correlate<D, FAM>(f, zz, sq, dW, dB);
        if (use_f && steps < cv_steps) bcv = __fmaf_rn(disc, cv_dot<D, FAM>(fo, dW, dB), bcv);
        if (use_g) ccv = __fmaf_rn(disc * gs, dt, ccv);
#pragma unroll
        for (int i = 0; i < D; ++i) xo[i] = x[i];
        euler<D, FAM>(f, x, dt, dW, dB);
        {
#pragma clang fp contract(off)
          tn = t + dt;
        }
        t = tn;
        ++steps;
#pragma unroll
        for (int i = 0; i < D; ++i) xl[i] = x[i];
        Jn = 0.0f;
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
          Jn = J;
          src.advance(f);
          tau = src.next_time();
        }
        active = t < T32;
      } else {
        row = false;
      }
    }
  }
  float v = 0.0f;
  if (valid) {
    v = payoff_raw<D>(job.pay, x) * job.pay.disc + bcv + jcv - comp * ccv;
    if (gamma) gamma[p] = v;
  }
  const double dv = valid ? (double)v : 0.0;
  chunk_store(dv, dv * dv, ovf, partials + 2 * chunk, cflags + chunk);
}

// Arguments of one control-variate launch.
struct CvLaunch {
  const uint16_t* fw;
  int32_t f_hidden;
  const uint16_t* gw;
  int32_t g_hidden;
  float rate, comp;
  int32_t cv_steps;
  double* partials;
  int32_t* cflags;
  float* gamma;
  hipStream_t stream;
};

constexpr int CV_THREADS = 256;

template <int D, int FAM>
int launch_cv_uniform_d(const KJob& job, const CvLaunch& a, int64_t nchunks) {
  const int64_t blocks = (nchunks + 3) / 4;
  if (blocks <= 0 || blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  const size_t lds = (size_t)cv_words(a.f_hidden) * 2;
  hipLaunchKernelGGL((k_cv_uniform<D, FAM>), dim3((unsigned)blocks), dim3(CV_THREADS), lds,
                     a.stream, job, a.fw, a.f_hidden, a.rate, a.cv_steps, a.partials, a.cflags,
                     a.gamma, nchunks);
  return hipGetLastError() == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

template <int D, int FAM>
int launch_cv_adapted_d(const KJob& job, const CvLaunch& a, int64_t nchunks) {
  const int64_t blocks = (nchunks + 3) / 4;
  if (blocks <= 0 || blocks > 0x7fffffffLL) return SDEMC_ERR_ARG;
  const size_t lds = (size_t)(cv_words(a.f_hidden) + cv_words(a.g_hidden)) * 2;
  hipLaunchKernelGGL((k_cv_adapted<D, FAM>), dim3((unsigned)blocks), dim3(CV_THREADS), lds,
                     a.stream, job, a.fw, a.f_hidden, a.gw, a.g_hidden, a.rate, a.comp,
                     a.cv_steps, a.partials, a.cflags, a.gamma, nchunks);
  return hipGetLastError() == hipSuccess ? SDEMC_OK : SDEMC_ERR_LAUNCH;
}

int launch_cv_uniform(const sdemc_job& job, const CvLaunch& a);
int launch_cv_adapted(const sdemc_job& job, const CvLaunch& a);

}  // namespace sdemc

// Warm-start network: per-amplitude features, BF16 MFMA forward and backward,
// weight gradients, Adam.  No counterpart in the reference, whose start is
// MBPT2 (setup_t_amplitudes, ccm_iter.f90:2329-2405): the network predicts the
// correction of the converged CCD amplitude over its MBPT2 value,
//   t(ab,ij) = t_MBPT2(ab,ij) + scale * y(features(ab,ij)),
// and the solver then iterates to the FP64 fixed point as before.
//
// Network 16 -> 64 -> 64 -> 64 -> 1, GELU (normal CDF as a polynomial, below).  Parameters are one flat
// FP32 vector (row-major W[k][o]):
//   W1[16][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
// The three hidden layers run on v_mfma_f32_16x16x16bf16 with BF16 weights and
// activations and FP32 accumulation; GELU, biases, the output layer and the
// loss are FP32.  Rounding to BF16 is nearest-even, at fixed points: the inputs,
// the hidden weights, each hidden activation after GELU and each
// pre-activation gradient that feeds a product.
//
// Layout trick: a layer is computed transposed, Z^T = W^T X^T, so that the
// MFMA result of one layer (lane l: sample l%16, features 4(l/16)+i of each
// 16-block) is exactly the B operand of the next; activations never leave
// registers inside the forward and backward passes.
#include "device.h"

#include <algorithm>
#include <cstdint>

namespace ccgpu {

typedef short mbf4 __attribute__((ext_vector_type(4)));
typedef float mf4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 16, kMlpH = 64;
constexpr int kOffW1 = 0, kOffB1 = 1024, kOffW2 = 1088, kOffB2 = 5184, kOffW3 = 5248, kOffB3 = 9344,
              kOffW4 = 9408, kOffB4 = 9472, kMlpParams = kMlpParamCount;
static_assert(kOffB4 + 1 == kMlpParamCount, "parameter layout");
constexpr int kLdW = 68;   // padded LDS row (bf16 elements) of the 64-wide weight copies
constexpr int kLdW1 = 20;  // padded LDS row of W1^T (16 inputs)

// round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
// GELU(z) = z Phi(z), Phi the normal CDF, as an odd polynomial on [-4, 4]:
// Phi(z) = 1/2 + zc P(zc^2), zc = z clamped to [-4, 4], P of degree 7.  Its
// error is at most 5.3e-5 (at the clamp, where 1 - Phi(4) = 3.2e-5), far below
// the BF16 rounding (2^-9 relative) that follows every activation, and it
// takes full-rate FMAs only (no exponential or reciprocal).  Coefficients: a
// least-squares fit on [0, 4] reweighted towards the largest error, rounded to
// FP32 one at a time with the others refitted.
#define PHI_C0 3.988475204e-01f
#define PHI_C1 -6.617541611e-02f
#define PHI_C2 9.664904326e-03f
#define PHI_C3 -1.048214734e-03f
#define PHI_C4 8.066917508e-05f
#define PHI_C5 -4.101028480e-06f
#define PHI_C6 1.217185570e-07f
#define PHI_C7 -1.580922393e-09f
__device__ __forceinline__ float phi_poly(float zc, float s) {
  float p = PHI_C7;
  p = fmaf(p, s, PHI_C6);
  p = fmaf(p, s, PHI_C5);
  p = fmaf(p, s, PHI_C4);
  p = fmaf(p, s, PHI_C3);
  p = fmaf(p, s, PHI_C2);
  p = fmaf(p, s, PHI_C1);
  p = fmaf(p, s, PHI_C0);
  return p;
}
__device__ __forceinline__ float gelu_f(float z) {
  const float zc = fminf(fmaxf(z, -4.f), 4.f);
  return z * fmaf(zc, phi_poly(zc, zc * zc), 0.5f);
}
// d/dz [z Phi(z)] = Phi(z) + z Phi'(z), Phi'(z) = P(s) + 2 s P'(s) inside the
// clamp, 0 outside
__device__ __forceinline__ float gelu_d(float z) {
  const float zc = fminf(fmaxf(z, -4.f), 4.f), s = zc * zc;
  const float p = phi_poly(zc, s);
  float q = 7.f * PHI_C7;
  q = fmaf(q, s, 6.f * PHI_C6);
  q = fmaf(q, s, 5.f * PHI_C5);
  q = fmaf(q, s, 4.f * PHI_C4);
  q = fmaf(q, s, 3.f * PHI_C3);
  q = fmaf(q, s, 2.f * PHI_C2);
  q = fmaf(q, s, PHI_C1);
  const float dphi = fabsf(z) < 4.f ? fmaf(2.f * s, q, p) : 0.f;
  return fmaf(zc, p, 0.5f) + z * dphi;
}
// two BF16 values per 32-bit register: (lo, hi)
__device__ __forceinline__ mbf4 pack4(uint32_t lo, uint32_t hi) {
  mbf4 r;
  r[0] = (short)(lo & 0xFFFF); r[1] = (short)(lo >> 16);
  r[2] = (short)(hi & 0xFFFF); r[3] = (short)(hi >> 16);
  return r;
}
__device__ __forceinline__ mbf4 lds4(const uint16_t *p) {
  uint2 v = *reinterpret_cast<const uint2 *>(p);
  mbf4 r;
  r[0] = (short)(v.x & 0xFFFF); r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xFFFF); r[3] = (short)(v.y >> 16);
  return r;
}
__device__ __forceinline__ mf4 mfma_bf16(mbf4 a, mbf4 b, mf4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Weights into LDS: W^T copies for the forward A operand (rows = outputs),
// W copies for the backward A operand (rows = inputs), all BF16.
struct MlpLds {
  uint16_t wt1[kMlpH * kLdW1];
  uint16_t wt2[kMlpH * kLdW], wt3[kMlpH * kLdW];
  uint16_t w2[kMlpH * kLdW], w3[kMlpH * kLdW];
  float b1[kMlpH], b2[kMlpH], b3[kMlpH], w4[kMlpH];
  float b4;
  float red[4];
};

__device__ void mlp_load_lds(const float *__restrict__ p, MlpLds &L) {
  for (int e = threadIdx.x; e < kMlpIn * kMlpH; e += blockDim.x) {
    int k = e / kMlpH, o = e % kMlpH;
    L.wt1[o * kLdW1 + k] = f2bf(p[kOffW1 + e]);
  }
  for (int e = threadIdx.x; e < kMlpH * kMlpH; e += blockDim.x) {
    int k = e / kMlpH, o = e % kMlpH;
    uint16_t v2 = f2bf(p[kOffW2 + e]), v3 = f2bf(p[kOffW3 + e]);
    L.wt2[o * kLdW + k] = v2; L.w2[k * kLdW + o] = v2;
    L.wt3[o * kLdW + k] = v3; L.w3[k * kLdW + o] = v3;
  }
  for (int e = threadIdx.x; e < kMlpH; e += blockDim.x) {
    L.b1[e] = p[kOffB1 + e]; L.b2[e] = p[kOffB2 + e]; L.b3[e] = p[kOffB3 + e]; L.w4[e] = p[kOffW4 + e];
  }
  if (threadIdx.x == 0) L.b4 = p[kOffB4];
  __syncthreads();
}

// Stores one lane's 16 BF16 values (features 16mt+4q+i of its sample, packed
// in pairs) into a feature-major [64][B] array.
__device__ __forceinline__ void store_t(uint16_t *dst, const uint32_t (&v)[4][2], int q, int64_t B, int64_t pos) {
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t w = v[mt][i >> 1];
      dst[(int64_t)(16 * mt + 4 * q + i) * B + pos] = (uint16_t)((i & 1) ? (w >> 16) : (w & 0xFFFF));
    }
}

// One hidden 64 -> 64 layer, transposed: outputs 16mt+4q+i of this lane's
// sample; bias added; h = bf16(gelu(z)) packed in pairs; with GD also gelu'(z).
template <bool GD>
__device__ __forceinline__ void mlp_layer64(const uint16_t *wt, const float *bias, const uint32_t (&hin)[4][2],
                                            uint32_t (&h)[4][2], float (&gd)[4][4], int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    mf4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = bias[16 * mt + 4 * q + i];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      acc = mfma_bf16(lds4(wt + (16 * mt + r) * kLdW + 16 * kb + 4 * q), pack4(hin[kb][0], hin[kb][1]), acc);
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o[i] = f2bf(gelu_f(acc[i]));
      if (GD) gd[mt][i] = gelu_d(acc[i]);
    }
    h[mt][0] = o[0] | (o[1] << 16);
    h[mt][1] = o[2] | (o[3] << 16);
  }
}

// Backward through one 64 -> 64 layer: dz_prev = bf16((W dz) * gelu'(z_prev)).
__device__ __forceinline__ void mlp_back64(const uint16_t *w, const uint32_t (&dz)[4][2], const float (&gd)[4][4],
                                           uint32_t (&dzprev)[4][2], int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    mf4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      acc = mfma_bf16(lds4(w + (16 * mt + r) * kLdW + 16 * kb + 4 * q), pack4(dz[kb][0], dz[kb][1]), acc);
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = f2bf(acc[i] * gd[mt][i]);
    dzprev[mt][0] = o[0] | (o[1] << 16);
    dzprev[mt][1] = o[2] | (o[3] << 16);
  }
}

// Forward of one wave's 16 samples: lane l holds inputs 4(l/16)..4(l/16)+3 of
// sample l%16; returns the output of the lane's sample (every lane of the
// sample holds it).  Weights and biases in LDS as in MlpLds.
__device__ __forceinline__ float mlp_forward16(const uint16_t *wt1, const float *b1, const uint16_t *wt2,
                                               const float *b2, const uint16_t *wt3, const float *b3, const float *w4,
                                               float b4, uint2 xin, int lane) {
  const int r = lane & 15, q = lane >> 4;
  uint32_t h1[4][2], h2[4][2], h3[4][2];
  float gd[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    mf4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = b1[16 * mt + 4 * q + i];
    acc = mfma_bf16(lds4(wt1 + (16 * mt + r) * kLdW1 + 4 * q), pack4(xin.x, xin.y), acc);
    uint32_t o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = f2bf(gelu_f(acc[i]));
    h1[mt][0] = o[0] | (o[1] << 16);
    h1[mt][1] = o[2] | (o[3] << 16);
  }
  mlp_layer64<false>(wt2, b2, h1, h2, gd, lane);
  mlp_layer64<false>(wt3, b3, h2, h3, gd, lane);
  float part = 0.f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    part += w4[16 * mt + 4 * q + 0] * bf2f((uint16_t)(h3[mt][0] & 0xFFFF));
    part += w4[16 * mt + 4 * q + 1] * bf2f((uint16_t)(h3[mt][0] >> 16));
    part += w4[16 * mt + 4 * q + 2] * bf2f((uint16_t)(h3[mt][1] & 0xFFFF));
    part += w4[16 * mt + 4 * q + 3] * bf2f((uint16_t)(h3[mt][1] >> 16));
  }
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  return part + b4;
}

// Forward (TRAIN: and backward) over samples at positions [0, npad), npad a
// multiple of 64; position p reads sample idx[p] (or p), positions >= n are
// padding (zero input, no loss, no gradient).  Inference writes y[p].
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_pass(const float *__restrict__ params, const uint16_t *__restrict__ feat,
                                                  const float *__restrict__ target, const int64_t *__restrict__ idx,
                                                  int64_t n, int64_t npad, float inv_n, float *__restrict__ yout,
                                                  MlpTrainBufs tb) {
  __shared__ MlpLds L;
  mlp_load_lds(params, L);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int64_t B = npad;
  float loss = 0.f;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < npad; base += (int64_t)gridDim.x * 64) {
    const int64_t pos = base + 16 * wave + r;
    const bool valid = pos < n;
    const int64_t smp = valid ? (idx ? idx[pos] : pos) : 0;
    // inputs: lane holds features 4q..4q+3 of its sample
    uint2 xin = make_uint2(0u, 0u);
    if (valid) xin = *reinterpret_cast<const uint2 *>(feat + smp * kMlpIn + 4 * q);
    if constexpr (!TRAIN) {
      const float y = mlp_forward16(L.wt1, L.b1, L.wt2, L.b2, L.wt3, L.b3, L.w4, L.b4, xin, lane);
      if (q == 0 && valid) yout[pos] = y;
      continue;
    }
    float gd1[4][4], gd2[4][4], gd3[4][4];
    uint32_t h1[4][2], h2[4][2], h3[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      mf4 acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = L.b1[16 * mt + 4 * q + i];
      acc = mfma_bf16(lds4(L.wt1 + (16 * mt + r) * kLdW1 + 4 * q), pack4(xin.x, xin.y), acc);
      uint32_t o[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        o[i] = f2bf(gelu_f(acc[i]));
        if (TRAIN) gd1[mt][i] = gelu_d(acc[i]);
      }
      h1[mt][0] = o[0] | (o[1] << 16);
      h1[mt][1] = o[2] | (o[3] << 16);
    }
    if constexpr (TRAIN) {
      tb.x1t[(int64_t)(4 * q + 0) * B + pos] = (uint16_t)(xin.x & 0xFFFF);
      tb.x1t[(int64_t)(4 * q + 1) * B + pos] = (uint16_t)(xin.x >> 16);
      tb.x1t[(int64_t)(4 * q + 2) * B + pos] = (uint16_t)(xin.y & 0xFFFF);
      tb.x1t[(int64_t)(4 * q + 3) * B + pos] = (uint16_t)(xin.y >> 16);
      store_t(tb.h1t, h1, q, B, pos);
    }
    mlp_layer64<TRAIN>(L.wt2, L.b2, h1, h2, gd2, lane);
    if constexpr (TRAIN) store_t(tb.h2t, h2, q, B, pos);
    mlp_layer64<TRAIN>(L.wt3, L.b3, h2, h3, gd3, lane);
    if constexpr (TRAIN) store_t(tb.h3t, h3, q, B, pos);
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      part += L.w4[16 * mt + 4 * q + 0] * bf2f((uint16_t)(h3[mt][0] & 0xFFFF));
      part += L.w4[16 * mt + 4 * q + 1] * bf2f((uint16_t)(h3[mt][0] >> 16));
      part += L.w4[16 * mt + 4 * q + 2] * bf2f((uint16_t)(h3[mt][1] & 0xFFFF));
      part += L.w4[16 * mt + 4 * q + 3] * bf2f((uint16_t)(h3[mt][1] >> 16));
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    const float y = part + L.b4;
    {
      const float d = valid ? y - target[smp] : 0.f;
      if (q == 0) loss += d * d;
      const float g = 2.f * d * inv_n;
      uint32_t dz3[4][2], dz2[4][2], dz1[4][2];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        uint32_t o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) o[i] = f2bf(L.w4[16 * mt + 4 * q + i] * g * gd3[mt][i]);
        dz3[mt][0] = o[0] | (o[1] << 16);
        dz3[mt][1] = o[2] | (o[3] << 16);
      }
      store_t(tb.dz3t, dz3, q, B, pos);
      mlp_back64(L.w3, dz3, gd2, dz2, lane);
      store_t(tb.dz2t, dz2, q, B, pos);
      mlp_back64(L.w2, dz2, gd1, dz1, lane);
      store_t(tb.dz1t, dz1, q, B, pos);
      if (q == 0) tb.g[pos] = g;
    }
  }
  if constexpr (TRAIN) {
    // fixed-order loss reduction: lanes 0..15 of each wave, then the 4 waves
#pragma unroll
    for (int s = 8; s > 0; s >>= 1) loss += __shfl_xor(loss, s);
    if (lane == 0) L.red[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) tb.loss_part[blockIdx.x] = ((L.red[0] + L.red[1]) + L.red[2]) + L.red[3];
  }
}

// Weight gradients of one sample chunk: grid (nsplit, 4 layers).  Layer 0..2:
// gW[k][o] = sum_s X[k][s] dZ[o][s] on MFMA (contraction over samples) and
// gb[o] = sum_s dZ[o][s]; layer 3: gw4[f] = sum_s h3[f][s] g[s], gb4 = sum_s g[s].
// Partials part[split][param] are summed in fixed order by k_mlp_reduce.
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpTrainBufs tb, int64_t B, int chunk, float *__restrict__ part) {
  const int split = blockIdx.x, layer = blockIdx.y;
  const int64_t s_begin = (int64_t)split * chunk;
  const int64_t s_end = s_begin + chunk < B ? s_begin + chunk : B;
  float *P = part + (int64_t)split * kMlpParams;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  if (layer == 3) {
    const int t = threadIdx.x;
    if (t < kMlpH) {
      float acc = 0.f;
      for (int64_t s = s_begin; s < s_end; ++s) acc += bf2f(tb.h3t[(int64_t)t * B + s]) * tb.g[s];
      P[kOffW4 + t] = acc;
    } else if (t == kMlpH) {
      float acc = 0.f;
      for (int64_t s = s_begin; s < s_end; ++s) acc += tb.g[s];
      P[kOffB4] = acc;
    }
    return;
  }
  const uint16_t *X = layer == 0 ? tb.x1t : layer == 1 ? tb.h1t : tb.h2t;
  const uint16_t *DZ = layer == 0 ? tb.dz1t : layer == 1 ? tb.dz2t : tb.dz3t;
  const int K = layer == 0 ? kMlpIn : kMlpH;
  const int offW = layer == 0 ? kOffW1 : layer == 1 ? kOffW2 : kOffW3;
  const int offB = layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : kOffB3;
  // K=64: wave w owns input rows 16w..16w+15 and all 4 output tiles;
  // K=16: every wave owns the single input tile and output tile w
  const int mt = K == kMlpH ? wave : 0;
  const int nt0 = K == kMlpH ? 0 : wave, nt1 = K == kMlpH ? 4 : wave + 1;
  mf4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = mf4{0.f, 0.f, 0.f, 0.f};
  for (int64_t s0 = s_begin; s0 < s_end; s0 += 16) {
    const mbf4 a = lds4(X + (int64_t)(16 * mt + r) * B + s0 + 4 * q);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (t >= nt0 && t < nt1) acc[t] = mfma_bf16(a, lds4(DZ + (int64_t)(16 * t + r) * B + s0 + 4 * q), acc[t]);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (t >= nt0 && t < nt1)
#pragma unroll
      for (int i = 0; i < 4; ++i) P[offW + (16 * mt + 4 * q + i) * kMlpH + 16 * t + r] = acc[t][i];
  if (threadIdx.x < kMlpH) {
    const int o = threadIdx.x;
    float sb = 0.f;
    for (int64_t s = s_begin; s < s_end; ++s) sb += bf2f(DZ[(int64_t)o * B + s]);
    P[offB + o] = sb;
  }
}

__global__ void k_mlp_reduce(const float *__restrict__ part, int nsplit, float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  float s = 0.f;
  for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * kMlpParams + p];
  grad[p] = s;
}

__global__ void k_mlp_loss_reduce(const float *__restrict__ part, int n, double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int k = 0; k < n; ++k) s += part[k];
  out[0] = s;
}

__global__ void k_mlp_adam(float *__restrict__ th, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.f - b1) * gp;
  const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  th[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// minibatch positions -> training-set samples (with replacement, splitmix64)
__global__ void k_mlp_sample(int64_t *__restrict__ idx, int64_t B, int64_t ntot, uint64_t seed) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < B; p += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(p + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[p] = (int64_t)(z % (uint64_t)ntot);
  }
}

// MBPT2 amplitude of unique element u (the formula of k_mbpt2)
__device__ __forceinline__ double2 mbpt2_of(const SolveDesc &d, int64_t u, int a, int b, int i, int j) {
  const double den = -d.fdiag[a].x - d.fdiag[b].x + d.fdiag[i].x + d.fdiag[j].x;
  if (fabs(den) == 0.0) return make_double2(0.0, 0.0);
  const double2 v = d.R0[u];
  return make_double2(v.x / den, v.y / den);
}

// Scale w = |D|/e_F between the amplitude correction and the network output:
// the network predicts the correction of the numerator, so the amplitude
// correction falls off like the MBPT2 amplitude itself (1/D) at momenta the
// training lattices did not reach; w = 0 (no correction) where D = 0.
__device__ __forceinline__ double corr_weight(double den, float ef) {
  return fabs(den) > 0.0 ? fabs(den) / (double)ef : 0.0;
}

__device__ __forceinline__ void labels_of(const SolveDesc &d, int64_t u, int &a, int &b, int &i, int &j) {
  const int c = d.t_ch[u];
  const int64_t loc = u - d.off_t[c];
  const int nh = d.nhh[c];
  const int rr = (int)(loc / nh), ss = (int)(loc % nh);
  a = d.pp_a[d.off_pp[c] + rr]; b = d.pp_b[d.off_pp[c] + rr];
  i = d.hh_i[d.off_hh[c] + ss]; j = d.hh_j[d.off_hh[c] + ss];
}

// Features of every unique amplitude (a<b, i<j) of a solve, BF16 [nT][16]:
//  0 |p|/kF, p = (k_a - k_b)/2     1 |p'|/kF, p' = (k_i - k_j)/2
//  2 |K|/kF, K = k_a + k_b          3 cos(p, p')        4 |p - p'|/kF
//  5-8 sz_a, sz_b, sz_i, sz_j       9 (tz_a + tz_b)/2   10 kF (fm^-1)
//  11-13 twist angles / pi          14 25 Re t_MBPT2     15 D/e_F, D = f_ii+f_jj-f_aa-f_bb
// With target != null also the regression target (Re t - Re t_MBPT2) w/scale of
// the amplitudes currently in T (w = |D|/e_F).
__device__ __forceinline__ void mlp_features_of(const SolveDesc &d, int64_t u, float kf, float ef, float tx,
                                                float ty, float tz, float (&f)[16], double2 &t2, double &den) {
  int a, b, i, j;
  labels_of(d, u, a, b, i, j);
  const double px = 0.5 * (d.kx[a] - d.kx[b]), py = 0.5 * (d.ky[a] - d.ky[b]), pz = 0.5 * (d.kz[a] - d.kz[b]);
  const double qx = 0.5 * (d.kx[i] - d.kx[j]), qy = 0.5 * (d.ky[i] - d.ky[j]), qz = 0.5 * (d.kz[i] - d.kz[j]);
  const double Kx = d.kx[a] + d.kx[b], Ky = d.ky[a] + d.ky[b], Kz = d.kz[a] + d.kz[b];
  const double np_ = sqrt(px * px + py * py + pz * pz), nq = sqrt(qx * qx + qy * qy + qz * qz);
  const double dx = px - qx, dy = py - qy, dz = pz - qz;
  t2 = mbpt2_of(d, u, a, b, i, j);
  den = -d.fdiag[a].x - d.fdiag[b].x + d.fdiag[i].x + d.fdiag[j].x;
  f[0] = (float)(np_ / kf);
  f[1] = (float)(nq / kf);
  f[2] = (float)(sqrt(Kx * Kx + Ky * Ky + Kz * Kz) / kf);
  f[3] = (np_ > 0.0 && nq > 0.0) ? (float)((px * qx + py * qy + pz * qz) / (np_ * nq)) : 0.f;
  f[4] = (float)(sqrt(dx * dx + dy * dy + dz * dz) / kf);
  f[5] = (float)d.sz[a]; f[6] = (float)d.sz[b]; f[7] = (float)d.sz[i]; f[8] = (float)d.sz[j];
  f[9] = 0.5f * (float)(d.tz[a] + d.tz[b]);
  f[10] = kf;
  f[11] = tx; f[12] = ty; f[13] = tz;
  f[14] = (float)(25.0 * t2.x);
  f[15] = (float)(den / ef);
}

__global__ void k_mlp_features(SolveDesc d, float kf, float ef, float tx, float ty, float tz, float scale,
                               uint16_t *__restrict__ feat, float *__restrict__ target) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    float f[16];
    double2 t2;
    double den;
    mlp_features_of(d, u, kf, ef, tx, ty, tz, f, t2, den);
#pragma unroll
    for (int k = 0; k < 16; ++k) feat[u * 16 + k] = f2bf(f[k]);
    if (target) target[u] = (float)((d.T[u].x - t2.x) * corr_weight(den, ef) / scale);
  }
}

// The learned start in one pass: features, the network, and
// t = t_MBPT2 + scale * y / w written to T (the bits of k_mlp_features,
// k_mlp_pass<false> and k_mlp_apply in a row).  A wave takes 64 samples: lane l
// computes the features of sample l once and leaves them in LDS; the network
// then runs over the 64 in four groups of 16, and lane l = 16 q + r keeps the
// output of group q (its own sample) and writes its amplitude.
struct MlpLdsF {
  uint16_t wt1[kMlpH * kLdW1];
  uint16_t wt2[kMlpH * kLdW], wt3[kMlpH * kLdW];
  float b1[kMlpH], b2[kMlpH], b3[kMlpH], w4[kMlpH];
  float b4;
  uint2 x[4][64][4];   // per wave: 64 samples x 16 BF16 inputs
};
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ p, SolveDesc d, float kf, float ef,
                                                   float tx, float ty, float tz, float scale) {
  __shared__ MlpLdsF L;
  for (int e = threadIdx.x; e < kMlpIn * kMlpH; e += blockDim.x) {
    const int k = e / kMlpH, o = e % kMlpH;
    L.wt1[o * kLdW1 + k] = f2bf(p[kOffW1 + e]);
  }
  for (int e = threadIdx.x; e < kMlpH * kMlpH; e += blockDim.x) {
    const int k = e / kMlpH, o = e % kMlpH;
    L.wt2[o * kLdW + k] = f2bf(p[kOffW2 + e]);
    L.wt3[o * kLdW + k] = f2bf(p[kOffW3 + e]);
  }
  for (int e = threadIdx.x; e < kMlpH; e += blockDim.x) {
    L.b1[e] = p[kOffB1 + e]; L.b2[e] = p[kOffB2 + e]; L.b3[e] = p[kOffB3 + e]; L.w4[e] = p[kOffW4 + e];
  }
  if (threadIdx.x == 0) L.b4 = p[kOffB4];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  // every workgroup runs the same number of rounds (the barriers below)
  for (int64_t base = (int64_t)blockIdx.x * 256; base < d.nT; base += (int64_t)gridDim.x * 256) {
    const int64_t u = base + 64 * wave + lane;
    const bool valid = u < d.nT;
    double2 t2 = make_double2(0.0, 0.0);
    double den = 0.0;
    uint32_t fb[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) fb[k] = 0u;
    if (valid) {
      float f[16];
      mlp_features_of(d, u, kf, ef, tx, ty, tz, f, t2, den);
#pragma unroll
      for (int k = 0; k < 16; ++k) fb[k] = f2bf(f[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      L.x[wave][lane][k] = make_uint2(fb[4 * k] | (fb[4 * k + 1] << 16), fb[4 * k + 2] | (fb[4 * k + 3] << 16));
    __syncthreads();
    float y = 0.f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const float yg = mlp_forward16(L.wt1, L.b1, L.wt2, L.b2, L.wt3, L.b3, L.w4, L.b4, L.x[wave][16 * g + r][q], lane);
      if (g == q) y = yg;
    }
    __syncthreads();
    if (valid) {
      const double w = corr_weight(den, ef);
      double2 t = t2;
      if (w != 0.0) t.x += (double)scale * (double)y / w;
      d.T[u] = t;
    }
  }
}

// t = t_MBPT2 + scale * y / w on the real part (w = |D|/e_F)
__global__ void k_mlp_apply(SolveDesc d, const float *__restrict__ y, float scale, float ef) {
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    int a, b, i, j;
    labels_of(d, u, a, b, i, j);
    double2 t = mbpt2_of(d, u, a, b, i, j);
    const double den = -d.fdiag[a].x - d.fdiag[b].x + d.fdiag[i].x + d.fdiag[j].x;
    const double w = corr_weight(den, ef);
    if (w != 0.0) t.x += (double)scale * (double)y[u] / w;
    d.T[u] = t;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned mlp_grid(int64_t npad) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(npad / 64, 2048));
}

void launch_mlp_forward(const float *params, const uint16_t *feat, int64_t n, float *y, hipStream_t st) {
  const int64_t npad = (n + 63) / 64 * 64;
  MlpTrainBufs none{};
  hipLaunchKernelGGL(k_mlp_pass<false>, dim3(mlp_grid(npad)), dim3(256), 0, st, params, feat, nullptr, nullptr, n,
                     npad, 0.f, y, none);
}

// One loss+gradient evaluation over positions [0, npad) (npad multiple of 64):
// grad (kMlpParams floats) and the summed squared error in *loss_sum (device).
void launch_mlp_loss_grad(const float *params, const uint16_t *feat, const float *target, const int64_t *idx,
                          int64_t n, int64_t npad, const MlpTrainBufs &tb, float *part, int max_split, float *grad,
                          double *loss_sum, hipStream_t st) {
  const unsigned grid = mlp_grid(npad);
  hipLaunchKernelGGL(k_mlp_pass<true>, dim3(grid), dim3(256), 0, st, params, feat, target, idx, n, npad,
                     1.0f / (float)n, nullptr, tb);
  int chunk = 512;
  int nsplit = (int)((npad + chunk - 1) / chunk);
  if (nsplit > max_split) {
    chunk = (int)(((npad + max_split - 1) / max_split + 15) / 16 * 16);
    nsplit = (int)((npad + chunk - 1) / chunk);
  }
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(nsplit, 4), dim3(256), 0, st, tb, npad, chunk, part);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, part, nsplit, grad);
  hipLaunchKernelGGL(k_mlp_loss_reduce, dim3(1), dim3(64), 0, st, tb.loss_part, (int)grid, loss_sum);
}

void launch_mlp_adam(float *th, const float *g, float *m, float *v, float lr, int step, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, th, g, m, v, lr, b1, b2, 1e-8f,
                     c1, c2);
}

void launch_mlp_sample(int64_t *idx, int64_t B, int64_t ntot, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_sample, dim3((unsigned)std::min<int64_t>((B + 255) / 256, 4096)), dim3(256), 0, st, idx, B,
                     ntot, seed);
}

void launch_mlp_features(const SolveDesc &d, float kf, float ef, const float tw[3], float scale, uint16_t *feat,
                         float *target, hipStream_t st) {
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((d.nT + 255) / 256, 8192));
  hipLaunchKernelGGL(k_mlp_features, dim3(g), dim3(256), 0, st, d, kf, ef, tw[0], tw[1], tw[2], scale, feat, target);
}

void launch_mlp_infer(const float *params, const SolveDesc &d, float kf, float ef, const float tw[3], float scale,
                      hipStream_t st) {
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((d.nT + 255) / 256, 2048));
  hipLaunchKernelGGL(k_mlp_infer, dim3(g), dim3(256), 0, st, params, d, kf, ef, tw[0], tw[1], tw[2], scale);
}

void launch_mlp_apply(const SolveDesc &d, const float *y, float scale, float ef, hipStream_t st) {
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((d.nT + 255) / 256, 8192));
  hipLaunchKernelGGL(k_mlp_apply, dim3(g), dim3(256), 0, st, d, y, scale, ef);
}

}  // namespace ccgpu

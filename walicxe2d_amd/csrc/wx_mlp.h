// wx_mlp.h -- device code of the cooling surrogate shared by the training
// kernels (wx_surrogate.hip) and the fused cooling pass (wx_kernels.hip).
//
// Matrix-core layout (v_mfma_f32_16x16x16_bf16, one wave64):
//   A (16x16, M x K): lane l holds A[l % 16][4*(l/16) + e], e = 0..3
//   B (16x16, K x N): lane l holds B[4*(l/16) + e][l % 16]
//   D (16x16, M x N): lane l holds D[4*(l/16) + e][l % 16]
// The network is evaluated transposed, Z = W * A_prev with samples as the
// N dimension: a layer's D tile t (output features 16t .. 16t+15) is then,
// lane for lane, the B operand of K-step t of the next layer, so the
// activations never leave registers.
#pragma once
#include <hip/hip_runtime.h>

#include "wx_physics.h"
#include "wx_surrogate.h"

namespace wx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) {   // 4 consecutive bf16 (8-byte aligned)
  const uint2 v = *reinterpret_cast<const uint2 *>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xFFFF);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xFFFF);
  r[3] = (short)(v.y >> 16);
  return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}


// SiLU and its derivative with the hardware exp2 and reciprocal (1 ulp each).
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.f + z * (1.f - s));
}

// Forward pass of 16 samples by one wave.  x: B operand of layer 1 (lanes
// 0-15 carry their sample's 4 standardised features, lanes 16-63 zeros).
// w*: bf16 weights as in SurDev; b1..b4: the FP32 biases (global memory or
// LDS copies).  Returns the D tile of layer 4 (rows 0, 1 = standardised
// outputs of sample l in lanes l < 16).  If zs / as are given, the
// pre-activations (D layout) and bf16 activations are kept for the backward
// pass.
__device__ __forceinline__ f32x4 mlp_fwd16(const uint16_t *__restrict__ w1, const uint16_t *__restrict__ w2,
                                           const uint16_t *__restrict__ w3, const uint16_t *__restrict__ w4,
                                           const float *__restrict__ b1, const float *__restrict__ b2,
                                           const float *__restrict__ b3, const float *__restrict__ b4, bf16x4 x,
                                           f32x4 (*zs)[4] = nullptr, bf16x4 (*as)[4] = nullptr) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, q = l >> 4;
  bf16x4 a[4];
  // layer 1: K = 4 features padded to 16
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(ld4(w1 + (16 * mt + r) * 16 + 4 * q), x, acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] += b1[16 * mt + 4 * q + e];
    if (zs) zs[0][mt] = acc;
    a[mt] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
    if (as) as[0][mt] = a[mt];
  }
  // layers 2 and 3: 64 -> 64
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const uint16_t *__restrict__ w = layer == 0 ? w2 : w3;
    const float *__restrict__ b = layer == 0 ? b2 : b3;
    bf16x4 an[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma16(ld4(w + (16 * mt + r) * 64 + 16 * kt + 4 * q), a[kt], acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] += b[16 * mt + 4 * q + e];
      if (zs) zs[layer + 1][mt] = acc;
      an[mt] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
      if (as) as[layer + 1][mt] = an[mt];
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) a[mt] = an[mt];
  }
  // layer 4: 64 -> 2 (rows padded to 16)
  f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) y = mfma16(ld4(w4 + r * 64 + 16 * kt + 4 * q), a[kt], y);
  if (q == 0) {
    y[0] += b4[0];
    y[1] += b4[1];
  }
  return y;
}

// The forward pass with the weights and biases of a model in global memory.
__device__ __forceinline__ f32x4 mlp_fwd16(const SurDev *__restrict__ M, bf16x4 x, f32x4 (*zs)[4] = nullptr,
                                           bf16x4 (*as)[4] = nullptr) {
  return mlp_fwd16(M->w1, M->w2, M->w3, M->w4, M->theta + SUR_B1, M->theta + SUR_B2, M->theta + SUR_B3,
                   M->theta + SUR_B4, x, zs, as);
}

// al = ALOSS / n^2 at (T, x_HII) from the table (bilinear in log10 T and
// u = log10((x + 1e-4) / (1 - x + 1e-4)) on log10 al).
__device__ __forceinline__ float sur_al(const SurDev *__restrict__ M, float lt, float x) {
  float ft = (lt - SUR_TLO) / SUR_DT;
  ft = fminf(fmaxf(ft, 0.f), (float)(SUR_NT - 1) - 1e-3f);
  const float xc = fminf(fmaxf(x, 0.f), 1.f);
  float fx = (log10f((xc + 1e-4f) / (1.f - xc + 1e-4f)) + 4.f) * 0.125f * (float)(SUR_NX - 1);
  fx = fmaxf(fx, 0.f);
  fx = fminf(fx, (float)(SUR_NX - 1) - 1e-3f);
  const int it = (int)ft, ix = (int)fx;
  const float wt = ft - (float)it, wx = fx - (float)ix;
  const float *t = M->altab + it * SUR_NX + ix;
  return exp10f((1.f - wt) * ((1.f - wx) * t[0] + wx * t[1]) + wt * ((1.f - wx) * t[SUR_NX] + wx * t[SUR_NX + 1]));
}

// Raw features of one cell (n, T, y0 = n_HI/n, dt in seconds); returns
// ce*dt of the tabulated relaxation and whether the cell is in the hull.
__device__ __forceinline__ bool sur_features(const SurDev *__restrict__ M, double n, double T, double y0, double dts,
                                             float f[4], float &cedt) {
  const float lt = log10f((float)T);
  const float x = (float)(1.0 - y0);
  const float al = sur_al(M, lt, x);
  const double boltzm = 1.3807e-16;
  cedt = (float)((2. * n * (double)al) / (3. * boltzm * T) * dts);
  f[0] = log10f((float)n);
  f[1] = lt;
  f[2] = x;
  f[3] = log10f(fmaxf(cedt, 1e-30f));
  bool in = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) in = in && f[k] >= M->lo[k] && f[k] <= M->hi[k];
  return in;
}

}  // namespace wx

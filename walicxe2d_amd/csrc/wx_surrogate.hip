// wx_surrogate.hip -- training and evaluation kernels of the cooling
// surrogate (the network of wx_mlp.h), and the exact coolingh update on a
// list of states that produces its training pairs.
//
//   k_sur_pack     FP32 parameters -> bf16 operand copies
//   k_sur_altab    al(T, x_HII) = ALOSS/n^2 table from the exact loss function
//   k_sur_feat     raw surrogate features of a list of states
//   k_sur_infer    forward pass of a batch (standardised in and out)
//   k_sur_grad     forward + backward of a batch, mean squared error; every
//                  workgroup accumulates its gradient in LDS and adds it to
//                  the global FP32 gradient
//   k_sur_adam     one bias-corrected Adam step of the FP32 parameters
//   k_cool_pairs   src/coolingh.f90:171-259 on (n, T, y0, dt) states
#include <hip/hip_runtime.h>

#include "wx_mlp.h"

namespace wx {

__global__ __launch_bounds__(256) void k_sur_pack(SurDev *__restrict__ M) {
  const float *th = M->theta;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 64 * 64; t += gridDim.x * blockDim.x) {
    M->w2[t] = f2bf(th[SUR_W2 + t]);
    M->w3[t] = f2bf(th[SUR_W3 + t]);
    if (t < 64 * 16) {
      const int o = t / 16, i = t % 16;
      M->w1[t] = i < 4 ? f2bf(th[SUR_W1 + o * 4 + i]) : (uint16_t)0;
      const int o4 = t / 64, i4 = t % 64;
      M->w4[t] = o4 < 2 ? f2bf(th[SUR_W4 + o4 * 64 + i4]) : (uint16_t)0;
    }
  }
}

__global__ __launch_bounds__(256) void k_sur_altab(SurDev *__restrict__ M) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < SUR_NT * SUR_NX; t += gridDim.x * blockDim.x) {
    const int it = t / SUR_NX, ix = t % SUR_NX;
    const double T = pow(10.0, (double)SUR_TLO + (double)SUR_DT * it);
    const double rr = pow(10.0, -4.0 + 8.0 * (double)ix / (double)(SUR_NX - 1));
    const double x = dmin(dmax((rr * (1.0 + 1e-4) - 1e-4) / (1.0 + rr), 0.0), 1.0);
    const double y = 1.0 - x;
    // aloss(x1 = old neutral fraction, den, dh0 = new neutral density, T) at
    // n = 1 and an unchanged fraction
    M->altab[t] = (float)log10(dmax(aloss(y, 1.0, y, T), 1e-300));
  }
}

__global__ __launch_bounds__(256) void k_sur_feat(const SurDev *__restrict__ M, const double *__restrict__ in,
                                                  int64_t n, float *__restrict__ f, float *__restrict__ cedt) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
    float ff[4], ce;
    const bool inh = sur_features(M, in[4 * s], in[4 * s + 1], in[4 * s + 2], in[4 * s + 3], ff, ce);
    (void)inh;
#pragma unroll
    for (int k = 0; k < 4; ++k) f[4 * s + k] = ff[k];
    cedt[s] = ce;
  }
}

// Standardised inputs x[n][4] -> standardised outputs y[n][2].
__global__ __launch_bounds__(256) void k_sur_infer(const SurDev *__restrict__ M, const float *__restrict__ x,
                                                   int64_t n, float *__restrict__ y) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  const int64_t ngroups = (n + 15) / 16;
  for (int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); g < ngroups; g += (int64_t)gridDim.x * 4) {
    const int64_t s = 16 * g + r;
    bf16x4 xb = {0, 0, 0, 0};
    if (q == 0 && s < n) xb = pack4(x[4 * s], x[4 * s + 1], x[4 * s + 2], x[4 * s + 3]);
    const f32x4 yt = mlp_fwd16(M, xb);
    if (q == 0 && s < n) {
      y[2 * s] = yt[0];
      y[2 * s + 1] = yt[1];
    }
  }
}

// Backward of one hidden layer for one wave's 16 samples (see k_sur_grad).
// d: this layer's output deltas (D layout, 4 tiles of 16 features);
// dl: LDS [out][16 samples] bf16 scratch of this wave; act_in: the layer's
// bf16 inputs [in][16]; w: bf16 weights [out][in]; writes the input-side
// deltas (before the SiLU derivative) to da when da != nullptr.
template <int IN_TILES>
__device__ __forceinline__ void layer_backward(float *g, int woff, int boff, int in_dim, const f32x4 d[4],
                                               uint16_t *dl, const uint16_t *act_in, const uint16_t *w,
                                               f32x4 *da) {
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      dl[(16 * mt + 4 * q + e) * 16 + r] = f2bf(d[mt][e]);
      atomicAdd(&g[boff + 16 * mt + 4 * q + e], d[mt][e]);
    }
  __syncthreads();
  // dW[o][k] = sum over samples of d[o][n] * a_in[k][n]
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < IN_TILES; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma16(ld4(dl + (16 * mt + r) * 16 + 4 * q), ld4(act_in + (16 * nt + r) * 16 + 4 * q), acc);
      const int k = 16 * nt + r;
      if (k < in_dim)
#pragma unroll
        for (int e = 0; e < 4; ++e) atomicAdd(&g[woff + (16 * mt + 4 * q + e) * in_dim + k], acc[e]);
    }
  if (da) {
    // da[k][n] = sum over o of W[o][k] * d[o][n]
    bf16x4 bd[4];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) bd[kt] = pack4(d[kt][0], d[kt][1], d[kt][2], d[kt][3]);
#pragma unroll
    for (int mt = 0; mt < IN_TILES; ++mt) {
      bf16x4 wa[4];
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int e = 0; e < 4; ++e) wa[kt][e] = (short)w[(16 * kt + 4 * q + e) * in_dim + 16 * mt + r];
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma16(wa[kt], bd[kt], acc);
      da[mt] = acc;
    }
  }
  __syncthreads();
}

// Gradient of the mean over a batch of sum_j w_j (y_j - t_j)^2.  Sample k
// of the batch (0 <= k < nb) is data row (off + k) mod nd.  Each wave
// handles 16 samples per pass; all waves of a workgroup make the same
// number of passes.
__global__ __launch_bounds__(256) void k_sur_grad(const SurDev *__restrict__ M, const float *__restrict__ x,
                                                  const float *__restrict__ t, int64_t off, int64_t nb, int64_t nd,
                                                  float w0, float w1, float *__restrict__ grad,
                                                  float *__restrict__ loss) {
  __shared__ float g[SUR_NPARAM];
  __shared__ __align__(16) uint16_t act[4][4][64 * 16];
  __shared__ __align__(16) uint16_t dls[4][64 * 16];
  __shared__ float lsum;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, r = l & 15, q = l >> 4;
  for (int i = tid; i < SUR_NPARAM; i += blockDim.x) g[i] = 0.f;
  if (tid == 0) lsum = 0.f;
  __syncthreads();
  const float inv = 1.f / (float)nb;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < nb; base += (int64_t)gridDim.x * 64) {
    const int64_t k = base + 16 * w + r;
    const bool valid = k < nb;
    const int64_t s = valid ? (off + k) % nd : 0;
    bf16x4 xb = {0, 0, 0, 0};
    if (q == 0 && valid) xb = pack4(x[4 * s], x[4 * s + 1], x[4 * s + 2], x[4 * s + 3]);
    f32x4 zs[3][4];
    bf16x4 as[3][4];
    const f32x4 y = mlp_fwd16(M, xb, zs, as);
#pragma unroll
    for (int e = 0; e < 4; ++e) act[w][0][(4 * q + e) * 16 + r] = (uint16_t)xb[e];
#pragma unroll
    for (int L = 0; L < 3; ++L)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int e = 0; e < 4; ++e) act[w][L + 1][(16 * mt + 4 * q + e) * 16 + r] = (uint16_t)as[L][mt][e];
    // output deltas: rows 0, 1 of the layer-4 tile (lanes q == 0)
    f32x4 d4[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) d4[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (q == 0 && valid) {
      const float e0 = y[0] - t[2 * s], e1 = y[1] - t[2 * s + 1];
      d4[0][0] = 2.f * w0 * e0 * inv;
      d4[0][1] = 2.f * w1 * e1 * inv;
      atomicAdd(&lsum, (w0 * e0 * e0 + w1 * e1 * e1) * inv);
    }
    __syncthreads();
    // layer 4 (out 2 padded to 16 = one tile; the other tiles are zero)
    f32x4 da[4];
    {
#pragma unroll
      for (int e = 0; e < 4; ++e) dls[w][(4 * q + e) * 16 + r] = f2bf(d4[0][e]);
      if (q == 0) {
        atomicAdd(&g[SUR_B4], d4[0][0]);
        atomicAdd(&g[SUR_B4 + 1], d4[0][1]);
      }
      __syncthreads();
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(ld4(&dls[w][r * 16 + 4 * q]), ld4(&act[w][3][(16 * nt + r) * 16 + 4 * q]), acc);
        if (q == 0) {
          atomicAdd(&g[SUR_W4 + 0 * 64 + 16 * nt + r], acc[0]);
          atomicAdd(&g[SUR_W4 + 1 * 64 + 16 * nt + r], acc[1]);
        }
      }
      const bf16x4 bd = pack4(d4[0][0], d4[0][1], d4[0][2], d4[0][3]);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        bf16x4 a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = (short)M->w4[(4 * q + e) * 64 + 16 * mt + r];
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        da[mt] = mfma16(a, bd, acc);
      }
      __syncthreads();
    }
    f32x4 d[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) d[mt][e] = da[mt][e] * dsilu(zs[2][mt][e]);
    layer_backward<4>(g, SUR_W3, SUR_B3, 64, d, dls[w], act[w][2], M->w3, da);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) d[mt][e] = da[mt][e] * dsilu(zs[1][mt][e]);
    layer_backward<4>(g, SUR_W2, SUR_B2, 64, d, dls[w], act[w][1], M->w2, da);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int e = 0; e < 4; ++e) d[mt][e] = da[mt][e] * dsilu(zs[0][mt][e]);
    layer_backward<1>(g, SUR_W1, SUR_B1, 4, d, dls[w], act[w][0], M->w1, nullptr);
  }
  __syncthreads();
  for (int i = tid; i < SUR_NPARAM; i += blockDim.x)
    if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
  if (tid == 0) atomicAdd(loss, lsum);
}

__global__ __launch_bounds__(256) void k_sur_adam(SurDev *__restrict__ M, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, int step, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < SUR_NPARAM; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    M->theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Exact coolingh update of (n, T, y0, dt[s]) states at rest: the conserved
// state is built so that uprim recovers T (src/uprim.f90:37-41), atomic runs,
// and (T1, y1) are read back the same way.
__global__ __launch_bounds__(256) void k_cool_pairs(const wx_params *__restrict__ p, const double *__restrict__ in,
                                                    int64_t n, double *__restrict__ out) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
    const double nn = in[4 * s], T = in[4 * s + 1], y0 = in[4 * s + 2], dts = in[4 * s + 3];
    double uu[5] = {nn, 0., 0., p->cv * (2. * nn - y0 * nn) * T / p->Tempsc, y0 * nn};
    atomic_h<5>(p, dts, uu);
    out[2 * s] = uu[3] / p->cv * p->Tempsc / (2. * uu[0] - uu[4]);
    out[2 * s + 1] = uu[4] / uu[0];
  }
}

static inline int hip_ok() { return hipGetLastError() == hipSuccess ? WX_OK : WX_EHIP; }

static inline int grid_for(int64_t n, int per) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

int launch_sur_pack(SurDev *M, hipStream_t stream) {
  hipLaunchKernelGGL(k_sur_pack, dim3(16), dim3(256), 0, stream, M);
  return hip_ok();
}

int launch_sur_altab(const wx_params *P, SurDev *M, hipStream_t stream) {
  (void)P;
  hipLaunchKernelGGL(k_sur_altab, dim3(grid_for(SUR_NT * SUR_NX, 256)), dim3(256), 0, stream, M);
  return hip_ok();
}

int launch_sur_feat(const SurDev *M, const double *in, int64_t n, float *f, float *cedt, hipStream_t stream) {
  if (n <= 0) return WX_OK;
  hipLaunchKernelGGL(k_sur_feat, dim3(grid_for(n, 256)), dim3(256), 0, stream, M, in, n, f, cedt);
  return hip_ok();
}

int launch_sur_infer(const SurDev *M, const float *x, int64_t n, float *y, hipStream_t stream) {
  if (n <= 0) return WX_OK;
  hipLaunchKernelGGL(k_sur_infer, dim3(grid_for(n, 64)), dim3(256), 0, stream, M, x, n, y);
  return hip_ok();
}

int launch_sur_grad(const SurDev *M, const float *x, const float *t, int64_t off, int64_t nb, int64_t nd, float w0,
                    float w1, float *grad, float *loss, hipStream_t stream) {
  if (nb <= 0) return WX_OK;
  hipLaunchKernelGGL(k_sur_grad, dim3(grid_for(nb, 64)), dim3(256), 0, stream, M, x, t, off, nb, nd, w0, w1, grad,
                     loss);
  return hip_ok();
}

int launch_sur_adam(SurDev *M, const float *grad, float *m, float *v, int step, float lr, hipStream_t stream) {
  hipLaunchKernelGGL(k_sur_adam, dim3(grid_for(SUR_NPARAM, 256)), dim3(256), 0, stream, M, grad, m, v, step, lr);
  return hip_ok();
}

int launch_cool_pairs(const wx_params *P, const double *in, int64_t n, double *out, hipStream_t stream) {
  if (n <= 0) return WX_OK;
  hipLaunchKernelGGL(k_cool_pairs, dim3(grid_for(n, 256)), dim3(256), 0, stream, P, in, n, out);
  return hip_ok();
}

}  // namespace wx

// wx_surrogate.h -- the cooling surrogate's device-side model and kernels.
//
// A 4 -> 64 -> 64 -> 64 -> 2 SiLU network replaces coolingh's per-cell
// update (src/coolingh.f90:171-259, `atomic`) for states inside its training
// hull.  Inputs (raw features, standardised with in_mu / in_sig):
//   log10 n, log10 T, x_HII = 1 - n_HI/n, log10(dt / t_cool)
// where t_cool = 3 k T / (2 n al) is the cooling time of the routine's own
// relaxation law (src/coolingh.f90:249-250) with al = ALOSS/n^2 read from a
// table over (log10 T, x_HII) that the device fills from the exact loss
// function (al does not depend on n: every term of ALOSS scales as n^2).
// Outputs (de-standardised with out_mu / out_sig):
//   c = log10(T1 / T1_table), where T1_table is the relaxation law with the
//   tabulated loss, so that log10(T1 / T) = log10(T1_table / T) + c;
//   dx = x_HII(new) - x_HII(old).
// Weights and hidden activations are bfloat16 on the matrix cores
// (v_mfma_f32_16x16x16_bf16), accumulation and SiLU are FP32.
#pragma once
#include <stdint.h>

#include "wx_device.h"

namespace wx {

constexpr int SUR_NPARAM = 8770;
// offsets into the FP32 parameter vector (row-major [out][in] weights)
constexpr int SUR_W1 = 0, SUR_B1 = 256, SUR_W2 = 320, SUR_B2 = 4416, SUR_W3 = 4480, SUR_B3 = 8576,
              SUR_W4 = 8640, SUR_B4 = 8768;
// al table: log10(al) at log10 T from SUR_TLO in SUR_NT steps of SUR_DT
// dex and at u = log10((x_HII + 1e-4) / (1 - x_HII + 1e-4)) in SUR_NX steps
// over [-4, 4], interpolated bilinearly: the loss varies over many decades
// with T, in proportion to the electron density (x_HII + 1e-4) n near
// neutral gas and to the neutral density near full ionisation
// (src/coolingh.f90:101-106)
constexpr int SUR_NT = 321, SUR_NX = 81;
constexpr float SUR_TLO = 1.0f, SUR_DT = 0.025f;

struct SurDev {
  float theta[SUR_NPARAM];
  // bf16 copies of the weights in the matrix-core operand order
  alignas(16) uint16_t w1[64 * 16];   // [out][in padded to 16]
  alignas(16) uint16_t w2[64 * 64];
  alignas(16) uint16_t w3[64 * 64];
  alignas(16) uint16_t w4[16 * 64];   // [out padded to 16][in]
  float in_mu[4], in_sig[4], out_mu[2], out_sig[2];
  float lo[4], hi[4];     // training hull of the raw features
  float tol;              // energy-residual tolerance
  int32_t pad[3];
  float altab[SUR_NT * SUR_NX];
  // counters: cells handled, cells sent to the exact routine
  unsigned long long cells, fallbacks;
};

int launch_sur_pack(SurDev *M, hipStream_t stream);
int launch_sur_altab(const wx_params *P, SurDev *M, hipStream_t stream);
int launch_sur_feat(const SurDev *M, const double *in, int64_t n, float *f, float *cedt, hipStream_t stream);
int launch_sur_infer(const SurDev *M, const float *x, int64_t n, float *y, hipStream_t stream);
int launch_sur_grad(const SurDev *M, const float *x, const float *t, int64_t off, int64_t nbatch, int64_t ndata,
                    float w0, float w1, float *grad, float *loss, hipStream_t stream);
int launch_sur_adam(SurDev *M, const float *grad, float *m, float *v, int step, float lr, hipStream_t stream);
int launch_cool_pairs(const wx_params *P, const double *in, int64_t n, double *out, hipStream_t stream);
// host side (wx_surrogate.cpp)
int sur_upload(SurDev *M, const wx_sur_model *m, const wx_params *P, hipStream_t stream);
int sur_counters(SurDev *M, hipStream_t stream, int reset, int64_t *cells, int64_t *fallbacks);
int launch_fillI_cool_sur(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                          const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                          DevScalars *S, SurDev *M, FlagOut fo, hipStream_t stream);

}  // namespace wx

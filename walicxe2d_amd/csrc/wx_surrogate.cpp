// wx_surrogate.cpp -- host side of the cooling surrogate: the wx_sur_*
// entry points (training pairs, features, inference, gradient, training)
// and the upload of a trained model into a context's or ensemble's device
// model.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "wx_runtime.h"
#include "wx_surrogate.h"

using namespace wx;

static_assert(sizeof(float) * SUR_NPARAM == sizeof(((wx_sur_model *)nullptr)->theta), "theta size");

namespace wx {

// Copies a trained model into the device model M (on `stream`), refreshes
// the bf16 operand copies, fills the loss table with the parameters P
// (device pointer) and clears the counters.
int sur_upload(SurDev *M, const wx_sur_model *m, const wx_params *P, hipStream_t stream) {
  HIPCHK(hipMemcpyAsync(M->theta, m->theta, sizeof m->theta, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->in_mu, m->in_mu, 4 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->in_sig, m->in_sig, 4 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->out_mu, m->out_mu, 2 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->out_sig, m->out_sig, 2 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->lo, m->lo, 4 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(M->hi, m->hi, 4 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(&M->tol, &m->tol, sizeof(float), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemsetAsync(&M->cells, 0, 2 * sizeof(unsigned long long), stream));
  int rc = launch_sur_pack(M, stream);
  if (!rc) rc = launch_sur_altab(P, M, stream);
  if (rc) return fail(rc, "surrogate model upload kernels failed");
  return WX_OK;
}

int sur_counters(SurDev *M, hipStream_t stream, int reset, int64_t *cells, int64_t *fallbacks) {
  unsigned long long c[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(c, &M->cells, sizeof c, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  if (cells) *cells = (int64_t)c[0];
  if (fallbacks) *fallbacks = (int64_t)c[1];
  if (reset) {
    HIPCHK(hipMemsetAsync(&M->cells, 0, 2 * sizeof(unsigned long long), stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  return WX_OK;
}

}  // namespace wx

namespace {

// A stream and device buffers for one standalone call; released on scope exit.
struct Scratch {
  hipStream_t s = nullptr;
  std::vector<void *> bufs;
  ~Scratch() {
    if (s) (void)hipStreamSynchronize(s);
    for (void *b : bufs) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
  }
  int init() {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return fail(WX_EHIP, "stream");
    return WX_OK;
  }
  template <class T>
  int alloc(T **p, size_t n) {
    if (hipMalloc((void **)p, (n ? n : 1) * sizeof(T)) != hipSuccess) return fail(WX_ENOMEM, "hipMalloc %zu", n * sizeof(T));
    bufs.push_back((void *)*p);
    return WX_OK;
  }
};

int model_on_device(Scratch &S, const wx_sur_model *m, const wx_params *p, SurDev **M) {
  int rc = S.alloc(M, 1);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(*M, 0, sizeof(SurDev), S.s));
  wx_params *dP = nullptr;
  if ((rc = S.alloc(&dP, 1))) return rc;
  wx_params pp{};
  if (p) pp = *p;
  HIPCHK(hipMemcpyAsync(dP, &pp, sizeof pp, hipMemcpyHostToDevice, S.s));
  if (m) return sur_upload(*M, m, dP, S.s);
  int rc2 = launch_sur_altab(dP, *M, S.s);
  return rc2 ? fail(rc2, "loss table kernel failed") : WX_OK;
}

}  // namespace

extern "C" {

int wx_sur_pairs(const wx_params *p, const double *states, int64_t n, double *out) {
  if (!p || (!states && n > 0) || (!out && n > 0) || n < 0) return fail(WX_EARG, "bad argument");
  if (p->neq < 5) return fail(WX_EARG, "coolingh pairs need neq >= 5");
  Scratch S;
  int rc = S.init();
  if (rc) return rc;
  wx_params *dP;
  double *din, *dout;
  if ((rc = S.alloc(&dP, 1)) || (rc = S.alloc(&din, (size_t)n * 4)) || (rc = S.alloc(&dout, (size_t)n * 2))) return rc;
  HIPCHK(hipMemcpyAsync(dP, p, sizeof *p, hipMemcpyHostToDevice, S.s));
  HIPCHK(hipMemcpyAsync(din, states, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice, S.s));
  if ((rc = launch_cool_pairs(dP, din, n, dout, S.s))) return fail(rc, "pairs kernel failed");
  HIPCHK(hipMemcpyAsync(out, dout, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipStreamSynchronize(S.s));
  return WX_OK;
}

int wx_sur_features(const wx_params *p, const double *states, int64_t n, float *feat, float *cedt) {
  if (!p || n < 0 || (n > 0 && (!states || !feat || !cedt))) return fail(WX_EARG, "bad argument");
  Scratch S;
  int rc = S.init();
  if (rc) return rc;
  SurDev *M;
  if ((rc = model_on_device(S, nullptr, p, &M))) return rc;
  // features do not depend on the network; the hull test inside is unused
  double *din;
  float *df, *dc;
  if ((rc = S.alloc(&din, (size_t)n * 4)) || (rc = S.alloc(&df, (size_t)n * 4)) || (rc = S.alloc(&dc, (size_t)n)))
    return rc;
  HIPCHK(hipMemcpyAsync(din, states, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice, S.s));
  if ((rc = launch_sur_feat(M, din, n, df, dc, S.s))) return fail(rc, "feature kernel failed");
  HIPCHK(hipMemcpyAsync(feat, df, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipMemcpyAsync(cedt, dc, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipStreamSynchronize(S.s));
  return WX_OK;
}

int wx_sur_infer(const wx_sur_model *m, const float *x, int64_t n, float *y) {
  if (!m || n < 0 || (n > 0 && (!x || !y))) return fail(WX_EARG, "bad argument");
  Scratch S;
  int rc = S.init();
  if (rc) return rc;
  SurDev *M;
  if ((rc = model_on_device(S, m, nullptr, &M))) return rc;
  float *dx, *dy;
  if ((rc = S.alloc(&dx, (size_t)n * 4)) || (rc = S.alloc(&dy, (size_t)n * 2))) return rc;
  HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, S.s));
  if ((rc = launch_sur_infer(M, dx, n, dy, S.s))) return fail(rc, "inference kernel failed");
  HIPCHK(hipMemcpyAsync(y, dy, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipStreamSynchronize(S.s));
  return WX_OK;
}

int wx_sur_grad(const wx_sur_model *m, const float *x, const float *t, int64_t n, float w0, float w1, float *loss,
                float *grad) {
  if (!m || n <= 0 || !x || !t || !grad) return fail(WX_EARG, "bad argument");
  Scratch S;
  int rc = S.init();
  if (rc) return rc;
  SurDev *M;
  if ((rc = model_on_device(S, m, nullptr, &M))) return rc;
  float *dx, *dt, *dg, *dl;
  if ((rc = S.alloc(&dx, (size_t)n * 4)) || (rc = S.alloc(&dt, (size_t)n * 2)) || (rc = S.alloc(&dg, SUR_NPARAM)) ||
      (rc = S.alloc(&dl, 1)))
    return rc;
  HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, S.s));
  HIPCHK(hipMemcpyAsync(dt, t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, S.s));
  HIPCHK(hipMemsetAsync(dg, 0, SUR_NPARAM * sizeof(float), S.s));
  HIPCHK(hipMemsetAsync(dl, 0, sizeof(float), S.s));
  if ((rc = launch_sur_grad(M, dx, dt, 0, n, n, w0, w1, dg, dl, S.s))) return fail(rc, "gradient kernel failed");
  HIPCHK(hipMemcpyAsync(grad, dg, SUR_NPARAM * sizeof(float), hipMemcpyDeviceToHost, S.s));
  float l = 0.f;
  HIPCHK(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipStreamSynchronize(S.s));
  if (loss) *loss = l;
  return WX_OK;
}

int wx_sur_train(wx_sur_model *m, const float *x, const float *t, int64_t n, int32_t batch, int32_t steps, float lr0,
                 float lr1, float w0, float w1, float *loss_hist) {
  if (!m || n <= 0 || !x || !t || batch <= 0 || steps < 0 || lr0 <= 0.f || lr1 <= 0.f)
    return fail(WX_EARG, "bad argument");
  Scratch S;
  int rc = S.init();
  if (rc) return rc;
  SurDev *M;
  if ((rc = model_on_device(S, m, nullptr, &M))) return rc;
  float *dx, *dt, *dg, *dm, *dv, *dl;
  if ((rc = S.alloc(&dx, (size_t)n * 4)) || (rc = S.alloc(&dt, (size_t)n * 2)) || (rc = S.alloc(&dg, SUR_NPARAM)) ||
      (rc = S.alloc(&dm, SUR_NPARAM)) || (rc = S.alloc(&dv, SUR_NPARAM)) || (rc = S.alloc(&dl, (size_t)steps + 1)))
    return rc;
  HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, S.s));
  HIPCHK(hipMemcpyAsync(dt, t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, S.s));
  HIPCHK(hipMemsetAsync(dm, 0, SUR_NPARAM * sizeof(float), S.s));
  HIPCHK(hipMemsetAsync(dv, 0, SUR_NPARAM * sizeof(float), S.s));
  HIPCHK(hipMemsetAsync(dl, 0, ((size_t)steps + 1) * sizeof(float), S.s));
  const int64_t nb = batch < n ? batch : n;
  for (int k = 0; k < steps; ++k) {
    const double frac = steps > 1 ? (double)k / (double)(steps - 1) : 0.0;
    const float lr = (float)((double)lr0 * std::pow((double)lr1 / (double)lr0, frac));
    const int64_t off = (int64_t)(((__int128)k * nb) % n);
    HIPCHK(hipMemsetAsync(dg, 0, SUR_NPARAM * sizeof(float), S.s));
    if ((rc = launch_sur_grad(M, dx, dt, off, nb, n, w0, w1, dg, dl + k, S.s))) return fail(rc, "gradient kernel");
    if ((rc = launch_sur_adam(M, dg, dm, dv, k + 1, lr, S.s))) return fail(rc, "adam kernel");
    if ((rc = launch_sur_pack(M, S.s))) return fail(rc, "pack kernel");
  }
  HIPCHK(hipMemcpyAsync(m->theta, M->theta, sizeof m->theta, hipMemcpyDeviceToHost, S.s));
  if (loss_hist && steps > 0)
    HIPCHK(hipMemcpyAsync(loss_hist, dl, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, S.s));
  HIPCHK(hipStreamSynchronize(S.s));
  return WX_OK;
}

}  // extern "C"

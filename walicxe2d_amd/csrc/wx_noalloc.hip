// wx_noalloc.hip -- wx_sur_infer_noalloc: the surrogate's inference for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// wx_sur_infer builds the device form of the model (the FP32 parameters,
// their bf16 copies in the matrix cores' operand order, the standardisation
// and hull) in memory it allocates per call.  This variant builds it in a
// fixed static device buffer instead, from a model the caller keeps in
// device memory, with kernels only (no allocation, no memset, no copy from
// host memory), on the caller's stream.  Every call on one GPU uses that
// GPU's buffer and leaves the last call's model in it; calls that overlap on
// one GPU (two streams, or two host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wx_runtime.h"
#include "wx_surrogate.h"

namespace wx {

// The scratch buffer: one device model (SurDev).  Each GPU the library runs
// on holds its own copy until the process ends.
constexpr size_t kNoAllocScratchBytes = 160 * 1024;
static_assert(sizeof(SurDev) <= kNoAllocScratchBytes, "scratch holds one device model");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch (the host-side name `scratch` is
// only a stand-in): looked up once per device, with that device current.
int scratch_ptr(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(WX_EARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(WX_EHIP, "no device address for the scratch buffer");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return WX_OK;
}

// Every field of the device model that inference reads, from the caller's
// model in device memory; the counters are cleared.  The bf16 operand copies
// follow (k_sur_pack).  The loss table is not used by inference.
__global__ __launch_bounds__(256) void k_sur_load(const wx_sur_model *__restrict__ m, SurDev *__restrict__ M) {
  const int stride = gridDim.x * blockDim.x;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < SUR_NPARAM; t += stride) M->theta[t] = m->theta[t];
  if (blockIdx.x == 0) {
    const int k = threadIdx.x;
    if (k < 4) {
      M->in_mu[k] = m->in_mu[k];
      M->in_sig[k] = m->in_sig[k];
      M->lo[k] = m->lo[k];
      M->hi[k] = m->hi[k];
    }
    if (k < 2) {
      M->out_mu[k] = m->out_mu[k];
      M->out_sig[k] = m->out_sig[k];
    }
    if (k == 0) {
      M->tol = m->tol;
      M->cells = 0;
      M->fallbacks = 0;
    }
  }
}

}  // namespace
}  // namespace wx

using namespace wx;

extern "C" int wx_sur_infer_noalloc(const wx_sur_model *d_model, const float *d_x, int64_t n, float *d_y,
                                    void *stream) {
  if (!d_model || n < 0 || (n > 0 && (!d_x || !d_y)))
    return fail(WX_EARG, "wx_sur_infer_noalloc: null pointer or negative n");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  void *p = nullptr;
  int rc = scratch_ptr(dev, &p);
  if (rc) return rc;
  SurDev *M = static_cast<SurDev *>(p);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_sur_load, dim3(35), dim3(256), 0, s, d_model, M);
  HIPCHK(hipGetLastError());
  if ((rc = launch_sur_pack(M, s))) return fail(rc, "wx_sur_infer_noalloc: pack kernel failed");
  if ((rc = launch_sur_infer(M, d_x, n, d_y, s))) return fail(rc, "wx_sur_infer_noalloc: inference kernel failed");
  return WX_OK;
}

// Host-only stand-in for the subset of the HIP runtime API that
// walicxe2d_amd/csrc/wx_host.cpp uses, so the host runtime (tree logic,
// list compilation, staging) can be built with g++ and run under
// AddressSanitizer on a machine without a GPU.  "Device" memory is host
// memory; copies are synchronous memcpy; events are no-ops.
//
// Graph capture is emulated: a kernel launch on a stream that the calling
// thread captures is recorded (its work and the pointers it was given), an
// executable graph replays the records, and every device or pinned
// allocation is tracked, so that freeing memory a live executable graph
// still references aborts at once (and a replay that touches freed memory
// is a heap-use-after-free under ASan).  Copies and memsets are refused on a
// capturing stream: the product captures kernel launches only.
// TEST SCAFFOLDING: never part of the product build.
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <utility>
#include <vector>

namespace hipsim {
struct Node {
  std::function<void()> run;
  std::vector<const void *> refs;
};
struct Graph {
  std::vector<Node> nodes;
};
inline std::mutex mu;
inline std::map<const char *, size_t> allocs;   // live allocations: base -> bytes
inline std::set<Graph *> execs;                  // live executable graphs
inline int captures = 0, replays = 0;            // counts, for the test's report
inline thread_local void *cap_stream = nullptr;  // stream this thread captures
inline thread_local Graph *cap_graph = nullptr;

inline void *track(void *p, size_t n) {
  if (p) {
    std::lock_guard<std::mutex> g(mu);
    allocs[(const char *)p] = n;
  }
  return p;
}
// Aborts when a live executable graph references memory in [p, p + size).
inline void release(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  auto it = allocs.find((const char *)p);
  if (it == allocs.end()) {
    fprintf(stderr, "hostsim: free of %p, which is not a live allocation\n", p);
    abort();
  }
  const char *lo = it->first, *hi = it->first + it->second;
  for (Graph *x : execs)
    for (const Node &nd : x->nodes)
      for (const void *r : nd.refs)
        if ((const char *)r >= lo && (const char *)r < hi) {
          fprintf(stderr, "hostsim: freeing %p (%zu bytes) that a live executable graph still references (%p)\n", p,
                  it->second, r);
          abort();
        }
  allocs.erase(it);
}
inline bool capturing(void *s) { return cap_graph && s == cap_stream; }
// A kernel launch: recorded when its stream is being captured by this
// thread, run at once otherwise.
inline void launch(void *s, std::vector<const void *> refs, std::function<void()> f) {
  if (capturing(s)) cap_graph->nodes.push_back(Node{std::move(f), std::move(refs)});
  else f();
}
}  // namespace hipsim

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2
struct int2 { int x, y; };

inline const char *hipGetErrorString(hipError_t) { return "hostsim"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
template <class T> inline hipError_t hipMalloc(T **p, size_t n) {
  *p = (T *)hipsim::track(calloc(1, n ? n : 1), n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p) { hipsim::release(p); free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned = 0) {
  *p = hipsim::track(calloc(1, n ? n : 1), n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipHostFree(void *p) { hipsim::release(p); free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st) {
  if (hipsim::capturing(st)) { fprintf(stderr, "hostsim: copy on a capturing stream\n"); abort(); }
  memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st) {
  if (hipsim::capturing(st)) { fprintf(stderr, "hostsim: memset on a capturing stream\n"); abort(); }
  memset(d, v, n);
  return hipSuccess;
}
// every stream a distinct handle (its address identifies it while capturing)
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = malloc(1); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = malloc(1); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = malloc(1); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
// launches run at once here, so a wait between streams has nothing to order
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { free(e); return hipSuccess; }
// Device queries, mapped pinned memory, blocking copies.
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
#define hipHostMallocMapped 2
#define hipHostMallocCoherent 4
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 4; return hipSuccess; }
inline hipError_t hipHostGetDevicePointer(void **d, void *h, unsigned) { *d = h; return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memmove(d, s, n); return hipSuccess; }
// Graph capture, instantiation, replay and destruction (see the top).
typedef hipsim::Graph *hipGraph_t;
typedef hipsim::Graph *hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 1 };
enum { hipErrorNotSupported = 801, hipErrorInvalidValue = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (hipsim::cap_graph) return hipErrorInvalidValue;
  hipsim::cap_graph = new hipsim::Graph;
  hipsim::cap_stream = s;
  std::lock_guard<std::mutex> l(hipsim::mu);
  ++hipsim::captures;
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!hipsim::cap_graph || s != hipsim::cap_stream) { *g = nullptr; return hipErrorInvalidValue; }
  *g = hipsim::cap_graph;
  hipsim::cap_graph = nullptr;
  hipsim::cap_stream = nullptr;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, void *, char *, size_t) {
  if (!g) return hipErrorInvalidValue;
  *x = new hipsim::Graph(*g);
  std::lock_guard<std::mutex> l(hipsim::mu);
  hipsim::execs.insert(*x);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (!x || hipsim::capturing(s)) return hipErrorInvalidValue;
  {
    std::lock_guard<std::mutex> l(hipsim::mu);
    if (!hipsim::execs.count(x)) { fprintf(stderr, "hostsim: launch of a destroyed graph\n"); abort(); }
    ++hipsim::replays;
  }
  for (const hipsim::Node &nd : x->nodes) nd.run();
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  if (!x) return hipSuccess;
  {
    std::lock_guard<std::mutex> l(hipsim::mu);
    hipsim::execs.erase(x);
  }
  delete x;
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }

// wx_runtime.cpp -- stream, pinned upload staging and event timing shared
// by the single-mesh context and the ensemble.
#include <cstring>

#include "wx_runtime.h"

namespace wx {

int Runtime::create_stream() {
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
    stream = nullptr;
    return fail(WX_EHIP, "hipStreamCreate failed");
  }
  return WX_OK;
}

void Runtime::destroy() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &st : stage) {
    if (st.ev) (void)hipEventDestroy(st.ev);
    if (st.buf) (void)hipHostFree(st.buf);
    if (st.hdesc) (void)hipHostFree(st.hdesc);
    st = Stage{};
  }
  prof_collect(this);
  for (auto e : evpool) (void)hipEventDestroy(e);
  evpool.clear();
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

// Mapped, coherent pinned memory and its device alias.
static int pinned_alloc(size_t bytes, char **host, char **dev) {
  *host = *dev = nullptr;
  if (hipHostMalloc((void **)host, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    *host = nullptr;
    return fail(WX_ENOMEM, "pinned staging %zu bytes", bytes);
  }
  if (hipHostGetDevicePointer((void **)dev, *host, 0) != hipSuccess) {
    (void)hipHostFree(*host);
    *host = nullptr;
    return fail(WX_EHIP, "device alias of the pinned staging");
  }
  return WX_OK;
}

// Opens the next staging buffer for `total` bytes (waits only if that
// buffer's previous copies have not finished).
int stage_begin(Runtime *r, size_t total) {
  r->stage_cur = (r->stage_cur + 1) % Runtime::NSTAGE;
  Runtime::Stage &st = r->stage[r->stage_cur];
  if (st.pending) {
    (void)hipEventSynchronize(st.ev);
    st.pending = false;
  }
  if (!st.ev && hipEventCreateWithFlags(&st.ev, hipEventDisableTiming) != hipSuccess)
    return fail(WX_EHIP, "hipEventCreate failed");
  if (st.cap < total) {
    if (st.buf) (void)hipHostFree(st.buf);
    st.buf = st.dbuf = nullptr;
    st.cap = 0;
    const size_t want = total + total / 2 + 4096;
    int rc = pinned_alloc(want, &st.buf, &st.dbuf);
    if (rc) return rc;
    st.cap = want;
  }
  st.used = 0;
  st.desc.clear();
  return WX_OK;
}

int stage_copy(Runtime *r, void *dst, const void *src, size_t bytes) {
  if (!bytes) return WX_OK;
  Runtime::Stage &st = r->stage[r->stage_cur];
  if (st.used + bytes > st.cap) return fail(WX_EARG, "staging overflow (%zu + %zu > %zu)", st.used, bytes, st.cap);
  char *p = st.buf + st.used;
  memcpy(p, src, bytes);
  st.desc.push_back(CopyDesc{dst, st.dbuf + st.used, (uint64_t)bytes});
  st.used += al256(bytes);
  return WX_OK;
}

char *stage_take(Runtime *r, size_t bytes) {
  Runtime::Stage &st = r->stage[r->stage_cur];
  if (st.used + bytes > st.cap) return nullptr;
  char *p = st.buf + st.used;
  st.used += al256(bytes);
  return p;
}

int stage_push(Runtime *r, void *dst, const char *host, size_t bytes) {
  if (!bytes) return WX_OK;
  Runtime::Stage &st = r->stage[r->stage_cur];
  if (host < st.buf || host + bytes > st.buf + st.cap) return fail(WX_EARG, "stage_push outside the open staging buffer");
  st.desc.push_back(CopyDesc{dst, st.dbuf + (host - st.buf), (uint64_t)bytes});
  return WX_OK;
}

const void *stage_dev(Runtime *r, const char *host) {
  Runtime::Stage &st = r->stage[r->stage_cur];
  if (host < st.buf || host >= st.buf + st.cap) return nullptr;
  return st.dbuf + (host - st.buf);
}

const void *stage_inplace(Runtime *r, const void *src, size_t bytes) {
  if (stage_begin(r, al256(bytes ? bytes : 1))) return nullptr;
  char *h = stage_take(r, bytes ? bytes : 1);
  if (!h) {
    fail(WX_EARG, "staging %zu bytes", bytes);
    return nullptr;
  }
  memcpy(h, src, bytes);
  return stage_dev(r, h);
}

int stage_end(Runtime *r) {
  Runtime::Stage &st = r->stage[r->stage_cur];
  const size_t n = st.desc.size();
  if (n) {
    if (st.dcap < n) {   // the buffer's previous round has finished (stage_begin)
      if (st.hdesc) (void)hipHostFree(st.hdesc);
      st.hdesc = st.ddesc = nullptr;
      st.dcap = 0;
      const size_t want = n + n / 2 + 16;
      char *h = nullptr, *d = nullptr;
      int rc = pinned_alloc(want * sizeof(CopyDesc), &h, &d);
      if (rc) return rc;
      st.hdesc = (CopyDesc *)h;
      st.ddesc = (CopyDesc *)d;
      st.dcap = want;
    }
    memcpy(st.hdesc, st.desc.data(), n * sizeof(CopyDesc));
    int rc = launch_h2d(st.ddesc, (int)n, r->stream);
    if (rc) return fail(rc, "upload launch failed: %s", hipGetErrorString(hipGetLastError()));
  }
  HIPCHK(hipEventRecord(st.ev, r->stream));
  st.pending = true;
  return WX_OK;
}

static hipEvent_t take_event(Runtime *r) {
  if (!r->evpool.empty()) {
    hipEvent_t e = r->evpool.back();
    r->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

ProfScope::ProfScope(Runtime *r_, int kid_, hipStream_t s_) : r(r_), kid(kid_), s(s_ ? s_ : r_->stream) {
  if (r->prof && kid >= 0) {
    a = take_event(r);
    (void)hipEventRecord(a, s);
  }
}

ProfScope::~ProfScope() {
  if (r->prof && a) {
    hipEvent_t b = take_event(r);
    (void)hipEventRecord(b, s);
    r->marks.push_back(Runtime::Mark{kid, a, b});
  }
}

HostTimer::~HostTimer() {
  if (!r->prof) return;
  r->prof_ms[slot] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  r->prof_n[slot] += 1;
}

void prof_collect(Runtime *r) {
  if (r->marks.empty()) return;
  (void)hipStreamSynchronize(r->stream);
  for (auto &m : r->marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
      r->prof_ms[m.kid] += ms;
      r->prof_n[m.kid] += 1;
    }
    r->evpool.push_back(m.a);
    r->evpool.push_back(m.b);
  }
  r->marks.clear();
}

int prof_report(Runtime *r, int32_t enable, double *ms, int64_t *counts, int32_t n) {
  prof_collect(r);
  for (int k = 0; k < NPROF && k < n; ++k) {
    if (ms) ms[k] = r->prof_ms[k];
    if (counts) counts[k] = r->prof_n[k];
  }
  if (enable >= 0) {
    r->prof = enable != 0;
    for (int k = 0; k < NPROF; ++k) {
      r->prof_ms[k] = 0;
      r->prof_n[k] = 0;
    }
  }
  return WX_OK;
}

WorkerPool::WorkerPool(int nthreads) {
  for (int k = 1; k < nthreads; ++k) threads_.emplace_back([this] { loop(); });
}

WorkerPool::~WorkerPool() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto &t : threads_) t.join();
}

void WorkerPool::loop() {
  uint64_t seen = 0;
  for (;;) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
    if (stop_) return;
    seen = gen_;
    ++active_;
    while (next_ < n_) {
      const int i = next_++;
      lk.unlock();
      (*fn_)(i);
      lk.lock();
    }
    if (--active_ == 0) done_cv_.notify_all();
  }
}

void WorkerPool::run(int n, const std::function<void(int)> &fn) {
  if (n <= 0) return;
  if (threads_.empty() || n == 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::unique_lock<std::mutex> lk(mu_);
  fn_ = &fn;
  n_ = n;
  next_ = 0;
  ++gen_;
  ++active_;   // the calling thread takes part
  cv_.notify_all();
  while (next_ < n_) {
    const int i = next_++;
    lk.unlock();
    fn(i);
    lk.lock();
  }
  --active_;
  done_cv_.wait(lk, [&] { return active_ == 0 && next_ >= n_; });
  fn_ = nullptr;
  n_ = 0;
}

}  // namespace wx

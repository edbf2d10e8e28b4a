// wx_host.cpp -- host runtime behind include/wx_capi.h.
//
// Owns one AMR mesh per context: the device tiles (two buffers that swap
// roles every step), the leaf list, the ghost-gather tables compiled from
// the reference's ordered boundary list, and -- for the standalone driver --
// the host tree (lp/leaf/icoord) with the reference's admesh logic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wx_device.h"

using namespace wx;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(WX_EHIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

template <class T>
struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  int ensure(size_t want) {
    if (want <= n && ptr) return WX_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size_t cap = std::max<size_t>(want, 1);
    if (hipMalloc(&ptr, cap * sizeof(T)) != hipSuccess) {
      ptr = nullptr;
      n = 0;
      return fail(WX_ENOMEM, "hipMalloc of %zu bytes failed", cap * sizeof(T));
    }
    n = cap;
    return WX_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
};

}  // namespace

struct wx_ctx {
  wx_params p{};
  int cap = 0;   // largest block ID
  int neq = 0;
  hipStream_t stream = nullptr;
  // device state
  DevBuf<double> U[2];
  int cur = 0;   // U[cur] holds the current state
  DevBuf<wx_params> dP;
  DevBuf<DevScalars> dS;
  DevBuf<LeafRec> dLeaves;
  DevBuf<BlockRec> dBinfo;
  DevBuf<int2> dOpsII, dOpsI;
  DevBuf<double> dStat;
  DevBuf<int8_t> dFlag;
  DevBuf<MeshOp> dMeshOps;
  DevBuf<double> dScratch;
  // host tree (reference arrays, 1-based IDs)
  std::vector<int32_t> lp, leaf, icoord;
  int nbmax = 0, nbleafs = 0;
  std::vector<int32_t> bounds;   // [e*3 + c]
  int ninternal = 0;
  // compiled tables
  std::vector<LeafRec> leaves;
  std::vector<BlockRec> binfo;
  std::vector<int2> opsII, opsI;
  std::vector<int> lidx;
  bool tables_valid = false;   // leaf list / binfo / ops match the tree and list
  bool bounds_external = false;
  // block statistics (CFL partials + refinement flag) are kept per block ID;
  // after a mesh change only the new leaves are recomputed
  bool stats_all_stale = true;
  std::vector<int> new_ids;
  std::vector<LeafRec> dirty;
  DevBuf<LeafRec> dDirty;
  // pinned host staging for uploads: two buffers, each guarded by an event,
  // so a table upload never waits for the device
  struct Stage { char *buf = nullptr; size_t cap = 0, used = 0; hipEvent_t ev = nullptr; bool pending = false; };
  Stage stage[2];
  int stage_cur = 0;
  // admesh scratch (flags per block ID)
  std::vector<uint8_t> up, dn;
  // per-kernel timing with HIP events on this context's stream
  bool prof = false;
  struct Mark { int kid; hipEvent_t a, b; };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> evpool;
  double prof_ms[8] = {0};
  int64_t prof_n[8] = {0};
  // pinned staging
  DevScalars *hS = nullptr;
  int8_t *hFlag = nullptr;
  size_t hFlagN = 0;

  int LP(int nb, int k) const { return lp[(size_t)nb * 8 + k - 1]; }
  void SETLP(int nb, int k, int v) { lp[(size_t)nb * 8 + k - 1] = v; }
  double *cur_u() { return U[cur].ptr; }
  double *other_u() { return U[cur ^ 1].ptr; }
};

namespace {

// Kernel IDs for the profiling counters.
enum { KID_STAGEA = 0, KID_STAGEB = 1, KID_FILLI = 2, KID_DTRED = 3, KID_STAT = 4, KID_MESH = 5 };
// Host-side phases, wall milliseconds (slots 6, 7 of the same counters):
// 6 = waiting for the step's flags + host admesh, 7 = list compile + upload.
enum { HID_ADMESH = 6, HID_TABLES = 7 };

struct HostTimer {
  wx_ctx *c;
  int slot;
  std::chrono::steady_clock::time_point t0;
  HostTimer(wx_ctx *c_, int slot_) : c(c_), slot(slot_), t0(std::chrono::steady_clock::now()) {}
  ~HostTimer();
};

hipEvent_t take_event(wx_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  wx_ctx *c;
  int kid;
  hipEvent_t a = nullptr;
  ProfScope(wx_ctx *c_, int kid_) : c(c_), kid(kid_) {
    if (c->prof) { a = take_event(c); (void)hipEventRecord(a, c->stream); }
  }
  ~ProfScope() {
    if (c->prof && a) {
      hipEvent_t b = take_event(c);
      (void)hipEventRecord(b, c->stream);
      c->marks.push_back(wx_ctx::Mark{kid, a, b});
    }
  }
};

HostTimer::~HostTimer() {
  if (!c->prof) return;
  c->prof_ms[slot] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->prof_n[slot] += 1;
}

void prof_collect(wx_ctx *c) {
  if (c->marks.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &m : c->marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
      c->prof_ms[m.kid] += ms;
      c->prof_n[m.kid] += 1;
    }
    c->evpool.push_back(m.a);
    c->evpool.push_back(m.b);
  }
  c->marks.clear();
}

// Opens the next staging buffer for `total` bytes (waits only if that buffer's
// previous copies have not finished).
int stage_begin(wx_ctx *c, size_t total) {
  c->stage_cur ^= 1;
  wx_ctx::Stage &st = c->stage[c->stage_cur];
  if (st.pending) {
    (void)hipEventSynchronize(st.ev);
    st.pending = false;
  }
  if (!st.ev && hipEventCreateWithFlags(&st.ev, hipEventDisableTiming) != hipSuccess)
    return fail(WX_EHIP, "hipEventCreate failed");
  if (st.cap < total) {
    if (st.buf) (void)hipHostFree(st.buf);
    st.buf = nullptr;
    st.cap = 0;
    size_t want = total + total / 2 + 4096;
    if (hipHostMalloc((void **)&st.buf, want) != hipSuccess) return fail(WX_ENOMEM, "pinned staging %zu bytes", want);
    st.cap = want;
  }
  st.used = 0;
  return WX_OK;
}

int stage_copy(wx_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!bytes) return WX_OK;
  wx_ctx::Stage &st = c->stage[c->stage_cur];
  char *p = st.buf + st.used;
  memcpy(p, src, bytes);
  st.used += (bytes + 255) & ~(size_t)255;
  HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, c->stream));
  return WX_OK;
}

void stage_end(wx_ctx *c) {
  wx_ctx::Stage &st = c->stage[c->stage_cur];
  (void)hipEventRecord(st.ev, c->stream);
  st.pending = true;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

namespace {

// ---------------------------------------------------------------------
// src/locatebounds.f90:29-540, MPIP forms, one rank.
void locatebounds(wx_ctx *c) {
  std::vector<uint8_t> bnd((size_t)(c->nbmax + 1) * 4, 0);
  c->bounds.clear();
  static const int opp[4] = {1, 0, 3, 2};
  auto add = [&](int a, int b, int code) {
    c->bounds.push_back(a);
    c->bounds.push_back(b);
    c->bounds.push_back(code);
  };
  for (int nb = 1; nb <= c->nbmax; ++nb) {
    if (!c->leaf[nb]) continue;
    for (int f = 0; f < 4; ++f) {
      if (bnd[(size_t)nb * 4 + f]) continue;
      const int nbr = c->LP(nb, 5 + f);
      if (nbr == -1) {
        bnd[(size_t)nb * 4 + f] = 1;
        add(nb, 0, -(f + 1));
        continue;
      }
      if (c->LP(nbr, 1) < c->LP(nb, 1)) continue;   // coarser side owns the face
      if (c->leaf[nbr]) {
        add(nb, nbr, 1 + 3 * f);
        bnd[(size_t)nb * 4 + f] = 1;
        bnd[(size_t)nbr * 4 + opp[f]] = 1;
      } else {
        bnd[(size_t)nb * 4 + f] = 1;
        int n1, n2;
        switch (f) {
          case 0: n1 = c->LP(c->LP(nbr, 3), 6); n2 = c->LP(n1, 8); break;   // :130-131
          case 1: n1 = c->LP(nbr, 3); n2 = c->LP(n1, 8); break;             // :237-238
          case 2: n1 = c->LP(c->LP(nbr, 3), 8); n2 = c->LP(n1, 6); break;   // :345-346
          default: n1 = c->LP(nbr, 3); n2 = c->LP(n1, 6); break;            // :452-453
        }
        add(nb, n1, 2 + 3 * f);
        add(nb, n2, 3 + 3 * f);
      }
    }
  }
  c->ninternal = (int)(c->bounds.size() / 3);
}

// ---------------------------------------------------------------------
// Compile the ordered list into per-segment gather ops (last writer wins,
// jet-imposition order folded into two flags).
struct Seg {
  int kind = GK_NONE, src = 0, off = 0, e = -1;
};

int compile_tables(wx_ctx *c) {
  const int cap = c->nbmax;   // IDs above nbmax hold no block
  c->leaves.clear();
  c->lidx.assign((size_t)cap + 1, -1);
  c->binfo.assign((size_t)cap + 1, BlockRec{0, 0, 0, 0});
  for (int nb = 1; nb <= c->nbmax; ++nb) {
    c->binfo[nb] = BlockRec{c->LP(nb, 1), c->icoord[(size_t)nb * 2], c->icoord[(size_t)nb * 2 + 1], 0};
    if (!c->leaf[nb]) continue;
    c->lidx[nb] = (int)c->leaves.size();
    LeafRec r{};
    r.id = nb;
    r.level = c->LP(nb, 1);
    r.icx = c->icoord[(size_t)nb * 2];
    r.icy = c->icoord[(size_t)nb * 2 + 1];
    r.lidx = (int)c->leaves.size();
    if (r.level < 1 || r.level > c->p.nlevs) return fail(WX_EARG, "block %d has level %d", nb, r.level);
    c->leaves.push_back(r);
  }
  const int nleaf = (int)c->leaves.size();
  std::vector<Seg> s2((size_t)nleaf * 8), s1((size_t)nleaf * 8);
  std::vector<int> ifirst((size_t)cap + 1, INT_MAX), ilast((size_t)cap + 1, -1);
  auto put = [&](std::vector<Seg> &T, int dst, int side, int half, int kind, int src, int off, int e) {
    if (dst < 1 || dst > cap) return;
    const int li = c->lidx[dst];
    if (li < 0) return;   // not a leaf: its data is never read again
    Seg &s = T[(size_t)li * 8 + side * 2 + half];
    s.kind = kind;
    s.src = src;
    s.off = off;
    s.e = e;
  };
  auto both = [&](int dst, int side, int kind, int src, int off, int e) {
    for (int h = 0; h < 2; ++h) { put(s2, dst, side, h, kind, src, off, e); put(s1, dst, side, h, kind, src, off, e); }
  };
  const bool outfyb_bug = (c->p.compat & 2) && c->p.bc[2] == 0;
  for (int e = 0; e < c->ninternal; ++e) {
    const int nbs = c->bounds[(size_t)e * 3], n1 = c->bounds[(size_t)e * 3 + 1], code = c->bounds[(size_t)e * 3 + 2];
    if (nbs < 1 || nbs > c->nbmax) return fail(WX_EARG, "bound %d: block %d out of range", e + 1, nbs);
    if (code < 0) {
      const int side = -code - 1;
      if (side > 3) return fail(WX_EARG, "bound %d: code %d", e + 1, code);
      const int kind = c->p.bc[side] ? GK_BOXREF : GK_BOXOUT;
      for (int h = 0; h < 2; ++h) put(s2, nbs, side, h, kind, nbs, 0, e);   // boundaryII.f90:38-141
      if (side == 2 && outfyb_bug) {
        const int k = e + 1;                                                // boundaryI.f90:87
        for (int h = 0; h < 2; ++h) put(s1, k, 2, h, GK_BOXOUT, k, 0, e);
      } else {
        for (int h = 0; h < 2; ++h) put(s1, nbs, side, h, kind, nbs, 0, e);
      }
    } else {
      if (n1 < 1 || n1 > c->nbmax) return fail(WX_EARG, "bound %d: neighbour %d out of range", e + 1, n1);
      switch (code) {
        case 1: both(nbs, 0, GK_COPY, n1, 0, e); both(n1, 1, GK_COPY, nbs, 0, e); break;
        case 4: both(nbs, 1, GK_COPY, n1, 0, e); both(n1, 0, GK_COPY, nbs, 0, e); break;
        case 7: both(nbs, 2, GK_COPY, n1, 0, e); both(n1, 3, GK_COPY, nbs, 0, e); break;
        case 10: both(nbs, 3, GK_COPY, n1, 0, e); both(n1, 2, GK_COPY, nbs, 0, e); break;
        case 2: case 3: case 5: case 6: case 8: case 9: case 11: case 12: {
          const int side = (code - 2) / 3;       // 0 L, 1 R, 2 B, 3 T
          const int h = (code - 2) % 3;          // 0: first fine block, 1: second
          const int fine_side = side ^ 1;        // the fine block's facing side
          put(s2, nbs, side, h, GK_RESTRICT, n1, 0, e);
          put(s1, nbs, side, h, GK_RESTRICT, n1, 0, e);
          both(n1, fine_side, GK_PROLONG, nbs, 8 * h, e);
          break;
        }
        default: return fail(WX_EARG, "bound %d: unknown code %d", e + 1, code);
      }
    }
    if (c->p.otherb) {   // impose_user_bc(nbs) after each entry
      ifirst[nbs] = std::min(ifirst[nbs], e);
      ilast[nbs] = std::max(ilast[nbs], e);
    }
  }
  auto encode = [&](const std::vector<Seg> &T, std::vector<int2> &out, bool full, const char *name) -> int {
    out.assign((size_t)nleaf * 8, int2{0, 0});
    for (int l = 0; l < nleaf; ++l) {
      const int dst = c->leaves[l].id;
      for (int s = 0; s < 8; ++s) {
        const Seg &g = T[(size_t)l * 8 + s];
        if (g.kind == GK_NONE) {
          if (full) return fail(WX_EARG, "%s: block %d ghost segment %d not covered by the list", name, dst, s);
          continue;
        }
        int y = g.kind | (g.off << GF_OFF_SHIFT);
        if (c->p.otherb) {
          if (ifirst[g.src] < g.e) y |= GF_SRC_IMPOSED;
          if (ilast[dst] >= g.e) y |= GF_DST_IMPOSED;
        }
        out[(size_t)l * 8 + s] = int2{g.src, y};
      }
    }
    return WX_OK;
  };
  int rc = encode(s2, c->opsII, true, "boundaryII");
  if (rc) return rc;
  rc = encode(s1, c->opsI, false, "boundaryI");
  if (rc) return rc;
  for (auto &r : c->leaves) r.flags = (c->p.otherb && ilast[r.id] >= 0) ? 1 : 0;
  return WX_OK;
}

int upload_tables(wx_ctx *c) {
  const size_t nleaf = c->leaves.size();
  int rc;
  if ((rc = c->dLeaves.ensure(nleaf))) return rc;
  if ((rc = c->dBinfo.ensure(c->binfo.size()))) return rc;
  if ((rc = c->dOpsII.ensure(nleaf * 8))) return rc;
  if ((rc = c->dOpsI.ensure(nleaf * 8))) return rc;
  const size_t bl = nleaf * sizeof(LeafRec), bo = nleaf * 8 * sizeof(int2), bb = c->binfo.size() * sizeof(BlockRec);
  if ((rc = stage_begin(c, al256(bl) + 2 * al256(bo) + al256(bb)))) return rc;
  if ((rc = stage_copy(c, c->dLeaves.ptr, c->leaves.data(), bl))) return rc;
  if ((rc = stage_copy(c, c->dOpsII.ptr, c->opsII.data(), bo))) return rc;
  if ((rc = stage_copy(c, c->dOpsI.ptr, c->opsI.data(), bo))) return rc;
  if ((rc = stage_copy(c, c->dBinfo.ptr, c->binfo.data(), bb))) return rc;
  stage_end(c);
  return WX_OK;
}

int ensure_tables(wx_ctx *c) {
  if (c->tables_valid) return WX_OK;
  HostTimer ht(c, HID_TABLES);
  if (!c->bounds_external) locatebounds(c);
  int rc = compile_tables(c);
  if (rc) return rc;
  rc = upload_tables(c);
  if (rc) return rc;
  c->tables_valid = true;
  return WX_OK;
}

int ensure_stats(wx_ctx *c) {
  int rc = ensure_tables(c);
  if (rc) return rc;
  if (c->stats_all_stale) {
    ProfScope ps(c, KID_STAT);
    rc = launch_blockstat(c->dP.ptr, c->neq, c->dLeaves.ptr, (int)c->leaves.size(), c->cur_u(),
                          c->dStat.ptr, c->dFlag.ptr, c->stream);
    if (rc) return fail(rc, "blockstat launch failed");
    c->stats_all_stale = false;
    c->new_ids.clear();
    return WX_OK;
  }
  if (c->new_ids.empty()) return WX_OK;
  c->dirty.clear();
  for (int id : c->new_ids)
    if (id >= 1 && id <= c->nbmax && c->lidx[id] >= 0) c->dirty.push_back(c->leaves[c->lidx[id]]);
  c->new_ids.clear();
  if (c->dirty.empty()) return WX_OK;
  if ((rc = c->dDirty.ensure(c->dirty.size()))) return rc;
  if ((rc = stage_begin(c, al256(c->dirty.size() * sizeof(LeafRec))))) return rc;
  if ((rc = stage_copy(c, c->dDirty.ptr, c->dirty.data(), c->dirty.size() * sizeof(LeafRec)))) return rc;
  stage_end(c);
  ProfScope ps(c, KID_STAT);
  rc = launch_blockstat(c->dP.ptr, c->neq, c->dDirty.ptr, (int)c->dirty.size(), c->cur_u(), c->dStat.ptr,
                        c->dFlag.ptr, c->stream);
  if (rc) return fail(rc, "blockstat launch failed");
  return WX_OK;
}

int sync_scalars(wx_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->hS, c->dS.ptr, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->hS->err) {
    int e = c->hS->err;
    return fail(WX_EHLLC, "Error in hllc (device flag %d): a Riemann problem had no valid star speed", e);
  }
  return WX_OK;
}

int run_stages(wx_ctx *c) {
  const int nleaf = (int)c->leaves.size();
  int rc;
  {
    ProfScope ps(c, KID_STAGEA);
    rc = launch_stageA(c->dP.ptr, c->neq, c->dLeaves.ptr, nleaf, c->cur_u(), c->other_u(), c->dS.ptr, c->stream);
  }
  if (!rc) {
    ProfScope ps(c, KID_STAGEB);
    rc = launch_stageB(c->dP.ptr, c->neq, c->dLeaves.ptr, nleaf, c->dOpsII.ptr, c->dBinfo.ptr,
                       c->cur_u(), c->other_u(), c->dS.ptr, c->stream);
  }
  if (!rc) {
    ProfScope ps(c, KID_FILLI);
    rc = launch_fillI_cool(c->dP.ptr, c->neq, c->dLeaves.ptr, nleaf, c->dOpsI.ptr, c->dBinfo.ptr,
                           c->cur_u(), c->other_u(), c->dStat.ptr, c->dFlag.ptr, c->dS.ptr, c->stream);
  }
  if (rc) return fail(rc, "stage launch failed: %s", hipGetErrorString(hipGetLastError()));
  c->cur ^= 1;
  c->stats_all_stale = false;   // the epilogue of the last kernel wrote every leaf's statistics
  c->new_ids.clear();
  return WX_OK;
}

// ---------------------------------------------------------------------
// Host tree logic: src/admesh.f90:4-111 (MPIP forms, one rank) with the
// pointer surgery of src/updatelpup.f90 and src/updatelpdown.f90.
void updatelpup(wx_ctx *c, int nb, int s1, int s2, int s3, int s4, std::vector<uint8_t> &up,
                std::vector<uint8_t> &dn) {
  const int nl2 = c->LP(nb, 1) + 1;
  c->nbleafs += 3;
  auto coarse_or_edge = [&](int n) { return n == -1 || c->leaf[n]; };
  int nsr;
  c->SETLP(s1, 1, nl2); c->SETLP(s1, 2, nb); c->SETLP(s1, 4, 1); c->SETLP(s1, 6, s2); c->SETLP(s1, 8, s3);
  if (coarse_or_edge(c->LP(nb, 5))) c->SETLP(s1, 5, c->LP(nb, 5));
  else { nsr = c->LP(c->LP(c->LP(nb, 5), 3), 6); c->SETLP(s1, 5, nsr); c->SETLP(nsr, 6, s1); }
  if (coarse_or_edge(c->LP(nb, 7))) c->SETLP(s1, 7, c->LP(nb, 7));
  else { nsr = c->LP(c->LP(c->LP(nb, 7), 3), 8); c->SETLP(s1, 7, nsr); c->SETLP(nsr, 8, s1); }

  c->SETLP(s2, 1, nl2); c->SETLP(s2, 2, nb); c->SETLP(s2, 4, 2); c->SETLP(s2, 5, s1); c->SETLP(s2, 8, s4);
  if (coarse_or_edge(c->LP(nb, 6))) c->SETLP(s2, 6, c->LP(nb, 6));
  else { nsr = c->LP(c->LP(nb, 6), 3); c->SETLP(s2, 6, nsr); c->SETLP(nsr, 5, s2); }
  if (coarse_or_edge(c->LP(nb, 7))) c->SETLP(s2, 7, c->LP(nb, 7));
  else { nsr = c->LP(c->LP(c->LP(c->LP(nb, 7), 3), 8), 6); c->SETLP(s2, 7, nsr); c->SETLP(nsr, 8, s2); }

  c->SETLP(s3, 1, nl2); c->SETLP(s3, 2, nb); c->SETLP(s3, 4, 3); c->SETLP(s3, 6, s4); c->SETLP(s3, 7, s1);
  if (coarse_or_edge(c->LP(nb, 5))) c->SETLP(s3, 5, c->LP(nb, 5));
  else { nsr = c->LP(c->LP(c->LP(c->LP(nb, 5), 3), 8), 6); c->SETLP(s3, 5, nsr); c->SETLP(nsr, 6, s3); }
  if (coarse_or_edge(c->LP(nb, 8))) c->SETLP(s3, 8, c->LP(nb, 8));
  else { nsr = c->LP(c->LP(nb, 8), 3); c->SETLP(s3, 8, nsr); c->SETLP(nsr, 7, s3); }

  c->SETLP(s4, 1, nl2); c->SETLP(s4, 2, nb); c->SETLP(s4, 4, 4); c->SETLP(s4, 5, s3); c->SETLP(s4, 7, s2);
  if (coarse_or_edge(c->LP(nb, 6))) c->SETLP(s4, 6, c->LP(nb, 6));
  else { nsr = c->LP(c->LP(c->LP(nb, 6), 3), 8); c->SETLP(s4, 6, nsr); c->SETLP(nsr, 5, s4); }
  if (coarse_or_edge(c->LP(nb, 8))) c->SETLP(s4, 8, c->LP(nb, 8));
  else { nsr = c->LP(c->LP(c->LP(nb, 8), 3), 6); c->SETLP(s4, 8, nsr); c->SETLP(nsr, 7, s4); }

  c->leaf[nb] = 0;
  c->leaf[s1] = c->leaf[s2] = c->leaf[s3] = c->leaf[s4] = 1;
  up[nb] = 0; dn[nb] = 0;
  dn[s1] = dn[s2] = dn[s3] = dn[s4] = 0;
  const int nx0 = c->icoord[(size_t)nb * 2] * 2, ny0 = c->icoord[(size_t)nb * 2 + 1] * 2;
  const int sons[4] = {s1, s2, s3, s4};
  for (int s = 0; s < 4; ++s) {   // src/refine.f90:227-261
    c->icoord[(size_t)sons[s] * 2] = nx0 + ((s & 1) ? NX : 0);
    c->icoord[(size_t)sons[s] * 2 + 1] = ny0 + ((s & 2) ? NY : 0);
  }
}

void updatelpdown(wx_ctx *c, int dad, int s1, int s2, int s3, int s4) {
  c->leaf[dad] = 1;
  c->leaf[s1] = c->leaf[s2] = c->leaf[s3] = c->leaf[s4] = 0;
  c->nbleafs -= 3;
  int nb1 = c->LP(dad, 5);
  if (nb1 != -1 && !c->leaf[nb1]) { nb1 = c->LP(c->LP(nb1, 3), 6); c->SETLP(nb1, 6, dad); c->SETLP(c->LP(nb1, 8), 6, dad); }
  nb1 = c->LP(dad, 6);
  if (nb1 != -1 && !c->leaf[nb1]) { nb1 = c->LP(nb1, 3); c->SETLP(nb1, 5, dad); c->SETLP(c->LP(nb1, 8), 5, dad); }
  nb1 = c->LP(dad, 7);
  if (nb1 != -1 && !c->leaf[nb1]) { nb1 = c->LP(c->LP(nb1, 3), 8); c->SETLP(nb1, 8, dad); c->SETLP(c->LP(nb1, 6), 8, dad); }
  nb1 = c->LP(dad, 8);
  if (nb1 != -1 && !c->leaf[nb1]) { nb1 = c->LP(nb1, 3); c->SETLP(nb1, 7, dad); c->SETLP(c->LP(nb1, 6), 7, dad); }
}

// Runs the tree part of admesh from per-ID flags; records data ops.
// The reference sweeps every block ID once per level (src/admesh.f90:29-105,
// src/critneighup.f90:14-57, src/critneighdown.f90:14-50); only flagged
// blocks ever act, so the sweeps here visit the flagged IDs, in ascending
// ID order within each level as the reference does.
int admesh_tree(wx_ctx *c, std::vector<uint8_t> &up, std::vector<uint8_t> &dn, std::vector<MeshOp> &ops,
                std::vector<int> &upl, std::vector<int> &dnl) {
  const int nlevs = c->p.nlevs;
  int nbmaxp = c->nbmax;
  const int nbmaxp1 = c->nbmax;
  // critneighup: closure of "refine my coarser neighbours", finest level first
  std::vector<std::vector<int>> byl((size_t)nlevs + 2);
  for (int nb : upl) byl[c->LP(nb, 1)].push_back(nb);
  for (int nl = nlevs - 1; nl >= 1; --nl)
    for (size_t k = 0; k < byl[nl].size(); ++k) {
      const int nb = byl[nl][k];
      for (int f = 5; f <= 8; ++f) {
        const int n1 = c->LP(nb, f);
        if (n1 != -1 && c->LP(n1, 1) < c->LP(nb, 1) && !up[n1]) {
          up[n1] = 1;
          byl[c->LP(n1, 1)].push_back(n1);
        }
      }
    }
  for (int nl = 1; nl <= nlevs - 1; ++nl) {         // refine by level, src/admesh.f90:29-69
    std::vector<int> &L = byl[nl];
    std::sort(L.begin(), L.end());
    L.erase(std::unique(L.begin(), L.end()), L.end());
    for (int nb : L) {
      if (nb > nbmaxp1 || !up[nb] || c->LP(nb, 1) != nl) continue;
      int s1, s2, s3, s4;
      if (c->LP(nb, 3) == 0) {
        s1 = nbmaxp + 1; s2 = s1 + 1; s3 = s1 + 2; s4 = s1 + 3;
        if (c->nbmax >= c->cap || nbmaxp + 4 > c->cap)
          return fail(WX_EBLOCKS, "nblocks exceeded limit %d %d", nbmaxp + 4, c->cap);
        c->SETLP(nb, 3, s1);
        nbmaxp += 4;
      } else {
        s1 = c->LP(nb, 3); s2 = c->LP(s1, 6); s3 = c->LP(s1, 8); s4 = c->LP(s3, 6);
      }
      c->nbmax = nbmaxp;
      ops.push_back(MeshOp{0, nb, s1, s2, s3, s4, {0, 0}});
      updatelpup(c, nb, s1, s2, s3, s4, up, dn);
    }
  }
  std::vector<int> cand;
  for (int nb : dnl) {                               // critneighdown, src/critneighdown.f90:14-50
    if (!dn[nb]) continue;
    for (int f = 5; f <= 8; ++f) {
      const int n1 = c->LP(nb, f);
      if (n1 != -1 && !c->leaf[n1]) dn[nb] = 0;
    }
    if (dn[nb] && nb <= nbmaxp1 && c->LP(nb, 4) == 1) cand.push_back(nb);
  }
  std::sort(cand.begin(), cand.end());
  for (int nl = nlevs; nl >= 2; --nl)                // coarsen by level, src/admesh.f90:75-105
    for (int nb : cand) {
      if (!(dn[nb] && c->LP(nb, 4) == 1 && c->LP(nb, 1) == nl)) continue;
      const int dad = c->LP(nb, 2), s1 = nb, s2 = c->LP(nb, 6), s3 = c->LP(nb, 8), s4 = c->LP(c->LP(nb, 6), 8);
      if (dn[s2] && dn[s3] && dn[s4]) {
        ops.push_back(MeshOp{1, dad, s1, s2, s3, s4, {0, 0}});
        updatelpdown(c, dad, s1, s2, s3, s4);
      }
    }
  // leave the scratch flags cleared for the next call
  for (int nl = 0; nl <= nlevs; ++nl)
    for (int nb : byl[nl]) up[nb] = 0;
  for (const MeshOp &o : ops)
    if (o.kind == 0) { up[o.a] = 0; dn[o.s1] = dn[o.s2] = dn[o.s3] = dn[o.s4] = 0; }
  for (int nb : dnl) dn[nb] = 0;
  return WX_OK;
}

int run_meshops(wx_ctx *c, const std::vector<MeshOp> &ops) {
  if (ops.empty()) return WX_OK;
  int rc = c->dMeshOps.ensure(ops.size());
  if (rc) return rc;
  if ((rc = stage_begin(c, al256(ops.size() * sizeof(MeshOp))))) return rc;
  if ((rc = stage_copy(c, c->dMeshOps.ptr, ops.data(), ops.size() * sizeof(MeshOp)))) return rc;
  stage_end(c);
  {
    ProfScope ps(c, KID_MESH);
    rc = launch_meshops(c->dP.ptr, c->neq, c->dMeshOps.ptr, (int)ops.size(), c->cur_u(), c->stream);
  }
  if (rc) return fail(rc, "meshops launch failed");
  return WX_OK;
}

int fetch_flags(wx_ctx *c) {
  const size_t n = (size_t)c->nbmax + 1;
  if (c->hFlagN < n) {
    if (c->hFlag) (void)hipHostFree(c->hFlag);
    c->hFlag = nullptr;
    c->hFlagN = 0;
    const size_t want = (size_t)c->cap + 1;
    HIPCHK(hipHostMalloc((void **)&c->hFlag, want));
    c->hFlagN = want;
  }
  HIPCHK(hipMemcpyAsync(c->hFlag, c->dFlag.ptr, n, hipMemcpyDeviceToHost, c->stream));
  return sync_scalars(c);
}

int do_admesh(wx_ctx *c, int *changed) {
  int rc = ensure_stats(c);
  if (rc) return rc;
  HostTimer ht(c, HID_ADMESH);
  rc = fetch_flags(c);
  if (rc) return rc;
  std::vector<uint8_t> &up = c->up, &dn = c->dn;
  if (up.size() != (size_t)c->cap + 1) {
    up.assign((size_t)c->cap + 1, 0);
    dn.assign((size_t)c->cap + 1, 0);
  }
  std::vector<int> upl, dnl;
  for (const LeafRec &r : c->leaves) {   // markref, src/markref.f90:37-44
    if (c->hFlag[r.id] == 1) { up[r.id] = 1; upl.push_back(r.id); }
    else if (c->hFlag[r.id] == 2) { dn[r.id] = 1; dnl.push_back(r.id); }
  }
  std::vector<MeshOp> ops;
  rc = admesh_tree(c, up, dn, ops, upl, dnl);
  if (rc) return rc;
  if (changed) *changed = (int)ops.size();
  if (!ops.empty()) {
    rc = run_meshops(c, ops);
    if (rc) return rc;
    for (const MeshOp &o : ops) {
      if (o.kind == 0) { c->new_ids.push_back(o.s1); c->new_ids.push_back(o.s2); c->new_ids.push_back(o.s3); c->new_ids.push_back(o.s4); }
      else c->new_ids.push_back(o.a);
    }
    c->tables_valid = false;
    c->bounds_external = false;
  }
  return WX_OK;
}

bool params_ok(const wx_params *p, std::string &why) {
  if (p->neq < 4 || p->neq > 6) { why = "neq must be 4..6"; return false; }
  if (p->nlevs < 1 || p->nlevs > WX_MAXLEV) { why = "nlevs out of range"; return false; }
  if (p->cooling == 1 && p->neq < 5) { why = "COOLINGH needs neq >= 5 (u(5) = n_HI)"; return false; }
  if (p->cooling < 0 || p->cooling > 2) { why = "cooling mode"; return false; }
  if (p->solver < 0 || p->solver > 2) { why = "solver"; return false; }
  if (p->limiter < -1 || p->limiter > 6) { why = "limiter"; return false; }
  if (p->nblocks < 1) { why = "nblocks"; return false; }
  for (int k = 0; k < 4; ++k)
    if (p->bc[k] < 0 || p->bc[k] > 1) { why = "bc must be 0 (outflow) or 1 (reflective)"; return false; }
  return true;
}

}  // namespace

// =====================================================================
extern "C" {

int wx_last_error(char *buf, int32_t len) {
  if (!buf || len <= 0) return WX_EARG;
  snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return WX_OK;
}

int wx_create(const wx_params *p, wx_ctx **out) {
  if (!p || !out) return fail(WX_EARG, "null argument");
  std::string why;
  if (!params_ok(p, why)) return fail(WX_EARG, "bad parameters: %s", why.c_str());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WX_EHIP, "no HIP device");
  wx_ctx *c = new wx_ctx();
  c->p = *p;
  c->cap = p->nblocks;
  c->neq = p->neq;
  int rc = WX_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(WX_EHIP, "hipStreamCreate failed");
  }
  const size_t tiles = (size_t)(c->cap + 1) * c->neq * NC;
  for (int k = 0; k < 2 && !rc; ++k) rc = c->U[k].ensure(tiles);
  if (!rc) rc = c->dP.ensure(1);
  if (!rc) rc = c->dS.ensure(1);
  if (!rc) rc = c->dStat.ensure((size_t)(c->cap + 1) * 2);
  if (!rc) rc = c->dFlag.ensure((size_t)c->cap + 1);
  if (rc) { wx_destroy(c); return rc; }
  for (int k = 0; k < 2; ++k) (void)hipMemsetAsync(c->U[k].ptr, 0, tiles * sizeof(double), c->stream);
  (void)hipMemcpyAsync(c->dP.ptr, p, sizeof(wx_params), hipMemcpyHostToDevice, c->stream);
  DevScalars z{};
  (void)hipMemcpyAsync(c->dS.ptr, &z, sizeof z, hipMemcpyHostToDevice, c->stream);
  if (hipHostMalloc((void **)&c->hS, sizeof(DevScalars)) != hipSuccess) { wx_destroy(c); return fail(WX_EHIP, "pinned alloc"); }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { wx_destroy(c); return fail(WX_EHIP, "init sync"); }
  c->lp.assign((size_t)(c->cap + 1) * 8, 0);
  c->leaf.assign((size_t)c->cap + 1, 0);
  c->icoord.assign((size_t)(c->cap + 1) * 2, 0);
  *out = c;
  return WX_OK;
}

int wx_destroy(wx_ctx *c) {
  if (!c) return WX_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &b : c->U) b.release();
  c->dP.release(); c->dS.release(); c->dLeaves.release(); c->dBinfo.release();
  c->dOpsII.release(); c->dOpsI.release(); c->dStat.release(); c->dFlag.release();
  c->dMeshOps.release(); c->dScratch.release(); c->dDirty.release();
  for (auto &st : c->stage) {
    if (st.ev) (void)hipEventDestroy(st.ev);
    if (st.buf) (void)hipHostFree(st.buf);
  }
  prof_collect(c);
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  if (c->hS) (void)hipHostFree(c->hS);
  if (c->hFlag) (void)hipHostFree(c->hFlag);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return WX_OK;
}

void *wx_stream(wx_ctx *c) { return c ? (void *)c->stream : nullptr; }

static int set_tree(wx_ctx *c, const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
                    int32_t nbmax, int32_t nbleafs, bool *changed = nullptr) {
  if (!lp || !leaf || !icoord) return fail(WX_EARG, "null tree array");
  if (nbmax < 0 || nbmax > c->cap || ld < nbmax) return fail(WX_EARG, "nbmax %d / ld %d vs capacity %d", nbmax, ld, c->cap);
  bool same = (nbmax == c->nbmax);
  for (int nb = 1; same && nb <= nbmax; ++nb) {
    for (int k = 1; k <= 8 && same; ++k) same = c->LP(nb, k) == lp[(size_t)(k - 1) * ld + (nb - 1)];
    same = same && (c->leaf[nb] == (leaf[nb - 1] != 0)) && c->icoord[(size_t)nb * 2] == icoord[nb - 1] &&
           c->icoord[(size_t)nb * 2 + 1] == icoord[(size_t)ld + (nb - 1)];
  }
  if (changed) *changed = !same;
  if (same) {
    c->nbleafs = nbleafs;
    return WX_OK;
  }
  std::fill(c->lp.begin(), c->lp.end(), 0);
  std::fill(c->leaf.begin(), c->leaf.end(), 0);
  std::fill(c->icoord.begin(), c->icoord.end(), 0);
  for (int nb = 1; nb <= nbmax; ++nb) {
    for (int k = 1; k <= 8; ++k) c->SETLP(nb, k, lp[(size_t)(k - 1) * ld + (nb - 1)]);
    c->leaf[nb] = leaf[nb - 1] != 0;
    c->icoord[(size_t)nb * 2] = icoord[(nb - 1)];
    c->icoord[(size_t)nb * 2 + 1] = icoord[(size_t)ld + (nb - 1)];
  }
  c->nbmax = nbmax;
  c->nbleafs = nbleafs;
  c->tables_valid = false;
  c->stats_all_stale = true;
  c->bounds_external = false;
  return WX_OK;
}

int wx_upload(wx_ctx *c, const double *u, int32_t layout, const int32_t *lp, const int32_t *leaf,
              const int32_t *icoord, int32_t ld, int32_t nbmax, int32_t nbleafs) {
  if (!c || !u) return fail(WX_EARG, "null argument");
  int rc = set_tree(c, lp, leaf, icoord, ld, nbmax, nbleafs);
  if (rc) return rc;
  c->stats_all_stale = true;
  const size_t tiles = (size_t)(c->cap + 1) * c->neq * NC;
  if (layout == WX_LAYOUT_BLOCK) {
    const size_t n = (size_t)(nbmax + 1) * c->neq * NC;
    HIPCHK(hipMemcpyAsync(c->cur_u(), u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else if (layout == WX_LAYOUT_FORTRAN) {
    const size_t n = (size_t)ld * c->neq * NC;
    if ((rc = c->dScratch.ensure(n))) return rc;
    HIPCHK(hipMemcpyAsync(c->dScratch.ptr, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (ld > c->cap) return fail(WX_EARG, "ld %d exceeds capacity %d", ld, c->cap);
    rc = launch_transpose(c->neq, ld, c->dScratch.ptr, c->cur_u(), 0, ld, c->stream);
    if (rc) return fail(rc, "transpose launch failed");
  } else {
    return fail(WX_EARG, "unknown layout %d", layout);
  }
  (void)tiles;
  HIPCHK(hipStreamSynchronize(c->stream));
  return WX_OK;
}

int wx_upload_tree(wx_ctx *c, const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
                   int32_t nbmax, int32_t nbleafs) {
  if (!c) return fail(WX_EARG, "null context");
  return set_tree(c, lp, leaf, icoord, ld, nbmax, nbleafs);
}

int wx_get_tree(wx_ctx *c, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld, int32_t *nbmax,
                int32_t *nbleafs) {
  if (!c) return fail(WX_EARG, "null context");
  if (ld < c->nbmax) return fail(WX_EARG, "ld %d < nbmax %d", ld, c->nbmax);
  for (int nb = 1; nb <= c->nbmax; ++nb) {
    if (lp) for (int k = 1; k <= 8; ++k) lp[(size_t)(k - 1) * ld + (nb - 1)] = c->LP(nb, k);
    if (leaf) leaf[nb - 1] = c->leaf[nb];
    if (icoord) {
      icoord[nb - 1] = c->icoord[(size_t)nb * 2];
      icoord[(size_t)ld + nb - 1] = c->icoord[(size_t)nb * 2 + 1];
    }
  }
  if (nbmax) *nbmax = c->nbmax;
  if (nbleafs) *nbleafs = c->nbleafs;
  return WX_OK;
}

int wx_download(wx_ctx *c, double *u, double *prim, int32_t layout, int32_t ld_in) {
  if (!c) return fail(WX_EARG, "null context");
  int rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)(c->nbmax + 1) * c->neq * NC;
  for (int pass = 0; pass < 2; ++pass) {
    double *dst = pass == 0 ? u : prim;
    if (!dst) continue;
    const double *src = c->cur_u();
    if (pass == 1) {
      if ((rc = c->dScratch.ensure((size_t)(c->cap + 1) * c->neq * NC))) return rc;
      rc = launch_calcprim(c->dP.ptr, c->neq, c->nbmax, c->cur_u(), c->dScratch.ptr, c->stream);
      if (rc) return fail(rc, "calcprim launch failed");
      src = c->dScratch.ptr;
    }
    if (layout == WX_LAYOUT_BLOCK) {
      HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    } else if (layout == WX_LAYOUT_FORTRAN) {
      const int ld = ld_in > 0 ? ld_in : c->nbmax;
      if (ld < c->nbmax || ld > c->cap) return fail(WX_EARG, "download: ld %d outside [nbmax %d, capacity %d]", ld, c->nbmax, c->cap);
      const size_t m = (size_t)ld * c->neq * NC;
      DevBuf<double> tmp;
      if ((rc = tmp.ensure(m))) return rc;
      rc = launch_transpose(c->neq, ld, src, tmp.ptr, 1, ld, c->stream);
      if (rc) { tmp.release(); return fail(rc, "transpose launch failed"); }
      hipError_t e1 = hipMemcpyAsync(dst, tmp.ptr, m * sizeof(double), hipMemcpyDeviceToHost, c->stream);
      hipError_t e2 = hipStreamSynchronize(c->stream);
      tmp.release();
      if (e1 != hipSuccess || e2 != hipSuccess) return fail(WX_EHIP, "download copy failed");
    } else {
      return fail(WX_EARG, "unknown layout %d", layout);
    }
  }
  return WX_OK;
}

int wx_set_bounds(wx_ctx *c, const int32_t *ib, int32_t ld, int32_t n) {
  if (!c || (!ib && n > 0) || n < 0 || ld < n) return fail(WX_EARG, "bad innerbounds");
  c->bounds.resize((size_t)n * 3);
  for (int e = 0; e < n; ++e)
    for (int k = 0; k < 3; ++k) c->bounds[(size_t)e * 3 + k] = ib[(size_t)k * ld + e];
  c->ninternal = n;
  c->bounds_external = true;
  c->tables_valid = false;
  int rc = compile_tables(c);
  if (rc) return rc;
  rc = upload_tables(c);
  if (rc) return rc;
  c->tables_valid = true;
  return WX_OK;
}

int wx_locatebounds(wx_ctx *c, int32_t *ninternal) {
  if (!c) return fail(WX_EARG, "null context");
  c->bounds_external = false;
  c->tables_valid = false;
  int rc = ensure_tables(c);
  if (rc) return rc;
  if (ninternal) *ninternal = c->ninternal;
  return WX_OK;
}

int wx_timestep(wx_ctx *c, double *dt) {
  if (!c || !dt) return fail(WX_EARG, "null argument");
  int rc = ensure_stats(c);
  if (rc) return rc;
  rc = launch_dtreduce(c->dP.ptr, c->dLeaves.ptr, c->dStat.ptr, (int)c->leaves.size(), c->dS.ptr, 0, c->stream);
  if (rc) return fail(rc, "dtreduce launch failed");
  rc = sync_scalars(c);
  if (rc) return rc;
  *dt = c->hS->dt;
  return WX_OK;
}

int wx_tstep(wx_ctx *c, double time, double dt) {
  if (!c) return fail(WX_EARG, "null context");
  int rc = ensure_tables(c);
  if (rc) return rc;
  // stage the scalars through pinned memory (stream-ordered)
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hS->time = time;
  c->hS->dt = dt;
  HIPCHK(hipMemcpyAsync(c->dS.ptr, c->hS, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return run_stages(c);
}

int wx_criteria(wx_ctx *c, int32_t *irefup, int32_t *irefdown) {
  if (!c || !irefup || !irefdown) return fail(WX_EARG, "null argument");
  int rc = ensure_stats(c);
  if (rc) return rc;
  rc = fetch_flags(c);
  if (rc) return rc;
  for (int nb = 1; nb <= c->cap; ++nb) { irefup[nb - 1] = 0; irefdown[nb - 1] = 0; }
  for (const LeafRec &r : c->leaves) {
    irefup[r.id - 1] = c->hFlag[r.id] == 1;
    irefdown[r.id - 1] = c->hFlag[r.id] == 2;
  }
  return WX_OK;
}

int wx_refine(wx_ctx *c, int32_t nb, int32_t s1, int32_t s2, int32_t s3, int32_t s4) {
  if (!c) return fail(WX_EARG, "null context");
  const int ids[5] = {nb, s1, s2, s3, s4};
  for (int k = 0; k < 5; ++k) if (ids[k] < 1 || ids[k] > c->cap) return fail(WX_EARG, "refine: block %d out of range", ids[k]);
  std::vector<MeshOp> ops{MeshOp{0, nb, s1, s2, s3, s4, {0, 0}}};
  c->stats_all_stale = true;
  c->tables_valid = false;
  return run_meshops(c, ops);
}

int wx_coarse(wx_ctx *c, int32_t dad, int32_t s1, int32_t s2, int32_t s3, int32_t s4) {
  if (!c) return fail(WX_EARG, "null context");
  const int ids[5] = {dad, s1, s2, s3, s4};
  for (int k = 0; k < 5; ++k) if (ids[k] < 1 || ids[k] > c->cap) return fail(WX_EARG, "coarse: block %d out of range", ids[k]);
  std::vector<MeshOp> ops{MeshOp{1, dad, s1, s2, s3, s4, {0, 0}}};
  c->stats_all_stale = true;
  c->tables_valid = false;
  return run_meshops(c, ops);
}

int wx_basegrid(wx_ctx *c, const double *uroots, int32_t nroots) {
  if (!c || !uroots) return fail(WX_EARG, "null argument");
  const int nrx = c->p.nbrootx, nry = c->p.nbrooty;
  if (nrx < 1 || nry < 1 || nrx * nry != nroots) return fail(WX_EARG, "nroots %d != %d x %d", nroots, nrx, nry);
  if (nroots > c->cap) return fail(WX_EBLOCKS, "capacity %d below root count %d", c->cap, nroots);
  std::fill(c->lp.begin(), c->lp.end(), 0);
  std::fill(c->leaf.begin(), c->leaf.end(), 0);
  std::fill(c->icoord.begin(), c->icoord.end(), 0);
  for (int iy = 0; iy < nry; ++iy)
    for (int ix = 0; ix < nrx; ++ix) {
      const int nb = 1 + ix + iy * nrx;
      c->SETLP(nb, 1, 1); c->SETLP(nb, 2, 0); c->SETLP(nb, 3, 0); c->SETLP(nb, 4, 1);
      c->SETLP(nb, 5, ix > 0 ? nb - 1 : -1);
      c->SETLP(nb, 6, ix < nrx - 1 ? nb + 1 : -1);
      c->SETLP(nb, 7, iy > 0 ? nb - nrx : -1);
      c->SETLP(nb, 8, iy < nry - 1 ? nb + nrx : -1);
      c->icoord[(size_t)nb * 2] = ix * NX;
      c->icoord[(size_t)nb * 2 + 1] = iy * NY;
      c->leaf[nb] = 1;
    }
  c->nbmax = nroots;
  c->nbleafs = nroots;
  const size_t n = (size_t)(nroots + 1) * c->neq * NC;
  HIPCHK(hipMemcpyAsync(c->cur_u(), uroots, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->tables_valid = false;
  c->stats_all_stale = true;
  c->bounds_external = false;
  return WX_OK;
}

int wx_admesh(wx_ctx *c, int32_t *changed) {
  if (!c) return fail(WX_EARG, "null context");
  int ch = 0;
  int rc = do_admesh(c, &ch);
  if (changed) *changed = ch;
  return rc;
}

int wx_set_time(wx_ctx *c, double time) {
  if (!c) return fail(WX_EARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hS->time = time;
  HIPCHK(hipMemcpyAsync(&c->dS.ptr->time, &c->hS->time, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return WX_OK;
}

int wx_step(wx_ctx *c, int32_t nsteps, int32_t do_admesh_flag, double *time, double *dt, double *zones) {
  if (!c || nsteps < 0) return fail(WX_EARG, "bad argument");
  double z = 0.0;
  for (int s = 0; s < nsteps; ++s) {
    int rc = ensure_stats(c);          // rebuilds tables after a mesh change
    if (rc) return rc;
    const int nleaf = (int)c->leaves.size();
    {
      ProfScope ps(c, KID_DTRED);
      rc = launch_dtreduce(c->dP.ptr, c->dLeaves.ptr, c->dStat.ptr, nleaf, c->dS.ptr, 1, c->stream);  // timestep; time += dt
    }
    if (rc) return fail(rc, "dtreduce launch failed");
    z += (double)nleaf * NX * NY;
    rc = run_stages(c);                // tstep
    if (rc) return rc;
    if (do_admesh_flag && c->p.nlevs > 1) {
      rc = do_admesh(c, nullptr);      // admesh
      if (rc) return rc;
    }
  }
  int rc = sync_scalars(c);
  if (rc) return rc;
  if (time) *time = c->hS->time;
  if (dt) *dt = c->hS->dt;
  if (zones) *zones = z;
  return WX_OK;
}

int wx_profile(wx_ctx *c, int32_t enable, double *ms, int64_t *counts, int32_t n) {
  if (!c) return fail(WX_EARG, "null context");
  prof_collect(c);
  for (int k = 0; k < 8 && k < n; ++k) {
    if (ms) ms[k] = c->prof_ms[k];
    if (counts) counts[k] = c->prof_n[k];
  }
  if (enable >= 0) {
    c->prof = enable != 0;
    for (int k = 0; k < 8; ++k) { c->prof_ms[k] = 0; c->prof_n[k] = 0; }
  }
  return WX_OK;
}

int wx_info(wx_ctx *c, int32_t *nbmax, int32_t *nbleafs, int32_t *ninternal) {
  if (!c) return fail(WX_EARG, "null context");
  if (nbmax) *nbmax = c->nbmax;
  if (nbleafs) *nbleafs = c->nbleafs;
  if (ninternal) *ninternal = c->ninternal;
  return WX_OK;
}

}  // extern "C"

// wx_runtime.h -- host-side pieces shared by the single-mesh context
// (wx_host.cpp) and the batched ensemble (wx_ensemble.cpp):
//   Mesh     the host tree of one AMR mesh (the reference's lp/leaf/icoord,
//            1-based block IDs), its ghost list and the gather tables
//            compiled from it, and the admesh tree logic (wx_mesh.cpp);
//   Runtime  one HIP stream with a ring of pinned upload staging buffers
//            (copied by one kernel per round, or read in place) and
//            per-kernel event timing (wx_runtime.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wx_device.h"

namespace wx {

// Records the message of the last failure (wx_last_error) and returns code.
int fail(int code, const char *fmt, ...);
const std::string &last_error();

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return ::wx::fail(WX_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

template <class T>
struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  int ensure(size_t want) {
    if (want <= n && ptr) return WX_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size_t cap = want > 1 ? want : 1;
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

// ---------------------------------------------------------------------
// One AMR mesh on the host (src/globals.f90:13-29 for one rank).
struct Mesh {
  wx_params p{};
  int cap = 0;   // largest block ID
  int neq = 0;
  std::vector<int32_t> lp, leaf, icoord;   // lp[nb*8 + k-1], leaf[nb], icoord[nb*2 + d]
  int nbmax = 0, nbleafs = 0;
  std::vector<int32_t> bounds;   // the ordered ghost list, [e*3 + c]
  std::vector<uint8_t> bndDone;  // locatebounds scratch
  int ninternal = 0;
  // gather tables compiled from the list
  std::vector<LeafRec> leaves;
  std::vector<BlockRec> binfo;
  std::vector<int2> opsII, opsI;
  bool opsI_same = false;        // fill I's ops equal fill II's (no OUTFYB entries)
  std::vector<int> lidx;
  std::vector<int> segE2, segE1, ifirst, ilast, segLate;   // compile_ops scratch
  bool tables_valid = false;     // leaves / binfo / ops match the tree and the list
  bool bounds_external = false;  // the list came from the host (wx_set_bounds)
  // block statistics (CFL partials + refinement flag) are kept per block ID;
  // after a mesh change only the new leaves need them recomputed
  bool stats_all_stale = true;
  // under DMC cooling: the state came from the host (basegrid, upload), so
  // the pressure carried from cooling to the next stage A must first be
  // taken from u, as the reference's calcprim does
  bool pc_stale = true;
  std::vector<int> new_ids;
  // admesh scratch flags per block ID (all zero between calls), and its lists
  std::vector<uint8_t> up, dn;
  std::vector<int> admUp, admDown;
  std::vector<std::vector<int>> admByLevel, admCand;

  // rows of lp / leaf / icoord written since the device copy of the tree
  // was last brought up to date (all of them while allDirty)
  std::vector<int32_t> dirtyRows;
  std::vector<uint8_t> rowDirty;
  bool allDirty = true;

  int LP(int nb, int k) const { return lp[(size_t)nb * 8 + k - 1]; }
  void touch(int nb) {
    if (allDirty || rowDirty[nb]) return;
    rowDirty[nb] = 1;
    dirtyRows.push_back(nb);
  }
  void SETLP(int nb, int k, int v) {
    lp[(size_t)nb * 8 + k - 1] = v;
    touch(nb);
  }
  void all_dirty() {
    for (int nb : dirtyRows) rowDirty[nb] = 0;
    dirtyRows.clear();
    allDirty = true;
  }
  void init(const wx_params &pp);
};

bool params_ok(const wx_params *p, std::string &why);
// src/locatebounds.f90:29-540 (MPIP forms, one rank) from the tree.
void locatebounds(Mesh *m);
// The ordered list -> per-leaf gather ops (last writer of every segment):
// compile_leaves (leaf and block records from the tree) then compile_ops
// (the ops from the list, and the leaves' imposition flags).
int compile_leaves(Mesh *m);
int compile_ops(Mesh *m);
int compile_tables(Mesh *m);
// locatebounds (unless the list is external) + compile_tables when stale.
int refresh_tables(Mesh *m);
// src/admesh.f90:4-111 from per-leaf flags (flag of leaves[l] at flags[l]:
// 1 refine, 2 coarsen); appends the data ops and records the new leaves.
int admesh_from_flags(Mesh *m, const int8_t *flag_of_leaf, std::vector<MeshOp> &ops);
// Generalised root layout of src/basegrid.f90:26-97 (nbrootx x nbrooty roots).
int basegrid_tree(Mesh *m);
// Tree from the reference's Fortran arrays; *changed tells whether it differs.
int set_tree(Mesh *m, const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
             int32_t nbmax, int32_t nbleafs, bool *changed);
int get_tree(const Mesh *m, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld, int32_t *nbmax,
             int32_t *nbleafs);

// ---------------------------------------------------------------------
// Kernel IDs for the profiling counters; slots 6 and 7 are host wall time.
enum { KID_STAGEA = 0, KID_STAGEB = 1, KID_FILLI = 2, KID_DTRED = 3, KID_STAT = 4, KID_MESH = 5,
       HID_ADMESH = 6, HID_TABLES = 7, KID_BUILD = 8, NPROF = 9 };

struct Runtime {
  hipStream_t stream = nullptr;
  // Upload staging: a ring of four mapped pinned buffers, each guarded by an
  // event.  The copies of one staging round are recorded as descriptors and
  // performed by one kernel launch (k_h2d reads the pinned memory directly)
  // at stage_end.
  struct Stage {
    char *buf = nullptr, *dbuf = nullptr;   // host pointer, device alias
    size_t cap = 0, used = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
    std::vector<CopyDesc> desc;
    CopyDesc *hdesc = nullptr, *ddesc = nullptr;
    size_t dcap = 0;
  };
  static constexpr int NSTAGE = 4;
  Stage stage[NSTAGE];
  int stage_cur = 0;
  // per-kernel timing with HIP events on this stream
  bool prof = false;
  struct Mark { int kid; hipEvent_t a, b; };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> evpool;
  double prof_ms[NPROF] = {0};
  int64_t prof_n[NPROF] = {0};

  int create_stream();
  void destroy();
};

int stage_begin(Runtime *r, size_t total);
int stage_copy(Runtime *r, void *dst, const void *src, size_t bytes);
// Reserves `bytes` of the open staging buffer for the caller to fill
// (nullptr if it does not fit); stage_push enqueues its copy to dst.
char *stage_take(Runtime *r, size_t bytes);
int stage_push(Runtime *r, void *dst, const char *host, size_t bytes);
// Device alias of host memory taken from the open staging buffer, for a
// kernel that reads it in place (launched before stage_end, which then
// guards the buffer until that kernel has run).
const void *stage_dev(Runtime *r, const char *host);
// stage_begin + a copy of `bytes` from src into the staging buffer; returns
// its device alias (nullptr on failure, with the error recorded).
const void *stage_inplace(Runtime *r, const void *src, size_t bytes);
// Launches the round's copies and records the buffer's event.
int stage_end(Runtime *r);
void prof_collect(Runtime *r);
int prof_report(Runtime *r, int32_t enable, double *ms, int64_t *counts, int32_t n);

// Events around the launches of its scope on stream s (the runtime's own
// stream when s is null).
struct ProfScope {
  Runtime *r;
  int kid;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(Runtime *r_, int kid_, hipStream_t s_ = nullptr);
  ~ProfScope();
};

struct HostTimer {
  Runtime *r;
  int slot;
  std::chrono::steady_clock::time_point t0;
  HostTimer(Runtime *r_, int slot_) : r(r_), slot(slot_), t0(std::chrono::steady_clock::now()) {}
  ~HostTimer();
};

// A fixed set of host worker threads for per-instance work (table
// compilation, admesh, staging).  run(n, fn) calls fn(i) for i in [0, n)
// across the workers and the calling thread and returns when all are done.
class WorkerPool {
 public:
  explicit WorkerPool(int nthreads);
  ~WorkerPool();
  int size() const { return (int)threads_.size() + 1; }
  void run(int n, const std::function<void(int)> &fn);

 private:
  void loop();
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)> *fn_ = nullptr;
  int n_ = 0, next_ = 0, active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace wx

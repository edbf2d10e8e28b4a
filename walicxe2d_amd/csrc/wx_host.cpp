// wx_host.cpp -- the single-mesh context behind include/wx_capi.h.
//
// Owns one AMR mesh: the device tiles (two buffers that swap roles every
// step), the block queue and ghost-gather tables compiled from the
// reference's ordered boundary list, and -- for the standalone driver -- the
// host tree (lp/leaf/icoord) with the reference's admesh logic (wx_mesh.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wx_runtime.h"
#include "wx_surrogate.h"

using namespace wx;

struct wx_ctx : Mesh, Runtime {
  // device state
  DevBuf<double> U[2];
  int cur = 0;   // U[cur] holds the current state
  DevBuf<wx_params> dP;
  DevBuf<DevScalars> dS;
  // the tables of the current mesh in one device arena, uploaded by one
  // copy per mesh change: queue header, instance offsets {0, nleaf}, leaf
  // records, ghost ops of fills II and I, block records
  DevBuf<char> dTab;
  QueueHdr *tQ = nullptr;
  int32_t *tOff = nullptr;
  LeafRec *tLeaves = nullptr;
  int2 *tOpsII = nullptr, *tOpsI = nullptr;
  BlockRec *tBinfo = nullptr;
  int32_t tLeafCap = 0;
  std::vector<BlockRec> tBinfoHost;   // what the arena's block records hold (empty: unknown)
  DevBuf<double> dStat;
  DevBuf<int8_t> dFlag;
  DevBuf<double> dScratch;
  std::vector<LeafRec> dirty;
  // pinned staging of the device scalars and the flags
  DevScalars *hS = nullptr, *hS_dev = nullptr;   // mapped: k_gather_flags writes it
  int8_t *hFlag = nullptr;
  size_t hFlagN = 0;
  // the flags by leaf position in mapped memory, written by the fill's
  // epilogue (flags_host: they are those of the current dFlag and leaf
  // order) or else by k_gather_flags
  int8_t *hFlagQ = nullptr, *hFlagQ_dev = nullptr;
  bool flags_host = false;
  // cooling surrogate (coolingh only), used when use_sur
  DevBuf<SurDev> dSur;
  bool use_sur = false;
  // under DMC cooling: the pressure the last cooling left in each leaf's
  // cells 0..17 (the reference's primit), carried into the next stage A
  // (Mesh::pc_stale: the state came from the host)
  DevBuf<double> dPc;
  double *pc() { return dPc.ptr; }
  // device table builder: the tree rows on the device and the builder's
  // per-ID scan results; used after a mesh change unless the list is
  // external (wx_set_bounds) or the host builder is selected
  DevBuf<TreeRow> dTree;
  DevBuf<LbInfo> dLbInfo;
  DevBuf<LbPart> dLbPart;
  DevBuf<int32_t> dLbLen;          // the list length it found
  std::vector<int32_t> treeIds;    // scratch: the rows of an upload
  hipStream_t side = nullptr;      // its stream (default priority, like the context's)
  hipEvent_t evTree = nullptr, evOps = nullptr;   // upload done; ops done
  int dev_builder = 3;             // wx_set_builder's mode: 0 host, 1 device, 2 device (many-workgroup scan), 3 by size
  bool ops_pending = false;        // launched; the context's stream has not waited for it yet
  bool tables_on_device = false;   // the current tables came from it (the list length is on the device)

  double *cur_u() { return U[cur].ptr; }
  double *other_u() { return U[cur ^ 1].ptr; }
  int nleaf() const { return (int)leaves.size(); }
  QueueRef queue() const { return QueueRef{tQ, tLeaves, tLeafCap}; }
};

namespace {

// The arena holds fixed regions for tLeafCap leaves (records, then the ops
// of fills II and I) and for every block ID (block records, last); it is
// reallocated only when the leaf count outgrows the regions.
int arena_fit(wx_ctx *c) {
  const size_t nleaf = c->leaves.size();
  if (c->tQ && nleaf <= (size_t)c->tLeafCap) return WX_OK;
  const size_t lcap = std::max<size_t>(1024, nleaf + nleaf / 4);
  const size_t total = 512 + al256(lcap * sizeof(LeafRec)) + 2 * al256(lcap * 8 * sizeof(int2)) +
                       al256(((size_t)c->cap + 1) * sizeof(BlockRec));
  int rc = c->dTab.ensure(total);
  if (rc) return rc;
  char *base = c->dTab.ptr;
  c->tQ = (QueueHdr *)base;
  c->tOff = (int32_t *)(base + 256);
  c->tLeaves = (LeafRec *)(base + 512);
  c->tOpsII = (int2 *)((char *)c->tLeaves + al256(lcap * sizeof(LeafRec)));
  c->tOpsI = (int2 *)((char *)c->tOpsII + al256(lcap * 8 * sizeof(int2)));
  c->tBinfo = (BlockRec *)((char *)c->tOpsI + al256(lcap * 8 * sizeof(int2)));
  c->tLeafCap = (int32_t)lcap;
  c->tBinfoHost.clear();   // new arena: every block record goes up
  return WX_OK;
}

// The leaf count from which the device builds the ops when the builder is
// chosen by size: below it the host compiles the list while the device runs
// the dt reduction and stage A, and that is the faster of the two (full
// cooling jet, 1144 leaves: 101 us/step with the host list, 107 with the
// device builder; point explosion, 5236 leaves: 202 and 192; 9939 leaves:
// 357 and 311).
constexpr size_t kDeviceBuilderLeaves = 3072;

bool use_device_builder(const wx_ctx *c) {
  if (c->dev_builder == 0 || c->bounds_external) return false;
  return c->dev_builder != 3 || c->leaves.size() >= kDeviceBuilderLeaves;
}

// Staged copies of the tree rows written since the device copy was last
// brought up to date (every row up to nbmax after a new tree), in runs of
// consecutive IDs; the staging round must be open with room for
// tree_bytes(c).
size_t tree_bytes(const wx_ctx *c) {
  const size_t n = c->allDirty ? (size_t)c->nbmax : c->dirtyRows.size();
  return al256(n * sizeof(TreeRow)) + n * 256;
}

int stage_tree(wx_ctx *c) {
  std::vector<int32_t> &ids = c->treeIds;
  ids.clear();
  if (c->allDirty)
    for (int nb = 1; nb <= c->nbmax; ++nb) ids.push_back(nb);
  else
    ids.assign(c->dirtyRows.begin(), c->dirtyRows.end());
  std::sort(ids.begin(), ids.end());
  int rc;
  for (size_t k = 0; k < ids.size();) {
    size_t e = k + 1;
    while (e < ids.size() && ids[e] == ids[e - 1] + 1) ++e;
    const size_t bytes = (e - k) * sizeof(TreeRow);
    TreeRow *hr = (TreeRow *)stage_take(c, bytes);
    if (!hr) return fail(WX_EARG, "staging %zu tree rows", e - k);
    for (size_t q = k; q < e; ++q) {
      const int nb = ids[q];
      if (nb < 1 || nb > c->cap) return fail(WX_EARG, "tree row %d outside 1..%d", nb, c->cap);
      TreeRow &r = hr[q - k];
      memcpy(r.lp, &c->lp[(size_t)nb * 8], sizeof r.lp);
      r.leaf = c->leaf[nb];
      r.icx = c->icoord[(size_t)nb * 2];
      r.icy = c->icoord[(size_t)nb * 2 + 1];
      r.pad = 0;
    }
    if ((rc = stage_push(c, c->dTree.ptr + ids[k], (const char *)hr, bytes))) return rc;
    k = e;
  }
  return WX_OK;
}

// First half of a table upload: the queue header, the leaf records and the
// block records (what the block statistics, the dt reduction and stage A
// read), and for the device builder the changed tree rows.
int upload_leaves(wx_ctx *c, bool tree) {
  int rc = arena_fit(c);
  if (rc) return rc;
  const size_t nleaf = c->leaves.size();
  const size_t bl = nleaf * sizeof(LeafRec), nb = c->binfo.size();
  // block records: only the runs that differ from what the arena holds (a
  // mesh change alters the records of a few new blocks)
  std::vector<BlockRec> &dev = c->tBinfoHost;
  const BlockRec *bi = c->binfo.data();
  std::vector<std::pair<size_t, size_t>> runs;   // [first, last) record indices
  size_t nchanged = 0;
  if (dev.size() != nb) {
    runs.emplace_back(0, nb);
    nchanged = nb;
  } else {
    for (size_t k = 0; k < nb;) {
      if (memcmp(&bi[k], &dev[k], sizeof(BlockRec)) == 0) { ++k; continue; }
      size_t e = k + 1;
      while (e < nb && memcmp(&bi[e], &dev[e], sizeof(BlockRec)) != 0) ++e;
      runs.emplace_back(k, e);
      nchanged += e - k;
      k = e;
    }
  }
  // unknown until the copies are launched: a failure on the way leaves the
  // shadow empty, so the next upload sends every record
  dev.clear();
  const size_t tb = tree ? tree_bytes(c) : 0;
  if ((rc = stage_begin(c, al256(512 + bl) + runs.size() * 256 + al256(nchanged * sizeof(BlockRec)) + tb))) return rc;
  char *h = stage_take(c, 512 + bl);
  if (!h) return fail(WX_EARG, "staging %zu bytes", 512 + bl);
  const QueueHdr qh{(int32_t)nleaf, 1, {0, 0}};
  const int32_t off[2] = {0, (int32_t)nleaf};
  memset(h, 0, 512);
  memcpy(h, &qh, sizeof qh);
  memcpy(h + 256, off, sizeof off);
  memcpy(h + 512, c->leaves.data(), bl);
  if ((rc = stage_push(c, c->dTab.ptr, h, 512 + bl))) return rc;
  for (const auto &r : runs) {
    const size_t bytes = (r.second - r.first) * sizeof(BlockRec);
    char *hb = stage_take(c, bytes);
    if (!hb) return fail(WX_EARG, "staging %zu bytes", bytes);
    memcpy(hb, bi + r.first, bytes);
    if ((rc = stage_push(c, c->tBinfo + r.first, hb, bytes))) return rc;
  }
  if (tree && (rc = stage_tree(c))) return rc;
  if ((rc = stage_end(c))) return rc;
  dev.assign(bi, bi + nb);
  if (tree) {   // the device copy of the tree is current once its copies are launched
    for (int id : c->dirtyRows) c->rowDirty[id] = 0;
    c->dirtyRows.clear();
    c->allDirty = false;
  }
  return WX_OK;
}

// Second half: the ops of both fills (fill I's only where they differ from
// fill II's, which the fill then reads instead), and the leaf records again
// when the list set imposition flags in them.
int upload_ops(wx_ctx *c) {
  const size_t nleaf = c->leaves.size();
  const size_t bo = nleaf * 8 * sizeof(int2), bl = nleaf * sizeof(LeafRec);
  const bool flags = c->p.otherb != 0, two = !c->opsI_same;
  int rc = stage_begin(c, (two ? 2 : 1) * al256(bo) + (flags ? al256(bl) : 0));
  if (rc) return rc;
  char *h2 = stage_take(c, bo), *h1 = two ? stage_take(c, bo) : nullptr;
  char *hl = flags ? stage_take(c, bl) : nullptr;
  if (!h2 || (two && !h1) || (flags && !hl)) return fail(WX_EARG, "staging %zu bytes", 2 * bo + bl);
  memcpy(h2, c->opsII.data(), bo);
  if ((rc = stage_push(c, c->tOpsII, h2, bo))) return rc;
  if (two) {
    memcpy(h1, c->opsI.data(), bo);
    if ((rc = stage_push(c, c->tOpsI, h1, bo))) return rc;
  }
  if (flags) {
    memcpy(hl, c->leaves.data(), bl);
    if ((rc = stage_push(c, c->tLeaves, hl, bl))) return rc;
  }
  return stage_end(c);
}

// After a mesh change the tables are rebuilt in two halves, so that a step
// can launch the statistics, the dt reduction and stage A (which need only
// the leaf records) before the ghost ops are ready.  With the device
// builder the first half also launches the builder, on its own stream
// after the upload; the second half makes the context's stream wait for
// it.  Otherwise the host locates and compiles the list in the second half
// while the device runs the first half's kernels.
int tables_leaves(wx_ctx *c) {
  HostTimer ht(c, HID_TABLES);
  c->flags_host = false;   // leaf positions change
  int rc = compile_leaves(c);
  const bool dev = use_device_builder(c);
  if (!rc) rc = upload_leaves(c, dev);
  if (rc || !dev) return rc;
  HIPCHK(hipEventRecord(c->evTree, c->stream));
  HIPCHK(hipStreamWaitEvent(c->side, c->evTree, 0));
  const LbOut out{c->dLbLen.ptr, c->tOpsII, c->tOpsI, c->tLeafCap};
  {
    ProfScope ps(c, KID_BUILD, c->side);
    rc = launch_build_tables(c->p, c->dTree.ptr, c->nbmax, c->nleaf(), c->dLbInfo.ptr, c->dLbPart.ptr, out,
                             c->dev_builder == 2, c->dS.ptr, c->side);
  }
  if (rc) return fail(rc, "table builder launch failed");
  HIPCHK(hipEventRecord(c->evOps, c->side));
  c->ops_pending = true;
  return WX_OK;
}

int tables_ops(wx_ctx *c) {
  HostTimer ht(c, HID_TABLES);
  if (c->ops_pending) {
    HIPCHK(hipStreamWaitEvent(c->stream, c->evOps, 0));
    c->ops_pending = false;
    c->opsI_same = !lb_outfyb(c->p);
    c->tables_on_device = true;
    c->tables_valid = true;
    return WX_OK;
  }
  c->tables_on_device = false;
  if (!c->bounds_external) locatebounds(c);
  int rc = compile_ops(c);
  if (rc) return rc;
  if ((rc = upload_ops(c))) return rc;
  c->tables_valid = true;
  return WX_OK;
}

int ensure_tables(wx_ctx *c) {
  if (c->tables_valid) return WX_OK;
  int rc = tables_leaves(c);
  return rc ? rc : tables_ops(c);
}

// The list length: the host's, or the device builder's (with a wait).
int list_length(wx_ctx *c, int32_t *n) {
  if (!c->tables_on_device || !c->tables_valid) {
    *n = c->ninternal;
    return WX_OK;
  }
  int32_t h = 0;
  HIPCHK(hipMemcpyAsync(&h, c->dLbLen.ptr, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *n = h;
  return WX_OK;
}

// Block statistics of the leaves that lack them (all, or the new leaves of
// the last mesh change); the leaf records must be current.
int update_stats(wx_ctx *c) {
  int rc;
  if (c->stats_all_stale || !c->new_ids.empty()) c->flags_host = false;   // blockstat writes flags
  if (c->stats_all_stale) {
    ProfScope ps(c, KID_STAT);
    rc = launch_blockstat(c->dP.ptr, c->neq, c->tLeaves, c->nleaf(), c->cur_u(), c->dStat.ptr,
                          c->dFlag.ptr, c->pc(), c->pc_stale ? 1 : 0, c->stream);
    c->pc_stale = false;
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
  const LeafRec *drec = (const LeafRec *)stage_inplace(c, c->dirty.data(), c->dirty.size() * sizeof(LeafRec));
  if (!drec) return WX_ENOMEM;
  {
    ProfScope ps(c, KID_STAT);
    rc = launch_blockstat(c->dP.ptr, c->neq, drec, (int)c->dirty.size(), c->cur_u(), c->dStat.ptr, c->dFlag.ptr,
                          c->pc(), c->pc_stale ? 1 : 0, c->stream);
  }
  if (rc) return fail(rc, "blockstat launch failed");
  return stage_end(c);
}

int ensure_stats(wx_ctx *c) {
  int rc = ensure_tables(c);
  return rc ? rc : update_stats(c);
}

// The device's error flags after a wait: a Riemann failure of the stages,
// or a tree the table builder could not read.
int check_scalars(wx_ctx *c) {
  if (c->hS->err) {
    int e = c->hS->err;
    return fail(WX_EHLLC, "Error in hllc (device flag %d): a Riemann problem had no valid star speed", e);
  }
  if (c->hS->mesh_err) {
    int e = c->hS->mesh_err;
    return fail(WX_EARG, "device table builder: the tree is inconsistent (code %d: %s)", e,
                e == LBE_POINTER ? "a pointer outside 1..nbmax" : e == LBE_FAMILY ? "a coarser neighbour's sons"
                : e == LBE_COUNT ? "leaf count" : "leaf level");
  }
  return WX_OK;
}

int sync_scalars(wx_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->hS, c->dS.ptr, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_scalars(c);
}

// tstep: stage A, then stage B and the fill (which need the ops tables).
// Stage A runs in 8-wave workgroups, the others in 4-wave ones: the fastest
// shapes on the full cooling jet (1162 leaves: stage A 14.4 ns per block at
// 512 threads, 16.0 at 256; stage B and the fill 2-5 ns slower at 384 or 512).
int run_stage_a(wx_ctx *c) {
  ProfScope ps(c, KID_STAGEA);
  int rc = launch_stageA(c->dP.ptr, c->neq, c->queue(), c->nleaf(), 512, 0, c->cur_u(), c->other_u(), c->pc(),
                         c->dS.ptr, c->stream);
  if (rc) return fail(rc, "stage launch failed: %s", hipGetErrorString(hipGetLastError()));
  return WX_OK;
}

int run_stages_bf(wx_ctx *c) {
  const int g = c->nleaf();   // one workgroup per queue entry
  int rc;
  {
    ProfScope ps(c, KID_STAGEB);
    rc = launch_stageB(c->dP.ptr, c->neq, c->queue(), g, 256, 0, c->tOpsII, c->tBinfo, c->cur_u(), c->other_u(),
                       c->dS.ptr, c->stream);
  }
  if (!rc) {
    ProfScope ps(c, KID_FILLI);
    const FlagOut fo{c->tOff, c->hFlagQ_dev, c->hS_dev, 1};
    const int2 *opsI = c->opsI_same ? c->tOpsII : c->tOpsI;
    if (c->use_sur)
      rc = launch_fillI_cool_sur(c->dP.ptr, c->neq, c->queue(), g, 256, 0, opsI, c->tBinfo, c->cur_u(),
                                 c->other_u(), c->dStat.ptr, c->dFlag.ptr, c->dS.ptr, c->dSur.ptr, fo, c->stream);
    else
      rc = launch_fillI_cool(c->dP.ptr, c->neq, c->queue(), g, 256, 0, opsI, c->tBinfo, c->cur_u(),
                             c->other_u(), c->dStat.ptr, c->dFlag.ptr, c->dS.ptr, fo, c->pc(), c->stream);
  }
  if (rc) return fail(rc, "stage launch failed: %s", hipGetErrorString(hipGetLastError()));
  c->cur ^= 1;
  c->flags_host = true;
  c->stats_all_stale = false;   // the epilogue of the last kernel wrote every leaf's statistics
  c->new_ids.clear();
  return WX_OK;
}

int run_stages(wx_ctx *c) {
  int rc = run_stage_a(c);
  return rc ? rc : run_stages_bf(c);
}

// The data ops of a mesh change; with `stats` (the ops of this context's own
// admesh, whose tree already holds the new leaves) the kernel also computes
// the new leaves' statistics, so they need no k_blockstat.
int run_meshops(wx_ctx *c, std::vector<MeshOp> &ops, bool stats) {
  if (ops.empty()) return WX_OK;
  if (stats)
    for (MeshOp &o : ops) {
      o.level = c->LP(o.kind == 0 ? o.s1 : o.a, 1);
      o.inst = 0;
    }
  // the kernel reads the op list in place from the pinned staging
  const MeshOp *dops = (const MeshOp *)stage_inplace(c, ops.data(), ops.size() * sizeof(MeshOp));
  if (!dops) return WX_ENOMEM;
  int rc;
  {
    ProfScope ps(c, KID_MESH);
    rc = launch_meshops(c->dP.ptr, c->neq, dops, (int)ops.size(), c->cur_u(), stats ? c->dStat.ptr : nullptr,
                        stats ? c->dFlag.ptr : nullptr, c->pc(), c->stream);
  }
  if (rc) return fail(rc, "meshops launch failed");
  if ((rc = stage_end(c))) return rc;
  if (stats) c->new_ids.clear();   // the new leaves' statistics are done
  return WX_OK;
}

// Flags by block ID (1 refine, 2 coarsen) into hFlag[0..nbmax].
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

// The flags of the leaves by leaf position and the device scalars, written
// by one kernel into mapped host memory; then the step's wait and the
// Riemann-failure check of sync_scalars.
int fetch_leaf_flags(wx_ctx *c) {
  const int n = c->nleaf();
  if (!c->flags_host) {   // else the last fill's epilogue wrote them
    int rc = launch_gather_flags(c->queue(), c->tOff, std::max(1, std::min((n + 255) / 256, 1024)), c->dFlag.ptr,
                                 c->hFlagQ_dev, c->dS.ptr, c->hS_dev, 1, c->stream);
    if (rc) return fail(rc, "flag gather launch failed");
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_scalars(c);
}

int do_admesh(wx_ctx *c, int *changed) {
  int rc = ensure_stats(c);
  if (rc) return rc;
  rc = fetch_leaf_flags(c);
  if (rc) return rc;
  HostTimer ht(c, HID_ADMESH);   // host work only (after the step's wait)
  std::vector<MeshOp> ops;
  rc = admesh_from_flags(c, c->hFlagQ, ops);
  if (rc) return rc;
  if (changed) *changed = (int)ops.size();
  return run_meshops(c, ops, true);
}

}  // namespace

// =====================================================================
extern "C" {

int wx_last_error(char *buf, int32_t len) {
  if (!buf || len <= 0) return WX_EARG;
  snprintf(buf, (size_t)len, "%s", last_error().c_str());
  return WX_OK;
}

int wx_create(const wx_params *p, wx_ctx **out) {
  if (!p || !out) return fail(WX_EARG, "null argument");
  std::string why;
  if (!params_ok(p, why)) return fail(WX_EARG, "bad parameters: %s", why.c_str());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WX_EHIP, "no HIP device");
  wx_ctx *c = new wx_ctx();
  c->init(*p);
  int rc = c->create_stream();
  if (rc) { delete c; return rc; }
  const size_t tiles = (size_t)(c->cap + 1) * c->neq * NC;
  for (int k = 0; k < 2 && !rc; ++k) rc = c->U[k].ensure(tiles);
  if (!rc) rc = c->dP.ensure(1);
  if (!rc) rc = c->dS.ensure(1);
  if (!rc) rc = c->dStat.ensure((size_t)(c->cap + 1) * 2);
  if (!rc) rc = c->dFlag.ensure((size_t)c->cap + 1);
  if (!rc) rc = c->dTree.ensure((size_t)c->cap + 1);
  if (!rc) rc = c->dLbInfo.ensure((size_t)c->cap + 1);
  if (!rc) rc = c->dLbPart.ensure((size_t)(c->cap + 256) / 256);
  if (!rc) rc = c->dLbLen.ensure(1);
  if (!rc && p->cooling == 2) rc = c->dPc.ensure((size_t)(c->cap + 1) * NC);
  if (rc) { wx_destroy(c); return rc; }
  if (c->dPc.ptr) (void)hipMemsetAsync(c->dPc.ptr, 0, c->dPc.n * sizeof(double), c->stream);
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
    c->side = nullptr;
    wx_destroy(c);
    return fail(WX_EHIP, "builder stream");
  }
  if (hipEventCreateWithFlags(&c->evTree, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->evOps, hipEventDisableTiming) != hipSuccess) {
    wx_destroy(c);
    return fail(WX_EHIP, "builder events");
  }
  for (int k = 0; k < 2; ++k) (void)hipMemsetAsync(c->U[k].ptr, 0, tiles * sizeof(double), c->stream);
  (void)hipMemcpyAsync(c->dP.ptr, p, sizeof(wx_params), hipMemcpyHostToDevice, c->stream);
  DevScalars z{};
  (void)hipMemcpyAsync(c->dS.ptr, &z, sizeof z, hipMemcpyHostToDevice, c->stream);
  if (hipHostMalloc((void **)&c->hS, sizeof(DevScalars), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&c->hFlagQ, (size_t)c->cap + 1, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    wx_destroy(c);
    return fail(WX_EHIP, "pinned alloc");
  }
  if (hipHostGetDevicePointer((void **)&c->hS_dev, c->hS, 0) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->hFlagQ_dev, c->hFlagQ, 0) != hipSuccess) {
    wx_destroy(c);
    return fail(WX_EHIP, "mapped pointers of the pinned mirrors");
  }
  memset(c->hS, 0, sizeof(DevScalars));
  memset(c->hFlagQ, 0, (size_t)c->cap + 1);
  if (hipStreamSynchronize(c->stream) != hipSuccess) { wx_destroy(c); return fail(WX_EHIP, "init sync"); }
  *out = c;
  return WX_OK;
}

int wx_destroy(wx_ctx *c) {
  if (!c) return WX_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (auto &b : c->U) b.release();
  c->dP.release(); c->dS.release(); c->dTab.release(); c->dStat.release(); c->dFlag.release();
  c->dScratch.release(); c->dSur.release(); c->dTree.release(); c->dLbInfo.release(); c->dLbPart.release();
  c->dLbLen.release(); c->dPc.release();
  if (c->evTree) (void)hipEventDestroy(c->evTree);
  if (c->evOps) (void)hipEventDestroy(c->evOps);
  if (c->side) (void)hipStreamDestroy(c->side);
  c->Runtime::destroy();
  if (c->hS) (void)hipHostFree(c->hS);
  if (c->hFlag) (void)hipHostFree(c->hFlag);
  if (c->hFlagQ) (void)hipHostFree(c->hFlagQ);
  delete c;
  return WX_OK;
}

void *wx_stream(wx_ctx *c) { return c ? (void *)c->stream : nullptr; }

int wx_upload(wx_ctx *c, const double *u, int32_t layout, const int32_t *lp, const int32_t *leaf,
              const int32_t *icoord, int32_t ld, int32_t nbmax, int32_t nbleafs) {
  if (!c || !u) return fail(WX_EARG, "null argument");
  int rc = set_tree(c, lp, leaf, icoord, ld, nbmax, nbleafs, nullptr);
  if (rc) return rc;
  c->stats_all_stale = true;
  c->pc_stale = true;
  if (layout == WX_LAYOUT_BLOCK) {
    const size_t n = (size_t)(nbmax + 1) * c->neq * NC;
    HIPCHK(hipMemcpyAsync(c->cur_u(), u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else if (layout == WX_LAYOUT_FORTRAN) {
    if (ld > c->cap) return fail(WX_EARG, "ld %d exceeds capacity %d", ld, c->cap);
    const size_t n = (size_t)ld * c->neq * NC;
    if ((rc = c->dScratch.ensure(n))) return rc;
    HIPCHK(hipMemcpyAsync(c->dScratch.ptr, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    rc = launch_transpose(c->neq, ld, c->dScratch.ptr, c->cur_u(), 0, ld, c->stream);
    if (rc) return fail(rc, "transpose launch failed");
  } else {
    return fail(WX_EARG, "unknown layout %d", layout);
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return WX_OK;
}

int wx_upload_tree(wx_ctx *c, const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
                   int32_t nbmax, int32_t nbleafs) {
  if (!c) return fail(WX_EARG, "null context");
  return set_tree(c, lp, leaf, icoord, ld, nbmax, nbleafs, nullptr);
}

int wx_get_tree(wx_ctx *c, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld, int32_t *nbmax,
                int32_t *nbleafs) {
  if (!c) return fail(WX_EARG, "null context");
  return get_tree(c, lp, leaf, icoord, ld, nbmax, nbleafs);
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
  return ensure_tables(c);
}

int wx_locatebounds(wx_ctx *c, int32_t *ninternal) {
  if (!c) return fail(WX_EARG, "null context");
  c->bounds_external = false;
  c->tables_valid = false;
  int rc = ensure_tables(c);
  if (rc) return rc;
  if (ninternal && (rc = list_length(c, ninternal))) return rc;
  return c->tables_on_device ? sync_scalars(c) : WX_OK;
}

int wx_set_builder(wx_ctx *c, int32_t on_device) {
  if (!c || on_device < 0 || on_device > 3) return fail(WX_EARG, "bad argument");
  if (c->ops_pending) return fail(WX_EARG, "wx_set_builder: a table build is in progress");
  if (on_device != c->dev_builder) {
    c->dev_builder = on_device;
    c->tables_valid = false;   // the next step builds them the new way
  }
  return WX_OK;
}

int wx_get_tables(wx_ctx *c, int32_t *nleaf, int32_t *leaves, int32_t *ops2, int32_t *ops1, int32_t cap) {
  if (!c || !nleaf) return fail(WX_EARG, "null argument");
  int rc = ensure_tables(c);
  if (rc) return rc;
  const int n = c->nleaf();
  *nleaf = n;
  if (n > cap) return WX_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = sync_scalars(c))) return rc;
  if (leaves) HIPCHK(hipMemcpy(leaves, c->tLeaves, (size_t)n * sizeof(LeafRec), hipMemcpyDeviceToHost));
  if (ops2) HIPCHK(hipMemcpy(ops2, c->tOpsII, (size_t)n * 8 * sizeof(int2), hipMemcpyDeviceToHost));
  if (ops1)
    HIPCHK(hipMemcpy(ops1, c->opsI_same ? c->tOpsII : c->tOpsI, (size_t)n * 8 * sizeof(int2), hipMemcpyDeviceToHost));
  return WX_OK;
}

// Jet source speed at a time the host sets (JET/jet.f90:55), computed here
// with the host C library's sin -- the function the reference calls -- and
// handed to the stages with that time.  (When the device advances the time
// itself, the dt reduction computes the speed.)  Caller has synchronised the
// stream, so the pinned scalars are free.
static int stage_jet_speed(wx_ctx *c, double time) {
  if (!c->p.otherb) return WX_OK;
  const double pi = 3.141592653589793;  // acos(-1.)
  c->hS->vjet = c->p.jet[6] + c->p.jet[7] * std::sin(2. * pi * time / c->p.jet[8]);
  c->hS->vjet_time = time;
  c->hS->vjet_ok = 1;
  HIPCHK(hipMemcpyAsync(&c->dS.ptr->vjet, &c->hS->vjet, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(&c->dS.ptr->vjet_ok, &c->hS->vjet_ok, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  return WX_OK;
}

int wx_timestep(wx_ctx *c, double *dt) {
  if (!c || !dt) return fail(WX_EARG, "null argument");
  int rc = ensure_stats(c);
  if (rc) return rc;
  rc = launch_dtreduce(c->dP.ptr, c->tLeaves, c->dStat.ptr, c->tOff, 1, c->dS.ptr, 0, c->stream);
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
  rc = stage_jet_speed(c, time);
  if (rc) return rc;
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
  std::vector<MeshOp> ops{MeshOp{0, nb, s1, s2, s3, s4, 0, 0}};
  c->stats_all_stale = true;
  c->tables_valid = false;
  return run_meshops(c, ops, false);
}

int wx_coarse(wx_ctx *c, int32_t dad, int32_t s1, int32_t s2, int32_t s3, int32_t s4) {
  if (!c) return fail(WX_EARG, "null context");
  const int ids[5] = {dad, s1, s2, s3, s4};
  for (int k = 0; k < 5; ++k) if (ids[k] < 1 || ids[k] > c->cap) return fail(WX_EARG, "coarse: block %d out of range", ids[k]);
  std::vector<MeshOp> ops{MeshOp{1, dad, s1, s2, s3, s4, 0, 0}};
  c->stats_all_stale = true;
  c->tables_valid = false;
  return run_meshops(c, ops, false);
}

int wx_basegrid(wx_ctx *c, const double *uroots, int32_t nroots) {
  if (!c || !uroots) return fail(WX_EARG, "null argument");
  if (c->p.nbrootx * c->p.nbrooty != nroots)
    return fail(WX_EARG, "nroots %d != %d x %d", nroots, c->p.nbrootx, c->p.nbrooty);
  int rc = basegrid_tree(c);
  if (rc) return rc;
  c->pc_stale = true;
  const size_t n = (size_t)(nroots + 1) * c->neq * NC;
  HIPCHK(hipMemcpyAsync(c->cur_u(), uroots, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
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
  int rc = stage_jet_speed(c, time);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return WX_OK;
}

int wx_step(wx_ctx *c, int32_t nsteps, int32_t do_admesh_flag, double *time, double *dt, double *zones) {
  if (!c || nsteps < 0) return fail(WX_EARG, "bad argument");
  double z = 0.0;
  for (int s = 0; s < nsteps; ++s) {
    // after a mesh change the ghost list is located and compiled while the
    // device runs the statistics, the dt reduction and stage A
    const bool split = !c->tables_valid;
    int rc = split ? tables_leaves(c) : WX_OK;
    if (!rc) rc = update_stats(c);
    if (rc) return rc;
    {
      ProfScope ps(c, KID_DTRED);      // timestep; time += dt
      rc = launch_dtreduce(c->dP.ptr, c->tLeaves, c->dStat.ptr, c->tOff, 1, c->dS.ptr, 1, c->stream);
    }
    if (rc) return fail(rc, "dtreduce launch failed");
    z += (double)c->nleaf() * NX * NY;
    rc = run_stage_a(c);               // tstep
    if (!rc && split) rc = tables_ops(c);
    if (!rc) rc = run_stages_bf(c);
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
  return prof_report(c, enable, ms, counts, n);
}

int wx_set_surrogate(wx_ctx *c, const wx_sur_model *m) {
  if (!c) return fail(WX_EARG, "null context");
  if (!m) {
    c->use_sur = false;
    return WX_OK;
  }
  if (c->p.cooling != 1 || c->neq < 5) return fail(WX_EARG, "the cooling surrogate replaces coolingh (cooling = 1, neq >= 5)");
  int rc = c->dSur.ensure(1);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(c->dSur.ptr, 0, sizeof(SurDev), c->stream));
  if ((rc = sur_upload(c->dSur.ptr, m, c->dP.ptr, c->stream))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->use_sur = true;
  return WX_OK;
}

int wx_sur_counters(wx_ctx *c, int32_t reset, int64_t *cells, int64_t *fallbacks) {
  if (!c) return fail(WX_EARG, "null context");
  if (!c->dSur.ptr) {
    if (cells) *cells = 0;
    if (fallbacks) *fallbacks = 0;
    return WX_OK;
  }
  return sur_counters(c->dSur.ptr, c->stream, reset, cells, fallbacks);
}

int wx_shape(wx_ctx *c, int32_t *neq, int32_t *nblocks, int32_t *nlevs) {
  if (!c) return fail(WX_EARG, "null context");
  if (neq) *neq = c->neq;
  if (nblocks) *nblocks = c->cap;
  if (nlevs) *nlevs = c->p.nlevs;
  return WX_OK;
}

int wx_info(wx_ctx *c, int32_t *nbmax, int32_t *nbleafs, int32_t *ninternal) {
  if (!c) return fail(WX_EARG, "null context");
  if (nbmax) *nbmax = c->nbmax;
  if (nbleafs) *nbleafs = c->nbleafs;
  if (ninternal) return list_length(c, ninternal);
  return WX_OK;
}

}  // extern "C"

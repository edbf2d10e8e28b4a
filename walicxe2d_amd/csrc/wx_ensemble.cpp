// wx_ensemble.cpp -- many independent AMR meshes ("instances") advanced
// together behind the wx_ens_* entry points of include/wx_capi.h.
//
// Every instance keeps its own host tree, ghost list, parameters, time and
// timestep (the reference's whole time loop, src/main.f90:69-106, run once
// per instance).  On the device all instances share one slot space: the
// block ID `id` of instance i lives in tile slot slotbase[i] + id, which
// also indexes its block record, leaf flags and ghost ops; instance i's leaf
// IDs are in the region starting at leafbase[i].  After a mesh change the
// leaf IDs of the instances whose mesh changed go up, with the ops and flags
// of the slots where they differ from what the device holds (a refinement
// changes those of a few blocks, while every later leaf moves in the list),
// and k_build_queue gathers the queue from the regions and the slots.
// Each stage is ONE launch over that queue (one workgroup per entry, or a
// persistent grid looping over it).  The step's launches (dt reduction and
// the three stages, the last of which also writes the flags for the host)
// can be captured in two hipGraphs (one per tile-buffer parity); as the
// kernels read the queue length from device memory and the buffers are
// allocated at full capacity up front, the graphs are re-captured only when
// the queue outgrows the span they cover.
// Per-instance host work (table compilation, admesh) runs on a pool of
// worker threads; wx_ens_create_groups splits the instances into groups
// (each of the above per group) stepped concurrently from host threads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "wx_runtime.h"
#include "wx_surrogate.h"

using namespace wx;

struct EnsGroup : Runtime {
  int n = 0, neq = 0;
  int ibase = 0;   // ensemble index of the group's first instance
  std::vector<Mesh> m;
  std::vector<int32_t> slotbase, leafbase;   // per instance
  size_t nslots = 0, nrec = 0;
  int ncu = 256;
  // device state
  DevBuf<double> U[2];
  int cur = 0;
  DevBuf<wx_params> dP;
  DevBuf<DevScalars> dS;
  DevBuf<QueueHdr> dQ;
  DevBuf<int32_t> dInstOff, dLeafBase, dSlotBase;
  DevBuf<int32_t> dRegIds;                // per-instance regions of leaf IDs
  DevBuf<LeafRec> dRec;                   // the compacted queue
  DevBuf<BlockRec> dBinfo;
  DevBuf<int32_t> dSlotFlag;              // per slot: the leaf flags
  DevBuf<int2> dSlotOpsII, dSlotOpsI;     // per slot: the ops of fills II and I
  DevBuf<int2> dOpsII, dOpsI;             // the queue's ops
  // no instance in the OUTFYB compat mode: fill I's ops equal fill II's for
  // every instance, so they are neither uploaded nor gathered, and the fill
  // reads fill II's (fixed by the parameters for the ensemble's life)
  bool opsI_same_all = false;
  // block records on the device per instance: IDs 0 .. binfo_dev[i] - 1 (an
  // ID's level and position never change once created, so after a mesh
  // change only the IDs above the old maximum go up)
  std::vector<size_t> binfo_dev;
  // per instance: the per-slot ops and flags the device holds, by local
  // block ID (local source IDs; {-1, -1} / -1 where nothing went up yet;
  // invalid while an upload is being prepared or after it failed), and the
  // IDs of the current upload's changed slots
  struct SlotShadow {
    std::vector<int2> ops2, ops1;
    std::vector<int32_t> flag;
    bool valid = false;
  };
  std::vector<SlotShadow> shadow;
  std::vector<std::vector<int32_t>> changed;
  const int2 *opsI() const { return opsI_same_all ? dOpsII.ptr : dOpsI.ptr; }
  DevBuf<double> dStat;
  DevBuf<int8_t> dFlag;
  // under DMC cooling: per slot, the pressure the last cooling left in cells
  // 0..17, carried into the next stage A (nullptr without DMC instances)
  DevBuf<double> dPc;
  // mapped pinned mirrors written by the fill's epilogue (or k_gather_flags)
  DevScalars *hS = nullptr, *hS_dev = nullptr;
  int8_t *hFlagQ = nullptr, *hFlagQ_dev = nullptr;
  // queue bookkeeping
  std::vector<int32_t> inst_off;   // n + 1 prefix sums of leaf counts
  int nleaf_total = 0;
  // host scratch
  std::vector<int> rcs;
  std::vector<std::string> errs;
  std::vector<std::vector<MeshOp>> iops;
  std::vector<MeshOp> ops;
  std::vector<LeafRec> dirty;
  std::unique_ptr<WorkerPool> pool;
  // launch shapes of the stage kernels (A, B, fillI): persistent workgroups
  // (0 = one workgroup per queue entry) and threads per workgroup (stage A
  // is fastest with 8 waves, the others with 4, on the jet ensembles)
  int grid[3] = {0, 0, 0};
  int bs[3] = {512, 256, 256};
  // queue entries the captured graphs launch workgroups for (the one-per-
  // entry kernels): recaptured when the queue outgrows it or shrinks to half
  int gspan = 0;
  // cooling surrogate for the coolingh instances, used when use_sur
  DevBuf<SurDev> dSur;
  bool use_sur = false;
  // graphs, one per tile-buffer parity
  bool use_graph = false;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};

  double *cur_u() { return U[cur].ptr; }
  double *other_u() { return U[cur ^ 1].ptr; }
  double *pc() { return dPc.ptr; }
  QueueRef queue() const { return QueueRef{dQ.ptr, dRec.ptr, (int32_t)dRec.n}; }
  size_t tile() const { return (size_t)neq * NC; }
};

namespace {

int collect_errors(EnsGroup *e, const std::vector<int> &which) {
  for (int i : which)
    if (e->rcs[i]) return fail(e->rcs[i], "instance %d: %s", e->ibase + i, e->errs[i].c_str());
  return WX_OK;
}

void drop_graphs(EnsGroup *e) {
  for (int k = 0; k < 2; ++k) {
    if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
    if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
    e->gexec[k] = nullptr;
    e->graph[k] = nullptr;
  }
}

// Workgroups of stage kernel k for the current queue, and whether they loop
// over it (persistent) or take one entry each.
struct Shape {
  int grid, persist;
};
Shape stage_grid(const EnsGroup *e, int k, bool for_graph) {
  if (e->grid[k] > 0) return {e->grid[k], 1};
  return {for_graph ? e->gspan : e->nleaf_total, 0};
}

// Queue span a graph is captured for: the queue length and an eighth more,
// so that small mesh changes replay the same graph.
int graph_span(const EnsGroup *e) {
  const size_t want = ((size_t)e->nleaf_total * 9 / 8 + 255) / 256 * 256;
  return (int)std::min(e->nrec, std::max<size_t>(want, 256));
}

// The slots of instance i whose ops (either fill's when `two`) or leaf
// flags differ from the device's copy: their IDs go to changed[i] and the
// shadow takes the new values.  The shadow stays invalid until the upload
// has been launched.
void slot_diff(EnsGroup *e, int i, bool two) {
  const Mesh &m = e->m[i];
  EnsGroup::SlotShadow &sh = e->shadow[i];
  if (!sh.valid) {   // nothing known of the device's copy: every leaf goes up
    sh.ops2.clear();
    sh.ops1.clear();
    sh.flag.clear();
  }
  sh.valid = false;
  const size_t need = (size_t)m.nbmax + 1;
  if (sh.flag.size() < need) {
    sh.ops2.resize(need * 8, int2{-1, -1});
    if (two) sh.ops1.resize(need * 8, int2{-1, -1});
    sh.flag.resize(need, -1);
  }
  std::vector<int32_t> &ch = e->changed[i];
  ch.clear();
  for (size_t l = 0; l < m.leaves.size(); ++l) {
    const int id = m.leaves[l].id;
    const int2 *o2 = &m.opsII[l * 8];
    int2 *s2 = &sh.ops2[(size_t)id * 8];
    bool diff = sh.flag[id] != m.leaves[l].flags || memcmp(o2, s2, 8 * sizeof(int2)) != 0;
    if (two) diff = diff || memcmp(&m.opsI[l * 8], &sh.ops1[(size_t)id * 8], 8 * sizeof(int2)) != 0;
    if (!diff) continue;
    memcpy(s2, o2, 8 * sizeof(int2));
    if (two) memcpy(&sh.ops1[(size_t)id * 8], &m.opsI[l * 8], 8 * sizeof(int2));
    sh.flag[id] = m.leaves[l].flags;
    ch.push_back(id);
  }
}

int upload_queue(EnsGroup *e, const std::vector<int> &stale, bool two);

// Recompiles the tables of instances whose mesh changed, uploads their leaf
// IDs, the changed slots' ops and flags, the new block records and the
// leaf-count prefix sums, and rebuilds the queue.  On failure the changed
// instances stay stale, with nothing assumed of the device's copy.
int ensure_queue(EnsGroup *e) {
  std::vector<int> stale;
  for (int i = 0; i < e->n; ++i)
    if (!e->m[i].tables_valid) stale.push_back(i);
  if (stale.empty()) return WX_OK;
  HostTimer ht(e, HID_TABLES);
  const bool two = !e->opsI_same_all;
  const int rc = upload_queue(e, stale, two);
  if (rc)
    for (int i : stale) {
      e->m[i].tables_valid = false;
      e->shadow[i].valid = false;
    }
  return rc;
}

int upload_queue(EnsGroup *e, const std::vector<int> &stale, bool two) {
  e->pool->run((int)stale.size(), [&](int k) {
    const int i = stale[k];
    e->rcs[i] = refresh_tables(&e->m[i]);
    if (e->rcs[i]) {
      e->errs[i] = last_error();
      e->shadow[i].valid = false;
      return;
    }
    slot_diff(e, i, two);
  });
  int rc = collect_errors(e, stale);
  if (rc) return rc;
  e->inst_off[0] = 0;
  for (int i = 0; i < e->n; ++i) e->inst_off[i + 1] = e->inst_off[i] + (int32_t)e->m[i].leaves.size();
  e->nleaf_total = e->inst_off[e->n];
  // block records from binfo_dev[i] on (every ID above the old maximum)
  auto bi_first = [&](int i) { return std::min(e->binfo_dev[i], e->m[i].binfo.size()); };
  std::vector<size_t> roff(stale.size() + 1, 0);   // first changed-slot record of each instance
  for (size_t k = 0; k < stale.size(); ++k) roff[k + 1] = roff[k] + e->changed[stale[k]].size();
  const size_t nchanged = roff.back();
  size_t total = al256(sizeof(QueueHdr)) + al256((size_t)(e->n + 1) * sizeof(int32_t)) +
                 (two ? 2 : 1) * al256(nchanged * sizeof(SlotOps));
  for (int i : stale)
    total += al256(e->m[i].leaves.size() * sizeof(int32_t)) +
             al256((e->m[i].binfo.size() - bi_first(i)) * sizeof(BlockRec));
  if ((rc = stage_begin(e, total))) return rc;
  SlotOps *r2 = (SlotOps *)stage_take(e, nchanged * sizeof(SlotOps));
  SlotOps *r1 = two ? (SlotOps *)stage_take(e, nchanged * sizeof(SlotOps)) : nullptr;
  if (!r2 || (two && !r1)) return fail(WX_EARG, "staging %zu slot records", nchanged);
  struct Chunk { char *ids, *bi; };
  std::vector<Chunk> chk(stale.size());
  for (size_t k = 0; k < stale.size(); ++k) {
    const Mesh &m = e->m[stale[k]];
    chk[k].ids = stage_take(e, m.leaves.size() * sizeof(int32_t));
    chk[k].bi = stage_take(e, (m.binfo.size() - bi_first(stale[k])) * sizeof(BlockRec));
    if (!chk[k].ids || !chk[k].bi) return fail(WX_EARG, "staging the regions of instance %d", e->ibase + stale[k]);
  }
  e->pool->run((int)stale.size(), [&](int k) {   // local IDs -> global slots
    const int i = stale[k];
    const Mesh &m = e->m[i];
    const int sb = e->slotbase[i];
    int32_t *ids = (int32_t *)chk[k].ids;
    for (size_t l = 0; l < m.leaves.size(); ++l) ids[l] = m.leaves[l].id;
    const EnsGroup::SlotShadow &sh = e->shadow[i];
    const std::vector<int32_t> &ch = e->changed[i];
    for (size_t c = 0; c < ch.size(); ++c) {
      const size_t id = (size_t)ch[c];
      SlotOps &o = r2[roff[k] + c];
      o.slot = sb + (int32_t)id;
      o.flag = sh.flag[id];
      for (int t = 0; t < 8; ++t) o.ops[t] = int2{sh.ops2[id * 8 + t].x + sb, sh.ops2[id * 8 + t].y};
      if (r1) {
        SlotOps &o1 = r1[roff[k] + c];
        o1.slot = o.slot;
        o1.flag = o.flag;
        for (int t = 0; t < 8; ++t) o1.ops[t] = int2{sh.ops1[id * 8 + t].x + sb, sh.ops1[id * 8 + t].y};
      }
    }
    const size_t b0 = bi_first(i);
    memcpy(chk[k].bi, m.binfo.data() + b0, (m.binfo.size() - b0) * sizeof(BlockRec));
  });
  if (nchanged) {   // read in place from the staging (stage_end guards it)
    rc = launch_slot_scatter((const SlotOps *)stage_dev(e, (const char *)r2), (int)nchanged, e->dSlotOpsII.ptr,
                             e->dSlotFlag.ptr, e->stream);
    if (!rc && two)
      rc = launch_slot_scatter((const SlotOps *)stage_dev(e, (const char *)r1), (int)nchanged, e->dSlotOpsI.ptr,
                               nullptr, e->stream);
    if (rc) return fail(rc, "slot scatter launch failed");
  }
  for (size_t k = 0; k < stale.size(); ++k) {
    const int i = stale[k];
    const Mesh &m = e->m[i];
    if ((rc = stage_push(e, e->dRegIds.ptr + e->leafbase[i], chk[k].ids, m.leaves.size() * sizeof(int32_t))))
      return rc;
    const size_t b0 = bi_first(i);
    if (m.binfo.size() > b0 &&
        (rc = stage_push(e, e->dBinfo.ptr + e->slotbase[i] + b0, chk[k].bi, (m.binfo.size() - b0) * sizeof(BlockRec))))
      return rc;
  }
  const QueueHdr qh{e->nleaf_total, e->n, {0, 0}};
  if ((rc = stage_copy(e, e->dQ.ptr, &qh, sizeof qh))) return rc;
  if ((rc = stage_copy(e, e->dInstOff.ptr, e->inst_off.data(), (size_t)(e->n + 1) * sizeof(int32_t)))) return rc;
  for (int i : stale) e->binfo_dev[i] = 0;   // unknown until the copies are launched
  if ((rc = stage_end(e))) return rc;
  for (int i : stale) {
    e->binfo_dev[i] = e->m[i].binfo.size();
    e->shadow[i].valid = true;
  }
  rc = launch_build_queue(e->dQ.ptr, e->dInstOff.ptr, e->dLeafBase.ptr, e->dSlotBase.ptr, e->dRegIds.ptr,
                          e->dBinfo.ptr, e->dSlotFlag.ptr, e->dSlotOpsII.ptr, two ? e->dSlotOpsI.ptr : nullptr,
                          e->dRec.ptr, e->dOpsII.ptr, two ? e->dOpsI.ptr : nullptr,
                          std::max(1, std::min((8 * e->nleaf_total + 255) / 256, 4096)), e->stream);
  if (rc) return fail(rc, "queue build launch failed");
  return WX_OK;
}

// Block statistics of new leaves (or of every leaf of an instance whose
// statistics are all stale), one launch for all instances.
int ensure_stats(EnsGroup *e) {
  int rc = ensure_queue(e);
  if (rc) return rc;
  e->dirty.clear();
  for (int i = 0; i < e->n; ++i) {
    Mesh &m = e->m[i];
    const int sb = e->slotbase[i];
    // the pad bit asks the statistics to start the carried DMC pressure
    // from u (a state from the host)
    const int init = m.pc_stale ? 1 : 0;
    auto add = [&](const LeafRec &r0) {
      LeafRec r = r0;
      r.id += sb;
      r.inst = i;
      r.pad = init;
      e->dirty.push_back(r);
    };
    if (m.stats_all_stale) {
      for (const LeafRec &r : m.leaves) add(r);
      m.stats_all_stale = false;
      m.pc_stale = false;
    } else {
      for (int id : m.new_ids)
        if (id >= 1 && id <= m.nbmax && m.lidx[id] >= 0) add(m.leaves[m.lidx[id]]);
    }
    m.new_ids.clear();
  }
  if (e->dirty.empty()) return WX_OK;
  const LeafRec *drec = (const LeafRec *)stage_inplace(e, e->dirty.data(), e->dirty.size() * sizeof(LeafRec));
  if (!drec) return WX_ENOMEM;
  {
    ProfScope ps(e, KID_STAT);
    rc = launch_blockstat(e->dP.ptr, e->neq, drec, (int)e->dirty.size(), e->cur_u(), e->dStat.ptr, e->dFlag.ptr,
                          e->pc(), 2, e->stream);
  }
  if (rc) return fail(rc, "blockstat launch failed");
  return stage_end(e);
}

// dt reduction (time += dt) and the three stages (the fill also writing the
// flags and scalars for the host), reading U[par] and writing U[par ^ 1];
// `prof` records kernel events.
int launch_step(EnsGroup *e, int par, bool for_graph) {
  const bool prof = !for_graph;
  double *Ucur = e->U[par].ptr, *Uoth = e->U[par ^ 1].ptr;
  const QueueRef q = e->queue();
  const Shape sh[3] = {stage_grid(e, 0, for_graph), stage_grid(e, 1, for_graph), stage_grid(e, 2, for_graph)};
  int rc;
  {
    ProfScope ps(e, prof ? KID_DTRED : -1);
    rc = launch_dtreduce(e->dP.ptr, e->dRec.ptr, e->dStat.ptr, e->dInstOff.ptr, e->n, e->dS.ptr, 1, e->stream);
  }
  if (!rc) {
    ProfScope ps(e, prof ? KID_STAGEA : -1);
    rc = launch_stageA(e->dP.ptr, e->neq, q, sh[0].grid, e->bs[0], sh[0].persist, Ucur, Uoth, e->pc(), e->dS.ptr,
                       e->stream);
  }
  if (!rc) {
    ProfScope ps(e, prof ? KID_STAGEB : -1);
    rc = launch_stageB(e->dP.ptr, e->neq, q, sh[1].grid, e->bs[1], sh[1].persist, e->dOpsII.ptr, e->dBinfo.ptr, Ucur, Uoth,
                       e->dS.ptr, e->stream);
  }
  if (!rc) {   // its epilogue also writes the flags (queue order) and the scalars for the host
    ProfScope ps(e, prof ? KID_FILLI : -1);
    const FlagOut fo{e->dInstOff.ptr, e->hFlagQ_dev, e->hS_dev, e->n};
    if (e->use_sur)
      rc = launch_fillI_cool_sur(e->dP.ptr, e->neq, q, sh[2].grid, e->bs[2], sh[2].persist, e->opsI(), e->dBinfo.ptr, Ucur,
                                 Uoth, e->dStat.ptr, e->dFlag.ptr, e->dS.ptr, e->dSur.ptr, fo, e->stream);
    else
      rc = launch_fillI_cool(e->dP.ptr, e->neq, q, sh[2].grid, e->bs[2], sh[2].persist, e->opsI(), e->dBinfo.ptr, Ucur,
                             Uoth, e->dStat.ptr, e->dFlag.ptr, e->dS.ptr, fo, e->pc(), e->stream);
  }
  if (rc) return fail(rc, "step launch failed: %s", hipGetErrorString(hipGetLastError()));
  return WX_OK;
}

int ensure_graphs(EnsGroup *e) {
  if (e->gexec[0] && e->gexec[1] && e->nleaf_total <= e->gspan && 2 * e->nleaf_total > e->gspan) return WX_OK;
  drop_graphs(e);
  e->gspan = graph_span(e);
  for (int par = 0; par < 2; ++par) {
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(e, par, true);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (ce != hipSuccess) return fail(WX_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
    e->graph[par] = g;
    HIPCHK(hipGraphInstantiate(&e->gexec[par], g, nullptr, nullptr, 0));
  }
  return WX_OK;
}

int check_scalars(EnsGroup *e) {
  for (int i = 0; i < e->n; ++i)
    if (e->hS[i].err)
      return fail(WX_EHLLC, "Error in hllc (instance %d, device flag %d): a Riemann problem had no valid star speed", i,
                  e->hS[i].err);
  return WX_OK;
}

// admesh of every instance from the flags the last gather wrote (queue order).
int admesh_all(EnsGroup *e, int *changed) {
  HIPCHK(hipStreamSynchronize(e->stream));
  HostTimer ht(e, HID_ADMESH);   // host work only (after the step's wait)
  int rc = check_scalars(e);
  if (rc) return rc;
  std::vector<int> all(e->n);
  for (int i = 0; i < e->n; ++i) all[i] = i;
  e->pool->run(e->n, [&](int i) {
    e->iops[i].clear();
    if (e->m[i].p.nlevs <= 1) { e->rcs[i] = 0; return; }
    e->rcs[i] = admesh_from_flags(&e->m[i], e->hFlagQ + e->inst_off[i], e->iops[i]);
    if (e->rcs[i]) e->errs[i] = last_error();
  });
  rc = collect_errors(e, all);
  if (rc) return rc;
  e->ops.clear();
  for (int i = 0; i < e->n; ++i) {
    const int sb = e->slotbase[i];
    const Mesh &m = e->m[i];
    for (MeshOp o : e->iops[i]) {
      o.level = m.LP(o.kind == 0 ? o.s1 : o.a, 1);   // the new leaves', for their statistics
      o.inst = i;
      o.a += sb; o.s1 += sb; o.s2 += sb; o.s3 += sb; o.s4 += sb;
      e->ops.push_back(o);
    }
  }
  if (changed) *changed = (int)e->ops.size();
  if (e->ops.empty()) return WX_OK;
  // the kernel reads the op list in place from the pinned staging
  const MeshOp *dops = (const MeshOp *)stage_inplace(e, e->ops.data(), e->ops.size() * sizeof(MeshOp));
  if (!dops) return WX_ENOMEM;
  {
    ProfScope ps(e, KID_MESH);
    rc = launch_meshops(e->dP.ptr, e->neq, dops, (int)e->ops.size(), e->cur_u(), e->dStat.ptr, e->dFlag.ptr,
                        e->pc(), e->stream);
  }
  if (rc) return fail(rc, "meshops launch failed");
  if ((rc = stage_end(e))) return rc;
  for (int i = 0; i < e->n; ++i)   // the kernel computed the new leaves' statistics
    if (!e->iops[i].empty()) e->m[i].new_ids.clear();
  return WX_OK;
}

int check_inst(EnsGroup *e, int inst) {
  if (!e) return fail(WX_EARG, "null ensemble");
  if (inst < 0 || inst >= e->n) return fail(WX_EARG, "instance %d outside [0, %d)", inst, e->n);
  return WX_OK;
}

}  // namespace

namespace {

int grp_destroy(EnsGroup *e);

int grp_create(const wx_params *p, int32_t n, int nthreads, EnsGroup **out) {
  EnsGroup *e = new EnsGroup();
  e->n = n;
  e->neq = p[0].neq;
  e->m.resize(n);
  e->slotbase.resize(n);
  e->leafbase.resize(n);
  e->inst_off.assign((size_t)n + 1, 0);
  e->rcs.assign(n, 0);
  e->errs.resize(n);
  e->iops.resize(n);
  e->binfo_dev.assign(n, 0);
  e->shadow.resize(n);
  e->changed.resize(n);
  e->opsI_same_all = true;
  for (int i = 0; i < n; ++i) e->opsI_same_all = e->opsI_same_all && !lb_outfyb(p[i]);
  for (int i = 0; i < n; ++i) {
    e->m[i].init(p[i]);
    e->slotbase[i] = (int32_t)e->nslots;
    e->leafbase[i] = (int32_t)e->nrec;
    e->nslots += (size_t)p[i].nblocks + 1;
    e->nrec += (size_t)p[i].nblocks;
  }
  if (e->nslots > (size_t)INT32_MAX / 8) { delete e; return fail(WX_EARG, "ensemble too large (%zu slots)", e->nslots); }
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || e->ncu < 1) e->ncu = 256;
  e->pool.reset(new WorkerPool(nthreads));
  int rc = e->create_stream();
  const size_t tiles = e->nslots * e->tile();
  for (int k = 0; k < 2 && !rc; ++k) rc = e->U[k].ensure(tiles);
  if (!rc) rc = e->dP.ensure(n);
  if (!rc) rc = e->dS.ensure(n);
  if (!rc) rc = e->dQ.ensure(1);
  if (!rc) rc = e->dInstOff.ensure((size_t)n + 1);
  if (!rc) rc = e->dLeafBase.ensure(n);
  if (!rc) rc = e->dSlotBase.ensure(n);
  if (!rc) rc = e->dRegIds.ensure(e->nrec);
  if (!rc) rc = e->dRec.ensure(e->nrec);
  if (!rc) rc = e->dBinfo.ensure(e->nslots);
  if (!rc) rc = e->dSlotFlag.ensure(e->nslots);
  if (!rc) rc = e->dSlotOpsII.ensure(e->nslots * 8);
  if (!rc && !e->opsI_same_all) rc = e->dSlotOpsI.ensure(e->nslots * 8);
  if (!rc) rc = e->dOpsII.ensure(e->nrec * 8);
  if (!rc && !e->opsI_same_all) rc = e->dOpsI.ensure(e->nrec * 8);
  if (!rc) rc = e->dStat.ensure(e->nslots * 2);
  if (!rc) rc = e->dFlag.ensure(e->nslots);
  bool dmc = false;
  for (int i = 0; i < n; ++i) dmc = dmc || p[i].cooling == 2;
  if (!rc && dmc) rc = e->dPc.ensure(e->nslots * NC);
  if (rc) { grp_destroy(e); return rc; }
  if (hipHostMalloc((void **)&e->hS, (size_t)n * sizeof(DevScalars), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&e->hFlagQ, e->nrec, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    grp_destroy(e);
    return fail(WX_ENOMEM, "pinned mirrors");
  }
  if (hipHostGetDevicePointer((void **)&e->hS_dev, e->hS, 0) != hipSuccess ||
      hipHostGetDevicePointer((void **)&e->hFlagQ_dev, e->hFlagQ, 0) != hipSuccess) {
    grp_destroy(e);
    return fail(WX_EHIP, "mapped pointers of the pinned mirrors");
  }
  memset(e->hS, 0, (size_t)n * sizeof(DevScalars));
  memset(e->hFlagQ, 0, e->nrec);
  for (int k = 0; k < 2; ++k) (void)hipMemsetAsync(e->U[k].ptr, 0, tiles * sizeof(double), e->stream);
  (void)hipMemsetAsync(e->dS.ptr, 0, (size_t)n * sizeof(DevScalars), e->stream);
  (void)hipMemsetAsync(e->dQ.ptr, 0, sizeof(QueueHdr), e->stream);
  (void)hipMemsetAsync(e->dInstOff.ptr, 0, ((size_t)n + 1) * sizeof(int32_t), e->stream);
  (void)hipMemsetAsync(e->dBinfo.ptr, 0, e->nslots * sizeof(BlockRec), e->stream);
  (void)hipMemsetAsync(e->dSlotFlag.ptr, 0, e->nslots * sizeof(int32_t), e->stream);
  if (e->dPc.ptr) (void)hipMemsetAsync(e->dPc.ptr, 0, e->dPc.n * sizeof(double), e->stream);
  (void)hipMemcpyAsync(e->dP.ptr, p, (size_t)n * sizeof(wx_params), hipMemcpyHostToDevice, e->stream);
  (void)hipMemcpyAsync(e->dLeafBase.ptr, e->leafbase.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  (void)hipMemcpyAsync(e->dSlotBase.ptr, e->slotbase.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  if (hipStreamSynchronize(e->stream) != hipSuccess) { grp_destroy(e); return fail(WX_EHIP, "init sync"); }
  *out = e;
  return WX_OK;
}

int grp_destroy(EnsGroup *e) {
  if (!e) return WX_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graphs(e);
  for (auto &b : e->U) b.release();
  e->dP.release(); e->dS.release(); e->dQ.release(); e->dInstOff.release(); e->dLeafBase.release();
  e->dSlotBase.release(); e->dRegIds.release(); e->dRec.release(); e->dBinfo.release(); e->dSlotFlag.release();
  e->dSlotOpsII.release(); e->dSlotOpsI.release(); e->dOpsII.release(); e->dOpsI.release();
  e->dStat.release(); e->dFlag.release(); e->dSur.release(); e->dPc.release();
  e->Runtime::destroy();
  if (e->hS) (void)hipHostFree(e->hS);
  if (e->hFlagQ) (void)hipHostFree(e->hFlagQ);
  delete e;
  return WX_OK;
}

int grp_basegrid(EnsGroup *e, int32_t inst, const double *uroots, int32_t nroots) {
  int rc = check_inst(e, inst);
  if (rc) return rc;
  Mesh &m = e->m[inst];
  if (!uroots || m.p.nbrootx * m.p.nbrooty != nroots)
    return fail(WX_EARG, "instance %d: nroots %d != %d x %d", e->ibase + inst, nroots, m.p.nbrootx, m.p.nbrooty);
  if ((rc = basegrid_tree(&m))) return rc;
  m.pc_stale = true;
  e->binfo_dev[inst] = 0;
  e->shadow[inst].valid = false;
  const size_t bytes = (size_t)(nroots + 1) * e->tile() * sizeof(double);
  HIPCHK(hipMemcpyAsync(e->cur_u() + (size_t)e->slotbase[inst] * e->tile(), uroots, bytes, hipMemcpyHostToDevice,
                        e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return WX_OK;
}

int grp_admesh(EnsGroup *e, int32_t *changed) {
  if (!e) return fail(WX_EARG, "null ensemble");
  int rc = ensure_stats(e);
  if (rc) return rc;
  rc = launch_gather_flags(e->queue(), e->dInstOff.ptr, std::max(1, std::min((e->nleaf_total + 255) / 256, 1024)),
                           e->dFlag.ptr,
                           e->hFlagQ_dev, e->dS.ptr, e->hS_dev, e->n, e->stream);
  if (rc) return fail(rc, "flag gather launch failed");
  int ch = 0;
  rc = admesh_all(e, &ch);
  if (changed) *changed = ch;
  return rc;
}

int grp_step(EnsGroup *e, int32_t nsteps, int32_t do_admesh, double *zone_updates) {
  if (!e || nsteps < 0) return fail(WX_EARG, "bad argument");
  for (int i = 0; i < e->n; ++i)
    if (e->m[i].nbleafs <= 0)
      return fail(WX_EARG, "instance %d has no mesh (call wx_ens_basegrid first)", e->ibase + i);
  double z = 0.0;
  int rc;
  for (int s = 0; s < nsteps; ++s) {
    if ((rc = ensure_stats(e))) return rc;
    if (e->use_graph && !e->prof) {
      if ((rc = ensure_graphs(e))) return rc;
      HIPCHK(hipGraphLaunch(e->gexec[e->cur], e->stream));
    } else {
      if ((rc = launch_step(e, e->cur, false))) return rc;
    }
    e->cur ^= 1;
    z += (double)e->nleaf_total * NX * NY;
    if (do_admesh) {
      if ((rc = admesh_all(e, nullptr))) return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  if ((rc = check_scalars(e))) return rc;
  if (zone_updates) *zone_updates = z;
  return WX_OK;
}

int grp_instance(EnsGroup *e, int32_t inst, double *time, double *dt, int32_t *nbmax, int32_t *nbleafs) {
  int rc = check_inst(e, inst);
  if (rc) return rc;
  DevScalars s{};
  HIPCHK(hipMemcpyAsync(&s, e->dS.ptr + inst, sizeof s, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (time) *time = s.time;
  if (dt) *dt = s.dt;
  if (nbmax) *nbmax = e->m[inst].nbmax;
  if (nbleafs) *nbleafs = e->m[inst].nbleafs;
  return WX_OK;
}

int grp_download(EnsGroup *e, int32_t inst, double *u) {
  int rc = check_inst(e, inst);
  if (rc) return rc;
  if (!u) return fail(WX_EARG, "null buffer");
  const size_t bytes = (size_t)(e->m[inst].nbmax + 1) * e->tile() * sizeof(double);
  HIPCHK(hipMemcpyAsync(u, e->cur_u() + (size_t)e->slotbase[inst] * e->tile(), bytes, hipMemcpyDeviceToHost,
                        e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  return WX_OK;
}

int grp_get_tree(EnsGroup *e, int32_t inst, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld,
                    int32_t *nbmax, int32_t *nbleafs) {
  int rc = check_inst(e, inst);
  if (rc) return rc;
  return get_tree(&e->m[inst], lp, leaf, icoord, ld, nbmax, nbleafs);
}

int grp_set_graph(EnsGroup *e, int32_t enable) {
  if (!e) return fail(WX_EARG, "null ensemble");
  e->use_graph = enable != 0;
  if (!e->use_graph) drop_graphs(e);
  return WX_OK;
}

int grp_set_launch(EnsGroup *e, int32_t kernel, int32_t grid, int32_t threads) {
  if (!e) return fail(WX_EARG, "null ensemble");
  if (kernel < 0 || kernel > 2 || grid < 0) return fail(WX_EARG, "kernel %d / grid %d", kernel, grid);
  if (threads != 256 && threads != 384 && threads != 512) return fail(WX_EARG, "threads %d (256, 384 or 512)", threads);
  e->grid[kernel] = grid;
  e->bs[kernel] = threads;
  drop_graphs(e);
  return WX_OK;
}

// Launch-shape autotuner: times each stage kernel over `reps` launches for
// every workgroup size (4, 6, 8 waves) and persistent grid (1, 2, 4, 8, 16
// workgroups per CU, and one workgroup per queue entry) on the current
// meshes, and keeps the fastest.  Stage A and the fill are idempotent given
// Uⁿ; stage B updates in place, so it runs on a scratch copy of Uⁿ.  The
// statistics the fill's epilogue overwrites are recomputed afterwards, so
// tuning leaves the results of later steps unchanged.
int grp_tune(EnsGroup *e, int32_t reps, int32_t *grids, int32_t *threads, double *us) {
  if (!e || reps < 1) return fail(WX_EARG, "bad argument");
  int rc = ensure_stats(e);
  if (rc) return rc;
  if (e->nleaf_total == 0) return WX_OK;
  HIPCHK(hipStreamSynchronize(e->stream));
  std::vector<DevScalars> saved(e->n);
  HIPCHK(hipMemcpy(saved.data(), e->dS.ptr, (size_t)e->n * sizeof(DevScalars), hipMemcpyDeviceToHost));
  const size_t tiles = e->nslots * e->tile();
  DevBuf<double> scratch;
  if ((rc = scratch.ensure(tiles))) return rc;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  const QueueRef q = e->queue();
  std::vector<int> cands;
  for (int w : {1, 2, 4, 8, 16})
    if (e->ncu * w < e->nleaf_total) cands.push_back(e->ncu * w);
  cands.push_back(e->nleaf_total);
  for (int k = 0; k < 3 && !rc; ++k) {
    double best = 1e30;
    int bestg = e->nleaf_total, bestb = 256;
    for (int bsz : {256, 384, 512}) {
      for (int g : cands) {
        const int pers = g != e->nleaf_total;
        float tot = 0.f;
        for (int r = 0; r <= reps && !rc; ++r) {   // r = 0 warms up
          if (k == 1 && hipMemcpyAsync(scratch.ptr, e->cur_u(), tiles * sizeof(double), hipMemcpyDeviceToDevice,
                                       e->stream) != hipSuccess) rc = fail(WX_EHIP, "tune copy");
          (void)hipEventRecord(a, e->stream);
          if (k == 0)
            rc = launch_stageA(e->dP.ptr, e->neq, q, g, bsz, pers, e->cur_u(), e->other_u(), e->pc(), e->dS.ptr,
                               e->stream);
          else if (k == 1)
            rc = launch_stageB(e->dP.ptr, e->neq, q, g, bsz, pers, e->dOpsII.ptr, e->dBinfo.ptr, scratch.ptr, e->other_u(),
                               e->dS.ptr, e->stream);
          else if (e->use_sur)
            rc = launch_fillI_cool_sur(e->dP.ptr, e->neq, q, g, bsz, pers, e->opsI(), e->dBinfo.ptr, e->cur_u(),
                                       scratch.ptr, e->dStat.ptr, e->dFlag.ptr, e->dS.ptr, e->dSur.ptr, FlagOut{},
                                       e->stream);
          else
            // (a trial: the carried pressure is not written)
            rc = launch_fillI_cool(e->dP.ptr, e->neq, q, g, bsz, pers, e->opsI(), e->dBinfo.ptr, e->cur_u(), scratch.ptr,
                                   e->dStat.ptr, e->dFlag.ptr, e->dS.ptr, FlagOut{}, nullptr, e->stream);
          (void)hipEventRecord(b, e->stream);
          (void)hipEventSynchronize(b);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, a, b);
          if (r > 0) tot += ms;
        }
        const double t = 1e3 * tot / reps;
        if (t < best) { best = t; bestg = g; bestb = bsz; }
      }
    }
    e->grid[k] = bestg == e->nleaf_total ? 0 : bestg;
    e->bs[k] = bestb;
    if (grids) grids[k] = bestg;
    if (threads) threads[k] = bestb;
    if (us) us[k] = best;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIPCHK(hipStreamSynchronize(e->stream));
  scratch.release();
  HIPCHK(hipMemcpy(e->dS.ptr, saved.data(), (size_t)e->n * sizeof(DevScalars), hipMemcpyHostToDevice));
  if (e->use_sur) HIPCHK(hipMemsetAsync(&e->dSur.ptr->cells, 0, 2 * sizeof(unsigned long long), e->stream));
  for (auto &m : e->m) m.stats_all_stale = true;
  drop_graphs(e);
  return rc;
}

int grp_set_surrogate(EnsGroup *e, const wx_sur_model *m) {
  if (!e) return fail(WX_EARG, "null ensemble");
  drop_graphs(e);
  if (!m) {
    e->use_sur = false;
    return WX_OK;
  }
  if (e->neq < 5) return fail(WX_EARG, "the cooling surrogate replaces coolingh (neq >= 5)");
  int rc = e->dSur.ensure(1);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(e->dSur.ptr, 0, sizeof(SurDev), e->stream));
  // the loss table depends only on the routine's constants, the same for
  // every instance (instance 0's parameters fill it)
  if ((rc = sur_upload(e->dSur.ptr, m, e->dP.ptr, e->stream))) return rc;
  HIPCHK(hipStreamSynchronize(e->stream));
  e->use_sur = true;
  return WX_OK;
}

int grp_sur_counters(EnsGroup *e, int32_t reset, int64_t *cells, int64_t *fallbacks) {
  if (!e) return fail(WX_EARG, "null ensemble");
  if (!e->dSur.ptr) {
    if (cells) *cells = 0;
    if (fallbacks) *fallbacks = 0;
    return WX_OK;
  }
  return sur_counters(e->dSur.ptr, e->stream, reset, cells, fallbacks);
}

int grp_profile(EnsGroup *e, int32_t enable, double *ms, int64_t *counts, int32_t n) {
  if (!e) return fail(WX_EARG, "null ensemble");
  return prof_report(e, enable, ms, counts, n);
}

int grp_info(EnsGroup *e, int32_t *ninst, int32_t *nleaf_total, int32_t *grids, int32_t *threads) {
  if (!e) return fail(WX_EARG, "null ensemble");
  if (ninst) *ninst = e->n;
  if (nleaf_total) *nleaf_total = e->nleaf_total;
  if (grids) for (int k = 0; k < 3; ++k) grids[k] = e->grid[k];
  if (threads) for (int k = 0; k < 3; ++k) threads[k] = e->bs[k];
  return WX_OK;
}

}  // namespace

// =====================================================================
// The ensemble: its instances split into groups of consecutive instances,
// each group with its own stream, queue, graphs and worker threads.  With
// more than one group, wx_ens_step drives every group from a host thread of
// its own, so that one group's host work (admesh, list compilation) runs
// while the others' kernels occupy the GPU, and the groups' kernels overlap.
struct wx_ens {
  int n = 0, dev = 0;
  std::vector<EnsGroup *> g;
  std::vector<int> gi, li;   // group of each instance and its index there
};

namespace {

int check_ens(wx_ens *e, int inst = 0) {
  if (!e) return fail(WX_EARG, "null ensemble");
  if (inst < 0 || inst >= e->n) return fail(WX_EARG, "instance %d outside [0, %d)", inst, e->n);
  return WX_OK;
}

// Runs f(group) for every group, concurrently when there are several (each
// on a host thread with the ensemble's device current); returns the first
// failure with its message.
template <class F>
int each_group(wx_ens *e, bool concurrent, F f) {
  const int G = (int)e->g.size();
  if (!concurrent || G == 1) {
    for (int k = 0; k < G; ++k) {
      int rc = f(k);
      if (rc) return rc;
    }
    return WX_OK;
  }
  std::vector<int> rcs(G, 0);
  std::vector<std::string> msg(G);
  std::vector<std::thread> th;
  for (int k = 1; k < G; ++k)
    th.emplace_back([&, k] {
      (void)hipSetDevice(e->dev);
      rcs[k] = f(k);
      if (rcs[k]) msg[k] = last_error();
    });
  rcs[0] = f(0);
  if (rcs[0]) msg[0] = last_error();
  for (auto &t : th) t.join();
  for (int k = 0; k < G; ++k)
    if (rcs[k]) return fail(rcs[k], "%s", msg[k].c_str());
  return WX_OK;
}

}  // namespace

extern "C" {

int wx_ens_create(const wx_params *p, int32_t n, wx_ens **out) { return wx_ens_create_groups(p, n, 1, out); }

int wx_ens_create_groups(const wx_params *p, int32_t n, int32_t groups, wx_ens **out) {
  if (!p || !out || n < 1) return fail(WX_EARG, "bad argument");
  if (groups < 1 || groups > n) return fail(WX_EARG, "groups %d outside [1, %d]", groups, n);
  std::string why;
  for (int i = 0; i < n; ++i) {
    if (!params_ok(p + i, why)) return fail(WX_EARG, "instance %d: bad parameters: %s", i, why.c_str());
    if (p[i].neq != p[0].neq) return fail(WX_EARG, "instance %d: neq %d differs from instance 0 (%d)", i, p[i].neq, p[0].neq);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WX_EHIP, "no HIP device");
  int nthreads = 8;
  if (const char *s = getenv("WX_HOST_THREADS")) nthreads = std::max(1, atoi(s));
  else if (const char *s2 = getenv("OMP_NUM_THREADS")) nthreads = std::max(1, atoi(s2));
  nthreads = std::min(nthreads, std::max(1, (int)std::thread::hardware_concurrency()));
  nthreads = std::min(nthreads, 32);
  wx_ens *e = new wx_ens();
  e->n = n;
  (void)hipGetDevice(&e->dev);
  e->gi.resize(n);
  e->li.resize(n);
  int first = 0;
  for (int k = 0; k < groups; ++k) {
    const int cnt = n / groups + (k < n % groups ? 1 : 0);
    EnsGroup *grp = nullptr;
    int rc = grp_create(p + first, cnt, std::max(1, nthreads / groups), &grp);
    if (rc) { wx_ens_destroy(e); return rc; }
    grp->ibase = first;
    e->g.push_back(grp);
    for (int i = 0; i < cnt; ++i) { e->gi[first + i] = k; e->li[first + i] = i; }
    first += cnt;
  }
  *out = e;
  return WX_OK;
}

int wx_ens_destroy(wx_ens *e) {
  if (!e) return WX_OK;
  for (EnsGroup *g : e->g) grp_destroy(g);
  delete e;
  return WX_OK;
}

int wx_ens_basegrid(wx_ens *e, int32_t inst, const double *uroots, int32_t nroots) {
  int rc = check_ens(e, inst);
  return rc ? rc : grp_basegrid(e->g[e->gi[inst]], e->li[inst], uroots, nroots);
}

int wx_ens_admesh(wx_ens *e, int32_t *changed) {
  int rc = check_ens(e);
  if (rc) return rc;
  int total = 0;
  rc = each_group(e, false, [&](int k) {
    int ch = 0;
    int r = grp_admesh(e->g[k], &ch);
    total += ch;
    return r;
  });
  if (changed) *changed = total;
  return rc;
}

int wx_ens_step(wx_ens *e, int32_t nsteps, int32_t do_admesh, double *zone_updates) {
  int rc = check_ens(e);
  if (rc) return rc;
  // concurrent groups, except while kernel events are recorded: then each
  // kernel's time is its own, not shared with another group's kernels
  std::vector<double> z(e->g.size(), 0.0);
  rc = each_group(e, !e->g[0]->prof, [&](int k) { return grp_step(e->g[k], nsteps, do_admesh, &z[k]); });
  if (rc) return rc;
  double zt = 0.0;
  for (double v : z) zt += v;
  if (zone_updates) *zone_updates = zt;
  return WX_OK;
}

int wx_ens_instance(wx_ens *e, int32_t inst, double *time, double *dt, int32_t *nbmax, int32_t *nbleafs) {
  int rc = check_ens(e, inst);
  return rc ? rc : grp_instance(e->g[e->gi[inst]], e->li[inst], time, dt, nbmax, nbleafs);
}

int wx_ens_download(wx_ens *e, int32_t inst, double *u) {
  int rc = check_ens(e, inst);
  return rc ? rc : grp_download(e->g[e->gi[inst]], e->li[inst], u);
}

int wx_ens_get_tree(wx_ens *e, int32_t inst, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld,
                    int32_t *nbmax, int32_t *nbleafs) {
  int rc = check_ens(e, inst);
  return rc ? rc : grp_get_tree(e->g[e->gi[inst]], e->li[inst], lp, leaf, icoord, ld, nbmax, nbleafs);
}

int wx_ens_set_graph(wx_ens *e, int32_t enable) {
  int rc = check_ens(e);
  return rc ? rc : each_group(e, false, [&](int k) { return grp_set_graph(e->g[k], enable); });
}

int wx_ens_set_launch(wx_ens *e, int32_t kernel, int32_t grid, int32_t threads) {
  int rc = check_ens(e);
  return rc ? rc : each_group(e, false, [&](int k) { return grp_set_launch(e->g[k], kernel, grid, threads); });
}

// Every group is tuned on its own meshes (one after the other, so that the
// timings see an otherwise idle GPU); the values returned are group 0's.
int wx_ens_tune(wx_ens *e, int32_t reps, int32_t *grids, int32_t *threads, double *us) {
  int rc = check_ens(e);
  return rc ? rc : each_group(e, false, [&](int k) {
    return k == 0 ? grp_tune(e->g[k], reps, grids, threads, us) : grp_tune(e->g[k], reps, nullptr, nullptr, nullptr);
  });
}

int wx_ens_set_surrogate(wx_ens *e, const wx_sur_model *m) {
  int rc = check_ens(e);
  return rc ? rc : each_group(e, false, [&](int k) { return grp_set_surrogate(e->g[k], m); });
}

int wx_ens_sur_counters(wx_ens *e, int32_t reset, int64_t *cells, int64_t *fallbacks) {
  int rc = check_ens(e);
  if (rc) return rc;
  int64_t c = 0, f = 0;
  rc = each_group(e, false, [&](int k) {
    int64_t ck = 0, fk = 0;
    int r = grp_sur_counters(e->g[k], reset, &ck, &fk);
    c += ck;
    f += fk;
    return r;
  });
  if (cells) *cells = c;
  if (fallbacks) *fallbacks = f;
  return rc;
}

// Accumulated over the groups (with concurrent groups, kernel times add up
// to more than the wall time).
int wx_ens_profile(wx_ens *e, int32_t enable, double *ms, int64_t *counts, int32_t n) {
  int rc = check_ens(e);
  if (rc) return rc;
  if (n < 0) return fail(WX_EARG, "n %d", n);
  n = std::min(n, (int32_t)NPROF);
  double mst[NPROF] = {};
  int64_t ct[NPROF] = {};
  rc = each_group(e, false, [&](int k) {
    double m1[NPROF] = {};
    int64_t c1[NPROF] = {};
    int r = grp_profile(e->g[k], enable, m1, c1, n);
    for (int j = 0; j < n; ++j) { mst[j] += m1[j]; ct[j] += c1[j]; }
    return r;
  });
  for (int j = 0; j < n; ++j) {
    if (ms) ms[j] = mst[j];
    if (counts) counts[j] = ct[j];
  }
  return rc;
}

int wx_ens_info(wx_ens *e, int32_t *ninst, int32_t *nleaf_total, int32_t *grids, int32_t *threads) {
  int rc = check_ens(e);
  if (rc) return rc;
  int32_t total = 0;
  for (EnsGroup *g : e->g) total += g->nleaf_total;
  if (ninst) *ninst = e->n;
  if (nleaf_total) *nleaf_total = total;
  return grp_info(e->g[0], nullptr, nullptr, grids, threads);
}

int wx_ens_shape(wx_ens *e, int32_t inst, int32_t *neq, int32_t *nblocks, int32_t *nlevs) {
  int rc = check_ens(e, inst);
  if (rc) return rc;
  const Mesh &m = e->g[e->gi[inst]]->m[e->li[inst]];
  if (neq) *neq = m.p.neq;
  if (nblocks) *nblocks = m.cap;
  if (nlevs) *nlevs = m.p.nlevs;
  return WX_OK;
}

int wx_ens_groups(wx_ens *e, int32_t *groups) {
  int rc = check_ens(e);
  if (rc) return rc;
  if (groups) *groups = (int32_t)e->g.size();
  return WX_OK;
}

}  // extern "C"

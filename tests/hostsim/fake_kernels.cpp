// Kernel launch stand-ins for the host-only build of wx_host.cpp.  The
// statistics kernels write refinement flags from a table the test sets
// (wxsim_set_flags), so the host's admesh can be driven with chosen flags;
// the mesh-op launches record the ops they were given.  Every launch goes
// through hipsim::launch with the pointers it was given, so a launch on a
// capturing stream becomes a graph node (hip/hip_runtime.h).  TEST
// SCAFFOLDING.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>
#include "../../walicxe2d_amd/csrc/wx_surrogate.h"
#include "lb_cpu.h"

// Flags by slot; with an ensemble split in groups, one table per group,
// the group of an instance read from its jet[11] (a field the jet-less
// parameters leave unused).
static std::vector<signed char> g_flags;
static std::vector<std::vector<signed char>> g_gflags;
static std::vector<wx::MeshOp> g_ops;
static std::mutex g_mu;

extern "C" void wxsim_set_flags(const signed char *f, int n) { g_flags.assign(f, f + n); }
extern "C" void wxsim_set_group_flags(int group, const signed char *f, int n) {
  if ((int)g_gflags.size() <= group) g_gflags.resize(group + 1);
  g_gflags[group].assign(f, f + n);
}
extern "C" int wxsim_ops(int *out, int cap) {
  std::lock_guard<std::mutex> g(g_mu);
  int n = (int)g_ops.size();
  for (int k = 0; k < n && k < cap; ++k) {
    out[6 * k] = g_ops[k].kind; out[6 * k + 1] = g_ops[k].a; out[6 * k + 2] = g_ops[k].s1;
    out[6 * k + 3] = g_ops[k].s2; out[6 * k + 4] = g_ops[k].s3; out[6 * k + 5] = g_ops[k].s4;
  }
  g_ops.clear();
  return n;
}

namespace wx {
static void write_flags(const wx_params *P, const LeafRec *leaves, int n, signed char *flag, double *stat) {
  for (int l = 0; l < n; ++l) {
    int id = leaves[l].id;
    const std::vector<signed char> &tab = g_gflags.empty() ? g_flags : g_gflags[(int)P[leaves[l].inst].jet[11]];
    flag[id] = id < (int)tab.size() ? tab[id] : 0;
    stat[2 * id] = 1.0;
    stat[2 * id + 1] = 1e30;
  }
}
int launch_build_queue(const QueueHdr *hdr, const int32_t *inst_off, const int32_t *leafbase,
                       const int32_t *slotbase, const int32_t *reg_ids, const BlockRec *binfo,
                       const int32_t *slot_flag, const int2 *slot_ops2, const int2 *slot_ops1, LeafRec *rec,
                       int2 *ops2, int2 *ops1, int, hipStream_t st) {
  hipsim::launch(st, {hdr, inst_off, leafbase, slotbase, reg_ids, binfo, slot_flag, slot_ops2, slot_ops1, rec, ops2,
                      ops1},
                 [=] {
    for (int l = 0; l < hdr->nleaf; ++l) {
      int i = 0;
      while (i + 1 < hdr->ninst && inst_off[i + 1] <= l) ++i;
      const int lp = l - inst_off[i];
      const size_t slot = (size_t)slotbase[i] + (size_t)reg_ids[(size_t)leafbase[i] + lp];
      const BlockRec b = binfo[slot];
      rec[l] = LeafRec{(int32_t)slot, b.level, b.icx, b.icy, slot_flag[slot], lp, i, 0};
      for (int k = 0; k < 8; ++k) {
        ops2[(size_t)l * 8 + k] = slot_ops2[slot * 8 + k];
        if (ops1) ops1[(size_t)l * 8 + k] = slot_ops1[slot * 8 + k];
      }
    }
  });
  return 0;
}
int launch_slot_scatter(const SlotOps *src, int n, int2 *ops, int32_t *flag, hipStream_t st) {
  hipsim::launch(st, {src, ops, flag}, [=] {
    for (int r = 0; r < n; ++r) {
      for (int k = 0; k < 8; ++k) ops[(size_t)src[r].slot * 8 + k] = src[r].ops[k];
      if (flag) flag[src[r].slot] = src[r].flag;
    }
  });
  return 0;
}
// A one-entry-per-workgroup launch must cover the whole queue.
static void check_shape(QueueRef q, int grid, int bs, int persist) {
  if (grid <= 0 || (bs != 256 && bs != 384 && bs != 512) || (!persist && grid < q.hdr->nleaf)) {
    fprintf(stderr, "bad stage launch: grid %d threads %d persist %d for %d queue entries\n", grid, bs, persist,
            q.hdr->nleaf);
    abort();
  }
}
int launch_stageA(const wx_params *P, int, QueueRef q, int grid, int bs, int persist, const double *U, double *Uh,
                  const double *Pc, DevScalars *S, hipStream_t st) {
  hipsim::launch(st, {P, q.hdr, q.rec, U, Uh, Pc, S}, [=] { check_shape(q, grid, bs, persist); });
  return 0;
}
int launch_stageB(const wx_params *P, int, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                  const BlockRec *binfo, double *U, double *Uh, DevScalars *S, hipStream_t st) {
  hipsim::launch(st, {P, q.hdr, q.rec, ops, binfo, U, Uh, S}, [=] { check_shape(q, grid, bs, persist); });
  return 0;
}
int launch_fillI_cool(const wx_params *P, int, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                      const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                      DevScalars *S, FlagOut fo, double *Pc, hipStream_t st) {
  hipsim::launch(st, {P, q.hdr, q.rec, ops, binfo, Uin, Uout, stat, flag, S, fo.inst_off, fo.flagq, fo.Sout, Pc}, [=] {
    check_shape(q, grid, bs, persist);
    write_flags(P, q.rec, q.hdr->nleaf, (signed char *)flag, stat);
    if (fo.flagq) {
      for (int l = 0; l < q.hdr->nleaf; ++l) fo.flagq[fo.inst_off[q.rec[l].inst] + q.rec[l].lidx] = flag[q.rec[l].id];
      for (int i = 0; i < fo.ninst; ++i) fo.Sout[i] = S[i];
    }
  });
  return 0;
}
int launch_blockstat(const wx_params *P, int, const LeafRec *leaves, int nleaf, const double *U, double *stat,
                     int8_t *flag, double *Pc, int, hipStream_t st) {
  hipsim::launch(st, {P, leaves, U, stat, flag, Pc}, [=] { write_flags(P, leaves, nleaf, (signed char *)flag, stat); });
  return 0;
}
int launch_dtreduce(const wx_params *P, const LeafRec *rec, const double *stat, const int32_t *inst_off, int ninst,
                    DevScalars *S, int adv, hipStream_t st) {
  hipsim::launch(st, {P, rec, stat, inst_off, S}, [=] {
    for (int i = 0; i < ninst; ++i) {
      S[i].dt = 1e-3;
      if (adv) S[i].time += S[i].dt;
    }
  });
  return 0;
}
int launch_gather_flags(QueueRef q, const int32_t *inst_off, int, const int8_t *flag, int8_t *flagq, const DevScalars *S,
                        DevScalars *Sout, int ninst, hipStream_t st) {
  hipsim::launch(st, {q.hdr, q.rec, inst_off, flag, flagq, S, Sout}, [=] {
    for (int l = 0; l < q.hdr->nleaf; ++l) flagq[inst_off[q.rec[l].inst] + q.rec[l].lidx] = flag[q.rec[l].id];
    if (Sout) for (int i = 0; i < ninst; ++i) Sout[i] = S[i];
  });
  return 0;
}
int launch_meshops(const wx_params *P, int, const MeshOp *ops, int nops, double *U, double *stat, int8_t *flag,
                   double *Pc, hipStream_t st) {
  hipsim::launch(st, {P, ops, U, stat, flag, Pc}, [=] {
    std::lock_guard<std::mutex> g(g_mu);
    for (int k = 0; k < nops; ++k) {
      g_ops.push_back(ops[k]);
      if (!stat) continue;
      LeafRec r[4] = {};
      const int ids[4] = {ops[k].s1, ops[k].s2, ops[k].s3, ops[k].s4};
      const int n = ops[k].kind == 0 ? 4 : 1;
      for (int s = 0; s < n; ++s) {
        r[s].id = ops[k].kind == 0 ? ids[s] : ops[k].a;
        r[s].level = ops[k].level;
        r[s].inst = ops[k].inst;
      }
      write_flags(P, r, n, (signed char *)flag, stat);
    }
  });
  return 0;
}
int launch_transpose(int, int, const double *, double *, int, int, hipStream_t) { return 0; }
int launch_h2d(const CopyDesc *d, int n, hipStream_t st) {
  hipsim::launch(st, {d}, [=] {
    for (int k = 0; k < n; ++k) memmove(d[k].dst, d[k].src, d[k].bytes);
  });
  return 0;
}
int launch_fillI_cool_sur(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                          const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                          DevScalars *S, SurDev *, FlagOut fo, hipStream_t st) {
  return launch_fillI_cool(P, neq, q, grid, bs, persist, ops, binfo, Uin, Uout, stat, flag, S, fo, nullptr, st);
}
int sur_upload(SurDev *, const wx_sur_model *, const wx_params *, hipStream_t) { return 0; }
int sur_counters(SurDev *, hipStream_t, int, int64_t *c, int64_t *f) {
  if (c) *c = 0;
  if (f) *f = 0;
  return 0;
}
int launch_calcprim(const wx_params *, int, int, const double *, double *, hipStream_t) { return 0; }
int launch_build_tables(const wx_params &p, const TreeRow *T, int nbmax, int nleaf_host, LbInfo *info, LbPart *part,
                        LbOut out, int, DevScalars *S, hipStream_t st) {
  const wx_params pp = p;
  hipsim::launch(st, {T, info, part, out.nlist, out.ops2, out.ops1, S},
                 [=] { lb_cpu_build(pp, T, nbmax, nleaf_host, info, out, S); });
  return 0;
}
}  // namespace wx

// Kernel launch stand-ins for the host-only build of wx_host.cpp.  The
// statistics kernels write refinement flags from a table the test sets
// (wxsim_set_flags), so the host's admesh can be driven with chosen flags;
// the mesh-op launches record the ops they were given.  TEST SCAFFOLDING.
#include <cstring>
#include <vector>
#include "../../walicxe2d_amd/csrc/wx_device.h"

static std::vector<signed char> g_flags;
static std::vector<wx::MeshOp> g_ops;

extern "C" void wxsim_set_flags(const signed char *f, int n) { g_flags.assign(f, f + n); }
extern "C" int wxsim_ops(int *out, int cap) {
  int n = (int)g_ops.size();
  for (int k = 0; k < n && k < cap; ++k) {
    out[6 * k] = g_ops[k].kind; out[6 * k + 1] = g_ops[k].a; out[6 * k + 2] = g_ops[k].s1;
    out[6 * k + 3] = g_ops[k].s2; out[6 * k + 4] = g_ops[k].s3; out[6 * k + 5] = g_ops[k].s4;
  }
  g_ops.clear();
  return n;
}

namespace wx {
static void write_flags(const LeafRec *leaves, int n, signed char *flag, double *stat) {
  for (int l = 0; l < n; ++l) {
    int id = leaves[l].id;
    flag[id] = id < (int)g_flags.size() ? g_flags[id] : 0;
    stat[2 * id] = 1.0;
    stat[2 * id + 1] = 1e30;
  }
}
int launch_stageA(const wx_params *, int, const LeafRec *, int, const double *, double *, DevScalars *, void *) { return 0; }
int launch_stageB(const wx_params *, int, const LeafRec *, int, const int2 *, const BlockRec *, double *,
                  const double *, DevScalars *, void *) { return 0; }
int launch_fillI_cool(const wx_params *, int, const LeafRec *leaves, int nleaf, const int2 *, const BlockRec *,
                      const double *, double *, double *stat, int8_t *flag, DevScalars *, void *) {
  write_flags(leaves, nleaf, (signed char *)flag, stat);
  return 0;
}
int launch_blockstat(const wx_params *, int, const LeafRec *leaves, int nleaf, const double *, double *stat,
                     int8_t *flag, void *) {
  write_flags(leaves, nleaf, (signed char *)flag, stat);
  return 0;
}
int launch_dtreduce(const wx_params *, const LeafRec *, const double *, int, DevScalars *S, int adv, void *) {
  S->dt = 1e-3;
  if (adv) S->time += S->dt;
  return 0;
}
int launch_meshops(const wx_params *, int, const MeshOp *ops, int nops, double *, void *) {
  for (int k = 0; k < nops; ++k) g_ops.push_back(ops[k]);
  return 0;
}
int launch_transpose(int, int, const double *, double *, int, int, void *) { return 0; }
int launch_calcprim(const wx_params *, int, int, const double *, double *, void *) { return 0; }
}  // namespace wx

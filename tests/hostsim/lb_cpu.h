// The device table builder (walicxe2d_amd/csrc/wx_lbuild.hip) run on the
// CPU: the same per-element functions (wx_lbuild.h) in the kernels' order,
// with the workgroup scans replaced by one sequential exclusive scan (their
// combined result).  Used by the host-only runtime build as the stand-in of
// launch_build_tables and by tables_check against the host compilation.
// TEST SCAFFOLDING.
#pragma once
#include "../../walicxe2d_amd/csrc/wx_lbuild.h"

namespace wx {

inline void lb_cpu_build(const wx_params &p, const TreeRow *T, int nbmax, int nleaf_host, LbInfo *info, LbOut out,
                         DevScalars *S) {
  const bool outfyb = lb_outfyb(p) != 0, otherb = p.otherb != 0;
  const lb::RowTree t{T};
  // k_lb_place1 (or k_lb_count + k_lb_scan + k_lb_place)
  int nl = 0, ne = 0;
  for (int id = 0; id <= nbmax; ++id) {
    int cl = 0, ce = 0;
    if (id >= 1 && T[id].leaf) {
      cl = 1;
      ce = lb::block_count(t, nbmax, id);
      if (ce < 0) {
        S->mesh_err = LBE_POINTER;
        ce = 0;
      }
    }
    info[id] = LbInfo{cl ? nl : -1, ne, ce, 0};
    nl += cl;
    ne += ce;
  }
  if (nl != nleaf_host) S->mesh_err = LBE_COUNT;
  *out.nlist = ne;
  // k_lb_ops (with the OUTFYB writes of fill I)
  const int bc[4] = {p.bc[0], p.bc[1], p.bc[2], p.bc[3]};
  for (int id = 1; id <= nbmax; ++id) {
    if (!T[id].leaf) continue;
    const int l = info[id].lidx;
    if (l < 0 || l >= out.leaf_cap) {
      S->mesh_err = LBE_COUNT;
      continue;
    }
    for (int s = 0; s < 4; ++s) {
      int2 o2[2], o1[2];
      int e1[2];
      const int rc = lb::side_ops(t, nbmax, info, bc[s], outfyb, otherb, id, s, o2, o1, e1);
      if (rc) {
        S->mesh_err = rc;
        continue;
      }
      const size_t k = (size_t)l * 8 + 2 * s;
      out.ops2[k] = o2[0];
      out.ops2[k + 1] = o2[1];
      if (outfyb) {
        if (s == 2)
          for (int h = 0; h < 2; ++h) (void)lb::late(t, nbmax, info, *out.nlist, otherb, id, e1[h], &o1[h]);
        out.ops1[k] = o1[0];
        out.ops1[k + 1] = o1[1];
      }
    }
  }
}

}  // namespace wx

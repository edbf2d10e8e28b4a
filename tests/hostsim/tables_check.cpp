// The product's ghost-list compilation (compile_ops in
// walicxe2d_amd/csrc/wx_mesh.cpp: fill II compiled once, fill I derived from
// it plus the OUTFYB writes) against a direct restatement of the list's
// sequential semantics: every entry applied in list order to both fills,
// the last writer of a segment winning (src/boundaryII.f90:35-514,
// src/boundaryI.f90:34-396, the OUTFYB write to block e + 1 at
// src/boundaryI.f90:87), imposition flags from the first and last list
// position of each block.  Random meshes (seeded refinement rounds through
// the product's admesh) for every combination of jet source, OUTFYB compat
// mode and bottom/left boundary kind.  The device table builder
// (walicxe2d_amd/csrc/wx_lbuild.hip, run on the CPU by lb_cpu.h) is checked
// against the host compilation on the same meshes, from a copy of the tree
// kept up to date only through the rows the tree code marks as written (as
// the runtime uploads them).  TEST SCAFFOLDING.
#include <climits>
#include <cstdio>
#include <random>
#include <vector>

#include "../../walicxe2d_amd/csrc/wx_runtime.h"
#include "lb_cpu.h"

using namespace wx;

namespace {

struct Tables {
  std::vector<int2> o2, o1;
  std::vector<int> flags;
};

int restated(const Mesh &m, Tables &t) {
  const int cap = m.nbmax;
  const int nleaf = (int)m.leaves.size();
  const size_t nseg = (size_t)nleaf * 8;
  t.o2.assign(nseg, int2{0, GK_NONE});
  t.o1.assign(nseg, int2{0, GK_NONE});
  std::vector<int> e2(nseg, -1), e1(nseg, -1), first(cap + 1, INT_MAX), last(cap + 1, -1);
  auto put = [&](std::vector<int2> &T, std::vector<int> &E, int dst, int side, int half, int kind, int src, int off,
                 int e) {
    if (dst < 1 || dst > cap || m.lidx[dst] < 0) return;
    const size_t k = (size_t)m.lidx[dst] * 8 + side * 2 + half;
    T[k] = int2{src, kind | (off << GF_OFF_SHIFT)};
    E[k] = e;
  };
  auto both = [&](int dst, int side, int kind, int src, int off, int e) {
    for (int h = 0; h < 2; ++h) {
      put(t.o2, e2, dst, side, h, kind, src, off, e);
      put(t.o1, e1, dst, side, h, kind, src, off, e);
    }
  };
  const bool bug = (m.p.compat & 2) && m.p.bc[2] == 0;
  for (int e = 0; e < m.ninternal; ++e) {
    const int nbs = m.bounds[(size_t)e * 3], n1 = m.bounds[(size_t)e * 3 + 1], code = m.bounds[(size_t)e * 3 + 2];
    if (code < 0) {
      const int side = -code - 1;
      const int kind = m.p.bc[side] ? GK_BOXREF : GK_BOXOUT;
      for (int h = 0; h < 2; ++h) put(t.o2, e2, nbs, side, h, kind, nbs, 0, e);
      if (side == 2 && bug)
        for (int h = 0; h < 2; ++h) put(t.o1, e1, e + 1, 2, h, GK_BOXOUT, e + 1, 0, e);
      else
        for (int h = 0; h < 2; ++h) put(t.o1, e1, nbs, side, h, kind, nbs, 0, e);
    } else if (code % 3 == 1) {   // same level: 1 L, 4 R, 7 B, 10 T
      const int side = (code - 1) / 3;
      both(nbs, side, GK_COPY, n1, 0, e);
      both(n1, side ^ 1, GK_COPY, nbs, 0, e);
    } else {                      // finer neighbour
      const int side = (code - 2) / 3, h = (code - 2) % 3;
      put(t.o2, e2, nbs, side, h, GK_RESTRICT, n1, 0, e);
      put(t.o1, e1, nbs, side, h, GK_RESTRICT, n1, 0, e);
      both(n1, side ^ 1, GK_PROLONG, nbs, 8 * h, e);
    }
    first[nbs] = std::min(first[nbs], e);
    last[nbs] = std::max(last[nbs], e);
  }
  t.flags.assign(nleaf, 0);
  if (m.p.otherb)
    for (int l = 0; l < nleaf; ++l) {
      const int dst = m.leaves[l].id;
      for (int s = 0; s < 8; ++s) {
        const size_t k = (size_t)l * 8 + s;
        for (int w = 0; w < 2; ++w) {
          int2 &op = w ? t.o1[k] : t.o2[k];
          const int e = w ? e1[k] : e2[k];
          if ((op.y & 15) == GK_NONE) continue;
          if (first[op.x] < e) op.y |= GF_SRC_IMPOSED;
          if (last[dst] >= e) op.y |= GF_DST_IMPOSED;
        }
      }
      t.flags[l] = last[dst] >= 0 ? 1 : 0;
    }
  return 0;
}

// The builder's copy of the tree brought up to date from the rows the tree
// code marked (every row up to nbmax after a new tree), as upload_tree does.
void update_copy(Mesh &m, std::vector<TreeRow> &dev) {
  auto put = [&](int nb) {
    TreeRow &r = dev[nb];
    for (int k = 0; k < 8; ++k) r.lp[k] = m.lp[(size_t)nb * 8 + k];
    r.leaf = m.leaf[nb];
    r.icx = m.icoord[(size_t)nb * 2];
    r.icy = m.icoord[(size_t)nb * 2 + 1];
    r.pad = 0;
  };
  if (m.allDirty)
    for (int nb = 1; nb <= m.nbmax; ++nb) put(nb);
  else
    for (int nb : m.dirtyRows) put(nb);
  for (int nb : m.dirtyRows) m.rowDirty[nb] = 0;
  m.dirtyRows.clear();
  m.allDirty = false;
}

// Builder output against the host compilation of the same tree; 0 if equal.
int check_builder(const Mesh &m, const std::vector<TreeRow> &dev, const char *tag) {
  for (int nb = 1; nb <= m.nbmax; ++nb) {
    bool same = dev[nb].leaf == m.leaf[nb] && dev[nb].icx == m.icoord[(size_t)nb * 2] &&
                dev[nb].icy == m.icoord[(size_t)nb * 2 + 1];
    for (int k = 0; k < 8; ++k) same = same && dev[nb].lp[k] == m.lp[(size_t)nb * 8 + k];
    if (!same) { fprintf(stderr, "%s: tree row %d missed by the written-row marks\n", tag, nb); return 1; }
  }
  const int nleaf = (int)m.leaves.size(), cap = nleaf + 3;
  std::vector<LbInfo> info((size_t)m.nbmax + 1);
  std::vector<int2> o2((size_t)cap * 8, int2{-7, -7}), o1((size_t)cap * 8, int2{-7, -7});
  int32_t nlist = -1;
  DevScalars S{};
  const LbOut out{&nlist, o2.data(), o1.data(), cap};
  lb_cpu_build(m.p, dev.data(), m.nbmax, nleaf, info.data(), out, &S);
  if (S.mesh_err) { fprintf(stderr, "%s: builder error %d\n", tag, S.mesh_err); return 1; }
  if (nlist != m.ninternal) { fprintf(stderr, "%s: list length %d vs %d\n", tag, nlist, m.ninternal); return 1; }
  for (int nb = 0; nb <= m.nbmax; ++nb)
    if (info[nb].lidx != m.lidx[nb]) { fprintf(stderr, "%s: leaf position of %d differs\n", tag, nb); return 1; }
  const bool two = lb_outfyb(m.p) != 0;
  if (two == m.opsI_same) { fprintf(stderr, "%s: fill I table presence differs\n", tag); return 1; }
  for (size_t k = 0; k < (size_t)nleaf * 8; ++k) {
    const int2 d1 = two ? o1[k] : o2[k];
    if (o2[k].x != m.opsII[k].x || o2[k].y != m.opsII[k].y || d1.x != m.opsI[k].x || d1.y != m.opsI[k].y) {
      fprintf(stderr, "%s: builder segment %zu of leaf %d: II (%d,%d)/(%d,%d) I (%d,%d)/(%d,%d)\n", tag, k % 8,
              m.leaves[k / 8].id, o2[k].x, o2[k].y, m.opsII[k].x, m.opsII[k].y, d1.x, d1.y, m.opsI[k].x, m.opsI[k].y);
      return 1;
    }
  }
  return 0;
}

}  // namespace

int main() {
  int cases = 0, outfyb_cases = 0;
  for (int variant = 0; variant < 16; ++variant)
    for (int seed = 1; seed <= 5; ++seed) {
      Mesh m;
      wx_params p{};
      p.neq = 5; p.nlevs = 3 + seed % 3; p.nbrootx = 3 + seed; p.nbrooty = 2 + (seed * 7) % 4; p.nblocks = 30000;
      p.cooling = 0;
      p.otherb = variant & 1;
      p.compat = (variant & 2) ? 3 : 1;
      p.bc[0] = (variant & 8) ? 1 : 0; p.bc[1] = 0; p.bc[2] = (variant & 4) ? 1 : 0; p.bc[3] = 1;
      m.init(p);
      basegrid_tree(&m);
      std::vector<TreeRow> dev((size_t)p.nblocks + 1, TreeRow{});
      std::mt19937 rng(1000 * variant + seed);
      for (int round = 0; round < 4; ++round) {
        if (refresh_tables(&m)) { fprintf(stderr, "refresh failed\n"); return 3; }
        char tag[96];
        snprintf(tag, sizeof tag, "variant %d seed %d round %d", variant, seed, round);
        update_copy(m, dev);
        if (check_builder(m, dev, tag)) return 1;
        std::vector<int8_t> fl(m.leaves.size(), 0);
        for (auto &f : fl) f = (rng() % 4 == 0) ? 1 : ((rng() % 3 == 0) ? 2 : 0);
        std::vector<MeshOp> ops;
        if (admesh_from_flags(&m, fl.data(), ops)) { fprintf(stderr, "admesh failed\n"); return 3; }
        if (refresh_tables(&m)) { fprintf(stderr, "refresh failed\n"); return 3; }
        {   // the leaf records' imposition flags as compile_leaves sets them alone
          std::vector<LeafRec> full = m.leaves;
          compile_leaves(&m);
          for (size_t l = 0; l < full.size(); ++l)
            if (memcmp(&full[l], &m.leaves[l], sizeof(LeafRec))) { fprintf(stderr, "compile_leaves record %zu differs\n", l); return 1; }
          m.leaves = full;
        }
        Tables t;
        restated(m, t);
        for (size_t k = 0; k < t.o2.size(); ++k)
          if (t.o2[k].x != m.opsII[k].x || t.o2[k].y != m.opsII[k].y || t.o1[k].x != m.opsI[k].x ||
              t.o1[k].y != m.opsI[k].y) {
            fprintf(stderr, "variant %d seed %d round %d: segment %zu of leaf %d differs: II (%d,%d)/(%d,%d) I (%d,%d)/(%d,%d)\n",
                    variant, seed, round, k % 8, m.leaves[k / 8].id, m.opsII[k].x, m.opsII[k].y, t.o2[k].x, t.o2[k].y,
                    m.opsI[k].x, m.opsI[k].y, t.o1[k].x, t.o1[k].y);
            return 1;
          }
        for (size_t l = 0; l < m.leaves.size(); ++l)
          if (m.leaves[l].flags != t.flags[l]) { fprintf(stderr, "leaf flag %zu differs\n", l); return 1; }
        bool differ = false;
        for (size_t k = 0; k < t.o2.size() && !differ; ++k) differ = t.o1[k].x != t.o2[k].x || t.o1[k].y != t.o2[k].y;
        if (differ != !m.opsI_same) { fprintf(stderr, "opsI_same flag wrong\n"); return 1; }
        outfyb_cases += differ;
        update_copy(m, dev);
        if (check_builder(m, dev, tag)) return 1;
        ++cases;
      }
    }
  printf("tables_check: %d meshes (%d with fill I != fill II) -- tables identical, device builder identical\n", cases,
         outfyb_cases);
  return 0;
}

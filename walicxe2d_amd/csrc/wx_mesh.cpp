// wx_mesh.cpp -- the host tree of one AMR mesh: the reference's ghost list
// (locatebounds), its compilation into per-leaf gather ops, and the admesh
// tree logic with the pointer surgery of updatelpup/updatelpdown.  Pure host
// code: no HIP calls.
#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wx_lbuild.h"
#include "wx_runtime.h"

namespace wx {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

const std::string &last_error() { return g_err; }

void Mesh::init(const wx_params &pp) {
  p = pp;
  cap = pp.nblocks;
  neq = pp.neq;
  lp.assign((size_t)(cap + 1) * 8, 0);
  leaf.assign((size_t)cap + 1, 0);
  icoord.assign((size_t)(cap + 1) * 2, 0);
  up.assign((size_t)cap + 1, 0);
  dn.assign((size_t)cap + 1, 0);
  rowDirty.assign((size_t)cap + 1, 0);
  dirtyRows.clear();
  allDirty = true;
  nbmax = nbleafs = ninternal = 0;
  tables_valid = false;
  bounds_external = false;
  stats_all_stale = true;
}

// ---------------------------------------------------------------------
// src/locatebounds.f90:29-540, MPIP forms, one rank.
void locatebounds(Mesh *c) {
  const int nbmax = c->nbmax;
  std::vector<uint8_t> &bnd = c->bndDone;   // face already in the list
  bnd.assign((size_t)(nbmax + 1) * 4, 0);
  std::vector<int32_t> &out = c->bounds;
  out.resize((size_t)c->nbleafs * 4 * 3 + 3);   // 4 entries per leaf; add() grows it beyond
  size_t n = 0;
  const int32_t *lp = c->lp.data(), *leaf = c->leaf.data();
  auto LP = [lp](int nb, int k) { return lp[(size_t)nb * 8 + k - 1]; };
  auto add = [&](int a, int b, int code) {
    if (n + 3 > out.size()) out.resize(out.size() * 2 + 12);
    out[n] = a;
    out[n + 1] = b;
    out[n + 2] = code;
    n += 3;
  };
  static const int opp[4] = {1, 0, 3, 2};
  for (int nb = 1; nb <= nbmax; ++nb) {
    if (!leaf[nb]) continue;
    for (int f = 0; f < 4; ++f) {
      if (bnd[(size_t)nb * 4 + f]) continue;
      const int nbr = LP(nb, 5 + f);
      if (nbr == -1) {
        bnd[(size_t)nb * 4 + f] = 1;
        add(nb, 0, -(f + 1));
        continue;
      }
      if (LP(nbr, 1) < LP(nb, 1)) continue;   // coarser side owns the face
      if (leaf[nbr]) {
        add(nb, nbr, 1 + 3 * f);
        bnd[(size_t)nb * 4 + f] = 1;
        bnd[(size_t)nbr * 4 + opp[f]] = 1;
      } else {
        bnd[(size_t)nb * 4 + f] = 1;
        int n1, n2;
        switch (f) {
          case 0: n1 = LP(LP(nbr, 3), 6); n2 = LP(n1, 8); break;   // :130-131
          case 1: n1 = LP(nbr, 3); n2 = LP(n1, 8); break;          // :237-238
          case 2: n1 = LP(LP(nbr, 3), 8); n2 = LP(n1, 6); break;   // :345-346
          default: n1 = LP(nbr, 3); n2 = LP(n1, 6); break;         // :452-453
        }
        add(nb, n1, 2 + 3 * f);
        add(nb, n2, 3 + 3 * f);
      }
    }
  }
  out.resize(n);
  c->ninternal = (int)(n / 3);
}

// ---------------------------------------------------------------------
// The host tree through the accessor of the list functions (wx_lbuild.h).
struct MeshTree {
  const int32_t *lp_, *leaf_;
  int lp(int nb, int k) const { return lp_[(size_t)nb * 8 + k]; }
  int leaf(int nb) const { return leaf_[nb]; }
};

// The leaf records (block-ID order, with the jet imposition flag of the
// list the tree gives), the leaf position of every block ID and the block
// records, from the tree.
int compile_leaves(Mesh *c) {
  const int cap = c->nbmax;   // IDs above nbmax hold no block
  c->lidx.assign((size_t)cap + 1, -1);
  c->binfo.resize((size_t)cap + 1);
  const int32_t *lp = c->lp.data(), *leaf = c->leaf.data(), *ic = c->icoord.data();
  int *lidx = c->lidx.data();
  BlockRec *bi = c->binfo.data();
  bi[0] = BlockRec{0, 0, 0, 0};
  int nleaf = 0;
  for (int nb = 1; nb <= cap; ++nb) {
    bi[nb] = BlockRec{lp[(size_t)nb * 8], ic[(size_t)nb * 2], ic[(size_t)nb * 2 + 1], 0};
    if (leaf[nb]) lidx[nb] = nleaf++;
  }
  c->leaves.resize((size_t)nleaf);
  LeafRec *rec = c->leaves.data();
  const bool otherb = c->p.otherb != 0;
  const MeshTree tree{lp, leaf};
  for (int nb = 1; nb <= cap; ++nb) {
    const int l = lidx[nb];
    if (l < 0) continue;
    LeafRec r{};
    r.id = nb;
    r.level = bi[nb].level;
    r.icx = bi[nb].icx;
    r.icy = bi[nb].icy;
    r.lidx = l;
    // impose_user_bc runs on the block in the fills when it makes list
    // entries (compile_ops sets the same flag from the list itself)
    if (otherb) r.flags = lb::block_count(tree, cap, nb) > 0 ? 1 : 0;
    if (r.level < 1 || r.level > c->p.nlevs) return fail(WX_EARG, "block %d has level %d", nb, r.level);
    rec[l] = r;
  }
  return WX_OK;
}

// Compile the ordered list into per-segment gather ops (last writer wins,
// jet-imposition order folded into two flags).  Ops are written in place:
// op.x = source block, op.y = kind | offset << GF_OFF_SHIFT, and the list
// position of the writer is kept beside each op for the imposition flags.
// Needs compile_leaves' records of the same tree; sets their imposition
// flags.  All scratch lives in the Mesh and is reused from call to call.
int compile_ops(Mesh *c) {
  const int cap = c->nbmax;
  const int nleaf = (int)c->leaves.size();
  const size_t nseg = (size_t)nleaf * 8;
  const bool otherb = c->p.otherb != 0;
  const bool outfyb_bug = lb_outfyb(c->p) != 0;
  // Fills I and II receive the same writes except for the OUTFYB entries of
  // boundaryI (src/boundaryI.f90:87: a bottom box entry writes block e + 1
  // instead of its own block): fill II is compiled from the list without
  // those entries, fill I starts as its copy, and then each of those entries
  // is applied where it is the last writer (list position of every
  // segment's writer kept in segE2 / segE1).
  const bool track = otherb || outfyb_bug;
  c->opsII.assign(nseg, int2{0, GK_NONE});
  c->opsI.resize(nseg);
  if (track) {
    c->segE2.assign(nseg, -1);
    c->segE1.resize(nseg);
  }
  if (otherb) {
    c->ifirst.assign((size_t)cap + 1, INT_MAX);
    c->ilast.assign((size_t)cap + 1, -1);
  }
  int2 *o2 = c->opsII.data(), *o1 = c->opsI.data();
  int *e2 = track ? c->segE2.data() : nullptr, *e1 = track ? c->segE1.data() : nullptr;
  const int *lidx = c->lidx.data();
  auto seg = [&](int dst, int side, int half) -> long {   // -1: not a leaf (its data is never read again)
    if (dst < 1 || dst > cap || lidx[dst] < 0) return -1;
    return (long)lidx[dst] * 8 + side * 2 + half;
  };
  auto put = [&](int dst, int side, int half, int kind, int src, int off, int e) {
    const long k = seg(dst, side, half);
    if (k < 0) return;
    o2[k] = int2{src, kind | (off << GF_OFF_SHIFT)};
    if (e2) e2[k] = e;
  };
  auto both = [&](int dst, int side, int kind, int src, int off, int e) {
    put(dst, side, 0, kind, src, off, e);
    put(dst, side, 1, kind, src, off, e);
  };
  std::vector<int> &late = c->segLate;   // the OUTFYB entries, in list order
  late.clear();
  for (int e = 0; e < c->ninternal; ++e) {
    const int nbs = c->bounds[(size_t)e * 3], n1 = c->bounds[(size_t)e * 3 + 1], code = c->bounds[(size_t)e * 3 + 2];
    if (nbs < 1 || nbs > c->nbmax) return fail(WX_EARG, "bound %d: block %d out of range", e + 1, nbs);
    if (code < 0) {
      const int side = -code - 1;
      if (side > 3) return fail(WX_EARG, "bound %d: code %d", e + 1, code);
      if (side == 2 && outfyb_bug) {
        late.push_back(e);
      } else {
        const int kind = c->p.bc[side] ? GK_BOXREF : GK_BOXOUT;
        both(nbs, side, kind, nbs, 0, e);   // boundaryII.f90:38-141, boundaryI.f90:38-107
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
          put(nbs, side, h, GK_RESTRICT, n1, 0, e);
          both(n1, fine_side, GK_PROLONG, nbs, 8 * h, e);
          break;
        }
        default: return fail(WX_EARG, "bound %d: unknown code %d", e + 1, code);
      }
    }
    if (otherb) {   // impose_user_bc(nbs) after each entry
      c->ifirst[nbs] = std::min(c->ifirst[nbs], e);
      c->ilast[nbs] = std::max(c->ilast[nbs], e);
    }
  }
  const int *ifirst = otherb ? c->ifirst.data() : nullptr, *ilast = otherb ? c->ilast.data() : nullptr;
  // imposition flags of the writer of segment k of the leaf with ID dst
  auto impose = [&](int2 &op, int e, int dst) {
    if (ifirst[op.x] < e) op.y |= GF_SRC_IMPOSED;
    if (ilast[dst] >= e) op.y |= GF_DST_IMPOSED;
  };
  if (otherb)
    for (int l = 0; l < nleaf; ++l) {
      const int dst = c->leaves[l].id;
      for (int s = 0; s < 8; ++s) {
        const size_t k = (size_t)l * 8 + s;
        if ((o2[k].y & 15) != GK_NONE) impose(o2[k], e2[k], dst);
      }
      c->leaves[l].flags = ilast[dst] >= 0 ? 1 : 0;
    }
  memcpy(o1, o2, nseg * sizeof(int2));
  if (track) memcpy(e1, e2, nseg * sizeof(int));
  for (int e : late) {
    const int nbs = c->bounds[(size_t)e * 3];
    const int kind = c->p.bc[2] ? GK_BOXREF : GK_BOXOUT;
    for (int h = 0; h < 2; ++h) {
      const long k2 = seg(nbs, 2, h);             // fill II: the block's own bottom ghosts
      if (k2 >= 0 && e > e2[k2]) {
        o2[k2] = int2{nbs, kind};
        e2[k2] = e;
        if (otherb) impose(o2[k2], e, nbs);
      }
      const int t = e + 1;                        // fill I: block e + 1 (src/boundaryI.f90:87)
      const long k1 = seg(t, 2, h);
      if (k1 >= 0 && e > e1[k1]) {
        o1[k1] = int2{t, GK_BOXOUT};
        e1[k1] = e;
        if (otherb) impose(o1[k1], e, t);
      }
    }
  }
  c->opsI_same = late.empty();
  // every segment of fill II is written by the list (boundaryI may leave
  // segments to the carried values)
  for (size_t k = 0; k < nseg; ++k)
    if ((o2[k].y & 15) == GK_NONE)
      return fail(WX_EARG, "boundaryII: block %d ghost segment %d not covered by the list",
                  c->leaves[k / 8].id, (int)(k % 8));
  return WX_OK;
}

// ---------------------------------------------------------------------
// Host tree logic: src/admesh.f90:4-111 (MPIP forms, one rank) with the
// pointer surgery of src/updatelpup.f90 and src/updatelpdown.f90.
static void updatelpup(Mesh *c, int nb, int s1, int s2, int s3, int s4, std::vector<uint8_t> &up,
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
  c->touch(nb);
  up[nb] = 0; dn[nb] = 0;
  dn[s1] = dn[s2] = dn[s3] = dn[s4] = 0;
  const int nx0 = c->icoord[(size_t)nb * 2] * 2, ny0 = c->icoord[(size_t)nb * 2 + 1] * 2;
  const int sons[4] = {s1, s2, s3, s4};
  for (int s = 0; s < 4; ++s) {   // src/refine.f90:227-261
    c->icoord[(size_t)sons[s] * 2] = nx0 + ((s & 1) ? NX : 0);
    c->icoord[(size_t)sons[s] * 2 + 1] = ny0 + ((s & 2) ? NY : 0);
  }
}

static void updatelpdown(Mesh *c, int dad, int s1, int s2, int s3, int s4) {
  c->leaf[dad] = 1;
  c->leaf[s1] = c->leaf[s2] = c->leaf[s3] = c->leaf[s4] = 0;
  c->touch(dad); c->touch(s1); c->touch(s2); c->touch(s3); c->touch(s4);
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
static int admesh_tree(Mesh *c, std::vector<uint8_t> &up, std::vector<uint8_t> &dn, std::vector<MeshOp> &ops,
                std::vector<int> &upl, std::vector<int> &dnl) {
  const int nlevs = c->p.nlevs;
  int nbmaxp = c->nbmax;
  const int nbmaxp1 = c->nbmax;
  // critneighup: closure of "refine my coarser neighbours", finest level first
  std::vector<std::vector<int>> &byl = c->admByLevel;   // scratch kept between calls
  byl.resize((size_t)nlevs + 2);
  for (auto &v : byl) v.clear();
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
      ops.push_back(MeshOp{0, nb, s1, s2, s3, s4, 0, 0});
      updatelpup(c, nb, s1, s2, s3, s4, up, dn);
    }
  }
  // critneighdown, src/critneighdown.f90:14-50: candidates (first sons that
  // keep their flag) by level, each level in ascending ID order.  The
  // coarsening below changes neither levels, sibling numbers nor flags, so
  // the candidates of every level are known before it starts.
  std::vector<std::vector<int>> &cand = c->admCand;
  cand.resize((size_t)nlevs + 2);
  for (auto &v : cand) v.clear();
  {
    const int32_t *lp = c->lp.data(), *leaf = c->leaf.data();
    uint8_t *dnp = dn.data();
    std::vector<int> &dl = dnl;
    if (!std::is_sorted(dl.begin(), dl.end())) std::sort(dl.begin(), dl.end());
    for (int nb : dl) {
      if (!dnp[nb]) continue;
      const int32_t *L = lp + (size_t)nb * 8;
      bool keep = true;
      for (int f = 4; f < 8; ++f) {
        const int n1 = L[f];
        if (n1 != -1 && !leaf[n1]) keep = false;
      }
      if (!keep) dnp[nb] = 0;
      else if (nb <= nbmaxp1 && L[3] == 1 && L[0] >= 2 && L[0] <= nlevs) cand[L[0]].push_back(nb);
    }
  }
  for (int nl = nlevs; nl >= 2; --nl)                // coarsen by level, src/admesh.f90:75-105
    for (int nb : cand[nl]) {
      const int dad = c->LP(nb, 2), s1 = nb, s2 = c->LP(nb, 6), s3 = c->LP(nb, 8), s4 = c->LP(c->LP(nb, 6), 8);
      if (dn[s2] && dn[s3] && dn[s4]) {
        ops.push_back(MeshOp{1, dad, s1, s2, s3, s4, 0, 0});
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

int compile_tables(Mesh *c) {
  int rc = compile_leaves(c);
  return rc ? rc : compile_ops(c);
}

int refresh_tables(Mesh *c) {
  if (c->tables_valid) return WX_OK;
  if (!c->bounds_external) locatebounds(c);
  int rc = compile_tables(c);
  if (rc) return rc;
  c->tables_valid = true;
  return WX_OK;
}

int admesh_from_flags(Mesh *c, const int8_t *flag_of_leaf, std::vector<MeshOp> &ops) {
  std::vector<int> &upl = c->admUp, &dnl = c->admDown;   // scratch kept between calls
  upl.clear();
  dnl.clear();
  const size_t n0 = ops.size();
  for (size_t l = 0; l < c->leaves.size(); ++l) {   // markref, src/markref.f90:37-44
    const int id = c->leaves[l].id;
    if (flag_of_leaf[l] == 1) { c->up[id] = 1; upl.push_back(id); }
    else if (flag_of_leaf[l] == 2) { c->dn[id] = 1; dnl.push_back(id); }
  }
  std::vector<MeshOp> mine;
  int rc = admesh_tree(c, c->up, c->dn, mine, upl, dnl);
  if (rc) return rc;
  for (const MeshOp &o : mine) {
    if (o.kind == 0) {
      c->new_ids.push_back(o.s1); c->new_ids.push_back(o.s2);
      c->new_ids.push_back(o.s3); c->new_ids.push_back(o.s4);
    } else {
      c->new_ids.push_back(o.a);
    }
  }
  if (!mine.empty()) {
    c->tables_valid = false;
    c->bounds_external = false;
  }
  ops.insert(ops.end(), mine.begin(), mine.end());
  (void)n0;
  return WX_OK;
}

int basegrid_tree(Mesh *c) {
  const int nrx = c->p.nbrootx, nry = c->p.nbrooty;
  const int nroots = nrx * nry;
  if (nrx < 1 || nry < 1) return fail(WX_EARG, "root layout %d x %d", nrx, nry);
  if (nroots > c->cap) return fail(WX_EBLOCKS, "capacity %d below root count %d", c->cap, nroots);
  c->all_dirty();
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
  c->tables_valid = false;
  c->stats_all_stale = true;
  c->bounds_external = false;
  c->new_ids.clear();
  return WX_OK;
}

int set_tree(Mesh *c, const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
             int32_t nbmax, int32_t nbleafs, bool *changed) {
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
  c->all_dirty();
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

int get_tree(const Mesh *c, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld, int32_t *nbmax,
             int32_t *nbleafs) {
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

}  // namespace wx

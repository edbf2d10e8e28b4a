// wx_lbuild.h -- the ordered ghost list and its per-segment gather ops,
// computed per element from the tree (the device table builder).
//
// src/locatebounds.f90:29-540 (MPIP forms, one rank) walks the leaves in
// ascending block ID and appends, for faces left, right, bottom, top:
//   * a box face:                      one entry (nb, 0, -face)
//   * a coarser neighbour:             none (the coarser block owns the face)
//   * a same-level leaf neighbour:     one entry (nb, nbr, 1 + 3 f), made by
//                                      the lower of the two IDs only (the
//                                      other block's face is then marked done)
//   * a same-level neighbour with sons: two entries, one per finer block
//                                      facing nb (:130-131, :237-238,
//                                      :345-346, :452-453)
// Every count is local to the block, so the list position of an entry is an
// exclusive scan of the per-block counts plus its rank within the block.
// The boundary routines (src/boundaryII.f90:35-514, src/boundaryI.f90:
// 34-396) apply the list in order; each ghost segment (side, half) of a leaf
// has exactly one writer in it, which writer() finds from the tree and the
// scanned positions without walking the list.  The OUTFYB entries of
// boundaryI (a bottom box entry at position e writes block e + 1,
// src/boundaryI.f90:87) are the one place where a segment of fill I can have
// a second, later writer (late()).
//
// The functions take the tree through an accessor (lp(nb, k) for
// lp(nb, k + 1) of the reference, leaf(nb)): the device's rows (RowTree) or
// the host's arrays.  They run in the builder kernels (wx_lbuild.hip), in
// the host's leaf records (the imposition flag) and in the host-only checks
// against the host compilation (tests/hostsim).
#pragma once
#include "wx_device.h"

#if defined(__HIPCC__)
#define WXL_FN __host__ __device__ __forceinline__
#define WXL_MFN __host__ __device__ __forceinline__
#else
#define WXL_FN static inline
#define WXL_MFN inline
#endif

namespace wx {
namespace lb {

// The tree as the device builder holds it: one row per block ID.
struct RowTree {
  const TreeRow *T;
  WXL_MFN int lp(int nb, int k) const { return T[nb].lp[k]; }
  WXL_MFN int leaf(int nb) const { return T[nb].leaf; }
};

WXL_FN bool valid(int nb, int nbmax) { return nb >= 1 && nb <= nbmax; }

// Number of list entries that leaf nb makes for face f (0 left, 1 right,
// 2 bottom, 3 top); -1 on a pointer outside the tree.
template <class Tr>
WXL_FN int face_count(const Tr &t, int nbmax, int nb, int f) {
  const int nbr = t.lp(nb, 4 + f);
  if (nbr == -1) return 1;
  if (!valid(nbr, nbmax)) return -1;
  if (t.lp(nbr, 0) < t.lp(nb, 0)) return 0;
  if (t.leaf(nbr)) return nb < nbr ? 1 : 0;
  return 2;
}

// Entries of block nb (0 for a non-leaf); -1 on a bad pointer.
template <class Tr>
WXL_FN int block_count(const Tr &t, int nbmax, int nb) {
  if (!t.leaf(nb)) return 0;
  int n = 0;
  for (int f = 0; f < 4; ++f) {
    const int c = face_count(t, nbmax, nb, f);
    if (c < 0) return -1;
    n += c;
  }
  return n;
}

// Rank within leaf nb's entries of the first entry of face f.
template <class Tr>
WXL_FN int face_first(const Tr &t, int nbmax, int nb, int f) {
  int n = 0;
  for (int g = 0; g < f; ++g) {
    const int c = face_count(t, nbmax, nb, g);
    n += c > 0 ? c : 0;
  }
  return n;
}

// The two finer blocks across face f of a same-level neighbour nbr that has
// sons (src/locatebounds.f90:130-131, 237-238, 345-346, 452-453).
template <class Tr>
WXL_FN bool fine_pair(const Tr &t, int nbmax, int nbr, int f, int &n1, int &n2) {
  const int s = t.lp(nbr, 2);
  if (!valid(s, nbmax)) return false;
  switch (f) {
    case 0: n1 = t.lp(s, 5); break;   // lp(lp(nbr,3),6)
    case 2: n1 = t.lp(s, 7); break;   // lp(lp(nbr,3),8)
    default: n1 = s; break;           // lp(nbr,3)
  }
  if (!valid(n1, nbmax)) return false;
  n2 = (f < 2) ? t.lp(n1, 7) : t.lp(n1, 5);   // lp(n1,8) for left/right, lp(n1,6) for bottom/top
  return valid(n2, nbmax);
}

// Writer of one ghost segment: source block, kind | offset << GF_OFF_SHIFT,
// list position; late: an OUTFYB entry, which boundaryI applies to another
// block (fill I gets nothing from it here).
struct Writer {
  int src, y, e;
  bool late;
};

// The writer of half h of side s of leaf X; returns 0 or an LBE_* code.
// info[] holds every block's scanned list offset and count; bc is the box
// boundary kind of side s; outfyb: the compat mode of src/boundaryI.f90:87
// with an outflow bottom.
template <class Tr>
WXL_FN int writer(const Tr &t, int nbmax, const LbInfo *info, int bc, bool outfyb, int X, int s, int h, Writer &w) {
  w.late = false;
  const int nbr = t.lp(X, 4 + s);
  if (nbr == -1) {   // box face (boundaryII.f90:38-141, boundaryI.f90:38-107)
    w.src = X;
    w.y = bc ? GK_BOXREF : GK_BOXOUT;
    w.e = info[X].off + face_first(t, nbmax, X, s);
    w.late = outfyb && s == 2;
    return 0;
  }
  if (!valid(nbr, nbmax)) return LBE_POINTER;
  if (t.lp(nbr, 0) < t.lp(X, 0)) {   // coarser neighbour C: one of its two entries across the opposite face
    const int C = nbr, f = s ^ 1;
    const int P = t.lp(C, 4 + f);
    if (!valid(P, nbmax)) return LBE_POINTER;
    int n1, n2;
    if (!fine_pair(t, nbmax, P, f, n1, n2)) return LBE_POINTER;
    const int j = X == n1 ? 0 : (X == n2 ? 1 : -1);
    if (j < 0) return LBE_FAMILY;
    w.src = C;
    w.y = GK_PROLONG | ((8 * j) << GF_OFF_SHIFT);
    w.e = info[C].off + face_first(t, nbmax, C, f) + j;
    return 0;
  }
  if (t.leaf(nbr)) {   // same-level leaf: the entry of the lower ID
    w.src = nbr;
    w.y = GK_COPY;
    w.e = X < nbr ? info[X].off + face_first(t, nbmax, X, s) : info[nbr].off + face_first(t, nbmax, nbr, s ^ 1);
    return 0;
  }
  int n1, n2;   // same level with sons: X's own entry for this half
  if (!fine_pair(t, nbmax, nbr, s, n1, n2)) return LBE_POINTER;
  w.src = h ? n2 : n1;
  w.y = GK_RESTRICT;
  w.e = info[X].off + face_first(t, nbmax, X, s) + h;
  return 0;
}

// Imposition flags of an op written at list position e into block dst from
// src: impose_user_bc runs after each entry on the entry's first block, so
// the source carried the jet if one of its own entries came before e, and
// the destination is imposed after the write if one of its entries comes at
// or after e.
WXL_FN int impose(const LbInfo *info, int src, int dst, int e) {
  int y = 0;
  if (info[src].cnt > 0 && info[src].off < e) y |= GF_SRC_IMPOSED;
  if (info[dst].cnt > 0 && info[dst].off + info[dst].cnt - 1 >= e) y |= GF_DST_IMPOSED;
  return y;
}

// The ops of side s (both halves) of leaf X for fills II and I (seg = 2 s +
// h); bc: the box boundary kind of side s.  e1[h]: the list position of
// fill I's writer of half h (-1: none, an OUTFYB entry of X's own).
template <class Tr>
WXL_FN int side_ops(const Tr &t, int nbmax, const LbInfo *info, int bc, bool outfyb, bool otherb, int X, int s,
                    int2 *o2, int2 *o1, int *e1) {
  for (int h = 0; h < 2; ++h) {
    Writer w;
    const int rc = writer(t, nbmax, info, bc, outfyb, X, s, h, w);
    if (rc) return rc;
    int y = w.y;
    if (otherb) y |= impose(info, w.src, X, w.e);
    o2[h] = int2{w.src, y};
    o1[h] = w.late ? int2{0, GK_NONE} : int2{w.src, y};
    e1[h] = w.late ? -1 : w.e;
  }
  return 0;
}

// The OUTFYB write to the bottom ghosts (half h) of leaf tb in fill I: the
// entry at list position e = tb - 1, when it is a bottom box entry
// (src/boundaryI.f90:87 writes block e + 1), replaces tb's own writer at
// position e1 if it comes later.  The block owning position e is the
// highest ID whose first entry is at or before e (the offsets are an
// exclusive scan over the IDs).  nlist: the list length.
template <class Tr>
WXL_FN bool late(const Tr &t, int nbmax, const LbInfo *info, int nlist, bool otherb, int tb, int e1, int2 *op1) {
  const int e = tb - 1;
  if (e < 0 || e >= nlist || e <= e1) return false;
  int lo = 1, hi = nbmax;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (info[mid].off <= e) lo = mid;
    else hi = mid - 1;
  }
  const int B = lo;
  if (!t.leaf(B) || t.lp(B, 6) != -1 || e != info[B].off + face_first(t, nbmax, B, 2)) return false;
  *op1 = int2{tb, GK_BOXOUT | (otherb ? impose(info, tb, tb, e) : 0)};
  return true;
}

}  // namespace lb
}  // namespace wx

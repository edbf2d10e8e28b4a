// wx_loadbalance.cpp -- the block placement plan of the reference's MPI load
// balancer (src/loadbalance.f90): each block's Hilbert key at the finest
// level, the blocks in key order, and the new block IDs that split them into
// nranks contiguous runs of equal leaf count.  One GPU holds every block, so
// the plan is what a sharded run would apply; it runs on the host, as on the
// reference's master rank (an integer sort of the tree, microseconds).
//
// The sort is the reference's median-of-three quicksort with its insertion
// cutoff (src/loadbalance.f90:320-400): fathers share their south-west son's
// key, and the order among equal keys decides the new IDs, so it is kept.
#include <cstdint>
#include <vector>

#include "wx_runtime.h"

namespace {

// Hilbert state machine of src/loadbalance.f90:274-283, indexed by
// state * 4 + (2 * x bit + y bit): next state in bits 2-3, key digit in 0-1.
constexpr uint8_t kHilbert[16] = {
    (1 << 2) | 0, (0 << 2) | 1, (2 << 2) | 3, (0 << 2) | 2,
    (0 << 2) | 0, (3 << 2) | 3, (1 << 2) | 1, (1 << 2) | 2,
    (2 << 2) | 2, (2 << 2) | 1, (0 << 2) | 3, (3 << 2) | 0,
    (3 << 2) | 2, (1 << 2) | 3, (3 << 2) | 1, (2 << 2) | 0,
};

int32_t hilbert_key(int32_t x, int32_t y, int bits) {
  uint32_t key = 0;
  unsigned st = 0;
  for (int b = bits - 1; b >= 0; --b) {
    unsigned e = kHilbert[st * 4 + ((((unsigned)x >> b) & 1u) << 1 | (((unsigned)y >> b) & 1u))];
    key = key << 2 | (e & 3u);
    st = e >> 2;
  }
  return (int32_t)key;
}

// Positions p[0..n-1] (0-based block slots) ordering key[] ascending, with
// the reference's tie order: ranges [lo, hi] of at most 5 elements go to
// straight insertion; larger ones are split around the median of lo, mid,
// hi, the larger part deferred on a stack.
void key_order(const std::vector<int32_t> &key, std::vector<int32_t> &p) {
  const int n = (int)key.size();
  p.resize(n);
  for (int k = 0; k < n; ++k) p[k] = k;
  auto K = [&](int pos) { return key[p[pos]]; };
  auto sw = [&](int a, int b) { int32_t t = p[a]; p[a] = p[b]; p[b] = t; };
  std::vector<int> todo;
  int lo = 0, hi = n - 1;
  for (;;) {
    if (hi - lo < 5) {
      for (int j = lo + 1; j <= hi; ++j) {
        const int32_t v = p[j], kv = key[v];
        int i = j;
        for (; i > lo && K(i - 1) > kv; --i) p[i] = p[i - 1];
        p[i] = v;
      }
      if (todo.empty()) return;
      hi = todo.back(); todo.pop_back();
      lo = todo.back(); todo.pop_back();
      continue;
    }
    sw((lo + hi) / 2, lo + 1);
    if (K(lo) > K(hi)) sw(lo, hi);
    if (K(lo + 1) > K(hi)) sw(lo + 1, hi);
    if (K(lo) > K(lo + 1)) sw(lo, lo + 1);
    const int32_t piv = p[lo + 1], kp = key[piv];
    int i = lo + 1, j = hi;
    for (;;) {
      while (K(++i) < kp) {}
      while (K(--j) > kp) {}
      if (j < i) break;
      sw(i, j);
    }
    p[lo + 1] = p[j];
    p[j] = piv;
    if (hi - i + 1 >= j - lo) { todo.push_back(i); todo.push_back(hi); hi = j - 1; }
    else                      { todo.push_back(lo); todo.push_back(j - 1); lo = i; }
  }
}

}  // namespace

extern "C" int wx_lb_plan(const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld,
                          int32_t nbmax, int32_t nlevs, int32_t nblocks, int32_t nranks, int32_t *hilbert,
                          int32_t *idnew) {
  if (!lp || !leaf || !icoord || !hilbert || !idnew || nbmax < 1 || ld < nbmax || nranks < 1 || nlevs < 1 ||
      nlevs > 14 || nblocks < 1)
    return wx::fail(WX_EARG, "wx_lb_plan: bad arguments (need nbmax >= 1, ld >= nbmax, nranks >= 1, "
                             "1 <= nlevs <= 14, nblocks >= 1)");
  std::vector<int32_t> key(nbmax);
  int nleaf = 0;
  for (int s = 0; s < nbmax; ++s) {
    const int lev = lp[s];                                 // lp(nb, 1)
    if (lev < 0 || lev > nlevs) return wx::fail(WX_EARG, "wx_lb_plan: block %d has level %d", s + 1, lev);
    const int32_t scale = 1 << (nlevs - lev);
    key[s] = hilbert_key(icoord[s] * scale / wx::NX, icoord[ld + s] * scale / wx::NY, nlevs + 1);
    hilbert[s] = key[s];
    nleaf += leaf[s] != 0;
  }
  std::vector<int32_t> order;
  key_order(key, order);
  // walk the key order: rank r takes IDs r*nblocks + 1, 2, ... and closes
  // once it holds nleaf / nranks leaves (src/loadbalance.f90:86-97)
  const int quota = nleaf / nranks;
  int rank = 0, held = 0, first = 0;
  for (int k = 0; k < nbmax; ++k) {
    const int s = order[k];
    const int32_t id = rank * nblocks + (k - first) + 1;
    idnew[s] = id;
    held += leaf[s] != 0;
    if (held >= quota && rank < nranks - 1) { ++rank; held = 0; first = k + 1; }
    if (id / (rank + 1) == nblocks)
      return wx::fail(WX_EBLOCKS, "wx_lb_plan: block %d would take ID %d, the capacity of rank %d", s + 1, id,
                      rank);
  }
  return WX_OK;
}

// wx_bin.cpp -- the reference's BIN checkpoint files (src/output.f90:39-67,
// written with OUTBIN; read back by the warm start of src/basegrid.f90:
// 167-251) from and into a context, through the public entry points only.
//
// Both files are unformatted streams, little-endian, no record markers:
//   pointers: int32 nbmin, nbmax, nbleafs, nblocks; lp(nbmin:nbmax, 1:8)
//             column-major; leaf(nbmin:nbmax) as 4-byte logicals;
//             icoord(nbmin:nbmax, 1:2) column-major
//   blocks:   for every leaf in ID order, u(nb, 1:neq, -1:18, -1:18) with the
//             variable fastest, then i, then j, as FP64.
// Under MPI each rank writes its own pair; rank r's block IDs start at
// nbmin = r*nblocks + 1 and lp holds global IDs, which the context keeps
// as they are (its capacity must cover the largest).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "wx_runtime.h"

using namespace wx;

namespace {

struct File {
  FILE *f = nullptr;
  ~File() { if (f) fclose(f); }
};

constexpr int NPAD = NP;   // 20 cells per padded row

}  // namespace


namespace {

// The OUTBIN pair of one mesh: tree arrays in Fortran layout (ld = nbmax)
// and the tiles [id][var][j][i] of IDs 0..nbmax.
int write_pair(const char *pointers_path, const char *blocks_path, int32_t nbmax, int32_t nbleafs, int32_t cap,
               int neq, const std::vector<int32_t> &lp, const std::vector<int32_t> &leaf,
               const std::vector<int32_t> &ic, const std::vector<double> &u, int32_t logical_true) {
  if (logical_true != -1 && logical_true != 1) return fail(WX_EARG, "logical_true %d (-1 ifort, 1 flang)", logical_true);
  const int ld = nbmax;
  {
    File out;
    if (!(out.f = fopen(pointers_path, "wb"))) return fail(WX_EIO, "cannot open %s", pointers_path);
    const int32_t head[4] = {1, nbmax, nbleafs, cap};
    std::vector<int32_t> lg(ld);
    for (int n = 0; n < ld; ++n) lg[n] = leaf[n] ? logical_true : 0;
    if (fwrite(head, 4, 4, out.f) != 4 || fwrite(lp.data(), 4, lp.size(), out.f) != lp.size() ||
        fwrite(lg.data(), 4, lg.size(), out.f) != lg.size() || fwrite(ic.data(), 4, ic.size(), out.f) != ic.size())
      return fail(WX_EIO, "short write to %s", pointers_path);
  }
  File out;
  if (!(out.f = fopen(blocks_path, "wb"))) return fail(WX_EIO, "cannot open %s", blocks_path);
  std::vector<double> blk((size_t)neq * NC);
  for (int nb = 1; nb <= nbmax; ++nb) {
    if (!leaf[nb - 1]) continue;
    const double *t = u.data() + (size_t)nb * neq * NC;   // [var][j][i]
    size_t o = 0;
    for (int j = 0; j < NPAD; ++j)
      for (int i = 0; i < NPAD; ++i)
        for (int k = 0; k < neq; ++k) blk[o++] = t[(size_t)k * NC + j * NPAD + i];
    if (fwrite(blk.data(), 8, blk.size(), out.f) != blk.size()) return fail(WX_EIO, "short write to %s", blocks_path);
  }
  return WX_OK;
}

}  // namespace

extern "C" {

int wx_write_bin(wx_ctx *c, const char *pointers_path, const char *blocks_path, int32_t logical_true) {
  if (!c || !pointers_path || !blocks_path) return fail(WX_EARG, "null argument");
  int32_t neq = 0, cap = 0, nbmax = 0, nbleafs = 0;
  int rc = wx_shape(c, &neq, &cap, nullptr);
  if (!rc) rc = wx_info(c, &nbmax, &nbleafs, nullptr);
  if (rc) return rc;
  if (nbmax < 1) return fail(WX_EARG, "no mesh to write");
  const int ld = nbmax;
  std::vector<int32_t> lp((size_t)8 * ld), leaf(ld), ic((size_t)2 * ld);
  int32_t nbm = 0, nbl = 0;
  if ((rc = wx_get_tree(c, lp.data(), leaf.data(), ic.data(), ld, &nbm, &nbl))) return rc;
  std::vector<double> u((size_t)(nbmax + 1) * neq * NC);
  if ((rc = wx_download(c, u.data(), nullptr, WX_LAYOUT_BLOCK, 0))) return rc;
  return write_pair(pointers_path, blocks_path, nbmax, nbleafs, cap, neq, lp, leaf, ic, u, logical_true);
}

int wx_ens_write_bin(wx_ens *e, int32_t inst, const char *pointers_path, const char *blocks_path,
                     int32_t logical_true) {
  if (!e || !pointers_path || !blocks_path) return fail(WX_EARG, "null argument");
  int32_t neq = 0, cap = 0, nbmax = 0, nbleafs = 0;
  int rc = wx_ens_shape(e, inst, &neq, &cap, nullptr);
  if (!rc) rc = wx_ens_instance(e, inst, nullptr, nullptr, &nbmax, &nbleafs);
  if (rc) return rc;
  if (nbmax < 1) return fail(WX_EARG, "instance %d has no mesh to write", inst);
  const int ld = nbmax;
  std::vector<int32_t> lp((size_t)8 * ld), leaf(ld), ic((size_t)2 * ld);
  int32_t nbm = 0, nbl = 0;
  if ((rc = wx_ens_get_tree(e, inst, lp.data(), leaf.data(), ic.data(), ld, &nbm, &nbl))) return rc;
  std::vector<double> u((size_t)(nbmax + 1) * neq * NC);
  if ((rc = wx_ens_download(e, inst, u.data()))) return rc;
  return write_pair(pointers_path, blocks_path, nbmax, nbleafs, cap, neq, lp, leaf, ic, u, logical_true);
}

int wx_read_bin(wx_ctx *c, int32_t nfiles, const char *const *pointers_paths, const char *const *blocks_paths) {
  if (!c || nfiles < 1 || !pointers_paths || !blocks_paths) return fail(WX_EARG, "bad argument");
  int32_t neq = 0, cap = 0;
  int rc = wx_shape(c, &neq, &cap, nullptr);
  if (rc) return rc;
  // global arrays over IDs 1..cap (Fortran layout, ld = cap)
  const int ld = cap;
  std::vector<int32_t> lp((size_t)8 * ld, 0), leaf(ld, 0), ic((size_t)2 * ld, 0);
  std::vector<double> u((size_t)(cap + 1) * neq * NC, 0.0);
  int32_t nbmax = 0, nbleafs = 0;
  std::vector<double> blk((size_t)neq * NC);
  for (int r = 0; r < nfiles; ++r) {
    File in;
    if (!(in.f = fopen(pointers_paths[r], "rb"))) return fail(WX_EIO, "cannot open %s", pointers_paths[r]);
    int32_t head[4];
    if (fread(head, 4, 4, in.f) != 4) return fail(WX_EIO, "%s: short header", pointers_paths[r]);
    const int32_t nb0 = head[0], nb1 = head[1];
    if (nb0 < 1 || nb1 < nb0 - 1 || nb1 > cap)
      return fail(WX_EBLOCKS, "%s: block IDs %d..%d outside the capacity %d", pointers_paths[r], nb0, nb1, cap);
    const int n = nb1 - nb0 + 1;
    std::vector<int32_t> a((size_t)8 * n), lg(n), b((size_t)2 * n);
    if (fread(a.data(), 4, a.size(), in.f) != a.size() || fread(lg.data(), 4, lg.size(), in.f) != lg.size() ||
        fread(b.data(), 4, b.size(), in.f) != b.size())
      return fail(WX_EIO, "%s: short file", pointers_paths[r]);
    for (int q = 0; q < n; ++q) {
      const int nb = nb0 + q;
      for (int k = 0; k < 8; ++k) lp[(size_t)k * ld + nb - 1] = a[(size_t)k * n + q];
      leaf[nb - 1] = lg[q] != 0;
      ic[nb - 1] = b[q];
      ic[(size_t)ld + nb - 1] = b[(size_t)n + q];
    }
    nbmax = std::max(nbmax, nb1);
    nbleafs += head[2];
    File bin;
    if (!(bin.f = fopen(blocks_paths[r], "rb"))) return fail(WX_EIO, "cannot open %s", blocks_paths[r]);
    for (int nb = nb0; nb <= nb1; ++nb) {
      if (!leaf[nb - 1]) continue;
      if (fread(blk.data(), 8, blk.size(), bin.f) != blk.size()) return fail(WX_EIO, "%s: short file", blocks_paths[r]);
      double *t = u.data() + (size_t)nb * neq * NC;
      size_t o = 0;
      for (int j = 0; j < NPAD; ++j)
        for (int i = 0; i < NPAD; ++i)
          for (int k = 0; k < neq; ++k) t[(size_t)k * NC + j * NPAD + i] = blk[o++];
    }
  }
  int32_t count = 0;
  for (int nb = 0; nb < nbmax; ++nb) count += leaf[nb];
  if (count != nbleafs) return fail(WX_EBLOCKS, "the files list %d leaves, their headers %d", count, nbleafs);
  return wx_upload(c, u.data(), WX_LAYOUT_BLOCK, lp.data(), leaf.data(), ic.data(), ld, nbmax, nbleafs);
}

}  // extern "C"

// wx_lbuild.hip -- the device table builder: the ghost-gather ops of both
// fills, built on the GPU from the tree rows after a mesh change
// (src/locatebounds.f90:29-540 and the list's application order in
// src/boundaryII.f90:35-514, src/boundaryI.f90:34-396).
//
// Two launches, on a stream of their own so that they run beside the dt
// reduction and stage A (which need only the leaf records the host
// uploads):
//   k_lb_place1  one workgroup: entries and leaves of every ID, their
//                exclusive scan (leaf position, first list position), the
//                list length
//   k_lb_ops     the 8 ghost ops of every leaf for fills II and I, with the
//                OUTFYB writes of fill I in that compat mode
// Above 2048 block IDs the scan takes three launches instead (k_lb_count,
// k_lb_scan, k_lb_place): one workgroup walking more IDs in chunks takes
// longer than the two extra launches.
// The per-element logic is in wx_lbuild.h, shared with the host.
#include <hip/hip_runtime.h>

#include "wx_lbuild.h"
#include "wx_runtime.h"

namespace wx {
namespace {

constexpr int LB_BS = 256;
constexpr int LB_ONE_WG = 2048;   // block IDs up to which one workgroup scans them

// Exclusive scan of (a, b) over a workgroup of BS threads; *ta, *tb: totals.
template <int BS>
__device__ void scan2(int a, int b, int *ea, int *eb, int *ta, int *tb) {
  __shared__ int wa[BS / 64], wb[BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = a, y = b;
  for (int d = 1; d < 64; d <<= 1) {
    const int px = __shfl_up(x, d, 64), py = __shfl_up(y, d, 64);
    if (lane >= d) {
      x += px;
      y += py;
    }
  }
  if (lane == 63) {
    wa[w] = x;
    wb[w] = y;
  }
  __syncthreads();
  int ba = 0, bb = 0, sa = 0, sb = 0;
  for (int k = 0; k < BS / 64; ++k) {
    if (k < w) {
      ba += wa[k];
      bb += wb[k];
    }
    sa += wa[k];
    sb += wb[k];
  }
  __syncthreads();   // the wave sums may be rewritten by the next call
  *ea = ba + x - a;
  *eb = bb + y - b;
  *ta = sa;
  *tb = sb;
}

__device__ __forceinline__ void lb_counts(const TreeRow *T, int nbmax, int id, int *nl, int *ne, DevScalars *S) {
  *nl = 0;
  *ne = 0;
  if (id < 1 || id > nbmax || !T[id].leaf) return;
  *nl = 1;
  const int c = lb::block_count(lb::RowTree{T}, nbmax, id);
  if (c < 0) {
    S->mesh_err = LBE_POINTER;
    return;
  }
  *ne = c;
}

__global__ __launch_bounds__(LB_BS) void k_lb_count(const TreeRow *__restrict__ T, int nbmax, LbPart *__restrict__ part,
                                                    DevScalars *S) {
  const int id = blockIdx.x * LB_BS + threadIdx.x;
  int nl, ne, el, ee, tl, te;
  lb_counts(T, nbmax, id, &nl, &ne, S);
  scan2<LB_BS>(nl, ne, &el, &ee, &tl, &te);
  if (threadIdx.x == 0) part[blockIdx.x] = LbPart{tl, te};
}

// One workgroup of 1024 threads: exclusive scan of the nblk workgroup sums
// in place, in chunks of 1024 with a carry; the list length to *nlist.
__global__ __launch_bounds__(1024) void k_lb_scan(LbPart *__restrict__ part, int nblk, int nleaf_host,
                                                  int32_t *nlist, DevScalars *S) {
  int cl = 0, ce = 0;
  for (int base = 0; base < nblk; base += 1024) {
    const int i = base + (int)threadIdx.x;
    const LbPart v = i < nblk ? part[i] : LbPart{0, 0};
    int el, ee, tl, te;
    scan2<1024>(v.leaves, v.entries, &el, &ee, &tl, &te);
    if (i < nblk) part[i] = LbPart{cl + el, ce + ee};
    cl += tl;
    ce += te;
  }
  if (threadIdx.x == 0) {
    if (cl != nleaf_host) S->mesh_err = LBE_COUNT;
    *nlist = ce;
  }
}

// One workgroup of 1024 threads over IDs 0..nbmax in chunks of 1024, the
// scan carried from chunk to chunk.
__global__ __launch_bounds__(1024) void k_lb_place1(const TreeRow *__restrict__ T, int nbmax, int nleaf_host,
                                                    LbInfo *__restrict__ info, int32_t *nlist, DevScalars *S) {
  int cl = 0, ce = 0;
  for (int base = 0; base <= nbmax; base += 1024) {
    const int id = base + (int)threadIdx.x;
    int nl, ne, el, ee, tl, te;
    lb_counts(T, nbmax, id, &nl, &ne, S);
    scan2<1024>(nl, ne, &el, &ee, &tl, &te);
    if (id <= nbmax) info[id] = LbInfo{nl ? cl + el : -1, ce + ee, ne, 0};
    cl += tl;
    ce += te;
  }
  if (threadIdx.x == 0) {
    if (cl != nleaf_host) S->mesh_err = LBE_COUNT;
    *nlist = ce;
  }
}

struct LbCfg {
  int nbmax, leaf_cap;
  int bc0, bc1, bc2, bc3;   // box boundary kinds of the left, right, bottom, top sides
  int outfyb, otherb;
};

__global__ __launch_bounds__(LB_BS) void k_lb_place(const TreeRow *__restrict__ T, int nbmax,
                                                    const LbPart *__restrict__ part, LbInfo *__restrict__ info,
                                                    DevScalars *S) {
  const int id = blockIdx.x * LB_BS + threadIdx.x;
  int nl, ne, el, ee, tl, te;
  lb_counts(T, nbmax, id, &nl, &ne, S);
  scan2<LB_BS>(nl, ne, &el, &ee, &tl, &te);
  if (id > nbmax) return;
  const LbPart b = part[blockIdx.x];
  info[id] = LbInfo{nl ? b.leaves + el : -1, b.entries + ee, ne, 0};
}

// Four threads per block ID, one per side.
__global__ __launch_bounds__(LB_BS) void k_lb_ops(const TreeRow *__restrict__ T, LbCfg cf,
                                                  const LbInfo *__restrict__ info, const int32_t *__restrict__ nlist,
                                                  int2 *__restrict__ ops2, int2 *__restrict__ ops1, DevScalars *S) {
  const int t = blockIdx.x * LB_BS + threadIdx.x;
  const int id = t >> 2, s = t & 3;
  if (id < 1 || id > cf.nbmax || !T[id].leaf) return;
  const int l = info[id].lidx;
  if (l < 0 || l >= cf.leaf_cap) {
    S->mesh_err = LBE_COUNT;
    return;
  }
  const lb::RowTree tr{T};
  int2 o2[2], o1[2];
  int e1[2];
  const int bcs = s == 0 ? cf.bc0 : s == 1 ? cf.bc1 : s == 2 ? cf.bc2 : cf.bc3;
  const int rc = lb::side_ops(tr, cf.nbmax, info, bcs, cf.outfyb != 0, cf.otherb != 0, id, s, o2, o1, e1);
  if (rc) {
    S->mesh_err = rc;
    return;
  }
  const size_t k = (size_t)l * 8 + 2 * s;
  ops2[k] = o2[0];
  ops2[k + 1] = o2[1];
  if (cf.outfyb) {
    if (s == 2) {
      const int n = *nlist;
      for (int h = 0; h < 2; ++h) (void)lb::late(tr, cf.nbmax, info, n, cf.otherb != 0, id, e1[h], &o1[h]);
    }
    ops1[k] = o1[0];
    ops1[k + 1] = o1[1];
  }
}

}  // namespace

int launch_build_tables(const wx_params &p, const TreeRow *T, int nbmax, int nleaf_host, LbInfo *info, LbPart *part,
                        LbOut out, int multi_wg_scan, DevScalars *S, hipStream_t stream) {
  if (nbmax < 0) return WX_EARG;
  LbCfg cf;
  cf.nbmax = nbmax;
  cf.leaf_cap = out.leaf_cap;
  cf.bc0 = p.bc[0];
  cf.bc1 = p.bc[1];
  cf.bc2 = p.bc[2];
  cf.bc3 = p.bc[3];
  cf.outfyb = lb_outfyb(p);
  cf.otherb = p.otherb != 0;
  if (nbmax + 1 <= LB_ONE_WG && !multi_wg_scan) {
    hipLaunchKernelGGL(k_lb_place1, dim3(1), dim3(1024), 0, stream, T, nbmax, nleaf_host, info, out.nlist, S);
  } else {
    const int nblk = (nbmax + 1 + LB_BS - 1) / LB_BS;
    hipLaunchKernelGGL(k_lb_count, dim3(nblk), dim3(LB_BS), 0, stream, T, nbmax, part, S);
    hipLaunchKernelGGL(k_lb_scan, dim3(1), dim3(1024), 0, stream, part, nblk, nleaf_host, out.nlist, S);
    hipLaunchKernelGGL(k_lb_place, dim3(nblk), dim3(LB_BS), 0, stream, T, nbmax, (const LbPart *)part, info, S);
  }
  const int g4 = (4 * (nbmax + 1) + LB_BS - 1) / LB_BS;
  hipLaunchKernelGGL(k_lb_ops, dim3(g4), dim3(LB_BS), 0, stream, T, cf, (const LbInfo *)info,
                     (const int32_t *)out.nlist, out.ops2, cf.outfyb ? out.ops1 : nullptr, S);
  return hipGetLastError() == hipSuccess ? WX_OK : WX_EHIP;
}

}  // namespace wx

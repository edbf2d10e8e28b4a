// wx_kernels.hip -- gfx950 kernels of the Walicxe-2D hot path.
//
// Data layout in HBM: one padded 20x20 tile per variable per block,
// [id][var][j][i] with i fastest, id = the reference's block ID.  Every
// kernel below assigns one 256-thread workgroup (4 wave64) to one leaf
// block and stages the tile in LDS.
//
// One step (src/tstep.f90:4-126) is three launches over the leaf list:
//   k_stageA      Uⁿ -> U^{n+1/2} interior (first-order fluxes, dt/2)
//   k_stageB      pulls the 2-layer ghost ring of U^{n+1/2} from the
//                 neighbours (boundaryII semantics), limited fluxes, update,
//                 viscosity; writes U^{n+1} interior and the U^{n+1/2} ring
//                 in place over Uⁿ (tstep.f90:85 copies the whole padded block)
//   k_fillI_cool  pulls ghost layer 1 (boundaryI), imposes the jet, applies
//                 cooling on cells 0..17, writes the final tile to the other
//                 buffer, and emits the next CFL partials and the refinement
//                 flag of the block (fused timestep/criteria epilogue).
#include <hip/hip_runtime.h>
#include "wx_device.h"
#include "wx_mlp.h"
#include "wx_physics.h"

namespace wx {

__device__ __forceinline__ int IDX(int i, int j) { return (j + 1) * NP + (i + 1); }

// Queue entry of workgroup b in a one-entry-per-workgroup launch of G
// workgroups.  The hardware deals workgroups to the 8 XCDs round robin
// (b % 8); within each run of 64 workgroups this hands XCD x the 8
// consecutive entries 8x..8x+7, so that neighbouring blocks, which share
// ghost data, are computed behind the same L2.  The last partial run keeps
// the identity, so the map is a bijection of [0, G).
__device__ __forceinline__ int xcd_entry(int b, int G) {
  if (b >= (G & ~63)) return b;
  return (b & ~63) | ((b & 7) << 3) | ((b >> 3) & 7);
}

// ---------------------------------------------------------------------
// Ghost gather: value of one non-corner ghost cell of block `dst` at
// (side, layer, pos) from the op that last wrote its segment.
//   side 0 left (i = 1-layer), 1 right (i = 16+layer), 2 bottom (j = 1-layer),
//   3 top (j = 16+layer); layer 1 is adjacent to the interior, 2 outer;
//   pos 1..16 runs along the side.
// Formulas: src/boundaryII.f90:143-500 and src/boundaryI.f90:109-382.
template <int NEQ>
__device__ __forceinline__ void src_cell(const wx_params *__restrict__ p, const double *__restrict__ A,
                                         const BlockRec *__restrict__ binfo, int b, int i, int j,
                                         bool imposed, double vjet, double *v) {
  const double *t = A + (size_t)b * NEQ * NC + IDX(i, j);
#pragma unroll
  for (int k = 0; k < NEQ; ++k) v[k] = t[k * NC];
  if (imposed) {
    BlockRec r = binfo[b];
    if (in_jet(p, r.level, r.icx, r.icy, i, j)) jet_state<NEQ>(p, vjet, v);
  }
}

template <int NEQ>
__device__ void ghost_gather(const wx_params *__restrict__ p, const double *__restrict__ A,
                             const BlockRec *__restrict__ binfo, int2 op, int dst, int dlev,
                             int dicx, int dicy, int side, int layer, int pos, double vjet,
                             double *g) {
  const int kind = op.y & 15;
  const int src = op.x;
  const bool simp = (op.y & GF_SRC_IMPOSED) != 0;
  const bool dimp = (op.y & GF_DST_IMPOSED) != 0;
  const int off = (op.y >> GF_OFF_SHIFT) & 15;
  const bool xside = side < 2;
  // destination cell
  int di, dj;
  if (side == 0) { di = 1 - layer; dj = pos; }
  else if (side == 1) { di = NX + layer; dj = pos; }
  else if (side == 2) { dj = 1 - layer; di = pos; }
  else { dj = NY + layer; di = pos; }
  if (kind == GK_COPY) {
    int si = di, sj = dj;
    if (side == 0) si = di + NX;
    else if (side == 1) si = di - NX;
    else if (side == 2) sj = dj + NY;
    else sj = dj - NY;
    src_cell<NEQ>(p, A, binfo, src, si, sj, simp, vjet, g);
  } else if (kind == GK_RESTRICT) {
    const int half = (pos - 1) >> 3;
    int i1, j1;
    if (xside) {
      int jj = pos - 8 * half;
      i1 = (side == 0) ? (2 * di - 1 + NX) : (2 * (di - NX) - 1);
      j1 = 2 * jj - 1;
    } else {
      int ii = pos - 8 * half;
      i1 = 2 * ii - 1;
      j1 = (side == 2) ? (2 * dj - 1 + NY) : (2 * (dj - NY) - 1);
    }
    double a[NEQ], b2[NEQ], c[NEQ], d[NEQ];
    src_cell<NEQ>(p, A, binfo, src, i1, j1, simp, vjet, a);
    src_cell<NEQ>(p, A, binfo, src, i1 + 1, j1, simp, vjet, b2);
    src_cell<NEQ>(p, A, binfo, src, i1, j1 + 1, simp, vjet, c);
    src_cell<NEQ>(p, A, binfo, src, i1 + 1, j1 + 1, simp, vjet, d);
#pragma unroll
    for (int k = 0; k < NEQ; ++k) g[k] = 0.25 * (((a[k] + b2[k]) + c[k]) + d[k]);
  } else if (kind == GK_PROLONG) {
    int si, sj;
    if (xside) {
      si = (side == 0) ? NX : 1;
      sj = (dj + 1) / 2 + off;
    } else {
      sj = (side == 2) ? NY : 1;
      si = (di + 1) / 2 + off;
    }
    src_cell<NEQ>(p, A, binfo, src, si, sj, simp, vjet, g);
  } else {  // GK_BOXOUT / GK_BOXREF : mirror own interior
    int si = di, sj = dj;
    if (side == 0) si = 1 - di;
    else if (side == 1) si = 2 * NX + 1 - di;
    else if (side == 2) sj = 1 - dj;
    else sj = 2 * NY + 1 - dj;
    src_cell<NEQ>(p, A, binfo, dst, si, sj, simp, vjet, g);
    if (kind == GK_BOXREF) {
      if (xside) g[1] = -g[1];
      else g[2] = -g[2];
    }
  }
  if (dimp && in_jet(p, dlev, dicx, dicy, di, dj)) jet_state<NEQ>(p, vjet, g);
}

// (Selects rather than branches: branches that assign i or j make the
// compiler keep the pair in scratch memory.)
__device__ __forceinline__ void ghost_coords(int side, int layer, int pos, int &i, int &j) {
  static_assert(NX == NY, "one edge index for both directions");
  const int edge = (side & 1) ? NX + layer : 1 - layer;
  const bool xside = side < 2;
  i = xside ? edge : pos;
  j = xside ? pos : edge;
}

// Flux sweep shared by both stages.  prim: LDS tile [NEQ][NC].  F: x-faces
// fi=0..16 on rows 1..16 at [k][(j-1)*17+fi]; G: y-faces fj=0..16 on columns
// 1..16 at [k][fj*16+(i-1)].  choice 2 reconstructs with the limiter.
// SOLVER / LIM fix the solver and limiter at compile time (-1 / -99: read
// from p); the arithmetic is the same either way.
template <int NEQ, int CHOICE, int SOLVER = -1, int LIM = -99>
__device__ __forceinline__ int flux_sweep_t(const wx_params *__restrict__ p, const double *prim,
                                            double *F, double *G) {
  int err = 0;
  const int lim = LIM != -99 ? LIM : p->limiter;
  for (int f = threadIdx.x; f < 544; f += blockDim.x) {
    // x-face fi on row j, or y-face fj on column i: the cells across the face
    // are `stride` apart, and a y-face's states are swapped into the x frame
    // (swapy, src/swapy.f90:4-13), so one solver call serves both
    const bool yf = f >= 272;
    int c0, stride;
    if (!yf) {
      c0 = IDX(f % 17, f / 17 + 1);
      stride = 1;
    } else {
      const int g = f - 272;
      c0 = IDX(g % 16 + 1, g / 16);
      stride = NP;
    }
    double ql[NEQ], qr[NEQ], ff[NEQ];
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      const int kk = (k == 1 || k == 2) && yf ? 3 - k : k;
      ql[k] = prim[kk * NC + c0];
      qr[k] = prim[kk * NC + c0 + stride];
    }
    if (CHOICE == 2) {
      double qll[NEQ], qrr[NEQ];
#pragma unroll
      for (int k = 0; k < NEQ; ++k) {
        const int kk = (k == 1 || k == 2) && yf ? 3 - k : k;
        qll[k] = prim[kk * NC + c0 - stride];
        qrr[k] = prim[kk * NC + c0 + 2 * stride];
      }
      limiter<NEQ>(lim, qll, ql, qr, qrr);
    }
    err |= riemann<NEQ, SOLVER>(p, ql, qr, ff);
    if (!yf) {
      const int fi = f % 17, j = f / 17 + 1;
#pragma unroll
      for (int k = 0; k < NEQ; ++k) F[k * 272 + (j - 1) * 17 + fi] = ff[k];
    } else {
      const int g = f - 272;
      const int i = g % 16 + 1, fj = g / 16;
#pragma unroll
      for (int k = 0; k < NEQ; ++k) G[k * 272 + fj * 16 + (i - 1)] = ff[(k == 1 || k == 2) ? 3 - k : k];
    }
  }
  return err;
}

// The sweep with the common configurations compiled in: HLLC, and in
// stage B the minmod or van Leer limiter; anything else reads p.
template <int NEQ, int CHOICE>
__device__ __forceinline__ int flux_sweep(const wx_params *__restrict__ p, const double *prim,
                                          double *F, double *G) {
  if (p->solver == 0) {
    if (CHOICE == 1) return flux_sweep_t<NEQ, CHOICE, 0>(p, prim, F, G);
    if (p->limiter == 1) return flux_sweep_t<NEQ, CHOICE, 0, 1>(p, prim, F, G);
    if (p->limiter == 2) return flux_sweep_t<NEQ, CHOICE, 0, 2>(p, prim, F, G);
  }
  return flux_sweep_t<NEQ, CHOICE>(p, prim, F, G);
}

// Cylindrical source terms, src/step.f90:49-76 (prims of the stage).
template <int NEQ>
__device__ __forceinline__ void cyl_source(const wx_params *__restrict__ p, const double *prim, int i,
                                           int j, int icy, int lev, double *S) {
  double radius = ((double)(j + icy) - 0.5) * p->dy[lev - 1];
  double rho = prim[0 * NC + IDX(i, j)];
  double v1 = prim[1 * NC + IDX(i, j)], v2 = prim[2 * NC + IDX(i, j)], pr = prim[3 * NC + IDX(i, j)];
  double H = 0.5 * rho * (v1 * v1 + v2 * v2) + (p->cv + 1.) * pr;
  S[0] = -(rho * v2 / radius);
  S[1] = -(rho * v1 * v2 / radius);
  S[2] = -(rho * v2 * v2 / radius);
  S[3] = -(H * v2 / radius);
#pragma unroll
  for (int k = 4; k < NEQ; ++k) S[k] = -(prim[k * NC + IDX(i, j)] * v2 / radius);
}

// ---------------------------------------------------------------------
// Stage A: src/tstep.f90:22-46 (hllcfluxes choice 1 + step with dt/2).
// The one-entry-per-workgroup form is compiled for 8 waves per SIMD (at 512
// threads that caps it at 64 VGPRs, with a few spills: 8.5 vs 9.2 ns per
// block at 74 VGPRs and 6 waves on the 16-variant cooling jet; the smaller
// workgroups are held to fewer waves by their LDS anyway).
template <int NEQ, int BS, bool PERSIST>
__global__ __launch_bounds__(BS, PERSIST ? 1 : 8) void k_stageA(const wx_params *__restrict__ P, QueueRef q,
                                                const double *__restrict__ U, double *__restrict__ Q,
                                                const double *__restrict__ Pc, DevScalars *__restrict__ Sa) {
  __shared__ double prim[NEQ * NC];
  __shared__ double F[NEQ * 272];
  __shared__ double G[NEQ * 272];
  // the first entry's record is read together with the queue length
  const int b0 = PERSIST ? (int)blockIdx.x : xcd_entry(blockIdx.x, gridDim.x);
  LeafRec Lnext = q.rec[min(b0, q.cap - 1)];
  const int nleaf = q.hdr->nleaf;
  // one queue entry per workgroup, or (PERSIST) entries blockIdx.x +
  // k * gridDim.x; the loop form costs registers, so each is its own kernel
  for (int l = b0; l < nleaf; l += gridDim.x) {
  const LeafRec L = Lnext;
  if (PERSIST && l + (int)gridDim.x < nleaf) Lnext = q.rec[l + gridDim.x];
  const wx_params *__restrict__ p = P + L.inst;
  DevScalars *__restrict__ S = Sa + L.inst;
  const double *ub = U + (size_t)L.id * NEQ * NC;
  // rows j = 0..17 are one contiguous range [NP, 19*NP) of every variable.
  // Under DMC the pressure of cells 0..17 is the one the last cooling left
  // (the reference's primit, src/coolingdmc.f90:27-38 and
  // src/tstep.f90:27-36), not the one u gives back.
  const bool carried = Pc && p->cooling == 2;
  const double *pcb = carried ? Pc + (size_t)L.id * NC : nullptr;
  for (int c = threadIdx.x; c < 18 * NP; c += blockDim.x) {
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = ub[k * NC + NP + c];
    uprim<NEQ>(p, q, uu, T);
    if (carried) {
      const int col = c % NP;
      if (col >= 1 && col <= 18) q[3] = pcb[NP + c];
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k) prim[k * NC + NP + c] = q[k];
  }
  __syncthreads();
  int err = flux_sweep<NEQ, 1>(p, prim, F, G);
  __syncthreads();
  if (threadIdx.x < 256) {   // one thread per interior cell
    const double dtm = S->dt / 2.;
    const double dtdx = dtm / p->dx[L.level - 1];
    const double dtdy = dtm / p->dy[L.level - 1];
    const int i = threadIdx.x % 16 + 1, j = threadIdx.x / 16 + 1;
    double Sc[NEQ];
    if (p->cylindrical) cyl_source<NEQ>(p, prim, i, j, L.icy, L.level, Sc);
    double *qb = Q + (size_t)L.id * NEQ * NC;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      double v = ub[k * NC + IDX(i, j)] - dtdx * (F[k * 272 + (j - 1) * 17 + i] - F[k * 272 + (j - 1) * 17 + i - 1])
                 - dtdy * (G[k * 272 + j * 16 + (i - 1)] - G[k * 272 + (j - 1) * 16 + (i - 1)]);
      if (p->cylindrical) v = v + dtm * Sc[k];
      qb[k * NC + IDX(i, j)] = v;
    }
  }
  if (err) atomicOr(&S->err, WX_EHLLC);
  __syncthreads();   // the next queue entry reuses the LDS tiles
   if (!PERSIST) break;   // one entry per workgroup: no loop
  }
}

// ---------------------------------------------------------------------
// Stage B: src/tstep.f90:53-88 with boundaryII pulled from the neighbours.
template <int NEQ, int BS, bool PERSIST>
__global__ __launch_bounds__(BS) void k_stageB(const wx_params *__restrict__ P, QueueRef q,
                                                const int2 *__restrict__ ops,
                                                const BlockRec *__restrict__ binfo,
                                                double *__restrict__ U, const double *__restrict__ Q,
                                                double *__restrict__ Qring, DevScalars *__restrict__ Sa) {
  __shared__ double prim[NEQ * NC];
  __shared__ double F[NEQ * 272];
  __shared__ double G[NEQ * 272];
  const int b0 = PERSIST ? (int)blockIdx.x : xcd_entry(blockIdx.x, gridDim.x);
  LeafRec Lnext = q.rec[min(b0, q.cap - 1)];
  const int nleaf = q.hdr->nleaf;
  for (int l = b0; l < nleaf; l += gridDim.x) {
  const int r = l;
  const LeafRec L = Lnext;
  if (PERSIST && l + (int)gridDim.x < nleaf) Lnext = q.rec[l + gridDim.x];
  const wx_params *__restrict__ p = P + L.inst;
  DevScalars *__restrict__ S = Sa + L.inst;
  const int b = L.id;
  const double time = S->time;
  const bool imp = (L.flags & 1) != 0;
  const double vjet = p->otherb ? jet_velocity_at(p, S, time) : 0.0;
  const double *qb = Q + (size_t)b * NEQ * NC;
  double *ub = U + (size_t)b * NEQ * NC;
  const int tid = threadIdx.x;
  // two-layer ghost ring: 4 sides x 2 layers x 16 positions = 128 cells,
  // gathered by waves of their own when the workgroup has them.  The op of
  // the ghost's segment and Uⁿ of the thread's interior cell (used by the
  // update after the sweep) are read first, so that their latency overlaps
  // the interior's primitive recovery.
  const int gt = BS >= 384 ? tid - 256 : tid;
  const bool ghost = gt >= 0 && gt < 128;
  const int side = gt >> 5, layer = ((gt >> 4) & 1) + 1, pos = (gt & 15) + 1;
  int2 op = make_int2(0, 0);
  if (ghost) op = ops[r * 8 + side * 2 + ((pos - 1) >> 3)];
  double u0[NEQ];
  if (tid < 256) {
    const int i = tid % 16 + 1, j = tid / 16 + 1;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) u0[k] = ub[k * NC + IDX(i, j)];
  }
  // interior of U^{n+1/2} (imposed when the block is in the source list)
  if (tid < 256) {
    const int i = tid % 16 + 1, j = tid / 16 + 1;
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = qb[k * NC + IDX(i, j)];
    if (imp && in_jet(p, L.level, L.icx, L.icy, i, j)) jet_state<NEQ>(p, vjet, uu);
    uprim<NEQ>(p, q, uu, T);
#pragma unroll
    for (int k = 0; k < NEQ; ++k) prim[k * NC + IDX(i, j)] = q[k];
  }
  double gh[NEQ];
  int gi = 0, gj = 0;
  if (ghost) {
    ghost_gather<NEQ>(p, Q, binfo, op, b, L.level, L.icx, L.icy, side, layer, pos, vjet, gh);
    ghost_coords(side, layer, pos, gi, gj);
    double q[NEQ], T;
    uprim<NEQ>(p, q, gh, T);
#pragma unroll
    for (int k = 0; k < NEQ; ++k) prim[k * NC + IDX(gi, gj)] = q[k];
  }
  __syncthreads();
  int err = flux_sweep<NEQ, 2>(p, prim, F, G);
  __syncthreads();
  const int i = tid % 16 + 1, j = tid / 16 + 1;
  double Sc[NEQ];
  if (tid < 256 && p->cylindrical) cyl_source<NEQ>(p, prim, i, j, L.icy, L.level, Sc);
  // the viscosity's stencil (src/viscosity.f90:6-25, on the finest two
  // levels, src/tstep.f90:82) reads Uⁿ of the interior and of ghost layer 1
  // from the primitive tile, which is free after the sweep
  const bool visc = L.level >= p->nlevs - 1;
  if (visc) {
    __syncthreads();   // the sweep's and the sources' reads of the primitives are done
    if (tid < 256)
#pragma unroll
      for (int k = 0; k < NEQ; ++k) prim[k * NC + IDX(i, j)] = u0[k];
    for (int g = tid; g < 64; g += blockDim.x) {
      int ri, rj;
      ghost_coords(g >> 4, 1, (g & 15) + 1, ri, rj);
#pragma unroll
      for (int k = 0; k < NEQ; ++k) prim[k * NC + IDX(ri, rj)] = ub[k * NC + IDX(ri, rj)];
    }
    __syncthreads();
  }
  double un[NEQ];
  if (tid < 256) {   // one thread per interior cell
    const double dt = S->dt;
    const double dtdx = dt / p->dx[L.level - 1];
    const double dtdy = dt / p->dy[L.level - 1];
    const double eta = p->eta;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) {
      const double *uk = prim + k * NC;
      double v = u0[k] - dtdx * (F[k * 272 + (j - 1) * 17 + i] - F[k * 272 + (j - 1) * 17 + i - 1])
                 - dtdy * (G[k * 272 + j * 16 + (i - 1)] - G[k * 272 + (j - 1) * 16 + (i - 1)]);
      if (p->cylindrical) v = v + dt * Sc[k];
      if (visc)   // no dt factor
        v = v + eta * (uk[IDX(i + 1, j)] + uk[IDX(i - 1, j)] + uk[IDX(i, j + 1)] + uk[IDX(i, j - 1)] - 4. * u0[k]);
      un[k] = v;
    }
  }
  // each thread writes the one interior cell of Uⁿ it read (the stencil's
  // reads of other cells were from the tile in LDS)
  if (tid < 256)
#pragma unroll
    for (int k = 0; k < NEQ; ++k) ub[k * NC + IDX(i, j)] = un[k];
  // the ring of U^{n+1/2} goes to its own tile's ring in the U^{n+1/2}
  // buffer (Qring is Q: this launch reads only interiors of Q), where the
  // fill leaves layer 2 as the step's final value (src/tstep.f90:85 copies
  // up's ring into u; boundaryI rewrites layer 1 only)
  if (ghost) {
    double *rb = Qring + (size_t)b * NEQ * NC;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) rb[k * NC + IDX(gi, gj)] = gh[k];
  }
  if (err) atomicOr(&S->err, WX_EHLLC);
  __syncthreads();   // the next queue entry reuses the LDS tiles
   if (!PERSIST) break;   // one entry per workgroup: no loop
  }
}

// ---------------------------------------------------------------------
// Block statistics: CFL partials (src/timestep.f90:29-41) and the
// refinement flag (src/criteria.f90:16-50) from a tile of primitives.
// Indexed by block ID: stat[2id] = min dx/(|u|+cs), dy/(|v|+cs);
// stat[2id+1] = DMC t_cool min; flag[id]: 1 refine, 2 coarsen, 0 neither.
template <int NEQ>
__device__ void block_stats(const wx_params *__restrict__ p, const double *prim, const LeafRec &L,
                            double *stat, int8_t *flag, const FlagOut &fo) {
  __shared__ double red[3][16];
  const int tid = threadIdx.x;
  const int i = tid % 16 + 1, j = tid / 16 + 1;
  const int lev = L.level;
  double dtp = 1.e30, dtc = 1e30, grad = 0.;
  if (tid < 256) {   // one thread per interior cell
    const double rho = prim[0 * NC + IDX(i, j)], vx = prim[1 * NC + IDX(i, j)];
    const double vy = prim[2 * NC + IDX(i, j)], pr = prim[3 * NC + IDX(i, j)];
    const double cs = sqrt(p->gamma * pr / rho);
    dtp = dmin(dtp, p->dx[lev - 1] / (fabs(vx) + cs));
    dtp = dmin(dtp, p->dy[lev - 1] / (fabs(vy) + cs));
    if (p->cooling == 2) {
      double T = (pr / rho) * p->Tempsc;
      if (T > 1e4) {
        double al = cooldmc(p, T) * (rho * rho);
        dtc = dmin(dtc, p->cv * pr / al);
      }
    }
    const double am = dmax(1e-15, fabs(pr));
    const double am2 = dmax(1e-15, fabs(rho));
    const double *P4 = prim + 3 * NC, *P1 = prim;
    // the reference's max of the four quotients |dP|/am is the quotient of
    // the largest |dP| (a correctly rounded division by am > 0 is monotonic),
    // bit for bit, at one division instead of four
    const double dp = dmax(dmax(dmax(fabs(P4[IDX(i + 1, j)] - pr), fabs(P4[IDX(i - 1, j)] - pr)),
                                fabs(P4[IDX(i, j + 1)] - pr)), fabs(P4[IDX(i, j - 1)] - pr));
    const double dr = dmax(dmax(dmax(fabs(P1[IDX(i + 1, j)] - rho), fabs(P1[IDX(i - 1, j)] - rho)),
                                fabs(P1[IDX(i, j + 1)] - rho)), fabs(P1[IDX(i, j - 1)] - rho));
    grad = dmax(dp / am, dr / am2);
  }
  for (int o = 32; o > 0; o >>= 1) {
    dtp = dmin(dtp, __shfl_xor(dtp, o));
    dtc = dmin(dtc, __shfl_xor(dtc, o));
    grad = dmax(grad, __shfl_xor(grad, o));
  }
  const int w = tid >> 6;
  if ((tid & 63) == 0) { red[0][w] = dtp; red[1][w] = dtc; red[2][w] = grad; }
  __syncthreads();
  if (tid == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      dtp = dmin(dtp, red[0][k]);
      dtc = dmin(dtc, red[1][k]);
      grad = dmax(grad, red[2][k]);
    }
    stat[2 * L.id] = dtp;
    stat[2 * L.id + 1] = dtc;
    int8_t fl = 0;
    if (grad >= p->tol) fl = (lev < p->nlevs) ? 1 : 0;
    else if (grad <= p->tol2) fl = 2;
    flag[L.id] = fl;
    if (fo.flagq) fo.flagq[fo.inst_off[L.inst] + L.lidx] = fl;
  }
}

// ---------------------------------------------------------------------
// boundaryI + cooling (+ CFL / criteria epilogue): src/tstep.f90:92-122.
// Cooling surrogate on the 18x18 cells of a tile (coolingh instances only):
// features per cell, the network per wave on groups of 16 cells (matrix
// cores), then per cell either the network's update or -- outside the
// training hull or when the energy residual exceeds the tolerance -- the
// exact routine.  The network's temperature output is a correction to the
// routine's relaxation law T1 = 10 + (T - 10) exp(-ce dt)
// (src/coolingh.f90:249-250) evaluated with the tabulated loss; the energy
// residual is the change of thermal energy (2 - y1) T1 that correction
// makes, relative to the old thermal energy (2 - y0) T.
// The network's weights and biases staged in LDS once per workgroup.
struct SurLds {
  uint16_t w1[64 * 16], w2[64 * 64], w3[64 * 64], w4[16 * 64];
  float b[64 * 3 + 2];
};

template <bool SUR>
struct SurSlot {
  using type = SurLds;
};
template <>
struct SurSlot<false> {
  using type = char;
};

__device__ void sur_stage(const SurDev *__restrict__ M, SurLds *L) {
  for (int t = threadIdx.x; t < 64 * 64 / 4; t += blockDim.x) {   // 8-byte chunks
    reinterpret_cast<uint2 *>(L->w2)[t] = reinterpret_cast<const uint2 *>(M->w2)[t];
    reinterpret_cast<uint2 *>(L->w3)[t] = reinterpret_cast<const uint2 *>(M->w3)[t];
    if (t < 64 * 16 / 4) {
      reinterpret_cast<uint2 *>(L->w1)[t] = reinterpret_cast<const uint2 *>(M->w1)[t];
      reinterpret_cast<uint2 *>(L->w4)[t] = reinterpret_cast<const uint2 *>(M->w4)[t];
    }
  }
  for (int t = threadIdx.x; t < 64; t += blockDim.x) {
    L->b[t] = M->theta[SUR_B1 + t];
    L->b[64 + t] = M->theta[SUR_B2 + t];
    L->b[128 + t] = M->theta[SUR_B3 + t];
  }
  if (threadIdx.x < 2) L->b[192 + threadIdx.x] = M->theta[SUR_B4 + threadIdx.x];
}

template <int NEQ>
__device__ void cool_surrogate(const wx_params *__restrict__ p, SurDev *__restrict__ M, const SurLds *W, double dts,
                               double *tile) {
  __shared__ float sx[324 * 4];
  __shared__ float sy[324 * 2];
  __shared__ float scedt[324];
  __shared__ uint8_t sinh[324];
  __shared__ unsigned int nfb;
  const int tid = threadIdx.x;
  if (tid == 0) nfb = 0u;
  for (int c = tid; c < 324; c += blockDim.x) {
    const int i = c % 18, j = c / 18;
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = tile[k * NC + IDX(i, j)];
    uprim<NEQ>(p, q, uu, T);
    float f[4], cedt;
    sinh[c] = sur_features(M, uu[0], T, uu[4] / uu[0], dts, f, cedt) ? 1 : 0;
    scedt[c] = cedt;
#pragma unroll
    for (int k = 0; k < 4; ++k) sx[4 * c + k] = (f[k] - M->in_mu[k]) / M->in_sig[k];
  }
  __syncthreads();
  {
    const int l = tid & 63, r = l & 15, qq = l >> 4;
    for (int g = tid >> 6; g < 21; g += (int)(blockDim.x >> 6)) {   // 21 groups of 16 cover the 324 cells
      const int c = 16 * g + r;
      bf16x4 xb = {0, 0, 0, 0};
      if (qq == 0 && c < 324) xb = pack4(sx[4 * c], sx[4 * c + 1], sx[4 * c + 2], sx[4 * c + 3]);
      const f32x4 y = mlp_fwd16(W->w1, W->w2, W->w3, W->w4, W->b, W->b + 64, W->b + 128, W->b + 192, xb);
      if (qq == 0 && c < 324) {
        sy[2 * c] = y[0];
        sy[2 * c + 1] = y[1];
      }
    }
  }
  __syncthreads();
  unsigned int mine = 0;
  for (int c = tid; c < 324; c += blockDim.x) {
    const int i = c % 18, j = c / 18;
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = tile[k * NC + IDX(i, j)];
    uprim<NEQ>(p, q, uu, T);
    const double y0 = uu[4] / uu[0];
    const float corr = sy[2 * c] * M->out_sig[0] + M->out_mu[0];
    const float dx = sy[2 * c + 1] * M->out_sig[1] + M->out_mu[1];
    double y1 = y0 - (double)dx;
    y1 = dmin(y1, 1.);
    y1 = dmax(y1, 0.0001);
    // the relaxation law with the tabulated loss, then the network's
    // correction log10(T1 / T1_table)
    double tt = 10. + (T - 10.) * (double)__expf(-scedt[c]);
    tt = dmax(tt, 0.1 * T);
    tt = dmin(tt, 10. * T);
    tt = dmax(tt, 10.);
    double t1 = tt * (double)exp10f(corr);
    t1 = dmax(t1, 0.1 * T);
    t1 = dmin(t1, 10. * T);
    t1 = dmax(t1, 10.);
    const double resid = (2. - y1) * fabs(t1 - tt) / ((2. - y0) * T);
    if (!sinh[c] || !(resid <= (double)M->tol)) {
      atomic_h<NEQ>(p, dts, uu);
      ++mine;
    } else {
      uu[4] = y1 * uu[0];
      uu[3] = p->cv * (2. * uu[0] - uu[4]) * t1 / p->Tempsc + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k) tile[k * NC + IDX(i, j)] = uu[k];
  }
  if (mine) atomicAdd(&nfb, mine);
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&M->fallbacks, (unsigned long long)nfb);
    atomicAdd(&M->cells, 324ull);
  }
}

// Cell c of the fill's per-cell passes: c < 256 the interior cell
// (c % 16 + 1, c / 16 + 1); 256 <= c < 320 the layer-1 ghost that thread
// c - 256 gathers (side (c - 256) / 16, position (c - 256) % 16 + 1).
__device__ __forceinline__ void cool_cell(int c, int &i, int &j) {
  if (c < 256) {
    i = c % 16 + 1;
    j = c / 16 + 1;
  } else {
    const int g = c - 256;
    ghost_coords(g >> 4, 1, (g & 15) + 1, i, j);
  }
}

// Waves per SIMD the plain one-entry-per-workgroup fill is compiled for: 5
// caps it at 96 VGPRs (nothing spilled on its hot path), which measured
// faster than 4 waves at 109 VGPRs (16.8 vs 17.8 ns per block on the
// 16-variant cooling jet; 6 waves spill).  The persistent form carries the
// next record as well and keeps 1024 / BS.
constexpr int kFillWaves = 5;

template <int NEQ, bool SUR, int BS, bool PERSIST>
__global__ __launch_bounds__(BS, (SUR ? 2 : (PERSIST ? 1024 / BS : kFillWaves))) void k_fillI_cool(const wx_params *__restrict__ P, QueueRef q,
                                                    const int2 *__restrict__ ops,
                                                    const BlockRec *__restrict__ binfo,
                                                    const double *__restrict__ Uin,
                                                    double *__restrict__ Uout,
                                                    double *__restrict__ stat,
                                                    int8_t *__restrict__ flag,
                                                    DevScalars *__restrict__ Sa,
                                                    FlagOut fo,
                                                    SurDev *__restrict__ M,
                                                    double *__restrict__ Pc) {
  __shared__ double tile[NEQ * NC];
  // the network in LDS, staged once per workgroup (no LDS without the surrogate)
  __shared__ __align__(16) typename SurSlot<SUR>::type Wbuf;
  SurLds *W = nullptr;
  if constexpr (SUR) {
    W = &Wbuf;
    sur_stage(M, W);
    __syncthreads();
  }
  const int b0 = PERSIST ? (int)blockIdx.x : xcd_entry(blockIdx.x, gridDim.x);
  LeafRec Lnext = q.rec[min(b0, q.cap - 1)];
  const int nleaf = q.hdr->nleaf;
  // the scalars are final for the step (dt reduction and stages A and B
  // are done; this kernel does not write them)
  if (fo.flagq && blockIdx.x == 0)
    for (int t = threadIdx.x; t < fo.ninst; t += blockDim.x) fo.Sout[t] = Sa[t];
  for (int l = b0; l < nleaf; l += gridDim.x) {
  const int r = l;
  const LeafRec L = Lnext;
  if (PERSIST && l + (int)gridDim.x < nleaf) Lnext = q.rec[l + gridDim.x];
  const wx_params *__restrict__ p = P + L.inst;
  DevScalars *__restrict__ S = Sa + L.inst;
  const int b = L.id;
  const int tid = threadIdx.x;
  const double time = S->time;
  const bool imp = (L.flags & 1) != 0;
  const double vjet = p->otherb ? jet_velocity_at(p, S, time) : 0.0;
  const double *ib = Uin + (size_t)b * NEQ * NC;
  double *ob = Uout + (size_t)b * NEQ * NC;
  // interior U^{n+1}; imposed where listed
  for (int c = tid; c < 256; c += blockDim.x) {
    const int i = c % 16 + 1, j = c / 16 + 1;
    double uu[NEQ];
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = ib[k * NC + IDX(i, j)];
    if (imp && in_jet(p, L.level, L.icx, L.icy, i, j)) jet_state<NEQ>(p, vjet, uu);
#pragma unroll
    for (int k = 0; k < NEQ; ++k) tile[k * NC + IDX(i, j)] = uu[k];
  }
  // ghost layer 1 (64 cells) from the list; a cell no entry writes keeps the
  // U^{n+1/2} ring that stage B left in this tile's ring of Uout, imposed
  // where listed.  Ghost g is gathered by the thread that cools it below:
  // thread g with 256 threads, thread 256 + g with more.
  const int gt = BS > 256 ? tid - 256 : tid;
  if (gt >= 0 && gt < 64) {
    const int side = gt >> 4, pos = (gt & 15) + 1;
    const int2 op = ops[r * 8 + side * 2 + ((pos - 1) >> 3)];
    int gi, gj;
    ghost_coords(side, 1, pos, gi, gj);
    double g[NEQ];
    if ((op.y & 15) != GK_NONE) {
      ghost_gather<NEQ>(p, Uin, binfo, op, b, L.level, L.icx, L.icy, side, 1, pos, vjet, g);
    } else {
#pragma unroll
      for (int k = 0; k < NEQ; ++k) g[k] = ob[k * NC + IDX(gi, gj)];
      if (imp && in_jet(p, L.level, L.icx, L.icy, gi, gj)) jet_state<NEQ>(p, vjet, g);
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k) tile[k * NC + IDX(gi, gj)] = g[k];
  }
  // ghost layer 2 keeps stage B's U^{n+1/2} ring (already in Uout); the
  // source is imposed on its cells where listed (before the cooling, which
  // writes other cells, so that the source speed is not held across it)
  if (imp)
    for (int g = tid; g < 64; g += blockDim.x) {
      int gi, gj;
      ghost_coords(g >> 4, 2, (g & 15) + 1, gi, gj);
      if (in_jet(p, L.level, L.icx, L.icy, gi, gj)) {
        double v[NEQ];
        jet_state<NEQ>(p, vjet, v);
#pragma unroll
        for (int k = 0; k < NEQ; ++k) ob[k * NC + IDX(gi, gj)] = v[k];
      }
    }
  // cooling over cells 0..n+1 (src/coolingh.f90:19-20, src/coolingdmc.f90:23-24)
  // except the four corner cells, which nothing reads: the interior (cell
  // c < 256) and the non-corner ghosts of layer 1 (c = 256 + the gather's
  // index).  Every cell is cooled by the thread that loaded it, so no wave
  // waits for another's loads.
  // Each cooled cell goes to Uout, and its primitives to the LDS tile for
  // the epilogue's 5-point stencils.
  const int mode = p->cooling;
  bool done = false;
  if constexpr (SUR && NEQ >= 5) {
    if (mode == 1) {
      __syncthreads();   // the surrogate's cells are spread differently
      cool_surrogate<NEQ>(p, M, W, S->dt * p->tsc, tile);
      __syncthreads();
      done = true;
    }
  }
  const double dts = S->dt * p->tsc;
  // cell c of this thread: cooled state to Uout, its primitives to the tile
  // (under DMC with the cooled pressure, carried into the next step,
  // src/coolingdmc.f90:27-38)
  auto cool_pass = [&](auto cool) {
    for (int c = tid; c < 320; c += blockDim.x) {
      int i, j;
      cool_cell(c, i, j);
      double uu[NEQ];
#pragma unroll
      for (int k = 0; k < NEQ; ++k) uu[k] = tile[k * NC + IDX(i, j)];
      const double pcool = cool(uu);
#pragma unroll
      for (int k = 0; k < NEQ; ++k) ob[k * NC + IDX(i, j)] = uu[k];
      double q[NEQ], T;
      uprim<NEQ>(p, q, uu, T);
      if (mode == 2) {
        q[3] = pcool;
        if (Pc) Pc[(size_t)L.id * NC + IDX(i, j)] = pcool;
      }
#pragma unroll
      for (int k = 0; k < NEQ; ++k) tile[k * NC + IDX(i, j)] = q[k];
    }
  };
  // one loop per cooling mode, so that what the compiler hoists out of one
  // mode's loop (constants of DMC's log10, kept in scratch memory at this
  // kernel's register budget) is set up only when that mode runs
  if (!done && mode == 1) {
    cool_pass([&](double *uu) {
      atomic_h<NEQ>(p, dts, uu);
      return 0.0;
    });
  } else if (!done && mode == 2) {
    cool_pass([&](double *uu) {
      double pcool;
      cool_dmc<NEQ>(p, dts, uu, pcool);
      return pcool;
    });
  } else {
    cool_pass([](double *) { return 0.0; });
  }
  __syncthreads();
  block_stats<NEQ>(p, tile, L, stat, flag, fo);
  __syncthreads();   // the next queue entry reuses the LDS tile
   if (!PERSIST) break;   // one entry per workgroup: no loop
  }
}

// Statistics only (after upload or a mesh change).
template <int NEQ>
// Under DMC (Pc != nullptr, the instance cooling with DMC) the pressure is
// the carried one, or (pc_init 1, or 2 and the record's pad bit 0) the
// block's state came from outside and the carried pressure starts as the
// one u gives (the reference's calcprim after basegrid or a warm start).
__global__ __launch_bounds__(256) void k_blockstat(const wx_params *__restrict__ P,
                                                   const LeafRec *__restrict__ leaves,
                                                   const double *__restrict__ U,
                                                   double *__restrict__ stat,
                                                   int8_t *__restrict__ flag,
                                                   double *__restrict__ Pc, int pc_init) {
  __shared__ double tile[NEQ * NC];
  const int l = blockIdx.x;
  const LeafRec L = leaves[l];
  const wx_params *__restrict__ p = P + L.inst;
  const double *ub = U + (size_t)L.id * NEQ * NC;
  const bool carried = Pc && p->cooling == 2;
  const bool init = pc_init == 1 || (pc_init == 2 && (L.pad & 1));
  double *pcb = carried ? Pc + (size_t)L.id * NC : nullptr;
  for (int c = threadIdx.x; c < 18 * 18; c += blockDim.x) {
    const int i = c % 18, j = c / 18;
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = ub[k * NC + IDX(i, j)];
    uprim<NEQ>(p, q, uu, T);
    if (carried) {
      if (init) pcb[IDX(i, j)] = q[3];
      else q[3] = pcb[IDX(i, j)];
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k) tile[k * NC + IDX(i, j)] = q[k];
  }
  __syncthreads();
  block_stats<NEQ>(p, tile, L, stat, flag, FlagOut{nullptr, nullptr, nullptr, 0});
}

// dt = min(cfl*dtp, 0.4*dtcool*Psc/tsc) (src/timestep.f90:50); optionally
// time += dt (src/main.f90:88).  Workgroup i reduces instance i's records.
__global__ __launch_bounds__(1024) void k_dtreduce(const wx_params *__restrict__ P,
                                                   const LeafRec *__restrict__ rec,
                                                   const double *__restrict__ stat,
                                                   const int32_t *__restrict__ inst_off,
                                                   DevScalars *__restrict__ Sa, int advance) {
  __shared__ double r0[16], r1[16];
  const int inst = blockIdx.x;
  const wx_params *__restrict__ p = P + inst;
  DevScalars *__restrict__ S = Sa + inst;
  const int n = inst_off[inst + 1] - inst_off[inst];
  const LeafRec *__restrict__ mine = rec + inst_off[inst];
  double a = 1.e30, c = 1e30;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int id = mine[k].id;
    a = dmin(a, stat[2 * id]);
    c = dmin(c, stat[2 * id + 1]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = dmin(a, __shfl_xor(a, o));
    c = dmin(c, __shfl_xor(c, o));
  }
  if ((threadIdx.x & 63) == 0) { r0[threadIdx.x >> 6] = a; r1[threadIdx.x >> 6] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 1; k < nw; ++k) { a = dmin(a, r0[k]); c = dmin(c, r1[k]); }
    S->dtp_min = a;
    S->dtcool_min = c;
    double dt = dmin(p->cfl * a, 0.4 * c * p->Psc / p->tsc);
    S->dt = dt;
    if (advance) S->time = S->time + dt;
    if (p->otherb) {   // the jet source speed of the step, once instead of per wave of the stages
      const double t = S->time;
      S->vjet = jet_velocity(p, t);
      S->vjet_time = t;
      S->vjet_ok = 1;
    }
  }
}

// Compacted queue entry l = record leafbase[i] + l - inst_off[i] of the
// per-instance regions, i the instance with inst_off[i] <= l < inst_off[i+1]
// (binary search); copies the record and its 8 + 8 ghost ops.
__global__ __launch_bounds__(256) void k_build_queue(const QueueHdr *__restrict__ hdr,
                                                     const int32_t *__restrict__ inst_off,
                                                     const int32_t *__restrict__ leafbase,
                                                     const int32_t *__restrict__ slotbase,
                                                     const int32_t *__restrict__ reg_ids,
                                                     const BlockRec *__restrict__ binfo,
                                                     const int32_t *__restrict__ slot_flag,
                                                     const int2 *__restrict__ slot_ops2,
                                                     const int2 *__restrict__ slot_ops1,
                                                     LeafRec *__restrict__ rec, int2 *__restrict__ ops2,
                                                     int2 *__restrict__ ops1) {
  const int n = hdr->nleaf, ni = hdr->ninst;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 8 * n; t += gridDim.x * blockDim.x) {
    const int l = t >> 3, k = t & 7;
    int lo = 0, hi = ni - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (inst_off[mid] <= l) lo = mid;
      else hi = mid - 1;
    }
    const int lp = l - inst_off[lo];
    const size_t slot = (size_t)slotbase[lo] + (size_t)reg_ids[(size_t)leafbase[lo] + lp];
    if (k == 0) {
      const BlockRec b = binfo[slot];
      rec[l] = LeafRec{(int32_t)slot, b.level, b.icx, b.icy, slot_flag[slot], lp, lo, 0};
    }
    ops2[(size_t)l * 8 + k] = slot_ops2[slot * 8 + k];
    if (ops1) ops1[(size_t)l * 8 + k] = slot_ops1[slot * 8 + k];
  }
}

// The changed per-slot ops (and, with flag != nullptr, leaf flags) of an
// upload, read in place from the pinned staging.
__global__ __launch_bounds__(256) void k_slot_scatter(const SlotOps *__restrict__ src, int n,
                                                      int2 *__restrict__ ops, int32_t *__restrict__ flag) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 8 * n; t += gridDim.x * blockDim.x) {
    const int r = t >> 3, k = t & 7;
    const size_t slot = (size_t)src[r].slot;
    ops[slot * 8 + k] = src[r].ops[k];
    if (flag && k == 0) flag[slot] = src[r].flag;
  }
}

// Refinement flags in queue order for the host's admesh, and the instances'
// scalars; both outputs may be pinned host memory (mapped), so that the host
// reads them after the stream synchronises with no copy launched.
__global__ __launch_bounds__(256) void k_gather_flags(QueueRef q, const int32_t *__restrict__ inst_off,
                                                      const int8_t *__restrict__ flag,
                                                      int8_t *__restrict__ flagq, const DevScalars *__restrict__ S,
                                                      DevScalars *__restrict__ Sout, int ninst) {
  const int n = q.hdr->nleaf;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < n; l += gridDim.x * blockDim.x) {
    const LeafRec r = q.rec[l];
    flagq[inst_off[r.inst] + r.lidx] = flag[r.id];   // by leaf position, whatever the queue order
  }
  if (Sout && blockIdx.x == 0)
    for (int i = threadIdx.x; i < ninst; i += blockDim.x) Sout[i] = S[i];
}

// ---------------------------------------------------------------------
// admesh data ops: src/refine.f90:30-42 (injection over the padded block)
// and src/coarse.f90:182-366 (MPIP branch: 2x2 means incl. ghosts, then
// outer ghost copies).  One workgroup per op.
template <int NEQ>
__global__ __launch_bounds__(256) void k_meshops(const wx_params *__restrict__ P, const MeshOp *__restrict__ ops,
                                                 double *__restrict__ U, double *__restrict__ stat,
                                                 int8_t *__restrict__ flag, double *__restrict__ Pc) {
  // workgroup 4 * o + s: son s of a refine op o; a coarse op's dad is done
  // by its first workgroup
  const MeshOp op = ops[blockIdx.x >> 2];
  const int part = blockIdx.x & 3;
  const int tid = threadIdx.x;
  __shared__ double d[NEQ * NC];   // coarse: the dad tile; then a tile of primitives for the statistics
  if (op.kind == 0) {
    const double *pb = U + (size_t)op.a * NEQ * NC;
    const int s = part;
    const int son = s == 0 ? op.s1 : s == 1 ? op.s2 : s == 2 ? op.s3 : op.s4;
    double *sb = U + (size_t)son * NEQ * NC;
    for (int cc = tid; cc < NC; cc += blockDim.x) {
      const int i = cc % NP - 1, j = cc / NP - 1;
      const int si = (i + 1) / 2 + ((s & 1) ? NX / 2 : 0);
      const int sj = (j + 1) / 2 + ((s & 2) ? NY / 2 : 0);
#pragma unroll
      for (int k = 0; k < NEQ; ++k) sb[k * NC + cc] = pb[k * NC + IDX(si, sj)];
    }
    // statistics of the son from its cells 0..17, which are the parent's
    // cells it was injected from (read back from the parent, not the son);
    // under DMC its carried pressure starts from them (src/refine.f90:231-258
    // recomputes the sons' primit)
    const wx_params *__restrict__ p = P + op.inst;
    double *pcb = (Pc && p->cooling == 2) ? Pc + (size_t)son * NC : nullptr;
    if (!stat && !pcb) return;
    for (int c = tid; c < 18 * 18; c += blockDim.x) {
      const int i = c % 18, j = c / 18;
      const int si = (i + 1) / 2 + ((s & 1) ? NX / 2 : 0);
      const int sj = (j + 1) / 2 + ((s & 2) ? NY / 2 : 0);
      double uu[NEQ], q[NEQ], T;
#pragma unroll
      for (int k = 0; k < NEQ; ++k) uu[k] = pb[k * NC + IDX(si, sj)];
      uprim<NEQ>(p, q, uu, T);
      if (pcb) pcb[IDX(i, j)] = q[3];
#pragma unroll
      for (int k = 0; k < NEQ; ++k) d[k * NC + IDX(i, j)] = q[k];
    }
    if (!stat) return;
    __syncthreads();
    LeafRec L{};
    L.id = son;
    L.level = op.level;
    L.inst = op.inst;
    block_stats<NEQ>(p, d, L, stat, flag, FlagOut{nullptr, nullptr, nullptr, 0});
  } else {
    if (part != 0) return;
    const int sons[4] = {op.s1, op.s2, op.s3, op.s4};
    // start from the current dad tile (corners the rules below do not set)
    const double *dbg = U + (size_t)op.a * NEQ * NC;
    for (int c = tid; c < NEQ * NC; c += blockDim.x) d[c] = dbg[c];
    __syncthreads();
    // dad cells (0..17)^2 from the four sons: son s covers i in [is0, is1]
    for (int c = tid; c < 18 * 18; c += blockDim.x) {
      const int i = c % 18, j = c / 18;
      const int sx = (i <= NX / 2) ? 0 : 1, sy = (j <= NY / 2) ? 0 : 1;
      const int s = sx + 2 * sy;
      const int li = i - sx * (NX / 2), lj = j - sy * (NY / 2);
      const double *sb = U + (size_t)sons[s] * NEQ * NC;
      const int i1 = 2 * li - 1, j1 = 2 * lj - 1;
#pragma unroll
      for (int k = 0; k < NEQ; ++k)
        d[k * NC + IDX(i, j)] = 0.25 * (((sb[k * NC + IDX(i1, j1)] + sb[k * NC + IDX(i1 + 1, j1)])
                                        + sb[k * NC + IDX(i1, j1 + 1)]) + sb[k * NC + IDX(i1 + 1, j1 + 1)]);
    }
    __syncthreads();
    // outer ghost copies (src/coarse.f90:349-358), same order
    if (tid < NEQ) {
      double *dk = d + tid * NC;
      for (int j = -1; j <= NY + 2; ++j) dk[IDX(-1, j)] = dk[IDX(0, j)];
      for (int j = -1; j <= NY + 2; ++j) dk[IDX(NX + 2, j)] = dk[IDX(NX + 1, j)];
      for (int i = -1; i <= NX + 2; ++i) dk[IDX(i, -1)] = dk[IDX(i, 0)];
      for (int i = -1; i <= NX + 2; ++i) dk[IDX(i, NY + 2)] = dk[IDX(i, NY + 1)];
      dk[IDX(-1, -1)] = dk[IDX(0, -1)];
      dk[IDX(NX + 2, -1)] = dk[IDX(NX + 2, 0)];
      dk[IDX(-1, NY + 2)] = dk[IDX(-1, NY + 1)];
      dk[IDX(NX + 2, NY + 2)] = dk[IDX(NX + 1, NY + 2)];
    }
    __syncthreads();
    double *db = U + (size_t)op.a * NEQ * NC;
    for (int c = tid; c < NEQ * NC; c += blockDim.x) db[c] = d[c];
    // statistics of the dad: primitives of its cells 0..17 in place; under
    // DMC its carried pressure starts from them (src/coarse.f90:178, 360
    // recompute the dad's primit)
    const wx_params *__restrict__ p = P + op.inst;
    double *pcb = (Pc && p->cooling == 2) ? Pc + (size_t)op.a * NC : nullptr;
    if (!stat && !pcb) return;
    __syncthreads();
    for (int c = tid; c < 18 * 18; c += blockDim.x) {
      const int i = c % 18, j = c / 18;
      double uu[NEQ], q[NEQ], T;
#pragma unroll
      for (int k = 0; k < NEQ; ++k) uu[k] = d[k * NC + IDX(i, j)];
      uprim<NEQ>(p, q, uu, T);
      if (pcb) pcb[IDX(i, j)] = q[3];
#pragma unroll
      for (int k = 0; k < NEQ; ++k) d[k * NC + IDX(i, j)] = q[k];
    }
    if (!stat) return;
    __syncthreads();
    LeafRec L{};
    L.id = op.a;
    L.level = op.level;
    L.inst = op.inst;
    block_stats<NEQ>(p, d, L, stat, flag, FlagOut{nullptr, nullptr, nullptr, 0});
  }
}

// Layout transposes between [id][var][j][i] and the reference's
// u(nb, ieq, i, j) with nb fastest (ld = number of block rows).
template <int NEQ>
__global__ void k_transpose(int cap, const double *__restrict__ src, double *__restrict__ dst,
                            int to_fortran, int ld) {
  const size_t n = (size_t)cap * NEQ * NC;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    if (to_fortran) {
      // t enumerates the Fortran array: nb fastest
      const size_t nb = t % cap, r = t / cap;
      const size_t k = r % NEQ, cell = r / NEQ;     // cell = (i+1) + 20*(j+1)
      dst[t] = src[((nb + 1) * NEQ + k) * NC + cell];
    } else {
      const size_t cell = t % NC, r = t / NC;
      const size_t k = r % NEQ, id = r / NEQ;       // id 0..cap-1 -> block id+1
      dst[((id + 1) * NEQ + k) * NC + cell] = src[(cell * NEQ + k) * (size_t)ld + id];
    }
  }
}

template <int NEQ>
__global__ void k_calcprim(const wx_params *__restrict__ p, int cap, const double *__restrict__ U,
                           double *__restrict__ prim) {
  const size_t n = (size_t)(cap + 1) * NC;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const size_t id = t / NC, c = t % NC;
    double uu[NEQ], q[NEQ], T;
#pragma unroll
    for (int k = 0; k < NEQ; ++k) uu[k] = U[(id * NEQ + k) * NC + c];
    if (uu[0] == 0.0 && uu[3] == 0.0) {
#pragma unroll
      for (int k = 0; k < NEQ; ++k) q[k] = 0.0;
    } else {
      uprim<NEQ>(p, q, uu, T);
    }
#pragma unroll
    for (int k = 0; k < NEQ; ++k) prim[(id * NEQ + k) * NC + c] = q[k];
  }
}

// ---------------------------------------------------------------------
// launch wrappers
#define WX_DISPATCH(NEQV, CALL)                  \
  switch (NEQV) {                                \
    case 4: { constexpr int N = 4; CALL; } break; \
    case 5: { constexpr int N = 5; CALL; } break; \
    case 6: { constexpr int N = 6; CALL; } break; \
    default: return WX_EARG;                     \
  }

static inline int hip_status() { return hipGetLastError() == hipSuccess ? WX_OK : WX_EHIP; }

int launch_build_queue(const QueueHdr *hdr, const int32_t *inst_off, const int32_t *leafbase,
                       const int32_t *slotbase, const int32_t *reg_ids, const BlockRec *binfo,
                       const int32_t *slot_flag, const int2 *slot_ops2, const int2 *slot_ops1, LeafRec *rec,
                       int2 *ops2, int2 *ops1, int grid, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  hipLaunchKernelGGL(k_build_queue, dim3(grid), dim3(256), 0, stream, hdr, inst_off, leafbase, slotbase, reg_ids,
                     binfo, slot_flag, slot_ops2, slot_ops1, rec, ops2, ops1);
  return hip_status();
}

int launch_slot_scatter(const SlotOps *src, int n, int2 *ops, int32_t *flag, hipStream_t stream) {
  if (n <= 0) return WX_OK;
  const int grid = min(max(1, (8 * n + 255) / 256), 1024);
  hipLaunchKernelGGL(k_slot_scatter, dim3(grid), dim3(256), 0, stream, src, n, ops, flag);
  return hip_status();
}

// Workgroup sizes the stage kernels are built for (4, 6 or 8 waves), and
// the one-entry-per-workgroup / persistent forms.
#define WX_BS_DISPATCH(BSV, CALL)                    \
  switch (BSV) {                                     \
    case 256: { constexpr int B = 256; CALL; } break; \
    case 384: { constexpr int B = 384; CALL; } break; \
    case 512: { constexpr int B = 512; CALL; } break; \
    default: return WX_EARG;                         \
  }
#define WX_PERSIST_DISPATCH(PV, CALL)                     \
  if (PV) { constexpr bool PS = true; CALL; }             \
  else { constexpr bool PS = false; CALL; }

int launch_stageA(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const double *U,
                  double *Q, const double *Pc, DevScalars *S, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  WX_PERSIST_DISPATCH(persist, WX_BS_DISPATCH(bs, WX_DISPATCH(neq, hipLaunchKernelGGL((k_stageA<N, B, PS>), dim3(grid), dim3(B), 0, stream, P, q, U, Q, Pc, S))));
  return hip_status();
}

int launch_stageB(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                  const BlockRec *binfo, double *U, double *Q, DevScalars *S, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  WX_PERSIST_DISPATCH(persist, WX_BS_DISPATCH(bs, WX_DISPATCH(neq, hipLaunchKernelGGL((k_stageB<N, B, PS>), dim3(grid), dim3(B), 0, stream, P, q, ops, binfo, U, Q, Q, S))));
  return hip_status();
}

int launch_fillI_cool(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                      const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                      DevScalars *S, FlagOut fo, double *Pc, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  WX_PERSIST_DISPATCH(persist, WX_BS_DISPATCH(bs, WX_DISPATCH(neq, hipLaunchKernelGGL((k_fillI_cool<N, false, B, PS>), dim3(grid), dim3(B), 0, stream, P, q, ops, binfo, Uin, Uout, stat, flag, S, fo, nullptr, Pc))));
  return hip_status();
}

int launch_fillI_cool_sur(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                          const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                          DevScalars *S, SurDev *M, FlagOut fo, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  switch (neq) {   // coolingh needs u(5) = n_HI
    case 5: WX_PERSIST_DISPATCH(persist, WX_BS_DISPATCH(bs, hipLaunchKernelGGL((k_fillI_cool<5, true, B, PS>), dim3(grid), dim3(B), 0, stream, P, q, ops, binfo, Uin, Uout, stat, flag, S, fo, M, nullptr))); break;
    case 6: WX_PERSIST_DISPATCH(persist, WX_BS_DISPATCH(bs, hipLaunchKernelGGL((k_fillI_cool<6, true, B, PS>), dim3(grid), dim3(B), 0, stream, P, q, ops, binfo, Uin, Uout, stat, flag, S, fo, M, nullptr))); break;
    default: return WX_EARG;
  }
  return hip_status();
}

int launch_blockstat(const wx_params *P, int neq, const LeafRec *leaves, int nleaf, const double *U,
                     double *stat, int8_t *flag, double *Pc, int pc_init, hipStream_t stream) {
  if (nleaf <= 0) return WX_OK;
  WX_DISPATCH(neq, hipLaunchKernelGGL(k_blockstat<N>, dim3(nleaf), dim3(256), 0, stream, P, leaves, U, stat, flag, Pc,
                                      pc_init));
  return hip_status();
}

int launch_dtreduce(const wx_params *P, const LeafRec *rec, const double *stat, const int32_t *inst_off,
                    int ninst, DevScalars *S, int advance, hipStream_t stream) {
  if (ninst <= 0) return WX_OK;
  hipLaunchKernelGGL(k_dtreduce, dim3(ninst), dim3(1024), 0, stream, P, rec, stat, inst_off, S, advance);
  return hip_status();
}

int launch_gather_flags(QueueRef q, const int32_t *inst_off, int grid, const int8_t *flag, int8_t *flagq,
                        const DevScalars *S,
                        DevScalars *Sout, int ninst, hipStream_t stream) {
  if (grid <= 0) return WX_OK;
  hipLaunchKernelGGL(k_gather_flags, dim3(grid), dim3(256), 0, stream, q, inst_off, flag, flagq, S, Sout, ninst);
  return hip_status();
}

int launch_meshops(const wx_params *P, int neq, const MeshOp *ops, int nops, double *U, double *stat, int8_t *flag,
                   double *Pc, hipStream_t stream) {
  if (nops <= 0) return WX_OK;
  WX_DISPATCH(neq, hipLaunchKernelGGL(k_meshops<N>, dim3(4 * nops), dim3(256), 0, stream, P, ops, U, stat, flag, Pc));
  return hip_status();
}

// Upload copies: workgroups (x, y) copy part x of copy y, in 16-byte words
// when both ends and the size allow, else 4-byte words, else bytes.
__global__ __launch_bounds__(256) void k_h2d(const CopyDesc *__restrict__ d, int base) {
  const CopyDesc c = d[base + blockIdx.y];
  const uint64_t al = (uint64_t)c.dst | (uint64_t)c.src | c.bytes;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  if ((al & 15) == 0) {
    const uint4 *s = (const uint4 *)c.src;
    uint4 *o = (uint4 *)c.dst;
    for (size_t i = t0; i < c.bytes / 16; i += st) o[i] = s[i];
  } else if ((al & 3) == 0) {
    const uint32_t *s = (const uint32_t *)c.src;
    uint32_t *o = (uint32_t *)c.dst;
    for (size_t i = t0; i < c.bytes / 4; i += st) o[i] = s[i];
  } else {
    const uint8_t *s = (const uint8_t *)c.src;
    uint8_t *o = (uint8_t *)c.dst;
    for (size_t i = t0; i < c.bytes; i += st) o[i] = s[i];
  }
}

int launch_h2d(const CopyDesc *d, int n, hipStream_t stream) {
  for (int base = 0; base < n; base += 65535) {
    const int m = n - base < 65535 ? n - base : 65535;
    hipLaunchKernelGGL(k_h2d, dim3(16, m), dim3(256), 0, stream, d, base);
  }
  return hip_status();
}

int launch_transpose(int neq, int cap, const double *src, double *dst, int to_fortran, int ld, hipStream_t stream) {
  WX_DISPATCH(neq, hipLaunchKernelGGL(k_transpose<N>, dim3(1024), dim3(256), 0, stream, cap, src, dst, to_fortran, ld));
  return hip_status();
}

int launch_calcprim(const wx_params *P, int neq, int cap, const double *U, double *prim, hipStream_t stream) {
  WX_DISPATCH(neq, hipLaunchKernelGGL(k_calcprim<N>, dim3(1024), dim3(256), 0, stream, P, cap, U, prim));
  return hip_status();
}

}  // namespace wx

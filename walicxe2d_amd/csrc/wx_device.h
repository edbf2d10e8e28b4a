// wx_device.h -- structures shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/wx_capi.h"

namespace wx {

constexpr int NX = 16;
constexpr int NY = 16;
constexpr int NP = 20;      // padded tile extent (nghost = 2)
constexpr int NC = NP * NP; // cells per padded tile

// Ghost-segment op kinds.  A destination block's non-corner ghost ring is 4
// sides x 2 halves ("segments", seg = side*2 + half; side 0 left, 1 right,
// 2 bottom, 3 top; half 0 covers positions 1..8, half 1 positions 9..16).
// Each segment is written by exactly one op per fill (the last writer of
// the reference's ordered list, src/boundaryI.f90:34-396 and
// src/boundaryII.f90:35-514); corners are never read by the scheme.
enum GhostKind : int {
  GK_NONE = 0,      // not written in this fill: keep the carried value
  GK_COPY = 1,      // same level: neighbour interior strip
  GK_RESTRICT = 2,  // finer neighbour: 2x2 means
  GK_PROLONG = 3,   // coarser neighbour: 0th-order injection
  GK_BOXOUT = 4,    // box boundary, outflow: mirror own interior
  GK_BOXREF = 5,    // box boundary, reflective: mirror + normal velocity flip
};
// op.y bit fields
constexpr int GF_SRC_IMPOSED = 1 << 4;   // source interior carried the jet at copy time
constexpr int GF_DST_IMPOSED = 1 << 5;   // destination imposed after this write
constexpr int GF_OFF_SHIFT = 8;          // PROLONG: 0 or 8 along the coarse side

// Per-leaf record: one entry of the block queue.  A workgroup takes queue
// entries l = blockIdx.x, blockIdx.x + gridDim.x, ... while l < nleaf.
struct LeafRec {
  int32_t id;      // tile slot (the block ID; + the instance's slot base in an ensemble)
  int32_t level;
  int32_t icx, icy;
  int32_t flags;   // bit0: block imposed by the jet source in this step's fills
  int32_t lidx;    // position in the instance's leaf list (block-ID order)
  int32_t inst;    // instance: selects wx_params and DevScalars
  int32_t pad;
};

// Queue header in device memory.  Kernels read the leaf count from here, so
// launches (and a captured graph of them) stay valid when the mesh changes.
struct QueueHdr {
  int32_t nleaf;
  int32_t ninst;
  int32_t pad[2];
};

// The compacted block queue as the kernels see it: entries 0 .. hdr->nleaf-1
// of rec (capacity `cap` entries, cap >= 1); the 8 ghost ops per fill of
// entry l are at ops[8 * l].  In an ensemble, k_build_queue gathers the
// instances' record and op regions into these compact arrays after a mesh
// change, so only the regions of instances whose mesh changed are uploaded.
struct QueueRef {
  const QueueHdr *hdr;
  const LeafRec *rec;
  int32_t cap;
};

// Per-ID record for source lookups (level and lower-left cell index).
struct BlockRec {
  int32_t level, icx, icy, pad;
};

// Device scalars of one context.
struct DevScalars {
  double time;
  double dt;
  double dtp_min;
  double dtcool_min;
  double vjet, vjet_time;   // jet source speed at vjet_time, valid when vjet_ok (set by k_dtreduce)
  int32_t err;
  int32_t vjet_ok;
  int32_t mesh_err;         // the device table builder met a tree it cannot read (LBE_* code)
  int32_t pad;
};

// One block of the AMR tree as the device table builder reads it
// (src/globals.f90:13-29): lp(nb, 1..8) = level, father, first son, sibling
// number, left, right, bottom, top neighbour; leaf(nb); icoord(nb, 1..2).
struct TreeRow {
  int32_t lp[8];
  int32_t leaf;
  int32_t icx, icy;
  int32_t pad;
};

// Per block ID, written by the table builder: leaf position (-1: not a
// leaf), position of its first entry in the ordered ghost list, and its
// number of entries.
struct LbInfo {
  int32_t lidx, off, cnt, pad;
};

// Scan partials of the table builder: leaves and list entries of 256 IDs.
struct LbPart {
  int32_t leaves, entries;
};

// mesh_err codes of the table builder
enum LbError : int32_t {
  LBE_POINTER = 1,    // a neighbour or son pointer outside 1..nbmax
  LBE_FAMILY = 2,     // a coarser neighbour's entry pair does not hold the block
  LBE_COUNT = 3,      // the leaf count differs from the host's
};

// Mesh data ops for admesh (src/refine.f90, src/coarse.f90).
struct MeshOp {
  int32_t kind;   // 0 refine (a = parent), 1 coarse (a = dad)
  int32_t a, s1, s2, s3, s4;
  int32_t level, inst;   // level and instance of the new leaves (sons, or the dad)
};

// Where the fill's epilogue also puts the refinement flags and the scalars
// for the host (both in mapped pinned memory): flagq[inst_off[rec.inst] +
// rec.lidx] = the entry's flag, Sout[0 .. ninst) = S.  flagq == nullptr: off.
struct FlagOut {
  const int32_t *inst_off;
  int8_t *flagq;
  DevScalars *Sout;
  int32_t ninst;
};

// One host-to-device copy of an upload: `bytes` from src (the device alias of
// mapped pinned host memory) to dst (device memory).
struct CopyDesc {
  void *dst;
  const void *src;
  uint64_t bytes;
};

// An upload's record of one slot whose ghost ops (or leaf flags) changed:
// the slot, its leaf flags (LeafRec::flags) and its 8 ops.
struct SlotOps {
  int32_t slot, flag;
  int2 ops[8];
};

// launch wrappers (wx_kernels.hip).  P and S are arrays indexed by
// LeafRec::inst (one element for a single mesh); `grid` is the number of
// workgroups of the persistent stage kernels.
// Builds the queue in instance order: entry inst_off[i] + l is leaf l of
// instance i, whose local block ID is reg_ids[leafbase[i] + l] and whose
// slot is slotbase[i] + that ID; its record comes from binfo and slot_flag,
// its ops from the per-slot tables (fill I's only when ops1 is set).
int launch_build_queue(const QueueHdr *hdr, const int32_t *inst_off, const int32_t *leafbase,
                       const int32_t *slotbase, const int32_t *reg_ids, const BlockRec *binfo,
                       const int32_t *slot_flag, const int2 *slot_ops2, const int2 *slot_ops1, LeafRec *rec,
                       int2 *ops2, int2 *ops1, int grid, hipStream_t stream);
// Writes n SlotOps records (src may be mapped pinned memory) into the
// per-slot op table ops (8 per slot) and, when flag is set, the flags.
int launch_slot_scatter(const SlotOps *src, int n, int2 *ops, int32_t *flag, hipStream_t stream);
// bs = threads per workgroup: 256, 384 or 512 (4, 6 or 8 waves).  persist
// = 0: workgroup g takes queue entry g (grid >= queue length); 1: workgroup
// g takes entries g, g + grid, ... (any grid).
// Pc (nullptr unless some instance cools with DMC): per block ID a padded
// tile of the pressure the last DMC cooling left in cells 0..17 (the
// reference's primit(nb,4,:,:), src/coolingdmc.f90:27-38); stage A and the
// statistics read it, the fill writes it, k_meshops starts it for new leaves.
int launch_stageA(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const double *U,
                  double *Q, const double *Pc, DevScalars *S, hipStream_t stream);
int launch_stageB(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                  const BlockRec *binfo, double *U, double *Q, DevScalars *S, hipStream_t stream);
int launch_fillI_cool(const wx_params *P, int neq, QueueRef q, int grid, int bs, int persist, const int2 *ops,
                      const BlockRec *binfo, const double *Uin, double *Uout, double *stat, int8_t *flag,
                      DevScalars *S, FlagOut fo, double *Pc, hipStream_t stream);
// Statistics of an explicit list of leaf records (one workgroup each).
// Under DMC they read the carried pressure, or start it from u: pc_init 1
// for every record, 2 for the records whose pad has bit 0 set, 0 for none.
int launch_blockstat(const wx_params *P, int neq, const LeafRec *leaves, int nleaf, const double *U,
                     double *stat, int8_t *flag, double *Pc, int pc_init, hipStream_t stream);
// Workgroup i reduces the CFL partials of instance i's queue entries
// [inst_off[i], inst_off[i+1]).
int launch_dtreduce(const wx_params *P, const LeafRec *rec, const double *stat, const int32_t *inst_off,
                    int ninst, DevScalars *S, int advance_time, hipStream_t stream);
// flagq[inst_off[rec.inst] + rec.lidx] = flag[rec.id] over the queue (the
// flags by leaf position, whatever order the queue holds); Sout = S.
int launch_gather_flags(QueueRef q, const int32_t *inst_off, int grid, const int8_t *flag, int8_t *flagq,
                        const DevScalars *S, DevScalars *Sout, int ninst, hipStream_t stream);
// The data parts of refine/coarse; with stat != nullptr also the block
// statistics of the new leaves (as k_blockstat would compute them).
int launch_meshops(const wx_params *P, int neq, const MeshOp *ops, int nops, double *U, double *stat, int8_t *flag,
                   double *Pc, hipStream_t stream);
// Performs the n copies described in d (itself in mapped host memory), one
// launch for all: a kernel reads the pinned staging directly.
int launch_h2d(const CopyDesc *d, int n, hipStream_t stream);
int launch_transpose(int neq, int cap, const double *src, double *dst, int to_fortran, int ld,
                     hipStream_t stream);

// The OUTFYB compat mode of src/boundaryI.f90:87 is in effect: a bottom box
// entry of fill I writes block e + 1 (e its list position) instead of its own.
inline int lb_outfyb(const wx_params &p) { return (p.compat & 2) && p.bc[2] == 0; }

// Where the device table builder writes: the list length, and the ops of
// fills II and I (8 per leaf position, capacity leaf_cap leaves; fill I's
// only in the OUTFYB mode, else the fill reads fill II's).
struct LbOut {
  int32_t *nlist;
  int2 *ops2, *ops1;
  int32_t leaf_cap;
};
// Builds the ghost ops of the tree T[0..nbmax] (wx_lbuild.hip): they equal
// locatebounds + compile_ops on the host; the leaf positions are those of
// compile_leaves.  info holds nbmax + 1 entries, part (nbmax + 256) / 256.
// nleaf_host: the leaf count the host's tree holds (a different count sets
// S->mesh_err).  multi_wg_scan: scan over many workgroups at any size (else
// only above 2048 IDs).
int launch_build_tables(const wx_params &p, const TreeRow *T, int nbmax, int nleaf_host, LbInfo *info, LbPart *part,
                        LbOut out, int multi_wg_scan, DevScalars *S, hipStream_t stream);
int launch_calcprim(const wx_params *P, int neq, int cap, const double *U, double *prim, hipStream_t stream);

}  // namespace wx

/*
 * wx_capi.h -- C-ABI of the MI355X (gfx950) implementation of Walicxe-2D's
 * per-block hydrodynamic update.
 *
 * The reference (esquivas/walicxe2d, Fortran 90) has no FFI: its "plugin"
 * mechanism is link-time shadowing of same-named subroutines from a problem
 * directory (JET/Makefile:8, `VPATH = . : ../src`), all sharing module
 * globals (src/globals.f90:8-33).  This header is the boundary a Fortran host
 * binds through `iso_c_binding` (see fortran/wx_gpu.f90) so that drop-in
 * `tstep` / `timestep` / `criteria` / `refine` / `coarse` routines call the
 * GPU path instead of the CPU loops.
 *
 * Conventions
 *   - every entry point returns int status: 0 = OK, otherwise a WX_E* code;
 *     wx_last_error() returns the text of the last failure;
 *   - plain pointers and sizes only; the library never keeps host pointers;
 *   - block IDs are the reference's 1-based IDs (lp/leaf/icoord indices);
 *   - one context = one AMR mesh resident on the current HIP device; calls
 *     on one context are not re-entrant and run on the context's stream;
 *   - wx_tstep and wx_step are asynchronous; wx_timestep, wx_criteria,
 *     wx_admesh and wx_download synchronise.
 */
#ifndef WX_CAPI_H
#define WX_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WX_MAXLEV 16

/* Status codes */
#define WX_OK        0
#define WX_EHLLC     1   /* HLLC/hybrid fell through (NaN star speed): src/hllcfluxes.f90:170-171 */
#define WX_EBLOCKS   2   /* block capacity exceeded: src/admesh.f90:44-47 */
#define WX_EARG      3   /* bad argument / inconsistent mesh */
#define WX_EHIP      4   /* HIP runtime failure */
#define WX_ENOMEM    5
#define WX_EIO       6   /* checkpoint file missing, short or inconsistent */

/* Layout selector for wx_upload / wx_download */
#define WX_LAYOUT_BLOCK   0  /* [id][var][j][i], i fastest, id 0..cap (id 0 unused) */
#define WX_LAYOUT_FORTRAN 1  /* u(nbmin:nbmaxtot, neq, -1:18, -1:18), block index fastest
                                (src/initmain.f90:76-78) */

/*
 * Run-time parameters: the reference's cpp switches (JET/Makefile:118-209)
 * and parameters.f90 constants (src/parameters.f90:12-89).
 */
typedef struct wx_params {
  int32_t neq;          /* 4 + passive scalars (4..6) */
  int32_t nlevs;        /* number of refinement levels */
  int32_t nbrootx, nbrooty;
  int32_t solver;       /* 0 HLLC (src/hllcfluxes.f90), 1 HLL (src/hllfluxes.f90),
                           2 HLL_HLLC (src/hll_hllcfluxes.f90) */
  int32_t limiter;      /* -1..6 (src/limiter.f90:30-85) */
  int32_t cooling;      /* 0 ADIABATIC, 1 COOLINGH (src/coolingh.f90),
                           2 COOLINGDMC (src/coolingdmc.f90) */
  int32_t otherb;       /* 1: jet source via impose_user_bc in both ghost fills */
  int32_t bc[4];        /* left,right,bottom,top: 0 OUTFLOW, 1 REFLECTIVE */
  int32_t cylindrical;  /* axisymmetric source terms (src/step.f90:49-76) */
  int32_t compat;       /* bit0: HLLC right-star typo (src/hllcfluxes.f90:157);
                           bit1: OUTFYB writes block = list position (src/boundaryI.f90:87) */
  int32_t nblocks;      /* block-ID capacity of the context */
  int32_t pad0;
  double cv, gamma, eta, cfl;
  double xmax, ymax;
  double Tempsc, Psc, tsc;
  double tol, tol2;     /* refinement thresholds (src/criteria.f90:11) */
  double dx[WX_MAXLEV], dy[WX_MAXLEV];
  double cooltab[82];   /* DMC table cooltab(2,41), column-major (src/dmc_module.f90:9) */
  double jet[12];       /* RJ, LJ, posx, posy, denJ, TempJ, vJ0, dvJ, tau, y, 2-y, unused
                           (JET/jet.f90:15-35, code units) */
} wx_params;

typedef struct wx_ctx wx_ctx;

/* Allocation of the block arrays and the DMC table: replaces
 * src/initmain.f90:52-98 (nbmin/nbmaxtot, allocate u/up/primit, cooltab). */
int wx_create(const wx_params *p, wx_ctx **out);
int wx_destroy(wx_ctx *ctx);

/* Host -> device copy of the mesh: u (layout per WX_LAYOUT_*), lp(nb,1:8),
 * leaf(nb), icoord(nb,1:2) in the reference's Fortran layouts
 * (src/globals.f90:13-18, column-major, ld = leading dimension = number of
 * rows of lp/icoord/leaf arrays, rows indexed by block ID - 1).
 * Replaces nothing in the reference (its arrays are already in memory);
 * this is the transpose into the device's [id][var][j][i] tiles. */
int wx_upload(wx_ctx *ctx, const double *u, int32_t layout, const int32_t *lp,
              const int32_t *leaf, const int32_t *icoord, int32_t ld,
              int32_t nbmax, int32_t nbleafs);
/* Device -> host copy of u and of the primitives recomputed from it
 * (calcprim, src/calcprim.f90:5-21).  prim may be NULL.  WX_LAYOUT_BLOCK
 * writes tiles 0..nbmax; WX_LAYOUT_FORTRAN writes ld block rows (ld >= nbmax,
 * the host array's leading dimension; ld <= 0 means nbmax).  The reference
 * reads its arrays in place, so this is the transpose out of the device tiles
 * that output.f90 (src/output.f90:39-67) needs before it writes. */
int wx_download(wx_ctx *ctx, double *u, double *prim, int32_t layout, int32_t ld);
/* Device -> host copy of the mesh tree kept by the context's own admesh. */
int wx_get_tree(wx_ctx *ctx, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld,
                int32_t *nbmax, int32_t *nbleafs);

/* Ghost-fill list from the host's locatebounds (src/locatebounds.f90:29-540):
 * innerbounds(ld,3) column-major, ninternal rows (codes -4..12, MPIP forms). */
int wx_set_bounds(wx_ctx *ctx, const int32_t *innerbounds, int32_t ld, int32_t ninternal);
/* Same list built by the library's own locatebounds from its tree (on the
 * device unless wx_set_builder selected the host; *ninternal then waits for
 * the device). */
int wx_locatebounds(wx_ctx *ctx, int32_t *ninternal);
/* Where the ghost-gather ops of the library's own list are built after a
 * mesh change (src/locatebounds.f90:29-540 and the list order of
 * src/boundaryII.f90:35-514, src/boundaryI.f90:34-396): 1 on the device,
 * from the tree rows changed since the last build, beside the dt reduction
 * and stage A; 0 on the host (the list itself, compiled, then uploaded);
 * 2 on the device with the scan over block IDs spread over many workgroups
 * at any mesh size (the form it takes above 2048 IDs); 3 (default) on the
 * device from 3072 leaves up, else on the host.  All give identical
 * tables.  A list set by wx_set_bounds is always compiled on the host. */
int wx_set_builder(wx_ctx *ctx, int32_t on_device);
/* The current gather tables, for checks: *nleaf leaf records (8 int32 each:
 * id, level, icx, icy, flags, leaf position, instance, 0) and the ops of
 * fills II and I (8 per leaf, 2 int32 each: source block, kind | flags).
 * Copies only when *nleaf <= cap; synchronising. */
int wx_get_tables(wx_ctx *ctx, int32_t *nleaf, int32_t *leaves, int32_t *ops2, int32_t *ops1, int32_t cap);

/* src/timestep.f90:7-61: CFL (+ DMC cooling-time) timestep over leaf
 * interiors.  Synchronising. */
int wx_timestep(wx_ctx *ctx, double *dt);

/* src/tstep.f90:4-126: one full step (stage A, boundaryII, stage B +
 * viscosity, boundaryI, cooling) at time = t^{n+1} (src/main.f90:88,95).
 * Uses the ghost list last set by wx_set_bounds / wx_locatebounds.
 * Asynchronous. */
int wx_tstep(wx_ctx *ctx, double time, double dt);

/* src/criteria.f90:4-53 over all leaves (markref, src/markref.f90:4-85):
 * irefup/irefdown indexed by block ID - 1, length nblocks.  Synchronising. */
int wx_criteria(wx_ctx *ctx, int32_t *irefup, int32_t *irefdown);

/* wx_upload_tree keeps the compiled tables and block statistics when the tree
 * is unchanged, so a host may call it every step. */
/* Data parts of src/refine.f90:30-42 (injection over the padded block) and
 * src/coarse.f90:182-366 (2x2 restriction incl. ghosts).  The caller keeps
 * its own tree (updatelpup/updatelpdown) and calls wx_upload_tree after. */
int wx_refine(wx_ctx *ctx, int32_t nb, int32_t s1, int32_t s2, int32_t s3, int32_t s4);
int wx_coarse(wx_ctx *ctx, int32_t dad, int32_t s1, int32_t s2, int32_t s3, int32_t s4);
int wx_upload_tree(wx_ctx *ctx, const int32_t *lp, const int32_t *leaf, const int32_t *icoord,
                   int32_t ld, int32_t nbmax, int32_t nbleafs);

/* Standalone driver (the library's own host tree):
 *   wx_basegrid  -- generalised root layout (src/basegrid.f90:26-97) with the
 *                   initial data taken from u (WX_LAYOUT_BLOCK, roots only);
 *   wx_admesh    -- src/admesh.f90:4-111 (markref, critneighup, refine by
 *                   level, critneighdown, coarsen by level);
 *   wx_step      -- one iteration of the time loop src/main.f90:87-98:
 *                   timestep, time += dt, tstep, admesh; time and dt stay
 *                   on the device and are returned only when asked. */
int wx_basegrid(wx_ctx *ctx, const double *uroots, int32_t nroots);
int wx_admesh(wx_ctx *ctx, int32_t *changed);
int wx_step(wx_ctx *ctx, int32_t nsteps, int32_t do_admesh, double *time, double *dt,
            double *zone_updates);
int wx_set_time(wx_ctx *ctx, double time);

/* Per-kernel timing with HIP events recorded on the context's stream around
 * every launch (no counterpart in the reference, which has no timers).
 * Returns the accumulated milliseconds / launch counts of kernels
 * 0 stage A, 1 stage B, 2 boundaryI+cooling, 3 dt reduction, 4 block
 * statistics, 5 refine/coarse ops; 6 host wall time running admesh on the
 * flags (after the wait for the step), 7 host wall time compiling and
 * uploading the ghost lists, 8 the device table builder (tree rows and
 * the builder's kernels)
 * (n entries); then, if enable >= 0, resets
 * the counters and switches recording on (1) or off (0). */
int wx_profile(wx_ctx *ctx, int32_t enable, double *ms, int64_t *counts, int32_t n);

/* Bookkeeping */
int wx_info(wx_ctx *ctx, int32_t *nbmax, int32_t *nbleafs, int32_t *ninternal);
/* neq, block-ID capacity and levels the context was created with. */
int wx_shape(wx_ctx *ctx, int32_t *neq, int32_t *nblocks, int32_t *nlevs);

/* BIN checkpoints.  wx_write_bin writes the reference's OUTBIN pair
 * (src/output.f90:39-67): `pointers` = int32 nbmin=1, nbmax, nbleafs,
 * nblocks; lp(1:nbmax,1:8), leaf(1:nbmax) as 4-byte logicals with .true. =
 * logical_true (-1 as ifort writes it and py/walicxe2d_utils.py:48 expects,
 * or 1 as flang writes it), icoord(1:nbmax,1:2), all column-major; `blocks` =
 * u(nb,1:neq,-1:18,-1:18) of every leaf in ID order, FP64 little-endian.
 * wx_read_bin is the warm start of src/basegrid.f90:167-251: it reads nfiles
 * such pairs (one per rank of an MPI run, block IDs kept as written; the
 * capacity must cover them) and replaces the context's mesh and state. */
int wx_write_bin(wx_ctx *ctx, const char *pointers_path, const char *blocks_path, int32_t logical_true);
int wx_read_bin(wx_ctx *ctx, int32_t nfiles, const char *const *pointers_paths, const char *const *blocks_paths);

/* Block placement plan of the MPI load balancer (src/loadbalance.f90:5-105)
 * for a mesh held by one rank (IDs 1..nbmax): hilbert(nb) = the Hilbert key
 * of the block's corner at the finest level over nlevs+1 bits per axis
 * (hilbert2d, :251-306; higher bits ignored as there), and idnew(nb) = the
 * block's new ID when the blocks, in key order (indexx, :320-400, with its
 * order among equal keys), are split into nranks runs: rank r gets IDs
 * r*nblocks + 1, 2, ... and closes after the block that brings its leaf count
 * to nbleafs/nranks.  lp(ld,8), leaf(ld), icoord(ld,2) column-major as in
 * wx_upload_tree; hilbert and idnew have nbmax entries.  Host only (no
 * device work, no context).  WX_EBLOCKS where the reference stops (an ID
 * reaching its rank's capacity).  The reference's remapping of lp, its data
 * exchange and its MPI collectives are not part of the plan. */
int wx_lb_plan(const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int32_t ld, int32_t nbmax,
               int32_t nlevs, int32_t nblocks, int32_t nranks, int32_t *hilbert, int32_t *idnew);

/*
 * Batched ensemble: n independent AMR meshes ("instances"), each with its own
 * parameters (p[0..n-1], same neq), tree, time and timestep, advanced
 * together.  The reference runs one mesh per process (the time loop
 * src/main.f90:69-106); an ensemble replaces n such runs.  Each stage is one
 * launch over the compacted queue of every instance's leaf blocks (one
 * workgroup per block, or a persistent grid that the autotuner may choose);
 * the step can be captured in a hipGraph; host work per instance
 * (list compilation, admesh) runs on worker threads (WX_HOST_THREADS, else
 * OMP_NUM_THREADS, at most 32).  Results of instance i are bit-identical to
 * running it alone in a wx_ctx with the same parameters and initial data.
 */
typedef struct wx_ens wx_ens;
int wx_ens_create(const wx_params *p, int32_t n, wx_ens **out);
/* The same with the instances split into `groups` groups of consecutive
 * instances (1 <= groups <= n), each with its own stream, queue, graphs and
 * worker threads (WX_HOST_THREADS / groups each).  wx_ens_step then steps
 * the groups concurrently from one host thread each: one group's admesh and
 * list compilation overlap the other groups' kernels, and the groups'
 * kernels overlap one another.  Results are bit-identical for any grouping.
 * wx_ens_create is the one-group case. */
int wx_ens_create_groups(const wx_params *p, int32_t n, int32_t groups, wx_ens **out);
int wx_ens_groups(wx_ens *ens, int32_t *groups);
/* neq, block-ID capacity and levels of instance inst. */
int wx_ens_shape(wx_ens *ens, int32_t inst, int32_t *neq, int32_t *nblocks, int32_t *nlevs);
/* The OUTBIN checkpoint pair of instance inst, as wx_write_bin. */
int wx_ens_write_bin(wx_ens *ens, int32_t inst, const char *pointers_path, const char *blocks_path,
                     int32_t logical_true);
int wx_ens_destroy(wx_ens *ens);
/* Root layout + initial data of instance inst (as wx_basegrid). */
int wx_ens_basegrid(wx_ens *ens, int32_t inst, const double *uroots, int32_t nroots);
/* admesh of every instance (src/admesh.f90:4-111), e.g. the initial
 * refinement passes of src/main.f90:60-63.  Synchronising. */
int wx_ens_admesh(wx_ens *ens, int32_t *changed);
/* nsteps iterations of src/main.f90:87-98 for every instance (timestep,
 * time += dt, tstep, admesh if do_admesh).  Returns after a device sync. */
int wx_ens_step(wx_ens *ens, int32_t nsteps, int32_t do_admesh, double *zone_updates);
int wx_ens_instance(wx_ens *ens, int32_t inst, double *time, double *dt, int32_t *nbmax, int32_t *nbleafs);
/* Tiles 0..nbmax of instance inst, WX_LAYOUT_BLOCK. */
int wx_ens_download(wx_ens *ens, int32_t inst, double *u);
int wx_ens_get_tree(wx_ens *ens, int32_t inst, int32_t *lp, int32_t *leaf, int32_t *icoord, int32_t ld,
                    int32_t *nbmax, int32_t *nbleafs);
/* hipGraph capture of the step's launches (1 on, 0 off). */
int wx_ens_set_graph(wx_ens *ens, int32_t enable);
/* Launch shape of stage kernel k (0 stage A, 1 stage B, 2 boundaryI+cooling):
 * persistent workgroups (0 = one per queue entry) and threads per workgroup
 * (256, 384 or 512). */
int wx_ens_set_launch(wx_ens *ens, int32_t kernel, int32_t grid, int32_t threads);
/* Autotuner: times the three stage kernels at every workgroup size and
 * several persistent grids on the current meshes (reps launches each) and
 * keeps the fastest, for each group; returns group 0's chosen grids, threads
 * and microseconds per launch (3 each).  Later steps give the same results
 * as untuned ones. */
int wx_ens_tune(wx_ens *ens, int32_t reps, int32_t *grids, int32_t *threads, double *us);
/* As wx_profile, summed over the groups.  While recording is on, wx_ens_step
 * steps the groups one after the other, so that each kernel's time is not
 * shared with another group's kernels. */
int wx_ens_profile(wx_ens *ens, int32_t enable, double *ms, int64_t *counts, int32_t n);
/* Instances, leaf blocks of all instances, group 0's launch shapes. */
int wx_ens_info(wx_ens *ens, int32_t *ninst, int32_t *nleaf_total, int32_t *grids, int32_t *threads);
int wx_last_error(char *buf, int32_t len);
/* HIP stream the context launches on (hipStream_t as void*). */
void *wx_stream(wx_ctx *ctx);

/*
 * Cooling surrogate.  No counterpart in the reference: a 4 -> 64 -> 64 ->
 * 64 -> 2 SiLU network (bf16 weights and activations on the matrix cores,
 * FP32 accumulation) stands in for coolingh's per-cell update
 * (src/coolingh.f90:171-259) on states inside its training hull, and the
 * exact routine runs wherever a state leaves the hull or the energy residual
 * exceeds `tol`.  Features: log10 n, log10 T, x_HII = 1 - n_HI/n,
 * log10(dt/t_cool) with t_cool from the routine's relaxation law
 * (src/coolingh.f90:249-250) and a device table of ALOSS/n^2; outputs:
 * c = log10(T1/T1_table), the correction to that relaxation law with the
 * tabulated loss, and x_HII(new) - x_HII.  The energy residual is the
 * change of thermal energy the correction makes, relative to the old one.
 */
typedef struct wx_sur_model {
  float theta[8770];   /* W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64],
                          W4[2][64], b4[2] (row-major [out][in]) */
  float in_mu[4], in_sig[4];    /* feature standardisation */
  float out_mu[2], out_sig[2];  /* output standardisation */
  float lo[4], hi[4];           /* training hull of the raw features */
  float tol;                    /* energy-residual tolerance (relative) */
  int32_t pad[3];
} wx_sur_model;

/* Exact coolingh update of n states at rest, states[s] = (n, T, y0 =
 * n_HI/n, dt in s), with the cooling parameters of p: out[s] = (T1, y1).
 * These are the training pairs. */
int wx_sur_pairs(const wx_params *p, const double *states, int64_t n, double *out);
/* Raw features of n states (as above), feat[s][4], and ce*dt of the
 * tabulated relaxation law, cedt[s]. */
int wx_sur_features(const wx_params *p, const double *states, int64_t n, float *feat, float *cedt);
/* Network on standardised inputs x[n][4] -> standardised outputs y[n][2]. */
int wx_sur_infer(const wx_sur_model *m, const float *x, int64_t n, float *y);
/* Mean over the n samples of w0 (y0 - t0)^2 + w1 (y1 - t1)^2 and its
 * gradient with respect to theta (8770 floats). */
int wx_sur_grad(const wx_sur_model *m, const float *x, const float *t, int64_t n, float w0, float w1,
                float *loss, float *grad);
/* `steps` Adam steps (lr decaying geometrically from lr0 to lr1) over
 * consecutive batches of the n samples (wrapping); updates m->theta and
 * writes each step's batch loss to loss_hist (may be NULL). */
int wx_sur_train(wx_sur_model *m, const float *x, const float *t, int64_t n, int32_t batch, int32_t steps,
                 float lr0, float lr1, float w0, float w1, float *loss_hist);
/* Use the surrogate in the cooling pass of coolingh instances (m = NULL:
 * exact cooling again), and read / reset the counts of cells it handled and
 * of cells sent to the exact routine. */
int wx_set_surrogate(wx_ctx *ctx, const wx_sur_model *m);
int wx_ens_set_surrogate(wx_ens *ens, const wx_sur_model *m);
int wx_sur_counters(wx_ctx *ctx, int32_t reset, int64_t *cells, int64_t *fallbacks);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): wx_sur_infer with the model (a wx_sur_model), the
 * inputs x[n][4] and the outputs y[n][2] all in device memory, launched on
 * `stream` (a hipStream_t; NULL = the default stream) and not waited for.
 * The device form of the model is built by kernels in a fixed static device
 * buffer of 160 KiB per GPU (held until the process ends) instead of memory
 * allocated per call; n needs no scratch and is not limited by it.  The
 * buffer keeps the last call's model; calls that overlap on one GPU race on
 * it.  Results equal wx_sur_infer's bit for bit. */
int wx_sur_infer_noalloc(const wx_sur_model *d_model, const float *d_x, int64_t n, float *d_y, void *stream);
int wx_ens_sur_counters(wx_ens *ens, int32_t reset, int64_t *cells, int64_t *fallbacks);

#ifdef __cplusplus
}
#endif
#endif

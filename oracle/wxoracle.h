/*
 * wxoracle.h -- CPU oracle for the Walicxe-2D per-block hydro update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (walicxe2d_amd/) never links, loads or calls it.
 *
 * It restates, in plain C, the reference algorithm of esquivas/walicxe2d
 * (Fortran 90, `-r8`, i.e. every `real` is FP64) with the reference's
 * evaluation order kept term by term, so that a compiler run with
 * -ffp-contract=off reproduces the Fortran arithmetic.  Each function cites
 * the reference file:line it follows.
 *
 * Storage: block IDs are 1-based as in the reference (slot 0 unused).  A
 * block is a padded 20x20 tile per variable, laid out [id][var][j][i] with
 * i fastest; cell (i,j) of the reference (i,j in -1..18) lives at
 * (j+1)*20 + (i+1).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - hydro (uprim/primu/primf/sound/limiters/HLL/HLLC/hybrid/step/viscosity/
 *     calcprim), boundaryI/II, coolingh, coolingdmc, timestep, criteria,
 *     refine, coarse, updatelpup: against the compiled reference (oracle/_ref)
 *     and committed golden vectors, plus closed forms / invariants;
 *   - locatebounds, updatelpdown, the admesh driver: not compilable from the
 *     reference unmodified (ifort `not(logical)` extension), pinned by
 *     geometric brute force (neighbour pointers recomputed from icoord/level)
 *     and list invariants (every leaf face covered exactly once).
 */
#ifndef WXORACLE_H
#define WXORACLE_H
#include <stdint.h>

#define WXO_NX 16
#define WXO_NY 16
#define WXO_NP 20          /* padded extent (nghost = 2) */
#define WXO_NCELL 400
#define WXO_MAXLEV 16
#define WXO_MAXEQ 8

/* Runtime replacement for the reference's cpp switches (JET/Makefile:118-209)
 * and parameters.f90 constants.  Same binary layout as wx_params in
 * include/wx_capi.h so one host-side structure can feed both. */
typedef struct wxo_params {
  int32_t neq;         /* 4 + passive scalars */
  int32_t nlevs;
  int32_t nbrootx, nbrooty;
  int32_t solver;      /* 0 HLLC, 1 HLL, 2 HLL_HLLC */
  int32_t limiter;     /* -1..6 (src/limiter.f90:30-85) */
  int32_t cooling;     /* 0 ADIABATIC, 1 COOLINGH, 2 COOLINGDMC */
  int32_t otherb;      /* 1: jet source imposed in the ghost fills */
  int32_t bc[4];       /* left,right,bottom,top: 0 outflow, 1 reflective */
  int32_t cylindrical;
  int32_t compat;      /* bit0 HLLC right-star typo, bit1 OUTFYB list-index write */
  int32_t nblocks;     /* block-ID capacity */
  int32_t pad0;
  double cv, gamma, eta, cfl;
  double xmax, ymax;
  double Tempsc, Psc, tsc;
  double tol, tol2;
  double dx[WXO_MAXLEV], dy[WXO_MAXLEV];
  double cooltab[82];  /* Fortran cooltab(2,41), column-major */
  double jet[12];      /* RJ, LJ, posx, posy, den, Temp(code), v0, dv, tau, yneutral, efac, - */
} wxo_params;

typedef struct wxo_mesh {
  int32_t cap;         /* largest usable block ID */
  int32_t nbmax, nbleafs, ninternal;
  double *u, *up, *primit;   /* [(cap+1)][neq][20][20] */
  int32_t *lp;               /* [(cap+1)][8] : lp(nb,k) at lp[nb*8+k-1] */
  int32_t *leaf;             /* [cap+1] */
  int32_t *icoord;           /* [(cap+1)][2] */
  int32_t *bounds;           /* innerbounds [4*(cap+1)][3] */
} wxo_mesh;

enum { WXO_OK = 0, WXO_ERR_HLLC = 1, WXO_ERR_BLOCKS = 2, WXO_ERR_ARG = 3 };

#ifdef __cplusplus
extern "C" {
#endif
void   wxo_uprim(const wxo_params *p, double *prim, const double *uu, double *T);
void   wxo_primu(const wxo_params *p, const double *prim, double *uu);
void   wxo_primf(const wxo_params *p, const double *prim, double *ff);
double wxo_average(int limiter, double a, double b);
void   wxo_limiter(const wxo_params *p, const double *pll, double *pl, double *pr, const double *prr);
int    wxo_riemann(const wxo_params *p, const double *pl, const double *pr, double *ff);
int    wxo_fluxes(const wxo_params *p, const wxo_mesh *m, int nb, int choice, double *f, double *g);
void   wxo_step(const wxo_params *p, wxo_mesh *m, int nb, double dt, const double *f, const double *g);
void   wxo_viscosity(const wxo_params *p, wxo_mesh *m, int nb);
void   wxo_calcprim(const wxo_params *p, const double *u1, double *prim1);
void   wxo_impose_jet(const wxo_params *p, const wxo_mesh *m, double *ublk, int nb, double time);
void   wxo_boundary(const wxo_params *p, wxo_mesh *m, double time, int which);
double wxo_aloss(double x1, double x2, double dt, double den, double dh0, double te0);
void   wxo_atomic(const wxo_params *p, double dt, double *uu);
double wxo_cooldmc(const wxo_params *p, double T);
void   wxo_coolingh(const wxo_params *p, wxo_mesh *m, double dt);
void   wxo_coolingdmc(const wxo_params *p, wxo_mesh *m, double dt);
double wxo_timestep(const wxo_params *p, const wxo_mesh *m);
int    wxo_tstep(const wxo_params *p, wxo_mesh *m, double time, double dt);
int    wxo_locatebounds(const wxo_params *p, wxo_mesh *m);
void   wxo_criteria(const wxo_params *p, const wxo_mesh *m, int nb, int32_t *irefup, int32_t *irefdown);
void   wxo_refine(const wxo_params *p, wxo_mesh *m, int nb, int s1, int s2, int s3, int s4);
void   wxo_updatelpup(const wxo_params *p, wxo_mesh *m, int nb, int s1, int s2, int s3, int s4,
                      int32_t *irefup, int32_t *irefdown);
void   wxo_coarse(const wxo_params *p, wxo_mesh *m, int dad, int s1, int s2, int s3, int s4);
void   wxo_updatelpdown(const wxo_params *p, wxo_mesh *m, int dad, int s1, int s2, int s3, int s4);
void   wxo_markref(const wxo_params *p, const wxo_mesh *m, int32_t *irefup, int32_t *irefdown);
void   wxo_critneighup(const wxo_params *p, const wxo_mesh *m, int32_t *irefup);
void   wxo_critneighdown(const wxo_params *p, const wxo_mesh *m, int32_t *irefdown);
int    wxo_admesh(const wxo_params *p, wxo_mesh *m);
int    wxo_admesh_flags(const wxo_params *p, wxo_mesh *m, int32_t *irefup, int32_t *irefdown);
void   wxo_basegrid(const wxo_params *p, wxo_mesh *m);
int    wxo_run(const wxo_params *p, wxo_mesh *m, int nsteps, double *time, double *dtlast,
               int do_admesh, double *zones);
/* Load-balance plan (src/loadbalance.f90): Hilbert keys, the index sort and
 * the new block IDs for nranks ranks of a mesh held on one rank. */
int32_t wxo_hilbert2d(int32_t x, int32_t y, int bits);
void   wxo_indexx(int n, const int32_t *arr, int32_t *indx);
int    wxo_lb_plan(const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int nbmax, int nlevs,
                   int nblocks, int nranks, int32_t *key, int32_t *idnew);
#ifdef __cplusplus
}
#endif
#endif

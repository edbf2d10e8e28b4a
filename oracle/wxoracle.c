/*
 * wxoracle.c -- CPU oracle (test infrastructure only; see wxoracle.h).
 *
 * Plain-C restatement of Walicxe-2D's hot path.  Each function names the
 * reference file:line it follows.  Evaluation order follows the Fortran
 * expressions term by term (left-to-right for same-precedence operators,
 * unary minus applied after `*`/`/` as Fortran parses `-a*b`), so that a
 * build with -ffp-contract=off gives the reference's arithmetic.
 */
#include "wxoracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

#define NX WXO_NX
#define NY WXO_NY
#define NP WXO_NP
#define NC WXO_NCELL
#define IDX(i, j) (((j) + 1) * NP + ((i) + 1))
#define AT(a, v, i, j) (a)[(v) * NC + IDX(i, j)]

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double *blk(const wxo_params *p, double *a, int nb) {
  return a + (size_t)nb * (size_t)p->neq * NC;
}
static inline int LP(const wxo_mesh *m, int nb, int k) { return m->lp[nb * 8 + k - 1]; }
static inline void SETLP(wxo_mesh *m, int nb, int k, int v) { m->lp[nb * 8 + k - 1] = v; }

/* ------------------------------------------------------------------ */
/* src/uprim.f90:5-63 : conserved -> primitive with floors, T by mode   */
void wxo_uprim(const wxo_params *p, double *prim, const double *uu, double *T) {
  const double cv = p->cv;
  double r = dmax(uu[0], 1e-15);                       /* :16 */
  prim[0] = r;
  prim[1] = uu[1] / r;
  prim[2] = uu[2] / r;
  for (int k = 4; k < p->neq; ++k) prim[k] = uu[k];     /* :22-24 PASSIVES */
  prim[3] = (uu[3] - 0.5 * r * (prim[1] * prim[1] + prim[2] * prim[2])) / cv;  /* :25 */
  prim[3] = dmax(prim[3], 1e-15);                      /* :26 */
  if (p->cooling == 0) {                               /* ADIABATIC :29-33 */
    *T = prim[3] / r;
  } else if (p->cooling == 1) {                        /* COOLINGH :35-43 */
    double dentot = (2.0 * r - prim[4]);
    dentot = dmax(dentot, 1e-15);
    *T = dmax(1.0, (prim[3] / dentot) * p->Tempsc);
    prim[3] = dentot * *T / p->Tempsc;
  } else {                                             /* COOLINGDMC :45-53 */
    r = dmax(r, 1e-15);
    *T = dmax(1.0, (prim[3] / (r)) * p->Tempsc);
    prim[3] = r * *T / p->Tempsc;
  }
}

/* src/primu.f90:4-19 */
void wxo_primu(const wxo_params *p, const double *prim, double *uu) {
  uu[0] = prim[0];
  uu[1] = prim[0] * prim[1];
  uu[2] = prim[0] * prim[2];
  uu[3] = 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]) + p->cv * prim[3];
  for (int k = 4; k < p->neq; ++k) uu[k] = prim[k];
}

/* src/primf.f90:5-23 */
void wxo_primf(const wxo_params *p, const double *prim, double *ff) {
  double etot = 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]) + p->cv * prim[3];
  ff[0] = prim[0] * prim[1];
  ff[1] = prim[0] * prim[1] * prim[1] + prim[3];
  ff[2] = prim[0] * prim[1] * prim[2];
  ff[3] = prim[1] * (etot + prim[3]);
  for (int k = 4; k < p->neq; ++k) ff[k] = prim[k] * prim[1];
}

/* src/sound.f90:4-13 */
static inline double sound(const wxo_params *p, double pr, double d) {
  return sqrt(p->gamma * pr / d);
}

/* src/swapy.f90:4-13 */
static inline void swapy(double *q) { double a = q[1]; q[1] = q[2]; q[2] = a; }

/* src/limiter.f90:26-87 : the `average` slope function, variants -1..6 */
double wxo_average(int limiter, double a, double b) {
  double s, c, d, av1, av2, r;
  switch (limiter) {
    case -1: return 0.0;                                        /* :30-32 */
    case 0: return 0.5 * (a + b);                               /* :35-37 */
    case 1:                                                     /* :40-45 minmod */
      s = copysign(1.0, a);
      return s * dmax(0.0, dmin(fabs(a), s * b));
    case 2:                                                     /* :47-54 Falle/van Leer */
      if (a * b <= 0.0) return 0.0;
      return a * b * (a + b) / (a * a + b * b);
    case 3: {                                                   /* :56-59 van Albada */
      const double delta = 1.e-7;
      return (a * (b * b + delta) + b * (a * a + delta)) / (a * a + b * b + delta);
    }
    case 4:                                                     /* :61-68 UMIST */
      s = copysign(1.0, a);
      c = 0.25 * a + 0.75 * b;
      d = 0.75 * a + 0.25 * b;
      r = dmin(dmin(dmin(2.0 * fabs(a), 2.0 * s * b), s * c), s * d);
      return s * dmax(0.0, r);
    case 5:                                                     /* :70-76 Woodward/MC */
      s = copysign(1.0, a);
      c = 0.5 * (a + b);
      r = dmin(dmin(2.0 * fabs(a), 2.0 * s * b), s * c);
      return s * dmax(0.0, r);
    case 6:                                                     /* :78-84 superbee */
      s = copysign(1.0, b);
      av1 = dmin(2.0 * fabs(b), s * a);
      av2 = dmin(fabs(b), 2.0 * s * a);
      return s * dmax(dmax(0.0, av1), av2);
    default: return 0.0;
  }
}

/* src/limiter.f90:7-24 : limited reconstruction, every variable (incl. passives) */
void wxo_limiter(const wxo_params *p, const double *pll, double *pl, double *pr, const double *prr) {
  for (int k = 0; k < p->neq; ++k) {
    double dl = pl[k] - pll[k];
    double dm = pr[k] - pl[k];
    double dr = prr[k] - pr[k];
    double al = wxo_average(p->limiter, dl, dm);
    double ar = wxo_average(p->limiter, dm, dr);
    pl[k] = pl[k] + al * 0.5;
    pr[k] = pr[k] - ar * 0.5;
  }
}

/* src/hllfluxes.f90:98-131 primfhll */
static int riemann_hll(const wxo_params *p, const double *pl, const double *pr, double *ff) {
  double fl[WXO_MAXEQ], fr[WXO_MAXEQ], ul[WXO_MAXEQ], ur[WXO_MAXEQ];
  double csl = sound(p, pl[3], pl[0]);
  double csr = sound(p, pr[3], pr[0]);
  double sr = dmax(pl[1] + csl, pr[1] + csr);
  double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { wxo_primf(p, pl, ff); return 0; }
  if (sr < 0) { wxo_primf(p, pr, ff); return 0; }
  wxo_primf(p, pl, fl); wxo_primf(p, pr, fr);
  wxo_primu(p, pl, ul); wxo_primu(p, pr, ur);
  for (int k = 0; k < p->neq; ++k)                                   /* :128 */
    ff[k] = (sr * fl[k] - sl * fr[k] + sl * sr * (ur[k] - ul[k])) / (sr - sl);
  return 0;
}

/* HLLC star-state flux on one side; src/hllcfluxes.f90:131-166.
 * side 0 = left (sst>=0), 1 = right (sst<=0). */
static void hllc_star(const wxo_params *p, const double *q, double s, double smu, double sst,
                      int right, double *ff) {
  double uu[WXO_MAXEQ], uuk[WXO_MAXEQ];
  double rhost = q[0] * (smu) / (s - sst);
  double ek = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]) + p->cv * q[3];
  uuk[0] = rhost;
  uuk[1] = rhost * sst;
  uuk[2] = rhost * q[2];
  /* :138 left; :157 right carries the reference's primr(4)/(primr(4)*srmur) typo */
  double den = (right && (p->compat & 1)) ? (q[3] * smu) : (q[0] * smu);
  uuk[3] = rhost * (ek / q[0] + (sst - q[1]) * (sst + q[3] / den));
  for (int k = 4; k < p->neq; ++k) uuk[k] = rhost * q[k] / q[0];   /* :141, :160 */
  wxo_primf(p, q, ff);
  wxo_primu(p, q, uu);
  for (int k = 0; k < p->neq; ++k) ff[k] = ff[k] + s * (uuk[k] - uu[k]);
}

/* src/hllcfluxes.f90:98-173 primfhllc ; returns WXO_ERR_HLLC on the :170 path */
static int riemann_hllc(const wxo_params *p, const double *pl, const double *pr, double *ff) {
  double csl = sound(p, pl[3], pl[0]);
  double csr = sound(p, pr[3], pr[0]);
  double sr = dmax(pl[1] + csl, pr[1] + csr);
  double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { wxo_primf(p, pl, ff); return 0; }
  if (sr < 0) { wxo_primf(p, pr, ff); return 0; }
  double slmul = sl - pl[1];
  double srmur = sr - pr[1];
  double rholul = pl[0] * pl[1];
  double rhorur = pr[0] * pr[1];
  double sst = (srmur * rhorur - slmul * rholul - pr[3] + pl[3]) / (srmur * pr[0] - slmul * pl[0]);
  if (sst >= 0.0) { hllc_star(p, pl, sl, slmul, sst, 0, ff); return 0; }
  if (sst <= 0.0) { hllc_star(p, pr, sr, srmur, sst, 1, ff); return 0; }
  return WXO_ERR_HLLC;
}

/* src/hll_hllcfluxes.f90:95-209 primfhll_hllc */
static int riemann_hybrid(const wxo_params *p, const double *pl, const double *pr, double *ff) {
  double fl[WXO_MAXEQ], fr[WXO_MAXEQ], ul[WXO_MAXEQ], ur[WXO_MAXEQ], fs[WXO_MAXEQ];
  const double eps = 1e-6;
  double dq1 = pr[1] - pl[1], dq2 = pr[2] - pl[2];
  double mag = sqrt(dq1 * dq1 + dq2 * dq2);
  double beta1, beta2;
  if (mag <= eps) { beta1 = 1.0; beta2 = 0.0; }
  else {
    double alpha1 = fabs(dq1) / mag;
    beta1 = sqrt(alpha1);
    beta2 = 1.0 - beta1;
  }
  double csl = sound(p, pl[3], pl[0]);
  double csr = sound(p, pr[3], pr[0]);
  double sr = dmax(pl[1] + csl, pr[1] + csr);
  double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { wxo_primf(p, pl, ff); return 0; }
  if (sr < 0) { wxo_primf(p, pr, ff); return 0; }
  wxo_primf(p, pl, fl); wxo_primf(p, pr, fr);
  wxo_primu(p, pl, ul); wxo_primu(p, pr, ur);
  double slmul = sl - pl[1];
  double srmur = sr - pr[1];
  double rholul = pl[0] * pl[1];
  double rhorur = pr[0] * pr[1];
  double sst = (srmur * rhorur - slmul * rholul - pr[3] + pl[3]) / (srmur * pr[0] - slmul * pl[0]);
  int side;
  if (sst >= 0.0) side = 0; else if (sst <= 0.0) side = 1; else return WXO_ERR_HLLC;
  if (side == 0) hllc_star(p, pl, sl, slmul, sst, 0, fs);
  else hllc_star(p, pr, sr, srmur, sst, 1, fs);
  for (int k = 0; k < p->neq; ++k)                                     /* :180-181, :201-202 */
    ff[k] = beta1 * (fs[k]) + beta2 * ((sr * fl[k] - sl * fr[k] + sl * sr * (ur[k] - ul[k])) / (sr - sl));
  return 0;
}

int wxo_riemann(const wxo_params *p, const double *pl, const double *pr, double *ff) {
  if (p->solver == 1) return riemann_hll(p, pl, pr, ff);
  if (p->solver == 2) return riemann_hybrid(p, pl, pr, ff);
  return riemann_hllc(p, pl, pr, ff);
}

/* src/hllcfluxes.f90:7-91 (same loop structure in hllfluxes/hll_hllcfluxes).
 * f, g : padded [neq][20][20]; faces i=0..nx, j=0..ny are written. */
int wxo_fluxes(const wxo_params *p, const wxo_mesh *m, int nb, int choice, double *f, double *g) {
  const int neq = p->neq;
  const double *pr = blk(p, m->primit, nb);
  double ql[WXO_MAXEQ], qr[WXO_MAXEQ], qll[WXO_MAXEQ], qrr[WXO_MAXEQ], ff[WXO_MAXEQ];
  int err = 0;
  for (int i = 0; i <= NX; ++i) {
    for (int j = 0; j <= NY; ++j) {
      /* x direction */
      for (int k = 0; k < neq; ++k) { ql[k] = AT(pr, k, i, j); qr[k] = AT(pr, k, i + 1, j); }
      if (choice == 2) {
        for (int k = 0; k < neq; ++k) { qll[k] = AT(pr, k, i - 1, j); qrr[k] = AT(pr, k, i + 2, j); }
        wxo_limiter(p, qll, ql, qr, qrr);
      }
      err |= wxo_riemann(p, ql, qr, ff);
      for (int k = 0; k < neq; ++k) AT(f, k, i, j) = ff[k];
      /* y direction */
      for (int k = 0; k < neq; ++k) { ql[k] = AT(pr, k, i, j); qr[k] = AT(pr, k, i, j + 1); }
      swapy(ql); swapy(qr);
      if (choice == 2) {
        for (int k = 0; k < neq; ++k) { qll[k] = AT(pr, k, i, j - 1); qrr[k] = AT(pr, k, i, j + 2); }
        swapy(qll); swapy(qrr);
        wxo_limiter(p, qll, ql, qr, qrr);
      }
      err |= wxo_riemann(p, ql, qr, ff);
      swapy(ff);
      for (int k = 0; k < neq; ++k) AT(g, k, i, j) = ff[k];
    }
  }
  return err;
}

/* src/step.f90:5-41 (+ cylindrical sources :49-76) */
void wxo_step(const wxo_params *p, wxo_mesh *m, int nb, double dt, const double *f, const double *g) {
  const int neq = p->neq;
  const int lev = LP(m, nb, 1);
  const double dtdx = dt / p->dx[lev - 1];
  const double dtdy = dt / p->dy[lev - 1];
  double *uo = blk(p, m->u, nb), *upb = blk(p, m->up, nb);
  const double *pr = blk(p, m->primit, nb);
  for (int i = 1; i <= NX; ++i) {
    for (int j = 1; j <= NY; ++j) {
      double S[WXO_MAXEQ];
      if (p->cylindrical) {
        double radius = ((double)(j + m->icoord[nb * 2 + 1]) - 0.5) * p->dy[lev - 1];
        double rho = AT(pr, 0, i, j);
        double H = 0.5 * rho * (AT(pr, 1, i, j) * AT(pr, 1, i, j) + AT(pr, 2, i, j) * AT(pr, 2, i, j))
                   + (p->cv + 1.) * AT(pr, 3, i, j);
        S[0] = -(rho * AT(pr, 2, i, j) / radius);
        S[1] = -(rho * AT(pr, 1, i, j) * AT(pr, 2, i, j) / radius);
        S[2] = -(rho * AT(pr, 2, i, j) * AT(pr, 2, i, j) / radius);
        S[3] = -(H * AT(pr, 2, i, j) / radius);
        for (int k = 4; k < neq; ++k) S[k] = -(AT(pr, k, i, j) * AT(pr, 2, i, j) / radius);
      }
      for (int k = 0; k < neq; ++k) {
        double v = AT(uo, k, i, j) - dtdx * (AT(f, k, i, j) - AT(f, k, i - 1, j))
                   - dtdy * (AT(g, k, i, j) - AT(g, k, i, j - 1));
        if (p->cylindrical) v = v + dt * S[k];
        AT(upb, k, i, j) = v;
      }
    }
  }
}

/* src/viscosity.f90:6-25 */
void wxo_viscosity(const wxo_params *p, wxo_mesh *m, int nb) {
  double *uo = blk(p, m->u, nb), *upb = blk(p, m->up, nb);
  for (int i = 1; i <= NX; ++i)
    for (int j = 1; j <= NY; ++j)
      for (int k = 0; k < p->neq; ++k)
        AT(upb, k, i, j) = AT(upb, k, i, j) + p->eta * (AT(uo, k, i + 1, j) + AT(uo, k, i - 1, j)
                            + AT(uo, k, i, j + 1) + AT(uo, k, i, j - 1) - 4. * AT(uo, k, i, j));
}

/* src/calcprim.f90:5-21 over one padded block */
void wxo_calcprim(const wxo_params *p, const double *u1, double *prim1) {
  double uu[WXO_MAXEQ], q[WXO_MAXEQ], T;
  for (int i = -1; i <= NX + 2; ++i)
    for (int j = -1; j <= NY + 2; ++j) {
      for (int k = 0; k < p->neq; ++k) uu[k] = AT(u1, k, i, j);
      wxo_uprim(p, q, uu, &T);
      for (int k = 0; k < p->neq; ++k) AT(prim1, k, i, j) = q[k];
    }
}

/* ------------------------------------------------------------------ */
/* Jet source: JET/jet.f90:41-77 impose_jet via src/user_mod.f90:73-84.
 * Parameters come from wxo_params.jet (RJ, LJ, posJ, denJ, TempJ, vJ0, dvJ,
 * tau, neutral fraction y, energy factor 2-y); writes u(6)=denJ only when
 * neq >= 6 (the reference writes it unconditionally). */
void wxo_impose_jet(const wxo_params *p, const wxo_mesh *m, double *ub, int nb, double time) {
  const double *J = p->jet;
  const double pi = acos(-1.);
  int lev = LP(m, nb, 1);
  double vJet = J[6] + J[7] * sin(2. * pi * time / J[8]);
  for (int i = -1; i <= NX + 2; ++i)
    for (int j = -1; j <= NY + 2; ++j) {
      double x = ((double)(i + m->icoord[nb * 2]) - 0.5) * p->dx[lev - 1];
      double y = ((double)(j + m->icoord[nb * 2 + 1]) - 0.5) * p->dy[lev - 1];
      if ((fabs(x - J[2]) <= J[1]) && (fabs(y - J[3]) <= J[0])) {
        AT(ub, 0, i, j) = J[4];
        AT(ub, 1, i, j) = J[4] * vJet;
        AT(ub, 2, i, j) = 0.;
        AT(ub, 3, i, j) = 0.5 * J[4] * (vJet * vJet) + p->cv * J[4] * J[10] * J[5];
        if (p->neq >= 5) AT(ub, 4, i, j) = J[4] * J[9];
        if (p->neq >= 6) AT(ub, 5, i, j) = J[4];
      }
    }
}

/* ------------------------------------------------------------------ */
/* Ghost fills, intra-rank part, MPIP variant (codes 1..12).
 *   which == 2 : src/boundaryII.f90:35-514 on `up`, two layers
 *   which == 1 : src/boundaryI.f90:34-396  on `u`,  one layer
 * The list is applied in order; compat bit1 keeps boundaryI.f90:87 (OUTFYB
 * writes block `nb` = list position instead of nbs). */
#define A(b, v, i, j) AT(blk(p, arr, b), v, i, j)
static void bc_box(const wxo_params *p, double *arr, int nbs, int side, int which, int listpos) {
  const int neq = p->neq;
  const int refl = p->bc[side];
  if (which == 2) {
    if (side == 0) {           /* boundaryII.f90:38-57 */
      for (int jj = 2, i = -1; i <= 0; ++i, --jj)
        for (int j = -1; j <= NY + 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j) = (refl && k == 1) ? -A(nbs, k, jj, j) : A(nbs, k, jj, j);
    } else if (side == 1) {    /* :62-81 */
      for (int jj = NX, i = NX + 1; i <= NX + 2; ++i, --jj)
        for (int j = -1; j <= NY + 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j) = (refl && k == 1) ? -A(nbs, k, jj, j) : A(nbs, k, jj, j);
    } else if (side == 2) {    /* :95-114 */
      for (int jj = 2, j = -1; j <= 0; ++j, --jj)
        for (int i = -1; i <= NX + 2; ++i)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j) = (refl && k == 2) ? -A(nbs, k, i, jj) : A(nbs, k, i, jj);
    } else {                   /* :119-138 */
      for (int jj = NY, j = NY + 1; j <= NY + 2; ++j, --jj)
        for (int i = -1; i <= NX + 2; ++i)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j) = (refl && k == 2) ? -A(nbs, k, i, jj) : A(nbs, k, i, jj);
    }
  } else {
    if (side == 0) {           /* boundaryI.f90:37-49 */
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, 0, j) = (refl && k == 1) ? -A(nbs, k, 1, j) : A(nbs, k, 1, j);
    } else if (side == 1) {    /* :54-65 */
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, NX + 1, j) = (refl && k == 1) ? -A(nbs, k, NX, j) : A(nbs, k, NX, j);
    } else if (side == 2) {    /* :77-88 */
      int tgt = nbs;
      if (!refl && (p->compat & 2)) tgt = listpos;     /* :87 writes u(nb,...) */
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k)
          A(tgt, k, i, 0) = (refl && k == 2) ? -A(tgt, k, i, 1) : A(tgt, k, i, 1);
    } else {                   /* :93-104 */
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, i, NY + 1) = (refl && k == 2) ? -A(nbs, k, i, NY) : A(nbs, k, i, NY);
    }
  }
}

static inline double mean4(const wxo_params *p, double *arr, int b, int k, int i1, int j1) {
  /* 0.25*sum(a(i1:i1+1, j1:j1+1)) in Fortran column-major order */
  return 0.25 * (((A(b, k, i1, j1) + A(b, k, i1 + 1, j1)) + A(b, k, i1, j1 + 1)) + A(b, k, i1 + 1, j1 + 1));
}

static void bc_entryII(const wxo_params *p, double *arr, int nbs, int n1, int code) {
  const int neq = p->neq;
  const int nx2 = NX / 2, ny2 = NY / 2;
  switch (code) {
    case 1:  /* boundaryII.f90:143-149 */
      for (int j = -1; j <= NY + 2; ++j)
        for (int k = 0; k < neq; ++k) {
          A(nbs, k, -1, j) = A(n1, k, NX - 1, j);
          A(nbs, k, 0, j) = A(n1, k, NX, j);
        }
      for (int j = -1; j <= NY + 2; ++j)
        for (int k = 0; k < neq; ++k) {
          A(n1, k, NX + 1, j) = A(nbs, k, 1, j);
          A(n1, k, NX + 2, j) = A(nbs, k, 2, j);
        }
      break;
    case 2: case 3: {  /* :151-197 (n1 half) and :200-230 (n2 half) */
      int off = (code == 3) ? ny2 : 0;
      for (int j = 1; j <= ny2; ++j)
        for (int i = -1; i <= 0; ++i)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j + off) = mean4(p, arr, n1, k, 2 * i - 1 + NX, 2 * j - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 2) {
          A(nbs, k, 0, 0) = mean4(p, arr, n1, k, NX - 1, -1);
          A(nbs, k, -1, 0) = mean4(p, arr, n1, k, NX - 3, -1);
          A(nbs, k, 0, -1) = A(nbs, k, 0, 0);
          A(nbs, k, -1, -1) = A(nbs, k, -1, 0);
        } else {
          A(nbs, k, 0, NY + 1) = mean4(p, arr, n1, k, NX - 1, NY + 1);
          A(nbs, k, -1, NY + 1) = mean4(p, arr, n1, k, NX - 3, NY + 1);
          A(nbs, k, 0, NY + 2) = A(nbs, k, 0, NY + 1);
          A(nbs, k, -1, NY + 2) = A(nbs, k, -1, NY + 1);
        }
      }
      for (int i = 1; i <= 2; ++i)
        for (int j = -1; j <= NY + 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(n1, k, i + NX, j) = A(nbs, k, (i + 1) / 2, (j + 1) / 2 + off);
      break;
    }
    case 4:  /* :233-239 */
      for (int j = -1; j <= NY + 2; ++j)
        for (int k = 0; k < neq; ++k) {
          A(nbs, k, NX + 1, j) = A(n1, k, 1, j);
          A(nbs, k, NX + 2, j) = A(n1, k, 2, j);
        }
      for (int j = -1; j <= NY + 2; ++j)
        for (int k = 0; k < neq; ++k) {
          A(n1, k, -1, j) = A(nbs, k, NX - 1, j);
          A(n1, k, 0, j) = A(nbs, k, NX, j);
        }
      break;
    case 5: case 6: {  /* :241-287 and :290-320 */
      int off = (code == 6) ? ny2 : 0;
      for (int i = NX + 1; i <= NX + 2; ++i)
        for (int j = 1; j <= ny2; ++j)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i, j + off) = mean4(p, arr, n1, k, 2 * (i - NX) - 1, 2 * j - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 5) {
          A(nbs, k, NX + 1, 0) = mean4(p, arr, n1, k, 1, -1);
          A(nbs, k, NX + 2, 0) = mean4(p, arr, n1, k, 3, -1);
          A(nbs, k, NX + 1, -1) = A(nbs, k, NX + 1, 0);
          A(nbs, k, NX + 2, -1) = A(nbs, k, NX + 2, 0);
        } else {
          A(nbs, k, NX + 1, NY + 1) = mean4(p, arr, n1, k, 1, NY + 1);
          A(nbs, k, NX + 2, NY + 1) = mean4(p, arr, n1, k, 3, NY + 1);
          A(nbs, k, NX + 1, NY + 2) = A(nbs, k, NX + 1, NY + 1);
          A(nbs, k, NX + 2, NY + 2) = A(nbs, k, NX + 2, NY + 1);
        }
      }
      for (int i = 1; i <= 2; ++i)
        for (int j = -1; j <= NY + 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(n1, k, i - 2, j) = A(nbs, k, (i + 1) / 2 + NX - 1, (j + 1) / 2 + off);
      break;
    }
    case 7:  /* :323-329 */
      for (int i = -1; i <= NX + 2; ++i)
        for (int k = 0; k < neq; ++k) {
          A(nbs, k, i, -1) = A(n1, k, i, NY - 1);
          A(nbs, k, i, 0) = A(n1, k, i, NY);
        }
      for (int i = -1; i <= NX + 2; ++i)
        for (int k = 0; k < neq; ++k) {
          A(n1, k, i, NY + 1) = A(nbs, k, i, 1);
          A(n1, k, i, NY + 2) = A(nbs, k, i, 2);
        }
      break;
    case 8: case 9: {  /* :331-377 and :380-410 */
      int off = (code == 9) ? nx2 : 0;
      for (int i = 1; i <= nx2; ++i)
        for (int j = -1; j <= 0; ++j)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i + off, j) = mean4(p, arr, n1, k, 2 * i - 1, 2 * j - 1 + NY);
      for (int k = 0; k < neq; ++k) {
        if (code == 8) {
          A(nbs, k, 0, 0) = mean4(p, arr, n1, k, -1, NY - 1);
          A(nbs, k, 0, -1) = mean4(p, arr, n1, k, -1, NY - 3);
          A(nbs, k, -1, 0) = A(nbs, k, 0, 0);
          A(nbs, k, -1, -1) = A(nbs, k, -1, 0);
        } else {
          A(nbs, k, NX + 1, 0) = mean4(p, arr, n1, k, NX + 1, NY - 1);
          A(nbs, k, NX + 1, -1) = mean4(p, arr, n1, k, NX + 1, NY - 3);
          A(nbs, k, NX + 2, 0) = A(nbs, k, NX + 1, 0);
          A(nbs, k, NX + 2, -1) = A(nbs, k, NX + 1, -1);
        }
      }
      for (int i = -1; i <= NX + 2; ++i)
        for (int j = 1; j <= 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(n1, k, i, j + NY) = A(nbs, k, (i + 1) / 2 + off, (j + 1) / 2);
      break;
    }
    case 10:  /* :413-419 (the reference writes nyp1:nxmax, equal since nx=ny) */
      for (int i = -1; i <= NX + 2; ++i)
        for (int k = 0; k < neq; ++k) {
          A(nbs, k, i, NY + 1) = A(n1, k, i, 1);
          A(nbs, k, i, NY + 2) = A(n1, k, i, 2);
        }
      for (int i = -1; i <= NX + 2; ++i)
        for (int k = 0; k < neq; ++k) {
          A(n1, k, i, -1) = A(nbs, k, i, NY - 1);
          A(n1, k, i, 0) = A(nbs, k, i, NY);
        }
      break;
    case 11: case 12: {  /* :421-467 and :470-500 */
      int off = (code == 12) ? nx2 : 0;
      for (int i = 1; i <= nx2; ++i)
        for (int j = NY + 1; j <= NY + 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(nbs, k, i + off, j) = mean4(p, arr, n1, k, 2 * i - 1, 2 * (j - NY) - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 11) {
          A(nbs, k, 0, NY + 1) = mean4(p, arr, n1, k, -1, 1);
          A(nbs, k, 0, NY + 2) = mean4(p, arr, n1, k, -1, 3);
          A(nbs, k, -1, NY + 1) = A(nbs, k, 0, NY + 1);
          A(nbs, k, -1, NY + 2) = A(nbs, k, 0, NY + 2);
        } else {
          A(nbs, k, NX + 1, NY + 1) = mean4(p, arr, n1, k, NX + 1, 1);
          A(nbs, k, NX + 1, NY + 2) = mean4(p, arr, n1, k, NX + 1, 3);
          A(nbs, k, NX + 2, NY + 1) = A(nbs, k, NX + 1, NY + 1);
          A(nbs, k, NX + 2, NY + 2) = A(nbs, k, NX + 1, NY + 2);
        }
      }
      for (int i = -1; i <= NX + 2; ++i)
        for (int j = 1; j <= 2; ++j)
          for (int k = 0; k < neq; ++k)
            A(n1, k, i, j - 2) = A(nbs, k, (i + 1) / 2 + off, (j + 1) / 2 + NY - 1);
      break;
    }
    default: break;
  }
}

static void bc_entryI(const wxo_params *p, double *arr, int nbs, int n1, int code) {
  const int neq = p->neq;
  const int nx2 = NX / 2, ny2 = NY / 2;
  switch (code) {
    case 1:  /* boundaryI.f90:109-115 */
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(nbs, k, 0, j) = A(n1, k, NX, j);
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(n1, k, NX + 1, j) = A(nbs, k, 1, j);
      break;
    case 2: case 3: {  /* :117-151 and :154-175 */
      int off = (code == 3) ? ny2 : 0;
      for (int j = 1; j <= ny2; ++j)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, 0, j + off) = mean4(p, arr, n1, k, NX - 1, 2 * j - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 2) A(nbs, k, 0, 0) = mean4(p, arr, n1, k, NX - 1, -1);
        else A(nbs, k, 0, NY + 1) = mean4(p, arr, n1, k, NX - 1, NY + 1);
      }
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(n1, k, NX + 1, j) = A(nbs, k, 1, (j + 1) / 2 + off);
      break;
    }
    case 4:  /* :178-184 */
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(nbs, k, NX + 1, j) = A(n1, k, 1, j);
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(n1, k, 0, j) = A(nbs, k, NX, j);
      break;
    case 5: case 6: {  /* :186-220 and :223-244 */
      int off = (code == 6) ? ny2 : 0;
      for (int j = 1; j <= ny2; ++j)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, NX + 1, j + off) = mean4(p, arr, n1, k, 1, 2 * j - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 5) A(nbs, k, NX + 1, 0) = mean4(p, arr, n1, k, 1, -1);
        else A(nbs, k, NX + 1, NY + 1) = mean4(p, arr, n1, k, 1, NY + 1);
      }
      for (int j = 0; j <= NY + 1; ++j)
        for (int k = 0; k < neq; ++k) A(n1, k, 0, j) = A(nbs, k, NX, (j + 1) / 2 + off);
      break;
    }
    case 7:  /* :247-253 */
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(nbs, k, i, 0) = A(n1, k, i, NY);
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(n1, k, i, NY + 1) = A(nbs, k, i, 1);
      break;
    case 8: case 9: {  /* :255-289 and :292-313 */
      int off = (code == 9) ? nx2 : 0;
      for (int i = 1; i <= nx2; ++i)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, i + off, 0) = mean4(p, arr, n1, k, 2 * i - 1, NY - 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 8) A(nbs, k, 0, 0) = mean4(p, arr, n1, k, -1, NY - 1);
        else A(nbs, k, NX + 1, 0) = mean4(p, arr, n1, k, NX + 1, NY - 1);
      }
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(n1, k, i, NY + 1) = A(nbs, k, (i + 1) / 2 + off, 1);
      break;
    }
    case 10:  /* :316-322 */
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(nbs, k, i, NY + 1) = A(n1, k, i, 1);
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(n1, k, i, 0) = A(nbs, k, i, NY);
      break;
    case 11: case 12: {  /* :324-358 and :361-382 */
      int off = (code == 12) ? nx2 : 0;
      for (int i = 1; i <= nx2; ++i)
        for (int k = 0; k < neq; ++k)
          A(nbs, k, i + off, NY + 1) = mean4(p, arr, n1, k, 2 * i - 1, 1);
      for (int k = 0; k < neq; ++k) {
        if (code == 11) A(nbs, k, 0, NY + 1) = mean4(p, arr, n1, k, -1, 1);
        else A(nbs, k, NX + 1, NY + 1) = mean4(p, arr, n1, k, NX + 1, 1);
      }
      for (int i = 0; i <= NX + 1; ++i)
        for (int k = 0; k < neq; ++k) A(n1, k, i, 0) = A(nbs, k, (i + 1) / 2 + off, NY);
      break;
    }
    default: break;
  }
}

void wxo_boundary(const wxo_params *p, wxo_mesh *m, double time, int which) {
  double *arr = (which == 2) ? m->up : m->u;
  for (int e = 0; e < m->ninternal; ++e) {
    int nbs = m->bounds[e * 3 + 0];
    int n1 = m->bounds[e * 3 + 1];
    int code = m->bounds[e * 3 + 2];
    if (code < 0) bc_box(p, arr, nbs, -code - 1, which, e + 1);
    else if (which == 2) bc_entryII(p, arr, nbs, n1, code);
    else bc_entryI(p, arr, nbs, n1, code);
    if (p->otherb) wxo_impose_jet(p, m, blk(p, arr, nbs), nbs, time);  /* II :506-511, I :388-393 */
  }
}
#undef A

/* ------------------------------------------------------------------ */
/* Cooling.  src/coolingh.f90:41-76 rates */
static double alpha_rec(double T) { return 2.55e-13 * pow(1.e4 / T, 0.79); }
static double colf(double T) { return 5.83e-11 * sqrt(T) * exp(-(157828. / T)); }
static double betah(double T) {
  double a = 157890. / T;
  return 1.133e-24 / sqrt(a) * (-0.0713 + 0.5 * log(a) + 0.640 * pow(a, -0.33333));
}

/* src/coolingh.f90:88-165 ALOSS (X2, DT unused as in the reference) */
double wxo_aloss(double x1, double x2, double dt, double den, double dh0, double te0) {
  (void)x2; (void)dt;
  const double XION = 2.179e-11, XO = 1.e-3;
  const double C0 = 0.5732, C1 = 1.8288e-5, C2 = -1.15822e-10, C3 = 9.4288e-16;
  const double D0 = 0.5856, D1 = 1.55083e-5, D2 = -9.669e-12, D3 = 5.716e-19;
  const double ENK = 118409., EN = 1.634E-11;
  double TE = dmax(te0, 1000.);
  double DH = den;
  double DHP = (1. - x1) * DH;
  double DE = DHP + 1.E-4 * DH;
  double DOI = XO * dh0;
  double DOII = XO * DHP;
  double OMEGA = 0.0;
  if (TE <= 55000.) OMEGA = C0 + TE * (C1 + TE * (C2 + TE * C3));
  if (TE >= 72000.) OMEGA = D0 + TE * (D1 + TE * (D2 + TE * D3));
  if (TE > 55000. && TE < 72000.) {
    double OMEGAL = C0 + TE * (C1 + TE * (C2 + TE * C3));
    double OMEGAH = D0 + TE * (D1 + TE * (D2 + TE * D3));
    double FRAC = (TE - 55000.) / 17000.;
    OMEGA = (1. - FRAC) * OMEGAL + FRAC * OMEGAH;
  }
  double QLA = 8.6287E-6 / (2. * sqrt(TE)) * OMEGA * exp(-(ENK / TE));
  double ECOLL = DE * dh0 * QLA * EN;
  ECOLL = dmax(ECOLL, 0.);
  double CION = 5.834E-11 * sqrt(TE) * exp(-(1.579E5 / TE));
  double EION = DE * dh0 * CION * XION;
  double EREC = DE * DHP * (betah(TE));
  EREC = dmax(EREC, 0.);
  double TM = 1. / TE;
  double T2 = TM * TM;
  double EOI = DE * DOI * pow(10., 1381465 * T2 - 12328.69 * TM - 19.82621);
  double EOII = DE * DOII * pow(10., -(2061075. * T2) - 14596.24 * TM - 19.01402);
  EOI = dmax(EOI, 0.);
  EOII = dmax(EOII, 0.);
  double BETAF = 1.3 * 1.42E-27 * pow(TE, 0.5);
  double HIICOOL = DE * DHP * BETAF;
  double EQUIL = (1.0455E-18 / pow(TE, 0.63)) * (1. - exp(-pow(TE * 1.E-5, 1.63))) * DE * den + HIICOOL;
  double FR = 0.0;
  if (TE <= 44770.) FR = 0.;
  if (TE >= 54770.) FR = 1.;
  if (TE > 44770. && TE < 54770.) {
    double EX2 = exp(-(2. * (TE - 49770.) / 500.));
    double TANH = (1. - EX2) / (1. + EX2);
    FR = 0.5 * (1. + TANH);
  }
  return ECOLL + EION + (EREC + 7.033 * (EOI + EOII)) * (1. - FR) + EQUIL * FR;
}

/* src/coolingh.f90:171-259 atomic (tau unused) */
void wxo_atomic(const wxo_params *p, double dt, double *uu) {
  const double xi = 1.e-4, boltzm = 1.3807e-16;
  double prim[WXO_MAXEQ], T;
  wxo_uprim(p, prim, uu, &T);
  double col = colf(T);
  double rec = alpha_rec(T);
  double y0 = uu[4] / uu[0];
  double dh = uu[0];
  double a = rec + col;
  double b = -((2. + xi) * rec + (1. + xi) * col);
  double c = (1. + xi) * rec;
  double d = sqrt(b * b - 4. * a * c);
  double g0 = (2. * a * y0 + b + d) / (2. * a * y0 + b - d);
  double e = exp(-(d * dh * dt));
  double y1 = (-b - d * (1. + g0 * e) / (1. - g0 * e)) / (2. * a);
  y1 = dmin(y1, 1.);
  y1 = dmax(y1, 0.0001);
  uu[4] = y1 * uu[0];
  double dh0 = uu[4];
  double al = wxo_aloss(y0, y1, dt, dh, dh0, T) / (dh * dh);
  double tprime = 10.;
  double ce = (2. * dh * al) / (3. * boltzm * T);
  double t1 = tprime + (T - tprime) * exp(-(ce * dt));
  t1 = dmax(t1, 0.1 * T);
  t1 = dmin(t1, 10. * T);
  t1 = dmax(t1, tprime);
  uu[3] = p->cv * (2. * uu[0] - uu[4]) * t1 / p->Tempsc + 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]);
}

/* src/dmc_module.f90:14-32 cooldmc; row 42 (T == 1e8 exactly, H9) clamps to row 41 */
double wxo_cooldmc(const wxo_params *p, double T) {
  if (T > 1e8) return 0.21e-26 * sqrt(T);
  int if1 = (int)(log10(T) * 10) - 39;
  int if2 = if1 + 1;
  if (if1 < 1) if1 = 1;
  if (if2 > 41) if2 = 41;
  double T0 = p->cooltab[2 * (if1 - 1)], c0 = p->cooltab[2 * (if1 - 1) + 1];
  double T1 = p->cooltab[2 * (if2 - 1)], c1 = p->cooltab[2 * (if2 - 1) + 1];
  if (T1 == T0) return c0;
  return (c1 - c0) * (T - T0) / (T1 - T0) + c0;
}

/* src/coolingh.f90:5-35 over cells 0..n+1 of every leaf */
void wxo_coolingh(const wxo_params *p, wxo_mesh *m, double dt) {
  double uu[WXO_MAXEQ];
  for (int nb = 1; nb <= m->nbmax; ++nb) {
    if (!m->leaf[nb]) continue;
    double *ub = blk(p, m->u, nb);
    for (int i = 0; i <= NX + 1; ++i)
      for (int j = 0; j <= NY + 1; ++j) {
        for (int k = 0; k < p->neq; ++k) uu[k] = AT(ub, k, i, j);
        wxo_atomic(p, dt, uu);
        for (int k = 0; k < p->neq; ++k) AT(ub, k, i, j) = uu[k];
      }
  }
}

/* src/coolingdmc.f90:6-54 over cells 0..n+1 of every leaf (updates primit there) */
void wxo_coolingdmc(const wxo_params *p, wxo_mesh *m, double dt) {
  const double Tmin = 1.e4;
  double uu[WXO_MAXEQ], q[WXO_MAXEQ], T;
  for (int nb = 1; nb <= m->nbmax; ++nb) {
    if (!m->leaf[nb]) continue;
    double *ub = blk(p, m->u, nb), *pb = blk(p, m->primit, nb);
    for (int i = 0; i <= NX + 1; ++i)
      for (int j = 0; j <= NY + 1; ++j) {
        for (int k = 0; k < p->neq; ++k) uu[k] = AT(ub, k, i, j);
        wxo_uprim(p, q, uu, &T);
        for (int k = 0; k < p->neq; ++k) AT(pb, k, i, j) = q[k];
        if (T > Tmin) {
          double Eth0 = p->cv * AT(pb, 3, i, j);
          double Aloss = wxo_cooldmc(p, T);
          double dens = AT(pb, 0, i, j);
          double Ce = (Aloss * (dens * dens)) / (Eth0 * p->Psc);
          AT(pb, 3, i, j) = AT(pb, 3, i, j) * exp(-(Ce * dt));
          AT(ub, 3, i, j) = AT(ub, 3, i, j) - Eth0 + p->cv * AT(pb, 3, i, j);
        }
      }
  }
}

/* src/timestep.f90:7-61 (single rank: the allreduce MIN is the identity) */
double wxo_timestep(const wxo_params *p, const wxo_mesh *m) {
  double dtp = 1.e30, dtcool = 1e30;
  for (int nb = 1; nb <= m->nbmax; ++nb) {
    if (!m->leaf[nb]) continue;
    int lev = LP(m, nb, 1);
    const double *pb = blk(p, m->primit, nb);
    for (int i = 1; i <= NX; ++i)
      for (int j = 1; j <= NY; ++j) {
        double cs = sound(p, AT(pb, 3, i, j), AT(pb, 0, i, j));
        dtp = dmin(dtp, p->dx[lev - 1] / (fabs(AT(pb, 1, i, j)) + cs));
        dtp = dmin(dtp, p->dy[lev - 1] / (fabs(AT(pb, 2, i, j)) + cs));
        if (p->cooling == 2) {
          double T = (AT(pb, 3, i, j) / AT(pb, 0, i, j)) * p->Tempsc;
          if (T > 1e4) {
            double aloss = wxo_cooldmc(p, T) * (AT(pb, 0, i, j) * AT(pb, 0, i, j));
            dtcool = dmin(dtcool, p->cv * AT(pb, 3, i, j) / aloss);
          }
        }
      }
  }
  return dmin(p->cfl * dtp, 0.4 * dtcool * p->Psc / p->tsc);
}

/* src/tstep.f90:4-126 */
int wxo_tstep(const wxo_params *p, wxo_mesh *m, double time, double dt) {
  const size_t bsz = (size_t)p->neq * NC;
  double *f = (double *)malloc(bsz * sizeof(double));
  double *g = (double *)malloc(bsz * sizeof(double));
  int err = 0;
  memset(f, 0, bsz * sizeof(double));
  memset(g, 0, bsz * sizeof(double));
  wxo_locatebounds(p, m);                                      /* :14 */
  double dtm = dt / 2.;
  for (int nb = 1; nb <= m->nbmax; ++nb) {                      /* :22-46 */
    if (!m->leaf[nb]) continue;
    err |= wxo_fluxes(p, m, nb, 1, f, g);
    wxo_step(p, m, nb, dtm, f, g);
  }
  wxo_boundary(p, m, time, 2);                                  /* :53 */
  for (int nb = 1; nb <= m->nbmax; ++nb) {                      /* :55-88 */
    if (!m->leaf[nb]) continue;
    wxo_calcprim(p, blk(p, m->up, nb), blk(p, m->primit, nb));
    err |= wxo_fluxes(p, m, nb, 2, f, g);
    wxo_step(p, m, nb, dt, f, g);
    if (LP(m, nb, 1) >= p->nlevs - 1) wxo_viscosity(p, m, nb);
    memcpy(blk(p, m->u, nb), blk(p, m->up, nb), bsz * sizeof(double));
  }
  wxo_boundary(p, m, time, 1);                                  /* :92 */
  if (p->cooling == 1) wxo_coolingh(p, m, dt * p->tsc);         /* :107-109 */
  if (p->cooling == 2) {
    wxo_coolingdmc(p, m, dt * p->tsc);                          /* :119-122 */
  } else {
    for (int nb = 1; nb <= m->nbmax; ++nb)                      /* :95-102, :110-117 */
      if (m->leaf[nb]) wxo_calcprim(p, blk(p, m->u, nb), blk(p, m->primit, nb));
  }
  free(f);
  free(g);
  return err;
}

/* ------------------------------------------------------------------ */
/* src/locatebounds.f90:29-540, MPIP branches, one rank (no external list). */
int wxo_locatebounds(const wxo_params *p, wxo_mesh *m) {
  (void)p;
  int cap = m->cap;
  uint8_t *bnd = (uint8_t *)calloc((size_t)(cap + 1) * 4, 1);
  int n = 0;
#define ADD(a, b, c) do { m->bounds[n * 3] = (a); m->bounds[n * 3 + 1] = (b); m->bounds[n * 3 + 2] = (c); ++n; } while (0)
  /* face f: 0 left(lp5), 1 right(lp6), 2 bottom(lp7), 3 top(lp8) */
  static const int opp[4] = {1, 0, 3, 2};
  for (int nb = 1; nb <= m->nbmax; ++nb) {
    if (!m->leaf[nb]) continue;
    for (int f = 0; f < 4; ++f) {
      if (bnd[nb * 4 + f]) continue;
      int nbr = LP(m, nb, 5 + f);
      if (nbr == -1) { bnd[nb * 4 + f] = 1; ADD(nb, 0, -(f + 1)); continue; }
      if (!(LP(m, nbr, 1) >= LP(m, nb, 1))) continue;
      if (m->leaf[nbr]) {
        ADD(nb, nbr, 1 + 3 * f);
        bnd[nb * 4 + f] = 1;
        bnd[nbr * 4 + opp[f]] = 1;
      } else {
        int n1, n2;
        bnd[nb * 4 + f] = 1;
        if (f == 0) { n1 = LP(m, LP(m, nbr, 3), 6); n2 = LP(m, n1, 8); }      /* :130-131 */
        else if (f == 1) { n1 = LP(m, nbr, 3); n2 = LP(m, n1, 8); }          /* :237-238 */
        else if (f == 2) { n1 = LP(m, LP(m, nbr, 3), 8); n2 = LP(m, n1, 6); } /* :345-346 */
        else { n1 = LP(m, nbr, 3); n2 = LP(m, n1, 6); }                      /* :452-453 */
        ADD(nb, n1, 2 + 3 * f);
        ADD(nb, n2, 3 + 3 * f);
      }
    }
  }
#undef ADD
  m->ninternal = n;
  free(bnd);
  return n;
}

/* src/criteria.f90:4-53 */
void wxo_criteria(const wxo_params *p, const wxo_mesh *m, int nb, int32_t *irefup, int32_t *irefdown) {
  const double floor_ = 1e-15;
  double grad = 0.;
  const double *q = blk(p, m->primit, nb);
  for (int i = 1; i <= NX; ++i)
    for (int j = 1; j <= NY; ++j) {
      double am = dmax(floor_, fabs(AT(q, 3, i, j)));
      double am2 = dmax(floor_, fabs(AT(q, 0, i, j)));
      double c = AT(q, 3, i, j);
      grad = dmax(grad, dmax(dmax(dmax(fabs(AT(q, 3, i + 1, j) - c) / am, fabs(AT(q, 3, i - 1, j) - c) / am),
                                  fabs(AT(q, 3, i, j + 1) - c) / am), fabs(AT(q, 3, i, j - 1) - c) / am));
      c = AT(q, 0, i, j);
      grad = dmax(grad, dmax(dmax(dmax(fabs(AT(q, 0, i + 1, j) - c) / am2, fabs(AT(q, 0, i - 1, j) - c) / am2),
                                  fabs(AT(q, 0, i, j + 1) - c) / am2), fabs(AT(q, 0, i, j - 1) - c) / am2));
      if (grad >= p->tol) {
        if (LP(m, nb, 1) < p->nlevs) { irefup[nb] = 1; irefdown[nb] = 0; }
        return;
      }
    }
  if (grad <= p->tol2) irefdown[nb] = 1;
}

/* src/refine.f90:30-42 (data) and :227-261 (calcprim, icoord) */
void wxo_refine(const wxo_params *p, wxo_mesh *m, int nb, int s1, int s2, int s3, int s4) {
  const int nx2 = NX / 2, ny2 = NY / 2;
  const double *pu = blk(p, m->u, nb);
  double *b1 = blk(p, m->u, s1), *b2 = blk(p, m->u, s2), *b3 = blk(p, m->u, s3), *b4 = blk(p, m->u, s4);
  for (int i = -1; i <= NX + 2; ++i)
    for (int j = -1; j <= NY + 2; ++j)
      for (int k = 0; k < p->neq; ++k) {
        AT(b1, k, i, j) = AT(pu, k, (i + 1) / 2, (j + 1) / 2);
        AT(b2, k, i, j) = AT(pu, k, (i + 1) / 2 + nx2, (j + 1) / 2);
        AT(b3, k, i, j) = AT(pu, k, (i + 1) / 2, (j + 1) / 2 + ny2);
        AT(b4, k, i, j) = AT(pu, k, (i + 1) / 2 + nx2, (j + 1) / 2 + ny2);
      }
  int nx0 = m->icoord[nb * 2] * 2, ny0 = m->icoord[nb * 2 + 1] * 2;
  int sons[4] = {s1, s2, s3, s4};
  int ox[4] = {0, NX, 0, NX}, oy[4] = {0, 0, NY, NY};
  for (int s = 0; s < 4; ++s) {
    wxo_calcprim(p, blk(p, m->u, sons[s]), blk(p, m->primit, sons[s]));
    m->icoord[sons[s] * 2] = nx0 + ox[s];
    m->icoord[sons[s] * 2 + 1] = ny0 + oy[s];
  }
}

/* src/updatelpup.f90:4-191 (MPIP pointer forms) */
void wxo_updatelpup(const wxo_params *p, wxo_mesh *m, int nb, int s1, int s2, int s3, int s4,
                    int32_t *irefup, int32_t *irefdown) {
  (void)p;
  int nl2 = LP(m, nb, 1) + 1;
  m->nbleafs += 3;
  int nsr;
  /* son1 :26-63 */
  SETLP(m, s1, 1, nl2); SETLP(m, s1, 2, nb); SETLP(m, s1, 4, 1); SETLP(m, s1, 6, s2); SETLP(m, s1, 8, s3);
  if (LP(m, nb, 5) == -1 || m->leaf[LP(m, nb, 5)]) SETLP(m, s1, 5, LP(m, nb, 5));
  else { nsr = LP(m, LP(m, LP(m, nb, 5), 3), 6); SETLP(m, s1, 5, nsr); SETLP(m, nsr, 6, s1); }
  if (LP(m, nb, 7) == -1 || m->leaf[LP(m, nb, 7)]) SETLP(m, s1, 7, LP(m, nb, 7));
  else { nsr = LP(m, LP(m, LP(m, nb, 7), 3), 8); SETLP(m, s1, 7, nsr); SETLP(m, nsr, 8, s1); }
  /* son2 :65-98 */
  SETLP(m, s2, 1, nl2); SETLP(m, s2, 2, nb); SETLP(m, s2, 4, 2); SETLP(m, s2, 5, s1); SETLP(m, s2, 8, s4);
  if (LP(m, nb, 6) == -1 || m->leaf[LP(m, nb, 6)]) SETLP(m, s2, 6, LP(m, nb, 6));
  else { nsr = LP(m, LP(m, nb, 6), 3); SETLP(m, s2, 6, nsr); SETLP(m, nsr, 5, s2); }
  if (LP(m, nb, 7) == -1 || m->leaf[LP(m, nb, 7)]) SETLP(m, s2, 7, LP(m, nb, 7));
  else { nsr = LP(m, LP(m, LP(m, LP(m, nb, 7), 3), 8), 6); SETLP(m, s2, 7, nsr); SETLP(m, nsr, 8, s2); }
  /* son3 :100-132 */
  SETLP(m, s3, 1, nl2); SETLP(m, s3, 2, nb); SETLP(m, s3, 4, 3); SETLP(m, s3, 6, s4); SETLP(m, s3, 7, s1);
  if (LP(m, nb, 5) == -1 || m->leaf[LP(m, nb, 5)]) SETLP(m, s3, 5, LP(m, nb, 5));
  else { nsr = LP(m, LP(m, LP(m, LP(m, nb, 5), 3), 8), 6); SETLP(m, s3, 5, nsr); SETLP(m, nsr, 6, s3); }
  if (LP(m, nb, 8) == -1 || m->leaf[LP(m, nb, 8)]) SETLP(m, s3, 8, LP(m, nb, 8));
  else { nsr = LP(m, LP(m, nb, 8), 3); SETLP(m, s3, 8, nsr); SETLP(m, nsr, 7, s3); }
  /* son4 :134-170 */
  SETLP(m, s4, 1, nl2); SETLP(m, s4, 2, nb); SETLP(m, s4, 4, 4); SETLP(m, s4, 5, s3); SETLP(m, s4, 7, s2);
  if (LP(m, nb, 6) == -1 || m->leaf[LP(m, nb, 6)]) SETLP(m, s4, 6, LP(m, nb, 6));
  else { nsr = LP(m, LP(m, LP(m, nb, 6), 3), 8); SETLP(m, s4, 6, nsr); SETLP(m, nsr, 5, s4); }
  if (LP(m, nb, 8) == -1 || m->leaf[LP(m, nb, 8)]) SETLP(m, s4, 8, LP(m, nb, 8));
  else { nsr = LP(m, LP(m, LP(m, nb, 8), 3), 6); SETLP(m, s4, 8, nsr); SETLP(m, nsr, 7, s4); }
  /* :175-189 */
  m->leaf[nb] = 0;
  m->leaf[s1] = m->leaf[s2] = m->leaf[s3] = m->leaf[s4] = 1;
  irefup[nb] = 0; irefdown[nb] = 0;
  irefdown[s1] = irefdown[s2] = irefdown[s3] = irefdown[s4] = 0;
}

/* src/coarse.f90:182-366 (MPIP branch, all sons local) */
void wxo_coarse(const wxo_params *p, wxo_mesh *m, int dad, int s1, int s2, int s3, int s4) {
  const int nx2 = NX / 2, ny2 = NY / 2;
  double *d = blk(p, m->u, dad);
  double *b1 = blk(p, m->u, s1), *b2 = blk(p, m->u, s2), *b3 = blk(p, m->u, s3), *b4 = blk(p, m->u, s4);
#define M4(b, k, i1, j1) (0.25 * (((AT(b, k, i1, j1) + AT(b, k, (i1) + 1, j1)) + AT(b, k, i1, (j1) + 1)) + AT(b, k, (i1) + 1, (j1) + 1)))
  for (int i = 0; i <= nx2; ++i)
    for (int j = 0; j <= ny2; ++j)
      for (int k = 0; k < p->neq; ++k) AT(d, k, i, j) = M4(b1, k, 2 * i - 1, 2 * j - 1);
  for (int i = 1; i <= nx2 + 1; ++i)
    for (int j = 0; j <= ny2; ++j)
      for (int k = 0; k < p->neq; ++k) AT(d, k, i + nx2, j) = M4(b2, k, 2 * i - 1, 2 * j - 1);
  for (int i = 0; i <= nx2; ++i)
    for (int j = 1; j <= ny2 + 1; ++j)
      for (int k = 0; k < p->neq; ++k) AT(d, k, i, j + ny2) = M4(b3, k, 2 * i - 1, 2 * j - 1);
  for (int i = 1; i <= nx2 + 1; ++i)
    for (int j = 1; j <= ny2 + 1; ++j)
      for (int k = 0; k < p->neq; ++k) AT(d, k, i + nx2, j + ny2) = M4(b4, k, 2 * i - 1, 2 * j - 1);
#undef M4
  for (int k = 0; k < p->neq; ++k) {                          /* :349-358 */
    for (int j = -1; j <= NY + 2; ++j) AT(d, k, -1, j) = AT(d, k, 0, j);
    for (int j = -1; j <= NY + 2; ++j) AT(d, k, NX + 2, j) = AT(d, k, NX + 1, j);
    for (int i = -1; i <= NX + 2; ++i) AT(d, k, i, -1) = AT(d, k, i, 0);
    for (int i = -1; i <= NX + 2; ++i) AT(d, k, i, NY + 2) = AT(d, k, i, NY + 1);
    AT(d, k, -1, -1) = AT(d, k, 0, -1);
    AT(d, k, NX + 2, -1) = AT(d, k, NX + 2, 0);
    AT(d, k, -1, NY + 2) = AT(d, k, -1, NY + 1);
    AT(d, k, NX + 2, NY + 2) = AT(d, k, NX + 1, NY + 2);
  }
  wxo_calcprim(p, d, blk(p, m->primit, dad));
}

/* src/updatelpdown.f90:4-59 */
void wxo_updatelpdown(const wxo_params *p, wxo_mesh *m, int dad, int s1, int s2, int s3, int s4) {
  (void)p;
  m->leaf[dad] = 1;
  m->leaf[s1] = m->leaf[s2] = m->leaf[s3] = m->leaf[s4] = 0;
  m->nbleafs -= 3;
  int nb1 = LP(m, dad, 5);
  if (nb1 != -1 && !m->leaf[nb1]) { nb1 = LP(m, LP(m, nb1, 3), 6); SETLP(m, nb1, 6, dad); SETLP(m, LP(m, nb1, 8), 6, dad); }
  nb1 = LP(m, dad, 6);
  if (nb1 != -1 && !m->leaf[nb1]) { nb1 = LP(m, nb1, 3); SETLP(m, nb1, 5, dad); SETLP(m, LP(m, nb1, 8), 5, dad); }
  nb1 = LP(m, dad, 7);
  if (nb1 != -1 && !m->leaf[nb1]) { nb1 = LP(m, LP(m, nb1, 3), 8); SETLP(m, nb1, 8, dad); SETLP(m, LP(m, nb1, 6), 8, dad); }
  nb1 = LP(m, dad, 8);
  if (nb1 != -1 && !m->leaf[nb1]) { nb1 = LP(m, nb1, 3); SETLP(m, nb1, 7, dad); SETLP(m, LP(m, nb1, 6), 7, dad); }
}

/* src/markref.f90:22-27: criteria over the leaves; flags by block ID,
 * cleared first (MPIP with one rank: the flag exchange :51-78 is a no-op) */
void wxo_markref(const wxo_params *p, const wxo_mesh *m, int32_t *up, int32_t *dn) {
  for (int nb = 0; nb <= m->cap; ++nb) { up[nb] = 0; dn[nb] = 0; }
  for (int nb = 1; nb <= m->nbmax; ++nb)
    if (m->leaf[nb] && LP(m, nb, 1) >= 1) wxo_criteria(p, m, nb, up, dn);
}

/* src/critneighup.f90:4-62: a block flagged for refinement at level nl
 * flags its coarser neighbours, levels nlevs-1 down to 1 */
void wxo_critneighup(const wxo_params *p, const wxo_mesh *m, int32_t *up) {
  for (int nl = p->nlevs - 1; nl >= 1; --nl)
    for (int nb = 1; nb <= m->nbmax; ++nb)
      if (up[nb] && LP(m, nb, 1) == nl)
        for (int f = 5; f <= 8; ++f) {
          int n1 = LP(m, nb, f);
          if (n1 != -1 && LP(m, n1, 1) < LP(m, nb, 1)) up[n1] = 1;
        }
}

/* src/critneighdown.f90:4-56: no coarsening next to a non-leaf neighbour */
void wxo_critneighdown(const wxo_params *p, const wxo_mesh *m, int32_t *dn) {
  (void)p;
  for (int nb = 1; nb <= m->nbmax; ++nb)
    if (dn[nb])
      for (int f = 5; f <= 8; ++f) {
        int n1 = LP(m, nb, f);
        if (n1 != -1 && !m->leaf[n1]) dn[nb] = 0;
      }
}

/* src/admesh.f90:4-111, MPIP forms with one rank.  Returns WXO_ERR_BLOCKS
 * on the :44-47 stop. */
int wxo_admesh(const wxo_params *p, wxo_mesh *m) {
  int n = m->cap + 1;
  int32_t *up = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  int32_t *dn = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  wxo_markref(p, m, up, dn);
  int err = wxo_admesh_flags(p, m, up, dn);
  free(up);
  free(dn);
  return err;
}

/* admesh after markref (:22), from given flags (irefup/irefdown by block ID). */
int wxo_admesh_flags(const wxo_params *p, wxo_mesh *m, int32_t *up, int32_t *dn) {
  int err = 0;
  int nbmaxp = m->nbmax, nbmaxp1 = m->nbmax;
  wxo_critneighup(p, m, up);                                  /* :24 */
  /* refine by level :28-69 */
  for (int nl = 1; nl <= p->nlevs - 1 && !err; ++nl)
    for (int nb = 1; nb <= nbmaxp1; ++nb) {
      if (!(LP(m, nb, 1) == nl && up[nb])) continue;
      int s1, s2, s3, s4;
      if (LP(m, nb, 3) == 0) {
        s1 = nbmaxp + 1; s2 = s1 + 1; s3 = s1 + 2; s4 = s1 + 3;
        SETLP(m, nb, 3, s1);
        nbmaxp += 4;
        if (m->nbmax >= m->cap || nbmaxp > m->cap) { err = WXO_ERR_BLOCKS; break; }
      } else {
        s1 = LP(m, nb, 3); s2 = LP(m, s1, 6); s3 = LP(m, s1, 8); s4 = LP(m, s3, 6);
      }
      m->nbmax = nbmaxp;
      wxo_refine(p, m, nb, s1, s2, s3, s4);
      wxo_updatelpup(p, m, nb, s1, s2, s3, s4, up, dn);
    }
  if (!err) {
    wxo_critneighdown(p, m, dn);                              /* :72 */
    /* coarsen by level :75-105 */
    for (int nl = p->nlevs; nl >= 2; --nl)
      for (int nb = 1; nb <= nbmaxp1; ++nb) {
        if (!(dn[nb] && LP(m, nb, 4) == 1 && LP(m, nb, 1) == nl)) continue;
        int dad = LP(m, nb, 2), s1 = nb, s2 = LP(m, nb, 6), s3 = LP(m, nb, 8), s4 = LP(m, LP(m, nb, 6), 8);
        if (dn[s2] && dn[s3] && dn[s4]) {
          wxo_coarse(p, m, dad, s1, s2, s3, s4);
          wxo_updatelpdown(p, m, dad, s1, s2, s3, s4);
        }
      }
  }
  return err;
}

/* Generalised root layout (nbrootx x nbrooty roots at level 1, x fastest);
 * the 1-D form is src/basegrid.f90:26-97.  Data (initial conditions) are
 * written by the caller. */
void wxo_basegrid(const wxo_params *p, wxo_mesh *m) {
  int nrx = p->nbrootx, nry = p->nbrooty;
  for (int iy = 0; iy < nry; ++iy)
    for (int ix = 0; ix < nrx; ++ix) {
      int nb = 1 + ix + iy * nrx;
      SETLP(m, nb, 1, 1); SETLP(m, nb, 2, 0); SETLP(m, nb, 3, 0); SETLP(m, nb, 4, 1);
      SETLP(m, nb, 5, ix > 0 ? nb - 1 : -1);
      SETLP(m, nb, 6, ix < nrx - 1 ? nb + 1 : -1);
      SETLP(m, nb, 7, iy > 0 ? nb - nrx : -1);
      SETLP(m, nb, 8, iy < nry - 1 ? nb + nrx : -1);
      m->icoord[nb * 2] = ix * NX;
      m->icoord[nb * 2 + 1] = iy * NY;
      m->leaf[nb] = 1;
    }
  m->nbmax = nrx * nry;
  m->nbleafs = nrx * nry;
}

/* src/main.f90:69-106 time loop body: timestep, time += dt, tstep, admesh. */
int wxo_run(const wxo_params *p, wxo_mesh *m, int nsteps, double *time, double *dtlast,
            int do_admesh, double *zones) {
  int err = 0;
  for (int s = 0; s < nsteps && !err; ++s) {
    double dt = wxo_timestep(p, m);
    *time = *time + dt;
    if (zones) *zones += (double)m->nbleafs * NX * NY;
    err = wxo_tstep(p, m, *time, dt);
    if (!err && do_admesh) err = wxo_admesh(p, m);
    *dtlast = dt;
  }
  return err;
}

/* ---------------------------------------------------------------------
 * Load-balance plan (src/loadbalance.f90).  Arrays here are C-indexed by
 * block ID (slot 0 unused): lp[nb*8 + k-1] = lp(nb,k), icoord[nb*2 + d].
 * ------------------------------------------------------------------- */

/* src/loadbalance.f90:251-306 (hilbert2d): two key bits per level, most
 * significant level first; the pair (x bit, y bit) as a digit 2*x+y walks a
 * 4-state table (state_diagram(digit, 0|1, state) = next state | key digit).
 * Bits of x and y at or above `bits` do not enter the key. */
static const int8_t lb_next_state[4][4] = {{1, 0, 2, 0}, {0, 3, 1, 1}, {2, 2, 0, 3}, {3, 1, 3, 2}};
static const int8_t lb_key_digit[4][4]  = {{0, 1, 3, 2}, {0, 3, 1, 2}, {2, 1, 3, 0}, {2, 3, 1, 0}};

int32_t wxo_hilbert2d(int32_t x, int32_t y, int bits) {
  int state = 0;
  uint32_t key = 0;
  for (int b = bits - 1; b >= 0; --b) {
    int d = (((x >> b) & 1) << 1) | ((y >> b) & 1);
    key |= (uint32_t)lb_key_digit[state][d] << (2 * b);
    state = lb_next_state[state][d];
  }
  return (int32_t)key;
}

/* src/loadbalance.f90:320-400 (indexx): indx[0..n-1] (1-based values) such
 * that arr[indx[k]-1] ascends; median-of-three quicksort that hands ranges
 * of fewer than 6 elements (ir - l < 5) to straight insertion, larger part
 * pushed.  The order among equal keys is this algorithm's and part of the
 * result.  Work indices below are 1-based as in the source. */
void wxo_indexx(int n, const int32_t *arr, int32_t *indx) {
  int32_t *ix = indx - 1;                       /* ix[1..n] */
  const int32_t *a = arr - 1;                   /* a[1..n]  */
  int stack[64], sp = 0, l = 1, ir = n;
  for (int k = 1; k <= n; ++k) ix[k] = k;
#define LB_SWAP(p, q) do { int32_t t_ = ix[p]; ix[p] = ix[q]; ix[q] = t_; } while (0)
  for (;;) {
    if (ir - l < 5) {
      for (int j = l + 1; j <= ir; ++j) {
        int32_t v = ix[j], av = a[v];
        int i = j - 1;
        while (i >= l && a[ix[i]] > av) { ix[i + 1] = ix[i]; --i; }
        ix[i + 1] = v;
      }
      if (sp == 0) return;
      ir = stack[--sp];
      l = stack[--sp];
    } else {
      int mid = (l + ir) / 2;
      LB_SWAP(mid, l + 1);
      if (a[ix[l]] > a[ix[ir]]) LB_SWAP(l, ir);
      if (a[ix[l + 1]] > a[ix[ir]]) LB_SWAP(l + 1, ir);
      if (a[ix[l]] > a[ix[l + 1]]) LB_SWAP(l, l + 1);
      int i = l + 1, j = ir;
      int32_t pv = ix[l + 1], ap = a[pv];
      for (;;) {
        do ++i; while (a[ix[i]] < ap);
        do --j; while (a[ix[j]] > ap);
        if (j < i) break;
        LB_SWAP(i, j);
      }
      ix[l + 1] = ix[j];
      ix[j] = pv;
      if (ir - i + 1 >= j - l) { stack[sp++] = i; stack[sp++] = ir; ir = j - 1; }
      else                     { stack[sp++] = l; stack[sp++] = j - 1; l = i; }
    }
  }
#undef LB_SWAP
}

/* src/loadbalance.f90:24-33 (block coordinates at the finest level and their
 * keys over nlevs+1 bits) and :73-105 (new IDs on the master) for blocks
 * 1..nbmax held by one rank (idold(j) = j): blocks in key order fill rank
 * 0, 1, ... with IDs r*nblocks + 1, 2, ...; a rank closes after the block
 * that brings its leaf count to nbleafs/nranks (the last rank takes the
 * rest).  key[nb], idnew[nb] for nb = 1..nbmax.  Returns WXO_ERR_BLOCKS
 * where the source stops (an ID reaching the rank's capacity). */
int wxo_lb_plan(const int32_t *lp, const int32_t *leaf, const int32_t *icoord, int nbmax, int nlevs,
                int nblocks, int nranks, int32_t *key, int32_t *idnew) {
  if (nbmax < 1 || nranks < 1 || nlevs < 1 || nlevs > 14) return WXO_ERR_ARG;
  int nleaf = 0;
  for (int nb = 1; nb <= nbmax; ++nb) {
    int sh = nlevs - lp[nb * 8];
    int32_t xh = icoord[nb * 2] * (1 << sh) / NX, yh = icoord[nb * 2 + 1] * (1 << sh) / NY;
    key[nb] = wxo_hilbert2d(xh, yh, nlevs + 1);
    nleaf += leaf[nb] != 0;
  }
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)nbmax);
  wxo_indexx(nbmax, key + 1, order);
  int per = nleaf / nranks, r = 1, cnt = 1, base = 0, err = WXO_OK;
  for (int k = 1; k <= nbmax; ++k) {
    int nb = order[k - 1];
    idnew[nb] = (r - 1) * nblocks + k - base;
    if (leaf[nb]) ++cnt;
    if (cnt > per && r < nranks) { ++r; cnt = 1; base = k; }
    if (idnew[nb] / r == nblocks) { err = WXO_ERR_BLOCKS; break; }
  }
  free(order);
  return err;
}

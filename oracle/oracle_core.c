/*
 * oracle_core.c -- context, DMDA global/local semantics, metrics, equation of
 * state and boundary conditions of the CPU restatement (test infrastructure
 * only; see cles_oracle.h).
 */
#include "oracle_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* DMGlobalToLocal for a single-rank DMDA: owned values are copied and the
 * periodic ghosts (width 2) wrap to the owned node with index (g+m)%m. */
void orc_g2l(const orc_ctx *c, const double *g, double *l, int nc) {
  #pragma omp parallel for schedule(static)
  for (int k = -c->gz; k < c->mz + c->gz; k++) {
    int kk = (k + c->mz) % c->mz;
    for (int j = -c->gy; j < c->my + c->gy; j++) {
      int jj = (j + c->my) % c->my;
      for (int i = -c->gx; i < c->mx + c->gx; i++) {
        int ii = (i + c->mx) % c->mx;
        const double *s = g + GI(c, kk, jj, ii) * nc;
        double *d = l + LI(c, k, j, i) * nc;
        for (int q = 0; q < nc; q++) d[q] = s[q];
      }
    }
  }
}

/* DMLocalToLocal(da, l, INSERT_VALUES, l): refreshes the ghosts of a local
 * array from its own owned region. */
void orc_l2l(const orc_ctx *c, double *l, int nc) {
  if (!(c->gx || c->gy || c->gz)) return;
  #pragma omp parallel for schedule(static)
  for (int k = -c->gz; k < c->mz + c->gz; k++) {
    int kk = (k + c->mz) % c->mz;
    for (int j = -c->gy; j < c->my + c->gy; j++) {
      int jj = (j + c->my) % c->my;
      for (int i = -c->gx; i < c->mx + c->gx; i++) {
        int ii = (i + c->mx) % c->mx;
        if (ii == i && jj == j && kk == k) continue;
        const double *s = l + LI(c, kk, jj, ii) * nc;
        double *d = l + LI(c, k, j, i) * nc;
        for (int q = 0; q < nc; q++) d[q] = s[q];
      }
    }
  }
}

/* FormMetrics, metrics.c:3-247 plus the G2L of metrics.c:695-707.
 * Metric triples are the cross products of half central coordinate
 * differences; aj = 1/det; boundary layers copy the adjacent layer in the
 * order i, j, k (metrics.c:138-246). */
static void form_metrics(orc_ctx *c) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  orc_v3 *csi = xcalloc(c->nG, sizeof(orc_v3));
  orc_v3 *eta = xcalloc(c->nG, sizeof(orc_v3));
  orc_v3 *zet = xcalloc(c->nG, sizeof(orc_v3));
  double *aj = xcalloc(c->nG, sizeof(double));
  const orc_v3 *coor = c->coor;
#define CO(k, j, i) coor[GI(c, k, j, i)]
  /* csi = de x dz (metrics.c:52-70) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 0; i < mx; i++) {
        double dxde = 0.5 * (CO(k, j + 1, i).x - CO(k, j - 1, i).x);
        double dyde = 0.5 * (CO(k, j + 1, i).y - CO(k, j - 1, i).y);
        double dzde = 0.5 * (CO(k, j + 1, i).z - CO(k, j - 1, i).z);
        double dxdz = 0.5 * (CO(k + 1, j, i).x - CO(k - 1, j, i).x);
        double dydz = 0.5 * (CO(k + 1, j, i).y - CO(k - 1, j, i).y);
        double dzdz = 0.5 * (CO(k + 1, j, i).z - CO(k - 1, j, i).z);
        orc_v3 *m = &csi[GI(c, k, j, i)];
        m->x = dyde * dzdz - dzde * dydz;
        m->y = -dxde * dzdz + dzde * dxdz;
        m->z = dxde * dydz - dyde * dxdz;
      }
  /* eta = dz x dc (metrics.c:73-92) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 0; j < my; j++)
      for (int i = 1; i < mx - 1; i++) {
        double dxdc = 0.5 * (CO(k, j, i + 1).x - CO(k, j, i - 1).x);
        double dydc = 0.5 * (CO(k, j, i + 1).y - CO(k, j, i - 1).y);
        double dzdc = 0.5 * (CO(k, j, i + 1).z - CO(k, j, i - 1).z);
        double dxdz = 0.5 * (CO(k + 1, j, i).x - CO(k - 1, j, i).x);
        double dydz = 0.5 * (CO(k + 1, j, i).y - CO(k - 1, j, i).y);
        double dzdz = 0.5 * (CO(k + 1, j, i).z - CO(k - 1, j, i).z);
        orc_v3 *m = &eta[GI(c, k, j, i)];
        m->x = dydz * dzdc - dzdz * dydc;
        m->y = -dxdz * dzdc + dzdz * dxdc;
        m->z = dxdz * dydc - dydz * dxdc;
      }
  /* zet = dc x de (metrics.c:95-113) */
  for (int k = 0; k < mz; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        double dxdc = 0.5 * (CO(k, j, i + 1).x - CO(k, j, i - 1).x);
        double dydc = 0.5 * (CO(k, j, i + 1).y - CO(k, j, i - 1).y);
        double dzdc = 0.5 * (CO(k, j, i + 1).z - CO(k, j, i - 1).z);
        double dxde = 0.5 * (CO(k, j + 1, i).x - CO(k, j - 1, i).x);
        double dyde = 0.5 * (CO(k, j + 1, i).y - CO(k, j - 1, i).y);
        double dzde = 0.5 * (CO(k, j + 1, i).z - CO(k, j - 1, i).z);
        orc_v3 *m = &zet[GI(c, k, j, i)];
        m->x = dydc * dzde - dzdc * dyde;
        m->y = -dxdc * dzde + dzdc * dxde;
        m->z = dxdc * dyde - dydc * dxde;
      }
  /* aj = 1/J (metrics.c:116-136) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        double dxdc = 0.5 * (CO(k, j, i + 1).x - CO(k, j, i - 1).x);
        double dydc = 0.5 * (CO(k, j, i + 1).y - CO(k, j, i - 1).y);
        double dzdc = 0.5 * (CO(k, j, i + 1).z - CO(k, j, i - 1).z);
        double dxde = 0.5 * (CO(k, j + 1, i).x - CO(k, j - 1, i).x);
        double dyde = 0.5 * (CO(k, j + 1, i).y - CO(k, j - 1, i).y);
        double dzde = 0.5 * (CO(k, j + 1, i).z - CO(k, j - 1, i).z);
        double dxdz = 0.5 * (CO(k + 1, j, i).x - CO(k - 1, j, i).x);
        double dydz = 0.5 * (CO(k + 1, j, i).y - CO(k - 1, j, i).y);
        double dzdz = 0.5 * (CO(k + 1, j, i).z - CO(k - 1, j, i).z);
        double a = dxdc * (dyde * dzdz - dzde * dydz) -
                   dydc * (dxde * dzdz - dzde * dxdz) +
                   dzdc * (dxde * dydz - dyde * dxdz);
        aj[GI(c, k, j, i)] = 1. / a;
      }
#undef CO
  /* boundary copies, metrics.c:138-246 */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++) {
      size_t n = GI(c, k, j, 0), nn = GI(c, k, j, 1);
      eta[n] = eta[nn]; zet[n] = zet[nn]; aj[n] = aj[nn];
    }
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++) {
      size_t n = GI(c, k, j, mx - 1), nn = GI(c, k, j, mx - 2);
      eta[n] = eta[nn]; zet[n] = zet[nn]; aj[n] = aj[nn];
    }
  for (int k = 0; k < mz; k++)
    for (int i = 0; i < mx; i++) {
      size_t n = GI(c, k, 0, i), nn = GI(c, k, 1, i);
      csi[n] = csi[nn]; zet[n] = zet[nn]; aj[n] = aj[nn];
    }
  for (int k = 0; k < mz; k++)
    for (int i = 0; i < mx; i++) {
      size_t n = GI(c, k, my - 1, i), nn = GI(c, k, my - 2, i);
      csi[n] = csi[nn]; zet[n] = zet[nn]; aj[n] = aj[nn];
    }
  for (int j = 0; j < my; j++)
    for (int i = 0; i < mx; i++) {
      size_t n = GI(c, 0, j, i), nn = GI(c, 1, j, i);
      eta[n] = eta[nn]; csi[n] = csi[nn]; aj[n] = aj[nn];
    }
  for (int j = 0; j < my; j++)
    for (int i = 0; i < mx; i++) {
      size_t n = GI(c, mz - 1, j, i), nn = GI(c, mz - 2, j, i);
      eta[n] = eta[nn]; csi[n] = csi[nn]; aj[n] = aj[nn];
    }
  orc_g2l(c, (double *)csi, (double *)c->lcsi, 3);
  orc_g2l(c, (double *)eta, (double *)c->leta, 3);
  orc_g2l(c, (double *)zet, (double *)c->lzet, 3);
  orc_g2l(c, aj, c->laj, 1);
  free(csi); free(eta); free(zet); free(aj);
}

orc_ctx *orc_create(const orc_params *p, const double *coor_global) {
  orc_ctx *c = xcalloc(1, sizeof(orc_ctx));
  c->p = *p;
  c->mx = p->mx; c->my = p->my; c->mz = p->mz;
  /* x periodic iff bctype[0]==7, y iff bctype[2]==7, z iff bctype[4]==7
   * (init.c:158-197, every branch) */
  c->gx = p->bctype[0] == 7 ? 2 : 0;
  c->gy = p->bctype[2] == 7 ? 2 : 0;
  c->gz = p->bctype[4] == 7 ? 2 : 0;
  c->LX = c->mx + 2 * c->gx; c->LY = c->my + 2 * c->gy; c->LZ = c->mz + 2 * c->gz;
  c->nG = (size_t)c->mx * c->my * c->mz;
  c->nL = (size_t)c->LX * c->LY * c->LZ;
  c->coor = xcalloc(c->nG, sizeof(orc_v3));
  memcpy(c->coor, coor_global, c->nG * sizeof(orc_v3));
  c->lcsi = xcalloc(c->nL, sizeof(orc_v3));
  c->leta = xcalloc(c->nL, sizeof(orc_v3));
  c->lzet = xcalloc(c->nL, sizeof(orc_v3));
  c->laj = xcalloc(c->nL, sizeof(double));
  c->nvert = xcalloc(c->nG, sizeof(double));
  c->lnvert = xcalloc(c->nL, sizeof(double));
  c->Q = xcalloc(c->nG, sizeof(orc_q));
  c->Qold = xcalloc(c->nG, sizeof(orc_q));
  c->Rhs = xcalloc(c->nG, sizeof(orc_q));
  c->lQ = xcalloc(c->nL, sizeof(orc_q));
  c->P = xcalloc(c->nG, sizeof(double));
  c->lP = xcalloc(c->nL, sizeof(double));
  c->Sigma = xcalloc(c->nG, sizeof(orc_v3));
  c->lSigma = xcalloc(c->nL, sizeof(orc_v3));
  c->CS = xcalloc(c->nG, sizeof(double));
  c->lCS = xcalloc(c->nL, sizeof(double));
  c->Nu_t = xcalloc(c->nG, sizeof(double));
  c->lNu_t = xcalloc(c->nL, sizeof(double));
  c->lUcat = xcalloc(c->nL, sizeof(orc_v3));
  c->fp1 = xcalloc(c->nL, sizeof(orc_q));
  c->fp2 = xcalloc(c->nL, sizeof(orc_q));
  c->fp3 = xcalloc(c->nL, sizeof(orc_q));
  c->conv = xcalloc(c->nG, sizeof(orc_q));
  c->visc = xcalloc(c->nG, sizeof(orc_q));
  c->al1 = xcalloc(c->nL, sizeof(orc_v3));
  c->al2 = xcalloc(c->nL, sizeof(orc_v3));
  c->al3 = xcalloc(c->nL, sizeof(orc_v3));
  c->ti = p->tistart;
  form_metrics(c);
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  free(c->coor); free(c->lcsi); free(c->leta); free(c->lzet); free(c->laj);
  free(c->nvert); free(c->lnvert); free(c->Q); free(c->Qold); free(c->Rhs);
  free(c->lQ); free(c->P); free(c->lP); free(c->Sigma); free(c->lSigma);
  free(c->CS); free(c->lCS); free(c->Nu_t); free(c->lNu_t); free(c->lUcat);
  free(c->fp1); free(c->fp2); free(c->fp3); free(c->conv); free(c->visc);
  free(c->al1); free(c->al2); free(c->al3);
  free(c->x_bp); free(c->y_bp); free(c->z_bp); free(c->nf_x); free(c->nf_y); free(c->nf_z);
  free(c->radvec); free(c->nv1); free(c->nv2); free(c->nv3); free(c->qvec); free(c->u_bp);
  free(c->uold_bp); free(c->ibl);
  free(c->Ucat); free(c->Q_sum); free(c->P_sum); free(c->U2_sum); free(c->Ucross_sum);
  free(c);
}

void orc_set_nvert(orc_ctx *c, const double *nvert) {
  memcpy(c->nvert, nvert, c->nG * sizeof(double));
  orc_g2l(c, c->nvert, c->lnvert, 1);
}

void orc_get_nvert(const orc_ctx *c, double *nvert) { memcpy(nvert, c->nvert, c->nG * sizeof(double)); }

void orc_set_time(orc_ctx *c, int ti) { c->ti = ti; }
int orc_get_time(const orc_ctx *c) { return c->ti; }

/* EqOfState, rhs.c:4820-4870: p from lQ on all owned nodes, then G2L. */
void orc_eq_of_state(orc_ctx *c) {
  const double gamma = c->p.gamma;
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < c->mz; k++)
    for (int j = 0; j < c->my; j++)
      for (int i = 0; i < c->mx; i++) {
        const orc_q *q = &c->lQ[LI(c, k, j, i)];
        c->P[GI(c, k, j, i)] =
            (gamma - 1) * (q->rhoE - 0.5 * (q->rhoU * q->rhoU + q->rhoV * q->rhoV +
                                            q->rhoW * q->rhoW) / q->rho);
      }
  orc_g2l(c, c->P, c->lP, 1);
}

/* Tail of SetInitialCondition, rhs.c:424-433. */
void orc_set_state(orc_ctx *c, const double *Q) {
  memcpy(c->Q, Q, c->nG * sizeof(orc_q));
  orc_g2l(c, (double *)c->Q, (double *)c->lQ, 5);
  memcpy(c->Qold, c->Q, c->nG * sizeof(orc_q));
  orc_eq_of_state(c);
}

void orc_get_state(const orc_ctx *c, double *Q, double *P) {
  if (Q) memcpy(Q, c->Q, c->nG * sizeof(orc_q));
  if (P) memcpy(P, c->P, c->nG * sizeof(double));
}

void orc_get_metrics(const orc_ctx *c, double *csi, double *eta, double *zet, double *aj) {
  for (int k = 0; k < c->mz; k++)
    for (int j = 0; j < c->my; j++)
      for (int i = 0; i < c->mx; i++) {
        size_t g = GI(c, k, j, i), l = LI(c, k, j, i);
        if (csi) memcpy(csi + 3 * g, &c->lcsi[l], 3 * sizeof(double));
        if (eta) memcpy(eta + 3 * g, &c->leta[l], 3 * sizeof(double));
        if (zet) memcpy(zet + 3 * g, &c->lzet[l], 3 * sizeof(double));
        if (aj) aj[g] = c->laj[l];
      }
}

void orc_get_les(const orc_ctx *c, double *cs, double *nu_t, double *sigma) {
  if (cs) memcpy(cs, c->CS, c->nG * sizeof(double));
  if (nu_t) memcpy(nu_t, c->Nu_t, c->nG * sizeof(double));
  if (sigma) memcpy(sigma, c->Sigma, c->nG * sizeof(orc_v3));
}

/* VecNorm(NORM_INFINITY) of Rhs and Q as RungeKutta takes them after the
 * last formRHS (solvers.c:251-252): the Rhs of that stage and its input Q. */
void orc_norms(const orc_ctx *c, double out[2]) {
  out[0] = c->nrm[0];
  out[1] = c->nrm[1];
}

/* What RungeKutta runs after FormBCS: G2L of Q, then EqOfState
 * (solvers.c:276-279). */
void orc_bcs_tail(orc_ctx *c) {
  orc_g2l(c, (double *)c->Q, (double *)c->lQ, 5);
  orc_eq_of_state(c);
}

/* ------------------------------------------------------------------------
 * FormBCS, rhs.c:477-1316.  q = global Q (already RK-updated), qold = Qold,
 * lq = stage-input local lQ, lp = local lP (modified in place, as in the
 * reference), p = global P.  The channel/wavy mass correction
 * (rhs.c:539-598) is not restated: it requires -channel/-wavy.
 * ------------------------------------------------------------------------ */
void orc_form_bcs(orc_ctx *c, double alfa) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  const orc_params *P = &c->p;
  const int *bc = P->bctype;
  orc_q *q = c->Q;
  const orc_q *qold = c->Qold, *lq = c->lQ;
  double *lp = c->lP, *p = c->P;
  const double *nvert = c->lnvert, *aj = c->laj;
  const orc_v3 *eta = c->leta, *zet = c->lzet;
  const double gam = P->gamma, Ma = P->Ma, Twall = P->Twall, dt = P->dt;
  const int isothermal = P->isothermal, twoD = P->twoD;
  const double pi = 3.14159265;
  int lys = 1, lye = my - 1, lzs = 1, lze = mz - 1;
#define Qg(k, j, i) q[GI(c, k, j, i)]
#define QO(k, j, i) qold[GI(c, k, j, i)]
#define LQ(k, j, i) lq[LI(c, k, j, i)]
#define LP(k, j, i) lp[LI(c, k, j, i)]
#define AJ(k, j, i) aj[LI(c, k, j, i)]
#define NV(k, j, i) nvert[LI(c, k, j, i)]
#define ETA(k, j, i) eta[LI(c, k, j, i)]
#define ZET(k, j, i) zet[LI(c, k, j, i)]
  /* i BC, rhs.c:600-637 */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++) {
      int i = 0;
      if (bc[0] == 3) {
        Qg(k, j, i).rho = Qg(k, j, i + 1).rho;
        Qg(k, j, i).rhoU = 0.;
        Qg(k, j, i).rhoV = Qg(k, j, i + 1).rhoV;
        Qg(k, j, i).rhoW = Qg(k, j, i + 1).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j, i + 1).rhoE;
      } else if (bc[0] == 51) {
        Qg(k, j, i).rho = 0.445; Qg(k, j, i).rhoU = 0.311; Qg(k, j, i).rhoV = 0.;
        Qg(k, j, i).rhoW = 0.; Qg(k, j, i).rhoE = 8.928;
      }
      i = mx - 1;
      if (bc[1] == 3) {
        Qg(k, j, i).rho = Qg(k, j, i - 1).rho;
        Qg(k, j, i).rhoU = 0.;
        Qg(k, j, i).rhoV = Qg(k, j, i - 1).rhoV;
        Qg(k, j, i).rhoW = Qg(k, j, i - 1).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j, i - 1).rhoE;
      } else if (bc[0] == 51) { /* tests bctype[0], rhs.c:627 */
        Qg(k, j, i).rho = 0.5; Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
        Qg(k, j, i).rhoW = 0.; Qg(k, j, i).rhoE = 1.4275;
      }
    }
  /* j BC, rhs.c:640-851 */
  for (int k = 0; k < mz; k++)
    for (int i = 0; i < mx; i++) {
      int j = 0;
      if (bc[2] == 3) {
        Qg(k, j, i).rho = Qg(k, j + 1, i).rho;
        Qg(k, j, i).rhoU = Qg(k, j + 1, i).rhoU;
        Qg(k, j, i).rhoV = 0.;
        Qg(k, j, i).rhoW = Qg(k, j + 1, i).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j + 1, i).rhoE;
      } else if (bc[2] == 10) {
        Qg(k, j, i).rho = Qg(k, j + 1, i).rho;
        Qg(k, j, i).rhoU = 0.;
        Qg(k, j, i).rhoV = Qg(k, j + 1, i).rhoV;
        Qg(k, j, i).rhoW = Qg(k, j + 1, i).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j + 1, i).rhoE;
      } else if (bc[2] == 1) {
        Qg(k, j, i).rhoU = 0.0; Qg(k, j, i).rhoV = 0.0; Qg(k, j, i).rhoW = 0.0;
        if (isothermal == 2 && alfa != 0 && NV(k, j, i) < 1.5) {
          /* Poinsot-Lele characteristic isothermal wall, rhs.c:692-730 */
          double a_sound = sqrt(Twall / (Ma * Ma));
          double beta = sqrt(ETA(k, j, i).y * ETA(k, j, i).y + ETA(k, j, i).x * ETA(k, j, i).x +
                             ETA(k, j, i).z * ETA(k, j, i).z);
          double Lambda0 = -a_sound * beta;
          double L0;
          if (k == 0 || k == mz - 1) {
            L0 = 0.;
          } else {
            L0 = Lambda0 * AJ(k, j, i) *
                 ((-0.5 * LP(k, j + 2, i) + 2 * LP(k, j + 1, i) - 1.5 * LP(k, j, i)) -
                  LQ(k, j, i).rho * a_sound / beta *
                      (ETA(k, j, i).y * (-0.5 * LQ(k, j + 2, i).rhoV / LQ(k, j + 2, i).rho +
                                         2. * LQ(k, j + 1, i).rhoV / LQ(k, j + 1, i).rho -
                                         1.5 * LQ(k, j, i).rhoV / LQ(k, j, i).rho) +
                       ETA(k, j, i).z * (-0.5 * LQ(k, j + 2, i).rhoW / LQ(k, j + 2, i).rho +
                                         2. * LQ(k, j + 1, i).rhoW / LQ(k, j + 1, i).rho -
                                         1.5 * LQ(k, j, i).rhoW / LQ(k, j, i).rho)));
            if (twoD != 1)
              L0 -= Lambda0 * AJ(k, j, i) * LQ(k, j, i).rho * a_sound / beta * ETA(k, j, i).x *
                    (-0.5 * LQ(k, j + 2, i).rhoU / LQ(k, j + 2, i).rho +
                     2. * LQ(k, j + 1, i).rhoU / LQ(k, j + 1, i).rho -
                     1.5 * LQ(k, j, i).rhoU / LQ(k, j, i).rho);
            Qg(k, j, i).rho = QO(k, j, i).rho - dt * alfa * L0 / a_sound / a_sound;
            LP(k, j, i) = Qg(k, j, i).rho * Twall / (gam * Ma * Ma);
          }
        } else {
          LP(k, j, i) = LP(k, j + 1, i);
          Qg(k, j, i).rho = LP(k, j, i) * (gam * Ma * Ma);
          Qg(k, j, i).rhoE = LP(k, j, i) / (gam - 1);
        }
      }
      j = my - 1;
      if (bc[3] == 3) {
        Qg(k, j, i).rho = Qg(k, j - 1, i).rho;
        Qg(k, j, i).rhoU = Qg(k, j - 1, i).rhoU;
        Qg(k, j, i).rhoV = 0.;
        Qg(k, j, i).rhoW = Qg(k, j - 1, i).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j - 1, i).rhoE;
      } else if (bc[3] == 10) {
        Qg(k, j, i).rho = Qg(k, j - 1, i).rho;
        Qg(k, j, i).rhoU = 0.;
        Qg(k, j, i).rhoV = Qg(k, j - 1, i).rhoV;
        Qg(k, j, i).rhoW = Qg(k, j - 1, i).rhoW;
        Qg(k, j, i).rhoE = Qg(k, j - 1, i).rhoE;
      } else if (bc[3] == 1) {
        Qg(k, j, i).rhoU = 0.0; Qg(k, j, i).rhoV = 0.0; Qg(k, j, i).rhoW = 0.0;
        if (isothermal == 2) {
          /* rhs.c:825-838 */
          double a_sound = sqrt(gam * LP(k, j - 1, i) / LQ(k, j - 1, i).rho);
          double Lambda0 = LQ(k, j - 1, i).rhoV / LQ(k, j - 1, i).rho - a_sound;
          double L0;
          if (k == 0 || k == mz - 1 || Lambda0 > 0.)
            L0 = 0.;
          else
            L0 = Lambda0 * AJ(k, j - 1, i) * ETA(k, j - 1, i).y *
                 ((LP(k, j - 1, i) - LP(k, j - 2, i)) -
                  LQ(k, j - 1, i).rho * a_sound *
                      (LQ(k, j, i).rhoV / LQ(k, j, i).rho - LQ(k, j - 1, i).rhoV / LQ(k, j - 1, i).rho));
          Qg(k, j, i).rho = QO(k, j, i).rho - dt * alfa * L0 / a_sound / a_sound;
          LP(k, j, i) = Qg(k, j, i).rho * Twall / (gam * Ma * Ma);
        } else if (isothermal == 1) {
          Qg(k, j, i).rho = LP(k, j, i) * (gam * Ma * Ma);
          LP(k, j, i) = LP(k, j - 1, i);
        } else {
          Qg(k, j, i).rho = Qg(k, j - 1, i).rho;
          LP(k, j, i) = LP(k, j - 1, i);
        }
        Qg(k, j, i).rhoE = LP(k, j, i) / (gam - 1.0);
      }
    }
  /* k BC, rhs.c:855-1088 */
  {
    const double u = Ma, rho = gam; /* rhs.c:855 */
    int ls = bc[2] == 1 ? lys : 0, le = bc[3] == 1 ? lye : my;
    for (int j = ls; j < le; j++)
      for (int i = 0; i < mx; i++) {
        int k = 0;
        if (bc[4] == 3) {
          Qg(k, j, i).rho = Qg(k + 1, j, i).rho;
          Qg(k, j, i).rhoU = Qg(k + 1, j, i).rhoU;
          Qg(k, j, i).rhoV = Qg(k + 1, j, i).rhoV;
          Qg(k, j, i).rhoW = 0.;
          Qg(k, j, i).rhoE = Qg(k + 1, j, i).rhoE;
        } else if (bc[4] == 52) {
          const orc_v3 *co = &c->coor[GI(c, k, j, i)];
          double AA = cos(pi / 2 * (co->y - 1.)) * cos(pi / 2 * (co->y - 1.));
          double a_sound = sqrt(gam * LP(k + 1, j, i) / LQ(k + 1, j, i).rho);
          double Lambda0 = LQ(k + 1, j, i).rhoW / LQ(k + 1, j, i).rho - a_sound;
          double L0 = Lambda0 * AJ(k + 1, j, i) * ZET(k + 1, j, i).z *
                      ((LP(k + 2, j, i) - LP(k + 1, j, i)) -
                       LQ(k + 1, j, i).rho * a_sound * (LQ(k + 1, j, i).rhoW / LQ(k + 1, j, i).rho - AA));
          double L1 = (gam - 1.) * L0, L4 = L0;
          double d0 = 1. / (a_sound * a_sound) * (L1 + 0.5 * (L4 + L0));
          Qg(k, j, i).rho = QO(k, j, i).rho - dt * alfa * d0;
          LP(k, j, i) = Qg(k, j, i).rho / (gam * Ma * Ma);
          Qg(k, j, i).rhoW = Qg(k, j, i).rho * AA;
          Qg(k, j, i).rhoU = 0.;
          Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoE = 0.5 * Qg(k, j, i).rhoW * Qg(k, j, i).rhoW / Qg(k, j, i).rho + LP(k, j, i) / (gam - 1);
        } else if (bc[4] == 55) {
          Qg(k, j, i).rho = rho;
          Qg(k, j, i).rhoW = rho * u;
          Qg(k, j, i).rhoE = 0.5 * rho * u * u + 1. / (gam - 1.);
        } else if (bc[4] == 44) {
          Qg(k, j, i).rho = Qg(k + 1, j, i).rho;
          Qg(k, j, i).rhoU = 0.;
          Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoW = Qg(k + 1, j, i).rhoW;
          Qg(k, j, i).rhoE = Qg(k + 1, j, i).rhoE;
        } else if (bc[4] == 45) {
          Qg(k, j, i).rho = 1.;
          Qg(k, j, i).rhoW = 1.;
          Qg(k, j, i).rhoE = .5 + 1. / (Ma * Ma * gam * (gam - 1.));
        } else if (bc[4] == 51) {
          Qg(k, j, i).rho = 0.445; Qg(k, j, i).rhoW = 0.311; Qg(k, j, i).rhoE = 8.928;
        } else if (bc[4] == 10) {
          Qg(k, j, i).rho = 1.; Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoW = 1.; Qg(k, j, i).rhoE = 1. / 0.4 + 0.5;
        } else if (bc[4] == 47) {
          /* power-law inflow with travelling-wave disturbances, rhs.c:904-945;
           * rhoU is left as the stage update made it below y = 8 */
          const orc_v3 *co = &c->coor[GI(c, k, j, i)];
          const double pw = 3.141592;
          const double c1 = 0.03, c2 = -0.0032, c3 = 0.02, c4 = -0.01, yp1 = 1.0, yp2 = 1.5, yp3 = 2.0;
          const double w0 = 0.1, w1 = 0.25, w2 = 0.125, w3 = 0.0625;
          const double phi1 = 0.0, phi2 = 0.1, phi3 = 0.15;
          const double b0 = 0.5 * pw, b1 = 0.75 * pw, b2 = 0.5 * pw, b3 = 0.25 * pw;
          const double yy = co->y, yplus = yy * 0.065 * P->ren;
          if (yy < 8.0) {
            Qg(k, j, i).rhoW = pow(yy / 8.0, 0.1428);
          } else {
            Qg(k, j, i).rhoW = 1.0;
            Qg(k, j, i).rhoU = 0.0;
          }
          if (yy < 8.0) {
            Qg(k, j, i).rhoW += c1 * yplus * exp(-yplus / 12.5) * sin(w0 * c->ti * dt) * cos(b0 * co->x);
            Qg(k, j, i).rhoV = c2 * yplus * yplus * exp(-yplus * yplus / 12.5 / 12.5) * sin(w0 * c->ti * dt) *
                               cos(b0 * co->x);
          } else {
            Qg(k, j, i).rhoW += c3 * yy / yp1 * exp(-yy / yp1) * sin(w1 * c->ti * dt) * cos(b1 * co->x + phi1) +
                                c3 * yy / yp2 * exp(-yy / yp2) * sin(w2 * c->ti * dt) * cos(b2 * co->x + phi2) +
                                c3 * yy / yp3 * exp(-yy / yp3) * sin(w3 * c->ti * dt) * cos(b3 * co->x + phi3);
            Qg(k, j, i).rhoV =
                c4 * yy * yy / yp1 / yp1 * exp(-yy * yy / yp1 / yp1) * sin(w1 * c->ti * dt) * cos(b1 * co->x + phi1) +
                c4 * yy * yy / yp2 / yp2 * exp(-yy * yy / yp2 / yp2) * sin(w2 * c->ti * dt) * cos(b2 * co->x + phi2) +
                c4 * yy * yy / yp3 / yp3 * exp(-yy * yy / yp3 / yp3) * sin(w3 * c->ti * dt) * cos(b3 * co->x + phi3);
          }
          Qg(k, j, i).rho = 1.0;
          Qg(k, j, i).rhoE = 0.5 * Qg(k, j, i).rhoW * Qg(k, j, i).rhoW + 1. / (Ma * Ma * gam * (gam - 1.));
        }
        k = mz - 1;
        if (bc[5] == 3) {
          Qg(k, j, i).rho = Qg(k - 1, j, i).rho;
          Qg(k, j, i).rhoU = Qg(k - 1, j, i).rhoU;
          Qg(k, j, i).rhoV = Qg(k - 1, j, i).rhoV;
          Qg(k, j, i).rhoW = 0.;
          Qg(k, j, i).rhoE = Qg(k - 1, j, i).rhoE;
        } else if (bc[5] == 44) {
          Qg(k, j, i).rho = Qg(k - 1, j, i).rho;
          Qg(k, j, i).rhoU = 0.;
          Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoW = Qg(k - 1, j, i).rhoW;
          Qg(k, j, i).rhoE = Qg(k - 1, j, i).rhoE;
        } else if (bc[5] == 45) {
          Qg(k, j, i) = Qg(k - 1, j, i);
        } else if (bc[5] == 49) {
          /* characteristic Riemann-invariant outflow, rhs.c:970-1020 */
          double Vn = Qg(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho;
          double a_sound = sqrt(gam * LP(k - 1, j, i) / Qg(k - 1, j, i).rho);
          double lambda1 = Vn + a_sound, lambda2 = Vn - a_sound;
          double Rplus, Rminus, s;
          if (Vn < 0.) {
            if (lambda1 < 0) {
              Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
              Rplus = Vn + 2 / Ma / (gam - 1.);
              Rminus = Vn - 2 / Ma / (gam - 1.);
              s = 1. / (Ma * Ma * gam);
            } else {
              Rplus = Vn + 2 * a_sound / (gam - 1.);
              Rminus = Vn - 2 / Ma / (gam - 1.);
              Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
              s = 1. / (Ma * Ma * gam);
            }
          } else {
            if (lambda2 >= 0) {
              Rplus = Vn + 2 * a_sound / (gam - 1.);
              Rminus = Vn - 2 * a_sound / (gam - 1.);
              Qg(k, j, i).rhoU = Qg(k - 1, j, i).rhoU;
              Qg(k, j, i).rhoV = Qg(k - 1, j, i).rho;
              s = LP(k - 1, j, i) / pow(Qg(k - 1, j, i).rho, gam);
            } else {
              Rminus = Vn - 2 / Ma / (gam - 1.);
              Rplus = Vn + 2 * a_sound / (gam - 1.);
              Qg(k, j, i).rhoU = Qg(k - 1, j, i).rhoU;
              Qg(k, j, i).rhoV = Qg(k - 1, j, i).rhoV;
              s = LP(k - 1, j, i) / pow(Qg(k - 1, j, i).rho, gam);
            }
          }
          Vn = 0.5 * (Rplus + Rminus);
          a_sound = 0.25 * (gam - 1.) * (Rplus - Rminus);
          double pOverrho = a_sound * a_sound / (gam);
          Qg(k, j, i).rho = pow(pOverrho / s, 1. / (gam - 2.));
          Qg(k, j, i).rhoW = Qg(k, j, i).rho * Vn;
          double pp = pOverrho * Qg(k, j, i).rho;
          Qg(k, j, i).rhoE = pp / (gam - 1.) + 0.5 * (Qg(k, j, i).rhoU * Qg(k, j, i).rhoU +
                                                  Qg(k, j, i).rhoV * Qg(k, j, i).rhoV +
                                                  Qg(k, j, i).rhoW * Qg(k, j, i).rhoW) * Qg(k, j, i).rho;
        } else if (bc[5] == 51) {
          Qg(k, j, i).rho = 0.5; Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoW = 0.; Qg(k, j, i).rhoE = 1.4275;
        } else if (bc[5] == 10) {
          Qg(k, j, i).rho = 1.; Qg(k, j, i).rhoU = 0.; Qg(k, j, i).rhoV = 0.;
          Qg(k, j, i).rhoW = 1.; Qg(k, j, i).rhoE = 1 / 0.4 + 0.5;
        } else if (bc[5] == 8) {
          /* subsonic non-reflecting outflow, rhs.c:1042-1085 */
          double a_sound = sqrt(gam * LP(k - 1, j, i) / LQ(k - 1, j, i).rho);
          double AA = Qg(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho;
          double Lam1 = AA, Lam2 = AA, Lam3 = AA, Lam4 = AA + a_sound;
          double L0 = 0, L1, L2, L3, L4;
          L1 = Lam1 * AJ(k - 1, j, i) * ZET(k - 1, j, i).z *
               (a_sound * a_sound * (LQ(k - 1, j, i).rho - LQ(k - 2, j, i).rho) -
                (LP(k, j, i) - LP(k - 1, j, i)));
          /* -TwoD (the global of main.c:36, 749), not -twoD: a different
           * option, not on this path, so it keeps its default 0 here */
          const int TwoD = 0;
          if (TwoD != 1)
            L2 = Lam2 * AJ(k - 1, j, i) * ZET(k - 1, j, i).z *
                 (LQ(k - 1, j, i).rhoU / Qg(k - 1, j, i).rho - LQ(k - 2, j, i).rhoU / Qg(k - 1, j, i).rho);
          else
            L2 = 0.;
          if (TwoD != 2)
            L3 = Lam3 * AJ(k - 1, j, i) * ZET(k - 1, j, i).z *
                 (LQ(k - 1, j, i).rhoV / Qg(k - 1, j, i).rho - LQ(k - 2, j, i).rhoV / Qg(k - 2, j, i).rho);
          else
            L3 = 0.;
          L4 = Lam4 * AJ(k - 1, j, i) * ZET(k - 1, j, i).z *
               ((LP(k - 1, j, i) - LP(k - 2, j, i)) +
                Qg(k - 1, j, i).rho * a_sound *
                    (LQ(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho - LQ(k - 2, j, i).rhoW / Qg(k - 2, j, i).rho));
          double d0 = 1. / (a_sound * a_sound) * (L1 + 0.5 * (L4 + L0));
          double d1 = 0.5 * (L0 + L4);
          double d2 = 0.5 / (Qg(k - 1, j, i).rho * a_sound) * (L4 - L0);
          (void)L2; (void)L3;
          Qg(k, j, i).rho = QO(k, j, i).rho - dt * alfa * d0;
          Qg(k, j, i).rhoE = QO(k, j, i).rhoE -
                             dt * alfa * (0.5 * d0 * (LQ(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho *
                                                      LQ(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho) +
                                          d1 / (gam - 1.) + LQ(k - 1, j, i).rhoW * d2);
          Qg(k, j, i).rhoW = QO(k, j, i).rhoW - dt * alfa *
                                                    (LQ(k - 1, j, i).rhoW / Qg(k - 1, j, i).rho * d0 +
                                                     Qg(k - 1, j, i).rho * d2);
        }
      }
  }
  /* periodic seams from the stage-input local vector, rhs.c:1107-1227 */
  if (bc[0] == 7 || bc[1] == 7) {
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        if (k > 0 && k < mz && j > 0 && j < my) { /* user->KM==mz, JM==my */
          Qg(k, j, 0) = LQ(k, j, -2);
          p[GI(c, k, j, 0)] = LP(k, j, -2);
        }
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        if (k > 0 && k < mz && j > 0 && j < my) {
          Qg(k, j, mx - 1) = LQ(k, j, mx + 1);
          p[GI(c, k, j, mx - 1)] = LP(k, j, mx + 1);
        }
  }
  if (bc[2] == 7 || bc[3] == 7) {
    for (int k = 0; k < mz; k++)
      for (int i = 0; i < mx; i++)
        if (k > 0 && k < mz) {
          Qg(k, 0, i) = LQ(k, -2, i);
          p[GI(c, k, 0, i)] = LP(k, -2, i);
        }
    for (int k = lzs; k < lze; k++)
      for (int i = 0; i < mx; i++)
        if (k > 0 && k < mz) {
          Qg(k, my - 1, i) = LQ(k, my + 1, i);
          p[GI(c, k, my - 1, i)] = LP(k, my + 1, i);
        }
  }
  if (bc[4] == 7 || bc[5] == 7) {
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        Qg(0, j, i) = LQ(-2, j, i);
        p[GI(c, 0, j, i)] = LP(-2, j, i);
      }
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        Qg(mz - 1, j, i) = LQ(mz + 1, j, i);
        p[GI(c, mz - 1, j, i)] = LP(mz + 1, j, i);
      }
  }
#undef Qg
#undef QO
#undef LQ
#undef LP
#undef AJ
#undef NV
#undef ETA
#undef ZET
  orc_g2l(c, (double *)c->Q, (double *)c->lQ, 5);
  orc_g2l(c, c->P, c->lP, 1);
}

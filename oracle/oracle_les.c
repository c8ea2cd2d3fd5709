/*
 * oracle_les.c -- dynamic Smagorinsky model and eddy viscosity of the CPU
 * restatement (test infrastructure only).
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

static const double LES_EPS = 0.000001; /* les.c:8 */

/* Compute_du_center + Compute_du_dxyz (les.c:31-63) at local index n. */
static void grad_at(const orc_ctx *c, const orc_v3 *um, const orc_v3 *up_i, const orc_v3 *up_j,
                    const orc_v3 *up_k, const orc_v3 *um_i, const orc_v3 *um_j, const orc_v3 *um_k,
                    size_t n, double g[9]) {
  (void)um;
  double dudc = (up_i->x - um_i->x) * 0.5, dvdc = (up_i->y - um_i->y) * 0.5, dwdc = (up_i->z - um_i->z) * 0.5;
  double dudz = (up_k->x - um_k->x) * 0.5, dvdz = (up_k->y - um_k->y) * 0.5, dwdz = (up_k->z - um_k->z) * 0.5;
  double dude = (up_j->x - um_j->x) * 0.5, dvde = (up_j->y - um_j->y) * 0.5, dwde = (up_j->z - um_j->z) * 0.5;
  const orc_v3 *cs = &c->lcsi[n], *et = &c->leta[n], *ze = &c->lzet[n];
  double ajc = c->laj[n];
  double csi0 = cs->x, csi1 = cs->y, csi2 = cs->z;
  double eta0 = et->x, eta1 = et->y, eta2 = et->z;
  double zet0 = ze->x, zet1 = ze->y, zet2 = ze->z;
  g[0] = (dudc * csi0 + dude * eta0 + dudz * zet0) * ajc; /* du_dx */
  g[1] = (dudc * csi1 + dude * eta1 + dudz * zet1) * ajc; /* du_dy */
  g[2] = (dudc * csi2 + dude * eta2 + dudz * zet2) * ajc; /* du_dz */
  g[3] = (dvdc * csi0 + dvde * eta0 + dvdz * zet0) * ajc; /* dv_dx */
  g[4] = (dvdc * csi1 + dvde * eta1 + dvdz * zet1) * ajc; /* dv_dy */
  g[5] = (dvdc * csi2 + dvde * eta2 + dvdz * zet2) * ajc; /* dv_dz */
  g[6] = (dwdc * csi0 + dwde * eta0 + dwdz * zet0) * ajc; /* dw_dx */
  g[7] = (dwdc * csi1 + dwde * eta1 + dwdz * zet1) * ajc; /* dw_dy */
  g[8] = (dwdc * csi2 + dwde * eta2 + dwdz * zet2) * ajc; /* dw_dz */
}

/* |S| with the trace term 2/3 div subtracted from the diagonal (les.c:268-274) */
static double sabs_of(const double g[9]) {
  double du_dx = g[0], du_dy = g[1], du_dz = g[2], dv_dx = g[3], dv_dy = g[4], dv_dz = g[5];
  double dw_dx = g[6], dw_dy = g[7], dw_dz = g[8];
  double Sdd = 2. / 3. * (du_dx + dv_dy + dw_dz);
  double Sxx = 0.5 * (du_dx + du_dx) - Sdd, Sxy = 0.5 * (du_dy + dv_dx), Sxz = 0.5 * (du_dz + dw_dx);
  double Syx = Sxy, Syy = 0.5 * (dv_dy + dv_dy) - Sdd, Syz = 0.5 * (dv_dz + dw_dy);
  double Szx = Sxz, Szy = Syz, Szz = 0.5 * (dw_dz + dw_dz) - Sdd;
  return sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz + Szx * Szx +
                     Szy * Szy + Szz * Szz));
}

/* The reference gradient loop (les.c:250-309) overwrites lUcat in place
 * while it walks k,j,i: when node (k,j,i) is visited its -1 neighbours that
 * are interior and not solid already hold the new velocity rho*u/rho of lQ,
 * while its +1 neighbours and every boundary/ghost/solid node still hold the
 * velocity of the last CalcUCat.  This two-array form states that exactly. */
static int in_loop(const orc_ctx *c, int k, int j, int i) {
  return k >= 1 && k <= c->mz - 2 && j >= 1 && j <= c->my - 2 && i >= 1 && i <= c->mx - 2 &&
         !(c->lnvert[LI(c, k, j, i)] > 1.1);
}
static orc_v3 unew_at(const orc_ctx *c, size_t n) {
  const orc_q *q = &c->lQ[n];
  orc_v3 u = {q->rhoU / q->rho, q->rhoV / q->rho, q->rhoW / q->rho};
  return u;
}

static void gradients_two_array(orc_ctx *c, orc_v3 *Ax, orc_v3 *Ay, orc_v3 *Az, double *S, int update_ucat) {
  const size_t sx = 1, sy = (size_t)c->LX, sz = (size_t)c->LX * c->LY;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 1; i < c->mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        if (c->lnvert[n] > 1.1) continue;
        orc_v3 um_i = in_loop(c, k, j, i - 1) ? unew_at(c, n - sx) : c->lUcat[n - sx];
        orc_v3 um_j = in_loop(c, k, j - 1, i) ? unew_at(c, n - sy) : c->lUcat[n - sy];
        orc_v3 um_k = in_loop(c, k - 1, j, i) ? unew_at(c, n - sz) : c->lUcat[n - sz];
        double g[9];
        grad_at(c, NULL, &c->lUcat[n + sx], &c->lUcat[n + sy], &c->lUcat[n + sz], &um_i, &um_j, &um_k, n, g);
        S[n] = sabs_of(g);
        Ax[n].x = g[0]; Ax[n].y = g[1]; Ax[n].z = g[2];
        Ay[n].x = g[3]; Ay[n].y = g[4]; Ay[n].z = g[5];
        Az[n].x = g[6]; Az[n].y = g[7]; Az[n].z = g[8];
      }
  if (update_ucat) {
    #pragma omp parallel for schedule(static)
    for (int k = 1; k < c->mz - 1; k++)
      for (int j = 1; j < c->my - 1; j++)
        for (int i = 1; i < c->mx - 1; i++) {
          size_t n = LI(c, k, j, i);
          if (c->lnvert[n] > 1.1) continue;
          c->lUcat[n] = unew_at(c, n);
        }
  }
}

static void export_grad(const orc_ctx *c, const orc_v3 *Ax, const orc_v3 *Ay, const orc_v3 *Az,
                        const double *S, double *A9, double *sabs) {
  for (int k = 0; k < c->mz; k++)
    for (int j = 0; j < c->my; j++)
      for (int i = 0; i < c->mx; i++) {
        size_t g = GI(c, k, j, i), n = LI(c, k, j, i);
        if (A9) {
          double *a = A9 + 9 * g;
          a[0] = Ax[n].x; a[1] = Ax[n].y; a[2] = Ax[n].z;
          a[3] = Ay[n].x; a[4] = Ay[n].y; a[5] = Ay[n].z;
          a[6] = Az[n].x; a[7] = Az[n].y; a[8] = Az[n].z;
        }
        if (sabs) sabs[g] = S[n];
      }
}

void orc_les_gradients(orc_ctx *c, double *A9, double *sabs) {
  orc_v3 *Ax = calloc(c->nL, sizeof(orc_v3)), *Ay = calloc(c->nL, sizeof(orc_v3)), *Az = calloc(c->nL, sizeof(orc_v3));
  double *S = calloc(c->nL, sizeof(double));
  gradients_two_array(c, Ax, Ay, Az, S, 0);
  export_grad(c, Ax, Ay, Az, S, A9, sabs);
  free(Ax); free(Ay); free(Az); free(S);
}

/* Literal serial restatement of les.c:250-309, run on a copy of lUcat. */
void orc_les_gradients_literal(orc_ctx *c, double *A9, double *sabs) {
  orc_v3 *Ax = calloc(c->nL, sizeof(orc_v3)), *Ay = calloc(c->nL, sizeof(orc_v3)), *Az = calloc(c->nL, sizeof(orc_v3));
  double *S = calloc(c->nL, sizeof(double));
  orc_v3 *ucat = malloc(c->nL * sizeof(orc_v3));
  memcpy(ucat, c->lUcat, c->nL * sizeof(orc_v3));
  const size_t sx = 1, sy = (size_t)c->LX, sz = (size_t)c->LX * c->LY;
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 1; i < c->mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        if (c->lnvert[n] > 1.1) continue;
        ucat[n] = unew_at(c, n);
        double g[9];
        grad_at(c, NULL, &ucat[n + sx], &ucat[n + sy], &ucat[n + sz], &ucat[n - sx], &ucat[n - sy], &ucat[n - sz], n, g);
        S[n] = sabs_of(g);
        Ax[n].x = g[0]; Ax[n].y = g[1]; Ax[n].z = g[2];
        Ay[n].x = g[3]; Ay[n].y = g[4]; Ay[n].z = g[5];
        Az[n].x = g[6]; Az[n].y = g[7]; Az[n].z = g[8];
      }
  export_grad(c, Ax, Ay, Az, S, A9, sabs);
  free(Ax); free(Ay); free(Az); free(S); free(ucat);
}

/* periodic face copies on a local array restricted to the interior ranges
 * of the other two directions (les.c:316-377, 1196-1256, 1378-1430) */
static void seam_copy_local(const orc_ctx *c, double *l, int nc) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  const int *bc = c->p.bctype;
#define CP(kd, jd, id_, ks, js, is)                                            \
  memcpy(l + LI(c, kd, jd, id_) * nc, l + LI(c, ks, js, is) * nc, nc * sizeof(double))
  if (bc[0] == 7 || bc[1] == 7) {
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++) CP(k, j, 0, k, j, -2);
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++) CP(k, j, mx - 1, k, j, mx + 1);
  }
  if (bc[2] == 7 || bc[3] == 7) {
    for (int k = 1; k < mz - 1; k++)
      for (int i = 1; i < mx - 1; i++) CP(k, 0, i, k, -2, i);
    for (int k = 1; k < mz - 1; k++)
      for (int i = 1; i < mx - 1; i++) CP(k, my - 1, i, k, my + 1, i);
  }
  if (bc[4] == 7 || bc[5] == 7) {
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) CP(0, j, i, -2, j, i);
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) CP(mz - 1, j, i, mz + 1, j, i);
  }
#undef CP
}

/* integrate_testfilter, les.c:80-131: weighted mean over the 3x3x3 box in
 * [k][j][i] order. */
static double integ_w(const double v[27], const double w[27]) {
  double vs = 0.0, ws = 0.0;
  for (int n = 0; n < 27; n++) { vs += w[n] * v[n]; ws += w[n]; }
  return vs / ws;
}
static double integ_t(const double v[27]) { /* integrate_t, les.c:66-78 */
  double s = 0.0;
  for (int n = 0; n < 27; n++) s += v[n] / 27.0;
  return s;
}

/* Homogeneous-direction averages of LM and MM (les.c:934-1116), after the
 * pointwise Cs: with -i and -k_homo_filter the mean over each j plane of the
 * interior fluid nodes, Cs = 0.5 <LM>_j / (<MM>_j + eps) on every interior
 * node; with -i alone the mean along each (k, j) line, set on the interior
 * fluid nodes; with all three, Cs = sum LM / sum MM over the interior (no
 * fluid test) on every owned node.  -j or -k alone average nothing (their
 * branches are commented out in the reference) but, like the others, drop
 * the 27-point means and the smoothing. */
static void homogeneous_cs(orc_ctx *c, const double *LM, const double *MM) {
  const orc_params *P = &c->p;
  const int mx = c->mx, my = c->my, mz = c->mz;
  if (P->i_homo_filter && P->k_homo_filter) {
    double *JL = calloc(my, sizeof(double)), *JM = calloc(my, sizeof(double));
    int *cnt = calloc(my, sizeof(int));
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++)
          if (c->lnvert[LI(c, k, j, i)] < 0.1) {
            JL[j] += LM[LI(c, k, j, i)];
            JM[j] += MM[LI(c, k, j, i)];
            cnt[j]++;
          }
    for (int j = 0; j < my; j++)
      if (cnt[j] > 0) { JL[j] /= (double)cnt[j]; JM[j] /= (double)cnt[j]; }
    for (int j = 1; j < my - 1; j++)
      for (int k = 1; k < mz - 1; k++)
        for (int i = 1; i < mx - 1; i++) c->CS[GI(c, k, j, i)] = 0.5 * JL[j] / (JM[j] + LES_EPS);
    free(JL); free(JM); free(cnt);
  }
  if (P->i_homo_filter && !P->k_homo_filter) {
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++) {
        double sl = 0.0, sm = 0.0;
        int cnt = 0;
        for (int i = 1; i < mx - 1; i++)
          if (c->lnvert[LI(c, k, j, i)] < 0.1) {
            sl += LM[LI(c, k, j, i)];
            sm += MM[LI(c, k, j, i)];
            cnt++;
          }
        if (cnt > 0) { sl /= (double)cnt; sm /= (double)cnt; }
        for (int i = 1; i < mx - 1; i++)
          if (c->lnvert[LI(c, k, j, i)] < 0.1) c->CS[GI(c, k, j, i)] = 0.5 * sl / (sm + LES_EPS);
      }
  }
  if (P->i_homo_filter && P->j_homo_filter && P->k_homo_filter) {
    double sl = 0.0, sm = 0.0;
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          sm += MM[LI(c, k, j, i)];
          sl += LM[LI(c, k, j, i)];
        }
    for (size_t n = 0; n < c->nG; n++) c->CS[n] = sl / sm;
  }
}

/* Compute_Smagorinsky_Constant_1, les.c:135-1472. */
static void smagorinsky(orc_ctx *c) {
  const orc_params *P = &c->p;
  const int mx = c->mx, my = c->my, mz = c->mz;
  /* les.c:139 reads the global tistart after main.c:890-891 has incremented
   * it (0 -> 1 for a fresh run, N -> N+1 for a restart): never 0 in the time
   * loop, so only les == 10 takes this branch */
  const int tistart_in_loop = P->tistart + 1;
  if ((c->ti < 2 && tistart_in_loop == 0 && !P->rstart) || P->les == 10) {
    for (size_t n = 0; n < c->nG; n++) c->CS[n] = 0.0;
    return;
  }
  if (P->les == 1) {
    for (size_t n = 0; n < c->nG; n++) c->CS[n] = 0.03;
    return;
  }
  const double max_cs = P->max_cs;
  const int homo = P->i_homo_filter || P->j_homo_filter || P->k_homo_filter;
  orc_v3 *Ax = calloc(c->nL, sizeof(orc_v3)), *Ay = calloc(c->nL, sizeof(orc_v3)), *Az = calloc(c->nL, sizeof(orc_v3));
  double *S = calloc(c->nL, sizeof(double)), *LM = calloc(c->nL, sizeof(double)), *MM = calloc(c->nL, sizeof(double));
  gradients_two_array(c, Ax, Ay, Az, S, 1);
  orc_l2l(c, (double *)Ax, 3);
  orc_l2l(c, (double *)Ay, 3);
  orc_l2l(c, (double *)Az, 3);
  orc_l2l(c, S, 1);
  seam_copy_local(c, (double *)Ax, 3);
  seam_copy_local(c, (double *)Ay, 3);
  seam_copy_local(c, (double *)Az, 3);

  /* test-filtered Germano identity contracted with S (les.c:413-803) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t n0 = LI(c, k, j, i);
        if (c->lnvert[n0] > 1.1) { LM[n0] = MM[n0] = 0; continue; }
        double w[27], u[27], v[27], wv[27], rho[27], ru[27], rv[27], rw[27];
        double ruu[27], ruv[27], ruw[27], rvv[27], rvw[27], rww[27];
        double Sab[9][27], SS[9][27];
        for (int p = -1; p <= 1; p++)
          for (int q = -1; q <= 1; q++)
            for (int r = -1; r <= 1; r++) {
              int R = r + 1, Qi = q + 1, Pi = p + 1;
              int a = (R * 3 + Qi) * 3 + Pi;
              size_t n = LI(c, k + r, j + q, i + p);
              w[a] = 1. / (fabs(R - 1.0) + 1) * 1. / (fabs(Qi - 1.0) + 1) * 1. / (fabs(Pi - 1.0) + 1);
              u[a] = c->lUcat[n].x; v[a] = c->lUcat[n].y; wv[a] = c->lUcat[n].z;
              ru[a] = c->lQ[n].rhoU; rv[a] = c->lQ[n].rhoV; rw[a] = c->lQ[n].rhoW; rho[a] = c->lQ[n].rho;
              ruu[a] = ru[a] * u[a]; ruv[a] = ru[a] * v[a]; ruw[a] = ru[a] * wv[a];
              rvv[a] = rv[a] * v[a]; rvw[a] = rv[a] * wv[a]; rww[a] = rw[a] * wv[a];
              double du_dx = Ax[n].x, du_dy = Ax[n].y, du_dz = Ax[n].z;
              double dv_dx = Ay[n].x, dv_dy = Ay[n].y, dv_dz = Ay[n].z;
              double dw_dx = Az[n].x, dw_dy = Az[n].y, dw_dz = Az[n].z;
              double Sxx = 0.5 * (du_dx + du_dx), Sxy = 0.5 * (du_dy + dv_dx), Sxz = 0.5 * (du_dz + dw_dx);
              double Syy = 0.5 * (dv_dy + dv_dy), Syz = 0.5 * (dv_dz + dw_dy);
              double Szz = 0.5 * (dw_dz + dw_dz);
              double Sv[9] = {Sxx, Sxy, Sxz, Sxy, Syy, Syz, Sxz, Syz, Szz};
              double Sn = S[n];
              for (int e = 0; e < 9; e++) { Sab[e][a] = Sv[e]; SS[e][a] = Sv[e] * Sn * rho[a]; }
            }
        double filter = pow(1. / c->laj[n0], 1. / 3.);
        double test_filter = P->testfilter_ik ? pow(5.0, 1. / 3.) * filter : 2.0 * filter;
        double _rho = integ_w(rho, w), _rhou = integ_w(ru, w), _rhov = integ_w(rv, w), _rhow = integ_w(rw, w);
        double Lij[3][3];
        Lij[0][0] = integ_w(ruu, w) - _rhou * _rhou / _rho;
        Lij[0][1] = integ_w(ruv, w) - _rhou * _rhov / _rho;
        Lij[0][2] = integ_w(ruw, w) - _rhou * _rhow / _rho;
        Lij[1][0] = integ_w(ruv, w) - _rhov * _rhou / _rho;
        Lij[1][1] = integ_w(rvv, w) - _rhov * _rhov / _rho;
        Lij[1][2] = integ_w(rvw, w) - _rhov * _rhow / _rho;
        Lij[2][0] = integ_w(ruw, w) - _rhow * _rhou / _rho;
        Lij[2][1] = integ_w(rvw, w) - _rhow * _rhov / _rho;
        Lij[2][2] = integ_w(rww, w) - _rhow * _rhow / _rho;
        double Sxyz[3][3], Sh[3][3], SSh[3][3];
        for (int e = 0; e < 9; e++) {
          Sxyz[e / 3][e % 3] = Sab[e][13];
          Sh[e / 3][e % 3] = integ_w(Sab[e], w);
          SSh[e / 3][e % 3] = integ_w(SS[e], w);
        }
        double S_hat = 0;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) S_hat += orc_sq(Sh[a][b]);
        S_hat = sqrt(2 * S_hat);
        double L_cat = 0.0, M_cat = 0.0;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            M_cat += -2.0 * _rho * orc_sq(test_filter) * S_hat * (Sh[a][b] * Sxyz[a][b] - 0.33 * Sh[a][a] * Sxyz[b][b]) +
                     2.0 * orc_sq(filter) * (SSh[a][b] * Sxyz[a][b] - 0.33 * SSh[a][a] * Sxyz[b][b]);
            L_cat += Lij[a][b] * Sxyz[a][b] - 0.33 * Sxyz[a][a] * Lij[b][b];
          }
        LM[n0] = L_cat;
        MM[n0] = M_cat;
      }
  orc_l2l(c, LM, 1);
  orc_l2l(c, MM, 1);
  if (c->dbg_lm) {
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          c->dbg_lm[GI(c, k, j, i)] = LM[LI(c, k, j, i)];
          c->dbg_mm[GI(c, k, j, i)] = MM[LI(c, k, j, i)];
        }
  }

  /* Cs = <LM>/(<MM>+eps) with 27-point means (les.c:889-930) */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t g = GI(c, k, j, i);
        if (c->lnvert[LI(c, k, j, i)] > 1.1) { c->CS[g] = 0; continue; }
        double LM0[27], MM0[27];
        for (int a = -1; a <= 1; a++)
          for (int b = -1; b <= 1; b++)
            for (int cc = -1; cc <= 1; cc++) {
              int idx = ((cc + 1) * 3 + (b + 1)) * 3 + (a + 1);
              size_t n = LI(c, k + cc, j + b, i + a);
              LM0[idx] = LM[n]; MM0[idx] = MM[n];
            }
        double LM_avg, MM_avg;
        if (homo || P->les == 3) { LM_avg = LM[LI(c, k, j, i)]; MM_avg = MM[LI(c, k, j, i)]; }
        else { LM_avg = integ_t(LM0); MM_avg = integ_t(MM0); }
        double C = LM_avg / (MM_avg + LES_EPS);
        c->CS[g] = P->les == 1 ? 0.01 : pmax(C, 0.0001);
      }
  homogeneous_cs(c, LM, MM);
  /* clipping on every owned node (les.c:1146-1161); the bctype clauses of
   * that test are constant-false ("!bctype==7") */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        size_t g = GI(c, k, j, i);
        double nv = c->lnvert[LI(c, k, j, i)];
        if (nv > 1.1 && nv < 7.) {
          c->CS[g] = 0.0;
        } else {
          if (nv > 0.1 && nv < 1.1) c->CS[g] = pmax(0.01, c->CS[g]);
          c->CS[g] = pmin(pmax(c->CS[g], 0.000), max_cs);
        }
      }
  orc_g2l(c, c->CS, c->lCS, 1);
  seam_copy_local(c, c->lCS, 1);
  if (P->les != 3 && !homo) {
    /* volume-weighted 27-point smoothing of Cs (les.c:1258-1337) */
    double *tmp = malloc(c->nL * sizeof(double));
    memcpy(tmp, c->lCS, c->nL * sizeof(double));
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          if (!(c->lnvert[LI(c, k, j, i)] < 0.1)) continue;
          double w[27], C[27];
          for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
              for (int cc = -1; cc <= 1; cc++) {
                int idx = ((cc + 1) * 3 + (b + 1)) * 3 + (a + 1);
                size_t n = LI(c, k + cc, j + b, i + a);
                w[idx] = 1. / c->laj[n];
                C[idx] = tmp[n];
              }
          c->CS[GI(c, k, j, i)] = integ_w(C, w);
        }
    free(tmp);
  }
  orc_g2l(c, c->CS, c->lCS, 1);
  seam_copy_local(c, c->lCS, 1);
  free(Ax); free(Ay); free(Az); free(S); free(LM); free(MM);
}

/* Compute_eddy_viscosity_LES, les.c:1477-1661: nu_t = Cs (Delta)^2 |S|.
 * The Vreman energy terms (beta, bbeta) it also forms are not read by the
 * right-hand side and are not restated. */
static void eddy_viscosity(orc_ctx *c) {
  const size_t sx = 1, sy = (size_t)c->LX, sz = (size_t)c->LX * c->LY;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 1; i < c->mx - 1; i++) {
        size_t n = LI(c, k, j, i), g = GI(c, k, j, i);
        if (c->lnvert[n] > 0.1) { c->Nu_t[g] = 0; continue; }
        const orc_v3 *u = c->lUcat;
        double gr[9];
        grad_at(c, NULL, &u[n + sx], &u[n + sy], &u[n + sz], &u[n - sx], &u[n - sy], &u[n - sz], n, gr);
        double Sabs = sabs_of(gr);
        double filter = pow(1. / c->laj[n], 1. / 3.);
        c->Nu_t[g] = c->CS[g] * orc_sq(filter) * Sabs;
      }
  orc_g2l(c, c->Nu_t, c->lNu_t, 1);
}

void orc_les_debug(orc_ctx *c, double *lm, double *mm) {
  c->dbg_lm = lm;
  c->dbg_mm = mm;
}

void orc_les_update(orc_ctx *c) {
  smagorinsky(c);
  eddy_viscosity(c);
}

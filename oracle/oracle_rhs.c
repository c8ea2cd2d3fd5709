/*
 * oracle_rhs.c -- convective and viscous right-hand side, MultAj, formRHS and
 * the SSP Runge-Kutta driver of the CPU restatement (test infrastructure).
 */
#include "oracle_internal.h"

#include <stdlib.h>
#include <string.h>

static const double BLANKWALL = 2.9, BLANK = 0.1;

/* ---------------- reconstruction helpers, rhs.c:2256-2568 -------------- */
double orc_minmod(double a, double b) { /* rhs.c:2256-2263 */
  if (a * b < 0) return 0.;
  return (a >= 0 ? 1.0 : -1.0) * pmin(fabs(a), fabs(b));
}
double orc_muscl_ql(double qm, double q0, double qp) { /* rhs.c:2266-2285 */
  double Splus = qp - q0, Sminus = q0 - qm;
  double S = orc_minmod(Splus, Sminus);
  return q0 + 0.5 * S;
}
double orc_muscl_qr(double qm, double q0, double qp) { /* rhs.c:2287-2306 */
  double Splus = qp - q0, Sminus = q0 - qm;
  double S = orc_minmod(Splus, Sminus);
  return q0 - 0.5 * S;
}
/* 3rd-order WENO, Shu procedure 2.2, truncated linear weights (rhs.c:2359-2409) */
double orc_weno_ql(double qm, double q0, double qp) {
  const double d0 = 0.6666667, d1 = 0.3333333, eps = 1e-6;
  double Qr0 = 0.5 * q0 + 0.5 * qp;
  double Qr1 = -0.5 * qm + 1.5 * q0;
  double beta0 = (qp - q0) * (qp - q0);
  double beta1 = (qm - q0) * (qm - q0);
  double alfa0 = d0 / ((eps + beta0) * (eps + beta0));
  double alfa1 = d1 / ((eps + beta1) * (eps + beta1));
  double alfasum = alfa0 + alfa1;
  double w0 = alfa0 / alfasum, w1 = alfa1 / alfasum;
  return w0 * Qr0 + w1 * Qr1;
}
double orc_weno_qr(double qm, double q0, double qp) { /* rhs.c:2411-2461 */
  const double d0 = 0.6666667, d1 = 0.3333333, eps = 1e-6;
  double Qr0 = 1.5 * q0 - 0.5 * qp;
  double Qr1 = 0.5 * qm + 0.5 * q0;
  double beta0 = (qp - q0) * (qp - q0);
  double beta1 = (qm - q0) * (qm - q0);
  double alfa0 = d0 / ((eps + beta0) * (eps + beta0));
  double alfa1 = d1 / ((eps + beta1) * (eps + beta1));
  double alfasum = alfa0 + alfa1;
  double w0 = alfa0 / alfasum, w1 = alfa1 / alfasum;
  return w0 * Qr0 + w1 * Qr1;
}
double orc_eno_qr(double qm, double q0, double qp, int k) { /* rhs.c:2498-2533 */
  if (k == 2) return 0.5 * q0 + 0.5 * qp;
  return -qm * 0.1666666666666666 + q0 * 0.8333333333333333 + qp * 0.3333333333333333;
}
double orc_eno_ql(double qm, double q0, double qp, int k) { /* rhs.c:2535-2568 */
  if (k == 2) return 0.5 * q0 + 0.5 * qm;
  return qm * 0.3333333333333333 + q0 * 0.8333333333333333 - qp * 0.1666666666666666;
}
/* skew-symmetric 4-point central product, rhs.c:2468-2474 */
double orc_fp_2(const double u[4], const double v[4]) {
  return (u[1] + u[2]) * (v[1] + v[2]) / 3.0 -
         1. / 24. * (u[0] * v[0] + u[0] * v[2] + u[1] * v[1] + u[1] * v[3] + u[2] * v[2] + u[2] * v[0] +
                     u[3] * v[1] + u[3] * v[3]);
}
static double fp_1(const double u[4], const double v[4]) { /* rhs.c:2476-2481 */
  return 1.0 / 12.0 * (-u[0] * v[0] + 7 * u[1] * v[1] + 7 * u[2] * v[2] - u[3] * v[3]);
}
double orc_fp4(const double u[4]) { /* rhs.c:2483-2489 */
  return 1.0 / 12.0 * (-u[0] + 7 * u[1] + 7 * u[2] - u[3]);
}

/* Rusanov (local Lax-Friedrichs) flux with gamma hard-coded to 1.4
 * (rhs.c:2596, 2779-2810).  s = {rho,rhoU,rhoV,rhoW,rhoE,p}. */
static void rusanov6(const double *L, const double *R, const double *h, double uconL,
                     double uconR, orc_q *F) {
  const double gamma = 1.4;
  double gii = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  double CsoundL = gamma * L[5] / L[0];
  double CsoundR = gamma * R[5] / R[0];
  double saL = fabs(uconL) + sqrt(fabs(CsoundL) * gii);
  double saR = fabs(uconR) + sqrt(fabs(CsoundR) * gii);
  double SA = pmax(saL, saR);
  F->rho = 0.5 * (L[0] * uconL + R[0] * uconR) - 0.5 * SA * (R[0] - L[0]);
  F->rhoU = 0.5 * ((L[1] * uconL + R[1] * uconR) + (L[5] * h[0] + R[5] * h[0])) - 0.5 * SA * (R[1] - L[1]);
  F->rhoV = 0.5 * ((L[2] * uconL + R[2] * uconR) + (L[5] * h[1] + R[5] * h[1])) - 0.5 * SA * (R[2] - L[2]);
  F->rhoW = 0.5 * ((L[3] * uconL + R[3] * uconR) + (L[5] * h[2] + R[5] * h[2])) - 0.5 * SA * (R[3] - L[3]);
  F->rhoE = 0.5 * ((L[5] + L[4]) * uconL + (R[5] + R[4]) * uconR) - 0.5 * SA * (R[4] - L[4]);
}
static inline double ucon_of(const double *s, const double *h) {
  return h[0] * (s[1] / s[0]) + h[1] * (s[2] / s[0]) + h[2] * (s[3] / s[0]);
}
/* ---------------------------------------------------------------------
 * Build-only schemes selectable beside the reference's (no counterpart in
 * the reference; the options weno5 and roe):
 *  - WENO5 of Jiang & Shu (J. Comput. Phys. 126, 1996, eqs. 2.10-2.17): the
 *    left state at face i+1/2 from q[i-2..i+2], candidates
 *    (2q0 - 7q1 + 11q2)/6, (-q1 + 5q2 + 2q3)/6, (2q2 + 5q3 - q4)/6, linear
 *    weights 0.1, 0.6, 0.3, smoothness indicators 13/12 (2nd difference)^2 +
 *    1/4 (1st difference)^2, alpha = d / (1e-6 + beta)^2; the right state
 *    mirrors it on q[i+3..i-1].  Applied to the point values as the
 *    reference applies its WENO3 (rhs.c:2359-2461).
 *  - The Roe flux (Roe, J. Comput. Phys. 43, 1981) in the contravariant form
 *    of the reference's Rusanov flux: F = (F(L) + F(R))/2 - |h|/2 sum_k
 *    |lambda_k| alpha_k r_k with Roe averages across the face normal h/|h|,
 *    gamma of the equation of state, and Harten's entropy fix on the two
 *    acoustic waves (|lambda| < delta: (lambda^2 + delta^2) / (2 delta),
 *    delta = 0.1 c).
 * --------------------------------------------------------------------- */
double orc_weno5_ql(const double q[5]) {
  const double eps = 1e-6, d0 = 0.1, d1 = 0.6, d2 = 0.3;
  const double b0 = 13.0 / 12.0 * (q[0] - 2 * q[1] + q[2]) * (q[0] - 2 * q[1] + q[2]) +
                    0.25 * (q[0] - 4 * q[1] + 3 * q[2]) * (q[0] - 4 * q[1] + 3 * q[2]);
  const double b1 = 13.0 / 12.0 * (q[1] - 2 * q[2] + q[3]) * (q[1] - 2 * q[2] + q[3]) +
                    0.25 * (q[1] - q[3]) * (q[1] - q[3]);
  const double b2 = 13.0 / 12.0 * (q[2] - 2 * q[3] + q[4]) * (q[2] - 2 * q[3] + q[4]) +
                    0.25 * (3 * q[2] - 4 * q[3] + q[4]) * (3 * q[2] - 4 * q[3] + q[4]);
  const double a0 = d0 / ((eps + b0) * (eps + b0)), a1 = d1 / ((eps + b1) * (eps + b1)),
               a2 = d2 / ((eps + b2) * (eps + b2));
  const double p0 = (2 * q[0] - 7 * q[1] + 11 * q[2]) / 6.0, p1 = (-q[1] + 5 * q[2] + 2 * q[3]) / 6.0,
               p2 = (2 * q[2] + 5 * q[3] - q[4]) / 6.0;
  return (a0 * p0 + a1 * p1 + a2 * p2) / (a0 + a1 + a2);
}

static double harten(double lam, double delta) {
  const double a = fabs(lam);
  return a < delta ? (lam * lam + delta * delta) / (2 * delta) : a;
}

/* physical contravariant flux of s = {rho,rhoU,rhoV,rhoW,rhoE,p} across h */
static void phys_flux(const double *s, const double *h, double F[5]) {
  const double U = (h[0] * s[1] + h[1] * s[2] + h[2] * s[3]) / s[0];
  F[0] = s[0] * U;
  F[1] = s[1] * U + s[5] * h[0];
  F[2] = s[2] * U + s[5] * h[1];
  F[3] = s[3] * U + s[5] * h[2];
  F[4] = (s[4] + s[5]) * U;
}

static void roe6(const double *L, const double *R, const double *h, double gamma, orc_q *Fo) {
  double FL[5], FR[5];
  phys_flux(L, h, FL);
  phys_flux(R, h, FR);
  const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  const double n[3] = {h[0] / hn, h[1] / hn, h[2] / hn};
  const double uL[3] = {L[1] / L[0], L[2] / L[0], L[3] / L[0]}, uR[3] = {R[1] / R[0], R[2] / R[0], R[3] / R[0]};
  const double HL = (L[4] + L[5]) / L[0], HR = (R[4] + R[5]) / R[0];
  const double sL = sqrt(L[0]), sR = sqrt(R[0]), ws = sL + sR;
  double u[3];
  for (int c = 0; c < 3; c++) u[c] = (sL * uL[c] + sR * uR[c]) / ws;
  const double H = (sL * HL + sR * HR) / ws, rho = sL * sR;
  const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double c2 = (gamma - 1) * (H - 0.5 * q2), c = sqrt(c2);
  const double Un = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
  const double dr = R[0] - L[0], dp = R[5] - L[5];
  double du[3];
  for (int k = 0; k < 3; k++) du[k] = uR[k] - uL[k];
  const double dUn = n[0] * du[0] + n[1] * du[1] + n[2] * du[2];
  const double a1 = (dp - rho * c * dUn) / (2 * c2), a2 = dr - dp / c2, a5 = (dp + rho * c * dUn) / (2 * c2);
  const double delta = 0.1 * c;
  const double l1 = harten(Un - c, delta), l2 = fabs(Un), l5 = harten(Un + c, delta);
  double dt[3];  /* tangential velocity jump */
  for (int k = 0; k < 3; k++) dt[k] = du[k] - dUn * n[k];
  const double udt = u[0] * dt[0] + u[1] * dt[1] + u[2] * dt[2];
  double D[5];
  D[0] = l1 * a1 + l2 * a2 + l5 * a5;
  for (int k = 0; k < 3; k++)
    D[1 + k] = l1 * a1 * (u[k] - c * n[k]) + l2 * (a2 * u[k] + rho * dt[k]) + l5 * a5 * (u[k] + c * n[k]);
  D[4] = l1 * a1 * (H - c * Un) + l2 * (a2 * 0.5 * q2 + rho * udt) + l5 * a5 * (H + c * Un);
  Fo->rho = 0.5 * (FL[0] + FR[0]) - 0.5 * hn * D[0];
  Fo->rhoU = 0.5 * (FL[1] + FR[1]) - 0.5 * hn * D[1];
  Fo->rhoV = 0.5 * (FL[2] + FR[2]) - 0.5 * hn * D[2];
  Fo->rhoW = 0.5 * (FL[3] + FR[3]) - 0.5 * hn * D[3];
  Fo->rhoE = 0.5 * (FL[4] + FR[4]) - 0.5 * hn * D[4];
}

void orc_roe(const double ql[5], double pl, const double qr[5], double pr, const double h[3], double gamma,
             double out[5]) {
  double L[6] = {ql[0], ql[1], ql[2], ql[3], ql[4], pl};
  double R[6] = {qr[0], qr[1], qr[2], qr[3], qr[4], pr};
  orc_q F;
  roe6(L, R, h, gamma, &F);
  out[0] = F.rho; out[1] = F.rhoU; out[2] = F.rhoV; out[3] = F.rhoW; out[4] = F.rhoE;
}

void orc_rusanov(const double ql[5], double pl, const double qr[5], double pr,
                 const double csih[3], double out[5]) {
  double L[6] = {ql[0], ql[1], ql[2], ql[3], ql[4], pl};
  double R[6] = {qr[0], qr[1], qr[2], qr[3], qr[4], pr};
  orc_q F;
  rusanov6(L, R, csih, ucon_of(L, csih), ucon_of(R, csih), &F);
  out[0] = F.rho; out[1] = F.rhoU; out[2] = F.rhoV; out[3] = F.rhoW; out[4] = F.rhoE;
}

static inline void load6(const orc_ctx *c, size_t n, double s[6]) {
  const orc_q *q = &c->lQ[n];
  s[0] = q->rho; s[1] = q->rhoU; s[2] = q->rhoV; s[3] = q->rhoW; s[4] = q->rhoE; s[5] = c->lP[n];
}
static inline size_t lstride(const orc_ctx *c, int d) {
  return d == 0 ? 1 : d == 1 ? (size_t)c->LX : (size_t)c->LX * c->LY;
}
static inline const orc_v3 *metric_of(const orc_ctx *c, int d) {
  return d == 0 ? c->lcsi : d == 1 ? c->leta : c->lzet;
}

static void recon_muscl(const orc_ctx *c, size_t nm, size_t n0, size_t n1, size_t n2, double *L, double *R) {
  double s0[6], s1[6], sm[6], sp[6];
  const double *nv = c->lnvert;
  load6(c, n0, s0); load6(c, n1, s1);
  if (nv[nm] > BLANKWALL) load6(c, n1, sm); else load6(c, nm, sm);
  if (nv[n2] > BLANKWALL) load6(c, n0, sp); else load6(c, n2, sp);
  for (int q = 0; q < 6; q++) {
    L[q] = orc_muscl_ql(sm[q], s0[q], s1[q]);
    R[q] = orc_muscl_qr(s0[q], s1[q], sp[q]);
  }
}

/* WENO3 with ENO fallback next to walls.  eno_left: use ENO for the left
 * state; eno_left_mask selects which of the 6 scalars (bit q) use ENO when
 * eno_left is false (the eta-face isothermal-wall case, rhs.c:2884-2891). */
static void recon_weno(const orc_ctx *c, size_t nm, size_t n0, size_t n1, size_t n2, int eno_left,
                       int eno_left_mask, double *L, double *R) {
  double sm[6], s0[6], s1[6], s2[6];
  const int order = c->p.eno_order;
  load6(c, nm, sm); load6(c, n0, s0); load6(c, n1, s1); load6(c, n2, s2);
  for (int q = 0; q < 6; q++) {
    if (eno_left || (eno_left_mask >> q & 1))
      L[q] = orc_eno_ql(s0[q], s1[q], s2[q], order);
    else
      L[q] = orc_weno_ql(sm[q], s0[q], s1[q]);
  }
  if (c->lnvert[n2] > BLANKWALL) {
    for (int q = 0; q < 6; q++) R[q] = orc_eno_qr(sm[q], s0[q], s1[q], order);
  } else {
    for (int q = 0; q < 6; q++) R[q] = orc_weno_qr(s0[q], s1[q], s2[q]);
  }
}

/* WENO5 on the six nodes nm2, nm, n0, n1, n2, n3 of a face (stride s) */
static void recon_weno5(const orc_ctx *c, size_t n0, size_t s, double *L, double *R) {
  double v[6][6];
  for (int t = 0; t < 6; t++) load6(c, n0 + (size_t)(t - 2) * s, v[t]);  /* t = node - n0 + 2 */
  for (int q = 0; q < 6; q++) {
    const double l[5] = {v[0][q], v[1][q], v[2][q], v[3][q], v[4][q]};
    const double r[5] = {v[5][q], v[4][q], v[3][q], v[2][q], v[1][q]};
    L[q] = orc_weno5_ql(l);
    R[q] = orc_weno5_ql(r);
  }
}
/* the WENO5 stencil of the face lies in the grid and holds no solid node */
static int weno5_fits(const orc_ctx *c, int id, int m, size_t n0, size_t s) {
  if (!c->p.weno5 || id < 2 || id > m - 4) return 0;
  const double *nv = c->lnvert;
  return !(nv[n0 - 2 * s] > BLANKWALL || nv[n0 - s] > BLANKWALL || nv[n0 + 2 * s] > BLANKWALL ||
           nv[n0 + 3 * s] > BLANKWALL);
}
static void riemann_flux(const orc_ctx *c, const double *L, const double *R, const double *h, double uconL,
                         double uconR, orc_q *F) {
  if (c->p.roe) roe6(L, R, h, c->p.gamma, F);
  else rusanov6(L, R, h, uconL, uconR, F);
}

/* One face of formConvection (rhs.c:2646-2814 xi, 2817-2971 eta, 2973-3121 zeta). */
static void face_conv(const orc_ctx *c, int d, int k, int j, int i, orc_q *F) {
  const orc_params *P = &c->p;
  const int *bc = P->bctype;
  const int m = d == 0 ? c->mx : d == 1 ? c->my : c->mz;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const size_t s = lstride(c, d);
  const size_t n0 = LI(c, k, j, i), n1 = n0 + s, nm = n0 - s, n2 = n0 + 2 * s;
  const double *nv = c->lnvert;
  const int iso = P->isothermal != 0;
  double L[6], R[6];
  if (d == 1 && j == 0 && iso && bc[2] == 1) {
    load6(c, n0, L); load6(c, n0, R);
  } else if (d == 1 && j == c->my - 2 && iso && bc[3] == 1) {
    load6(c, n1, L); load6(c, n1, R);
  } else if (P->fluxorder == 1 || id == 0 || id == m - 2 || (nv[n0] + nv[n1]) > BLANKWALL) {
    load6(c, n0, L); load6(c, n1, R);
  } else if (P->fluxorder == 2 || (d != 0 && j < 30 && k < 7)) {
    recon_muscl(c, nm, n0, n1, n2, L, R);
  } else if (weno5_fits(c, id, m, n0, s) && !(d == 1 && j == 1 && iso && bc[2] == 1)) {
    recon_weno5(c, n0, s, L, R);
  } else {
    int eno_left, mask = 0;
    if (d == 0) {
      eno_left = nv[nm] > BLANKWALL && (i + 2) < c->mx;
    } else if (d == 1) {
      eno_left = nv[nm] > BLANKWALL;
      if (!eno_left && j == 1 && iso && bc[2] == 1) mask = (1 << 0) | (1 << 4) | (1 << 5);
    } else {
      eno_left = nv[nm] > BLANKWALL;
    }
    recon_weno(c, nm, n0, n1, n2, eno_left, mask, L, R);
  }
  const orc_v3 *mt = metric_of(c, d);
  double h[3] = {0.5 * (mt[n0].x + mt[n1].x), 0.5 * (mt[n0].y + mt[n1].y), 0.5 * (mt[n0].z + mt[n1].z)};
  double uconL, uconR;
  if (id == m - 2 && bc[2 * d + 1] == 1)
    uconL = ucon_of(R, h);
  else
    uconL = ucon_of(L, h);
  if (id == 0 && bc[2 * d] == 1)
    uconR = uconL;
  else
    uconR = ucon_of(R, h);
  riemann_flux(c, L, R, h, uconL, uconR, F);
}

static void residual(const orc_ctx *c, orc_q *out, int les_forcing) {
  const size_t sx = 1, sy = (size_t)c->LX, sz = (size_t)c->LX * c->LY;
  const orc_q *f1 = c->fp1, *f2 = c->fp2, *f3 = c->fp3;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 1; i < c->mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        orc_q *o = &out[GI(c, k, j, i)];
        if (c->lnvert[n] < BLANK) {
          o->rho = f1[n].rho - f1[n - sx].rho + f2[n].rho - f2[n - sy].rho + f3[n].rho - f3[n - sz].rho;
          /* -twoD 1: no rhoU residual, rhs.c:3168-3173, 4215-4220 */
          if (c->p.twoD != 1)
            o->rhoU = f1[n].rhoU - f1[n - sx].rhoU + f2[n].rhoU - f2[n - sy].rhoU + f3[n].rhoU - f3[n - sz].rhoU;
          else
            o->rhoU = 0.;
          o->rhoV = f1[n].rhoV - f1[n - sx].rhoV + f2[n].rhoV - f2[n - sy].rhoV + f3[n].rhoV - f3[n - sz].rhoV;
          o->rhoW = f1[n].rhoW - f1[n - sx].rhoW + f2[n].rhoW - f2[n - sy].rhoW + f3[n].rhoW - f3[n - sz].rhoW;
          o->rhoE = f1[n].rhoE - f1[n - sx].rhoE + f2[n].rhoE - f2[n - sy].rhoE + f3[n].rhoE - f3[n - sz].rhoE;
          if (les_forcing) { /* channel forcing, rhs.c:4230-4235 (ratio==0 -> minus zero) */
            const orc_q *q = &c->lQ[n];
            o->rhoW = o->rhoW - c->p.ratio / c->laj[n] * q->rho / c->p.dt;
            o->rhoE = o->rhoE - c->p.ratio / c->laj[n] * q->rhoW / c->p.dt;
          }
        } else {
          o->rho = o->rhoU = o->rhoV = o->rhoW = o->rhoE = 0.;
        }
      }
}

/* -twoD 1 skips the xi faces (rhs.c:1487-1714, 2646-2814, 3465-3706): their
 * flux Vecs are fresh VecDuplicates, zero where no face writes them. */
static void zero_xi_faces(orc_ctx *c, orc_q *F, orc_v3 *al) {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 0; i < c->mx - 1; i++) {
        const size_t n = LI(c, k, j, i);
        F[n].rho = F[n].rhoU = F[n].rhoV = F[n].rhoW = F[n].rhoE = 0.;
        if (al) al[n].x = al[n].y = al[n].z = 0.;
      }
}

/* formConvection, rhs.c:2570-3224 (non-LES). Writes c->conv. */
void orc_form_convection(orc_ctx *c) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  if (c->p.twoD == 1) zero_xi_faces(c, c->fp1, NULL);
  else {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 0; i < mx - 1; i++) face_conv(c, 0, k, j, i, &c->fp1[LI(c, k, j, i)]);
  }
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) face_conv(c, 1, k, j, i, &c->fp2[LI(c, k, j, i)]);
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) face_conv(c, 2, k, j, i, &c->fp3[LI(c, k, j, i)]);
  residual(c, c->conv, 0);
}

/* Central flux of formConvectionLES for one face (rhs.c:3481-3559 xi,
 * 3731-3802 eta, 3964-4033 zeta). */
static void face_central(const orc_ctx *c, int d, int id, int m, size_t nm, size_t n0, size_t n1,
                         size_t n2, const double *h, orc_q *Fc) {
  const double *nv = c->lnvert, *p = c->lP;
  const orc_q *q = c->lQ;
  int four = 0;
  if (id > 0 && id < m - 2) {
    double sum = d == 0 ? nv[n0] + nv[n1] + nv[n2] + nv[nm] : nv[nm] + nv[n0] + nv[n1] + nv[n2];
    four = sum < 2.9;
  }
  if (four && (c->p.central == 2 || c->p.central == 3)) {
    const orc_v3 *mt = metric_of(c, d);
    double r[4], u[4], v[4], w[4], ep[4], pp[4], U[4];
    size_t nn[4] = {nm, n0, n1, n2};
    for (int a = 0; a < 4; a++) {
      const orc_q *Q = &q[nn[a]];
      r[a] = Q->rho; u[a] = Q->rhoU; v[a] = Q->rhoV; w[a] = Q->rhoW;
      pp[a] = p[nn[a]];
      ep[a] = d == 2 ? Q->rhoE + pp[a] : pp[a] + Q->rhoE;
      U[a] = 1. / Q->rho * (mt[nn[a]].x * Q->rhoU + mt[nn[a]].y * Q->rhoV + mt[nn[a]].z * Q->rhoW);
    }
    if (c->p.central == 2) {
      Fc->rho = orc_fp_2(r, U);
      Fc->rhoU = orc_fp_2(u, U) + orc_fp4(pp) * h[0];
      Fc->rhoV = orc_fp_2(v, U) + orc_fp4(pp) * h[1];
      Fc->rhoW = orc_fp_2(w, U) + orc_fp4(pp) * h[2];
      Fc->rhoE = orc_fp_2(ep, U);
    } else {
      Fc->rho = fp_1(r, U);
      Fc->rhoU = fp_1(u, U) + orc_fp4(pp) * h[0];
      Fc->rhoV = fp_1(v, U) + orc_fp4(pp) * h[1];
      Fc->rhoW = fp_1(w, U) + orc_fp4(pp) * h[2];
      Fc->rhoE = fp_1(ep, U);
    }
  } else {
    double rc = 0.5 * (q[n0].rho + q[n1].rho);
    double uc = 0.5 * (q[n0].rhoU + q[n1].rhoU);
    double vc = 0.5 * (q[n0].rhoV + q[n1].rhoV);
    double wc = 0.5 * (q[n0].rhoW + q[n1].rhoW);
    double ec = 0.5 * (q[n0].rhoE + q[n1].rhoE);
    double Uc = (h[0] * uc + h[1] * vc + h[2] * wc) / rc;
    Fc->rho = rc * Uc;
    Fc->rhoU = uc * Uc + 0.5 * (p[n0] + p[n1]) * h[0];
    Fc->rhoV = vc * Uc + 0.5 * (p[n0] + p[n1]) * h[1];
    Fc->rhoW = wc * Uc + 0.5 * (p[n0] + p[n1]) * h[2];
    Fc->rhoE = Uc * (ec + 0.5 * (p[n0] + p[n1]));
  }
}

/* One face of formConvectionLES: central + upwind blended by sigma. */
static void face_conv_les(const orc_ctx *c, int d, int k, int j, int i, orc_q *F) {
  const orc_params *P = &c->p;
  const int *bc = P->bctype;
  const int m = d == 0 ? c->mx : d == 1 ? c->my : c->mz;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const size_t s = lstride(c, d);
  const size_t n0 = LI(c, k, j, i), n1 = n0 + s, nm = n0 - s, n2 = n0 + 2 * s;
  const double *nv = c->lnvert;
  const orc_v3 *mt = metric_of(c, d);
  double h[3] = {0.5 * (mt[n0].x + mt[n1].x), 0.5 * (mt[n0].y + mt[n1].y), 0.5 * (mt[n0].z + mt[n1].z)};
  const orc_v3 *sg = &c->lSigma[n0];
  double sigma = d == 0 ? sg->x : d == 1 ? sg->y : sg->z;
  orc_q Fc, Fw;
  face_central(c, d, id, m, nm, n0, n1, n2, h, &Fc);
  /* zeta-face wall test reads nvert[k+1][j][1] (rhs.c:4046) */
  double nvpair = d == 2 ? nv[n0] + nv[LI(c, k + 1, j, 1)] : nv[n0] + nv[n1];
  double L[6], R[6];
  if (P->fluxorder == 1 || ((id == 0 || id == m - 2) && bc[2 * d] != 7) || nvpair > BLANKWALL) {
    load6(c, n0, L); load6(c, n1, R);
  } else if (P->fluxorder == 2) {
    recon_muscl(c, nm, n0, n1, n2, L, R);
  } else if (weno5_fits(c, id, m, n0, s)) {
    recon_weno5(c, n0, s, L, R);
  } else {
    int eno_left = d == 0 ? (nv[nm] > BLANKWALL && (i + 2) < c->mx) : nv[nm] > BLANKWALL;
    recon_weno(c, nm, n0, n1, n2, eno_left, 0, L, R);
  }
  riemann_flux(c, L, R, h, ucon_of(L, h), ucon_of(R, h), &Fw);
  F->rho = Fw.rho * (1.0 - sigma) + Fc.rho * sigma;
  F->rhoU = Fw.rhoU * (1.0 - sigma) + Fc.rhoU * sigma;
  F->rhoV = Fw.rhoV * (1.0 - sigma) + Fc.rhoV * sigma;
  F->rhoW = Fw.rhoW * (1.0 - sigma) + Fc.rhoW * sigma;
  F->rhoE = Fw.rhoE * (1.0 - sigma) + Fc.rhoE * sigma;
}

/* compute_ri / sigma_cent, rhs.c:2334-2357 */
static double compute_ri(double fR, double f0, double fL, double ks) {
  double dfL = fL - f0, dfR = f0 - fR;
  return (fabs(2 * dfL * dfR) + ks) / (dfL * dfL + dfR * dfR + ks);
}
static double sigma_cent(double fL, double f0, double fR, double fRR, double ks) {
  double rL = compute_ri(fL, f0, fR, ks);
  double rR = compute_ri(f0, fR, fRR, ks);
  double r = pmin(rR, rL), rd = 0.5, rmax = pmax(rd, r);
  return 0.5 + 0.5 * tanh(3.0 * (r - rd) / rmax) / tanh(3.0);
}

double orc_sigma_cent(double fL, double f0, double fR, double fRR, double ks) {
  return sigma_cent(fL, f0, fR, fRR, ks);
}

/* Sigma at istage 0, rhs.c:3327-3444 (parity for -calc_sigma: tanh from libm). */
static void form_sigma(orc_ctx *c) {
  const orc_params *P = &c->p;
  const int mx = c->mx, my = c->my, mz = c->mz;
  orc_v3 *sg = c->Sigma;
  for (size_t n = 0; n < c->nG; n++) { sg[n].x = P->s_x; sg[n].y = P->s_y; sg[n].z = P->s_z; }
  if (P->calc_sigma) {
    const double *p = c->lP;
    const double ks = P->sigma_thrsh;
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          orc_v3 *s = &sg[GI(c, k, j, i)];
          if (P->abc != 0 && k > P->abc) s->x = 0.;
          else if (i < mx - 4 && i > 2)
            s->x = sigma_cent(p[LI(c, k, j, i - 1)], p[LI(c, k, j, i)], p[LI(c, k, j, i + 1)], p[LI(c, k, j, i + 2)], ks);
          s->x = pmin(P->s_x, s->x);
          s->x = pmax(P->s_x_min, s->x);
        }
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          orc_v3 *s = &sg[GI(c, k, j, i)];
          if (P->abc != 0 && k > P->abc) s->z = 0.;
          else if (P->bctype[4] == 7 || (k < mz - 4 && k > 4))
            s->z = sigma_cent(p[LI(c, k - 1, j, i)], p[LI(c, k, j, i)], p[LI(c, k + 1, j, i)], p[LI(c, k + 2, j, i)], ks);
          s->z = pmin(P->s_z, s->z);
          s->z = pmax(P->s_z_min, s->z);
        }
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          orc_v3 *s = &sg[GI(c, k, j, i)];
          if (P->abc != 0 && k > P->abc) s->y = 0.;
          else if (j < my - 4 && j > 4)
            s->y = sigma_cent(p[LI(c, k, j - 1, i)], p[LI(c, k, j, i)], p[LI(c, k, j + 1, i)], p[LI(c, k, j + 2, i)], ks);
          s->y = pmin(P->s_y, s->y);
          s->y = pmax(P->s_y_min, s->y);
        }
    orc_g2l(c, (double *)sg, (double *)c->lSigma, 3);
    const orc_v3 *ls = c->lSigma;
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          orc_v3 *s = &sg[GI(c, k, j, i)];
#define LS(kk, jj, ii) ls[LI(c, kk, jj, ii)]
          s->x *= sqrt(LS(k, j - 1, i).x * LS(k, j, i).x) * sqrt(LS(k - 1, j, i).x * LS(k + 1, j, i).x) *
                  sqrt(LS(k, j, i - 1).x * LS(k, j, i + 1).x);
          s->y *= sqrt(LS(k, j - 1, i).y * LS(k, j, i).y) * sqrt(LS(k - 1, j, i).y * LS(k + 1, j, i).y) *
                  sqrt(LS(k, j, i - 1).y * LS(k, j, i + 1).y);
          s->z *= sqrt(LS(k, j - 1, i).z * LS(k, j, i).z) * sqrt(LS(k - 1, j, i).z * LS(k + 1, j, i).z) *
                  sqrt(LS(k, j, i - 1).z * LS(k, j, i + 1).z);
#undef LS
        }
  }
  orc_g2l(c, (double *)sg, (double *)c->lSigma, 3);
  {
    const orc_v3 *ls = c->lSigma;
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          orc_v3 *s = &sg[GI(c, k, j, i)];
          if (j < my - 4 && j > 4 &&
              ls[LI(c, k, j - 1, i)].y + ls[LI(c, k, j + 1, i)].y - 2. * ls[LI(c, k, j, i)].y < 0.0)
            s->y = ls[LI(c, k, j - 1, i)].y * ls[LI(c, k, j, i)].y * ls[LI(c, k, j, i)].y;
          if (k < mz - 4 && k > 4 &&
              ls[LI(c, k - 1, j, i)].z + ls[LI(c, k + 1, j, i)].z - 2. * ls[LI(c, k, j, i)].z < 0.0)
            s->z = ls[LI(c, k - 1, j, i)].z * ls[LI(c, k + 1, j, i)].z * ls[LI(c, k, j, i)].z;
          s->z = pmin(s->z, s->y);
          s->y = s->z;
        }
  }
  orc_g2l(c, (double *)sg, (double *)c->lSigma, 3);
}

/* formConvectionLES, rhs.c:3231-4291. Writes c->conv. */
void orc_form_convection_les(orc_ctx *c, int istage) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  if (istage == 0) form_sigma(c);
  if (c->p.twoD == 1) zero_xi_faces(c, c->fp1, NULL);
  else {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 0; i < mx - 1; i++) face_conv_les(c, 0, k, j, i, &c->fp1[LI(c, k, j, i)]);
  }
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) face_conv_les(c, 1, k, j, i, &c->fp2[LI(c, k, j, i)]);
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) face_conv_les(c, 2, k, j, i, &c->fp3[LI(c, k, j, i)]);
  residual(c, c->conv, 1);
}

/* CalcUCat, rhs.c:1322-1365: Cartesian velocity on owned nodes, then G2L. */
void orc_calc_ucat(orc_ctx *c) {
  if (!c->Ucat) c->Ucat = malloc(c->nG * sizeof(orc_v3));
  orc_v3 *ucat = c->Ucat;
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < c->mz; k++)
    for (int j = 0; j < c->my; j++)
      for (int i = 0; i < c->mx; i++) {
        const orc_q *q = &c->lQ[LI(c, k, j, i)];
        orc_v3 *u = &ucat[GI(c, k, j, i)];
        if (c->p.twoD != 1) u->x = q->rhoU / q->rho; else u->x = 0.; /* rhs.c:1349 */
        u->y = q->rhoV / q->rho;
        u->z = q->rhoW / q->rho;
      }
  orc_g2l(c, (const double *)ucat, (double *)c->lUcat, 3);
}

/* Do_averaging, les.c:1664-1747 (period 1 without -fish): Q_sum += Q and
 * P_sum += P (VecAXPY with alpha 1), then over every owned node the squares
 * and cross products of the global Ucat of the last CalcUCat. */
void orc_do_averaging(orc_ctx *c) {
  if (!c->Q_sum) {
    c->Q_sum = calloc(c->nG, sizeof(orc_q));
    c->P_sum = calloc(c->nG, sizeof(double));
    c->U2_sum = calloc(c->nG, sizeof(orc_v3));
    c->Ucross_sum = calloc(c->nG, sizeof(orc_v3));
  }
  if (!c->Ucat) c->Ucat = calloc(c->nG, sizeof(orc_v3)); /* no CalcUCat yet: Ucat is zero */
  for (size_t n = 0; n < c->nG; n++) {
    double *qs = (double *)&c->Q_sum[n];
    const double *q = (const double *)&c->Q[n];
    for (int v = 0; v < 5; v++) qs[v] = qs[v] + q[v];
    c->P_sum[n] = c->P_sum[n] + c->P[n];
  }
  for (size_t n = 0; n < c->nG; n++) {
    const orc_v3 u = c->Ucat[n];
    c->U2_sum[n].x += u.x * u.x;
    c->U2_sum[n].y += u.y * u.y;
    c->U2_sum[n].z += u.z * u.z;
    c->Ucross_sum[n].x += u.x * u.y; /* uv */
    c->Ucross_sum[n].y += u.y * u.z; /* vw */
    c->Ucross_sum[n].z += u.z * u.x; /* wu */
  }
}

void orc_get_averages(const orc_ctx *c, double *q_sum, double *p_sum, double *u2_sum, double *ucross_sum) {
  const size_t n = c->nG;
  if (q_sum) { if (c->Q_sum) memcpy(q_sum, c->Q_sum, n * sizeof(orc_q)); else memset(q_sum, 0, n * sizeof(orc_q)); }
  if (p_sum) { if (c->P_sum) memcpy(p_sum, c->P_sum, n * sizeof(double)); else memset(p_sum, 0, n * sizeof(double)); }
  if (u2_sum) { if (c->U2_sum) memcpy(u2_sum, c->U2_sum, n * sizeof(orc_v3)); else memset(u2_sum, 0, n * sizeof(orc_v3)); }
  if (ucross_sum) {
    if (c->Ucross_sum) memcpy(ucross_sum, c->Ucross_sum, n * sizeof(orc_v3));
    else memset(ucross_sum, 0, n * sizeof(orc_v3));
  }
}

/* ---------------------------- formViscous ------------------------------- */
static inline int ib_node(double nv) { return nv > 0.5 && nv < 7.; } /* solid, innerblank */
static inline double Tof(const orc_ctx *c, size_t n) { return c->lP[n] / c->lQ[n].rho; }

/* Cross derivative across direction t of the face (A,B):
 * one-sided when a neighbour row holds immersed-boundary nodes.
 * The pressure term of the xi-face zeta derivative repeats the centre node
 * in its fourth term (rhs.c:1544-1545); q8 selects that form. */
static void cross_deriv(const orc_ctx *c, size_t A, size_t B, size_t t, int q8, double du[3], double *dT) {
  const double *nv = c->lnvert;
  const orc_v3 *u = c->lUcat;
  if (ib_node(nv[A + t]) || ib_node(nv[B + t])) {
    du[0] = (u[B].x + u[A].x - u[B - t].x - u[A - t].x) * 0.5;
    du[1] = (u[B].y + u[A].y - u[B - t].y - u[A - t].y) * 0.5;
    du[2] = (u[B].z + u[A].z - u[B - t].z - u[A - t].z) * 0.5;
    if (q8)
      *dT = (Tof(c, B) + Tof(c, A) - Tof(c, B - t) - Tof(c, A)) * 0.5;
    else
      *dT = (Tof(c, B) + Tof(c, A) - Tof(c, B - t) - Tof(c, A - t)) * 0.5;
  } else if (ib_node(nv[A - t]) || ib_node(nv[B - t])) {
    du[0] = (u[B + t].x + u[A + t].x - u[B].x - u[A].x) * 0.5;
    du[1] = (u[B + t].y + u[A + t].y - u[B].y - u[A].y) * 0.5;
    du[2] = (u[B + t].z + u[A + t].z - u[B].z - u[A].z) * 0.5;
    *dT = (Tof(c, B + t) + Tof(c, A + t) - Tof(c, B) - Tof(c, A)) * 0.5;
  } else {
    du[0] = (u[B + t].x + u[A + t].x - u[B - t].x - u[A - t].x) * 0.25;
    du[1] = (u[B + t].y + u[A + t].y - u[B - t].y - u[A - t].y) * 0.25;
    du[2] = (u[B + t].z + u[A + t].z - u[B - t].z - u[A - t].z) * 0.25;
    *dT = (Tof(c, B + t) + Tof(c, A + t) - Tof(c, B - t) - Tof(c, A - t)) * 0.25;
  }
}

/* One viscous face (rhs.c:1487-1707 xi, 1719-1918 eta, 1926-2121 zeta). */
static void face_visc(const orc_ctx *c, int d, int k, int j, int i, orc_q *F, orc_v3 *alpha) {
  const orc_params *P = &c->p;
  const int *bc = P->bctype;
  const size_t s[3] = {1, (size_t)c->LX, (size_t)c->LX * c->LY};
  const size_t A = LI(c, k, j, i), B = A + s[d];
  const orc_v3 *u = c->lUcat;
  double dc[3], de[3], dz[3], dTc, dTe, dTz; /* d/dxi, d/deta, d/dzeta of (u,v,w) and T=p/rho */
  double *dd[3] = {dc, de, dz};
  double *dT[3] = {&dTc, &dTe, &dTz};
  dd[d][0] = u[B].x - u[A].x;
  dd[d][1] = u[B].y - u[A].y;
  dd[d][2] = u[B].z - u[A].z;
  *dT[d] = Tof(c, B) - Tof(c, A);
  for (int t = 0; t < 3; t++) {
    if (t == d) continue;
    cross_deriv(c, A, B, s[t], d == 0 && t == 2, dd[t], dT[t]);
  }
  const orc_v3 *cs = c->lcsi, *et = c->leta, *ze = c->lzet;
  double csi0 = 0.5 * (cs[A].x + cs[B].x), csi1 = 0.5 * (cs[A].y + cs[B].y), csi2 = 0.5 * (cs[A].z + cs[B].z);
  double eta0 = 0.5 * (et[A].x + et[B].x), eta1 = 0.5 * (et[A].y + et[B].y), eta2 = 0.5 * (et[A].z + et[B].z);
  double zet0 = 0.5 * (ze[A].x + ze[B].x), zet1 = 0.5 * (ze[A].y + ze[B].y), zet2 = 0.5 * (ze[A].z + ze[B].z);
  double m0, m1, m2, g11, g21, g31;
  if (d == 0) {
    m0 = csi0; m1 = csi1; m2 = csi2;
    g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2;
    g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2;
    g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2;
  } else if (d == 1) {
    m0 = eta0; m1 = eta1; m2 = eta2;
    g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2;
    g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2;
    g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2;
  } else {
    m0 = zet0; m1 = zet1; m2 = zet2;
    g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2;
    g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2;
    g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2;
  }
  const double dudc = dc[0], dvdc = dc[1], dwdc = dc[2];
  const double dude = de[0], dvde = de[1], dwde = de[2];
  const double dudz = dz[0], dvdz = dz[1], dwdz = dz[2];
  double r11 = dudc * csi0 + dude * eta0 + dudz * zet0;
  double r21 = dvdc * csi0 + dvde * eta0 + dvdz * zet0;
  double r31 = dwdc * csi0 + dwde * eta0 + dwdz * zet0;
  double r12 = dudc * csi1 + dude * eta1 + dudz * zet1;
  double r22 = dvdc * csi1 + dvde * eta1 + dvdz * zet1;
  double r32 = dwdc * csi1 + dwde * eta1 + dwdz * zet1;
  double r13 = dudc * csi2 + dude * eta2 + dudz * zet2;
  double r23 = dvdc * csi2 + dvde * eta2 + dvdz * zet2;
  double r33 = dwdc * csi2 + dwde * eta2 + dwdz * zet2;
  double ajc = 0.5 * (c->laj[A] + c->laj[B]);
  const orc_q *q = c->lQ;
  const double *p = c->lP;
  double nu = 1. / P->ren;
  /* LES viscosity averages rho at [i] and [i+1] on every face (rhs.c:1611,1835,2034) */
  if (P->les) nu = 1. / P->ren * 0.5 * (q[A].rho + q[A + 1].rho);
  if (P->sutherland) {
    const double C_sutherland = 110.4, Tref = P->Tref;
    double Tr = P->gamma * P->Ma * P->Ma * (p[A] + p[B]) / (q[A].rho + q[B].rho);
    if (Tr < 0)
      nu = 1. / P->ren;
    else
      nu = 1. / P->ren * (1 + C_sutherland / Tref) / (Tr + C_sutherland / Tref) * orc_pow15_cr(Tr);
  }
  double div = 2. / 3. * (dudc * csi0 + dude * eta0 + dudz * zet0 + dvdc * csi1 + dvde * eta1 + dvdz * zet1 +
                          dwdc * csi2 + dwde * eta2 + dwdz * zet2);
  double fu = (g11 * dudc + g21 * dude + g31 * dudz + r11 * m0 + r21 * m1 + r31 * m2) * ajc * (nu);
  double fv = (g11 * dvdc + g21 * dvde + g31 * dvdz + r12 * m0 + r22 * m1 + r32 * m2) * ajc * (nu);
  double fw = (g11 * dwdc + g21 * dwde + g31 * dwdz + r13 * m0 + r23 * m1 + r33 * m2) * ajc * (nu);
  if (d == 2) { fu = 0.0 + fu; fv = 0.0 + fv; fw = 0.0 + fw; } /* "+=" into a zeroed Vec, rhs.c:2048 */
  fu -= div * m0 * ajc * nu;
  fv -= div * m1 * ajc * nu;
  fw -= div * m2 * ajc * nu;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const int m = d == 0 ? c->mx : d == 1 ? c->my : c->mz;
  orc_v3 uh;
  if (id == 0 && bc[2 * d] == 1) uh = u[A];
  else if (id == m - 2 && bc[2 * d + 1] == 1) uh = u[B];
  else {
    uh.x = 0.5 * (u[B].x + u[A].x);
    uh.y = 0.5 * (u[B].y + u[A].y);
    uh.z = 0.5 * (u[B].z + u[A].z);
  }
  const double gamma = P->gamma, pr = P->Pr;
  double fe = (fu * uh.x + fv * uh.y + fw * uh.z) +
              (g11 * dTc + g21 * dTe + g31 * dTz) * ajc * (gamma * nu / pr / (gamma - 1.0));
  if (P->les) {
    const double CI = 0.06; /* rhs.c:13 */
    const orc_v3 *ls = c->lSigma;
    double sA = d == 0 ? ls[A].x : d == 1 ? ls[A].y : ls[A].z;
    double sB = d == 0 ? ls[B].x : d == 1 ? ls[B].y : ls[B].z;
    double nu_t = 0.5 * (c->lNu_t[A] * sA + c->lNu_t[B] * sB);
    double rho_c = 0.5 * (q[A].rho + q[B].rho);
    double sigma_c = 0.5 * (sA + sB);
    if ((bc[2 * d] == 1 && id == 0) || (bc[2 * d + 1] == 1 && id == m - 2)) nu_t = 0;
    double tx = (g11 * dudc + g21 * dude + g31 * dudz + r11 * m0 + r21 * m1 + r31 * m2 - div * m0) * ajc * (nu_t * rho_c);
    double ty = (g11 * dvdc + g21 * dvde + g31 * dvdz + r12 * m0 + r22 * m1 + r32 * m2 - div * m1) * ajc * (nu_t * rho_c);
    double tz = (g11 * dwdc + g21 * dwde + g31 * dwdz + r13 * m0 + r23 * m1 + r33 * m2 - div * m2) * ajc * (nu_t * rho_c);
    double S2 = .5 * ((r11 - div) * (r11 - div) + r12 * r12 + r13 * r13 + r21 * r21 + (r22 - div) * (r22 - div) +
                      r23 * r23 + r31 * r31 + r32 * r32 + (r33 - div) * (r33 - div));
    double filter = pow(1. / c->laj[A], 1. / 3.);
    double Ssx = -S2 * orc_sq(filter) * CI * 2.0 / 3.0 * m0 * rho_c * sigma_c;
    double Ssy = -S2 * orc_sq(filter) * CI * 2.0 / 3.0 * m1 * rho_c * sigma_c;
    double Ssz = -S2 * orc_sq(filter) * CI * 2.0 / 3.0 * m2 * rho_c * sigma_c;
    fu += tx + Ssx;
    fv += ty + Ssy;
    fw += tz + Ssz;
    alpha->x = tx + Ssx;
    alpha->y = ty + Ssy;
    alpha->z = tz + Ssz;
  }
  F->rho = 0.;
  F->rhoU = fu; F->rhoV = fv; F->rhoW = fw; F->rhoE = fe;
}

/* formViscous, rhs.c:1367-2254. Writes c->visc. */
void orc_form_viscous(orc_ctx *c) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  const orc_params *P = &c->p;
  if (P->twoD == 1) zero_xi_faces(c, c->fp1, c->al1);
  else {
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 0; i < mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        face_visc(c, 0, k, j, i, &c->fp1[n], &c->al1[n]);
      }
  }
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        face_visc(c, 1, k, j, i, &c->fp2[n], &c->al2[n]);
      }
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        face_visc(c, 2, k, j, i, &c->fp3[n], &c->al3[n]);
      }
  const size_t sx = 1, sy = (size_t)c->LX, sz = (size_t)c->LX * c->LY;
  const orc_q *f1 = c->fp1, *f2 = c->fp2, *f3 = c->fp3;
  const orc_v3 *a1 = c->al1, *a2 = c->al2, *a3 = c->al3;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t n = LI(c, k, j, i);
        orc_q *o = &c->visc[GI(c, k, j, i)];
        o->rho = 0.;
        if (P->twoD != 1) /* rhs.c:2135-2140 */
          o->rhoU = f1[n].rhoU - f1[n - sx].rhoU + f2[n].rhoU - f2[n - sy].rhoU + f3[n].rhoU - f3[n - sz].rhoU;
        else
          o->rhoU = 0.;
        o->rhoV = f1[n].rhoV - f1[n - sx].rhoV + f2[n].rhoV - f2[n - sy].rhoV + f3[n].rhoV - f3[n - sz].rhoV;
        o->rhoW = f1[n].rhoW - f1[n - sx].rhoW + f2[n].rhoW - f2[n - sy].rhoW + f3[n].rhoW - f3[n - sz].rhoW;
        o->rhoE = f1[n].rhoE - f1[n - sx].rhoE + f2[n].rhoE - f2[n - sy].rhoE + f3[n].rhoE - f3[n - sz].rhoE;
        if (P->les && j > 2 && j < my - 2 && c->lnvert[n] < 0.5) {
          /* LES energy term, rhs.c:2153-2197 */
          const orc_q *q = c->lQ;
          double rL = (q[n].rho + q[n - sx].rho) * 0.5, rR = (q[n + sx].rho + q[n].rho) * 0.5;
          double rD = (q[n].rho + q[n - sy].rho) * 0.5, rU = (q[n + sy].rho + q[n].rho) * 0.5;
          double rB = (q[n].rho + q[n - sz].rho) * 0.5, rF = (q[n + sz].rho + q[n].rho) * 0.5;
          double u0 = q[n].rhoU / q[n].rho, u1 = q[n].rhoV / q[n].rho, u2 = q[n].rhoW / q[n].rho;
          o->rhoE += u0 * (rR * a1[n].x - rL * a1[n - sx].x + rU * a2[n].x - rD * a2[n - sy].x + rF * a3[n].x -
                           rB * a3[n - sz].x) +
                     u1 * (rR * a1[n].y - rL * a1[n - sx].y + rU * a2[n].y - rD * a2[n - sy].y + rF * a3[n].y -
                           rB * a3[n - sz].y) +
                     u2 * (rR * a1[n].z - rL * a1[n - sx].z + rU * a2[n].z - rD * a2[n - sy].z + rF * a3[n].z -
                           rB * a3[n - sz].z);
        }
      }
}

/* formRHS + MultAj, rhs.c:4703-4736 and 4575-4701. */
void orc_form_rhs(orc_ctx *c, int istage, double *rhs_out) {
  const orc_params *P = &c->p;
  memset(c->conv, 0, c->nG * sizeof(orc_q));
  memset(c->visc, 0, c->nG * sizeof(orc_q));
  if (P->les)
    orc_form_convection_les(c, istage);
  else
    orc_form_convection(c);
  if (!P->invicid) {
    orc_calc_ucat(c);
    orc_form_viscous(c);
  }
  const double *cv = (const double *)c->conv, *vs = (const double *)c->visc;
  double *r = (double *)c->Rhs;
  #pragma omp parallel for schedule(static)
  for (size_t n = 0; n < 5 * c->nG; n++) r[n] = vs[n] - cv[n]; /* VecWAXPY(Rhs,-1,Conv,Visc) */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k < c->mz - 1; k++)
    for (int j = 1; j < c->my - 1; j++)
      for (int i = 1; i < c->mx - 1; i++) {
        orc_q *o = &c->Rhs[GI(c, k, j, i)];
        size_t n = LI(c, k, j, i);
        if (c->lnvert[n] < BLANK) {
          double a = c->laj[n];
          o->rho = o->rho * a; o->rhoU = o->rhoU * a; o->rhoV = o->rhoV * a;
          o->rhoW = o->rhoW * a; o->rhoE = o->rhoE * a;
        } else {
          o->rho = o->rhoU = o->rhoV = o->rhoW = o->rhoE = 0.;
        }
      }
  /* VecNorm(Rhs) and VecNorm(Q) right after formRHS, before the stage update
   * (solvers.c:249-252): |Q| is the norm of the stage input on every owned node */
  {
    double a = 0, b = 0;
    const double *q = (const double *)c->Q;
    for (size_t n = 0; n < 5 * c->nG; n++) {
      const double x = fabs(r[n]), y = fabs(q[n]);
      if (x > a) a = x;
      if (y > b) b = y;
    }
    c->nrm[0] = a;
    c->nrm[1] = b;
  }
  if (rhs_out) memcpy(rhs_out, c->Rhs, c->nG * sizeof(orc_q));
}

/* RungeKutta, solvers.c:223-294. */
void orc_runge_kutta(orc_ctx *c) {
  const orc_params *P = &c->p;
  double alfa[3] = {1., 1 / 4., 2 / 3.}, beta[3] = {1., 3 / 4., 1 / 3.};
  if (P->fluxorder < 3) { alfa[1] = 1 / 2.; beta[1] = 1 / 2.; }
  for (int istage = 0; istage < P->fluxorder; istage++) {
    orc_form_rhs(c, istage, NULL);
    double *q = (double *)c->Q;
    const double *qo = (const double *)c->Qold, *r = (const double *)c->Rhs;
    const size_t N = 5 * c->nG;
    if (istage == 0) {
      const double a = alfa[istage] * P->dt * P->st;
      #pragma omp parallel for schedule(static)
      for (size_t n = 0; n < N; n++) q[n] = a * r[n] + qo[n]; /* VecWAXPY */
    } else {
      const double al = beta[istage], be = alfa[istage] * P->dt * P->st, ga = alfa[istage];
      #pragma omp parallel for schedule(static)
      for (size_t n = 0; n < N; n++) q[n] = al * qo[n] + be * r[n] + ga * q[n]; /* VecAXPBYPCZ */
    }
    orc_form_bcs(c, alfa[istage]);
    orc_g2l(c, (double *)c->Q, (double *)c->lQ, 5);
    orc_eq_of_state(c);
  }
  if (c->immersed) { /* solvers.c:279-286 */
    orc_ibm_interpolation(c, 1);
    orc_form_bcs(c, 0.);
  }
}

void orc_flow_step(orc_ctx *c) {
  const orc_params *P = &c->p;
  c->ti += 1;
  if (P->les && !P->invicid && P->les != 10) orc_les_update(c);
  orc_runge_kutta(c);
  memcpy(c->Qold, c->Q, c->nG * sizeof(orc_q)); /* main.c:971 */
}

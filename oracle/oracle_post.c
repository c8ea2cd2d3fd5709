/*
 * oracle_post.c -- CPU restatement of the post-processor's Q-criterion
 * (data.c:4815-5053) with its eigenvalue routines jacobi (data.c:5057-5170)
 * and eigsrt (data.c:5172-5203).  Test infrastructure only: the parity
 * reference of the device kernel; the product library never includes, links
 * or calls it.  Arrays are in the global layout of cles_oracle.h (i fastest,
 * metric and state triples interleaved).  Compiled with -ffp-contract=off,
 * in the reference's operation order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cles_oracle.h"

/* jacobi (data.c:5057-5170) on a[n][n] followed by eigsrt (data.c:5172-5203),
 * the eigenvalue estimates d only (the eigenvector matrix v does not reach
 * them).  The loops are the reference's, bounds included: the updates of the
 * rows j < ip - 1 and ip + 1 <= j < iq - 1 are empty for n = 3, the test of
 * i > 4 is never true inside i < 5, and after five sweeps the last estimates
 * stand (the reference prints "too many iterations" and returns them).
 * abs() of a double there is the floating-point abs: the reference compiles
 * as C++ (variables.h includes <vector>), where stdlib.h declares abs(double).
 * Pinned in the tests by a diagonal matrix (sorted diagonal), matrices with
 * one off-diagonal pair (the exact eigenvalues of that 2x2 block), the trace
 * invariant, and the linear-field lambda2 closed forms. */
void orc_jacobi3(const double *ain, double *dout) {
  const int n = 3;
  double a[3][3], d[3], b[3], z2[3];
  int i, ip, iq, j;
  memcpy(a, ain, sizeof(a));
  for (ip = 0; ip < n; ip++) {
    b[ip] = a[ip][ip];
    d[ip] = b[ip];
    z2[ip] = 0.;
  }
  for (i = 0; i < 5; i++) {
    double sm = 0., tresh;
    for (ip = 0; ip < n - 1; ip++)
      for (iq = ip + 1; iq < n; iq++) sm = sm + fabs(a[ip][iq]);
    if (sm == 0.) break;
    if (i < 4)
      tresh = 0.2 * sm / (n * n);
    else
      tresh = 0.;
    for (ip = 0; ip < n - 1; ip++) {
      for (iq = ip + 1; iq < n; iq++) {
        double g = 100. * fabs(a[ip][iq]), h, t, c, s, tau, theta;
        if ((i > 4) && (fabs(d[ip]) + g == fabs(d[ip])) && (fabs(d[iq]) + g == fabs(d[iq]))) {
          a[ip][iq] = 0.;
        } else if (fabs(a[ip][iq]) > tresh) {
          h = d[iq] - d[ip];
          if (fabs(h) + g == fabs(h)) {
            t = a[ip][iq] / h;
          } else {
            theta = 0.5 * h / a[ip][iq];
            t = 1. / (fabs(theta) + sqrt(1. + theta * theta));
            if (theta < 0.) t = -t;
          }
          c = 1. / sqrt(1 + t * t);
          s = t * c;
          tau = s / (1. + c);
          h = t * a[ip][iq];
          z2[ip] = z2[ip] - h;
          z2[iq] = z2[iq] + h;
          d[ip] = d[ip] - h;
          d[iq] = d[iq] + h;
          a[ip][iq] = 0.;
          for (j = 0; j < ip - 1; j++) {
            g = a[j][ip];
            h = a[j][iq];
            a[j][ip] = g - s * (h + g * tau);
            a[j][iq] = h + s * (g - h * tau);
          }
          for (j = ip + 1; j < iq - 1; j++) {
            g = a[ip][j];
            h = a[j][iq];
            a[ip][j] = g - s * (h + g * tau);
            a[j][iq] = h + s * (g - h * tau);
          }
          for (j = iq + 1; j < n; j++) {
            g = a[ip][j];
            h = a[iq][j];
            a[ip][j] = g - s * (h + g * tau);
            a[iq][j] = h + s * (g - h * tau);
          }
        }
      }
    }
    for (ip = 0; ip < n; ip++) {
      b[ip] = b[ip] + z2[ip];
      d[ip] = b[ip];
      z2[ip] = 0.;
    }
  }
  /* eigsrt: selection into descending order */
  for (i = 0; i < n - 1; i++) {
    int k = i;
    double p = d[i];
    for (j = i + 1; j < n; j++)
      if (d[j] > p) {
        k = j;
        p = d[j];
      }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
    }
  }
  dout[0] = d[0];
  dout[1] = d[1];
  dout[2] = d[2];
}

/* QCriteria (data.c:4815-5053): u = rhoU / rho on every node; on the nodes
 * 1 .. m-2 of each direction, velocity differences in index space (one-sided
 * at node 1 or beside a node with nvert > 2 on the low side, at node m-2 or
 * beside one on the high side, else central), the face metrics averaged to
 * the node times aj, d_ab = du_a/dx_b, s = d + d^T, w = d - d^T,
 * Q = (wo - so) / 2 with so = |s|^2 terms and wo = |w|^2 terms as written,
 * and lambda2 the middle eigenvalue estimate of ss + ww, the products
 * "s^2" and "w^2" as the reference writes them (ss12 = s12 s11 +
 * s22 s12 s32 s13 and its like).  qc enters holding P: the reference writes Q
 * into P, so the boundary nodes keep the pressure; l2 is 0 there (Phi).
 * Pinned in the tests by the closed form Q = (|Omega|^2 - |S|^2) / 2 on
 * linear fields, the finite-difference choice on a quadratic field, and
 * 4 x the middle eigenvalue of S^2 + Omega^2 for fields in a plane. */
void orc_q_criteria(int mx, int my, int mz, const double *Q, const double *csi, const double *eta,
                    const double *zet, const double *aj, const double *nvert, double *qc, double *l2) {
  const long long sy = mx, sz = (long long)mx * my, N = sz * mz;
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)N);
  for (long long n = 0; n < N; n++) {
    u[3 * n] = Q[5 * n + 1] / Q[5 * n];
    u[3 * n + 1] = Q[5 * n + 2] / Q[5 * n];
    u[3 * n + 2] = Q[5 * n + 3] / Q[5 * n];
    l2[n] = 0.;
  }
#define U(k, j, i, c) u[3 * ((k) * sz + (j) * sy + (i)) + (c)]
#define NV(k, j, i) nvert[(k) * sz + (j) * sy + (i)]
#define M(a, k, j, i, c) a[3 * ((k) * sz + (j) * sy + (i)) + (c)]
  for (long long k = 1; k < mz - 1; k++)
    for (long long j = 1; j < my - 1; j++)
      for (long long i = 1; i < mx - 1; i++) {
        double uc[3], ue[3], uz[3], cs[3], et[3], ze[3];
        int c;
        const double ajn = aj[k * sz + j * sy + i];
        for (c = 0; c < 3; c++) {
          if (i == 1 || NV(k, j, i - 1) > 2.)
            uc[c] = (U(k, j, i + 1, c) - U(k, j, i, c));
          else if (i == mx - 2 || NV(k, j, i + 1) > 2.)
            uc[c] = (U(k, j, i, c) - U(k, j, i - 1, c));
          else
            uc[c] = (U(k, j, i + 1, c) - U(k, j, i - 1, c)) * 0.5;
          if (j == 1 || NV(k, j - 1, i) > 2)
            ue[c] = (U(k, j + 1, i, c) - U(k, j, i, c));
          else if (j == my - 2 || NV(k, j + 1, i) > 2)
            ue[c] = (U(k, j, i, c) - U(k, j - 1, i, c));
          else
            ue[c] = (U(k, j + 1, i, c) - U(k, j - 1, i, c)) * 0.5;
          if (k == 1 || NV(k - 1, j, i) > 2)
            uz[c] = (U(k + 1, j, i, c) - U(k, j, i, c));
          else if (k == mz - 2 || NV(k + 1, j, i) > 2)
            uz[c] = (U(k, j, i, c) - U(k - 1, j, i, c));
          else
            uz[c] = (U(k + 1, j, i, c) - U(k - 1, j, i, c)) * 0.5;
          cs[c] = (M(csi, k, j, i, c) + M(csi, k, j, i - 1, c)) * 0.5 * ajn;
          et[c] = (M(eta, k, j, i, c) + M(eta, k, j - 1, i, c)) * 0.5 * ajn;
          ze[c] = (M(zet, k, j, i, c) + M(zet, k - 1, j, i, c)) * 0.5 * ajn;
        }
        double dd[3][3];
        int r;
        for (r = 0; r < 3; r++)
          for (c = 0; c < 3; c++) dd[r][c] = uc[r] * cs[c] + ue[r] * et[c] + uz[r] * ze[c];
        const double s11 = dd[0][0] + dd[0][0], s12 = dd[0][1] + dd[1][0], s13 = dd[0][2] + dd[2][0];
        const double s21 = s12, s22 = dd[1][1] + dd[1][1], s23 = dd[1][2] + dd[2][1];
        const double s31 = s13, s32 = s23, s33 = dd[2][2] + dd[2][2];
        const double so = (s11 * s11 + s22 * s22 + s33 * s33) / 4. + (s12 * s12 + s13 * s13 + s23 * s23) / 2.;
        const double ss11 = s11 * s11 + s21 * s12 + s31 * s13;
        const double ss12 = s12 * s11 + s22 * s12 * s32 * s13;
        const double ss13 = s13 * s11 + s23 * s12 * s33 * s13;
        const double ss21 = s11 * s21 + s21 * s22 + s31 * s23;
        const double ss22 = s12 * s21 + s22 * s22 * s32 * s23;
        const double ss23 = s13 * s21 + s23 * s22 * s33 * s23;
        const double ss31 = s11 * s31 + s21 * s32 + s31 * s33;
        const double ss32 = s12 * s31 + s22 * s32 * s32 * s33;
        const double ss33 = s13 * s31 + s23 * s32 * s33 * s33;
        const double w11 = 0, w12 = dd[0][1] - dd[1][0], w13 = dd[0][2] - dd[2][0];
        const double w21 = -w12, w22 = 0., w23 = dd[1][2] - dd[2][1];
        const double w31 = -w13, w32 = -w23, w33 = 0.;
        const double wo = (w12 * w12 + w13 * w13 + w23 * w23) / 2.;
        const double ww11 = w11 * w11 + w21 * w12 + w31 * w13;
        const double ww12 = w12 * w11 + w22 * w12 * w32 * w13;
        const double ww13 = w13 * w11 + w23 * w12 * w33 * w13;
        const double ww21 = w11 * w21 + w21 * w22 + w31 * w23;
        const double ww22 = w12 * w21 + w22 * w22 * w32 * w23;
        const double ww23 = w13 * w21 + w23 * w22 * w33 * w23;
        const double ww31 = w11 * w31 + w21 * w32 + w31 * w33;
        const double ww32 = w12 * w31 + w22 * w32 * w32 * w33;
        const double ww33 = w13 * w31 + w23 * w32 * w33 * w33;
        const double a[9] = {ss11 + ww11, ss12 + ww12, ss13 + ww13, ss21 + ww21, ss22 + ww22,
                             ss23 + ww23, ss31 + ww31, ss32 + ww32, ss33 + ww33};
        double d[3];
        orc_jacobi3(a, d);
        l2[k * sz + j * sy + i] = d[1];
        qc[k * sz + j * sy + i] = (wo - so) / 2.;
      }
#undef U
#undef NV
#undef M
  free(u);
}

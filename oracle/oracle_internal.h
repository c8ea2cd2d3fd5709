/* Internal definitions of the CPU restatement (test infrastructure only). */
#ifndef CLES_ORACLE_INTERNAL_H
#define CLES_ORACLE_INTERNAL_H
#include "cles_oracle.h"
#include <math.h>
#include <stddef.h>

struct orc_ctx {
  orc_params p;
  int mx, my, mz;
  int gx, gy, gz;        /* ghost width per direction: 2 if periodic else 0 */
  int LX, LY, LZ;        /* local (ghosted) extents */
  size_t nG, nL;
  orc_v3 *coor;          /* global coordinates */
  orc_v3 *lcsi, *leta, *lzet;
  double *laj;
  double *nvert, *lnvert;
  orc_q *Q, *Qold, *Rhs, *lQ;
  double *P, *lP;
  orc_v3 *Sigma, *lSigma;
  double *CS, *lCS, *Nu_t, *lNu_t;
  orc_v3 *lUcat;
  orc_v3 *Ucat;          /* global Ucat of the last CalcUCat (kept for Do_averaging) */
  /* Do_averaging running sums (les.c:1664-1747), allocated on first use */
  orc_q *Q_sum;
  double *P_sum;
  orc_v3 *U2_sum, *Ucross_sum;
  int ti;
  double nrm[2];         /* |Rhs|inf, |Q|inf of the last formRHS and its input */
  /* per-call temporaries the reference obtains with VecDuplicate */
  orc_q *fp1, *fp2, *fp3, *conv, *visc;
  orc_v3 *al1, *al2, *al3;
  double *dbg_lm, *dbg_mm; /* caller-owned copies of LM/MM for tests */
  /* immersed body (one surface, ibi = 0) */
  int immersed;
  int n_v, n_elmt;
  double *x_bp, *y_bp, *z_bp, *nf_x, *nf_y, *nf_z, *radvec;
  int *nv1, *nv2, *nv3;
  orc_v3 *qvec, *u_bp, *uold_bp;
  orc_ibm_info *ibl;     /* interpolation list in traversal order */
  int n_ibl;
};

#define GI(c, k, j, i) ((((size_t)(k)) * (size_t)(c)->my + (size_t)(j)) * (size_t)(c)->mx + (size_t)(i))
#define LI(c, k, j, i)                                                          \
  ((((size_t)((k) + (c)->gz)) * (size_t)(c)->LY + (size_t)((j) + (c)->gy)) *   \
       (size_t)(c)->LX +                                                        \
   (size_t)((i) + (c)->gx))

void orc_g2l(const orc_ctx *c, const double *g, double *l, int nc);
void orc_l2l(const orc_ctx *c, double *l, int nc);

/* pow(x, 1.5) and pow(x, 2.) of the reference, evaluated as their exact
 * values rounded once (the C library's pow misrounds x^1.5 and x^2 in about
 * 0.1% of arguments; the exactly rounded value is the one both this
 * restatement and the device path use).  x^1.5 = x*s + x*(sqrt(x) - s) with
 * s = RN(sqrt(x)): fma gives x - s*s and the low half of x*s exactly, so the
 * sum carries ~104 bits before its single rounding. */
static inline double orc_pow15_cr(double x) {
  if (x == 0.0) return 0.0;
  const double s = sqrt(x);
  const double r = fma(-s, s, x);
  const double hi = x * s;
  const double lo = fma(x, s, -hi);
  const double corr = x * (r / (2.0 * s));
  return hi + (lo + corr);
}
static inline double orc_sq(double x) { return x * x; }

/* PetscMin / PetscMax as macros in the reference (petscmath.h) */
static inline double pmin(double a, double b) { return a < b ? a : b; }
static inline double pmax(double a, double b) { return a < b ? b : a; }

void orc_form_convection(orc_ctx *c);           /* rhs.c:2570 */
void orc_form_convection_les(orc_ctx *c, int istage); /* rhs.c:3231 */
void orc_calc_ucat(orc_ctx *c);                 /* rhs.c:1322 */
void orc_form_viscous(orc_ctx *c);              /* rhs.c:1367 */

#endif

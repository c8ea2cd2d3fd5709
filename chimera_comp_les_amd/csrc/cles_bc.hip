// cles_bc.hip -- boundary nodes of a stage: the stage update with a zero
// right-hand side (RungeKutta applies VecWAXPY / VecAXPBYPCZ to every owned
// node, solvers.c:256-271, and MultAj leaves boundary rows at zero), the
// boundary conditions of FormBCS (rhs.c:477-1316) and the equation of state
// on the boundary nodes (rhs.c:4853-4865).  Interior nodes are written by the
// stage kernel.
//
// Periodic seams (rhs.c:1107-1227) copy the *stage-input* state at the
// wrapped node: the reference reads its local vector lQ, which still holds
// the stage input when FormBCS runs.  The i, j and k passes of the reference
// overwrite each other on edges and corners in that order; each boundary node
// here takes the value of the last pass that covers it.
#include <cstring>

#include "cles_internal.hpp"
#include "cles_seam.hpp"

namespace cles {

namespace {

struct BndArgs {
  Grid g;
  const double *q, *p, *qold;  // stage input (lQ, lP) and Qold
  double *qo, *po;             // stage output (Q, P)
  const double *aj, *eta, *zet, *nv, *coor;
  int bc[6];
  int iso, twoD, has_ib;
  double gamma, Ma, Twall, dt, alfa;
  double ren;  // Reynolds number (bctype 47 inflow)
  int ti;      // time step of the stage (bctype 47 inflow)
  double a_rhs, a_old, a_cur;
  int istage;
  int periodic;  // any periodic direction
  unsigned long long *norms;
  long long nb;  // number of boundary nodes
};

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// stage update with a zero right-hand side on every boundary node, plus the
// boundary share of |Q_in|_inf
__global__ __launch_bounds__(256) void bnd_update_kernel(BndArgs a) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid &g = a.g;
  const long long np = g.np;
  double mq = 0.;
  if (b < a.nb) {
    int k, j, i;
    bnd_decode(g, b, k, j, i);
    const long long n = nid(g, k, j, i);
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const double qi = a.q[n + q * np];
      mq = fmax(mq, fabs(qi));
      a.qo[n + q * np] = a.istage == 0 ? a.a_rhs * 0.0 + a.qold[n + q * np]
                                       : a.a_old * a.qold[n + q * np] + a.a_rhs * 0.0 + a.a_cur * qi;
    }
  }
  if (a.norms) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mq = fmax(mq, __shfl_xor(mq, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&a.norms[64 + (blockIdx.x & 63)], dbits(mq));
  }
}

// accessors: O = stage output (global Q), I = stage input (lQ), P = lP
struct Acc {
  const BndArgs &a;
  __device__ double &O(int k, int j, int i, int c) const { return a.qo[nid(a.g, k, j, i) + c * a.g.np]; }
  __device__ double I(int k, int j, int i, int c) const { return a.q[nid(a.g, k, j, i) + c * a.g.np]; }
  __device__ double OLD(int k, int j, int i, int c) const { return a.qold[nid(a.g, k, j, i) + c * a.g.np]; }
  __device__ double P(int k, int j, int i) const { return a.p[nid(a.g, k, j, i)]; }
  __device__ double AJ(int k, int j, int i) const { return a.aj[nid(a.g, k, j, i)]; }
  __device__ double ETA(int k, int j, int i, int c) const { return a.eta[nid(a.g, k, j, i) + c * a.g.np]; }
  __device__ double ZET(int k, int j, int i, int c) const { return a.zet[nid(a.g, k, j, i) + c * a.g.np]; }
  __device__ double NV(int k, int j, int i) const { return a.has_ib ? a.nv[nid(a.g, k, j, i)] : 0.0; }
};
enum { R = 0, U = 1, V = 2, W = 3, E = 4 };

// i boundaries, rhs.c:600-637
__global__ __launch_bounds__(256) void bc_i_kernel(BndArgs a) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.my * g.mz) return;
  const int j = (int)(t % g.my), k = (int)(t / g.my);
  Acc A{a};
  int i = 0;
  if (a.bc[0] == 3) {
    A.O(k, j, i, R) = A.O(k, j, i + 1, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = A.O(k, j, i + 1, V);
    A.O(k, j, i, W) = A.O(k, j, i + 1, W);
    A.O(k, j, i, E) = A.O(k, j, i + 1, E);
  } else if (a.bc[0] == 51) {
    A.O(k, j, i, R) = 0.445; A.O(k, j, i, U) = 0.311; A.O(k, j, i, V) = 0.; A.O(k, j, i, W) = 0.;
    A.O(k, j, i, E) = 8.928;
  }
  i = g.mx - 1;
  if (a.bc[1] == 3) {
    A.O(k, j, i, R) = A.O(k, j, i - 1, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = A.O(k, j, i - 1, V);
    A.O(k, j, i, W) = A.O(k, j, i - 1, W);
    A.O(k, j, i, E) = A.O(k, j, i - 1, E);
  } else if (a.bc[0] == 51) {  // the reference tests bctype[0] here (rhs.c:627)
    A.O(k, j, i, R) = 0.5; A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.; A.O(k, j, i, W) = 0.;
    A.O(k, j, i, E) = 1.4275;
  }
}

// j boundaries, rhs.c:640-851.  lP is modified in place by the reference
// only at the wall node being set and only read there within this pass, so
// the modified value lives in a register (lpw).
__global__ __launch_bounds__(256) void bc_j_kernel(BndArgs a) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.mx * g.mz) return;
  const int i = (int)(t % g.mx), k = (int)(t / g.mx);
  Acc A{a};
  const double gam = a.gamma, Ma = a.Ma, Twall = a.Twall, dt = a.dt, alfa = a.alfa;
  int j = 0;
  if (a.bc[2] == 3) {
    A.O(k, j, i, R) = A.O(k, j + 1, i, R);
    A.O(k, j, i, U) = A.O(k, j + 1, i, U);
    A.O(k, j, i, V) = 0.;
    A.O(k, j, i, W) = A.O(k, j + 1, i, W);
    A.O(k, j, i, E) = A.O(k, j + 1, i, E);
  } else if (a.bc[2] == 10) {
    A.O(k, j, i, R) = A.O(k, j + 1, i, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = A.O(k, j + 1, i, V);
    A.O(k, j, i, W) = A.O(k, j + 1, i, W);
    A.O(k, j, i, E) = A.O(k, j + 1, i, E);
  } else if (a.bc[2] == 1) {
    A.O(k, j, i, U) = 0.0; A.O(k, j, i, V) = 0.0; A.O(k, j, i, W) = 0.0;
    if (a.iso == 2 && alfa != 0 && A.NV(k, j, i) < 1.5) {
      // characteristic isothermal no-slip wall (Poinsot-Lele), rhs.c:692-730
      const double a_sound = sqrt(Twall / (Ma * Ma));
      const double beta = sqrt(A.ETA(k, j, i, 1) * A.ETA(k, j, i, 1) + A.ETA(k, j, i, 0) * A.ETA(k, j, i, 0) +
                               A.ETA(k, j, i, 2) * A.ETA(k, j, i, 2));
      const double Lambda0 = -a_sound * beta;
      if (!(k == 0 || k == g.mz - 1)) {
        double L0 = Lambda0 * A.AJ(k, j, i) *
                    ((-0.5 * A.P(k, j + 2, i) + 2 * A.P(k, j + 1, i) - 1.5 * A.P(k, j, i)) -
                     A.I(k, j, i, R) * a_sound / beta *
                         (A.ETA(k, j, i, 1) * (-0.5 * A.I(k, j + 2, i, V) / A.I(k, j + 2, i, R) +
                                               2. * A.I(k, j + 1, i, V) / A.I(k, j + 1, i, R) -
                                               1.5 * A.I(k, j, i, V) / A.I(k, j, i, R)) +
                          A.ETA(k, j, i, 2) * (-0.5 * A.I(k, j + 2, i, W) / A.I(k, j + 2, i, R) +
                                               2. * A.I(k, j + 1, i, W) / A.I(k, j + 1, i, R) -
                                               1.5 * A.I(k, j, i, W) / A.I(k, j, i, R))));
        if (a.twoD != 1)
          L0 -= Lambda0 * A.AJ(k, j, i) * A.I(k, j, i, R) * a_sound / beta * A.ETA(k, j, i, 0) *
                (-0.5 * A.I(k, j + 2, i, U) / A.I(k, j + 2, i, R) + 2. * A.I(k, j + 1, i, U) / A.I(k, j + 1, i, R) -
                 1.5 * A.I(k, j, i, U) / A.I(k, j, i, R));
        A.O(k, j, i, R) = A.OLD(k, j, i, R) - dt * alfa * L0 / a_sound / a_sound;
      }
    } else {
      const double lpw = A.P(k, j + 1, i);
      A.O(k, j, i, R) = lpw * (gam * Ma * Ma);
      A.O(k, j, i, E) = lpw / (gam - 1);
    }
  }
  j = g.my - 1;
  if (a.bc[3] == 3) {
    A.O(k, j, i, R) = A.O(k, j - 1, i, R);
    A.O(k, j, i, U) = A.O(k, j - 1, i, U);
    A.O(k, j, i, V) = 0.;
    A.O(k, j, i, W) = A.O(k, j - 1, i, W);
    A.O(k, j, i, E) = A.O(k, j - 1, i, E);
  } else if (a.bc[3] == 10) {
    A.O(k, j, i, R) = A.O(k, j - 1, i, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = A.O(k, j - 1, i, V);
    A.O(k, j, i, W) = A.O(k, j - 1, i, W);
    A.O(k, j, i, E) = A.O(k, j - 1, i, E);
  } else if (a.bc[3] == 1) {
    A.O(k, j, i, U) = 0.0; A.O(k, j, i, V) = 0.0; A.O(k, j, i, W) = 0.0;
    double lpw;
    if (a.iso == 2) {  // rhs.c:825-838
      const double a_sound = sqrt(gam * A.P(k, j - 1, i) / A.I(k, j - 1, i, R));
      const double Lambda0 = A.I(k, j - 1, i, V) / A.I(k, j - 1, i, R) - a_sound;
      double L0;
      if (k == 0 || k == g.mz - 1 || Lambda0 > 0.)
        L0 = 0.;
      else
        L0 = Lambda0 * A.AJ(k, j - 1, i) * A.ETA(k, j - 1, i, 1) *
             ((A.P(k, j - 1, i) - A.P(k, j - 2, i)) -
              A.I(k, j - 1, i, R) * a_sound * (A.I(k, j, i, V) / A.I(k, j, i, R) - A.I(k, j - 1, i, V) / A.I(k, j - 1, i, R)));
      A.O(k, j, i, R) = A.OLD(k, j, i, R) - dt * alfa * L0 / a_sound / a_sound;
      lpw = A.O(k, j, i, R) * Twall / (gam * Ma * Ma);
    } else if (a.iso == 1) {
      A.O(k, j, i, R) = A.P(k, j, i) * (gam * Ma * Ma);
      lpw = A.P(k, j - 1, i);
    } else {
      A.O(k, j, i, R) = A.O(k, j - 1, i, R);
      lpw = A.P(k, j - 1, i);
    }
    A.O(k, j, i, E) = lpw / (gam - 1.0);
  }
}

// k boundaries, rhs.c:855-1088, for j in [ls, le)
__global__ __launch_bounds__(256) void bc_k_kernel(BndArgs a, int ls, int le) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.mx * (le - ls)) return;
  const int i = (int)(t % g.mx), j = ls + (int)(t / g.mx);
  Acc A{a};
  const double gam = a.gamma, Ma = a.Ma, dt = a.dt, alfa = a.alfa;
  const double pi = 3.14159265;
  const double u = Ma, rho = gam;  // rhs.c:855
  int k = 0;
  if (a.bc[4] == 3) {
    A.O(k, j, i, R) = A.O(k + 1, j, i, R);
    A.O(k, j, i, U) = A.O(k + 1, j, i, U);
    A.O(k, j, i, V) = A.O(k + 1, j, i, V);
    A.O(k, j, i, W) = 0.;
    A.O(k, j, i, E) = A.O(k + 1, j, i, E);
  } else if (a.bc[4] == 52) {  // subsonic NRBC inlet with a cosine profile, rhs.c:868-889
    const double y = a.coor[nid(g, k, j, i) + g.np];
    const double AA = cos(pi / 2 * (y - 1.)) * cos(pi / 2 * (y - 1.));
    const double a_sound = sqrt(gam * A.P(k + 1, j, i) / A.I(k + 1, j, i, R));
    const double Lambda0 = A.I(k + 1, j, i, W) / A.I(k + 1, j, i, R) - a_sound;
    const double L0 = Lambda0 * A.AJ(k + 1, j, i) * A.ZET(k + 1, j, i, 2) *
                      ((A.P(k + 2, j, i) - A.P(k + 1, j, i)) -
                       A.I(k + 1, j, i, R) * a_sound * (A.I(k + 1, j, i, W) / A.I(k + 1, j, i, R) - AA));
    const double L1 = (gam - 1.) * L0, L4 = L0;
    const double d0 = 1. / (a_sound * a_sound) * (L1 + 0.5 * (L4 + L0));
    A.O(k, j, i, R) = A.OLD(k, j, i, R) - dt * alfa * d0;
    const double lpw = A.O(k, j, i, R) / (gam * Ma * Ma);
    A.O(k, j, i, W) = A.O(k, j, i, R) * AA;
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = 0.;
    A.O(k, j, i, E) = 0.5 * A.O(k, j, i, W) * A.O(k, j, i, W) / A.O(k, j, i, R) + lpw / (gam - 1);
  } else if (a.bc[4] == 55) {
    A.O(k, j, i, R) = rho;
    A.O(k, j, i, W) = rho * u;
    A.O(k, j, i, E) = 0.5 * rho * u * u + 1. / (gam - 1.);
  } else if (a.bc[4] == 44) {
    A.O(k, j, i, R) = A.O(k + 1, j, i, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = 0.;
    A.O(k, j, i, W) = A.O(k + 1, j, i, W);
    A.O(k, j, i, E) = A.O(k + 1, j, i, E);
  } else if (a.bc[4] == 45) {
    A.O(k, j, i, R) = 1.;
    A.O(k, j, i, W) = 1.;
    A.O(k, j, i, E) = .5 + 1. / (Ma * Ma * gam * (gam - 1.));
  } else if (a.bc[4] == 51) {
    A.O(k, j, i, R) = 0.445; A.O(k, j, i, W) = 0.311; A.O(k, j, i, E) = 8.928;
  } else if (a.bc[4] == 10) {
    A.O(k, j, i, R) = 1.; A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.; A.O(k, j, i, W) = 1.;
    A.O(k, j, i, E) = 1. / 0.4 + 0.5;
  } else if (a.bc[4] == 47) {
    // power-law inflow profile plus travelling waves in x and time,
    // rhs.c:904-945 (with its own pi = 3.141592); below y = 8 rhoU keeps the
    // stage-updated value.  Operation order as written there.
    const long long n = nid(g, k, j, i);
    const double x = a.coor[n], yy = a.coor[n + g.np];
    const double pw = 3.141592;
    const double c1 = 0.03, c2 = -0.0032, c3 = 0.02, c4 = -0.01, y1 = 1.0, y2 = 1.5, y3 = 2.0;
    const double w0 = 0.1, w1 = 0.25, w2 = 0.125, w3 = 0.0625, h1 = 0.0, h2 = 0.1, h3 = 0.15;
    const double b0 = 0.5 * pw, b1 = 0.75 * pw, b2 = 0.5 * pw, b3 = 0.25 * pw;
    const double yplus = yy * 0.065 * a.ren;
    const int ti = a.ti;
    double rw;
    if (yy < 8.0) {
      rw = pow(yy / 8.0, 0.1428);
      rw += c1 * yplus * exp(-yplus / 12.5) * sin(w0 * ti * dt) * cos(b0 * x);
      A.O(k, j, i, V) = c2 * yplus * yplus * exp(-yplus * yplus / 12.5 / 12.5) * sin(w0 * ti * dt) * cos(b0 * x);
    } else {
      rw = 1.0;
      A.O(k, j, i, U) = 0.0;
      rw += c3 * yy / y1 * exp(-yy / y1) * sin(w1 * ti * dt) * cos(b1 * x + h1) +
            c3 * yy / y2 * exp(-yy / y2) * sin(w2 * ti * dt) * cos(b2 * x + h2) +
            c3 * yy / y3 * exp(-yy / y3) * sin(w3 * ti * dt) * cos(b3 * x + h3);
      A.O(k, j, i, V) = c4 * yy * yy / y1 / y1 * exp(-yy * yy / y1 / y1) * sin(w1 * ti * dt) * cos(b1 * x + h1) +
                        c4 * yy * yy / y2 / y2 * exp(-yy * yy / y2 / y2) * sin(w2 * ti * dt) * cos(b2 * x + h2) +
                        c4 * yy * yy / y3 / y3 * exp(-yy * yy / y3 / y3) * sin(w3 * ti * dt) * cos(b3 * x + h3);
    }
    A.O(k, j, i, W) = rw;
    A.O(k, j, i, R) = 1.0;
    A.O(k, j, i, E) = 0.5 * rw * rw + 1. / (Ma * Ma * gam * (gam - 1.));
  }
  k = g.mz - 1;
  if (a.bc[5] == 3) {
    A.O(k, j, i, R) = A.O(k - 1, j, i, R);
    A.O(k, j, i, U) = A.O(k - 1, j, i, U);
    A.O(k, j, i, V) = A.O(k - 1, j, i, V);
    A.O(k, j, i, W) = 0.;
    A.O(k, j, i, E) = A.O(k - 1, j, i, E);
  } else if (a.bc[5] == 44) {
    A.O(k, j, i, R) = A.O(k - 1, j, i, R);
    A.O(k, j, i, U) = 0.;
    A.O(k, j, i, V) = 0.;
    A.O(k, j, i, W) = A.O(k - 1, j, i, W);
    A.O(k, j, i, E) = A.O(k - 1, j, i, E);
  } else if (a.bc[5] == 45) {
#pragma unroll
    for (int c = 0; c < 5; c++) A.O(k, j, i, c) = A.O(k - 1, j, i, c);
  } else if (a.bc[5] == 49) {  // Riemann-invariant outflow, rhs.c:970-1020
    double Vn = A.O(k - 1, j, i, W) / A.O(k - 1, j, i, R);
    double a_sound = sqrt(gam * A.P(k - 1, j, i) / A.O(k - 1, j, i, R));
    const double lambda1 = Vn + a_sound, lambda2 = Vn - a_sound;
    double Rplus, Rminus, s;
    if (Vn < 0.) {
      if (lambda1 < 0) {
        A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.;
        Rplus = Vn + 2 / Ma / (gam - 1.);
        Rminus = Vn - 2 / Ma / (gam - 1.);
        s = 1. / (Ma * Ma * gam);
      } else {
        Rplus = Vn + 2 * a_sound / (gam - 1.);
        Rminus = Vn - 2 / Ma / (gam - 1.);
        A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.;
        s = 1. / (Ma * Ma * gam);
      }
    } else {
      if (lambda2 >= 0) {
        Rplus = Vn + 2 * a_sound / (gam - 1.);
        Rminus = Vn - 2 * a_sound / (gam - 1.);
        A.O(k, j, i, U) = A.O(k - 1, j, i, U);
        A.O(k, j, i, V) = A.O(k - 1, j, i, R);  // rho, as in rhs.c:996
        s = A.P(k - 1, j, i) / pow(A.O(k - 1, j, i, R), gam);
      } else {
        Rminus = Vn - 2 / Ma / (gam - 1.);
        Rplus = Vn + 2 * a_sound / (gam - 1.);
        A.O(k, j, i, U) = A.O(k - 1, j, i, U);
        A.O(k, j, i, V) = A.O(k - 1, j, i, V);
        s = A.P(k - 1, j, i) / pow(A.O(k - 1, j, i, R), gam);
      }
    }
    Vn = 0.5 * (Rplus + Rminus);
    a_sound = 0.25 * (gam - 1.) * (Rplus - Rminus);
    const double pOverrho = a_sound * a_sound / (gam);
    A.O(k, j, i, R) = pow(pOverrho / s, 1. / (gam - 2.));
    A.O(k, j, i, W) = A.O(k, j, i, R) * Vn;
    const double pp = pOverrho * A.O(k, j, i, R);
    A.O(k, j, i, E) = pp / (gam - 1.) + 0.5 * (A.O(k, j, i, U) * A.O(k, j, i, U) + A.O(k, j, i, V) * A.O(k, j, i, V) +
                                              A.O(k, j, i, W) * A.O(k, j, i, W)) * A.O(k, j, i, R);
  } else if (a.bc[5] == 51) {
    A.O(k, j, i, R) = 0.5; A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.; A.O(k, j, i, W) = 0.;
    A.O(k, j, i, E) = 1.4275;
  } else if (a.bc[5] == 10) {
    A.O(k, j, i, R) = 1.; A.O(k, j, i, U) = 0.; A.O(k, j, i, V) = 0.; A.O(k, j, i, W) = 1.;
    A.O(k, j, i, E) = 1 / 0.4 + 0.5;
  } else if (a.bc[5] == 8) {  // subsonic non-reflecting outflow, rhs.c:1042-1085
    const double a_sound = sqrt(gam * A.P(k - 1, j, i) / A.I(k - 1, j, i, R));
    const double AA = A.O(k - 1, j, i, W) / A.O(k - 1, j, i, R);
    const double Lam1 = AA, Lam4 = AA + a_sound;
    const double L0 = 0;
    const double L1 = Lam1 * A.AJ(k - 1, j, i) * A.ZET(k - 1, j, i, 2) *
                      (a_sound * a_sound * (A.I(k - 1, j, i, R) - A.I(k - 2, j, i, R)) - (A.P(k, j, i) - A.P(k - 1, j, i)));
    const double L4 = Lam4 * A.AJ(k - 1, j, i) * A.ZET(k - 1, j, i, 2) *
                      ((A.P(k - 1, j, i) - A.P(k - 2, j, i)) +
                       A.O(k - 1, j, i, R) * a_sound *
                           (A.I(k - 1, j, i, W) / A.O(k - 1, j, i, R) - A.I(k - 2, j, i, W) / A.O(k - 2, j, i, R)));
    const double d0 = 1. / (a_sound * a_sound) * (L1 + 0.5 * (L4 + L0));
    const double d1 = 0.5 * (L0 + L4);
    const double d2 = 0.5 / (A.O(k - 1, j, i, R) * a_sound) * (L4 - L0);
    A.O(k, j, i, R) = A.OLD(k, j, i, R) - dt * alfa * d0;
    A.O(k, j, i, E) = A.OLD(k, j, i, E) -
                      dt * alfa * (0.5 * d0 * (A.I(k - 1, j, i, W) / A.O(k - 1, j, i, R) * A.I(k - 1, j, i, W) / A.O(k - 1, j, i, R)) +
                                   d1 / (gam - 1.) + A.I(k - 1, j, i, W) * d2);
    A.O(k, j, i, W) = A.OLD(k, j, i, W) - dt * alfa * (A.I(k - 1, j, i, W) / A.O(k - 1, j, i, R) * d0 + A.O(k - 1, j, i, R) * d2);
  }
}

// periodic seams from the stage input and the equation of state on every
// boundary node
__device__ __forceinline__ void seam_eos_node(const BndArgs &a, long long b) {
  seam_copy_eos(a.g, a.q, a.qo, a.po, a.gamma, a.periodic, b, a.nb);
}

__global__ __launch_bounds__(256) void seam_eos_kernel(BndArgs a) {
  seam_eos_node(a, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// A fully periodic grid: every boundary node ends as the seam copy of the
// stage input and no boundary condition runs in between, so the stage update
// of bnd_update_kernel is dead there.  This one pass takes its boundary share
// of |Q_in|_inf (same nodes, same norm slots; the max is order-free) and then
// the seam copy and EOS of seam_eos_kernel.
__global__ __launch_bounds__(256) void seam_eos_norm_kernel(BndArgs a) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a.norms) {
    double mq = 0.;
    if (b < a.nb) {
      int k, j, i;
      bnd_decode(a.g, b, k, j, i);
      const long long n = nid(a.g, k, j, i);
#pragma unroll
      for (int q = 0; q < 5; q++) mq = fmax(mq, fabs(a.q[n + q * a.g.np]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mq = fmax(mq, __shfl_xor(mq, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&a.norms[64 + (blockIdx.x & 63)], dbits(mq));
  }
  seam_eos_node(a, b);
}

// batched seams + EOS over several grids: bstart[r] is the first workgroup of
// grid r (256 boundary nodes per workgroup)
__global__ __launch_bounds__(256) void seam_batch_kernel(const BndArgs *__restrict__ tab, const int *__restrict__ bstart,
                                                         int nreq) {
  const int b = blockIdx.x;
  int lo = 0, hi = nreq;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bstart[mid] <= b) lo = mid; else hi = mid;
  }
  seam_eos_node(tab[lo], (long long)(b - bstart[lo]) * 256 + threadIdx.x);
}

// periodic seams after the interpolation of the immersed body: Q and P both
// come from the snapshot taken before FormBCS(0) (the reference reads the
// ghost entries of lQ and lP, which FormBCS never writes), and no equation of
// state follows (solvers.c:283-285)
__global__ __launch_bounds__(256) void seam_post_kernel(BndArgs a) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid &g = a.g;
  const long long np = g.np;
  if (b >= a.nb) return;
  int k, j, i;
  bnd_decode(g, b, k, j, i);
  const long long s = seam_src(g, k, j, i);
  if (s < 0) return;
  const long long n = nid(g, k, j, i);
#pragma unroll
  for (int q = 0; q < 5; q++) a.qo[n + q * np] = a.q[s + q * np];
  a.po[n] = a.p[s];
}

struct EosArgs {
  const double *q;
  double *p;
  long long n, np;
  double gamma;
};
__global__ void eos_kernel(EosArgs a) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.n) return;
  const double r = a.q[n], ru = a.q[n + a.np], rv = a.q[n + 2 * a.np], rw = a.q[n + 3 * a.np], e = a.q[n + 4 * a.np];
  a.p[n] = (a.gamma - 1) * (e - 0.5 * (ru * ru + rv * rv + rw * rw) / r);
}

}  // namespace

bool supported_bc(int side, int b) {
  // 1 (no-slip wall) on any side: FormBCS has no branch for it outside j
  // (the nodes keep the stage update), the face fluxes treat it on every side
  // (rhs.c:1658-1683, 1867-1889, 2065-2094, 2779-2787, 3088-3096)
  if (b == 7 || b == 0 || b == 1) return true;
  if (side < 2) return b == 3 || b == 51;
  if (side < 4) return b == 3 || b == 10 || b == 1;
  if (side == 4) return b == 3 || b == 52 || b == 55 || b == 44 || b == 45 || b == 51 || b == 10 || b == 47;
  return b == 3 || b == 44 || b == 45 || b == 49 || b == 51 || b == 10 || b == 8;
}

int launch_boundary(cles_ctx *c, int in, int out, int istage, double alfa, const double coef[3]) {
  const cles_params &p = c->p;
  for (int q = 0; q < 6; q++)
    if (!supported_bc(q, p.bctype[q]))
      return set_error(c, CLES_EINVAL, "bctype[%d]=%d is not on this path", q, p.bctype[q]);
  BndArgs a;
  a.g = c->g;
  a.q = c->Q[in];
  a.p = c->P[in];
  a.qold = c->Q[c->cur];
  a.qo = c->Q[out];
  a.po = c->P[out];
  a.aj = c->aj; a.eta = c->eta; a.zet = c->zet; a.nv = c->nvert; a.coor = c->coor;
  for (int q = 0; q < 6; q++) a.bc[q] = p.bctype[q];
  a.iso = p.isothermal; a.twoD = p.twoD; a.has_ib = c->has_ib;
  a.gamma = p.gamma; a.Ma = p.Ma; a.Twall = p.Twall; a.dt = p.dt; a.alfa = alfa;
  a.ren = p.ren; a.ti = c->ti;
  a.a_rhs = coef[0]; a.a_old = coef[1]; a.a_cur = coef[2];
  a.istage = istage;
  a.periodic = c->g.px || c->g.py || c->g.pz;
  a.norms = c->dnorm;
  const long long mx = c->g.mx, my = c->g.my, mz = c->g.mz;
  a.nb = 2 * mx * my + 2 * mx * (mz - 2) + 2 * (my - 2) * (mz - 2);
  const int bs = 256;
  auto grid = [&](long long n) { return dim3((unsigned)((n + bs - 1) / bs)); };
  bool all7 = true;  // periodic on all six sides: no boundary-condition kernel runs
  for (int q = 0; q < 6; q++) all7 = all7 && p.bctype[q] == 7;
  if (all7) {
    hipLaunchKernelGGL(seam_eos_norm_kernel, grid(a.nb), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
    return CLES_OK;
  }
  hipLaunchKernelGGL(bnd_update_kernel, grid(a.nb), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  auto np7 = [&](int q) { return p.bctype[q] != 7 && p.bctype[q] != 0; };
  if (np7(0) || np7(1)) {
    hipLaunchKernelGGL(bc_i_kernel, grid(my * mz), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  if (np7(2) || np7(3)) {
    hipLaunchKernelGGL(bc_j_kernel, grid(mx * mz), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  if (np7(4) || np7(5)) {
    const int ls = p.bctype[2] == 1 ? 1 : 0, le = p.bctype[3] == 1 ? (int)my - 1 : (int)my;
    hipLaunchKernelGGL(bc_k_kernel, grid(mx * (le - ls)), dim3(bs), 0, c->stream, a, ls, le);
    CLES_HIP(c, hipGetLastError());
  }
  hipLaunchKernelGGL(seam_eos_kernel, grid(a.nb), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}

size_t seam_args_bytes() { return sizeof(BndArgs); }

// arguments of the seam + EOS pass of a fully periodic grid; returns the
// number of 256-node workgroups it needs
int seam_batch_args(cles_ctx *c, int in, int out, void *dst) {
  const cles_params &p = c->p;
  BndArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.q = c->Q[in];
  a.p = c->P[in];
  a.qold = c->Q[c->cur];
  a.qo = c->Q[out];
  a.po = c->P[out];
  a.aj = c->aj; a.eta = c->eta; a.zet = c->zet; a.nv = c->nvert; a.coor = c->coor;
  for (int q = 0; q < 6; q++) a.bc[q] = p.bctype[q];
  a.gamma = p.gamma; a.Ma = p.Ma; a.Twall = p.Twall; a.dt = p.dt;
  a.ren = p.ren; a.ti = c->ti;
  a.periodic = 1;
  const long long mx = c->g.mx, my = c->g.my, mz = c->g.mz;
  a.nb = 2 * mx * my + 2 * mx * (mz - 2) + 2 * (my - 2) * (mz - 2);
  memcpy(dst, &a, sizeof(a));
  return (int)((a.nb + 255) / 256);
}

hipError_t launch_seam_batch(const void *tab, const int *bstart, int nreq, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(seam_batch_kernel, dim3(nblocks), dim3(256), 0, s, (const BndArgs *)tab, bstart, nreq);
  return hipGetLastError();
}

int launch_boundary_post(cles_ctx *c, int buf) {
  const cles_params &p = c->p;
  const long long np = c->g.np;
  BndArgs a;
  a.g = c->g;
  a.q = c->qsnap;
  a.p = c->qsnap + 5 * np;
  a.qold = c->Q[c->cur];
  a.qo = c->Q[buf];
  a.po = c->P[buf];
  a.aj = c->aj; a.eta = c->eta; a.zet = c->zet; a.nv = c->nvert; a.coor = c->coor;
  for (int q = 0; q < 6; q++) a.bc[q] = p.bctype[q];
  a.iso = p.isothermal; a.twoD = p.twoD; a.has_ib = c->has_ib;
  a.gamma = p.gamma; a.Ma = p.Ma; a.Twall = p.Twall; a.dt = p.dt; a.alfa = 0.;
  a.ren = p.ren; a.ti = c->ti;
  a.a_rhs = 0.; a.a_old = 0.; a.a_cur = 0.;
  a.istage = 0;
  a.periodic = c->g.px || c->g.py || c->g.pz;
  a.norms = nullptr;
  const long long mx = c->g.mx, my = c->g.my, mz = c->g.mz;
  a.nb = 2 * mx * my + 2 * mx * (mz - 2) + 2 * (my - 2) * (mz - 2);
  const int bs = 256;
  auto grid = [&](long long n) { return dim3((unsigned)((n + bs - 1) / bs)); };
  auto np7 = [&](int q) { return p.bctype[q] != 7 && p.bctype[q] != 0; };
  if (np7(0) || np7(1)) {
    hipLaunchKernelGGL(bc_i_kernel, grid(my * mz), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  if (np7(2) || np7(3)) {
    hipLaunchKernelGGL(bc_j_kernel, grid(mx * mz), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  if (np7(4) || np7(5)) {
    const int ls = p.bctype[2] == 1 ? 1 : 0, le = p.bctype[3] == 1 ? (int)my - 1 : (int)my;
    hipLaunchKernelGGL(bc_k_kernel, grid(mx * (le - ls)), dim3(bs), 0, c->stream, a, ls, le);
    CLES_HIP(c, hipGetLastError());
  }
  if (a.periodic) {
    hipLaunchKernelGGL(seam_post_kernel, grid(a.nb), dim3(bs), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  return CLES_OK;
}

int launch_eos_all(cles_ctx *c, int buf) {
  EosArgs a;
  a.q = c->Q[buf];
  a.p = c->P[buf];
  a.n = (long long)c->g.mx * c->g.my * c->g.mz;
  a.np = c->g.np;
  a.gamma = c->p.gamma;
  const int bs = 256;
  hipLaunchKernelGGL(eos_kernel, dim3((unsigned)((a.n + bs - 1) / bs)), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}

}  // namespace cles

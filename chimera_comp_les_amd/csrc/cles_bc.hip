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
#include "cles_internal.hpp"

namespace cles {

namespace {

struct BndArgs {
  Grid g;
  const double *q, *p, *qold;
  double *qo, *po;
  double a_rhs, a_old, a_cur;
  int istage;
  double gamma;
  int bc_nonperiodic;  // non-periodic BC passes ran before this kernel
  unsigned long long *norms;
  long long nb;  // number of boundary nodes
};

// linear boundary index -> (k, j, i), every boundary node exactly once
__device__ __forceinline__ void bnd_decode(const Grid &g, long long b, int &k, int &j, int &i) {
  const long long mx = g.mx, my = g.my, mz = g.mz;
  const long long pk = mx * my;              // k = 0 and k = mz-1 planes
  const long long pj = mx * (mz - 2);        // j = 0 and j = my-1, 1<=k<=mz-2
  const long long pi = (my - 2) * (mz - 2);  // i = 0 and i = mx-1, interior j,k
  if (b < 2 * pk) {
    k = b < pk ? 0 : (int)(mz - 1);
    long long r = b % pk;
    j = (int)(r / mx);
    i = (int)(r % mx);
    return;
  }
  b -= 2 * pk;
  if (b < 2 * pj) {
    j = b < pj ? 0 : (int)(my - 1);
    long long r = b % pj;
    k = 1 + (int)(r / mx);
    i = (int)(r % mx);
    return;
  }
  b -= 2 * pj;
  i = b < pi ? 0 : (int)(mx - 1);
  long long r = b % pi;
  k = 1 + (int)(r / (my - 2));
  j = 1 + (int)(r % (my - 2));
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Source node of the periodic seam copy that finally owns (k,j,i), or -1.
__device__ __forceinline__ long long seam_src(const Grid &g, int k, int j, int i) {
  if (g.pz && (k == 0 || k == g.mz - 1)) return nid(g, k == 0 ? g.mz - 2 : 1, j, i);
  if (g.py) {
    if (j == 0 && k >= 1) return nid(g, k, g.my - 2, i);
    if (j == g.my - 1 && k >= 1 && k <= g.mz - 2) return nid(g, k, 1, i);
  }
  if (g.px && (i == 0 || i == g.mx - 1) && k >= 1 && j >= 1) return nid(g, k, j, i == 0 ? g.mx - 2 : 1);
  return -1;
}

__global__ void boundary_kernel(BndArgs a) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid &g = a.g;
  const long long np = g.np;
  double mq = 0.;
  if (b < a.nb) {
    int k, j, i;
    bnd_decode(g, b, k, j, i);
    const long long n = nid(g, k, j, i);
    double qn[5];
#pragma unroll
    for (int q = 0; q < 5; q++) mq = fmax(mq, fabs(a.q[n + q * np]));
    if (a.bc_nonperiodic) {
#pragma unroll
      for (int q = 0; q < 5; q++) qn[q] = a.qo[n + q * np];
    } else if (a.istage == 0) {
#pragma unroll
      for (int q = 0; q < 5; q++) qn[q] = a.a_rhs * 0.0 + a.qold[n + q * np];
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) qn[q] = a.a_old * a.qold[n + q * np] + a.a_rhs * 0.0 + a.a_cur * a.q[n + q * np];
    }
    const long long s = seam_src(g, k, j, i);
    if (s >= 0) {
#pragma unroll
      for (int q = 0; q < 5; q++) qn[q] = a.q[s + q * np];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) a.qo[n + q * np] = qn[q];
    a.po[n] = (a.gamma - 1) * (qn[4] - 0.5 * (qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]) / qn[0]);
  }
  if (a.norms) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mq = fmax(mq, __shfl_xor(mq, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&a.norms[64 + (blockIdx.x & 63)], dbits(mq));
  }
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

int launch_boundary(cles_ctx *c, int in, int out, int istage, double alfa, const double coef[3]) {
  (void)alfa;
  const cles_params &p = c->p;
  for (int q = 0; q < 6; q++) {
    if (p.bctype[q] != 7 && p.bctype[q] != 0)
      return set_error(c, CLES_EINVAL, "bctype[%d]=%d: only periodic (7) boundaries run on this path", q, p.bctype[q]);
  }
  BndArgs a;
  a.g = c->g;
  a.q = c->Q[in];
  a.p = c->P[in];
  a.qold = c->Q[c->cur];
  a.qo = c->Q[out];
  a.po = c->P[out];
  a.a_rhs = coef[0]; a.a_old = coef[1]; a.a_cur = coef[2];
  a.istage = istage;
  a.gamma = p.gamma;
  a.bc_nonperiodic = 0;
  a.norms = c->dnorm;
  const long long mx = c->g.mx, my = c->g.my, mz = c->g.mz;
  a.nb = 2 * mx * my + 2 * mx * (mz - 2) + 2 * (my - 2) * (mz - 2);
  const int bs = 256;
  const long long nblk = (a.nb + bs - 1) / bs;
  hipLaunchKernelGGL(boundary_kernel, dim3((unsigned)nblk), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
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

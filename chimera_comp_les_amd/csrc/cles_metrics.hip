// cles_metrics.hip -- curvilinear metrics on the device (FormMetrics,
// metrics.c:3-247): csi = d(eta) x d(zeta), eta = d(zeta) x d(xi),
// zet = d(xi) x d(eta) from half central coordinate differences, aj = 1/J,
// then the boundary layers copy their neighbour layer in the order of the
// reference (i, then j, then k).  Setup-time only.
#include "cles_internal.hpp"

namespace cles {

namespace {

struct MetArgs {
  Grid g;
  const double *coor;  // 3 planes
  double *csi, *eta, *zet, *aj;
};

__device__ __forceinline__ void co(const MetArgs &a, int k, int j, int i, double &x, double &y, double &z) {
  const long long n = nid(a.g, k, j, i);
  x = a.coor[n];
  y = a.coor[n + a.g.np];
  z = a.coor[n + 2 * a.g.np];
}

__global__ void metrics_kernel(MetArgs a) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid &g = a.g;
  const long long N = (long long)g.mx * g.my * g.mz;
  if (n >= N) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  const long long np = g.np;
  const bool ii = i >= 1 && i <= g.mx - 2, ji = j >= 1 && j <= g.my - 2, ki = k >= 1 && k <= g.mz - 2;
  double xp, yp, zp, xm, ym, zm;
  double dxdc = 0, dydc = 0, dzdc = 0, dxde = 0, dyde = 0, dzde = 0, dxdz = 0, dydz = 0, dzdz = 0;
  if (ii) {
    co(a, k, j, i + 1, xp, yp, zp);
    co(a, k, j, i - 1, xm, ym, zm);
    dxdc = 0.5 * (xp - xm); dydc = 0.5 * (yp - ym); dzdc = 0.5 * (zp - zm);
  }
  if (ji) {
    co(a, k, j + 1, i, xp, yp, zp);
    co(a, k, j - 1, i, xm, ym, zm);
    dxde = 0.5 * (xp - xm); dyde = 0.5 * (yp - ym); dzde = 0.5 * (zp - zm);
  }
  if (ki) {
    co(a, k + 1, j, i, xp, yp, zp);
    co(a, k - 1, j, i, xm, ym, zm);
    dxdz = 0.5 * (xp - xm); dydz = 0.5 * (yp - ym); dzdz = 0.5 * (zp - zm);
  }
  double c0 = 0, c1 = 0, c2 = 0, e0 = 0, e1 = 0, e2 = 0, z0 = 0, z1 = 0, z2 = 0, jac = 0;
  if (ki && ji) {  // metrics.c:52-70
    c0 = dyde * dzdz - dzde * dydz;
    c1 = -dxde * dzdz + dzde * dxdz;
    c2 = dxde * dydz - dyde * dxdz;
  }
  if (ki && ii) {  // metrics.c:73-92
    e0 = dydz * dzdc - dzdz * dydc;
    e1 = -dxdz * dzdc + dzdz * dxdc;
    e2 = dxdz * dydc - dydz * dxdc;
  }
  if (ji && ii) {  // metrics.c:95-113
    z0 = dydc * dzde - dzdc * dyde;
    z1 = -dxdc * dzde + dzdc * dxde;
    z2 = dxdc * dyde - dydc * dxde;
  }
  if (ii && ji && ki) {  // metrics.c:116-136
    double d = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) + dzdc * (dxde * dydz - dyde * dxdz);
    jac = 1. / d;
  }
  a.csi[n] = c0; a.csi[n + np] = c1; a.csi[n + 2 * np] = c2;
  a.eta[n] = e0; a.eta[n + np] = e1; a.eta[n + 2 * np] = e2;
  a.zet[n] = z0; a.zet[n + np] = z1; a.zet[n + 2 * np] = z2;
  a.aj[n] = jac;
}

// one boundary copy pass: dir 0/1/2 (i/j/k), side 0 (low) / 1 (high)
struct CopyArgs {
  MetArgs m;
  int dir, side;
};
__device__ __forceinline__ void cp3(double *v, long long d, long long s, long long np) {
  v[d] = v[s]; v[d + np] = v[s + np]; v[d + 2 * np] = v[s + 2 * np];
}
__global__ void metric_copy_kernel(CopyArgs a) {
  const Grid &g = a.m.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = g.np;
  int i, j, k;
  if (a.dir == 0) {
    if (t >= (long long)g.my * g.mz) return;
    j = (int)(t % g.my); k = (int)(t / g.my);
    i = a.side ? g.mx - 1 : 0;
    const long long d = nid(g, k, j, i), s = nid(g, k, j, a.side ? g.mx - 2 : 1);
    cp3(a.m.eta, d, s, np); cp3(a.m.zet, d, s, np); a.m.aj[d] = a.m.aj[s];
  } else if (a.dir == 1) {
    if (t >= (long long)g.mx * g.mz) return;
    i = (int)(t % g.mx); k = (int)(t / g.mx);
    j = a.side ? g.my - 1 : 0;
    const long long d = nid(g, k, j, i), s = nid(g, k, a.side ? g.my - 2 : 1, i);
    cp3(a.m.csi, d, s, np); cp3(a.m.zet, d, s, np); a.m.aj[d] = a.m.aj[s];
  } else {
    if (t >= (long long)g.mx * g.my) return;
    i = (int)(t % g.mx); j = (int)(t / g.mx);
    k = a.side ? g.mz - 1 : 0;
    const long long d = nid(g, k, j, i), s = nid(g, a.side ? g.mz - 2 : 1, j, i);
    cp3(a.m.eta, d, s, np); cp3(a.m.csi, d, s, np); a.m.aj[d] = a.m.aj[s];
  }
}

// squared LES filter width (pow(1/aj, 1/3))^2 of rhs.c:1693 and les.c:1574,
// formed once per node with the same two pow calls the reference makes

// The nine metrics and aj (values 0..8 the csi, eta, zet rows, 9 aj) packed
// in blocks of 8 consecutive nodes: block b holds five 128-byte lines, line p
// the pairs (2p, 2p + 1) of its 8 nodes, so value c of node n sits at
// ((n / 8) * 5 + c / 2) * 16 + (n % 8) * 2 + c % 2.  A face reads the ten
// values of each of its nodes with five 16-byte loads, a box node its metric
// row with two, and neighbouring nodes share the lines as in the planes.  The
// planes stay the metrics every other kernel reads.
struct PackArgs {
  Grid g;
  const double *m[4];  // csi, eta, zet (3 planes each), aj
  double *pack;
};
__global__ void metric_pack_kernel(PackArgs a) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = (long long)g.mx * g.my * g.mz, np = g.np;
  if (t >= N) return;
  const int i = (int)(t % g.mx), j = (int)((t / g.mx) % g.my), k = (int)(t / ((long long)g.mx * g.my));
  const long long n = nid(g, k, j, i);
#pragma unroll
  for (int c = 0; c < 10; c++)
    a.pack[((n >> 3) * 5 + (c >> 1)) * 16 + (n & 7) * 2 + (c & 1)] = c < 9 ? a.m[c / 3][(c % 3) * np + n] : a.m[3][n];
}

}  // namespace

int launch_metrics(cles_ctx *c) {
  MetArgs a{c->g, c->coor, c->csi, c->eta, c->zet, c->aj};
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  const int bs = 256;
  hipLaunchKernelGGL(metrics_kernel, dim3((unsigned)((N + bs - 1) / bs)), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  for (int dir = 0; dir < 3; dir++)
    for (int side = 0; side < 2; side++) {
      CopyArgs ca{a, dir, side};
      long long n = dir == 0 ? (long long)c->g.my * c->g.mz : dir == 1 ? (long long)c->g.mx * c->g.mz : (long long)c->g.mx * c->g.my;
      hipLaunchKernelGGL(metric_copy_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, c->stream, ca);
      CLES_HIP(c, hipGetLastError());
    }
  if (c->mpack) {  // after the boundary copies, which complete the metric planes
    PackArgs pa{c->g, {c->csi, c->eta, c->zet, c->aj}, c->mpack};
    hipLaunchKernelGGL(metric_pack_kernel, dim3((unsigned)((N + bs - 1) / bs)), dim3(bs), 0, c->stream, pa);
    CLES_HIP(c, hipGetLastError());
  }
  return CLES_OK;
}

}  // namespace cles

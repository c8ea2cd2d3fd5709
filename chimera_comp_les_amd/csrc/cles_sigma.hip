// cles_sigma.hip -- the shock sensor of formConvectionLES with -calc_sigma
// (rhs.c:3327-3444, compute_ri / sigma_cent rhs.c:2334-2357): the weight of
// the central flux in the hybrid (1 - sigma) F_upwind + sigma F_central,
// formed from the pressure of the first stage's input once per time step.
//
// Three node passes, each one thread per node, with a pass boundary wherever
// the reference refreshes its local vector (DMGlobalToLocal):
//   1. raw sensor per direction on the interior (ranges of rhs.c:3339-3378),
//      clipped to [s_min, s], the boundary nodes keeping (s_x, s_y, s_z);
//   2. smoothing by the geometric means of the raw neighbours
//      (rhs.c:3392-3405), written to the second buffer;
//   3. the curvature test and sigma_y = sigma_z = min (rhs.c:3419-3437).
// Reads beyond an owned node of a periodic direction wrap, which is the value
// DMGlobalToLocal gives the reference's ghost layer.
#include "cles_internal.hpp"

namespace cles {

namespace {

struct SigArgs {
  Grid g;
  const double *p;  // stage-input pressure
  double *a, *b;    // 3 planes each: the sensor buffers
  double s[3], smin[3];
  double ks;  // -sigma_thrsh
  int abc;    // -ABC: sigma = 0 for k > ABC
  int zper;   // bctype[4] == 7: the z sensor on every interior k
};

__device__ __forceinline__ double compute_ri(double fR, double f0, double fL, double ks) {
  const double dfL = fL - f0, dfR = f0 - fR;
  return (fabs(2 * dfL * dfR) + ks) / (dfL * dfL + dfR * dfR + ks);
}
__device__ __forceinline__ double sigma_cent(double fL, double f0, double fR, double fRR, double ks) {
  const double rL = compute_ri(fL, f0, fR, ks);
  const double rR = compute_ri(f0, fR, fRR, ks);
  const double r = rR < rL ? rR : rL, rd = 0.5, rmax = rd < r ? r : rd;
  return 0.5 + 0.5 * tanh(3.0 * (r - rd) / rmax) / tanh(3.0);
}
__device__ __forceinline__ bool node_of(const Grid &g, long long n, int &k, int &j, int &i) {
  if (n >= (long long)g.mx * g.my * g.mz) return false;
  i = (int)(n % g.mx);
  j = (int)((n / g.mx) % g.my);
  k = (int)(n / g.sz);
  return true;
}
__device__ __forceinline__ bool inner(const Grid &g, int k, int j, int i) {
  return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
__device__ __forceinline__ double clip(double v, double s, double smin) {
  v = s < v ? s : v;        // PetscMin(s, v)
  return smin < v ? v : smin;  // PetscMax(s_min, v)
}

// pass 1: raw sensor into a
__global__ __launch_bounds__(256) void sigma_raw_kernel(SigArgs a) {
  const Grid &g = a.g;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int k, j, i;
  if (!node_of(g, n, k, j, i)) return;
  const long long np = g.np;
  double s[3] = {a.s[0], a.s[1], a.s[2]};
  if (inner(g, k, j, i)) {
    const bool absorb = a.abc != 0 && k > a.abc;
    if (absorb) s[0] = 0.;
    else if (i < g.mx - 4 && i > 2)
      s[0] = sigma_cent(a.p[nidw(g, k, j, i - 1)], a.p[n], a.p[nidw(g, k, j, i + 1)], a.p[nidw(g, k, j, i + 2)], a.ks);
    if (absorb) s[2] = 0.;
    else if (a.zper || (k < g.mz - 4 && k > 4))
      s[2] = sigma_cent(a.p[nidw(g, k - 1, j, i)], a.p[n], a.p[nidw(g, k + 1, j, i)], a.p[nidw(g, k + 2, j, i)], a.ks);
    if (absorb) s[1] = 0.;
    else if (j < g.my - 4 && j > 4)
      s[1] = sigma_cent(a.p[nidw(g, k, j - 1, i)], a.p[n], a.p[nidw(g, k, j + 1, i)], a.p[nidw(g, k, j + 2, i)], a.ks);
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = clip(s[d], a.s[d], a.smin[d]);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) a.a[n + d * np] = s[d];
}

// pass 2: smoothing of the raw sensor a into b (the j pair is j-1 and j)
__global__ __launch_bounds__(256) void sigma_smooth_kernel(SigArgs a) {
  const Grid &g = a.g;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int k, j, i;
  if (!node_of(g, n, k, j, i)) return;
  const long long np = g.np;
  if (!inner(g, k, j, i)) {
#pragma unroll
    for (int d = 0; d < 3; d++) a.b[n + d * np] = a.a[n + d * np];
    return;
  }
  const long long jm = nidw(g, k, j - 1, i), km = nidw(g, k - 1, j, i), kp = nidw(g, k + 1, j, i);
  const long long im = nidw(g, k, j, i - 1), ip = nidw(g, k, j, i + 1);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double *S = a.a + d * np;
    a.b[n + d * np] = S[n] * (sqrt(S[jm] * S[n]) * sqrt(S[km] * S[kp]) * sqrt(S[im] * S[ip]));
  }
}

// pass 3: curvature test on src into dst, then sigma_y = sigma_z = min
__global__ __launch_bounds__(256) void sigma_curv_kernel(SigArgs a, const double *src, double *dst) {
  const Grid &g = a.g;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int k, j, i;
  if (!node_of(g, n, k, j, i)) return;
  const long long np = g.np;
  double sx = src[n], sy = src[n + np], sz = src[n + 2 * np];
  if (inner(g, k, j, i)) {
    const double *Y = src + np, *Z = src + 2 * np;
    if (j < g.my - 4 && j > 4) {
      const double ym = Y[nidw(g, k, j - 1, i)], yp = Y[nidw(g, k, j + 1, i)], y0 = Y[n];
      if (ym + yp - 2. * y0 < 0.0) sy = ym * y0 * y0;
    }
    if (k < g.mz - 4 && k > 4) {
      const double zm = Z[nidw(g, k - 1, j, i)], zp = Z[nidw(g, k + 1, j, i)], z0 = Z[n];
      if (zm + zp - 2. * z0 < 0.0) sz = zm * zp * z0;
    }
    sz = sz < sy ? sz : sy;
    sy = sz;
  }
  dst[n] = sx;
  dst[n + np] = sy;
  dst[n + 2 * np] = sz;
}

}  // namespace

// The sensor of the state in buffer buf into c->sgm (calc_sigma only), on
// stream s.
int launch_sigma_on(cles_ctx *c, int buf, hipStream_t s) {
  const cles_params &p = c->p;
  SigArgs a;
  a.g = c->g;
  a.p = c->P[buf];
  a.a = c->sgm;
  a.b = c->sgm_tmp;
  a.s[0] = p.s_x; a.s[1] = p.s_y; a.s[2] = p.s_z;
  a.smin[0] = p.s_x_min; a.smin[1] = p.s_y_min; a.smin[2] = p.s_z_min;
  a.ks = p.sigma_thrsh;
  a.abc = p.abc;
  a.zper = p.bctype[4] == 7;
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  const dim3 grid((unsigned)((N + 255) / 256)), blk(256);
  hipLaunchKernelGGL(sigma_raw_kernel, grid, blk, 0, s, a);
  CLES_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(sigma_smooth_kernel, grid, blk, 0, s, a);
  CLES_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(sigma_curv_kernel, grid, blk, 0, s, a, (const double *)c->sgm_tmp, c->sgm);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}

int launch_sigma(cles_ctx *c, int buf) { return launch_sigma_on(c, buf, c->stream); }

}  // namespace cles

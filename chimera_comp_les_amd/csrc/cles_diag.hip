// cles_diag.hip -- domain integrals of the conserved variables on the unique
// interior nodes (1..m-2 in every direction), volume weighted by 1/aj:
// mass, the three momenta, total energy and kinetic energy.  With
// Rhs = aj (Visc - Conv) the stage update changes sum(Q / aj) only by the
// boundary fluxes, so on periodic grids these sums are conserved to
// round-off.  Two passes with a fixed summation order, so the result does not
// depend on scheduling.
#include "cles_internal.hpp"

namespace cles {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

struct RedArgs {
  Grid g;
  const double *q, *aj;
  double *part;  // [6][kRedBlocks]
};

__global__ __launch_bounds__(kRedThreads) void integrals_kernel(RedArgs a) {
  __shared__ double sh[6][kRedThreads];
  const Grid &g = a.g;
  const long long ni = g.mx - 2, nj = g.my - 2, nk = g.mz - 2;
  const long long N = ni * nj * nk;
  const long long np = g.np;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long long t = (long long)blockIdx.x * kRedThreads + threadIdx.x; t < N; t += (long long)kRedBlocks * kRedThreads) {
    const int i = 1 + (int)(t % ni), j = 1 + (int)((t / ni) % nj), k = 1 + (int)(t / (ni * nj));
    const long long n = nid(g, k, j, i);
    const double vol = 1.0 / a.aj[n];
    const double r = a.q[n], ru = a.q[n + np], rv = a.q[n + 2 * np], rw = a.q[n + 3 * np], e = a.q[n + 4 * np];
    acc[0] += r * vol;
    acc[1] += ru * vol;
    acc[2] += rv * vol;
    acc[3] += rw * vol;
    acc[4] += e * vol;
    acc[5] += 0.5 * (ru * ru + rv * rv + rw * rw) / r * vol;
  }
#pragma unroll
  for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
#pragma unroll
      for (int c = 0; c < 6; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 6; c++) a.part[c * kRedBlocks + blockIdx.x] = sh[c][0];
}

__global__ __launch_bounds__(kRedThreads) void integrals_final_kernel(const double *part, double *out) {
  __shared__ double sh[6][kRedThreads];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < kRedBlocks; b += kRedThreads)
#pragma unroll
    for (int c = 0; c < 6; c++) acc[c] += part[c * kRedBlocks + b];
#pragma unroll
  for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
#pragma unroll
      for (int c = 0; c < 6; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 6; c++) out[c] = sh[c][0];
}

// Do_averaging, les.c:1664-1747 (period 1 without -fish): Q_sum += Q and
// P_sum += P (VecAXPY with alpha 1: y + x), then the squares and cross
// products of the Ucat of the last CalcUCat (u = rhoU / rho of that stage
// input, rhs.c:1346-1350; zero before any CalcUCat) on every owned node.
struct AvgArgs {
  long long n, np;
  const double *q, *p, *qu;  // state, pressure, CalcUCat input (or null)
  double *s;                 // 12 planes
  int twoD;                  // 1: CalcUCat sets u = 0 (rhs.c:1349)
};

__global__ __launch_bounds__(256) void averaging_kernel(AvgArgs a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const long long np = a.np;
  double *s = a.s;
#pragma unroll
  for (int c = 0; c < 5; c++) s[c * np + t] = s[c * np + t] + a.q[c * np + t];
  s[5 * np + t] = s[5 * np + t] + a.p[t];
  double u = 0., v = 0., w = 0.;
  if (a.qu) {
    const double r = a.qu[t];
    u = a.twoD == 1 ? 0. : a.qu[np + t] / r;
    v = a.qu[2 * np + t] / r;
    w = a.qu[3 * np + t] / r;
  }
  s[6 * np + t] += u * u;
  s[7 * np + t] += v * v;
  s[8 * np + t] += w * w;
  s[9 * np + t] += u * v;   // uv
  s[10 * np + t] += v * w;  // vw
  s[11 * np + t] += w * u;  // wu
}

}  // namespace

int launch_averaging(cles_ctx *c) {
  if (!c->avg) return set_error(c, CLES_ESTATE, "averaging buffers not allocated (-averaging 0)");
  AvgArgs a;
  a.n = (long long)c->g.mx * c->g.my * c->g.mz;
  a.np = c->g.np;
  a.q = c->Q[c->cur];
  a.p = c->P[c->cur];
  a.qu = c->ucat_src >= 0 ? c->Q[c->ucat_src] : nullptr;
  a.s = c->avg;
  a.twoD = c->p.twoD;
  hipLaunchKernelGGL(averaging_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}
}  // namespace cles

using namespace cles;

extern "C" int cles_integrals(cles_ctx *c, double out[6]) {
  if (!c || !out) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  double *buf = nullptr;
  CLES_HIP(c, hipMalloc((void **)&buf, (6 * kRedBlocks + 6) * sizeof(double)));
  RedArgs a{c->g, c->Q[c->cur], c->aj, buf};
  hipLaunchKernelGGL(integrals_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(integrals_final_kernel, dim3(1), dim3(kRedThreads), 0, c->stream, buf, buf + 6 * kRedBlocks);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, buf + 6 * kRedBlocks, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return set_error(c, CLES_EDEVICE, "cles_integrals: %s", hipGetErrorString(e));
  return CLES_OK;
}


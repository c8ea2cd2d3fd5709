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

// QCriteria (data.c:4815-5053), one thread per node.  Boundary nodes keep
// the pressure (the reference writes Q into P on nodes 1..m-2 only) and a
// lambda2 of 0 (Phi).  Velocity differences in index space: one-sided at
// node 1 or beside a low-side node with nvert > 2, at node m-2 or beside a
// high-side one, else central; the face metrics averaged to the node times
// aj; d_ab = du_a/dx_b, s = d + d^T, w = d - d^T; Q = (wo - so) / 2 and
// lambda2 the middle of jacobi's estimates for ss + ww, all in the
// reference's operation order (this file is built without contraction).
struct QcritArgs {
  Grid g;
  const double *q, *p, *csi, *eta, *zet, *aj, *nv;
  double *qc, *l2;
};

// jacobi (data.c:5057-5170) for n = 3 and eigsrt (data.c:5172-5203), the
// middle estimate.  Only the diagonal and the upper triangle are read.  With
// n = 3 the reference's update loops for j < ip - 1 and ip + 1 <= j < iq - 1
// are empty, so a rotation in (0,1) also updates a02 and a12 (its j = iq + 1
// loop) and the rotations in (0,2) and (1,2) update nothing else.  At most
// five sweeps, the last estimates standing.
__device__ double jacobi3_middle(double a00, double a01, double a02, double a11, double a12, double a22) {
  double b0 = a00, b1 = a11, b2 = a22;
  double d0 = b0, d1 = b1, d2 = b2;
  double z0 = 0., z1 = 0., z2 = 0.;
  for (int sweep = 0; sweep < 5; sweep++) {
    double sm = 0.;
    sm = sm + fabs(a01);
    sm = sm + fabs(a02);
    sm = sm + fabs(a12);
    if (sm == 0.) break;
    const double tresh = sweep < 4 ? 0.2 * sm / 9 : 0.;
    double s = 0., tau = 0.;
    auto rotate = [&](double &apq, double &dp, double &dq, double &zp, double &zq) {
      const double g = 100. * fabs(apq);
      if (!(fabs(apq) > tresh)) return false;
      double h = dq - dp, t;
      if (fabs(h) + g == fabs(h)) {
        t = apq / h;
      } else {
        const double theta = 0.5 * h / apq;
        t = 1. / (fabs(theta) + sqrt(1. + theta * theta));
        if (theta < 0.) t = -t;
      }
      const double c = 1. / sqrt(1. + t * t);
      s = t * c;
      tau = s / (1. + c);
      h = t * apq;
      zp = zp - h;
      zq = zq + h;
      dp = dp - h;
      dq = dq + h;
      apq = 0.;
      return true;
    };
    if (rotate(a01, d0, d1, z0, z1)) {
      const double g = a02, h = a12;
      a02 = g - s * (h + g * tau);
      a12 = h + s * (g - h * tau);
    }
    rotate(a02, d0, d2, z0, z2);
    rotate(a12, d1, d2, z1, z2);
    b0 = b0 + z0;
    d0 = b0;
    z0 = 0.;
    b1 = b1 + z1;
    d1 = b1;
    z1 = 0.;
    b2 = b2 + z2;
    d2 = b2;
    z2 = 0.;
  }
  // eigsrt: selection into descending order; only the middle is kept
  double p = d0;
  int k = 0;
  if (d1 > p) {
    k = 1;
    p = d1;
  }
  if (d2 > p) {
    k = 2;
    p = d2;
  }
  if (k == 1) {
    d1 = d0;
  } else if (k == 2) {
    d2 = d0;
  }
  return d2 > d1 ? d2 : d1;
}

__device__ __forceinline__ void qcrit_node(const QcritArgs &a, long long n, int i, int j, int k, double &qv,
                                           double &lv) {
  const Grid &g = a.g;
  const long long np = g.np, sy = g.sy, sz = g.sz;
  const double *q = a.q;
  auto vel = [&](long long m, int c) { return q[(c + 1) * np + m] / q[m]; };
  const double *nv = a.nv;
  // index-space differences of u, v, w along i (uc), j (ue), k (uz)
  double uc[3], ue[3], uz[3];
  const int ci = (i == 1 || nv[n - 1] > 2.) ? 1 : (i == g.mx - 2 || nv[n + 1] > 2.) ? -1 : 0;
  const int cj = (j == 1 || nv[n - sy] > 2) ? 1 : (j == g.my - 2 || nv[n + sy] > 2) ? -1 : 0;
  const int ck = (k == 1 || nv[n - sz] > 2) ? 1 : (k == g.mz - 2 || nv[n + sz] > 2) ? -1 : 0;
  auto diff = [&](int side, long long st, int c) {
    if (side == 1) return (vel(n + st, c) - vel(n, c));
    if (side == -1) return (vel(n, c) - vel(n - st, c));
    return (vel(n + st, c) - vel(n - st, c)) * 0.5;
  };
#pragma unroll
  for (int c = 0; c < 3; c++) {
    uc[c] = diff(ci, 1, c);
    ue[c] = diff(cj, sy, c);
    uz[c] = diff(ck, sz, c);
  }
  const double ajn = a.aj[n];
  double cs[3], et[3], ze[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cs[c] = (a.csi[c * np + n] + a.csi[c * np + n - 1]) * 0.5 * ajn;
    et[c] = (a.eta[c * np + n] + a.eta[c * np + n - sy]) * 0.5 * ajn;
    ze[c] = (a.zet[c * np + n] + a.zet[c * np + n - sz]) * 0.5 * ajn;
  }
  double dd[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) dd[r][c] = uc[r] * cs[c] + ue[r] * et[c] + uz[r] * ze[c];
  const double s11 = dd[0][0] + dd[0][0], s12 = dd[0][1] + dd[1][0], s13 = dd[0][2] + dd[2][0];
  const double s21 = s12, s22 = dd[1][1] + dd[1][1], s23 = dd[1][2] + dd[2][1];
  const double s31 = s13, s32 = s23, s33 = dd[2][2] + dd[2][2];
  const double so = (s11 * s11 + s22 * s22 + s33 * s33) / 4. + (s12 * s12 + s13 * s13 + s23 * s23) / 2.;
  // the reference's "S^2" and "W^2" (data.c:4981-4990, 5005-5014), upper triangle
  const double ss11 = s11 * s11 + s21 * s12 + s31 * s13;
  const double ss12 = s12 * s11 + s22 * s12 * s32 * s13;
  const double ss13 = s13 * s11 + s23 * s12 * s33 * s13;
  const double ss22 = s12 * s21 + s22 * s22 * s32 * s23;
  const double ss23 = s13 * s21 + s23 * s22 * s33 * s23;
  const double ss33 = s13 * s31 + s23 * s32 * s33 * s33;
  const double w11 = 0., w12 = dd[0][1] - dd[1][0], w13 = dd[0][2] - dd[2][0];
  const double w21 = -w12, w22 = 0., w23 = dd[1][2] - dd[2][1];
  const double w31 = -w13, w32 = -w23, w33 = 0.;
  const double wo = (w12 * w12 + w13 * w13 + w23 * w23) / 2.;
  const double ww11 = w11 * w11 + w21 * w12 + w31 * w13;
  const double ww12 = w12 * w11 + w22 * w12 * w32 * w13;
  const double ww13 = w13 * w11 + w23 * w12 * w33 * w13;
  const double ww22 = w12 * w21 + w22 * w22 * w32 * w23;
  const double ww23 = w13 * w21 + w23 * w22 * w33 * w23;
  const double ww33 = w13 * w31 + w23 * w32 * w33 * w33;
  lv = jacobi3_middle(ss11 + ww11, ss12 + ww12, ss13 + ww13, ss22 + ww22, ss23 + ww23, ss33 + ww33);
  qv = (wo - so) / 2.;
}

__global__ __launch_bounds__(256) void qcrit_kernel(QcritArgs a) {
  const Grid &g = a.g;
  const long long N = (long long)g.mx * g.my * g.mz;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / ((long long)g.mx * g.my));
  // Both outputs are stored once, at the end, for every node: a boundary
  // path with stores of its own let the compiler merge one of them with the
  // interior's into a store of per-lane base, whose base it then set wrongly
  // for some boundary lanes (hipcc of ROCm 7.2, gfx950).
  double qv = a.p[n], lv = 0.;
  if (i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2) qcrit_node(a, n, i, j, k, qv, lv);
  a.qc[n] = qv;
  a.l2[n] = lv;
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

extern "C" int cles_q_criteria(cles_ctx *c, double *q, double *lambda2) {
  if (!c || !q || !lambda2) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  if (c->g.mx < 3 || c->g.my < 3 || c->g.mz < 3)
    return set_error(c, CLES_EINVAL, "Q criterion: a grid of at least 3 nodes per direction is needed");
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  double *buf = nullptr;
  CLES_HIP(c, hipMalloc((void **)&buf, 2 * (size_t)N * sizeof(double)));
  QcritArgs a{c->g, c->Q[c->cur], c->P[c->cur], c->csi, c->eta, c->zet, c->aj, c->nvert, buf, buf + N};
  hipLaunchKernelGGL(qcrit_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(q, buf, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(lambda2, buf + N, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return set_error(c, CLES_EDEVICE, "cles_q_criteria: %s", hipGetErrorString(e));
  return CLES_OK;
}

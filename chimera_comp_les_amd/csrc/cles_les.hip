// cles_les.hip -- dynamic Smagorinsky model (Compute_Smagorinsky_Constant_1,
// les.c:135-1472) and eddy viscosity (Compute_eddy_viscosity_LES,
// les.c:1477-1661), once per time step before the Runge-Kutta stages.
//
// Velocity semantics.  The reference keeps a persistent lUcat written by
// CalcUCat at every RHS evaluation; the Smagorinsky gradient loop then
// overwrites it node by node in k,j,i order, so at node (k,j,i) its -1
// neighbours that are interior and not solid already hold the current
// velocity while its +1 neighbours and all boundary/solid nodes still hold
// the velocity of the last CalcUCat (the input of the previous step's last
// stage).  The device keeps that stage-input state resident (one of the
// three RK buffers) and evaluates both velocities from it, so no velocity
// field is ever stored.
//
// Kernels: gradients -> 27-point test filter (L_ij M_ij contraction) ->
// 27-point <LM>/<MM> + clipping -> volume-weighted 27-point smoothing and
// nu_t (a separate nu_t kernel on the paths that run no test filter).  Each
// is one thread per node with i-contiguous loads.
#include "cles_internal.hpp"

namespace cles {

namespace {

constexpr double LES_EPS = 0.000001;  // les.c:8

struct LesArgs {
  Grid g;
  const double *q;   // current state (lQ)
  const double *qs;  // state of the last CalcUCat, or null (lUcat still zero)
  const double *csi, *eta, *zet, *aj, *nv;
  int has_ib;
  double *grad;  // 7 planes: S11 S12 S13 S22 S23 S33 (no trace removal) and |S|
  double *ua;    // 3 planes: lUcat after the gradient loop
  double *lm, *mm;
  double *lmq, *mmq;    // LM/27, MM/27
  double *sabs;         // |S| for nu_t, formed by the test filter
  const double *ivol;   // 1/aj
  double *cs, *cs_out, *nut;
  const double *fil2;   // per-node (pow(1/aj, 1/3))^2, formed once at setup
  const double *tfil2;  // per-node squared test-filter width
  int les, testfilter_ik, ucat_updated;
  int twoD;  // 1: the last CalcUCat set u = 0 (rhs.c:1349); the loop's own velocities keep it
  double max_cs;
};

__device__ __forceinline__ double nvat(const LesArgs &a, long long n) { return a.has_ib ? a.nv[n] : 0.0; }
__device__ __forceinline__ bool interior(const Grid &g, int k, int j, int i) {
  return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
// node visited (and overwritten) by the gradient loop of les.c:250-258
__device__ __forceinline__ bool in_loop(const LesArgs &a, int k, int j, int i) {
  return interior(a.g, k, j, i) && !(nvat(a, nid(a.g, k, j, i)) > 1.1);
}
__device__ __forceinline__ void vel(const double *q, long long n, long long np, double u[3]) {
  const double r = q[n];
  u[0] = q[n + np] / r;
  u[1] = q[n + 2 * np] / r;
  u[2] = q[n + 3 * np] / r;
}
__device__ __forceinline__ void u_stale(const LesArgs &a, long long n, double u[3]) {
  if (a.qs) {
    vel(a.qs, n, a.g.np, u);
    if (a.twoD == 1) u[0] = 0.0;
  } else {
    u[0] = u[1] = u[2] = 0.0;
  }
}
// Compute_du_center + Compute_du_dxyz (les.c:31-63)
__device__ __forceinline__ void grad9(const LesArgs &a, long long n, const double up_i[3], const double um_i[3],
                                      const double up_j[3], const double um_j[3], const double up_k[3],
                                      const double um_k[3], double gr[9]) {
  const long long np = a.g.np;
  double dudc = (up_i[0] - um_i[0]) * 0.5, dvdc = (up_i[1] - um_i[1]) * 0.5, dwdc = (up_i[2] - um_i[2]) * 0.5;
  double dudz = (up_k[0] - um_k[0]) * 0.5, dvdz = (up_k[1] - um_k[1]) * 0.5, dwdz = (up_k[2] - um_k[2]) * 0.5;
  double dude = (up_j[0] - um_j[0]) * 0.5, dvde = (up_j[1] - um_j[1]) * 0.5, dwde = (up_j[2] - um_j[2]) * 0.5;
  const double csi0 = a.csi[n], csi1 = a.csi[n + np], csi2 = a.csi[n + 2 * np];
  const double eta0 = a.eta[n], eta1 = a.eta[n + np], eta2 = a.eta[n + 2 * np];
  const double zet0 = a.zet[n], zet1 = a.zet[n + np], zet2 = a.zet[n + 2 * np];
  const double ajc = a.aj[n];
  gr[0] = (dudc * csi0 + dude * eta0 + dudz * zet0) * ajc;
  gr[1] = (dudc * csi1 + dude * eta1 + dudz * zet1) * ajc;
  gr[2] = (dudc * csi2 + dude * eta2 + dudz * zet2) * ajc;
  gr[3] = (dvdc * csi0 + dvde * eta0 + dvdz * zet0) * ajc;
  gr[4] = (dvdc * csi1 + dvde * eta1 + dvdz * zet1) * ajc;
  gr[5] = (dvdc * csi2 + dvde * eta2 + dvdz * zet2) * ajc;
  gr[6] = (dwdc * csi0 + dwde * eta0 + dwdz * zet0) * ajc;
  gr[7] = (dwdc * csi1 + dwde * eta1 + dwdz * zet1) * ajc;
  gr[8] = (dwdc * csi2 + dwde * eta2 + dwdz * zet2) * ajc;
}
__device__ __forceinline__ double sabs_of(const double g[9]) {
  double Sdd = 2. / 3. * (g[0] + g[4] + g[8]);
  double Sxx = 0.5 * (g[0] + g[0]) - Sdd, Sxy = 0.5 * (g[1] + g[3]), Sxz = 0.5 * (g[2] + g[6]);
  double Syx = Sxy, Syy = 0.5 * (g[4] + g[4]) - Sdd, Syz = 0.5 * (g[5] + g[7]);
  double Szx = Sxz, Szy = Syz, Szz = 0.5 * (g[8] + g[8]) - Sdd;
  return sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz + Szx * Szx + Szy * Szy +
                     Szz * Szz));
}

// Velocity gradients of the loop les.c:250-309 (mixed current / previous
// velocities, see above), stored as the six strain-rate components the test
// filter uses (S_ab = (du_a/dx_b + du_b/dx_a)/2, les.c:520-522) and |S| with
// the trace term removed (les.c:268-274).  Also stores the velocity lUcat
// holds once the loop has finished, read by the test filter and by nu_t.
// Node-range kernels: each XCD (blockIdx % 8) walks a contiguous range of
// blocks, so the k +- 1 neighbours of its nodes stay in its own L2.
__device__ __forceinline__ long long xcd_node(int per_xcd) {
  const int b = blockIdx.x;
  return (long long)((b & 7) * per_xcd + (b >> 3)) * blockDim.x + threadIdx.x;
}

__global__ __launch_bounds__(256) void les_grad_kernel(LesArgs a, int per_xcd) {
  const Grid &g = a.g;
  const long long n = xcd_node(per_xcd);
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  const long long np = g.np;
  const bool loop = a.ucat_updated && in_loop(a, k, j, i);
  double u[3];
  if (loop) vel(a.q, n, np, u); else u_stale(a, n, u);
  a.ua[n] = u[0];
  a.ua[n + np] = u[1];
  a.ua[n + 2 * np] = u[2];
  if (!a.ucat_updated) return;
  if (!loop) {
#pragma unroll
    for (int c = 0; c < 7; c++) a.grad[n + c * np] = 0.0;
    return;
  }
  double up_i[3], up_j[3], up_k[3], um_i[3], um_j[3], um_k[3];
  u_stale(a, n + 1, up_i);
  u_stale(a, n + g.sy, up_j);
  u_stale(a, n + g.sz, up_k);
  if (in_loop(a, k, j, i - 1)) vel(a.q, n - 1, np, um_i); else u_stale(a, n - 1, um_i);
  if (in_loop(a, k, j - 1, i)) vel(a.q, n - g.sy, np, um_j); else u_stale(a, n - g.sy, um_j);
  if (in_loop(a, k - 1, j, i)) vel(a.q, n - g.sz, np, um_k); else u_stale(a, n - g.sz, um_k);
  double gr[9];
  grad9(a, n, up_i, um_i, up_j, um_j, up_k, um_k, gr);
  a.grad[n] = 0.5 * (gr[0] + gr[0]);
  a.grad[n + np] = 0.5 * (gr[1] + gr[3]);
  a.grad[n + 2 * np] = 0.5 * (gr[2] + gr[6]);
  a.grad[n + 3 * np] = 0.5 * (gr[4] + gr[4]);
  a.grad[n + 4 * np] = 0.5 * (gr[5] + gr[7]);
  a.grad[n + 5 * np] = 0.5 * (gr[8] + gr[8]);
  a.grad[n + 6 * np] = sabs_of(gr);
}

// gradient component as seen by the test filter: the local vectors lSx..lSz
// hold the periodic face copies of les.c:316-377 and zeros on every other
// boundary node
__device__ __forceinline__ long long seam_image(const Grid &g, int k, int j, int i) {
  const bool bi = i == 0 || i == g.mx - 1, bj = j == 0 || j == g.my - 1, bk = k == 0 || k == g.mz - 1;
  if (bi && !bj && !bk && g.px) return nid(g, k, j, i == 0 ? g.mx - 2 : 1);
  if (bj && !bi && !bk && g.py) return nid(g, k, j == 0 ? g.my - 2 : 1, i);
  if (bk && !bi && !bj && g.pz) return nid(g, k == 0 ? g.mz - 2 : 1, j, i);
  return -1;
}

// Test filter (integrate_testfilter, les.c:80-131, and the L_ij M_ij
// contraction of les.c:505-887).  A workgroup owns an 8x8x4 tile of interior
// nodes and stages the node data of the tile plus a one-node halo in LDS, in
// two passes over one 8 x 600 buffer: first rho, rhoU, rhoV, rhoW and the lUcat
// velocity, then the six strain-rate components and |S| as the filter sees
// them (periodic face copies of lSx..lSz, les.c:316-377; zero on other
// boundary nodes; |S| has no periodic copy), with rho kept.  Every running
// sum still adds its 27 terms in the reference's [k][j][i] order.  The
// weights 1/(|dk|+1) * 1/(|dj|+1) * 1/(|di|+1) are 1, 1/2, 1/4 or 1/8, so
// w * x is exact and fma(w, x, s) rounds once, exactly as s + w * x does
// (unless |w x| falls below 2^-1022, where the product itself would round);
// their sum is 8 exactly, so vsum / wsum is vsum * 0.125 (the same exact
// quotient rounded once).
constexpr int FTX = 8, FTY = 8, FTZ = 4;
constexpr double FILTER_INV_WSUM = 0.125;
#ifndef CLES_FILTER_XZY
#define CLES_FILTER_XZY 1
#endif
constexpr int FBX = FTX + 2, FBY = FTY + 2, FBZ = FTZ + 2, FNB = FBX * FBY * FBZ;
static_assert(FTX * FTY * FTZ == 256, "one node per thread");

__device__ __forceinline__ double filter_w(int R, int Qd, int Pd) {
  return (R == 1 ? 1.0 : 0.5) * (Qd == 1 ? 1.0 : 0.5) * (Pd == 1 ? 1.0 : 0.5);
}

__global__ __launch_bounds__(256) void les_filter_kernel(LesArgs a, int ntx, int nty, int ntiles, int per_xcd) {
  __shared__ double B[8][FNB];
  const Grid &g = a.g;
  const long long np = g.np;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);  // contiguous tiles per XCD
  if (tile >= ntiles) return;
#if CLES_FILTER_XZY
  // x fastest, then z, then y (the tiles in flight share their z halos)
  const int ntz = ntiles / (ntx * nty);
  const int i0 = 1 + (tile % ntx) * FTX, k0 = 1 + ((tile / ntx) % ntz) * FTZ, j0 = 1 + (tile / (ntx * ntz)) * FTY;
#else
  const int i0 = 1 + (tile % ntx) * FTX, j0 = 1 + ((tile / ntx) % nty) * FTY, k0 = 1 + (tile / (ntx * nty)) * FTZ;
#endif
  const int t = threadIdx.x;
  const int lx = t % FTX, ly = (t / FTX) % FTY, lz = t / (FTX * FTY);
  const int i = i0 + lx, j = j0 + ly, k = k0 + lz;
  const bool valid = i <= g.mx - 2 && j <= g.my - 2 && k <= g.mz - 2;
  const long long n0 = valid ? nid(g, k, j, i) : 0;
  const bool act = valid && !(nvat(a, n0) > 1.1);
  const int c0 = ((lz + 1) * FBY + (ly + 1)) * FBX + (lx + 1);  // slot of the node itself
  // ---- pass 1: density, momentum, velocity
  for (int w = t; w < FNB; w += 256) {
    const int ii = i0 - 1 + w % FBX, jj = j0 - 1 + (w / FBX) % FBY, kk = k0 - 1 + w / (FBX * FBY);
    if (ii > g.mx - 1 || jj > g.my - 1 || kk > g.mz - 1) continue;
    const long long m = nid(g, kk, jj, ii);
    B[0][w] = a.q[m];
    B[1][w] = a.q[m + np];
    B[2][w] = a.q[m + 2 * np];
    B[3][w] = a.q[m + 3 * np];
    B[4][w] = a.ua[m];
    B[5][w] = a.ua[m + np];
    B[6][w] = a.ua[m + 2 * np];
  }
  __syncthreads();
  double s_rho = 0, s_ru = 0, s_rv = 0, s_rw = 0, s_ruu = 0, s_ruv = 0, s_ruw = 0, s_rvv = 0, s_rvw = 0, s_rww = 0;
  if (act) {
#pragma unroll 1
    for (int R = 0; R < 3; R++)
#pragma unroll
      for (int Qd = 0; Qd < 3; Qd++)
#pragma unroll
        for (int Pd = 0; Pd < 3; Pd++) {
          const int o = c0 + ((R - 1) * FBY + (Qd - 1)) * FBX + (Pd - 1);
          const double w = filter_w(R, Qd, Pd);
          const double rho = B[0][o], ru = B[1][o], rv = B[2][o], rw = B[3][o];
          const double u0 = B[4][o], u1 = B[5][o], u2 = B[6][o];
          s_rho = fma(w, rho, s_rho);
          s_ru = fma(w, ru, s_ru);
          s_rv = fma(w, rv, s_rv);
          s_rw = fma(w, rw, s_rw);
          s_ruu = fma(w, ru * u0, s_ruu);
          s_ruv = fma(w, ru * u1, s_ruv);
          s_ruw = fma(w, ru * u2, s_ruw);
          s_rvv = fma(w, rv * u1, s_rvv);
          s_rvw = fma(w, rv * u2, s_rvw);
          s_rww = fma(w, rw * u2, s_rww);
        }
    // |S| of the velocity lUcat holds after the gradient loop, the input of
    // nu_t (les.c:1536-1573): its six neighbours are in the window
    double up_i[3], um_i[3], up_j[3], um_j[3], up_k[3], um_k[3];
    constexpr int sj = FBX, sk = FBX * FBY;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      up_i[c] = B[4 + c][c0 + 1];
      um_i[c] = B[4 + c][c0 - 1];
      up_j[c] = B[4 + c][c0 + sj];
      um_j[c] = B[4 + c][c0 - sj];
      up_k[c] = B[4 + c][c0 + sk];
      um_k[c] = B[4 + c][c0 - sk];
    }
    double gr[9];
    grad9(a, n0, up_i, um_i, up_j, um_j, up_k, um_k, gr);
    a.sabs[n0] = sabs_of(gr);
  }
  __syncthreads();
  // ---- pass 2: strain rates and |S| as the filter sees them (rho stays in B[0])
  for (int w = t; w < FNB; w += 256) {
    const int ii = i0 - 1 + w % FBX, jj = j0 - 1 + (w / FBX) % FBY, kk = k0 - 1 + w / (FBX * FBY);
    if (ii > g.mx - 1 || jj > g.my - 1 || kk > g.mz - 1) continue;
    long long m = nid(g, kk, jj, ii);
    double S = 0.0;
    if (interior(g, kk, jj, ii)) {
      S = a.grad[m + 6 * np];
    } else {
      m = seam_image(g, kk, jj, ii);
    }
#pragma unroll
    for (int c = 0; c < 6; c++) B[1 + c][w] = m >= 0 ? a.grad[m + c * np] : 0.0;
    B[7][w] = S;
  }
  __syncthreads();
  if (!act) {
    if (valid) {
      a.lm[n0] = 0;
      a.mm[n0] = 0;
      a.lmq[n0] = 0;
      a.mmq[n0] = 0;
    }
    return;
  }
  double sS[6] = {0, 0, 0, 0, 0, 0}, sSS[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
  for (int R = 0; R < 3; R++)
#pragma unroll
    for (int Qd = 0; Qd < 3; Qd++)
#pragma unroll
      for (int Pd = 0; Pd < 3; Pd++) {
        const int o = c0 + ((R - 1) * FBY + (Qd - 1)) * FBX + (Pd - 1);
        const double w = filter_w(R, Qd, Pd);
        const double rho = B[0][o], S = B[7][o];
#pragma unroll
        for (int c = 0; c < 6; c++) {
          const double Sv = B[1 + c][o];
          sS[c] = fma(w, Sv, sS[c]);
          sSS[c] = fma(w, Sv * S * rho, sSS[c]);
        }
      }
  double Sx[6];
#pragma unroll
  for (int c = 0; c < 6; c++) Sx[c] = B[1 + c][c0];
  const double iw = FILTER_INV_WSUM;
  const double _rho = s_rho * iw, _rhou = s_ru * iw, _rhov = s_rv * iw, _rhow = s_rw * iw;
  // full 3x3 tensors in row-major order from the six symmetric components
  const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  double Lij[9];
  Lij[0] = s_ruu * iw - _rhou * _rhou / _rho;
  Lij[1] = s_ruv * iw - _rhou * _rhov / _rho;
  Lij[2] = s_ruw * iw - _rhou * _rhow / _rho;
  Lij[3] = s_ruv * iw - _rhov * _rhou / _rho;
  Lij[4] = s_rvv * iw - _rhov * _rhov / _rho;
  Lij[5] = s_rvw * iw - _rhov * _rhow / _rho;
  Lij[6] = s_ruw * iw - _rhow * _rhou / _rho;
  Lij[7] = s_rvw * iw - _rhow * _rhov / _rho;
  Lij[8] = s_rww * iw - _rhow * _rhow / _rho;
  double Sh6[6], SSh6[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    Sh6[c] = sS[c] * iw;
    SSh6[c] = sSS[c] * iw;
  }
  double S_hat = 0;
#pragma unroll
  for (int c = 0; c < 9; c++) S_hat += Sh6[sym[c]] * Sh6[sym[c]];
  S_hat = sqrt(2 * S_hat);
  double L_cat = 0.0, M_cat = 0.0;
  const double tf2 = a.tfil2[n0], f2 = a.fil2[n0];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int ab = sym[p * 3 + q], aa = sym[p * 4], bb = sym[q * 4];
      M_cat += -2.0 * _rho * tf2 * S_hat * (Sh6[ab] * Sx[ab] - 0.33 * Sh6[aa] * Sx[bb]) +
               2.0 * f2 * (SSh6[ab] * Sx[ab] - 0.33 * SSh6[aa] * Sx[bb]);
      L_cat += Lij[p * 3 + q] * Sx[ab] - 0.33 * Sx[aa] * Lij[q * 4];
    }
  a.lm[n0] = L_cat;
  a.mm[n0] = M_cat;
  // the terms of the 27-point means of les_cs_kernel, each divided once here
  a.lmq[n0] = L_cat / 27.0;
  a.mmq[n0] = M_cat / 27.0;
}

// The 27-point means of les_cs_kernel and les_smooth_kernel run over tiles of
// every node (boundary layers included): a 16x4x4 block of nodes and its
// one-node halo staged in LDS, blocks placed contiguously per XCD (16-node
// rows: an 18-node halo row spans fewer partly used 128 B lines than 8x8x4).  Only
// interior nodes read the halo, and their neighbours always lie in the grid.
#ifndef CLES_MEAN_TILE
#define CLES_MEAN_TILE 16, 4, 4
#endif
constexpr int MT_[3] = {CLES_MEAN_TILE};
constexpr int MTX = MT_[0], MTY = MT_[1], MTZ = MT_[2];
constexpr int MBX = MTX + 2, MBY = MTY + 2, MBZ = MTZ + 2, MNB = MBX * MBY * MBZ;
static_assert(MTX * MTY * MTZ == 256, "one node per thread");

struct NodeTile {
  int i0, j0, k0;  // first node of the block
  int i, j, k;     // this thread's node
  int c0;          // its LDS slot
  bool valid;
};
__device__ __forceinline__ NodeTile node_tile(const Grid &g, int tile, int ntx, int nty) {
  NodeTile T;
  T.i0 = (tile % ntx) * MTX;
  T.j0 = ((tile / ntx) % nty) * MTY;
  T.k0 = (tile / (ntx * nty)) * MTZ;
  const int t = threadIdx.x, lx = t % MTX, ly = (t / MTX) % MTY, lz = t / (MTX * MTY);
  T.i = T.i0 + lx;
  T.j = T.j0 + ly;
  T.k = T.k0 + lz;
  T.c0 = ((lz + 1) * MBY + (ly + 1)) * MBX + (lx + 1);
  T.valid = T.i < g.mx && T.j < g.my && T.k < g.mz;
  return T;
}
// node of halo slot w, or false when it lies outside the grid
__device__ __forceinline__ bool halo_node(const Grid &g, const NodeTile &T, int w, int &kk, int &jj, int &ii) {
  ii = T.i0 - 1 + w % MBX;
  jj = T.j0 - 1 + (w / MBX) % MBY;
  kk = T.k0 - 1 + w / (MBX * MBY);
  return ii >= 0 && jj >= 0 && kk >= 0 && ii < g.mx && jj < g.my && kk < g.mz;
}

// <LM>/<MM> and clipping on every owned node (les.c:889-930, 1146-1161)
__global__ __launch_bounds__(256) void les_cs_kernel(LesArgs a, int ntx, int nty, int ntiles, int per_xcd) {
  __shared__ double B[2][MNB];
  const Grid &g = a.g;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const NodeTile T = node_tile(g, tile, ntx, nty);
  if (a.les != 3) {
    for (int w = threadIdx.x; w < MNB; w += 256) {
      int kk, jj, ii;
      if (!halo_node(g, T, w, kk, jj, ii)) continue;
      const long long m = nid(g, kk, jj, ii);
      B[0][w] = a.lmq[m];  // lm[m] / 27.0 (zero on boundary nodes, never written)
      B[1][w] = a.mmq[m];
    }
    __syncthreads();
  }
  if (!T.valid) return;
  const int i = T.i, j = T.j, k = T.k;
  const long long n = nid(g, k, j, i);
  const double nv = nvat(a, n);
  double cs = a.cs[n];
  if (interior(g, k, j, i)) {
    if (nv > 1.1) {
      cs = 0;
    } else {
      double LM_avg = 0.0, MM_avg = 0.0;
      if (a.les == 3) {
        LM_avg = a.lm[n];
        MM_avg = a.mm[n];
      } else {
#pragma unroll
        for (int c = -1; c <= 1; c++)
#pragma unroll
          for (int b = -1; b <= 1; b++)
#pragma unroll
            for (int aa = -1; aa <= 1; aa++) {
              const int o = T.c0 + (c * MBY + b) * MBX + aa;
              LM_avg += B[0][o];
              MM_avg += B[1][o];
            }
      }
      const double C = LM_avg / (MM_avg + LES_EPS);
      cs = a.les == 1 ? 0.01 : (C < 0.0001 ? 0.0001 : C);
    }
  }
  if (nv > 1.1 && nv < 7.) {
    cs = 0.0;
  } else {
    if (nv > 0.1 && nv < 1.1) cs = 0.01 < cs ? cs : 0.01;
    double t = cs < 0.000 ? 0.000 : cs;
    cs = t < a.max_cs ? t : a.max_cs;
  }
  a.cs[n] = cs;
}

// value of the local vector lCS at an owned node: periodic face copies
// (les.c:1196-1256), everything else the global CS
__device__ __forceinline__ double lcs_at(const LesArgs &a, int k, int j, int i) {
  const long long s = seam_image(a.g, k, j, i);
  return a.cs[s >= 0 ? s : nid(a.g, k, j, i)];
}

// volume-weighted smoothing of Cs over 27 nodes (les.c:1258-1337), then
// nu_t = Cs Delta^2 |S| of the same node (les.c:1536-1573) from the smoothed
// Cs and the |S| the test filter formed
__global__ __launch_bounds__(256) void les_smooth_kernel(LesArgs a, int ntx, int nty, int ntiles, int per_xcd) {
  __shared__ double B[2][MNB];
  const Grid &g = a.g;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const NodeTile T = node_tile(g, tile, ntx, nty);
  for (int w = threadIdx.x; w < MNB; w += 256) {
    int kk, jj, ii;
    if (!halo_node(g, T, w, kk, jj, ii)) continue;
    B[0][w] = a.ivol[nid(g, kk, jj, ii)];  // 1. / aj
    B[1][w] = lcs_at(a, kk, jj, ii);
  }
  __syncthreads();
  if (!T.valid) return;
  const long long n = nid(g, T.k, T.j, T.i);
  const bool in = interior(g, T.k, T.j, T.i);
  const double nv = nvat(a, n);
  double cs;
  if (!in || !(nv < 0.1) || a.les == 3) {
    cs = a.cs[n];
  } else {
    double vs = 0.0, ws = 0.0;
#pragma unroll
    for (int c = -1; c <= 1; c++)
#pragma unroll
      for (int b = -1; b <= 1; b++)
#pragma unroll
        for (int aa = -1; aa <= 1; aa++) {
          const int o = T.c0 + (c * MBY + b) * MBX + aa;
          const double w = B[0][o];
          vs += w * B[1][o];
          ws += w;
        }
    cs = vs / ws;
  }
  a.cs_out[n] = cs;
  a.nut[n] = !in || nv > 0.1 ? 0.0 : cs * a.fil2[n] * a.sabs[n];
}

// nu_t = Cs Delta^2 |S| (les.c:1536-1573)
__global__ __launch_bounds__(256) void les_nut_kernel(LesArgs a, int per_xcd) {
  const Grid &g = a.g;
  const long long n = xcd_node(per_xcd);
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  if (!interior(g, k, j, i) || nvat(a, n) > 0.1) {
    a.nut[n] = 0.0;
    return;
  }
  const long long np = g.np;
  auto ld = [&](long long m, double u[3]) {
    u[0] = a.ua[m];
    u[1] = a.ua[m + np];
    u[2] = a.ua[m + 2 * np];
  };
  double up_i[3], up_j[3], up_k[3], um_i[3], um_j[3], um_k[3];
  ld(n + 1, up_i);
  ld(n + g.sy, up_j);
  ld(n + g.sz, up_k);
  ld(n - 1, um_i);
  ld(n - g.sy, um_j);
  ld(n - g.sz, um_k);
  double gr[9];
  grad9(a, n, up_i, um_i, up_j, um_j, up_k, um_k, gr);
  const double Sabs = sabs_of(gr);
  a.nut[n] = a.cs[n] * a.fil2[n] * Sabs;
}

__global__ __launch_bounds__(256) void fill_kernel(double *x, long long n, double v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

}  // namespace

int launch_les(cles_ctx *c) {
  const cles_params &p = c->p;
  LesArgs a;
  a.g = c->g;
  a.q = c->Q[c->cur];
  a.qs = c->ucat_src >= 0 ? c->Q[c->ucat_src] : nullptr;
  a.csi = c->csi; a.eta = c->eta; a.zet = c->zet; a.aj = c->aj; a.nv = c->nvert;
  a.has_ib = c->has_ib;
  a.grad = c->grad; a.lm = c->lm; a.mm = c->mm; a.lmq = c->lmq; a.mmq = c->mmq; a.ivol = c->ivol; a.sabs = c->sabs; a.ua = c->ua; a.fil2 = c->fil2; a.tfil2 = c->tfil2;
  a.cs = c->cs; a.cs_out = c->cs_alt; a.nut = c->nut;
  a.les = p.les; a.testfilter_ik = p.testfilter_ik; a.max_cs = p.max_cs; a.twoD = p.twoD;
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  const int bs = 256;
  const unsigned gN = (unsigned)((N + bs - 1) / bs);
  const int per_xcd = (int)((gN + 7) / 8);  // blocks per XCD of the node-range kernels
  const unsigned gX = 8u * per_xcd;
  if (c->ucat_les_mixed)
    return set_error(c, CLES_ESTATE, "LES model called twice without a right-hand-side evaluation in between");
  // les.c:139 tests the global tistart, which main.c:890-891 has incremented
  // before the time loop (a fresh run 0 -> 1, a restart from N -> N + 1): in
  // the loop it is never 0, so the Cs = 0 early-out is taken for les == 10
  // only, and the first step of a fresh run computes the dynamic model from
  // an lUcat that no CalcUCat has written yet (zero)
  const int tistart_in_loop = p.tistart + 1;
  const bool early = (c->ti < 2 && tistart_in_loop == 0 && !p.rstart) || p.les == 10;
  a.ucat_updated = 0;
  if (early || p.les == 1) {
    // les.c:139-151: Cs = 0 or 0.03 everywhere; lUcat keeps the velocity of
    // the last CalcUCat
    hipLaunchKernelGGL(fill_kernel, dim3(gN), dim3(bs), 0, c->stream, c->cs, N, early ? 0.0 : 0.03);
    CLES_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(les_grad_kernel, dim3(gX), dim3(bs), 0, c->stream, a, per_xcd);
    CLES_HIP(c, hipGetLastError());
  } else {
    a.ucat_updated = 1;
    hipLaunchKernelGGL(les_grad_kernel, dim3(gX), dim3(bs), 0, c->stream, a, per_xcd);
    CLES_HIP(c, hipGetLastError());
    {
      const int ntx = (c->g.mx - 2 + FTX - 1) / FTX, nty = (c->g.my - 2 + FTY - 1) / FTY,
                ntz = (c->g.mz - 2 + FTZ - 1) / FTZ;
      const int ntiles = ntx * nty * ntz, per_xcd = (ntiles + 7) / 8;
      hipLaunchKernelGGL(les_filter_kernel, dim3(8 * per_xcd), dim3(256), 0, c->stream, a, ntx, nty, ntiles, per_xcd);
    }
    CLES_HIP(c, hipGetLastError());
    {
      const int ntx = (c->g.mx + MTX - 1) / MTX, nty = (c->g.my + MTY - 1) / MTY, ntz = (c->g.mz + MTZ - 1) / MTZ;
      const int ntiles = ntx * nty * ntz, per_xcd = (ntiles + 7) / 8;
      hipLaunchKernelGGL(les_cs_kernel, dim3(8 * per_xcd), dim3(256), 0, c->stream, a, ntx, nty, ntiles, per_xcd);
      CLES_HIP(c, hipGetLastError());
      hipLaunchKernelGGL(les_smooth_kernel, dim3(8 * per_xcd), dim3(256), 0, c->stream, a, ntx, nty, ntiles,
                         per_xcd);
      CLES_HIP(c, hipGetLastError());
    }
    double *t = c->cs;
    c->cs = c->cs_alt;
    c->cs_alt = t;
    a.cs = c->cs;
    c->ucat_les_mixed = 1;
  }
  if (!a.ucat_updated) {  // otherwise les_smooth_kernel has formed nu_t
    hipLaunchKernelGGL(les_nut_kernel, dim3(gX), dim3(bs), 0, c->stream, a, per_xcd);
    CLES_HIP(c, hipGetLastError());
  }
  return CLES_OK;
}

}  // namespace cles

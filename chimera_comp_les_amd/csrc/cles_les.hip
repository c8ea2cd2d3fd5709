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
// field is ever stored: after the loop, lUcat holds the current velocity on
// the nodes the loop visits and the stale one elsewhere (ua_at below).
//
// Kernels: gradients (strain rates, |S| of the loop, and the |S| of nu_t,
// which reads lUcat after the loop) -> 27-point test filter (L_ij M_ij
// contraction) -> 27-point <LM>/<MM> + clipping -> volume-weighted 27-point
// smoothing and nu_t (a separate nu_t kernel on the paths that run no test
// filter).  The gradient and node-range kernels are one thread per node with
// i-contiguous loads; the filter and the means stage node tiles in LDS.
//
// Precision: the model's fields (strain rates, |S|, LM, MM) and its
// arithmetic are in the type T of the form: double for the FP64 drop-in,
// float for the FP32-bulk form (the model is not on the list of quantities
// that form keeps in FP64); Cs and nu_t are stored in double for the stage
// kernel.
#include <type_traits>

#include "cles_internal.hpp"

namespace cles {

namespace {

constexpr double LES_EPS = 0.000001;  // les.c:8

template <typename T>
struct LesArgs {
  Grid g;
  const double *q;   // current state (lQ)
  const double *qs;  // state of the last CalcUCat, or null (lUcat still zero)
  // batched steps: the address of qs read from a device cell that the
  // captured step updates (null: use qs)
  const double *const *qs_cell;
  const double *csi, *eta, *zet, *aj, *nv;
  const double *mpack;  // the nine metrics and aj in blocks of 8 nodes (as in the stage kernel)
  int has_ib;
  // the gradient kernel's eight values per node in blocks of 8 nodes in
  // double, 16 in float ([np/B][4][B][2]: pairs (S11, S12), (S13, S22),
  // (S23, S33) (no trace removal), (|S|, |S| of the velocity lUcat holds
  // after the loop, for nu_t)), so that readers take two values per load
  T *grad;
  T *lm, *mm;
  const double *ivol;   // 1/aj
  double *cs, *cs_out, *nut;
  const double *fil2;   // per-node (pow(1/aj, 1/3))^2, formed once at setup
  const double *tfil2;  // per-node squared test-filter width
  int les, testfilter_ik;
  int twoD;  // 1: the last CalcUCat set u = 0 (rhs.c:1349); the loop's own velocities keep it
  double max_cs;
  int ih, jh, kh;  // -i/j/k_homo_filter
  double *homo;    // homogeneous sums: 5 per (k, j) line, 3 per j plane, 2 global
};

template <typename T> __device__ __forceinline__ double nvat(const LesArgs<T> &a, long long n) {
  return a.has_ib ? a.nv[n] : 0.0;
}
__device__ __forceinline__ bool interior(const Grid &g, int k, int j, int i) {
  return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
// node visited (and overwritten) by the gradient loop of les.c:250-258
template <typename T> __device__ __forceinline__ bool in_loop(const LesArgs<T> &a, int k, int j, int i) {
  return interior(a.g, k, j, i) && !(nvat(a, nid(a.g, k, j, i)) > 1.1);
}
template <typename T> __device__ __forceinline__ void vel(const double *q, long long n, long long np, T u[3]) {
  const double ir = rcp64(q[n]);
  u[0] = T(q[n + np] * ir);
  u[1] = T(q[n + 2 * np] * ir);
  u[2] = T(q[n + 3 * np] * ir);
}
// velocity of the last CalcUCat (zero before the first)
template <typename T> __device__ __forceinline__ void u_stale(const LesArgs<T> &a, long long n, T u[3]) {
  const double *qs = a.qs_cell ? *a.qs_cell : a.qs;
  if (qs) {
    vel(qs, n, a.g.np, u);
    if (a.twoD == 1) u[0] = T(0);
  } else {
    u[0] = u[1] = u[2] = T(0);
  }
}
// velocity lUcat holds at node (k, j, i) after the gradient loop
template <typename T> __device__ __forceinline__ void ua_at(const LesArgs<T> &a, int k, int j, int i, T u[3]) {
  const long long n = nid(a.g, k, j, i);
  if (in_loop(a, k, j, i)) vel(a.q, n, a.g.np, u); else u_stale(a, n, u);
}
// Compute_du_center + Compute_du_dxyz (les.c:31-63)
template <typename T>
__device__ __forceinline__ void grad9(const LesArgs<T> &a, long long n, const T up_i[3], const T um_i[3],
                                      const T up_j[3], const T um_j[3], const T up_k[3], const T um_k[3], T gr[9]) {
  const long long np = a.g.np;
  T dudc = (up_i[0] - um_i[0]) * T(0.5), dvdc = (up_i[1] - um_i[1]) * T(0.5), dwdc = (up_i[2] - um_i[2]) * T(0.5);
  T dudz = (up_k[0] - um_k[0]) * T(0.5), dvdz = (up_k[1] - um_k[1]) * T(0.5), dwdz = (up_k[2] - um_k[2]) * T(0.5);
  T dude = (up_j[0] - um_j[0]) * T(0.5), dvde = (up_j[1] - um_j[1]) * T(0.5), dwde = (up_j[2] - um_j[2]) * T(0.5);
  // the double form reads the blocked metric pack (five 16-byte loads; its
  // gradients 0.94 -> 0.89 ms at 256^3), the float form the planes (0.76 ms;
  // the pack, or a run-time choice between the two, made it slower)
  T csi0, csi1, csi2, eta0, eta1, eta2, zet0, zet1, zet2, ajc;
  if constexpr (sizeof(T) == 8) {
    const double2 *pk = reinterpret_cast<const double2 *>(a.mpack + (n >> 3) * 80 + (n & 7) * 2);
    const double2 x0 = pk[0], x1 = pk[8], x2 = pk[16], x3 = pk[24], x4 = pk[32];
    csi0 = x0.x; csi1 = x0.y; csi2 = x1.x;
    eta0 = x1.y; eta1 = x2.x; eta2 = x2.y;
    zet0 = x3.x; zet1 = x3.y; zet2 = x4.x;
    ajc = x4.y;
  } else {
    csi0 = T(a.csi[n]); csi1 = T(a.csi[n + np]); csi2 = T(a.csi[n + 2 * np]);
    eta0 = T(a.eta[n]); eta1 = T(a.eta[n + np]); eta2 = T(a.eta[n + 2 * np]);
    zet0 = T(a.zet[n]); zet1 = T(a.zet[n + np]); zet2 = T(a.zet[n + 2 * np]);
    ajc = T(a.aj[n]);
  }
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
// |S| with the trace term 2/3 div subtracted from the diagonal (les.c:268-274)
template <typename T> __device__ __forceinline__ T sabs_of(const T g[9]) {
  T Sdd = T(2. / 3.) * (g[0] + g[4] + g[8]);
  T Sxx = T(0.5) * (g[0] + g[0]) - Sdd, Sxy = T(0.5) * (g[1] + g[3]), Sxz = T(0.5) * (g[2] + g[6]);
  T Syx = Sxy, Syy = T(0.5) * (g[4] + g[4]) - Sdd, Syz = T(0.5) * (g[5] + g[7]);
  T Szx = Sxz, Szy = Syz, Szz = T(0.5) * (g[8] + g[8]) - Sdd;
  return sqrt(T(2.0) * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz + Szx * Szx +
                        Szy * Szy + Szz * Szz));
}

// pair p (values 2p, 2p + 1) of node n in the gradient pack, and value c:
// blocks of GB nodes, so that one pair of a block fills a 128-byte line
template <typename T> struct Pair2;
template <> struct Pair2<double> { using type = double2; };
template <> struct Pair2<float> { using type = float2; };
template <typename T> __device__ __forceinline__ typename Pair2<T>::type *gpair(T *pack, long long n, int p) {
  constexpr int GB = 64 / sizeof(T);
  return reinterpret_cast<typename Pair2<T>::type *>(pack + ((n / GB) * 4 + p) * (2 * GB) + (n % GB) * 2);
}
template <typename T> __device__ __forceinline__ T gval(const T *pack, long long n, int c) {
  constexpr int GB = 64 / sizeof(T);
  return pack[((n / GB) * 4 + (c >> 1)) * (2 * GB) + (n % GB) * 2 + (c & 1)];
}

// Node-range kernels: each XCD (blockIdx % 8) walks a contiguous range of
// blocks, so the k +- 1 neighbours of its nodes stay in its own L2.
__device__ __forceinline__ long long xcd_node(int per_xcd) {
  const int b = blockIdx.x;
  return (long long)((b & 7) * per_xcd + (b >> 3)) * blockDim.x + threadIdx.x;
}

// Velocity gradients of the loop les.c:250-309 (mixed current / previous
// velocities, see above), stored as the six strain-rate components the test
// filter uses (S_ab = (du_a/dx_b + du_b/dx_a)/2, les.c:520-522) and |S|; and
// the |S| of nu_t (les.c:1536-1573), whose neighbours hold lUcat after the
// loop: the current velocity where the loop has visited them.
template <typename T>
__global__ __launch_bounds__(256) void les_grad_kernel(LesArgs<T> a, int per_xcd) {
  const Grid &g = a.g;
  const long long n = xcd_node(per_xcd);
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  const long long np = g.np;
  using P2 = typename Pair2<T>::type;
  if (!in_loop(a, k, j, i)) {
#pragma unroll
    for (int p = 0; p < 4; p++) *gpair(a.grad, n, p) = P2{T(0), T(0)};
    return;
  }
  T up_i[3], up_j[3], up_k[3], um_i[3], um_j[3], um_k[3];
  u_stale(a, n + 1, up_i);
  u_stale(a, n + g.sy, up_j);
  u_stale(a, n + g.sz, up_k);
  ua_at(a, k, j, i - 1, um_i);
  ua_at(a, k, j - 1, i, um_j);
  ua_at(a, k - 1, j, i, um_k);
  T gr[9];
  grad9(a, n, up_i, um_i, up_j, um_j, up_k, um_k, gr);
  *gpair(a.grad, n, 0) = P2{T(0.5) * (gr[0] + gr[0]), T(0.5) * (gr[1] + gr[3])};
  *gpair(a.grad, n, 1) = P2{T(0.5) * (gr[2] + gr[6]), T(0.5) * (gr[4] + gr[4])};
  *gpair(a.grad, n, 2) = P2{T(0.5) * (gr[5] + gr[7]), T(0.5) * (gr[8] + gr[8])};
  const T s_loop = sabs_of(gr);
  // nu_t: the +1 neighbours as lUcat holds them after the loop
  if (in_loop(a, k, j, i + 1)) vel(a.q, n + 1, np, up_i);
  if (in_loop(a, k, j + 1, i)) vel(a.q, n + g.sy, np, up_j);
  if (in_loop(a, k + 1, j, i)) vel(a.q, n + g.sz, np, up_k);
  grad9(a, n, up_i, um_i, up_j, um_j, up_k, um_k, gr);
  *gpair(a.grad, n, 3) = P2{s_loop, sabs_of(gr)};
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
// contraction of les.c:505-887).  The weights 1/(|dk|+1) * 1/(|dj|+1) *
// 1/(|di|+1) are 1, 1/2, 1/4 or 1/8, the product of one factor per axis, so
// the 27-point weighted sum of a quantity is 0.5 P(k-1) + P(k) + 0.5 P(k+1)
// with P(k) the 9-point weighted sum over plane k; the weights sum to 8, so
// vsum / wsum is vsum * 0.125.  A workgroup owns a 16 x 16 column of
// interior nodes and walks a chunk of kc planes: each plane's node data with
// a one-node halo is staged in LDS once, every thread forms the plane sums
// of its node from it, and keeps the sums of the two planes before in
// registers.  The node data and sums of the contraction:
//   L_ij S_ij: sums of rho, rhoU, rhoV, rhoW and the six rho u_a u_b
//     products (staged: rho, rhoU, rhoV, rhoW and the velocity lUcat holds
//     after the gradient loop); writes LM;
//   M_ij S_ij: sums of rho, the six strain-rate components as the filter
//     sees them (periodic face copies of lSx..lSz, les.c:316-377, zero on
//     other boundary nodes) and their products with rho |S| (|S| without a
//     periodic copy); writes MM.
// PART 2 forms both (22 sums); PART 0 the L part and PART 1 the M part, two
// launches that share no data.  In double the 22 sums need 221 VGPRs (2
// waves per SIMD) and each part fits 128 (4 waves); in float all 22 fit 128.
// Each sum is formed with the same operations, in the same order, in every
// part.
constexpr int FTX = 16, FTY = 16;
constexpr int FBX = FTX + 2, FBY = FTY + 2, FNB = FBX * FBY;
constexpr double FILTER_INV_WSUM = 0.125;
static_assert(FTX * FTY == 256, "one column per thread");

template <typename T, int PART>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void les_filter_kernel(LesArgs<T> a, int ntx, int nty, int kc, int ntiles, int per_xcd) {
  constexpr bool DL = PART != 1, DM = PART != 0;  // forms L_ij S_ij, M_ij S_ij
  constexpr int QS = DL ? 7 : 1;                   // first staged strain-rate component
  constexpr int NQ = DM ? QS + 7 : 7;              // staged values per node
  constexpr int PS = DL ? 10 : 1;                  // first strain-rate sum
  constexpr int NS = DM ? PS + 12 : 10;            // plane sums per node
  __shared__ T B[NQ][FNB];
  const Grid &g = a.g;
  const long long np = g.np;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);  // contiguous tiles per XCD
  if (tile >= ntiles) return;
  // x fastest, then y, then the plane chunk
  const int i0 = 1 + (tile % ntx) * FTX, j0 = 1 + ((tile / ntx) % nty) * FTY;
  const int k0 = 1 + (tile / (ntx * nty)) * kc;
  const int k1 = k0 + kc - 1 < g.mz - 2 ? k0 + kc - 1 : g.mz - 2;  // last plane of the chunk
  const int t = threadIdx.x;
  const int lx = t % FTX, ly = t / FTX;
  const int i = i0 + lx, j = j0 + ly;
  const bool col = i <= g.mx - 2 && j <= g.my - 2;
  const int c0 = (ly + 1) * FBX + (lx + 1);  // slot of the column's node
  T Pm[NS], P0[NS], Sx0[6];
  bool act0 = false;
  for (int kk = k0 - 1; kk <= k1 + 1; kk++) {
    // ---- stage plane kk
    for (int w = t; w < FNB; w += 256) {
      const int ii = i0 - 1 + w % FBX, jj = j0 - 1 + w / FBX;
      if (ii > g.mx - 1 || jj > g.my - 1) continue;
      const long long m = nid(g, kk, jj, ii);
      const double r = a.q[m];
      B[0][w] = T(r);
      if (DL) {
        const double ru = a.q[m + np], rv = a.q[m + 2 * np], rw = a.q[m + 3 * np];
        B[1][w] = T(ru);
        B[2][w] = T(rv);
        B[3][w] = T(rw);
        T u[3];
        if (in_loop(a, kk, jj, ii)) {
          const double ir = rcp64(r);
          u[0] = T(ru * ir);
          u[1] = T(rv * ir);
          u[2] = T(rw * ir);
        } else {
          u_stale(a, m, u);
        }
        B[4][w] = u[0];
        B[5][w] = u[1];
        B[6][w] = u[2];
      }
      if (DM) {
        using P2 = typename Pair2<T>::type;
        if (interior(g, kk, jj, ii)) {
          const P2 x0 = *gpair(a.grad, m, 0), x1 = *gpair(a.grad, m, 1), x2 = *gpair(a.grad, m, 2), x3 = *gpair(a.grad, m, 3);
          B[QS][w] = x0.x;
          B[QS + 1][w] = x0.y;
          B[QS + 2][w] = x1.x;
          B[QS + 3][w] = x1.y;
          B[QS + 4][w] = x2.x;
          B[QS + 5][w] = x2.y;
          B[QS + 6][w] = x3.x;
        } else {
          const long long ms = seam_image(g, kk, jj, ii);
          P2 x0{T(0), T(0)}, x1{T(0), T(0)}, x2{T(0), T(0)};
          if (ms >= 0) {
            x0 = *gpair(a.grad, ms, 0);
            x1 = *gpair(a.grad, ms, 1);
            x2 = *gpair(a.grad, ms, 2);
          }
          B[QS][w] = x0.x;
          B[QS + 1][w] = x0.y;
          B[QS + 2][w] = x1.x;
          B[QS + 3][w] = x1.y;
          B[QS + 4][w] = x2.x;
          B[QS + 5][w] = x2.y;
          B[QS + 6][w] = T(0);  // |S| without a periodic copy
        }
      }
    }
    const bool act1 = col && kk >= 1 && kk <= g.mz - 2 && !(nvat(a, nid(g, kk, j, i)) > 1.1);
    // the node's own strain rates (it is interior when active): without the
    // M part they are read here, and used when the node's sums are complete,
    // one plane later
    T Sx1[6];
    if (!DM) {
      const long long n1 = act1 ? nid(g, kk, j, i) : 0;
      using P2 = typename Pair2<T>::type;
      P2 x0{T(0), T(0)}, x1{T(0), T(0)}, x2{T(0), T(0)};
      if (act1) {
        x0 = *gpair(a.grad, n1, 0);
        x1 = *gpair(a.grad, n1, 1);
        x2 = *gpair(a.grad, n1, 2);
      }
      Sx1[0] = x0.x;
      Sx1[1] = x0.y;
      Sx1[2] = x1.x;
      Sx1[3] = x1.y;
      Sx1[4] = x2.x;
      Sx1[5] = x2.y;
    }
    __syncthreads();
    // ---- the plane sums of this column's node
    T P1[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) P1[q] = T(0);
    // one node of the 3x3 neighbourhood at a time: its LDS reads are in
    // flight together, not those of a row or of all nine nodes (registers:
    // a row at a time spills in the M part of the double form, and runs the
    // LES model 0.06 ms slower at 256^3)
#pragma unroll 1
    for (int Qd = 0; Qd < 3; Qd++)
#pragma unroll 1
      for (int Pd = 0; Pd < 3; Pd++) {
        const int o = c0 + (Qd - 1) * FBX + (Pd - 1);
        const T w = T((Qd == 1 ? 1.0 : 0.5) * (Pd == 1 ? 1.0 : 0.5));
        const T rho = B[0][o];
        P1[0] = fma(w, rho, P1[0]);
        if (DL) {
          const T ru = B[1][o], rv = B[2][o], rw = B[3][o];
          const T u0 = B[4][o], u1 = B[5][o], u2 = B[6][o];
          P1[1] = fma(w, ru, P1[1]);
          P1[2] = fma(w, rv, P1[2]);
          P1[3] = fma(w, rw, P1[3]);
          P1[4] = fma(w, ru * u0, P1[4]);
          P1[5] = fma(w, ru * u1, P1[5]);
          P1[6] = fma(w, ru * u2, P1[6]);
          P1[7] = fma(w, rv * u1, P1[7]);
          P1[8] = fma(w, rv * u2, P1[8]);
          P1[9] = fma(w, rw * u2, P1[9]);
        }
        if (DM) {
          const T rS = rho * B[QS + 6][o];
#pragma unroll
          for (int c = 0; c < 6; c++) {
            const T Sv = B[QS + c][o];
            P1[PS + c] = fma(w, Sv, P1[PS + c]);
            P1[PS + 6 + c] = fma(w, Sv * rS, P1[PS + 6 + c]);
          }
        }
      }
    if (DM) {
#pragma unroll
      for (int c = 0; c < 6; c++) Sx1[c] = B[QS + c][c0];
    }
    __syncthreads();  // the next plane overwrites the buffer
    // ---- node (kk - 1, j, i): its 27-point sums, L_ij and M_ij
    const int k = kk - 1;
    if (col && k >= k0) {
      const long long n0 = nid(g, k, j, i);
      if (!act0) {
        if (DL) a.lm[n0] = T(0);
        if (DM) a.mm[n0] = T(0);
      } else {
        T sm[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) sm[q] = fma(T(0.5), Pm[q], P0[q]) + T(0.5) * P1[q];
        const T iw = T(FILTER_INV_WSUM);
        const T _rho = sm[0] * iw;
        // full 3x3 tensors in row-major order from the six symmetric components
        const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        if (DL) {
          const T _rhou = sm[1] * iw, _rhov = sm[2] * iw, _rhow = sm[3] * iw;
          const T irho = T(1) / _rho;
          T Lij[9];
          Lij[0] = sm[4] * iw - _rhou * _rhou * irho;
          Lij[1] = sm[5] * iw - _rhou * _rhov * irho;
          Lij[2] = sm[6] * iw - _rhou * _rhow * irho;
          Lij[3] = sm[5] * iw - _rhov * _rhou * irho;
          Lij[4] = sm[7] * iw - _rhov * _rhov * irho;
          Lij[5] = sm[8] * iw - _rhov * _rhow * irho;
          Lij[6] = sm[6] * iw - _rhow * _rhou * irho;
          Lij[7] = sm[8] * iw - _rhow * _rhov * irho;
          Lij[8] = sm[9] * iw - _rhow * _rhow * irho;
          T L_cat = 0;
#pragma unroll
          for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
              const int ab = sym[p * 3 + q], aa = sym[p * 4];
              L_cat += Lij[p * 3 + q] * Sx0[ab] - T(0.33) * Sx0[aa] * Lij[q * 4];
            }
          a.lm[n0] = L_cat;
        }
        if (DM) {
          T Sh6[6], SSh6[6];
#pragma unroll
          for (int c = 0; c < 6; c++) {
            Sh6[c] = sm[PS + c] * iw;
            SSh6[c] = sm[PS + 6 + c] * iw;
          }
          T S_hat = 0;
#pragma unroll
          for (int c = 0; c < 9; c++) S_hat += Sh6[sym[c]] * Sh6[sym[c]];
          S_hat = sqrt(T(2) * S_hat);
          T M_cat = 0;
          const T tf2 = T(a.tfil2[n0]), f2 = T(a.fil2[n0]);
          const T m1 = T(-2.0) * _rho * tf2 * S_hat, m2 = T(2.0) * f2;
#pragma unroll
          for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
              const int ab = sym[p * 3 + q], aa = sym[p * 4], bb = sym[q * 4];
              M_cat += m1 * (Sh6[ab] * Sx0[ab] - T(0.33) * Sh6[aa] * Sx0[bb]) +
                       m2 * (SSh6[ab] * Sx0[ab] - T(0.33) * SSh6[aa] * Sx0[bb]);
            }
          a.mm[n0] = M_cat;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NS; q++) {
      Pm[q] = P0[q];
      P0[q] = P1[q];
    }
#pragma unroll
    for (int c = 0; c < 6; c++) Sx0[c] = Sx1[c];
    act0 = act1;
  }
}

// The 27-point means of les_cs_kernel and les_smooth_kernel run over tiles of
// every node (boundary layers included): a 16x4x4 block of nodes and its
// one-node halo staged in LDS, blocks placed contiguously per XCD (16-node
// rows: an 18-node halo row spans fewer partly used 128 B lines than 8x8x4).  Only
// interior nodes read the halo, and their neighbours always lie in the grid.
constexpr int MTX = 16, MTY = 4, MTZ = 4;
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

// Homogeneous-direction averages of LM and MM (les.c:934-1116).  One thread
// per (k, j) line sums along i (the fluid nodes, with their count, and every
// node), one thread per j plane sums its lines over k, one thread the planes.
// homo layout: line sums [5][mz][my], plane means/sums [3][my], global [2].
template <typename T>
__global__ __launch_bounds__(256) void homo_line_kernel(LesArgs<T> a) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long L = (long long)g.my * g.mz;
  if (t >= L) return;
  const int j = (int)(t % g.my), k = (int)(t / g.my);
  double sl = 0, sm = 0, cnt = 0, al = 0, am = 0;
  if (k >= 1 && k <= g.mz - 2 && j >= 1 && j <= g.my - 2) {
    for (int i = 1; i <= g.mx - 2; i++) {
      const long long n = nid(g, k, j, i);
      const double l = double(a.lm[n]), m = double(a.mm[n]);
      if (nvat(a, n) < 0.1) {
        sl += l;
        sm += m;
        cnt += 1;
      }
      al += l;
      am += m;
    }
  }
  double *h = a.homo;
  h[t] = sl;
  h[L + t] = sm;
  h[2 * L + t] = cnt;
  h[3 * L + t] = al;
  h[4 * L + t] = am;
}
template <typename T>
__global__ __launch_bounds__(256) void homo_plane_kernel(LesArgs<T> a) {
  const Grid &g = a.g;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= g.my) return;
  const long long L = (long long)g.my * g.mz;
  const double *h = a.homo;
  double sl = 0, sm = 0, cnt = 0, al = 0, am = 0;
  for (int k = 1; k <= g.mz - 2; k++) {
    const long long t = (long long)k * g.my + j;
    sl += h[t];
    sm += h[L + t];
    cnt += h[2 * L + t];
    al += h[3 * L + t];
    am += h[4 * L + t];
  }
  double *P = a.homo + 5 * L;
  if (cnt > 0) {
    sl /= cnt;
    sm /= cnt;
  }
  P[j] = sl;                  // <LM>_j over the fluid nodes of plane j
  P[g.my + j] = sm;           // <MM>_j
  P[2 * g.my + j] = al;       // sum of LM over the plane's interior
  P[3 * g.my + j] = am;
}
template <typename T>
__global__ void homo_global_kernel(LesArgs<T> a) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Grid &g = a.g;
  const long long L = (long long)g.my * g.mz;
  const double *P = a.homo + 5 * L;
  double al = 0, am = 0;
  for (int j = 1; j <= g.my - 2; j++) {
    al += P[2 * g.my + j];
    am += P[3 * g.my + j];
  }
  a.homo[5 * L + 4 * g.my] = al;
  a.homo[5 * L + 4 * g.my + 1] = am;
}
// Cs after the homogeneous averages of the reference, in its order: the
// (i, k) plane mean on interior nodes, the i-line mean on interior fluid
// nodes, the global ratio on every node
template <typename T>
__device__ __forceinline__ double homogeneous_cs(const LesArgs<T> &a, int k, int j, int i, double nv, double cs) {
  const Grid &g = a.g;
  const long long L = (long long)g.my * g.mz;
  const bool in = interior(g, k, j, i);
  const double *P = a.homo + 5 * L;
  if (a.ih && a.kh && in) cs = 0.5 * P[j] / (P[g.my + j] + LES_EPS);
  if (a.ih && !a.kh && in && nv < 0.1) {
    const long long t = (long long)k * g.my + j;
    double sl = a.homo[t], sm = a.homo[L + t];
    const double cnt = a.homo[2 * L + t];
    if (cnt > 0) {
      sl /= cnt;
      sm /= cnt;
    }
    cs = 0.5 * sl / (sm + LES_EPS);
  }
  if (a.ih && a.jh && a.kh) cs = P[4 * g.my] / P[4 * g.my + 1];
  return cs;
}

// <LM>/<MM> (integrate_t: the 27-point mean, les.c:66-78) and clipping on
// every owned node (les.c:889-930, 1146-1161)
template <typename T>
__global__ __launch_bounds__(256) void les_cs_kernel(LesArgs<T> a, int ntx, int nty, int ntiles, int per_xcd) {
  __shared__ T B[2][MNB];
  const Grid &g = a.g;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const NodeTile Tn = node_tile(g, tile, ntx, nty);
  const bool homo = a.ih || a.jh || a.kh;
  if (a.les != 3 && !homo) {
    for (int w = threadIdx.x; w < MNB; w += 256) {
      int kk, jj, ii;
      if (!halo_node(g, Tn, w, kk, jj, ii)) continue;
      const long long m = nid(g, kk, jj, ii);
      B[0][w] = interior(g, kk, jj, ii) ? a.lm[m] : T(0);  // LM is zero off the loop's nodes
      B[1][w] = interior(g, kk, jj, ii) ? a.mm[m] : T(0);
    }
    __syncthreads();
  }
  if (!Tn.valid) return;
  const int i = Tn.i, j = Tn.j, k = Tn.k;
  const long long n = nid(g, k, j, i);
  const double nv = nvat(a, n);
  double cs = a.cs[n];
  if (interior(g, k, j, i)) {
    if (nv > 1.1) {
      cs = 0;
    } else {
      T LM_avg = 0, MM_avg = 0;
      if (a.les == 3 || homo) {
        LM_avg = a.lm[n];
        MM_avg = a.mm[n];
      } else {
#pragma unroll
        for (int c = -1; c <= 1; c++)
#pragma unroll
          for (int b = -1; b <= 1; b++)
#pragma unroll
            for (int aa = -1; aa <= 1; aa++) {
              const int o = Tn.c0 + (c * MBY + b) * MBX + aa;
              LM_avg += B[0][o];
              MM_avg += B[1][o];
            }
        LM_avg *= T(1.0 / 27.0);
        MM_avg *= T(1.0 / 27.0);
      }
      const double C = double(LM_avg) / (double(MM_avg) + LES_EPS);
      cs = a.les == 1 ? 0.01 : (C < 0.0001 ? 0.0001 : C);
    }
  }
  if (homo) cs = homogeneous_cs(a, k, j, i, nv, cs);
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
template <typename T> __device__ __forceinline__ double lcs_at(const LesArgs<T> &a, int k, int j, int i) {
  const long long s = seam_image(a.g, k, j, i);
  return a.cs[s >= 0 ? s : nid(a.g, k, j, i)];
}

// volume-weighted smoothing of Cs over 27 nodes (les.c:1258-1337), then
// nu_t = Cs Delta^2 |S| of the same node (les.c:1536-1573) from the smoothed
// Cs and the |S| of the gradient kernel
template <typename T>
__global__ __launch_bounds__(256) void les_smooth_kernel(LesArgs<T> a, int ntx, int nty, int ntiles, int per_xcd) {
  __shared__ double B[2][MNB];
  const Grid &g = a.g;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;
  const NodeTile Tn = node_tile(g, tile, ntx, nty);
  for (int w = threadIdx.x; w < MNB; w += 256) {
    int kk, jj, ii;
    if (!halo_node(g, Tn, w, kk, jj, ii)) continue;
    B[0][w] = a.ivol[nid(g, kk, jj, ii)];  // 1. / aj
    B[1][w] = lcs_at(a, kk, jj, ii);
  }
  __syncthreads();
  if (!Tn.valid) return;
  const long long n = nid(g, Tn.k, Tn.j, Tn.i);
  const bool in = interior(g, Tn.k, Tn.j, Tn.i);
  const double nv = nvat(a, n);
  double cs;
  if (!in || !(nv < 0.1) || a.les == 3 || a.ih || a.jh || a.kh) {
    cs = a.cs[n];
  } else {
    double vs = 0.0, ws = 0.0;
#pragma unroll
    for (int c = -1; c <= 1; c++)
#pragma unroll
      for (int b = -1; b <= 1; b++)
#pragma unroll
        for (int aa = -1; aa <= 1; aa++) {
          const int o = Tn.c0 + (c * MBY + b) * MBX + aa;
          const double w = B[0][o];
          vs += w * B[1][o];
          ws += w;
        }
    cs = vs * rcp64(ws);
  }
  a.cs_out[n] = cs;
  a.nut[n] = !in || nv > 0.1 ? 0.0 : cs * a.fil2[n] * double(gval(a.grad, n, 7));
}

// The same two 27-point means as plane marches (used for the dynamic model
// without homogeneous directions; the node tiles above serve the others): a workgroup owns a 16 x 16 column of nodes
// (boundary layers included) and walks a chunk of kc planes; each plane's two
// node values with a one-node halo are staged in LDS once, every thread forms
// the 9-point sums of its node over the plane, and the node's 27-point sum is
// P(k-1) + P(k) + P(k+1) from the sums of three planes kept in registers.
//   PART 0: Cs from the means of LM and MM (les_cs_kernel);
//   PART 1: the volume-weighted mean of Cs, then nu_t (les_smooth_kernel).
template <typename T, int PART>
__global__ __launch_bounds__(256) void les_mean_kernel(LesArgs<T> a, int ntx, int nty, int kc, int ntiles, int per_xcd) {
  using V = typename std::conditional<PART == 0, T, double>::type;
  __shared__ V B[2][FNB];
  const Grid &g = a.g;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);  // contiguous tiles per XCD
  if (tile >= ntiles) return;
  const int i0 = (tile % ntx) * FTX, j0 = ((tile / ntx) % nty) * FTY;
  const int k0 = (tile / (ntx * nty)) * kc;
  const int k1 = k0 + kc - 1 < g.mz - 1 ? k0 + kc - 1 : g.mz - 1;  // last plane of the chunk
  const int t = threadIdx.x;
  const int lx = t % FTX, ly = t / FTX;
  const int i = i0 + lx, j = j0 + ly;
  const bool col = i < g.mx && j < g.my;
  const int c0 = (ly + 1) * FBX + (lx + 1);  // slot of the column's node
  V Pm[2] = {V(0), V(0)}, P0[2] = {V(0), V(0)};
  // plane kk's slots in registers: loaded one plane ahead, so that their
  // loads are in flight during the sums of the plane before
  constexpr int NSL = (FNB + 255) / 256;
  V r0[NSL], r1[NSL];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int s = 0; s < NSL; s++) {
      const int w = t + 256 * s;
      r0[s] = V(0);
      r1[s] = V(0);
      const int ii = i0 - 1 + w % FBX, jj = j0 - 1 + w / FBX;
      if (w >= FNB || kk < 0 || kk > g.mz - 1 || ii < 0 || jj < 0 || ii > g.mx - 1 || jj > g.my - 1) continue;
      const long long m = nid(g, kk, jj, ii);
      if (PART == 0) {
        const bool in = interior(g, kk, jj, ii);
        r0[s] = in ? a.lm[m] : T(0);  // LM is zero off the loop's nodes
        r1[s] = in ? a.mm[m] : T(0);
      } else {
        r0[s] = a.ivol[m];  // 1. / aj
        r1[s] = lcs_at(a, kk, jj, ii);
      }
    }
  };
  fetch(k0 - 1);
  for (int kk = k0 - 1; kk <= k1 + 1; kk++) {
    const bool st = kk >= 0 && kk <= g.mz - 1;  // uniform
    if (st) {
#pragma unroll
      for (int s = 0; s < NSL; s++) {
        const int w = t + 256 * s;
        if (w < FNB) {
          B[0][w] = r0[s];
          B[1][w] = r1[s];
        }
      }
    }
    __syncthreads();
    if (kk <= k1) fetch(kk + 1);
    V P1[2] = {V(0), V(0)};
    if (st && col && i >= 1 && j >= 1 && i <= g.mx - 2 && j <= g.my - 2) {
#pragma unroll
      for (int b = -1; b <= 1; b++)
#pragma unroll
        for (int aa = -1; aa <= 1; aa++) {
          const int o = c0 + b * FBX + aa;
          if (PART == 0) {
            P1[0] += B[0][o];
            P1[1] += B[1][o];
          } else {
            const double w = B[0][o];
            P1[0] += w * B[1][o];
            P1[1] += w;
          }
        }
    }
    __syncthreads();  // the next plane overwrites the buffer
    const int k = kk - 1;
    if (col && k >= k0 && k <= k1) {
      const long long n = nid(g, k, j, i);
      const bool in = interior(g, k, j, i);
      const double nv = nvat(a, n);
      if (PART == 0) {
        double cs = a.cs[n];
        if (in) {
          if (nv > 1.1) {
            cs = 0;
          } else if (a.les == 1) {
            cs = 0.01;
          } else {
            const T LM_avg = ((Pm[0] + P0[0]) + P1[0]) * T(1.0 / 27.0);
            const T MM_avg = ((Pm[1] + P0[1]) + P1[1]) * T(1.0 / 27.0);
            const double C = double(LM_avg) / (double(MM_avg) + LES_EPS);
            cs = C < 0.0001 ? 0.0001 : C;
          }
        }
        if (nv > 1.1 && nv < 7.) {
          cs = 0.0;
        } else {
          if (nv > 0.1 && nv < 1.1) cs = 0.01 < cs ? cs : 0.01;
          double tt = cs < 0.000 ? 0.000 : cs;
          cs = tt < a.max_cs ? tt : a.max_cs;
        }
        a.cs[n] = cs;
      } else {
        double cs;
        if (!in || !(nv < 0.1)) {
          cs = a.cs[n];
        } else {
          const double vs = (Pm[0] + P0[0]) + P1[0], ws = (Pm[1] + P0[1]) + P1[1];
          cs = vs * rcp64(ws);
        }
        a.cs_out[n] = cs;
        a.nut[n] = !in || nv > 0.1 ? 0.0 : cs * a.fil2[n] * double(gval(a.grad, n, 7));
      }
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      Pm[q] = P0[q];
      P0[q] = P1[q];
    }
  }
}

// nu_t = Cs Delta^2 |S| (les.c:1536-1573) on the paths without the gradient
// loop (constant Cs, Cs = 0): lUcat is the velocity of the last CalcUCat
template <typename T>
__global__ __launch_bounds__(256) void les_nut_kernel(LesArgs<T> a, int per_xcd) {
  const Grid &g = a.g;
  const long long n = xcd_node(per_xcd);
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  if (!interior(g, k, j, i) || nvat(a, n) > 0.1) {
    a.nut[n] = 0.0;
    return;
  }
  T up_i[3], up_j[3], up_k[3], um_i[3], um_j[3], um_k[3];
  u_stale(a, n + 1, up_i);
  u_stale(a, n + g.sy, up_j);
  u_stale(a, n + g.sz, up_k);
  u_stale(a, n - 1, um_i);
  u_stale(a, n - g.sy, um_j);
  u_stale(a, n - g.sz, um_k);
  T gr[9];
  grad9(a, n, up_i, um_i, up_j, um_j, up_k, um_k, gr);
  a.nut[n] = a.cs[n] * a.fil2[n] * double(sabs_of(gr));
}

__global__ __launch_bounds__(256) void fill_kernel(double *x, long long n, double v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

template <typename T>
int launch_les_t(cles_ctx *c, hipStream_t stream, const double *const *qs_cell) {
  const cles_params &p = c->p;
  LesArgs<T> a;
  a.g = c->g;
  a.q = c->Q[c->cur];
  a.qs = c->ucat_src >= 0 ? c->Q[c->ucat_src] : nullptr;
  a.qs_cell = qs_cell;
  a.csi = c->csi; a.eta = c->eta; a.zet = c->zet; a.aj = c->aj; a.nv = c->nvert;
  a.mpack = c->mpack;
  a.has_ib = c->has_ib;
  a.grad = (T *)c->grad; a.lm = (T *)c->lm; a.mm = (T *)c->mm;
  a.ivol = c->ivol; a.fil2 = c->fil2; a.tfil2 = c->tfil2;
  a.cs = c->cs; a.cs_out = c->cs_alt; a.nut = c->nut;
  a.les = p.les; a.testfilter_ik = p.testfilter_ik; a.max_cs = p.max_cs; a.twoD = p.twoD;
  a.ih = p.i_homo_filter; a.jh = p.j_homo_filter; a.kh = p.k_homo_filter;
  a.homo = c->homo;
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
  if (early || p.les == 1) {
    // les.c:139-151: Cs = 0 or 0.03 everywhere; lUcat keeps the velocity of
    // the last CalcUCat
    hipLaunchKernelGGL(fill_kernel, dim3(gN), dim3(bs), 0, stream, c->cs, N, early ? 0.0 : 0.03);
    CLES_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(les_nut_kernel<T>, dim3(gX), dim3(bs), 0, stream, a, per_xcd);
    CLES_HIP(c, hipGetLastError());
    return CLES_OK;
  }
  hipLaunchKernelGGL(les_grad_kernel<T>, dim3(gX), dim3(bs), 0, stream, a, per_xcd);
  CLES_HIP(c, hipGetLastError());
  {
    // chunks of 32 planes, or fewer where that leaves fewer workgroups than
    // the chip holds at once (256 CUs x 4): a chunk is a serial march
    const int ntx = (c->g.mx - 2 + FTX - 1) / FTX, nty = (c->g.my - 2 + FTY - 1) / FTY;
    int kc = 32;
    while (kc > 4 && ntx * nty * ((c->g.mz - 2 + kc - 1) / kc) < 1024) kc /= 2;
    const int ntiles = ntx * nty * ((c->g.mz - 2 + kc - 1) / kc), pxf = (ntiles + 7) / 8;
    if constexpr (sizeof(T) == 8) {
      hipLaunchKernelGGL((les_filter_kernel<T, 0>), dim3(8 * pxf), dim3(256), 0, stream, a, ntx, nty, kc, ntiles, pxf);
      CLES_HIP(c, hipGetLastError());
      hipLaunchKernelGGL((les_filter_kernel<T, 1>), dim3(8 * pxf), dim3(256), 0, stream, a, ntx, nty, kc, ntiles, pxf);
    } else {
      hipLaunchKernelGGL((les_filter_kernel<T, 2>), dim3(8 * pxf), dim3(256), 0, stream, a, ntx, nty, kc, ntiles, pxf);
    }
  }
  CLES_HIP(c, hipGetLastError());
  if (a.ih) {  // the averaged branches (-j or -k alone average nothing)
    const long long L = (long long)c->g.my * c->g.mz;
    hipLaunchKernelGGL(homo_line_kernel<T>, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, stream, a);
    CLES_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(homo_plane_kernel<T>, dim3((unsigned)((c->g.my + 255) / 256)), dim3(256), 0, stream, a);
    CLES_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(homo_global_kernel<T>, dim3(1), dim3(64), 0, stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  {
    const int ntx = (c->g.mx + MTX - 1) / MTX, nty = (c->g.my + MTY - 1) / MTY, ntz = (c->g.mz + MTZ - 1) / MTZ;
    const int ntiles = ntx * nty * ntz, pxm = (ntiles + 7) / 8;
    if (a.ih || a.jh || a.kh || a.les == 3) {
      hipLaunchKernelGGL(les_cs_kernel<T>, dim3(8 * pxm), dim3(256), 0, stream, a, ntx, nty, ntiles, pxm);
      CLES_HIP(c, hipGetLastError());
      hipLaunchKernelGGL(les_smooth_kernel<T>, dim3(8 * pxm), dim3(256), 0, stream, a, ntx, nty, ntiles, pxm);
      CLES_HIP(c, hipGetLastError());
    } else {
      const int mtx = (c->g.mx + FTX - 1) / FTX, mty = (c->g.my + FTY - 1) / FTY;
      // 8-plane chunks (at 256^3, each plane's loads issued during the sums
      // of the plane before: 8 planes 2.19 ms of LES model, 16 2.21, 32
      // 2.24), fewer where a launch would hold fewer than 1024 workgroups
      int kc = 8;
      while (kc > 4 && mtx * mty * ((c->g.mz + kc - 1) / kc) < 1024) kc /= 2;
      const int nt = mtx * mty * ((c->g.mz + kc - 1) / kc), px = (nt + 7) / 8;
      hipLaunchKernelGGL((les_mean_kernel<T, 0>), dim3(8 * px), dim3(256), 0, stream, a, mtx, mty, kc, nt, px);
      CLES_HIP(c, hipGetLastError());
      hipLaunchKernelGGL((les_mean_kernel<T, 1>), dim3(8 * px), dim3(256), 0, stream, a, mtx, mty, kc, nt, px);
      CLES_HIP(c, hipGetLastError());
    }
  }
  double *t = c->cs;
  c->cs = c->cs_alt;
  c->cs_alt = t;
  c->ucat_les_mixed = 1;
  return CLES_OK;
}

}  // namespace

int launch_les(cles_ctx *c) {
  return c->p.precision == CLES_PREC_FP32BULK ? launch_les_t<float>(c, c->stream, nullptr)
                                              : launch_les_t<double>(c, c->stream, nullptr);
}

// The LES model on stream s with the stale state read through the device
// cell qs_cell (batched, graph-captured steps)
int launch_les_on(cles_ctx *c, hipStream_t s, const double *const *qs_cell) {
  return c->p.precision == CLES_PREC_FP32BULK ? launch_les_t<float>(c, s, qs_cell)
                                              : launch_les_t<double>(c, s, qs_cell);
}

}  // namespace cles

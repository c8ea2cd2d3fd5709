// cles_stage.hip -- the fused right-hand-side + SSP-RK stage kernel.
//
// One launch per RK stage evaluates, for every interior node of the grid,
//   Conv = sum_d F_d(n) - F_d(n-1)        formConvection / formConvectionLES
//   Visc = sum_d G_d(n) - G_d(n-1) (+LES)  formViscous
//   Rhs  = aj (Visc - Conv) or 0           formRHS + MultAj
// and then either stores Rhs (parity mode) or applies the stage update of
// RungeKutta (solvers.c:256-271) and the equation of state (rhs.c:4853-4865)
// in registers, so the only HBM traffic is the stage input, the metrics and
// the stage output.
//
// Work decomposition: a workgroup owns a tile of interior nodes (one node
// per thread; 8x4x3 nodes in 128 threads by default, seven other shapes).
// The primitive variables (u, v, w, p/rho) of the tile plus a one-node halo
// are staged in LDS first.  Then, one direction at a time: the convective
// fluxes of the direction's faces (8x4x3: at most 128, one pass of the
// workgroup) go to an LDS face buffer and each node adds F(hi) - F(lo); the
// viscous fluxes of the same faces follow, in the same pass or through the
// same buffer.  For the central flux of the LES forms the direction's node
// data (state, p and the contravariant velocity) are staged in LDS as well,
// so each node's division is done once.  Faces are computed once per
// workgroup instead of twice per node.  Tiles are numbered so that each XCD
// (blockIdx % 8) receives a contiguous slab of tiles and keeps its
// neighbours in its own L2.  The kernel is bound by load latency, so loads
// are issued in groups and ahead of the phases that use them.
#include <cstring>
#include <type_traits>

#include "cles_internal.hpp"
#include "cles_seam.hpp"

namespace cles {

namespace {

constexpr int NT = 256;    // threads per workgroup of the larger tile shapes
// occupancy target of the stage kernels (waves per SIMD the register
// allocation must allow); 3 spills and runs slower
constexpr int kStageWaves = 2;
constexpr int NSLOT = 64;  // spread of the inf-norm atomics

// contravariant velocity (m . rhoU) / rho of the four-point central flux
__device__ __forceinline__ double contra(double m0, double m1, double m2, double ru, double rv, double rw, double rd) {
  return (m0 * ru + m1 * rv + m2 * rw) * rcp64(rd);
}

// Tile of interior nodes owned by one workgroup, with the face counts of its
// convective/viscous face buffer and its primitive window (one-node halo).
template <int X, int Y, int Z, int N = NT>
struct Tile {
  static_assert(X * Y * Z <= N, "at most one node per thread");
  static_assert(N % 64 == 0, "whole waves");
  static constexpr int TX = X, TY = Y, TZ = Z, NN = X * Y * Z, NT = N;
  static constexpr int NFI = (X + 1) * Y * Z, NFJ = X * (Y + 1) * Z, NFK = X * Y * (Z + 1);
  static constexpr int NF = NFI + NFJ + NFK;
  static constexpr int WX = X + 2, WY = Y + 2, WZ = Z + 2, NW = WX * WY * WZ;
};
// tile shapes the launcher can pick (index = cles_ctx::tile, default 7)
typedef Tile<32, 4, 2> TileA;
typedef Tile<16, 4, 4> TileB;
typedef Tile<16, 8, 2> TileC;
typedef Tile<64, 2, 2> TileD;
// shapes whose face count per direction fits one pass of the workgroup
// (6x6x6: 252 faces; 8x5x5: at most 240), at 216 and 200 nodes per workgroup
typedef Tile<6, 6, 6> TileE;
typedef Tile<8, 5, 5> TileF;
// one- and two-wave workgroups: no wave waits at a barrier for another, at
// the price of more halo work per node (4x4x3: at most 64 faces; 8x4x3: at most 128)
typedef Tile<4, 4, 3, 64> TileG;
typedef Tile<8, 4, 3, 128> TileH;

struct StageArgs {
  Grid g;
  const double *q, *p;  // stage input (5 planes, 1 plane)
  const double *qold;   // Qold (5 planes)
  double *qo, *po;      // stage output
  double *rhs;          // parity mode output (5 planes)
  const double *csi, *eta, *zet, *aj, *nv, *nut, *fil2;
  const double *mpack;  // LES: the nine metrics and aj in blocks of 8 nodes ([np/8][5][8][2])
  int bc[6];
  int fluxorder, eno, central, iso, has_ib;
  double ren, Ma, Pr, gamma, Tref, ratio, dt;
  // uniform parts of the viscosity, formed once on the host with the same
  // operations: 1/ren, 110.4/Tref, (1/ren) (1 + 110.4/Tref)
  double inv_ren, suth_b, suth_a;
  // gamma Ma^2 (Sutherland's temperature) and the heat-flux factor
  // gamma / (Pr (gamma - 1)), formed once on the host
  double gm2, kheat;
  // reference state of the FP32-bulk form: the float arithmetic carries the
  // deviations from it, the reference parts of the fluxes stay in double
  double rho_ref, p_ref, e_ref, h_ref, t_ref;
  int suth;
  int twoD;  // -twoD 1: no xi-direction fluxes, u = 0 in CalcUCat, no rhoU residual
  SigmaSpec sig;
  const double *sgm;  // -calc_sigma: the sensor (3 planes), else null and sig applies
  int weno5, roe;     // build-only schemes (read by the SCH = 1 instantiations only)
  double a_rhs, a_old, a_cur;
  int istage;
  int ntx, nty, ntz, per_xcd;
  unsigned long long *norms;  // [2][NSLOT]
};

template <typename T>
__device__ __forceinline__ void load6(const StageArgs &a, long long n, T s[6]) {
  const long long np = a.g.np;
  s[0] = T(a.q[n]);
  s[1] = T(a.q[n + np]);
  s[2] = T(a.q[n + 2 * np]);
  s[3] = T(a.q[n + 3 * np]);
  s[4] = T(a.q[n + 4 * np]);
  s[5] = T(a.p[n]);
}
// 32-bit byte offsets of a node in a plane (a plane holds fewer than 2^29
// nodes, checked at cles_create): loads and stores address a uniform plane
// base plus the offset, the scalar-base form of the global memory
// instructions, instead of forming a 64-bit address per access.
__device__ __forceinline__ unsigned node_off(const Grid &g, int k, int j, int i) {
  return ((unsigned)k * (unsigned)g.sz + (unsigned)j * (unsigned)g.sy + (unsigned)i) * 8u;
}
// first value of a node in the blocked metric pack (value pair p of the node
// at + 16 p), from the node's byte offset in a plane
__device__ __forceinline__ const double *mpack_at(const double *pack, unsigned off) {
  const long long n = (long long)(off >> 3);
  return pack + (n >> 3) * 80 + (n & 7) * 2;
}
__device__ __forceinline__ double ldo(const double *plane, unsigned off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(plane) + off);
}
__device__ __forceinline__ void sto(double *plane, unsigned off, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(plane) + off) = v;
}

// Node data of the current face direction staged in LDS by the workgroup:
// S[c * ns + o], c = 0..6 for rho, rhoU, rhoV, rhoW, rhoE, p and the
// contravariant velocity of the direction; o0 is the slot of the face's lower
// node, st the slot step along the direction.  S == nullptr: read global memory.
struct Sv {
  const double *S;
  int ns, o0, st;
  bool on;  // staged (S valid)
};

// the six state values of node n, which sits dq nodes from the face's lower node
template <typename T>
__device__ __forceinline__ void load6v(const StageArgs &a, const Sv &sv, long long n, int dq, T s[6]) {
  if (sv.on) {
    const int o = sv.o0 + dq * sv.st;
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = T(sv.S[c * sv.ns + o]);
  } else {
    load6(a, n, s);
  }
}

__device__ __forceinline__ double nvat(const StageArgs &a, long long n) { return a.has_ib ? a.nv[n] : 0.0; }
__device__ __forceinline__ const double *metric(const StageArgs &a, int d) {
  return d == 0 ? a.csi : d == 1 ? a.eta : a.zet;
}
__device__ __forceinline__ bool interior(const Grid &g, int k, int j, int i) {
  return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
__device__ __forceinline__ double sigma_at(const StageArgs &a, int d, int k, int j, int i) {
  if (a.sgm) return a.sgm[nid(a.g, k, j, i) + d * a.g.np];
  return interior(a.g, k, j, i) ? a.sig.in[d] : a.sig.bnd[d];
}

template <typename T>
__device__ __forceinline__ void recon_muscl(const StageArgs &a, const Sv &sv, long long nm, long long n0, long long n1,
                                            long long n2, T *L, T *R) {
  T s0[6], s1[6], sm[6], sp[6];
  load6v(a, sv, n0, 0, s0);
  load6v(a, sv, n1, 1, s1);
  if (nvat(a, nm) > 2.9) load6v(a, sv, n1, 1, sm); else load6v(a, sv, nm, -1, sm);
  if (nvat(a, n2) > 2.9) load6v(a, sv, n0, 0, sp); else load6v(a, sv, n2, 2, sp);
#pragma unroll
  for (int q = 0; q < 6; q++) {
    L[q] = muscl_ql(sm[q], s0[q], s1[q]);
    R[q] = muscl_qr(s0[q], s1[q], sp[q]);
  }
}

template <typename T>
__device__ __forceinline__ void recon_weno(const StageArgs &a, const Sv &sv, long long nm, long long n0, long long n1,
                                           long long n2, bool eno_left, int eno_mask, T *L, T *R) {
  T sm[6], s0[6], s1[6], s2[6];
  load6v(a, sv, nm, -1, sm);
  load6v(a, sv, n0, 0, s0);
  load6v(a, sv, n1, 1, s1);
  load6v(a, sv, n2, 2, s2);
  const int order = a.eno;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    if (eno_left || (eno_mask >> q & 1))
      L[q] = eno_ql(s0[q], s1[q], s2[q], order);
    else
      L[q] = weno_ql(sm[q], s0[q], s1[q]);
  }
  if (nvat(a, n2) > 2.9) {
#pragma unroll
    for (int q = 0; q < 6; q++) R[q] = eno_qr(sm[q], s0[q], s1[q], order);
  } else {
#pragma unroll
    for (int q = 0; q < 6; q++) R[q] = weno_qr(s0[q], s1[q], s2[q]);
  }
}

template <typename T>
__device__ __forceinline__ void face_metric(const StageArgs &a, int d, long long n0, long long n1, T h[3]) {
  const double *m = metric(a, d);
  const long long np = a.g.np;
  h[0] = T(0.5 * (m[n0] + m[n1]));
  h[1] = T(0.5 * (m[n0 + np] + m[n1 + np]));
  h[2] = T(0.5 * (m[n0 + 2 * np] + m[n1 + 2 * np]));
}

__device__ __forceinline__ double ucon_d(const double *s, const double *h) {
  const double inv = rcp64(s[0]);
  return h[0] * (s[1] * inv) + h[1] * (s[2] * inv) + h[2] * (s[3] * inv);
}

// FP32-bulk Rusanov flux on double face states: the float arithmetic sees
// the deviations rho - rho_ref, E - e_ref, p - p_ref and the state jumps;
// rho_ref and e_ref + p_ref = h_ref times the mean contravariant velocity and
// p_ref times the face metric (the GCL part) are added in double.  Same
// algebra as rusanov() (rhs.c:2596-2640).
__device__ void rusanov_split(const StageArgs &a, const double *L, const double *R, const double *hd, double uconL,
                              double uconR, double F[5]) {
  const float h0 = (float)hd[0], h1 = (float)hd[1], h2 = (float)hd[2];
  const float gii = h0 * h0 + h1 * h1 + h2 * h2;
  const float CsoundL = 1.4f * (float)L[5] * rcp32((float)L[0]);
  const float CsoundR = 1.4f * (float)R[5] * rcp32((float)R[0]);
  const float uL = (float)uconL, uR = (float)uconR;
  const float saL = fabsf(uL) + sqrtf(fabsf(CsoundL) * gii);
  const float saR = fabsf(uR) + sqrtf(fabsf(CsoundR) * gii);
  const float SA = saL < saR ? saR : saL;
  const float rL = (float)(L[0] - a.rho_ref), rR = (float)(R[0] - a.rho_ref);
  const float pL = (float)(L[5] - a.p_ref), pR = (float)(R[5] - a.p_ref);
  const float eL = (float)(L[4] - a.e_ref), eR = (float)(R[4] - a.e_ref);
  float jump[5];
#pragma unroll
  for (int q = 0; q < 5; q++) jump[q] = (float)(R[q] - L[q]);
  const double um = 0.5 * (uconL + uconR);
  F[0] = a.rho_ref * um + (double)(0.5f * (rL * uL + rR * uR) - 0.5f * SA * jump[0]);
  const float hh[3] = {h0, h1, h2};
#pragma unroll
  for (int c = 0; c < 3; c++)
    F[1 + c] = a.p_ref * hd[c] +
               (double)(0.5f * (((float)L[1 + c] * uL + (float)R[1 + c] * uR) + (pL * hh[c] + pR * hh[c])) -
                        0.5f * SA * jump[1 + c]);
  F[4] = a.h_ref * um + (double)(0.5f * ((pL + eL) * uL + (pR + eR) * uR) - 0.5f * SA * jump[4]);
}

__device__ __forceinline__ void load6d(const StageArgs &a, long long n, double s[6]) { load6<double>(a, n, s); }

// FP32-bulk reconstruction: WENO / ENO / MUSCL are equivariant under a common
// shift of their three inputs, so they run in float on the deviations from
// the node n0 value and the shift is added back in double.
__device__ void recon_split(const StageArgs &a, const Sv &sv, bool muscl, long long nm, long long n0, long long n1,
                            long long n2, bool eno_left, int eno_mask, double L[6], double R[6]) {
  double sm[6], s0[6], s1[6], s2[6];
  load6v(a, sv, n0, 0, s0);
  load6v(a, sv, n1, 1, s1);
  if (muscl) {
    if (nvat(a, nm) > 2.9) load6v(a, sv, n1, 1, sm); else load6v(a, sv, nm, -1, sm);
    if (nvat(a, n2) > 2.9) load6v(a, sv, n0, 0, s2); else load6v(a, sv, n2, 2, s2);
  } else {
    load6v(a, sv, nm, -1, sm);
    load6v(a, sv, n2, 2, s2);
  }
  const bool solid2 = nvat(a, n2) > 2.9;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const float dm = (float)(sm[q] - s0[q]), d1 = (float)(s1[q] - s0[q]), d2 = (float)(s2[q] - s0[q]);
    float l, r;
    if (muscl) {
      l = muscl_ql(dm, 0.0f, d1);
      r = muscl_qr(0.0f, d1, d2);
    } else {
      l = (eno_left || (eno_mask >> q & 1)) ? eno_ql(0.0f, d1, d2, a.eno) : weno_ql(dm, 0.0f, d1);
      r = solid2 ? eno_qr(dm, 0.0f, d1, a.eno) : weno_qr(0.0f, d1, d2);
    }
    L[q] = s0[q] + (double)l;
    R[q] = s0[q] + (double)r;
  }
}

// WENO5 (option weno5, no reference counterpart) where the face's six-node
// stencil lies in the grid and holds no solid node; the reference's WENO3 /
// ENO elsewhere
__device__ __forceinline__ bool weno5_fits(const StageArgs &a, int id, int m, long long n0, long long s) {
  if (!a.weno5 || id < 2 || id > m - 4) return false;
  return !(nvat(a, n0 - 2 * s) > 2.9 || nvat(a, n0 - s) > 2.9 || nvat(a, n0 + 2 * s) > 2.9 ||
           nvat(a, n0 + 3 * s) > 2.9);
}
__device__ __forceinline__ void recon_weno5(const StageArgs &a, long long n0, long long s, double *L, double *R) {
  double v[6][6];
#pragma unroll
  for (int t = 0; t < 6; t++) load6(a, n0 + (t - 2) * s, v[t]);
#pragma unroll
  for (int q = 0; q < 6; q++) {
    L[q] = weno5_ql(v[0][q], v[1][q], v[2][q], v[3][q], v[4][q]);
    R[q] = weno5_ql(v[5][q], v[4][q], v[3][q], v[2][q], v[1][q]);
  }
}

// formConvection face (rhs.c:2646-2814, 2817-2971, 2973-3121).  SCH 1: the
// build-only schemes of the options weno5 / roe may replace WENO3 / Rusanov.
template <typename T, int SCH>
__device__ void conv_face_plain(const StageArgs &a, const Sv &sv, int d, int k, int j, int i, double F[5]) {
  const Grid &g = a.g;
  const int m = d == 0 ? g.mx : d == 1 ? g.my : g.mz;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const long long s = d == 0 ? 1 : d == 1 ? g.sy : g.sz;
  const long long n0 = nid(g, k, j, i), n1 = n0 + s;
  const bool iso = a.iso != 0;
  constexpr bool F32 = sizeof(T) == 4;
  typedef typename std::conditional<F32, double, T>::type S;  // face-state type
  S L[6], R[6];
  if (d == 1 && j == 0 && iso && a.bc[2] == 1) {
    load6v(a, sv, n0, 0, L);
    load6v(a, sv, n0, 0, R);
  } else if (d == 1 && j == g.my - 2 && iso && a.bc[3] == 1) {
    load6v(a, sv, n1, 1, L);
    load6v(a, sv, n1, 1, R);
  } else if (a.fluxorder == 1 || id == 0 || id == m - 2 || (nvat(a, n0) + nvat(a, n1)) > 2.9) {
    load6v(a, sv, n0, 0, L);
    load6v(a, sv, n1, 1, R);
  } else {
    const long long nm = n0 - s, n2 = n0 + 2 * s;
    const bool muscl = a.fluxorder == 2 || (d != 0 && j < 30 && k < 7);
    bool eno_left = false;
    int mask = 0;
    if (!muscl) {
      eno_left = nvat(a, nm) > 2.9;
      if (d == 0) eno_left = eno_left && (i + 2) < g.mx;
      if (d == 1 && !eno_left && j == 1 && iso && a.bc[2] == 1) mask = (1 << 0) | (1 << 4) | (1 << 5);
    }
    if constexpr (F32) {
      recon_split(a, sv, muscl, nm, n0, n1, n2, eno_left, mask, L, R);
    } else {
      bool done = false;
      if constexpr (SCH != 0) {
        if (!muscl && mask == 0 && weno5_fits(a, id, m, n0, s)) {
          recon_weno5(a, n0, s, L, R);
          done = true;
        }
      }
      if (!done) {
        if (muscl) recon_muscl(a, sv, nm, n0, n1, n2, L, R);
        else recon_weno(a, sv, nm, n0, n1, n2, eno_left, mask, L, R);
      }
    }
  }
  S h[3];
  face_metric(a, d, n0, n1, h);
  if constexpr (SCH != 0 && !F32) {
    if (a.roe) {
      roe_flux(L, R, h, a.gamma, F);
      return;
    }
  }
  if constexpr (F32) {
    const double uconL = (id == m - 2 && a.bc[2 * d + 1] == 1) ? ucon_d(R, h) : ucon_d(L, h);
    const double uconR = (id == 0 && a.bc[2 * d] == 1) ? uconL : ucon_d(R, h);
    rusanov_split(a, L, R, h, uconL, uconR, F);
  } else {
    T uconL = (id == m - 2 && a.bc[2 * d + 1] == 1) ? ucon_of(R, h) : ucon_of(L, h);
    T uconR = (id == 0 && a.bc[2 * d] == 1) ? uconL : ucon_of(R, h);
    T Ft[5];
    rusanov(L, R, h, uconL, uconR, Ft);
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = Ft[q];
  }
}

// CEN: sigma = 1 on every face (no -calc_sigma, s_x = s_y = s_z = 1), where
// the blend reduces to the central flux and the upwind part is not compiled
template <typename T, int SCH, bool CEN = false>
__device__ void conv_face_les(const StageArgs &a, const Sv &sv, int d, int k, int j, int i, double F[5],
                              const double *hp);

// FP32-bulk four-point central flux (sigma = 1).  fp_2(c, U) = c * fp4(U) for
// a constant c, so rho, rhoE + p and the pressure are split into reference
// parts, whose fluxes rho_ref fp4(U), h_ref fp4(U) and p_ref h use the
// contravariant velocity in double, and float deviations for the rest.  hp:
// the face's metric averages when the caller has them, else null.
template <bool CEN>
__device__ bool conv_face_les_split(const StageArgs &a, const Sv &sv, int d, int k, int j, int i, double F[5],
                                    const double *hp) {
  const Grid &g = a.g;
  const int m = d == 0 ? g.mx : d == 1 ? g.my : g.mz;
  const int id = d == 0 ? i : d == 1 ? j : k;
  if (!(id > 0 && id < m - 2 && (a.central == 2 || a.central == 3))) return false;
  if (!CEN && sigma_at(a, d, k, j, i) != 1.0) return false;
  const long long s = d == 0 ? 1 : d == 1 ? g.sy : g.sz;
  const long long n0 = nid(g, k, j, i), n1 = n0 + s, nm = n0 - s, n2 = n0 + 2 * s;
  const double sum = d == 0 ? nvat(a, n0) + nvat(a, n1) + nvat(a, n2) + nvat(a, nm)
                            : nvat(a, nm) + nvat(a, n0) + nvat(a, n1) + nvat(a, n2);
  if (!(sum < 2.9)) return false;
  const long long np = g.np;
  const double *mt = metric(a, d);
  double hd[3];
  if (hp) {
    hd[0] = hp[0];
    hd[1] = hp[1];
    hd[2] = hp[2];
  } else {
    face_metric(a, d, n0, n1, hd);
  }
  float r[4], u[4], v[4], w[4], ep[4], pp[4], U[4];
  double Ud[4];
  const long long nn[4] = {nm, n0, n1, n2};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const long long n = nn[q];
    double rd, ru, rv, rw, ed, pd;
    if (sv.on) {
      const int o = sv.o0 + (q - 1) * sv.st;
      rd = sv.S[o]; ru = sv.S[sv.ns + o]; rv = sv.S[2 * sv.ns + o]; rw = sv.S[3 * sv.ns + o];
      ed = sv.S[4 * sv.ns + o]; pd = sv.S[5 * sv.ns + o];
      Ud[q] = sv.S[6 * sv.ns + o];
    } else {
      rd = a.q[n]; ru = a.q[n + np]; rv = a.q[n + 2 * np]; rw = a.q[n + 3 * np];
      ed = a.q[n + 4 * np]; pd = a.p[n];
      Ud[q] = contra(mt[n], mt[n + np], mt[n + 2 * np], ru, rv, rw, rd);
    }
    U[q] = (float)Ud[q];
    r[q] = (float)(rd - a.rho_ref);
    u[q] = (float)ru;
    v[q] = (float)rv;
    w[q] = (float)rw;
    pp[q] = (float)(pd - a.p_ref);
    ep[q] = (float)((ed + pd) - a.h_ref);
  }
  const float h0 = (float)hd[0], h1 = (float)hd[1], h2 = (float)hd[2];
  float Fc[5];
  if (a.central == 2) {
    Fc[0] = fp_2(r, U);
    Fc[1] = fp_2(u, U) + fp4(pp) * h0;
    Fc[2] = fp_2(v, U) + fp4(pp) * h1;
    Fc[3] = fp_2(w, U) + fp4(pp) * h2;
    Fc[4] = fp_2(ep, U);
  } else {
    Fc[0] = fp_1(r, U);
    Fc[1] = fp_1(u, U) + fp4(pp) * h0;
    Fc[2] = fp_1(v, U) + fp4(pp) * h1;
    Fc[3] = fp_1(w, U) + fp4(pp) * h2;
    Fc[4] = fp_1(ep, U);
  }
  const double fu = fp4(Ud);
  F[0] = a.rho_ref * fu + (double)Fc[0];
  F[1] = a.p_ref * hd[0] + (double)Fc[1];
  F[2] = a.p_ref * hd[1] + (double)Fc[2];
  F[3] = a.p_ref * hd[2] + (double)Fc[3];
  F[4] = a.h_ref * fu + (double)Fc[4];
  return true;
}

// formConvectionLES face (rhs.c:3466-3706, 3714-3937, 3947-4168).  The
// FP32-bulk form uses the split central flux above and, for the rare faces
// it does not cover (boundaries, immersed nodes, sigma < 1), this function in
// double.  SCH 1: the upwind part may use the weno5 / roe schemes.  hp: the
// face's metric averages when the caller has them, else null.
template <typename T, int SCH, bool CEN>
__device__ void conv_face_les(const StageArgs &a, const Sv &sv, int d, int k, int j, int i, double F[5],
                              const double *hp) {
  if constexpr (sizeof(T) == 4) {
    // the double fallback reads global memory: the staged velocity is the float form's
    if (!conv_face_les_split<CEN>(a, sv, d, k, j, i, F, hp))
      conv_face_les<double, 0, CEN>(a, Sv{nullptr, 0, 0, 0, false}, d, k, j, i, F, hp);
    return;
  }
  const Grid &g = a.g;
  const int m = d == 0 ? g.mx : d == 1 ? g.my : g.mz;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const long long s = d == 0 ? 1 : d == 1 ? g.sy : g.sz;
  const long long n0 = nid(g, k, j, i), n1 = n0 + s;
  const long long np = g.np;
  T h[3];
  if (hp) {
    h[0] = T(hp[0]);
    h[1] = T(hp[1]);
    h[2] = T(hp[2]);
  } else {
    face_metric(a, d, n0, n1, h);
  }
  T Fc[5];
  bool four = false;
  if (id > 0 && id < m - 2 && (a.central == 2 || a.central == 3)) {
    const long long nm = n0 - s, n2 = n0 + 2 * s;
    double sum = d == 0 ? nvat(a, n0) + nvat(a, n1) + nvat(a, n2) + nvat(a, nm)
                        : nvat(a, nm) + nvat(a, n0) + nvat(a, n1) + nvat(a, n2);
    if (sum < 2.9) {
      four = true;
      const double *mt = metric(a, d);
      T r[4], u[4], v[4], w[4], ep[4], pp[4], U[4];
      const long long nn[4] = {nm, n0, n1, n2};
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long long n = nn[q];
        T e;
        if (sv.on) {
          const int o = sv.o0 + (q - 1) * sv.st;
          r[q] = T(sv.S[o]);
          u[q] = T(sv.S[sv.ns + o]);
          v[q] = T(sv.S[2 * sv.ns + o]);
          w[q] = T(sv.S[3 * sv.ns + o]);
          e = T(sv.S[4 * sv.ns + o]);
          pp[q] = T(sv.S[5 * sv.ns + o]);
          U[q] = T(sv.S[6 * sv.ns + o]);
        } else {
          r[q] = T(a.q[n]);
          u[q] = T(a.q[n + np]);
          v[q] = T(a.q[n + 2 * np]);
          w[q] = T(a.q[n + 3 * np]);
          e = T(a.q[n + 4 * np]);
          pp[q] = T(a.p[n]);
          U[q] = T(contra(mt[n], mt[n + np], mt[n + 2 * np], double(u[q]), double(v[q]), double(w[q]), double(r[q])));
        }
        ep[q] = d == 2 ? e + pp[q] : pp[q] + e;
      }
      if (a.central == 2) {
        Fc[0] = fp_2(r, U);
        Fc[1] = fp_2(u, U) + fp4(pp) * h[0];
        Fc[2] = fp_2(v, U) + fp4(pp) * h[1];
        Fc[3] = fp_2(w, U) + fp4(pp) * h[2];
        Fc[4] = fp_2(ep, U);
      } else {
        Fc[0] = fp_1(r, U);
        Fc[1] = fp_1(u, U) + fp4(pp) * h[0];
        Fc[2] = fp_1(v, U) + fp4(pp) * h[1];
        Fc[3] = fp_1(w, U) + fp4(pp) * h[2];
        Fc[4] = fp_1(ep, U);
      }
    }
  }
  if (!four) {
    T s0[6], s1[6];
    load6v(a, sv, n0, 0, s0);
    load6v(a, sv, n1, 1, s1);
    T rc = T(0.5) * (s0[0] + s1[0]);
    T uc = T(0.5) * (s0[1] + s1[1]);
    T vc = T(0.5) * (s0[2] + s1[2]);
    T wc = T(0.5) * (s0[3] + s1[3]);
    T ec = T(0.5) * (s0[4] + s1[4]);
    T Uc = (h[0] * uc + h[1] * vc + h[2] * wc) * T(rcp64(rc));
    T pc = T(0.5) * (s0[5] + s1[5]);
    Fc[0] = rc * Uc;
    Fc[1] = uc * Uc + pc * h[0];
    Fc[2] = vc * Uc + pc * h[1];
    Fc[3] = wc * Uc + pc * h[2];
    Fc[4] = Uc * (ec + pc);
  }
  if constexpr (CEN) {
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = Fc[q];
    return;
  }
  const T sigma = T(sigma_at(a, d, k, j, i));
  if (sigma == T(1)) {
    // (1 - sigma) * F_upwind == 0 for every finite upwind flux: the blend of
    // rhs.c:3698-3702 reduces to the central flux exactly.
#pragma unroll
    for (int q = 0; q < 5; q++) F[q] = Fc[q];
    return;
  }
  // upwind part; periodic faces 0 and m-2 read ghost nodes (wrapped)
  const long long nm = nidw(g, k - (d == 2), j - (d == 1), i - (d == 0));
  const long long n2 = nidw(g, k + 2 * (d == 2), j + 2 * (d == 1), i + 2 * (d == 0));
  const double nvpair = d == 2 ? nvat(a, n0) + nvat(a, nid(g, k + 1, j, 1)) : nvat(a, n0) + nvat(a, n1);
  T L[6], R[6];
  const Sv gv{nullptr, 0, 0, 0, false};  // wrapped neighbours: read global memory
  if (a.fluxorder == 1 || ((id == 0 || id == m - 2) && a.bc[2 * d] != 7) || nvpair > 2.9) {
    load6(a, n0, L);
    load6(a, n1, R);
  } else if (a.fluxorder == 2) {
    recon_muscl(a, gv, nm, n0, n1, n2, L, R);
  } else {
    bool done = false;
    if constexpr (SCH != 0) {
      if (weno5_fits(a, id, m, n0, s)) {
        recon_weno5(a, n0, s, L, R);
        done = true;
      }
    }
    if (!done) {
      bool eno_left = nvat(a, nm) > 2.9;
      if (d == 0) eno_left = eno_left && (i + 2) < g.mx;
      recon_weno(a, gv, nm, n0, n1, n2, eno_left, 0, L, R);
    }
  }
  T Fw[5];
  bool roe_done = false;
  if constexpr (SCH != 0) {
    if (a.roe) {
      roe_flux(L, R, h, a.gamma, Fw);
      roe_done = true;
    }
  }
  if (!roe_done) rusanov(L, R, h, ucon_of(L, h), ucon_of(R, h), Fw);
#pragma unroll
  for (int q = 0; q < 5; q++) F[q] = Fw[q] * (T(1.0) - sigma) + Fc[q] * sigma;
}

// instantiate the double version used by the float fallback
template __device__ void conv_face_les<double, 0, false>(const StageArgs &, const Sv &, int, int, int, int,
                                                         double *, const double *);
template __device__ void conv_face_les<double, 0, true>(const StageArgs &, const Sv &, int, int, int, int,
                                                        double *, const double *);

// ------------------------------ viscous ----------------------------------
struct Win {  // LDS window of primitives with a one-node halo
  int i0, j0, k0;  // global coordinates of window node (0,0,0)
};
template <class TL>
__device__ __forceinline__ int widx(const Win &w, int k, int j, int i) {
  return ((k - w.k0) * TL::WY + (j - w.j0)) * TL::WX + (i - w.i0);
}
__device__ __forceinline__ bool ibn(double nv) { return nv > 0.5 && nv < 7.; }

// one-sided cross derivatives next to immersed-boundary rows (rhs.c:1498-1575)
template <class TL, typename T>
__device__ __forceinline__ void cross_deriv(const StageArgs &a, const T *W, int wA, int wB, int wt,
                                            long long gA, long long gB, long long gt, bool q8, T du[3], T &dT) {
  constexpr int NW = TL::NW;
  const T *U = W, *V = W + NW, *Wz = W + 2 * NW, *Tm = W + 3 * NW;
  int mode = 2;
  if (a.has_ib) {
    if (ibn(a.nv[gA + gt]) || ibn(a.nv[gB + gt])) mode = 0;
    else if (ibn(a.nv[gA - gt]) || ibn(a.nv[gB - gt])) mode = 1;
  }
  if (mode == 0) {
    du[0] = (U[wB] + U[wA] - U[wB - wt] - U[wA - wt]) * T(0.5);
    du[1] = (V[wB] + V[wA] - V[wB - wt] - V[wA - wt]) * T(0.5);
    du[2] = (Wz[wB] + Wz[wA] - Wz[wB - wt] - Wz[wA - wt]) * T(0.5);
    dT = q8 ? (Tm[wB] + Tm[wA] - Tm[wB - wt] - Tm[wA]) * T(0.5)
            : (Tm[wB] + Tm[wA] - Tm[wB - wt] - Tm[wA - wt]) * T(0.5);
  } else if (mode == 1) {
    du[0] = (U[wB + wt] + U[wA + wt] - U[wB] - U[wA]) * T(0.5);
    du[1] = (V[wB + wt] + V[wA + wt] - V[wB] - V[wA]) * T(0.5);
    du[2] = (Wz[wB + wt] + Wz[wA + wt] - Wz[wB] - Wz[wA]) * T(0.5);
    dT = (Tm[wB + wt] + Tm[wA + wt] - Tm[wB] - Tm[wA]) * T(0.5);
  } else {
    du[0] = (U[wB + wt] + U[wA + wt] - U[wB - wt] - U[wA - wt]) * T(0.25);
    du[1] = (V[wB + wt] + V[wA + wt] - V[wB - wt] - V[wA - wt]) * T(0.25);
    du[2] = (Wz[wB + wt] + Wz[wA + wt] - Wz[wB - wt] - Wz[wA - wt]) * T(0.25);
    dT = (Tm[wB + wt] + Tm[wA + wt] - Tm[wB - wt] - Tm[wA - wt]) * T(0.25);
  }
}

// Global-memory inputs of one viscous face (its lower node A, upper node B):
// the nine metric components, aj, rho and p at both nodes, rho at A + 1 and
// (LES) nu_t at both nodes and the filter width at A.  Loaded in one place so
// that a caller can issue the loads early and keep them in flight across the
// phases before the face is evaluated.
struct VisIn {
  double m[9][2];  // csi0..2, eta0..2, zet0..2
  double aj[2], r[2], p[2], nut[2], rA1, f2;
};
// part 1: the metrics and aj; part 2: the rest; 3: both
template <bool LES, int PART = 3>
__device__ __forceinline__ void visc_load(const StageArgs &a, int d, int k, int j, int i, VisIn &v) {
  const Grid &g = a.g;
  const long long np = g.np;
  const long long gA = nid(g, k, j, i), gB = gA + (d == 0 ? 1 : d == 1 ? g.sy : g.sz);
  const double *mt[3] = {a.csi, a.eta, a.zet};
  if (PART & 1) {
#pragma unroll
    for (int c = 0; c < 9; c++) {
      v.m[c][0] = mt[c / 3][gA + (c % 3) * np];
      v.m[c][1] = mt[c / 3][gB + (c % 3) * np];
    }
    v.aj[0] = a.aj[gA];
    v.aj[1] = a.aj[gB];
  }
  if (!(PART & 2)) return;
  v.r[0] = a.q[gA];
  v.r[1] = a.q[gB];
  if (a.suth) {
    v.p[0] = a.p[gA];
    v.p[1] = a.p[gB];
  }
  if (LES) {
    v.rA1 = a.q[gA + 1];
    v.nut[0] = a.nut[gA];
    v.nut[1] = a.nut[gB];
    v.f2 = a.fil2[gA];
  }
}

// Viscous coefficients of a face from its densities, pressures and LES
// inputs (rhs.c:1616-1635, 1681-1697): cf[0] = ajc nu, cf[1] = ajc nu_t rho
// (the SGS stress factor), cf[2] = the factor of S2 in the isotropic SGS
// term.  ajc: the face's aj average; fin: nu_t at A and B, the filter width
// at A, rho at A + 1.
template <typename T, bool LES, bool CEN>
__device__ __forceinline__ void visc_coef(const StageArgs &a, int d, int k, int j, int i, double ajd, double rA,
                                          double rB, double pA, double pB, const double fin[4], double cf[3]) {
  const T ajc = T(ajd);
  T nu = T(a.inv_ren);
  if (LES) nu = T(a.inv_ren * 0.5 * (rA + fin[3]));  // rho at [i] and [i+1] on every face
  if (a.suth) {
    // Sutherland (rhs.c:1628-1635): Tr = gamma Ma^2 p_face / rho_face,
    // nu = 1/ren (1 + C/Tref) / (Tr + C/Tref) Tr^1.5
    const double Tr = a.gm2 * (pA + pB) * rcp64(rA + rB);
    nu = Tr < 0 ? T(a.inv_ren) : T(a.suth_a * rcp64(Tr + a.suth_b) * (Tr * sqrt(Tr)));
  }
  cf[0] = double(ajc * nu);
  if (LES) {
    const Grid &g = a.g;
    const int id = d == 0 ? i : d == 1 ? j : k;
    const int m = d == 0 ? g.mx : d == 1 ? g.my : g.mz;
    constexpr double kIso = 0.06 * 2.0 / 3.0;  // 2/3 CI, CI = 0.06 (rhs.c:13)
    const T sA = CEN ? T(1) : T(sigma_at(a, d, k, j, i));
    const T sB = CEN ? T(1) : T(sigma_at(a, d, k + (d == 2), j + (d == 1), i + (d == 0)));
    T nu_t = T(0.5) * (T(fin[0]) * sA + T(fin[1]) * sB);
    const T rho_c = T(0.5 * (rA + rB));
    const T sigma_c = T(0.5) * (sA + sB);
    if ((a.bc[2 * d] == 1 && id == 0) || (a.bc[2 * d + 1] == 1 && id == m - 2)) nu_t = T(0);
    cf[1] = double(ajc * (nu_t * rho_c));
    // isotropic part -2/3 CI rho Delta^2 S2 sigma, without the face Jacobian (rhs.c:1693-1697)
    cf[2] = double(T(fin[2]) * T(kIso) * rho_c * sigma_c);
  }
}

// formViscous face (rhs.c:1487-1707, 1719-1918, 1926-2121) from the LDS
// window and the face data fm[c * fs]: c = 0..8 the metric averages (csi,
// eta, zet), 9..11 the coefficients of visc_coef.  With r_ab = du_a/dx_b
// (the face's velocity gradient through its metrics) the reference's
// contraction g_t1 du_a/dt + r_ba m_b - div m_a equals
// m_b (r_ab + r_ba) - div m_a, and its divergence term 2/3 (r_11 + r_22 +
// r_33); the heat flux g_t1 dT/dt is m_b (dT/dt m_t,b).  G[0..3] = rhoU..rhoE
// flux, al = LES stress (alpha) of the face.
template <class TL, typename T, bool LES>
__device__ void visc_face_s(const StageArgs &a, const T *W, const Win &win, int d, int k, int j, int i,
                            const double *fm, int fs, T G[4], T al[3]) {
  constexpr int WX = TL::WX, WY = TL::WY, NW = TL::NW;
  const Grid &g = a.g;
  const long long gs[3] = {1, g.sy, g.sz};
  const int ws[3] = {1, WX, WX * WY};
  const long long gA = nid(g, k, j, i), gB = gA + gs[d];
  const int wA = widx<TL>(win, k, j, i), wB = wA + ws[d];
  const T *U = W, *V = W + NW, *Wz = W + 2 * NW, *Tm = W + 3 * NW;
  T dd[3][3], dT[3];  // dd[t][a]: d u_a / d(computational direction t)
  dd[d][0] = U[wB] - U[wA];
  dd[d][1] = V[wB] - V[wA];
  dd[d][2] = Wz[wB] - Wz[wA];
  dT[d] = Tm[wB] - Tm[wA];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    if (t == d) continue;
    cross_deriv<TL>(a, W, wA, wB, ws[t], gA, gB, gs[t], d == 0 && t == 2, dd[t], dT[t]);
  }
  T mt[3][3];  // mt[t][b]: face metric row t (csi, eta, zet), component b
#pragma unroll
  for (int t = 0; t < 3; t++)
#pragma unroll
    for (int b = 0; b < 3; b++) mt[t][b] = T(fm[(3 * t + b) * fs]);
  const T m0 = mt[d][0], m1 = mt[d][1], m2 = mt[d][2];
  T r[3][3];
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int b = 0; b < 3; b++) r[p][b] = dd[0][p] * mt[0][b] + dd[1][p] * mt[1][b] + dd[2][p] * mt[2][b];
  const T div = T(2. / 3.) * (r[0][0] + r[1][1] + r[2][2]);
  const T s01 = r[0][1] + r[1][0], s02 = r[0][2] + r[2][0], s12 = r[1][2] + r[2][1];
  const T bu = m0 * (r[0][0] + r[0][0] - div) + m1 * s01 + m2 * s02;
  const T bv = m0 * s01 + m1 * (r[1][1] + r[1][1] - div) + m2 * s12;
  const T bw = m0 * s02 + m1 * s12 + m2 * (r[2][2] + r[2][2] - div);
  const T cm = T(fm[9 * fs]);
  T fu = bu * cm, fv = bv * cm, fw = bw * cm;
  const int id = d == 0 ? i : d == 1 ? j : k;
  const int m = d == 0 ? g.mx : d == 1 ? g.my : g.mz;
  T uh0, uh1, uh2;
  if (id == 0 && a.bc[2 * d] == 1) {
    uh0 = U[wA]; uh1 = V[wA]; uh2 = Wz[wA];
  } else if (id == m - 2 && a.bc[2 * d + 1] == 1) {
    uh0 = U[wB]; uh1 = V[wB]; uh2 = Wz[wB];
  } else {
    uh0 = T(0.5) * (U[wB] + U[wA]);
    uh1 = T(0.5) * (V[wB] + V[wA]);
    uh2 = T(0.5) * (Wz[wB] + Wz[wA]);
  }
  // heat flux with the conductivity gamma nu / (Pr (gamma - 1)) (rhs.c:1666-1670)
  T gT = T(0);
#pragma unroll
  for (int b = 0; b < 3; b++) gT += mt[d][b] * (dT[0] * mt[0][b] + dT[1] * mt[1][b] + dT[2] * mt[2][b]);
  const T fe = (fu * uh0 + fv * uh1 + fw * uh2) + gT * (cm * T(a.kheat));
  if (LES) {
    const T ct = T(fm[10 * fs]);
    const T tx = bu * ct, ty = bv * ct, tz = bw * ct;
    const T S2 = T(.5) * ((r[0][0] - div) * (r[0][0] - div) + r[0][1] * r[0][1] + r[0][2] * r[0][2] +
                          r[1][0] * r[1][0] + (r[1][1] - div) * (r[1][1] - div) + r[1][2] * r[1][2] +
                          r[2][0] * r[2][0] + r[2][1] * r[2][1] + (r[2][2] - div) * (r[2][2] - div));
    const T ss = S2 * T(fm[11 * fs]);
    al[0] = tx - ss * m0;
    al[1] = ty - ss * m1;
    al[2] = tz - ss * m2;
    fu += al[0];
    fv += al[1];
    fw += al[2];
  }
  G[0] = fu;
  G[1] = fv;
  G[2] = fw;
  G[3] = fe;
}

// formViscous face from its global-memory inputs v (visc_load): the same
// arithmetic as the staged form, visc_coef then visc_face_s.
template <class TL, typename T, bool LES>
__device__ void visc_face(const StageArgs &a, const T *W, const Win &win, int d, int k, int j, int i, const VisIn &v,
                          T G[4], T al[3]) {
  double fa[12];
#pragma unroll
  for (int c = 0; c < 9; c++) fa[c] = 0.5 * (v.m[c][0] + v.m[c][1]);
  const double fin[4] = {LES ? v.nut[0] : 0.0, LES ? v.nut[1] : 0.0, LES ? v.f2 : 0.0, LES ? v.rA1 : 0.0};
  visc_coef<T, LES, false>(a, d, k, j, i, 0.5 * (v.aj[0] + v.aj[1]), v.r[0], v.r[1], a.suth ? v.p[0] : 0.0,
                           a.suth ? v.p[1] : 0.0, fin, fa + 9);
  visc_face_s<TL, T, LES>(a, W, win, d, k, j, i, fa, 1, G, al);
}

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}

// face decode: f in [0, NF) -> direction and local coordinates
template <class TL>
__device__ __forceinline__ void decode_face(int f, int &d, int &lx, int &ly, int &lz) {
  constexpr int TX = TL::TX, TY = TL::TY, NFI = TL::NFI, NFJ = TL::NFJ;
  if (f < NFI) {
    d = 0; lx = f % (TX + 1); ly = (f / (TX + 1)) % TY; lz = f / ((TX + 1) * TY);
  } else if (f < NFI + NFJ) {
    f -= NFI;
    d = 1; lx = f % TX; ly = (f / TX) % (TY + 1); lz = f / (TX * (TY + 1));
  } else {
    f -= NFI + NFJ;
    d = 2; lx = f % TX; ly = (f / TX) % TY; lz = f / (TX * TY);
  }
}
template <class TL> __device__ __forceinline__ int fi_of(int lx, int ly, int lz) {
  return (lz * TL::TY + ly) * (TL::TX + 1) + lx;
}
template <class TL> __device__ __forceinline__ int fj_of(int lx, int ly, int lz) {
  return TL::NFI + (lz * (TL::TY + 1) + ly) * TL::TX + lx;
}
template <class TL> __device__ __forceinline__ int fk_of(int lx, int ly, int lz) {
  return TL::NFI + TL::NFJ + (lz * TL::TY + ly) * TL::TX + lx;
}

// One workgroup's tile of one grid; b only spreads the norm atomics.
template <class TL, typename T, bool LES, bool VISC, int MODE, int SCH = 0, bool CEN = false>
__device__ __forceinline__ void stage_tile(const StageArgs &a, int tile, int b) {
  constexpr int TX = TL::TX, TY = TL::TY, TZ = TL::TZ, WX = TL::WX, WY = TL::WY, NW = TL::NW;
  // The convective and viscous flux of a face are computed in one pass (one
  // barrier pair per direction) except in the LES forms, where two passes
  // through a smaller face buffer measured faster.
  constexpr bool MERGE = VISC && !LES;
  // face buffer: the convective flux (5) and, computed in the same pass, the
  // viscous flux (4) and the LES stresses (3) of each face of one direction
  constexpr int FV = VISC ? (LES ? 7 : 4) : 0;
  constexpr int FC = MERGE ? 5 + FV : (VISC ? (LES ? 7 : 5) : 5);
  constexpr int VO = MERGE ? 5 : 0;  // first viscous slot
  constexpr int NFD = TL::NFI > TL::NFJ ? (TL::NFI > TL::NFK ? TL::NFI : TL::NFK) : (TL::NFJ > TL::NFK ? TL::NFJ : TL::NFK);
  // staged node box of direction d: the tile widened by two nodes on both
  // sides along d (the four-point stencils of its faces)
  constexpr int NCH = LES ? 7 : 6;
  constexpr int NS0 = (TX + 4) * TY * TZ, NS1 = TX * (TY + 4) * TZ, NS2 = TX * TY * (TZ + 4);
  constexpr int NSM = NS0 > NS1 ? (NS0 > NS2 ? NS0 : NS2) : (NS1 > NS2 ? NS1 : NS2);
  // Staged form (LES, every face of a direction in one pass of the
  // workgroup): per direction, one load phase fills the node box in LDS
  // (state, p, contravariant velocity from the metric row of the blocked
  // metric pack) and gives each face's thread its nine metric averages and aj
  // (from the pack) and viscous inputs in registers, later its viscous
  // coefficients, so the face passes read LDS and registers only; the flux
  // buffer Fs then reuses the box's memory.  Elsewhere (larger tiles) the
  // face routines read global memory.
  constexpr bool ONEPASS = TL::NFI <= TL::NT && TL::NFJ <= TL::NT && TL::NFK <= TL::NT;
  constexpr int NFM = VISC ? 12 : 3;
  constexpr int NSF = NCH * NSM > FC * NFD ? NCH * NSM : FC * NFD;
  constexpr bool STG = LES && ONEPASS && 8 * NSF + (VISC ? (int)sizeof(T) * 4 * NW : 0) <= 160 * 1024 * TL::NT / 512;
  __shared__ double SF[STG ? NSF : FC * NFD];
  __shared__ T Ws[VISC ? 4 * NW : 1];
  double(*Fs)[NFD] = reinterpret_cast<double(*)[NFD]>(SF);
  double *Ss = SF;
  const Grid &g = a.g;
  if (tile >= a.ntx * a.nty * a.ntz) return;
  // x fastest, then z, then y: the tiles in flight on an XCD share their z
  // halos (the deepest: a 3-node tile with +-1 and +-2 stencils) in its L2
  const int tix = tile % a.ntx, tiz = (tile / a.ntx) % a.ntz, tiy = tile / (a.ntx * a.ntz);
  const int i0 = 1 + tix * TX, j0 = 1 + tiy * TY, k0 = 1 + tiz * TZ;
  const int t = threadIdx.x;
  const long long np = g.np;

  // Faces are processed one direction at a time (convective, then viscous),
  // so the face buffer holds one direction's faces; the per-node sums run in
  // the order i, j, k exactly as a single expression over all six faces would.
  // threads beyond the tile's nodes take part in the face phases only (node
  // coordinates clamped to 0 keep their LDS reads in range)
  const bool tn = t < TL::NN;
  const int lx = tn ? t % TX : 0, ly = tn ? (t / TX) % TY : 0, lz = tn ? t / (TX * TY) : 0;
  const int gi = i0 + lx, gj = j0 + ly, gk = k0 + lz;
  const bool valid = tn && gi <= g.mx - 2 && gj <= g.my - 2 && gk <= g.mz - 2;
  const long long n = valid ? nid(g, gk, gj, gi) : 0;
  double conv[5] = {0., 0., 0., 0., 0.};
  double visc[5] = {0., 0., 0., 0., 0.};
  // LES energy term (rhs.c:2153-2197): face-density weighted alpha sums
  const bool les_e = LES && VISC && valid && gj > 2 && gj < g.my - 2 && !(a.has_ib && !(a.nv[n] < 0.5));
  T E0 = T(0), E1 = T(0), E2 = T(0);
  // the node's update inputs; the staged form takes them during the last
  // direction (the state from its box, aj and Qold loaded with its box)
  double ajn = 0., qin[5] = {0., 0., 0., 0., 0.}, qol[5] = {0., 0., 0., 0., 0.};
  const unsigned no = (unsigned)n * 8u;
  // Qold is the stage input itself in the first stage: its values are qin
  const bool qold_in = a.qold == a.q;
  auto load_update = [&]() {
    if (!valid) return;
    ajn = ldo(a.aj, no);
    if (MODE != 0 && !qold_in) {
#pragma unroll
      for (int q = 0; q < 5; q++) qol[q] = ldo(a.qold + q * np, no);
    }
  };
  Win win{i0 - 1, j0 - 1, k0 - 1};
  constexpr int NIT = (NW + TL::NT - 1) / TL::NT;
  double qv[NIT][5];
  bool ok[NIT];
  if (VISC) {
    // stage u, v, w, p/rho of the tile and a one-node halo; the loads of all
    // slots are issued before any LDS store, so they are in flight together
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int w = t + it * TL::NT;
      const int ii = win.i0 + w % WX, jj = win.j0 + (w / WX) % WY, kk = win.k0 + w / (WX * WY);
      ok[it] = w < NW && ii <= g.mx - 1 && jj <= g.my - 1 && kk <= g.mz - 1;
      if (ok[it]) {
        const unsigned mo = node_off(g, kk, jj, ii);
#pragma unroll
        for (int c = 0; c < 4; c++) qv[it][c] = ldo(a.q + c * np, mo);
        qv[it][4] = ldo(a.p, mo);
      }
    }
  }
  // the window's LDS stores: in the staged form they follow the first
  // direction's loads, so both load phases are in flight together
  auto store_window = [&]() {
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int w = t + it * TL::NT;
      if (!ok[it]) continue;
      const double ir = rcp64(qv[it][0]);
      Ws[w] = a.twoD == 1 ? T(0) : T(qv[it][1] * ir);  // CalcUCat, rhs.c:1349
      Ws[NW + w] = T(qv[it][2] * ir);
      Ws[2 * NW + w] = T(qv[it][3] * ir);
      Ws[3 * NW + w] = T(qv[it][4] * ir - a.t_ref);  // t_ref = 0 in FP64: only gradients of p/rho are used
    }
  };
  if (VISC && !STG) {
    store_window();
    __syncthreads();  // the window is read by other threads' viscous faces
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {  // unrolled: d is a constant in the face routines
    const int nfd = d == 0 ? TL::NFI : d == 1 ? TL::NFJ : TL::NFK;
    const int ex = TX + (d == 0), ey = TY + (d == 1);  // face-grid extents of direction d
    // faces of this node: lo at the node, hi one step along d
    const int lo = (lz * ey + ly) * ex + lx;
    const int hi = lo + (d == 0 ? 1 : d == 1 ? ex : ex * ey);
    // -twoD 1 computes no xi face (rhs.c:1487-1714, 2646-2814, 3465-3706):
    // their fluxes and LES stresses are the zeros of fresh Vecs
    const bool skip_d = d == 0 && a.twoD == 1;
    if constexpr (STG) {
      // ---- staged form: thread t owns face t of direction d
      constexpr int ns = NSM;
      const int bx = TX + 4 * (d == 0), by = TY + 4 * (d == 1), bz = TZ + 4 * (d == 2);
      const int ox = i0 - 2 * (d == 0), oy = j0 - 2 * (d == 1), oz = k0 - 2 * (d == 2);
      const int bst = d == 0 ? 1 : d == 1 ? bx : bx * by;  // box slot step along d
      const int fx = t % ex, fy = (t / ex) % ey, fz = t / (ex * ey);
      const int fi = i0 + fx - (d == 0), fj = j0 + fy - (d == 1), fk = k0 + fz - (d == 2);
      const bool fval = t < nfd && fi <= g.mx - 2 && fj <= g.my - 2 && fk <= g.mz - 2;
      // ---- one load phase: the node box (the tile widened by two nodes
      // along d: state, p and the metric row of the contravariant velocity),
      // then the face's metrics and viscous inputs; all in flight together
      constexpr int NBI = (NSM + TL::NT - 1) / TL::NT;
      const int nb = bx * by * bz;
      const double *mtd = metric(a, d);
      double bv[NBI][9];
      bool bok[NBI];
#pragma unroll
      for (int it = 0; it < NBI; it++) {
        const int w = t + it * TL::NT;
        const int ii = ox + w % bx, jj = oy + (w / bx) % by, kk = oz + w / (bx * by);
        bok[it] = w < nb && ii >= 0 && jj >= 0 && kk >= 0 && ii <= g.mx - 1 && jj <= g.my - 1 && kk <= g.mz - 1;
        if (bok[it]) {
          const unsigned mo = node_off(g, kk, jj, ii);
#pragma unroll
          for (int c = 0; c < 5; c++) bv[it][c] = ldo(a.q + c * np, mo);
          bv[it][5] = ldo(a.p, mo);
          // metric row d (values 3d .. 3d + 2) from the pack: two loads
          const double *pk = mpack_at(a.mpack, mo);
          if (d == 0) {
            const double2 x = *reinterpret_cast<const double2 *>(pk);
            bv[it][6] = x.x;
            bv[it][7] = x.y;
            bv[it][8] = pk[16];
          } else if (d == 1) {
            bv[it][6] = pk[17];
            const double2 x = *reinterpret_cast<const double2 *>(pk + 32);
            bv[it][7] = x.x;
            bv[it][8] = x.y;
          } else {
            const double2 x = *reinterpret_cast<const double2 *>(pk + 48);
            bv[it][6] = x.x;
            bv[it][7] = x.y;
            bv[it][8] = pk[64];
          }
        }
      }
      const bool fload = fval && !skip_d;
      double fm[NFM][2];
      double fin[4] = {0., 0., 0., 0.};  // LES: nu_t at A and B, filter width at A, rho at A + 1
      if (fload) {
        const unsigned oA = node_off(g, fk, fj, fi);
        const unsigned oB = oA + 8u * (unsigned)(d == 0 ? 1 : d == 1 ? g.sy : g.sz);
        if (VISC) {
          // the nine metrics and aj of nodes A and B from the metric pack:
          // five 16-byte loads per node
          const double2 *pA = reinterpret_cast<const double2 *>(mpack_at(a.mpack, oA));
          const double2 *pB = reinterpret_cast<const double2 *>(mpack_at(a.mpack, oB));
#pragma unroll
          for (int c = 0; c < 5; c++) {
            const double2 xa = pA[8 * c], xb = pB[8 * c];
            fm[2 * c][0] = xa.x;
            fm[2 * c + 1][0] = xa.y;
            fm[2 * c][1] = xb.x;
            fm[2 * c + 1][1] = xb.y;
          }
          if (LES) {
            fin[0] = ldo(a.nut, oA);
            fin[1] = ldo(a.nut, oB);
            fin[2] = ldo(a.fil2, oA);
            if (!a.suth) fin[3] = ldo(a.q, oA + 8u);
          }
        } else {
#pragma unroll
          for (int c = 0; c < 3; c++) {
            fm[c][0] = ldo(mtd + c * np, oA);
            fm[c][1] = ldo(mtd + c * np, oB);
          }
        }
      }
      if (d == 2) load_update();
      if (VISC && d == 0) store_window();
#pragma unroll
      for (int it = 0; it < NBI; it++) {
        if (!bok[it]) continue;
        const int w = t + it * TL::NT;
#pragma unroll
        for (int c = 0; c < 6; c++) Ss[c * ns + w] = bv[it][c];
        Ss[6 * ns + w] = contra(bv[it][6], bv[it][7], bv[it][8], bv[it][1], bv[it][2], bv[it][3], bv[it][0]);
      }
      // the face's averages (9 metrics and aj, or the metric row of d), then
      // its viscous coefficients: kept by the thread, which computes both
      // fluxes of its face
      double fa[NFM];
      if (fload) {
        constexpr int NAV = VISC ? 10 : 3;
#pragma unroll
        for (int c = 0; c < NAV; c++) fa[c] = 0.5 * (fm[c][0] + fm[c][1]);
      }
      __syncthreads();
      // ---- convective flux of the face, and its viscous coefficients from
      // the box's densities and pressures
      double F[5] = {0., 0., 0., 0., 0.};
      double cf[3];
      if (fload) {
        const Sv sv{Ss, ns, ((fk - oz) * by + (fj - oy)) * bx + (fi - ox), bst, true};
        double h[3];
#pragma unroll
        for (int c = 0; c < 3; c++) h[c] = fa[VISC ? 3 * d + c : c];
        conv_face_les<T, SCH, CEN>(a, sv, d, fk, fj, fi, F, h);
        if (VISC) {
          visc_coef<T, LES, CEN>(a, d, fk, fj, fi, fa[9], Ss[sv.o0], Ss[sv.o0 + bst], Ss[5 * ns + sv.o0],
                            Ss[5 * ns + sv.o0 + bst], fin, cf);
#pragma unroll
          for (int q = 0; q < 3; q++) fa[9 + q] = cf[q];
        }
      }
      // face densities of the LES energy term: rho at n - st, n, n + st along d
      T rl = T(0), rh = T(0);
      if (les_e) {
        const int on = ((gk - oz) * by + (gj - oy)) * bx + (gi - ox);
        const double r0 = Ss[on];
        rl = T((r0 + Ss[on - bst]) * 0.5);
        rh = T((Ss[on + bst] + r0) * 0.5);
      }
      if (d == 2 && valid) {
        const int on = ((gk - oz) * by + (gj - oy)) * bx + (gi - ox);
#pragma unroll
        for (int q = 0; q < 5; q++) qin[q] = Ss[q * ns + on];
      }
      __syncthreads();  // the box is read: the flux buffer takes its memory
      if (fval) {
#pragma unroll
        for (int q = 0; q < 5; q++) Fs[q][t] = F[q];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 5; q++) conv[q] = d == 0 ? Fs[q][hi] - Fs[q][lo] : conv[q] + Fs[q][hi] - Fs[q][lo];
      __syncthreads();
      if (VISC) {
        // ---- viscous flux of the face from the window and its face data
        if (fval) {
          T G[4] = {T(0), T(0), T(0), T(0)}, al[3] = {T(0), T(0), T(0)};
          if (!skip_d) visc_face_s<TL, T, LES>(a, Ws, win, d, fk, fj, fi, fa, 1, G, al);
#pragma unroll
          for (int q = 0; q < 4; q++) Fs[q][t] = G[q];
          if (LES) {
#pragma unroll
            for (int q = 0; q < 3; q++) Fs[4 + q][t] = al[q];
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++)
          visc[q + 1] = d == 0 ? Fs[q][hi] - Fs[q][lo] : visc[q + 1] + Fs[q][hi] - Fs[q][lo];
        if (les_e) {
          const T p0 = rh * T(Fs[4][hi]), p1 = rh * T(Fs[5][hi]), p2 = rh * T(Fs[6][hi]);
          E0 = d == 0 ? p0 - rl * T(Fs[4][lo]) : E0 + p0 - rl * T(Fs[4][lo]);
          E1 = d == 0 ? p1 - rl * T(Fs[5][lo]) : E1 + p1 - rl * T(Fs[5][lo]);
          E2 = d == 0 ? p2 - rl * T(Fs[6][lo]) : E2 + p2 - rl * T(Fs[6][lo]);
        }
        __syncthreads();
      }
      continue;
    }
    // Two-pass forms whose faces of every direction fit one pass of the
    // workgroup: the non-metric inputs of this thread's viscous face (f = t)
    // are loaded right after its convective face, and stay in flight through
    // the convective sums (barriers wait for LDS traffic only); loading the
    // metric part there too spills registers.
    constexpr bool PF = VISC && !MERGE && ONEPASS;
    VisIn vpre;
    // face densities of the LES energy term: rho at n, n - st, n + st along d
    double rle[3];
    if (les_e) {
      const long long st = d == 0 ? 1 : d == 1 ? g.sy : g.sz;
      rle[0] = a.q[n];
      rle[1] = a.q[n - st];
      rle[2] = a.q[n + st];
    }
    // ---- convective fluxes of direction d -> LDS
    for (int f = t; f < nfd; f += TL::NT) {
      const int fx = f % ex, fy = (f / ex) % ey, fz = f / (ex * ey);
      const int fi = i0 + fx - (d == 0), fj = j0 + fy - (d == 1), fk = k0 + fz - (d == 2);
      if (!(fi <= g.mx - 2 && fj <= g.my - 2 && fk <= g.mz - 2)) continue;
      if (skip_d) {
#pragma unroll
        for (int q = 0; q < FC; q++) Fs[q][f] = 0.0;
        continue;
      }
      const Sv sv{nullptr, 0, 0, 0, false};
      double F[5];
      if (LES) conv_face_les<T, SCH>(a, sv, d, fk, fj, fi, F, nullptr);
      else conv_face_plain<T, SCH>(a, sv, d, fk, fj, fi, F);
#pragma unroll
      for (int q = 0; q < 5; q++) Fs[q][f] = F[q];
      if (MERGE) {
        // ---- viscous flux of the same face
        T G[4], al[3] = {T(0), T(0), T(0)};
        VisIn vi;
        visc_load<LES>(a, d, fk, fj, fi, vi);
        visc_face<TL, T, LES>(a, Ws, win, d, fk, fj, fi, vi, G, al);
#pragma unroll
        for (int q = 0; q < 4; q++) Fs[VO + q][f] = G[q];
        if (LES) {
#pragma unroll
          for (int q = 0; q < 3; q++) Fs[VO + 4 + q][f] = al[q];
        }
      }
    }
    if (PF && t < nfd) {
      const int fx = t % ex, fy = (t / ex) % ey, fz = t / (ex * ey);
      const int fi = i0 + fx - (d == 0), fj = j0 + fy - (d == 1), fk = k0 + fz - (d == 2);
      if (fi <= g.mx - 2 && fj <= g.my - 2 && fk <= g.mz - 2) visc_load<LES, 2>(a, d, fk, fj, fi, vpre);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 5; q++) conv[q] = d == 0 ? Fs[q][hi] - Fs[q][lo] : conv[q] + Fs[q][hi] - Fs[q][lo];
    if (!VISC) {
      __syncthreads();
      continue;
    }
    if (!MERGE) {
      __syncthreads();
      // ---- viscous fluxes of direction d -> LDS
      for (int f = t; f < nfd; f += TL::NT) {
        const int fx = f % ex, fy = (f / ex) % ey, fz = f / (ex * ey);
        const int fi = i0 + fx - (d == 0), fj = j0 + fy - (d == 1), fk = k0 + fz - (d == 2);
        if (!(fi <= g.mx - 2 && fj <= g.my - 2 && fk <= g.mz - 2)) continue;
        if (skip_d) {
#pragma unroll
          for (int q = 0; q < FC; q++) Fs[q][f] = 0.0;
          continue;
        }
        T G[4], al[3] = {T(0), T(0), T(0)};
        if (PF) {
          visc_load<LES, 1>(a, d, fk, fj, fi, vpre);
          visc_face<TL, T, LES>(a, Ws, win, d, fk, fj, fi, vpre, G, al);
        } else {
          VisIn vi;
          visc_load<LES>(a, d, fk, fj, fi, vi);
          visc_face<TL, T, LES>(a, Ws, win, d, fk, fj, fi, vi, G, al);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) Fs[q][f] = G[q];
        if (LES) {
#pragma unroll
          for (int q = 0; q < 3; q++) Fs[4 + q][f] = al[q];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
      visc[q + 1] = d == 0 ? Fs[VO + q][hi] - Fs[VO + q][lo] : visc[q + 1] + Fs[VO + q][hi] - Fs[VO + q][lo];
    if (les_e) {
      // face densities (r_n + r_n-1)/2, (r_n+1 + r_n)/2 along d
      const T rl = T((rle[0] + rle[1]) * 0.5), rh = T((rle[2] + rle[0]) * 0.5);
      const T p0 = rh * T(Fs[VO + 4][hi]), p1 = rh * T(Fs[VO + 5][hi]), p2 = rh * T(Fs[VO + 6][hi]);
      E0 = d == 0 ? p0 - rl * T(Fs[VO + 4][lo]) : E0 + p0 - rl * T(Fs[VO + 4][lo]);
      E1 = d == 0 ? p1 - rl * T(Fs[VO + 5][lo]) : E1 + p1 - rl * T(Fs[VO + 5][lo]);
      E2 = d == 0 ? p2 - rl * T(Fs[VO + 6][lo]) : E2 + p2 - rl * T(Fs[VO + 6][lo]);
    }
    __syncthreads();
  }
  if (a.twoD == 1) {  // rhs.c:2135-2140, 3168-3173, 4215-4220
    conv[1] = 0.;
    visc[1] = 0.;
  }
  if (!STG && valid) {
    load_update();
#pragma unroll
    for (int q = 0; q < 5; q++) qin[q] = ldo(a.q + q * np, no);
  }
  if (LES && a.ratio != 0.0 && valid) {
    conv[3] = conv[3] - a.ratio / ajn * qin[0] / a.dt;
    conv[4] = conv[4] - a.ratio / ajn * qin[3] / a.dt;
  }
  if (les_e) {
    const double ir = rcp64(qin[0]);
    const T u0 = T(qin[1] * ir), u1 = T(qin[2] * ir), u2 = T(qin[3] * ir);
    const T e = u0 * E0 + u1 * E1 + u2 * E2;
    visc[4] = visc[4] + double(e);
  }
  // ---- Rhs = aj (Visc - Conv) (MultAj, rhs.c:4611-4629)
  double rhs[5] = {0., 0., 0., 0., 0.};
  if (valid) {
    const bool fluid = !a.has_ib || a.nv[n] < 0.1;
#pragma unroll
    for (int q = 0; q < 5; q++) {
      double r = visc[q] - conv[q];
      rhs[q] = fluid ? r * ajn : 0.;
    }
  }
  if (a.norms) {
    // every lane of the wave takes part; lanes without a node contribute 0
    double mr = 0., mq = 0.;
#pragma unroll
    for (int q = 0; q < 5; q++) {
      mr = fmax(mr, fabs(rhs[q]));
      mq = fmax(mq, fabs(qin[q]));
    }
    mr = wave_max(mr);
    mq = wave_max(mq);
    if ((t & 63) == 0) {
      const int slot = (b * 4 + (t >> 6)) & (NSLOT - 1);
      atomicMax(&a.norms[slot], dbits(mr));
      atomicMax(&a.norms[NSLOT + slot], dbits(mq));
    }
  }
  if (!valid) return;
  if (MODE == 0) {
#pragma unroll
    for (int q = 0; q < 5; q++) a.rhs[n + q * np] = rhs[q];
    return;
  }
  // ---- stage update, VecWAXPY / VecAXPBYPCZ (solvers.c:256-271)
  double qn[5];
  if (qold_in) {
#pragma unroll
    for (int q = 0; q < 5; q++) qol[q] = qin[q];
  }
  if (a.istage == 0) {
#pragma unroll
    for (int q = 0; q < 5; q++) qn[q] = a.a_rhs * rhs[q] + qol[q];
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) qn[q] = a.a_old * qol[q] + a.a_rhs * rhs[q] + a.a_cur * qin[q];
  }
#pragma unroll
  for (int q = 0; q < 5; q++) sto(a.qo + q * np, no, qn[q]);
  // EqOfState (rhs.c:4853-4857)
  a.po[n] = (a.gamma - 1) * (qn[4] - 0.5 * (qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]) * rcp64(qn[0]));
}

template <class TL, typename T, bool LES, bool VISC, int MODE, int SCH = 0, bool CEN = false>
__global__ __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(kStageWaves))) void stage_kernel(StageArgs a) {
  // XCD-aware tile numbering: blockIdx % 8 selects the XCD, each XCD walks a
  // contiguous range of tiles and keeps their shared halos in its own L2.
  const int b = blockIdx.x;
  stage_tile<TL, T, LES, VISC, MODE, SCH, CEN>(a, (b & 7) * a.per_xcd + (b >> 3), b);
}

// Batched stage: the workgroups of one launch cover the tiles of several
// independent grids; tstart[r] is the first workgroup of grid r (prefix sum
// of the workgroup counts, 8 per_xcd each, tstart[nreq] = total).
template <class TL, typename T, bool LES, bool VISC, int MODE, bool CEN = false>
__global__ __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(kStageWaves))) void stage_batch_kernel(const StageArgs *__restrict__ tab, const int *__restrict__ tstart,
                                                         int nreq) {
  const int b = blockIdx.x;
  int lo = 0, hi = nreq;  // largest r with tstart[r] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tstart[mid] <= b) lo = mid; else hi = mid;
  }
  // the grid's own XCD-aware numbering (every grid starts on a multiple of 8)
  const int lb = b - tstart[lo];
  stage_tile<TL, T, LES, VISC, MODE, 0, CEN>(tab[lo], (lb & 7) * tab[lo].per_xcd + (lb >> 3), b);
}

// Persistent batched stages (serving): the workgroups pull items from eight
// queues, their own XCD's first (blockIdx % 8), then the others.  Items are
// queued stage by stage, so an item waits only for items handed out before
// it; every one of those is held by a running workgroup, whose own waits
// again point only backwards, so the waits always end.  A wait is still
// bounded: after spin_limit polls it sets the abort flag and pushes every
// queue head past its end, the other waits see the flag within 32 polls,
// every workgroup stops taking items, the kernel exits and the host reports
// the failure.
template <class TL, typename T, bool LES, bool VISC>
__global__ __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(kStageWaves))) void stage_persist_kernel(PersistArgs a) {
  __shared__ int s_item;
  const StageArgs *tab = (const StageArgs *)a.stage;
  const int home = blockIdx.x & 7;
  for (int qq = 0; qq < 8; qq++) {
    const int x = (home + qq) & 7;
    const int n = a.nitems[x];
    const PersistPhase *ph0 = a.phase + (size_t)x * a.pstride;
    for (;;) {
      __syncthreads();  // the previous item's LDS is free
      if (threadIdx.x == 0) {
        const unsigned h = __hip_atomic_fetch_add(&a.head[x * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_item = h < (unsigned)n ? (int)h : -1;
      }
      __syncthreads();
      int it = s_item;
      if (it < 0) break;
      int lo = 0, hi = a.nph;  // the last phase starting at or before it
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ph0[mid].istart <= it) lo = mid; else hi = mid;
      }
      const PersistPhase ph = ph0[lo];
      if (ph.need) {
        if (threadIdx.x == 0) {
          unsigned spins = 0;
          while (__hip_atomic_load(&a.done[ph.member * kCtrStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ph.need) {
            ++spins;
            if (spins > a.spin_limit ||
                ((spins & 31) == 0 && __hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
              // give up: flag the failure and push every queue past its end
              __hip_atomic_store(a.abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              for (int y = 0; y < 8; y++)
                __hip_atomic_fetch_add(&a.head[y * kCtrStride], 1u << 30, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              s_item = -1;
              break;
            }
            __builtin_amdgcn_s_sleep(8);
          }
          // the producers' stores (released before their counter increments)
          // become visible to this workgroup
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        it = s_item;
        if (it < 0) break;
      }
      const int local = it - ph.istart;
      const int nti = (ph.nt + ph.grain - 1) / ph.grain;
      if (local < nti) {
        const StageArgs &sa = tab[(size_t)ph.stage * a.cap + ph.member];
        const int tb = ph.t0 + local * ph.grain;
        const int te = tb + ph.grain < ph.t0 + ph.nt ? tb + ph.grain : ph.t0 + ph.nt;
        for (int t = tb; t < te; t++) {
          if (t != tb) __syncthreads();
          stage_tile<TL, T, LES, VISC, 1>(sa, t, it);
        }
      } else {
        const SeamArgs &sm = a.seam[(size_t)ph.stage * a.cap + ph.member];
        const long long b0 = (long long)(ph.b0 + local - nti) * kSeamItem;
        for (int o = threadIdx.x; o < kSeamItem; o += TL::NT)
          seam_copy_eos(sm.g, sm.q, sm.qo, sm.po, sm.gamma, 1, b0 + o, sm.nb);
      }
      // every wave's stores have reached the L2 before the workgroup
      // releases them and counts the item
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&a.done[ph.member * kCtrStride], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

template <class TL>
void tile_counts(const Grid &g, StageArgs &a) {
  a.ntx = (g.mx - 2 + TL::TX - 1) / TL::TX;
  a.nty = (g.my - 2 + TL::TY - 1) / TL::TY;
  a.ntz = (g.mz - 2 + TL::TZ - 1) / TL::TZ;
  a.per_xcd = (a.ntx * a.nty * a.ntz + 7) / 8;
}

template <class TL, typename T, bool LES, bool VISC, int SCH = 0, bool CEN = false>
hipError_t launch_t(const StageArgs &a, int mode, int nblocks, hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL((stage_kernel<TL, T, LES, VISC, 0, SCH, CEN>), dim3(nblocks), dim3(TL::NT), 0, s, a);
  else
    hipLaunchKernelGGL((stage_kernel<TL, T, LES, VISC, 1, SCH, CEN>), dim3(nblocks), dim3(TL::NT), 0, s, a);
  return hipGetLastError();
}

// sigma = 1 on every face: no shock sensor field and s_x = s_y = s_z = 1
// inside and on the boundary
bool sigma_is_one(const StageArgs &a) {
  if (a.sgm) return false;
  for (int d = 0; d < 3; d++)
    if (a.sig.in[d] != 1.0 || a.sig.bnd[d] != 1.0) return false;
  return true;
}

// the build-only schemes (weno5 / roe, FP64 only) in the default 8x4x3 tile
hipError_t launch_schemes(const cles_params &p, StageArgs &a, int mode, hipStream_t s) {
  tile_counts<TileH>(a.g, a);
  const int nblocks = a.per_xcd * 8;
  const bool les = p.les != 0, visc = !p.invicid;
  if (les) return launch_t<TileH, double, true, true, 1>(a, mode, nblocks, s);
  return visc ? launch_t<TileH, double, false, true, 1>(a, mode, nblocks, s)
              : launch_t<TileH, double, false, false, 1>(a, mode, nblocks, s);
}

// every (precision, LES, viscous) variant of one tile shape
template <class TL>
hipError_t launch_shape(const cles_params &p, StageArgs &a, int mode, hipStream_t s) {
  tile_counts<TL>(a.g, a);
  const int nblocks = a.per_xcd * 8;
  const bool les = p.les != 0, visc = !p.invicid;
  if constexpr (std::is_same<TL, TileH>::value) {
    // the default tile of the LES forms: central flux only where sigma = 1
    if (les && sigma_is_one(a)) {
      if (p.precision == CLES_PREC_FP32BULK)
        return visc ? launch_t<TL, float, true, true, 0, true>(a, mode, nblocks, s)
                    : launch_t<TL, float, true, false, 0, true>(a, mode, nblocks, s);
      return visc ? launch_t<TL, double, true, true, 0, true>(a, mode, nblocks, s)
                  : launch_t<TL, double, true, false, 0, true>(a, mode, nblocks, s);
    }
  }
  if (p.precision == CLES_PREC_FP32BULK) {
    if (les) return visc ? launch_t<TL, float, true, true>(a, mode, nblocks, s) : launch_t<TL, float, true, false>(a, mode, nblocks, s);
    return visc ? launch_t<TL, float, false, true>(a, mode, nblocks, s) : launch_t<TL, float, false, false>(a, mode, nblocks, s);
  }
  if (les) return visc ? launch_t<TL, double, true, true>(a, mode, nblocks, s) : launch_t<TL, double, true, false>(a, mode, nblocks, s);
  return visc ? launch_t<TL, double, false, true>(a, mode, nblocks, s) : launch_t<TL, double, false, false>(a, mode, nblocks, s);
}

template <class TL, typename T, bool LES, bool VISC, bool CEN = false>
hipError_t launch_batch_t(const StageArgs *tab, const int *tstart, int nreq, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((stage_batch_kernel<TL, T, LES, VISC, 1, CEN>), dim3(nblocks), dim3(TL::NT), 0, s, tab, tstart,
                     nreq);
  return hipGetLastError();
}

template <class TL>
hipError_t launch_batch_shape(int fp32, int visc, int les, int cen, const StageArgs *tab, const int *tstart, int nreq,
                              int nblocks, hipStream_t s) {
  if (les) {  // the LES forms use the 8x4x3 tile
    if (cen) {
      if (fp32) return visc ? launch_batch_t<TileH, float, true, true, true>(tab, tstart, nreq, nblocks, s)
                            : launch_batch_t<TileH, float, true, false, true>(tab, tstart, nreq, nblocks, s);
      return visc ? launch_batch_t<TileH, double, true, true, true>(tab, tstart, nreq, nblocks, s)
                  : launch_batch_t<TileH, double, true, false, true>(tab, tstart, nreq, nblocks, s);
    }
    if (fp32) return visc ? launch_batch_t<TileH, float, true, true>(tab, tstart, nreq, nblocks, s)
                          : launch_batch_t<TileH, float, true, false>(tab, tstart, nreq, nblocks, s);
    return visc ? launch_batch_t<TileH, double, true, true>(tab, tstart, nreq, nblocks, s)
                : launch_batch_t<TileH, double, true, false>(tab, tstart, nreq, nblocks, s);
  }
  if (fp32) return visc ? launch_batch_t<TL, float, false, true>(tab, tstart, nreq, nblocks, s)
                        : launch_batch_t<TL, float, false, false>(tab, tstart, nreq, nblocks, s);
  return visc ? launch_batch_t<TL, double, false, true>(tab, tstart, nreq, nblocks, s)
              : launch_batch_t<TL, double, false, false>(tab, tstart, nreq, nblocks, s);
}

template <class TL, typename T, bool LES, bool VISC>
int persist_blocks_t() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, stage_persist_kernel<TL, T, LES, VISC>, TL::NT, 0) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return per_cu * cus;
}

template <class TL, typename T, bool LES, bool VISC>
hipError_t persist_t(const PersistArgs &a, hipStream_t s) {
  const int nb = persist_blocks_t<TL, T, LES, VISC>();
  if (nb < 8) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((stage_persist_kernel<TL, T, LES, VISC>), dim3(nb & ~7), dim3(TL::NT), 0, s, a);
  return hipGetLastError();
}

// (precision, model) -> instantiation; the LES forms use the 8x4x3 tile
template <class TL>
hipError_t persist_shape(int fp32, int visc, int les, const PersistArgs &a, hipStream_t s) {
  if (les) {
    if (fp32) return visc ? persist_t<TileH, float, true, true>(a, s) : persist_t<TileH, float, true, false>(a, s);
    return visc ? persist_t<TileH, double, true, true>(a, s) : persist_t<TileH, double, true, false>(a, s);
  }
  if (fp32) return visc ? persist_t<TL, float, false, true>(a, s) : persist_t<TL, float, false, false>(a, s);
  return visc ? persist_t<TL, double, false, true>(a, s) : persist_t<TL, double, false, false>(a, s);
}

template <class TL>
int persist_blocks_shape(int fp32, int visc, int les) {
  if (les) {
    if (fp32) return visc ? persist_blocks_t<TileH, float, true, true>() : persist_blocks_t<TileH, float, true, false>();
    return visc ? persist_blocks_t<TileH, double, true, true>() : persist_blocks_t<TileH, double, true, false>();
  }
  if (fp32) return visc ? persist_blocks_t<TL, float, false, true>() : persist_blocks_t<TL, float, false, false>();
  return visc ? persist_blocks_t<TL, double, false, true>() : persist_blocks_t<TL, double, false, false>();
}

}  // namespace

void fill_stage_args(cles_ctx *c, int in, int out, int istage, const double coef[3], StageArgs &a) {
  const cles_params &p = c->p;
  a.g = c->g;
  a.q = c->Q[in];
  a.p = c->P[in];
  a.qold = c->Q[c->cur];
  a.qo = out >= 0 ? c->Q[out] : nullptr;
  a.po = out >= 0 ? c->P[out] : nullptr;
  a.rhs = c->rhs;
  a.csi = c->csi; a.eta = c->eta; a.zet = c->zet; a.aj = c->aj; a.nv = c->nvert; a.nut = c->nut;
  a.fil2 = c->fil2;
  a.mpack = c->mpack;
  for (int q = 0; q < 6; q++) a.bc[q] = p.bctype[q];
  a.fluxorder = p.fluxorder; a.eno = p.eno_order; a.central = p.central; a.iso = p.isothermal;
  a.has_ib = c->has_ib;
  a.ren = p.ren; a.Ma = p.Ma; a.Pr = p.Pr; a.gamma = p.gamma; a.Tref = p.Tref; a.ratio = p.ratio; a.dt = p.dt;
  {
    const double C_sutherland = 110.4;  // rhs.c:1410
    a.inv_ren = 1. / p.ren;
    a.suth_b = C_sutherland / p.Tref;
    a.suth_a = 1. / p.ren * (1 + C_sutherland / p.Tref);
  }
  a.gm2 = p.gamma * p.Ma * p.Ma;
  a.kheat = p.gamma / (p.Pr * (p.gamma - 1.0));
  const bool f32 = p.precision == CLES_PREC_FP32BULK;
  a.rho_ref = f32 ? c->ref_rho : 0.0;
  a.p_ref = f32 ? c->ref_p : 0.0;
  a.e_ref = a.p_ref / (p.gamma - 1);
  a.h_ref = a.e_ref + a.p_ref;
  a.t_ref = f32 ? c->ref_p / c->ref_rho : 0.0;
  a.suth = p.sutherland;
  a.twoD = p.twoD;
  a.sig = c->sig;
  a.sgm = c->sgm;
  a.weno5 = p.weno5;
  a.roe = p.roe;
  a.a_rhs = coef[0]; a.a_old = coef[1]; a.a_cur = coef[2];
  a.istage = istage;
  a.norms = c->dnorm;
}

size_t stage_args_bytes() { return sizeof(StageArgs); }

int stage_batch_args(cles_ctx *c, int in, int out, int istage, const double coef[3], int tile, void *dst) {
  StageArgs a;
  fill_stage_args(c, in, out, istage, coef, a);
  a.norms = nullptr;
  switch (tile) {
    case 1: tile_counts<TileB>(a.g, a); break;
    case 2: tile_counts<TileC>(a.g, a); break;
    case 3: tile_counts<TileD>(a.g, a); break;
    case 4: tile_counts<TileE>(a.g, a); break;
    case 5: tile_counts<TileF>(a.g, a); break;
    case 6: tile_counts<TileG>(a.g, a); break;
    case 7: tile_counts<TileH>(a.g, a); break;
    default: tile_counts<TileA>(a.g, a); break;
  }
  memcpy(dst, &a, sizeof(a));
  return 8 * a.per_xcd;
}

#define CLES_TILE_SWITCH(tile, CALL)          \
  switch (tile) {                             \
    case 1: return CALL(TileB);               \
    case 2: return CALL(TileC);               \
    case 3: return CALL(TileD);               \
    case 4: return CALL(TileE);               \
    case 5: return CALL(TileF);               \
    case 6: return CALL(TileG);               \
    case 7: return CALL(TileH);               \
    default: return CALL(TileA);              \
  }

hipError_t launch_stage_persist(int tile, int fp32, int visc, int les, const PersistArgs &a, hipStream_t s) {
#define CALL(TL) persist_shape<TL>(fp32, visc, les, a, s)
  if (les) tile = 7;
  CLES_TILE_SWITCH(tile, CALL)
#undef CALL
}

int persist_blocks(int tile, int fp32, int visc, int les) {
#define CALL(TL) persist_blocks_shape<TL>(fp32, visc, les)
  if (les) tile = 7;
  CLES_TILE_SWITCH(tile, CALL)
#undef CALL
}

int stage_tiles(cles_ctx *c, int tile) {
  StageArgs a;
  a.g = c->g;
#define CALL(TL) (tile_counts<TL>(a.g, a), a.ntx * a.nty * a.ntz)
  CLES_TILE_SWITCH(tile, CALL)
#undef CALL
}

hipError_t launch_stage_batch(int tile, int fp32, int visc, int les, int cen, const void *tab, const int *tstart,
                              int nreq, int nblocks, hipStream_t s) {
  const StageArgs *t = (const StageArgs *)tab;
  switch (tile) {
    case 1: return launch_batch_shape<TileB>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 2: return launch_batch_shape<TileC>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 3: return launch_batch_shape<TileD>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 4: return launch_batch_shape<TileE>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 5: return launch_batch_shape<TileF>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 6: return launch_batch_shape<TileG>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    case 7: return launch_batch_shape<TileH>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
    default: return launch_batch_shape<TileA>(fp32, visc, les, cen, t, tstart, nreq, nblocks, s);
  }
}

int launch_stage(cles_ctx *c, int in, int out, int istage, int mode, const double coef[3]) {
  StageArgs a;
  fill_stage_args(c, in, out, istage, coef, a);
  hipError_t e;
  if (c->p.weno5 || c->p.roe) {
    CLES_HIP(c, launch_schemes(c->p, a, mode, c->stream));
    return CLES_OK;
  }
  switch (c->tile) {
    case 1: e = launch_shape<TileB>(c->p, a, mode, c->stream); break;
    case 2: e = launch_shape<TileC>(c->p, a, mode, c->stream); break;
    case 3: e = launch_shape<TileD>(c->p, a, mode, c->stream); break;
    case 4: e = launch_shape<TileE>(c->p, a, mode, c->stream); break;
    case 5: e = launch_shape<TileF>(c->p, a, mode, c->stream); break;
    case 6: e = launch_shape<TileG>(c->p, a, mode, c->stream); break;
    case 7: e = launch_shape<TileH>(c->p, a, mode, c->stream); break;
    default: e = launch_shape<TileA>(c->p, a, mode, c->stream); break;
  }
  CLES_HIP(c, e);
  return CLES_OK;
}

}  // namespace cles

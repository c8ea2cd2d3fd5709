// cles_internal.hpp -- device data structures and shared device functions of
// the gfx950 right-hand-side / time-advance path.
//
// Device layout (MI355X-first, differs from the host DMDA layout on purpose):
//   every field is stored as structure-of-arrays planes over the mx*my*mz
//   owned nodes, i fastest, each plane padded to a multiple of 64 nodes
//   (512 B) so that consecutive lanes read consecutive 8-byte words.
//   There are no ghost layers: a read at index -2,-1,m,m+1 of a periodic
//   direction wraps to (idx+m)%m, which is exactly the value the reference's
//   DMGlobalToLocal puts into its ghost layer (single-rank DMDA, width 2).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cles.h"

namespace cles {

constexpr int kMaxStages = 3;

struct Grid {
  int mx, my, mz;
  int px, py, pz;  // periodic flags per direction (bctype==7, init.c:158-197)
  long long sy;    // mx
  long long sz;    // mx*my
  long long np;    // plane stride (padded node count)
};

__host__ __device__ __forceinline__ int wrapc(int i, int m, int per) {
  if (per) {
    if (i < 0) i += m;
    else if (i >= m) i -= m;
  }
  return i;
}
__host__ __device__ __forceinline__ long long nid(const Grid &g, int k, int j, int i) {
  return (long long)k * g.sz + (long long)j * g.sy + i;
}
__host__ __device__ __forceinline__ long long nidw(const Grid &g, int k, int j, int i) {
  return nid(g, wrapc(k, g.mz, g.pz), wrapc(j, g.my, g.py), wrapc(i, g.mx, g.px));
}

// 1/x without an IEEE division: the hardware reciprocal refined by two Newton
// steps (float: one), a few instructions instead of the ~11 of a correctly
// rounded division and within an ulp of the quotient.  x is a density, a
// temperature or a sum of positive weights where it is used: finite, non-zero.
__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ float rcp32(float x) {
  float y = __builtin_amdgcn_rcpf(x);
  return fmaf(y, fmaf(-x, y, 1.0f), y);
}
__device__ __forceinline__ double rcpT(double x) { return rcp64(x); }
__device__ __forceinline__ float rcpT(float x) { return rcp32(x); }

// ---------------------------------------------------------------------------
// Reconstruction primitives, in the operation order of rhs.c:2256-2568
// (WENO weights normalised with one reciprocal instead of four divisions).
// T is the flux arithmetic type (double for the FP64 drop-in, float for the
// FP32-bulk form).
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T pmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T pmax(T a, T b) { return a < b ? b : a; }

template <typename T> __device__ __forceinline__ T minmod(T a, T b) {
  if (a * b < T(0)) return T(0);
  return (a >= T(0) ? T(1) : T(-1)) * pmin(fabs(a), fabs(b));
}
template <typename T> __device__ __forceinline__ T muscl_ql(T qm, T q0, T qp) {
  T S = minmod(qp - q0, q0 - qm);
  return q0 + T(0.5) * S;
}
template <typename T> __device__ __forceinline__ T muscl_qr(T qm, T q0, T qp) {
  T S = minmod(qp - q0, q0 - qm);
  return q0 - T(0.5) * S;
}
// WENO3 with the truncated linear weights d0 = 0.6666667, d1 = 0.3333333 and
// eps = 1e-6 (rhs.c:2359-2461): w0 = (d0 / b0^2) / (d0 / b0^2 + d1 / b1^2)
// with b = eps + beta, formed as A / (A + B), A = d0 b1^2, B = d1 b0^2.
template <typename T> __device__ __forceinline__ T weno_mix(T Qr0, T Qr1, T beta0, T beta1) {
  const T d0 = T(0.6666667), d1 = T(0.3333333), eps = T(1e-6);
  const T e0 = eps + beta0, e1 = eps + beta1;
  const T A = d0 * (e1 * e1), B = d1 * (e0 * e0);
  return (A * Qr0 + B * Qr1) * rcpT(A + B);
}
template <typename T> __device__ __forceinline__ T weno_ql(T qm, T q0, T qp) {
  T Qr0 = T(0.5) * q0 + T(0.5) * qp;
  T Qr1 = T(-0.5) * qm + T(1.5) * q0;
  return weno_mix(Qr0, Qr1, (qp - q0) * (qp - q0), (qm - q0) * (qm - q0));
}
template <typename T> __device__ __forceinline__ T weno_qr(T qm, T q0, T qp) {
  T Qr0 = T(1.5) * q0 - T(0.5) * qp;
  T Qr1 = T(0.5) * qm + T(0.5) * q0;
  return weno_mix(Qr0, Qr1, (qp - q0) * (qp - q0), (qm - q0) * (qm - q0));
}
// WENO5 of Jiang & Shu (1996): the left state at face i+1/2 from q[i-2..i+2];
// alpha_k = d_k / (1e-6 + beta_k)^2 normalised with one reciprocal (the
// weights multiplied through by the product of the squared denominators).
// No counterpart in the reference: the option weno5.
__device__ __forceinline__ double weno5_ql(double q0, double q1, double q2, double q3, double q4) {
  const double eps = 1e-6;
  const double s0 = q0 - 2 * q1 + q2, t0 = q0 - 4 * q1 + 3 * q2;
  const double s1 = q1 - 2 * q2 + q3, t1 = q1 - q3;
  const double s2 = q2 - 2 * q3 + q4, t2 = 3 * q2 - 4 * q3 + q4;
  const double e0 = eps + (13.0 / 12.0 * s0 * s0 + 0.25 * t0 * t0);
  const double e1 = eps + (13.0 / 12.0 * s1 * s1 + 0.25 * t1 * t1);
  const double e2 = eps + (13.0 / 12.0 * s2 * s2 + 0.25 * t2 * t2);
  const double f0 = e0 * e0, f1 = e1 * e1, f2 = e2 * e2;
  const double A0 = 0.1 * (f1 * f2), A1 = 0.6 * (f0 * f2), A2 = 0.3 * (f0 * f1);
  const double p0 = (2 * q0 - 7 * q1 + 11 * q2) * (1.0 / 6.0), p1 = (-q1 + 5 * q2 + 2 * q3) * (1.0 / 6.0),
               p2 = (2 * q2 + 5 * q3 - q4) * (1.0 / 6.0);
  return (A0 * p0 + A1 * p1 + A2 * p2) * rcp64(A0 + A1 + A2);
}

// Roe flux (Roe 1981) across the face metric h in the contravariant form of
// rusanov(); Roe averages, gamma of the equation of state, Harten's entropy
// fix (delta = 0.1 c) on the two acoustic waves.  s = {rho, rhoU, rhoV, rhoW,
// rhoE, p}.  No counterpart in the reference: the option roe.
__device__ __forceinline__ void roe_flux(const double *L, const double *R, const double *h, double gamma,
                                         double *F) {
  const double irL = rcp64(L[0]), irR = rcp64(R[0]);
  const double uL[3] = {L[1] * irL, L[2] * irL, L[3] * irL}, uR[3] = {R[1] * irR, R[2] * irR, R[3] * irR};
  const double UL = h[0] * uL[0] + h[1] * uL[1] + h[2] * uL[2], UR = h[0] * uR[0] + h[1] * uR[1] + h[2] * uR[2];
  const double gii = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  const double hn = sqrt(gii), ihn = rcp64(hn);
  const double n[3] = {h[0] * ihn, h[1] * ihn, h[2] * ihn};
  const double HL = (L[4] + L[5]) * irL, HR = (R[4] + R[5]) * irR;
  const double sL = sqrt(L[0]), sR = sqrt(R[0]), iws = rcp64(sL + sR);
  double u[3];
#pragma unroll
  for (int c = 0; c < 3; c++) u[c] = (sL * uL[c] + sR * uR[c]) * iws;
  const double H = (sL * HL + sR * HR) * iws, rho = sL * sR;
  const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double c2 = (gamma - 1) * (H - 0.5 * q2), c = sqrt(c2), ic2 = rcp64(c2);
  const double Un = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
  const double dr = R[0] - L[0], dp = R[5] - L[5];
  double du[3];
#pragma unroll
  for (int k = 0; k < 3; k++) du[k] = uR[k] - uL[k];
  const double dUn = n[0] * du[0] + n[1] * du[1] + n[2] * du[2];
  const double a1 = (dp - rho * c * dUn) * (0.5 * ic2), a2 = dr - dp * ic2, a5 = (dp + rho * c * dUn) * (0.5 * ic2);
  const double delta = 0.1 * c;
  auto fix = [&](double lam) {
    const double a = fabs(lam);
    return a < delta ? (lam * lam + delta * delta) * rcp64(2 * delta) : a;
  };
  const double l1 = fix(Un - c), l2 = fabs(Un), l5 = fix(Un + c);
  double dt[3];
#pragma unroll
  for (int k = 0; k < 3; k++) dt[k] = du[k] - dUn * n[k];
  const double udt = u[0] * dt[0] + u[1] * dt[1] + u[2] * dt[2];
  const double half = 0.5 * hn;
  F[0] = 0.5 * (L[0] * UL + R[0] * UR) - half * (l1 * a1 + l2 * a2 + l5 * a5);
#pragma unroll
  for (int k = 0; k < 3; k++)
    F[1 + k] = 0.5 * ((L[1 + k] * UL + L[5] * h[k]) + (R[1 + k] * UR + R[5] * h[k])) -
               half * (l1 * a1 * (u[k] - c * n[k]) + l2 * (a2 * u[k] + rho * dt[k]) + l5 * a5 * (u[k] + c * n[k]));
  F[4] = 0.5 * ((L[4] + L[5]) * UL + (R[4] + R[5]) * UR) -
         half * (l1 * a1 * (H - c * Un) + l2 * (a2 * 0.5 * q2 + rho * udt) + l5 * a5 * (H + c * Un));
}

template <typename T> __device__ __forceinline__ T eno_qr(T qm, T q0, T qp, int k) {
  if (k == 2) return T(0.5) * q0 + T(0.5) * qp;
  return -qm * T(0.1666666666666666) + q0 * T(0.8333333333333333) + qp * T(0.3333333333333333);
}
template <typename T> __device__ __forceinline__ T eno_ql(T qm, T q0, T qp, int k) {
  if (k == 2) return T(0.5) * q0 + T(0.5) * qm;
  return qm * T(0.3333333333333333) + q0 * T(0.8333333333333333) - qp * T(0.1666666666666666);
}
// skew-symmetric 4-point product, rhs.c:2468-2474
template <typename T> __device__ __forceinline__ T fp_2(const T u[4], const T v[4]) {
  return (u[1] + u[2]) * (v[1] + v[2]) * T(1.0 / 3.0) -
         T(1. / 24.) * (u[0] * v[0] + u[0] * v[2] + u[1] * v[1] + u[1] * v[3] + u[2] * v[2] + u[2] * v[0] +
                        u[3] * v[1] + u[3] * v[3]);
}
template <typename T> __device__ __forceinline__ T fp_1(const T u[4], const T v[4]) {
  return T(1.0 / 12.0) * (-u[0] * v[0] + T(7) * u[1] * v[1] + T(7) * u[2] * v[2] - u[3] * v[3]);
}
template <typename T> __device__ __forceinline__ T fp4(const T u[4]) {
  return T(1.0 / 12.0) * (-u[0] + T(7) * u[1] + T(7) * u[2] - u[3]);
}

// Rusanov flux, gamma hard-coded to 1.4 as in rhs.c:2596 / 3256.
template <typename T>
__device__ __forceinline__ void rusanov(const T *L, const T *R, const T *h, T uconL, T uconR, T *F) {
  const T gamma = T(1.4);
  T gii = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  T CsoundL = gamma * L[5] * rcpT(L[0]);
  T CsoundR = gamma * R[5] * rcpT(R[0]);
  T saL = fabs(uconL) + sqrt(fabs(CsoundL) * gii);
  T saR = fabs(uconR) + sqrt(fabs(CsoundR) * gii);
  T SA = pmax(saL, saR);
  F[0] = T(0.5) * (L[0] * uconL + R[0] * uconR) - T(0.5) * SA * (R[0] - L[0]);
  F[1] = T(0.5) * ((L[1] * uconL + R[1] * uconR) + (L[5] * h[0] + R[5] * h[0])) - T(0.5) * SA * (R[1] - L[1]);
  F[2] = T(0.5) * ((L[2] * uconL + R[2] * uconR) + (L[5] * h[1] + R[5] * h[1])) - T(0.5) * SA * (R[2] - L[2]);
  F[3] = T(0.5) * ((L[3] * uconL + R[3] * uconR) + (L[5] * h[2] + R[5] * h[2])) - T(0.5) * SA * (R[3] - L[3]);
  F[4] = T(0.5) * ((L[5] + L[4]) * uconL + (R[5] + R[4]) * uconR) - T(0.5) * SA * (R[4] - L[4]);
}
template <typename T> __device__ __forceinline__ T ucon_of(const T *s, const T *h) {
  const T inv = rcpT(s[0]);
  return h[0] * (s[1] * inv) + h[1] * (s[2] * inv) + h[2] * (s[3] * inv);
}

// Analytic hybrid-flux weight when -calc_sigma is 0 (rhs.c:3328-3444):
// Sigma = (s_x, s_y, s_z) on every node, then on interior nodes the
// curvature test never fires on a uniform field and sigma_y = sigma_z =
// min(s_z, s_y).
struct SigmaSpec {
  double bnd[3];
  double in[3];
};

// ---------------------------------------------------------------------------
// Host-side context.
// ---------------------------------------------------------------------------
}  // namespace cles

// Device data of the immersed body (one surface).
struct cles_ibm_dev {
  int nv, ne;
  double *vx, *vy, *vz;  // vertices
  int *tv;               // 3 planes of ne: vertex ids of each triangle
  double *nf;            // 3 planes of ne: unit normals
  double *qv, *rad;      // bounding spheres
  double *u, *uold;      // 3 planes of nv: surface velocity now / one step ago
  int n_ib;              // interpolation nodes, in the reference's list order
  int *node, *cell, *imode;
  int *don;              // 9 planes: donor (i,j,k) x 3
  long long *dn;         // 3 planes: donor node ids
  double *info;          // 11 planes: pmin xyz, d_s, d_i, cs1..3, cr1..3
  double *nrm;           // 3 planes: wall normal used by the interpolation
  double *w;             // per-sweep work planes
  double *carry;         // Pa carried from one sweep to the next
  double *frec, *fout;   // surface-force records (n_ib of them) and outputs, made at the first call
};

struct cles_batch;

struct cles_ctx {
  cles_params p;
  cles_batch *batch;  // the batch the context is a member of, or null
  cles::Grid g;
  int dev;
  hipStream_t stream;
  // static geometry
  double *coor;             // 3 planes
  double *csi, *eta, *zet;  // 3 planes each
  double *aj, *nvert;       // 1 plane each
  double *fil2;             // (pow(1/aj, 1/3))^2: squared LES filter width per node
  double *mpack;            // LES: csi, eta, zet and aj in blocks of 8 nodes, [np/8][5][8][2]
  double *tfil2;            // squared test-filter width (2 or 5^(1/3) filter widths)
  double ref_rho, ref_p;    // FP32-bulk reference state (node (1,1,1) of the uploaded state)
  int tile;                 // stage-kernel tile shape 0..7 (cles_set_tile), default 7: 8x4x3 in 128 threads
  int has_ib;               // any nvert != 0
  int immersed;             // a surface was classified: RK steps end with the interpolation
  cles_ibm_dev *ibm;
  double *qsnap;            // 6 planes: Q and P before FormBCS(0) of an immersed step
  double *avg;              // -averaging: 12 planes, Q_sum 5, P_sum 1, u^2 3, (uv, vw, wu) 3
  // state: three conserved-variable buffers with their pressures
  double *Q[3], *P[3];
  int cur;       // buffer holding Q^n (= Qold)
  int ucat_src;  // buffer whose velocity the reference's lUcat holds (-1: none yet)
  int ucat_les_mixed;  // the LES model has overwritten interior lUcat since the last RHS
  double *rhs;         // 5 planes, written by cles_form_rhs
  double *xfer;        // 5 planes (+ghosts): host<->device layout staging
  size_t xfer_elems;
  // LES
  double *cs, *cs_alt, *nut;
  // fields of the dynamic model in the precision of the form (double, or
  // float for FP32-bulk; the pointers then address float planes of np)
  double *grad;    // the gradient kernel's 8 values per node in blocks of B = 8 (double) or 16 (float) nodes
                   // ([np/B][4][B][2] of the form's type): pairs (S11, S12), (S13, S22), (S23, S33), (|S|, |S| of nu_t)
  double *lm, *mm;
  double *homo;           // -i_homo_filter: line, plane and global sums of LM and MM
  double *sgm, *sgm_tmp;  // -calc_sigma: 3 planes each, the shock sensor (sgm is the one the fluxes read)
  double *ivol;       // 1/aj, the smoothing weights
  cles::SigmaSpec sig;
  // reductions
  unsigned long long *dnorm;  // [2] bit patterns of non-negative doubles
  double host_norm[2];
  int ti;
  int have_state;
  size_t bytes;
  char err[256];
};

namespace cles {
// launchers implemented in the .hip translation units
int launch_metrics(cles_ctx *c);  // also fills mpack when the context has one
int launch_eos_all(cles_ctx *c, int buf);
int launch_stage(cles_ctx *c, int in, int out, int istage, int mode, const double coef[3]);
int launch_boundary(cles_ctx *c, int in, int out, int istage, double alfa, const double coef[3]);
// bctype b is implemented on side q (0..5 = i-, i+, j-, j+, k-, k+)
bool supported_bc(int side, int b);
int launch_les(cles_ctx *c);
// -calc_sigma: the shock sensor of the state in buffer buf into c->sgm
int launch_sigma(cles_ctx *c, int buf);
// the same on another stream (batched steps); the LES model on stream s with
// its stale state read through the device cell qs_cell
int launch_sigma_on(cles_ctx *c, int buf, hipStream_t s);
int launch_les_on(cles_ctx *c, hipStream_t s, const double *const *qs_cell);
int launch_averaging(cles_ctx *c);
// node coordinates, interleaved x y z in the natural order (global layout)
int get_coor(cles_ctx *c, double *xyz);
// batched stages of several periodic grids (serving)
size_t stage_args_bytes();
int stage_batch_args(cles_ctx *c, int in, int out, int istage, const double coef[3], int tile, void *dst);
// cen: every member has sigma = 1 on every face (the central instance of the LES forms)
hipError_t launch_stage_batch(int tile, int fp32, int visc, int les, int cen, const void *tab, const int *tstart, int nreq,
                              int nblocks,
                              hipStream_t s);
size_t seam_args_bytes();
int seam_batch_args(cles_ctx *c, int in, int out, void *dst);
hipError_t launch_seam_batch(const void *tab, const int *bstart, int nreq, int nblocks, hipStream_t s);

// Persistent batched stages: one launch per (precision, model) group covers
// every stage of every member of the group.  The items (a run of tiles, or
// kSeamItem boundary nodes) sit in eight queues, one per XCD; an item of
// stage s of a member waits until the member's completion counter shows
// every item of its stage s-1 done.
constexpr int kSeamItem = 512;
// the queue heads and completion counters sit 256 B apart: polls and
// increments of different counters go to different memory channels
constexpr int kCtrStride = 64;
struct PersistPhase {  // the share of one (stage, member) in one queue
  int member, stage;
  int t0, nt, grain;  // tiles [t0, t0 + nt), grain tiles per item
  int b0, nb;         // seam items [b0, b0 + nb)
  unsigned need;      // done[member] the items of the phase wait for
  int istart;         // first item of the phase in the queue (tiles, then seams)
};
struct SeamArgs {
  Grid g;
  const double *q;  // stage input
  double *qo, *po;  // stage output
  double gamma;
  long long nb;     // boundary nodes
};
struct PersistArgs {
  const void *stage;          // stage arguments [stage][member slot]
  const SeamArgs *seam;       // [stage][member slot]
  const PersistPhase *phase;  // [queue][pstride], the first nph of each queue used
  const int *nitems;          // [queue]
  int nph, pstride, cap;
  unsigned *head;             // [queue * kCtrStride] next item (zero at launch)
  unsigned *done;             // [member slot * kCtrStride] items done (zero at launch)
  unsigned *abort;            // set when a wait gave up (the heads are then pushed past the end)
  unsigned spin_limit;        // polls of a dependency wait before it gives up
};
hipError_t launch_stage_persist(int tile, int fp32, int visc, int les, const PersistArgs &a, hipStream_t s);
// workgroups of launch_stage_persist that fit on the device at once
int persist_blocks(int tile, int fp32, int visc, int les);
int stage_tiles(cles_ctx *c, int tile);
int rk_coef(const cles_params &p, int istage, double coef[3]);
size_t ctx_footprint(const cles_params *p);
// immersed body
int launch_ibm_interp(cles_ctx *c, int buf, int qold_buf, int add_dundt);
int launch_snapshot(cles_ctx *c, int buf);
int launch_boundary_post(cles_ctx *c, int buf);
void ibm_free(cles_ctx *c);
int set_error(cles_ctx *c, int code, const char *fmt, ...);
// A context in a batch: CLES_ESTATE while the batch runs; a call that
// changes the context (modifies) marks the batch for a rebuild.
int member_guard(cles_ctx *c, bool modifies);
// leave the batch the context is a member of (waits for the batch)
void member_leave(cles_ctx *c);
}  // namespace cles

#define CLES_HIP(c, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return cles::set_error((c), CLES_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call,      \
                             hipGetErrorString(e_));                                          \
  } while (0)

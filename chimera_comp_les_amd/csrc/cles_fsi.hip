// cles_fsi.hip -- surface forces of the immersed body: Calc_forces_SI2
// (fsi.c:1312-2470) for one body on one grid block.
//
// The reference walks its interpolation list once and updates some forty
// running sums statement by statement, so its result depends on the list
// order and on the order of the updates inside each face.  Here a parallel
// gather kernel forms, for every list node, everything local to the node
// (bounds test, viscous stress from the last CalcUCat velocity with the
// one-sided differences of fsi.c:1486-1640, node pressure, and per face the
// area vector and the face position); one thread then replays the
// accumulation in list order with the reference's update order, which keeps
// the sums bit-identical.  The function-scope r_x, r_y, r_z of the reference
// keep the last face's values across nodes; they start at 0 here.
#include <mutex>

#include "cles_internal.hpp"

namespace cles {
namespace {

// per-node record of the gather kernel
constexpr int kRecHead = 11;            // inb, p, Txx Tyy Tzz Tzy Tzx Tyx, nfy, nfz, sb
constexpr int kRecFace = 7;             // present, A x y z, face position x y z
constexpr int kRec = kRecHead + 6 * kRecFace;

struct ForceArgs {
  Grid g;
  int n_ib, invicid;
  const int *node, *cell;
  const double *info;   // d_s is plane 3
  const double *nf;     // 3 planes of triangles
  int ne;
  const double *q;      // CalcUCat input (or null: lUcat still zero)
  int twoD;             // 1: CalcUCat sets u = 0 (rhs.c:1349)
  const double *p, *nvert, *coor, *csi, *eta, *zet, *aj;
  double rei;
  double *rec;          // kRec per list node
};

__device__ __forceinline__ void ucat_at(const ForceArgs &a, long long n, double u[3]) {
  if (!a.q) {
    u[0] = u[1] = u[2] = 0.;
    return;
  }
  const long long np = a.g.np;
  const double r = a.q[n];
  u[0] = a.twoD == 1 ? 0. : a.q[np + n] / r;
  u[1] = a.q[2 * np + n] / r;
  u[2] = a.q[3 * np + n] / r;
}

__global__ __launch_bounds__(256) void force_gather_kernel(ForceArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_ib) return;
  const Grid &g = a.g;
  const long long np = g.np;
  double *rec = a.rec + (long long)m * kRec;
  const long long n = a.node[m];
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  const int mx = g.mx, my = g.my, mz = g.mz;
  // fsi.c:1478 on one rank: lxs = 1, lxe = mx-1, ys = 0, lye = my-1, lzs = 1, lze = mz-1
  const bool inb = i >= 1 && i < mx - 1 && j >= 0 && j < my - 1 && k >= 1 && k < mz - 1;
  rec[0] = inb ? 1. : 0.;
  if (!inb) return;
  auto NV = [&](int kk, int jj, int ii) { return a.nvert[nid(g, kk, jj, ii)]; };
  auto AJ = [&](int kk, int jj, int ii) { return a.aj[nid(g, kk, jj, ii)]; };
  auto MT = [&](const double *mt, int kk, int jj, int ii, int c) { return mt[c * np + nid(g, kk, jj, ii)]; };
  double T[6] = {0., 0., 0., 0., 0., 0.};  // Txx Tyy Tzz Tzy Tzx Tyx
  if (!a.invicid) {
    double dwdz, dvdz, dudz, dwdy, dvdy, dudy, dwdx, dvdx, dudx, g1, g2, g3;
    double up[3], um[3];
    // k differences (fsi.c:1486-1547)
    if (NV(k + 1, j, i) < 2.5 && NV(k - 1, j, i) < 2.5 && k + 1 < mz - 1 && k - 1 > 0) {
      g1 = 0.25 * ((MT(a.zet, k, j, i, 0) + MT(a.zet, k - 1, j, i, 0)) * (AJ(k, j, i) + AJ(k - 1, j, i)));
      g2 = 0.25 * ((MT(a.zet, k, j, i, 1) + MT(a.zet, k - 1, j, i, 1)) * (AJ(k, j, i) + AJ(k - 1, j, i)));
      g3 = 0.25 * ((MT(a.zet, k, j, i, 2) + MT(a.zet, k - 1, j, i, 2)) * (AJ(k, j, i) + AJ(k - 1, j, i)));
      ucat_at(a, nid(g, k + 1, j, i), up);
      ucat_at(a, nid(g, k - 1, j, i), um);
      dwdz = (up[2] - um[2]) / 2. * g3;
      dvdz = (up[1] - um[1]) / 2. * g3;
      dudz = (up[0] - um[0]) / 2. * g3;
      dwdy = (up[2] - um[2]) / 2. * g2;
      dvdy = (up[1] - um[1]) / 2. * g2;
      dudy = (up[0] - um[0]) / 2. * g2;
      dwdx = (up[2] - um[2]) / 2. * g1;
      dvdx = (up[1] - um[1]) / 2. * g1;
      dudx = (up[0] - um[0]) / 2. * g1;
    } else if ((NV(k + 1, j, i) < 2.5 && k + 1 < mz - 1) || (NV(k - 1, j, i) < 2.5 && k - 1 > 0)) {
      const bool fw = NV(k + 1, j, i) < 2.5 && k + 1 < mz - 1;  // forward from the node, else backward
      const int kb = fw ? k : k - 1;
      g1 = MT(a.zet, kb, j, i, 0) * AJ(kb, j, i);
      g2 = MT(a.zet, kb, j, i, 1) * AJ(kb, j, i);
      g3 = MT(a.zet, kb, j, i, 2) * AJ(kb, j, i);
      ucat_at(a, nid(g, kb + 1, j, i), up);
      ucat_at(a, nid(g, kb, j, i), um);
      dwdz = (up[2] - um[2]) * g3;
      dvdz = (up[1] - um[1]) * g3;
      dudz = (up[0] - um[0]) * g3;
      dwdy = (up[2] - um[2]) * g2;
      dvdy = (up[1] - um[1]) * g2;
      dudy = (up[0] - um[0]) * g2;
      dwdx = (up[2] - um[2]) * g1;
      dvdx = (up[1] - um[1]) * g1;
      dudx = (up[0] - um[0]) * g1;
    } else {
      dwdz = dvdz = dudz = dwdy = dvdy = dudy = dwdx = dvdx = dudx = 0.;
    }
    // j differences added (fsi.c:1550-1593), then i differences (fsi.c:1594-1637)
#pragma unroll
    for (int dir = 1; dir >= 0; dir--) {
      const int dj = dir == 1, di = dir == 0;
      const int idx = dir == 1 ? j : i, lim = dir == 1 ? my : mx;
      const double *mt = dir == 1 ? a.eta : a.csi;
      const bool hp = NV(k, j + dj, i + di) < 2.5 && idx + 1 < lim - 1;
      const bool hm = NV(k, j - dj, i - di) < 2.5 && idx - 1 > 0;
      if (hp && hm) {
        g1 = 0.25 * (MT(mt, k, j, i, 0) + MT(mt, k, j - dj, i - di, 0)) * (AJ(k, j, i) + AJ(k, j - dj, i - di));
        g2 = 0.25 * (MT(mt, k, j, i, 1) + MT(mt, k, j - dj, i - di, 1)) * (AJ(k, j, i) + AJ(k, j - dj, i - di));
        g3 = 0.25 * (MT(mt, k, j, i, 2) + MT(mt, k, j - dj, i - di, 2)) * (AJ(k, j, i) + AJ(k, j - dj, i - di));
        ucat_at(a, nid(g, k, j + dj, i + di), up);
        ucat_at(a, nid(g, k, j - dj, i - di), um);
        dwdz += (up[2] - um[2]) / 2. * g3;
        dvdz += (up[1] - um[1]) / 2. * g3;
        dudz += (up[0] - um[0]) / 2. * g3;
        dwdy += (up[2] - um[2]) / 2. * g2;
        dvdy += (up[1] - um[1]) / 2. * g2;
        dudy += (up[0] - um[0]) / 2. * g2;
        dwdx += (up[2] - um[2]) / 2. * g1;
        dvdx += (up[1] - um[1]) / 2. * g1;
        dudx += (up[0] - um[0]) / 2. * g1;
      } else if (hp || hm) {
        const int jb = hp ? j : j - dj, ib = hp ? i : i - di;  // lower node of the one-sided difference
        g1 = MT(mt, k, jb, ib, 0) * AJ(k, jb, ib);
        g2 = MT(mt, k, jb, ib, 1) * AJ(k, jb, ib);
        g3 = MT(mt, k, jb, ib, 2) * AJ(k, jb, ib);
        ucat_at(a, nid(g, k, jb + dj, ib + di), up);
        ucat_at(a, nid(g, k, jb, ib), um);
        dwdz += (up[2] - um[2]) * g3;
        dvdz += (up[1] - um[1]) * g3;
        dudz += (up[0] - um[0]) * g3;
        dwdy += (up[2] - um[2]) * g2;
        dvdy += (up[1] - um[1]) * g2;
        dudy += (up[0] - um[0]) * g2;
        dwdx += (up[2] - um[2]) * g1;
        dvdx += (up[1] - um[1]) * g1;
        dudx += (up[0] - um[0]) * g1;
      }
    }
    T[0] = a.rei * (dudx + dudx);
    T[1] = a.rei * (dvdy + dvdy);
    T[2] = a.rei * (dwdz + dwdz);
    T[3] = a.rei * (dwdy + dvdz);
    T[4] = a.rei * (dwdx + dudz);
    T[5] = a.rei * (dvdx + dudy);
  }
  const long long nn = nid(g, k, j, i);
  rec[1] = a.p[nn];
#pragma unroll
  for (int q = 0; q < 6; q++) rec[2 + q] = T[q];
  const int e = a.cell[m];
  rec[8] = a.nf[a.ne + e];
  rec[9] = a.nf[2 * a.ne + e];
  rec[10] = a.info[3 * a.n_ib + m];
  // faces k+, k-, j+, j-, i+, i- (fsi.c:1660-2290)
#pragma unroll
  for (int f = 0; f < 6; f++) {
    double *fr = rec + kRecHead + f * kRecFace;
    const int d = f / 2;
    const bool plus = (f & 1) == 0;
    const int s = plus ? 1 : -1;
    const int dk = d == 0 ? s : 0, dj = d == 1 ? s : 0, di = d == 2 ? s : 0;
    const double *mt = d == 0 ? a.zet : d == 1 ? a.eta : a.csi;
    const long long n1 = nid(g, k + dk, j + dj, i + di);
    if (d == 1 && !plus && j == 0) {  // fsi.c:1921-1927: the node's own metric and coordinates
      fr[0] = 1.;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        fr[1 + c] = mt[c * np + nn];
        fr[4 + c] = a.coor[c * np + nn];
      }
    } else if (a.nvert[n1] < 0.9) {
      fr[0] = 1.;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        fr[1 + c] = 0.5 * (mt[c * np + nn] + mt[c * np + n1]);
        fr[4 + c] = 0.5 * (a.coor[c * np + nn] + a.coor[c * np + n1]);
      }
    } else {
      fr[0] = 0.;
    }
  }
}

// One thread replays the list in order (the accumulation is the
// reference's, statement by statement).
__global__ void force_sum_kernel(const double *rec, int n_ib, double X_c, double Y_c, double Z_c, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double Cp_px = 0, Cp_py = 0, Cp_pz = 0, Cs_px = 0, Cs_py = 0, Cs_pz = 0;
  double Cp_nx = 0, Cp_ny = 0, Cp_nz = 0, Cs_nx = 0, Cs_ny = 0, Cs_nz = 0;
  double Ap_x = 0, Ap_y = 0, Ap_z = 0, An_x = 0, An_y = 0, An_z = 0;
  double Iap_x = 0, Iap_y = 0, Iap_z = 0, Ian_x = 0, Ian_y = 0, Ian_z = 0;
  double M_px = 0, M_py = 0, M_pz = 0, M_nx = 0, M_ny = 0, M_nz = 0;
  double Mdpdn_px = 0, Mdpdn_py = 0, Mdpdn_pz = 0, Mdpdn_nx = 0, Mdpdn_ny = 0, Mdpdn_nz = 0;
  double r_x = 0, r_y = 0, r_z = 0;
  for (int m = 0; m < n_ib; m++) {
    const double *rc = rec + (long long)m * kRec;
    if (rc[0] == 0.) continue;
    const double p = rc[1];
    const double Txx = rc[2], Tyy = rc[3], Tzz = rc[4], Tzy = rc[5], Tzx = rc[6], Tyx = rc[7];
    const double nfy = rc[8], nfz = rc[9], sb = rc[10];
    double MFdpdn_px = 0, MFdpdn_py = 0, MFdpdn_pz = 0, MFdpdn_nx = 0, MFdpdn_ny = 0, MFdpdn_nz = 0;
    for (int f = 0; f < 6; f++) {
      const double *fr = rc + kRecHead + f * kRecFace;
      if (fr[0] == 0.) continue;
      const double A_x = fr[1], A_y = fr[2], A_z = fr[3];
      r_x = fr[4] - X_c;
      r_y = fr[5] - Y_c;
      r_z = fr[6] - Z_c;
      double MF_x = 0., MF_y = 0., MF_z = 0.;
      if ((f & 1) == 0) {
        Cp_pz += -p * A_z; MF_z += -p * A_z; Ap_z += A_z;
        Iap_x -= r_y * A_z; Iap_y -= -r_x * A_z;
        MFdpdn_pz += sb * (nfy * r_z - nfz * r_y) * A_z;
        Cp_py += -p * A_y; MF_y += -p * A_y; Ap_y += A_y;
        Iap_x -= -r_z * A_y; Iap_z -= r_x * A_y;
        MFdpdn_py += sb * (nfy * r_z - nfz * r_y) * A_y;
        Cp_px += -p * A_x; MF_x += -p * A_x; Ap_x += A_x;
        Iap_z -= -r_y * A_x; Iap_y -= -(-r_z * A_x);
        MFdpdn_px += sb * (nfy * r_z - nfz * r_y) * A_x;
        Cs_pz += Tzz * A_z; Cs_py += Tzy * A_z; Cs_px += Tzx * A_z;
        MF_z += Tzz * A_z; MF_y += Tzy * A_z; MF_x += Tzx * A_z;
        Cs_pz += Tzy * A_y; Cs_py += Tyy * A_y; Cs_px += Tyx * A_y;
        MF_z += Tzy * A_y; MF_y += Tyy * A_y; MF_x += Tyx * A_y;
        Cs_pz += Tzx * A_x; Cs_py += Tyx * A_x; Cs_px += Txx * A_x;
        MF_z += Tzx * A_x; MF_y += Tyx * A_x; MF_x += Txx * A_x;
        M_px -= r_y * MF_z - r_z * MF_y;
        M_py -= -(r_x * MF_z - r_z * MF_x);
        M_pz -= r_x * MF_y - r_y * MF_x;
      } else {
        Cp_nz += p * A_z; MF_z += p * A_z; An_z += A_z;
        Ian_x -= r_y * A_z; Ian_y -= -r_x * A_z;
        MFdpdn_nz += -sb * (nfy * r_z - nfz * r_y) * A_z;
        Cp_ny += p * A_y; MF_y += p * A_y; An_y += A_y;
        Ian_x -= -r_z * A_y; Ian_z -= r_x * A_y;
        MFdpdn_ny += -sb * (nfy * r_z - nfz * r_y) * A_y;
        Cp_nx += p * A_x; MF_x += p * A_x; An_x += A_x;
        Ian_z -= -r_y * A_x; Ian_y -= -(-r_z * A_x);
        MFdpdn_nx += -sb * (nfy * r_z - nfz * r_y) * A_x;
        Cs_nz -= Tzz * A_z; Cs_ny -= Tzy * A_z; Cs_nx -= Tzx * A_z;
        MF_z -= Tzz * A_z; MF_y -= Tzy * A_z; MF_x -= Tzx * A_z;
        Cs_nz -= Tzy * A_y; Cs_ny -= Tyy * A_y; Cs_nx -= Tyx * A_y;
        MF_z -= Tzy * A_y; MF_y -= Tyy * A_y; MF_x -= Tyx * A_y;
        Cs_nz -= Tzx * A_x; Cs_ny -= Tyx * A_x; Cs_nx -= Txx * A_x;
        MF_z -= Tzx * A_x; MF_y -= Tyx * A_x; MF_x -= Txx * A_x;
        M_nx -= r_y * MF_z - r_z * MF_y;
        M_ny -= -(r_x * MF_z - r_z * MF_x);
        M_nz -= r_x * MF_y - r_y * MF_x;
      }
    }
    Mdpdn_px -= r_y * MFdpdn_pz - r_z * MFdpdn_py;
    Mdpdn_py -= -(r_x * MFdpdn_pz - r_z * MFdpdn_px);
    Mdpdn_pz -= r_x * MFdpdn_py - r_y * MFdpdn_px;
    Mdpdn_nx -= r_y * MFdpdn_nz - r_z * MFdpdn_ny;
    Mdpdn_ny -= -(r_x * MFdpdn_nz - r_z * MFdpdn_nx);
    Mdpdn_nz -= r_x * MFdpdn_ny - r_y * MFdpdn_nx;
  }
  // totals (fsi.c:2300-2420, one rank)
  const double F_px = Cp_px + Cs_px, F_py = Cp_py + Cs_py, F_pz = Cp_pz + Cs_pz;
  const double F_nx = Cp_nx + Cs_nx, F_ny = Cp_ny + Cs_ny, F_nz = Cp_nz + Cs_nz;
  out[0] = F_px + F_nx; out[1] = F_py + F_ny; out[2] = F_pz + F_nz;
  out[3] = Cp_px + Cp_nx; out[4] = Cp_py + Cp_ny; out[5] = Cp_pz + Cp_nz;
  out[6] = 0.5 * (Ap_x + An_x); out[7] = 0.5 * (Ap_y + An_y); out[8] = 0.5 * (Ap_z + An_z);
  out[9] = Ap_x + Ap_y + Ap_z;
  out[10] = M_px + M_nx; out[11] = M_py + M_ny; out[12] = M_pz + M_nz;
  out[13] = Mdpdn_px + Mdpdn_nx; out[14] = Mdpdn_py + Mdpdn_ny; out[15] = Mdpdn_pz + Mdpdn_nz;
  out[16] = 0.5 * (Iap_x - Ian_x); out[17] = 0.5 * (Iap_y - Ian_y); out[18] = 0.5 * (Iap_z - Ian_z);
}

}  // namespace
}  // namespace cles

using namespace cles;

namespace cles {
namespace {

// Gather the per-node records into rec (kRec per list node) and fold them
// into dout (CLES_FORCE_N), then copy the outputs to the host.
int forces_run(cles_ctx *c, const double center[3], double out[CLES_FORCE_N], double *rec, double *dout) {
  cles_ibm_dev *b = c->ibm;
  ForceArgs a;
  a.g = c->g;
  a.n_ib = b->n_ib;
  a.invicid = c->p.invicid;
  a.node = b->node; a.cell = b->cell; a.info = b->info; a.nf = b->nf; a.ne = b->ne;
  // lUcat as the body sees it at the start of a step: the velocity of the
  // last CalcUCat (its stage input), zero before the first
  a.q = c->ucat_src >= 0 ? c->Q[c->ucat_src] : nullptr;
  a.twoD = c->p.twoD;
  a.p = c->P[c->cur];
  a.nvert = c->nvert; a.coor = c->coor; a.csi = c->csi; a.eta = c->eta; a.zet = c->zet; a.aj = c->aj;
  a.rei = 1. / c->p.ren;
  a.rec = rec;
  hipError_t e = hipSuccess;
  hipLaunchKernelGGL(force_gather_kernel, dim3((unsigned)((b->n_ib + 255) / 256)), dim3(256), 0, c->stream, a);
  e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(force_sum_kernel, dim3(1), dim3(64), 0, c->stream, rec, b->n_ib, center[0], center[1],
                       center[2], dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, CLES_FORCE_N * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_error(c, CLES_EDEVICE, "surface forces: %s", hipGetErrorString(e));
  return CLES_OK;
}

}  // namespace
}  // namespace cles

extern "C" int cles_ibm_forces(cles_ctx *c, const double center[3], double out[CLES_FORCE_N]) {
  if (!c || !center || !out) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  cles_ibm_dev *b = c->ibm;
  if (!b) return set_error(c, CLES_ESTATE, "no immersed body (cles_ibm_search)");
  for (int q = 0; q < CLES_FORCE_N; q++) out[q] = 0.;
  if (b->n_ib == 0) return CLES_OK;
  // work buffers of the body's list, made at the first call and released
  // with the body (a new cles_ibm_search makes a new list)
  if (!b->frec) {
    if (hipMalloc((void **)&b->frec, (size_t)b->n_ib * kRec * sizeof(double)) != hipSuccess) {
      b->frec = nullptr;
      return set_error(c, CLES_ENOMEM, "force work buffers");
    }
  }
  if (!b->fout) {
    if (hipMalloc((void **)&b->fout, CLES_FORCE_N * sizeof(double)) != hipSuccess) {
      b->fout = nullptr;
      return set_error(c, CLES_ENOMEM, "force work buffers");
    }
  }
  return forces_run(c, center, out, b->frec, b->fout);
}

// ---------------------------------------------------------------------------
// Allocation-free form.  The records and outputs live in a file-scope device
// array: every GPU the library runs on holds its own copy of it (about 7 MB)
// until the process ends.  Its device address is looked up once per device
// with that device current; host code passes that address, never the
// array's name, to the kernels and copies.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchDoubles = (size_t)CLES_FORCE_NOALLOC_MAX_NODES * cles::kRec + CLES_FORCE_N;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

static int scratch_address(cles_ctx *c, double **ptr) {
  if (c->dev < 0 || c->dev >= kMaxDevices) return cles::set_error(c, CLES_EINVAL, "device %d out of range", c->dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->dev]) {
    int cur = -1;
    CLES_HIP(c, hipGetDevice(&cur));
    if (cur != c->dev) CLES_HIP(c, hipSetDevice(c->dev));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != c->dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return cles::set_error(c, CLES_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[c->dev] = (double *)p;
  }
  *ptr = g_scratch_ptr[c->dev];
  return CLES_OK;
}

extern "C" int cles_ibm_forces_noalloc(cles_ctx *c, const double center[3], double out[CLES_FORCE_N]) {
  if (!c || !center || !out) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return cles::set_error(c, CLES_ESTATE, "no state uploaded");
  cles_ibm_dev *b = c->ibm;
  if (!b) return cles::set_error(c, CLES_ESTATE, "no immersed body (cles_ibm_search)");
  for (int q = 0; q < CLES_FORCE_N; q++) out[q] = 0.;
  if (b->n_ib == 0) return CLES_OK;
  if (b->n_ib > CLES_FORCE_NOALLOC_MAX_NODES)
    return cles::set_error(c, CLES_EINVAL,
                           "cles_ibm_forces_noalloc: %d interpolation nodes, more than the %d its static "
                           "scratch holds (use cles_ibm_forces)", b->n_ib, CLES_FORCE_NOALLOC_MAX_NODES);
  double *base = nullptr;
  int r = scratch_address(c, &base);
  if (r) return r;
  // every record field the fold reads is written by the gather of this call
  return cles::forces_run(c, center, out, base, base + (size_t)CLES_FORCE_NOALLOC_MAX_NODES * cles::kRec);
}

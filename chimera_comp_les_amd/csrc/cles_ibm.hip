// cles_ibm.hip -- immersed body on gfx950: classification of the grid nodes
// against a closed triangulated surface by ray casting, the interpolation
// stencils of the boundary nodes, and the interpolation sweeps that run after
// the Runge-Kutta stages.  Restates ibm_search_advanced (ibm_iso.c:87-795)
// with point_cell_advanced (:1284), BoundingSphere (:1491), nearestcell1
// (:1581), ICP (:1875), InterceptionPoint (:1975), the geometry helpers of
// compgeom.c and ibm_interpolation_advanced (ibm_iso.c:2572-2880) with the
// wall models noslip / freeslip (wallfunction.c:12-46).  One body (ibi = 0),
// no wall function, no thin-body pass.
//
// Classification and stencils are integer / selection results; every
// floating-point expression that decides them is written in the reference's
// operation order and the library is built without contraction, so they are
// bit-identical to the reference's.  The interpolation carries its pressure
// variable Pa from one list node to the next (the reference sets it only in
// the first sweep, ibm_iso.c:2743); that serial chain is evaluated by one
// workgroup: an affine-map prefix scan estimates the entry value of every
// chunk of the list, exact serial replays of the chunks are then iterated to
// the bit-identical serial result.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cles_internal.hpp"
#include "glibc_rand.hpp"

namespace cles {
namespace {

constexpr int kNc = 40;         // bucket grid ncx = ncy = ncz (ibm_iso.c:104)
constexpr int kMaxSeed = 1024;  // ray directions tried per node before giving up
constexpr int kMaxHits = 48;    // crossings kept per ray
constexpr int kChainThreads = 1024;

enum { F_HITS = 0, F_SEED, F_NEAREST, F_STENCIL, F_NFLAGS };

struct Surf {
  int nv, ne;
  const double *vx, *vy, *vz;  // vertices
  const int *tv;               // 3 planes of ne: vertex ids
  const double *nf;            // 3 planes of ne: unit normals
  const double *qv, *rad;      // bounding spheres: 3 planes of ne, ne
};

struct P3 {
  double x, y, z;
};
struct P2 {
  double x, y;
};

__device__ __forceinline__ P3 vtx(const Surf &s, int n) { return P3{s.vx[n], s.vy[n], s.vz[n]}; }
__device__ __forceinline__ double dist(P3 a, P3 b) {
  return sqrt((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z));
}

// ---------------------------------------------------------------- geometry

// intsect_triangle, compgeom.c:17-53
__device__ int intsect(const double o[3], const double d[3], P3 a, P3 b, P3 c, double &t) {
  const double e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
  const double e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
  const double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
  const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if (det > -0.00000001 && det < 0.00000001) return 0;
  const double inv = 1.0 / det;
  const double tv[3] = {o[0] - a.x, o[1] - a.y, o[2] - a.z};
  const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
  if (u < 0.0 || u > 1.0) return 0;
  const double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
  const double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
  if (v < 0.0 || u + v > 1.0) return 0;
  t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
  return 1;
}

// ISSameSide2D (compgeom.c:55-103), effective test only: its first t1/t2
// assignment is overwritten unconditionally
__device__ int same_side(P2 p, P2 p1, P2 p2, P2 p3) {
  const double t1 = (p.x - p1.x) * (p2.y - p1.y) - (p.y - p1.y) * (p2.x - p1.x);
  const double t2 = (p3.x - p1.x) * (p2.y - p1.y) - (p3.y - p1.y) * (p2.x - p1.x);
  const double lt = sqrt((p1.x - p2.x) * (p1.x - p2.x) + (p1.y - p2.y) * (p1.y - p2.y));
  if (fabs(t1 / lt) < 1.e-10) return 1;
  if (t1 > 0) return t2 > 0 ? 1 : -1;
  return t2 < 0 ? 1 : -1;
}

// ISInsideTriangle2D + ISPointInTriangle, compgeom.c:105-157
__device__ int point_in_triangle(P3 p, P3 a, P3 b, P3 c, double nx, double ny, double nz) {
  P2 q, qa, qb, qc;
  if (fabs(nz) >= fabs(nx) && fabs(nz) >= fabs(ny)) {
    q = P2{p.x, p.y}; qa = P2{a.x, a.y}; qb = P2{b.x, b.y}; qc = P2{c.x, c.y};
  } else if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz)) {
    q = P2{p.z, p.y}; qa = P2{a.z, a.y}; qb = P2{b.z, b.y}; qc = P2{c.z, c.y};
  } else {
    q = P2{p.x, p.z}; qa = P2{a.x, a.z}; qb = P2{b.x, b.z}; qc = P2{c.x, c.z};
  }
  int ls = same_side(q, qa, qb, qc);
  if (ls < 0) return ls;
  ls = same_side(q, qb, qc, qa);
  if (ls < 0) return ls;
  return same_side(q, qc, qa, qb);
}

// Dis_P_Line, compgeom.c:160-186
__device__ void dis_p_line(P3 p, P3 p1, P3 p2, P3 &po, double &d) {
  const double dx21 = p2.x - p1.x, dy21 = p2.y - p1.y, dz21 = p2.z - p1.z;
  const double dx31 = p.x - p1.x, dy31 = p.y - p1.y, dz31 = p.z - p1.z;
  const double t = (dx31 * dx21 + dy31 * dy21 + dz31 * dz21) / (dx21 * dx21 + dy21 * dy21 + dz21 * dz21);
  if (t < 0) {
    po = p1;
    d = sqrt(dx31 * dx31 + dy31 * dy31 + dz31 * dz31);
  } else if (t > 1) {
    po = p2;
    d = sqrt((p.x - po.x) * (p.x - po.x) + (p.y - po.y) * (p.y - po.y) + (p.z - po.z) * (p.z - po.z));
  } else {
    po = P3{p1.x + t * dx21, p1.y + t * dy21, p1.z + t * dz21};
    d = sqrt((p.x - po.x) * (p.x - po.x) + (p.y - po.y) * (p.y - po.y) + (p.z - po.z) * (p.z - po.z));
  }
}

// triangle_intp / triangle_intp2, compgeom.c:188-229
__device__ void tri_bary(P2 p, P2 p1, P2 p2, P2 p3, double &c1, double &c2, double &c3) {
  const double x13 = p1.x - p3.x, y13 = p1.y - p3.y;
  const double x23 = p2.x - p3.x, y23 = p2.y - p3.y;
  const double xp3 = p.x - p3.x, yp3 = p.y - p3.y;
  const double a = x13 * y23 - x23 * y13;
  c1 = (y23 * xp3 - x23 * yp3) / a;
  c2 = (-y13 * xp3 + x13 * yp3) / a;
  c3 = 1. - c1 - c2;
}

// barycentric weights in the plane of the largest normal component
// (ibm_iso.c:1696-1719, :1935-1957)
__device__ void bary_by_normal(P3 q, P3 a, P3 b, P3 c, double nx, double ny, double nz, double &c1, double &c2,
                               double &c3) {
  if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz))
    tri_bary(P2{q.y, q.z}, P2{a.y, a.z}, P2{b.y, b.z}, P2{c.y, c.z}, c1, c2, c3);
  else if (fabs(ny) >= fabs(nx) && fabs(ny) >= fabs(nz))
    tri_bary(P2{q.x, q.z}, P2{a.x, a.z}, P2{b.x, b.z}, P2{c.x, c.z}, c1, c2, c3);
  else if (fabs(nz) >= fabs(ny) && fabs(nz) >= fabs(nx))
    tri_bary(P2{q.y, q.x}, P2{a.y, a.x}, P2{b.y, b.x}, P2{c.y, c.x}, c1, c2, c3);
}

// ------------------------------------------------------------ surface prep

// unit normals from the vertex order (ibm_io.c:1716-1731) and bounding
// spheres (BoundingSphere, ibm_iso.c:1491-1579), one thread per triangle
__global__ __launch_bounds__(256) void surface_kernel(Surf s, double *nf, double *qv, double *rad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= s.ne) return;
  const int ne = s.ne;
  const int n1 = s.tv[e], n2 = s.tv[ne + e], n3 = s.tv[2 * ne + e];
  const P3 p1 = vtx(s, n1), p2 = vtx(s, n2), p3 = vtx(s, n3);
  {
    const double dx12 = p2.x - p1.x, dy12 = p2.y - p1.y, dz12 = p2.z - p1.z;
    const double dx13 = p3.x - p1.x, dy13 = p3.y - p1.y, dz13 = p3.z - p1.z;
    const double nx = dy12 * dz13 - dz12 * dy13;
    const double ny = -dx12 * dz13 + dz12 * dx13;
    const double nz = dx12 * dy13 - dy12 * dx13;
    const double dr = sqrt(nx * nx + ny * ny + nz * nz);
    nf[e] = nx / dr;
    nf[ne + e] = ny / dr;
    nf[2 * ne + e] = nz / dr;
  }
  const double l12 = dist(p1, p2), l23 = dist(p2, p3), l31 = dist(p3, p1);
  P3 pa, pb, pc;
  if (l12 > l23) {
    if (l12 > l31) { pa = p1; pb = p2; pc = p3; }
    else { pa = p3; pb = p1; pc = p2; }
  } else {
    if (l31 < l23) { pa = p2; pb = p3; pc = p1; }
    else { pa = p3; pb = p1; pc = p2; }
  }
  const P3 p0{0, 0, 0};
  const P3 pf{0.5 * (pa.x + pb.x), 0.5 * (pa.y + pb.y), 0.5 * (pa.z + pb.z)};
  const P3 pu{pa.x - pf.x, pa.y - pf.y, pa.z - pf.z};
  const P3 pv{pc.x - pf.x, pc.y - pf.y, pc.z - pf.z};
  const P3 pt{pu.y * pv.z - pu.z * pv.y, pu.z * pv.x - pu.x * pv.z, pu.x * pv.y - pu.y * pv.x};
  const P3 pd{pt.y * pu.z - pt.z * pu.y, pt.z * pu.x - pt.x * pu.z, pt.x * pu.y - pt.y * pu.x};
  double gama = -(dist(pu, p0) * dist(pu, p0) - dist(pv, p0) * dist(pv, p0));
  const P3 pw{pv.x - pu.x, pv.y - pu.y, pv.z - pu.z};
  double lamda = 2 * (pd.x * pw.x + pd.y * pw.y + pd.z * pw.z);
  gama /= lamda;
  lamda = gama < 0 ? 0 : gama;
  const P3 q{pf.x + lamda * pd.x, pf.y + lamda * pd.y, pf.z + lamda * pd.z};
  qv[e] = q.x;
  qv[ne + e] = q.y;
  qv[2 * ne + e] = q.z;
  rad[e] = dist(q, pa);
}

// --------------------------------------------------------- classification

struct ClassArgs {
  Grid g;
  const double *coor;
  double *nvert;
  Surf s;
  double xmin, xmax, ymin, ymax, zmin, zmax, dcx, dcy, dcz;
  const int *col_start, *col_tri, *col_kmin, *col_kmax;  // per bucket column, unique triangles
  const double *sdraw;                                   // rand()/(RAND_MAX+1)-0.5 for srand(1..)
  int *flags;
  unsigned long long *first_bad;
};

// point_cell_advanced, ibm_iso.c:1284-1396: crossings of the ray from p to a
// point above the body in p's bucket column; a repeated crossing (two hits at
// the same parameter on non-opposite faces) retries with the next seed.
// Returns 4 (inside), 0 (outside), -1 (seeds exhausted), -2 (too many hits).
__device__ int ray_cast(const ClassArgs &a, P3 p, int ic, int jc, int kc) {
  const int col = jc * kNc + ic;
  const int b0 = a.col_start[col], b1 = a.col_start[col + 1];
  double ht[kMaxHits], hx[kMaxHits], hy[kMaxHits], hz[kMaxHits];
  const double xpc = a.dcx * (ic + 0.5) + a.xmin;
  const double ypc = a.dcy * (jc + 0.5) + a.ymin;
  for (int seed = 1; seed <= kMaxSeed; seed++) {
    const double s = a.sdraw[seed - 1];
    double dir[3] = {(xpc + s * a.dcx) - p.x, (ypc + s * a.dcy) - p.y, (a.zmax + 0.2) - p.z};
    const double sn = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    dir[0] /= sn;
    dir[1] /= sn;
    dir[2] /= sn;
    const double orig[3] = {p.x, p.y, p.z};
    int nh = 0;
    bool singular = false;
    for (int q = b0; q < b1 && !singular; q++) {
      const int kl = a.col_kmin[q] > kc ? a.col_kmin[q] : kc;
      if (!(kl < a.col_kmax[q])) continue;  // not in a bucket at or above kc
      const int e = a.col_tri[q], ne = a.s.ne;
      double t = 0;
      const int hit = intsect(orig, dir, vtx(a.s, a.s.tv[e]), vtx(a.s, a.s.tv[ne + e]), vtx(a.s, a.s.tv[2 * ne + e]), t);
      if (hit > 0 && t > 0) {
        if (nh == kMaxHits) return -2;
        const double nx = a.s.nf[e], ny = a.s.nf[ne + e], nz = a.s.nf[2 * ne + e];
        for (int z = 0; z < nh; z++) {
          const double ndotn = hx[z] * nx + hy[z] * ny + hz[z] * nz;
          if (fabs(t - ht[z]) < 1.e-8 && ndotn > -0.97) singular = true;
        }
        ht[nh] = t;
        hx[nh] = nx;
        hy[nh] = ny;
        hz[nh] = nz;
        nh++;
      }
    }
    if (!singular) return (nh & 1) ? 4 : 0;
  }
  return -1;
}

__device__ __forceinline__ void flag_node(int *flags, unsigned long long *first_bad, int f, long long n) {
  atomicOr(&flags[f], 1);
  atomicMin(first_bad, (unsigned long long)n);
}

// ibm_iso.c:234-250: reset, bounding-box test, ray casting
__global__ __launch_bounds__(256) void classify_kernel(ClassArgs a) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid &g = a.g;
  if (n >= (long long)g.mx * g.my * g.mz) return;
  double nv = a.nvert[n];
  if (nv < 5.) nv = 0;
  const P3 p{a.coor[n], a.coor[n + g.np], a.coor[n + 2 * g.np]};
  if (p.x > a.xmin && p.x < a.xmax && p.y > a.ymin && p.y < a.ymax && p.z > a.zmin && p.z < a.zmax) {
    const int ic = (int)floor((p.x - a.xmin) / a.dcx);
    const int jc = (int)floor((p.y - a.ymin) / a.dcy);
    const int kc = (int)floor((p.z - a.zmin) / a.dcz);
    const int r = ray_cast(a, p, ic, jc, kc);
    if (r < 0) {
      flag_node(a.flags, a.first_bad, r == -1 ? F_SEED : F_HITS, n);
    } else {
      nv = pmax(nv, (double)r);
      if (nv < 0) nv = 0;
    }
  }
  a.nvert[n] = nv;
}

// near the immersed boundary, ibm_iso.c:349-381 (reads the classification,
// writes the marked copy)
__global__ __launch_bounds__(256) void mark_kernel(Grid g, const double *snap, double *nvert) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  double nv = snap[n];
  if (nv < 0) nv = 0;
  const int ip = i < g.mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
  const int jp = j < g.my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
  const int kp = k < g.mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
  if ((int)(nv + 0.5) != 4)
    for (int kk = km; kk < kp + 1; kk++)
      for (int jj = jm; jj < jp + 1; jj++)
        for (int ii = im; ii < ip + 1; ii++)
          if ((int)(snap[nid(g, kk, jj, ii)] + 0.5) == 4) nv = pmax(2.0, nv);
  nvert[n] = nv;
}

// islanding, ibm_iso.c:397-486: a fluid node with more than three
// boundary/solid face neighbours joins the boundary.  The reference's second
// pass reads the same lNvert as the first and changes nothing more.
__global__ __launch_bounds__(256) void island_kernel(Grid g, const double *snap, double *nvert, int lo) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)g.mx * g.my * g.mz) return;
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  if (i < lo || i >= g.mx - lo || j < lo || j >= g.my - lo || k < lo || k >= g.mz - lo) return;
  double nv = nvert[n];
  if (nv < 0) nv = 0;
  const int ip = i < g.mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
  const int jp = j < g.my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
  const int kp = k < g.mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
  int around = 0;
  if ((int)(nv + 0.5) < 0.5) {
    for (int kk = km; kk < kp + 1; kk++)
      if ((int)(snap[nid(g, kk, j, i)] + 0.5) > 1 && kk != k) around++;
    for (int jj = jm; jj < jp + 1; jj++)
      if ((int)(snap[nid(g, k, jj, i)] + 0.5) > 1 && jj != j) around++;
    for (int ii = im; ii < ip + 1; ii++)
      if ((int)(snap[nid(g, k, j, ii)] + 0.5) > 1 && ii != i) around++;
  }
  if (around > 3) nv = pmax(2.0, nv);
  nvert[n] = nv;
}

// phase change against the previous classification, ibm_iso.c:495-507
__global__ __launch_bounds__(256) void phase_kernel(long long N, const double *nvert_o, double *nvert) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  if (nvert_o[n] > 2.5 && nvert[n] < 0.5) nvert[n] = 2;
}

// back to 1 (boundary) and 3 (solid), ibm_iso.c:761-769
__global__ __launch_bounds__(256) void relabel_kernel(long long N, double *nvert) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double nv = nvert[n];
  if ((int)(nv + 0.5) == 2) nv = 1 + 0 / 1000.0;
  if ((int)(nv + 0.5) == 4) nv = 3 + 0 / 1000.0;
  nvert[n] = nv;
}

// --------------------------------------------------------- list compaction

constexpr int kCompactPer = 16;                 // nodes per thread
constexpr int kCompactBlock = 256 * kCompactPer;  // nodes per workgroup

__device__ __forceinline__ bool in_list(const Grid &g, const double *nvert, long long n, int iso) {
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  if (iso == 2) {
    if (k >= g.mz - 1) return false;  // ibm_iso.c:525 loops k < lze
  } else {
    if (i < 1 || i >= g.mx - 1 || j < 1 || j >= g.my - 1 || k < 1 || k >= g.mz - 1) return false;
  }
  return (int)(nvert[n] + 0.5) == 2;
}

__global__ __launch_bounds__(256) void list_count_kernel(Grid g, const double *nvert, int iso, long long N, int *counts) {
  __shared__ int s[256];
  const long long base = (long long)blockIdx.x * kCompactBlock + (long long)threadIdx.x * kCompactPer;
  int c = 0;
  for (int q = 0; q < kCompactPer; q++) {
    const long long n = base + q;
    if (n < N && in_list(g, nvert, n, iso)) c++;
  }
  s[threadIdx.x] = c;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = s[0];
}

// writes list position n_ib-1-rank (traversal order of the prepended list)
__global__ __launch_bounds__(256) void list_write_kernel(Grid g, const double *nvert, int iso, long long N,
                                                         const int *offsets, int n_ib, int *node) {
  __shared__ int s[256];
  const long long base = (long long)blockIdx.x * kCompactBlock + (long long)threadIdx.x * kCompactPer;
  int c = 0;
  for (int q = 0; q < kCompactPer; q++) {
    const long long n = base + q;
    if (n < N && in_list(g, nvert, n, iso)) c++;
  }
  s[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < 256; t++) {
      const int v = s[t];
      s[t] = run;
      run += v;
    }
  }
  __syncthreads();
  int r = offsets[blockIdx.x] + s[threadIdx.x];
  for (int q = 0; q < kCompactPer; q++) {
    const long long n = base + q;
    if (n < N && in_list(g, nvert, n, iso)) node[n_ib - 1 - r++] = (int)n;
  }
}

// ------------------------------------------------------------------ stencils

struct StencilArgs {
  Grid g;
  const double *coor;
  Surf s;
  int n_ib, iso, twoD;
  const int *node;
  int *cell, *imode, *don;  // don: 9 planes (i1 j1 k1 i2 j2 k2 i3 j3 k3)
  double *info;             // 11 planes: pmin xyz, d_s, d_i, cs1..3, cr1..3
  long long *dn;            // 3 planes: donor node ids
  double *nrm;              // 3 planes: unit normal used by the interpolation
  int *flags;
  unsigned long long *first_bad;
};

__constant__ int kIcpTri[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 4, 5}, {4, 5, 8}, {4, 7, 8}, {4, 6, 7}, {3, 4, 6}, {3, 4, 0}};

__device__ __forceinline__ P3 node_xyz(const StencilArgs &a, long long n) {
  return P3{a.coor[n], a.coor[n + a.g.np], a.coor[n + 2 * a.g.np]};
}

__global__ __launch_bounds__(128) void stencil_kernel(StencilArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_ib) return;
  const Grid &g = a.g;
  const Surf &s = a.s;
  const int ne = s.ne;
  const long long n = a.node[m];
  const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
  const P3 p = node_xyz(a, n);
  // nearestcell1, ibm_iso.c:1581-1727
  double dmin = 1.e20;
  int cmin = -100;
  P3 pmin{0, 0, 0};
  for (int e = 0; e < ne; e++) {
    const P3 qc{s.qv[e], s.qv[ne + e], s.qv[2 * ne + e]};
    if (!(dist(p, qc) - s.rad[e] < dmin)) continue;
    const int n1 = s.tv[e], n2 = s.tv[ne + e], n3 = s.tv[2 * ne + e];
    const double nfx = s.nf[e], nfy = s.nf[ne + e], nfz = s.nf[2 * ne + e];
    const P3 p1 = vtx(s, n1), p2 = vtx(s, n2), p3 = vtx(s, n3);
    double tf = (p.x - p1.x) * nfx + (p.y - p1.y) * nfy + (p.z - p1.z) * nfz;
    if (fabs(tf) < 1.e-10) tf = 1.e-10;
    if (tf >= 0) {
      const P3 pj{p.x - tf * nfx, p.y - tf * nfy, p.z - tf * nfz};
      if (point_in_triangle(pj, p1, p2, p3, nfx, nfy, nfz) == 1) {
        if (tf < dmin) { dmin = tf; pmin = pj; cmin = e; }
      } else {
        P3 po;
        double d;
        dis_p_line(p, p1, p2, po, d);
        if (d < dmin) { dmin = d; pmin = po; cmin = e; }
        dis_p_line(p, p2, p3, po, d);
        if (d < dmin) { dmin = d; pmin = po; cmin = e; }
        dis_p_line(p, p3, p1, po, d);
        if (d < dmin) { dmin = d; pmin = po; cmin = e; }
      }
    }
  }
  const int nib = a.n_ib;
  double cs1 = 0, cs2 = 0, cs3 = 0;
  if (cmin < 0) {
    flag_node(a.flags, a.first_bad, F_NEAREST, n);
    cmin = 0;
  } else {
    bary_by_normal(pmin, vtx(s, s.tv[cmin]), vtx(s, s.tv[ne + cmin]), vtx(s, s.tv[2 * ne + cmin]), s.nf[cmin],
                   s.nf[ne + cmin], s.nf[2 * ne + cmin], cs1, cs2, cs3);
  }
  // InterceptionPoint, ibm_iso.c:1975-2469: faces i-1, i+1, j-1, j+1, k-1,
  // k+1; a face that needs a node outside the grid is skipped
  double nfx = p.x - pmin.x, nfy = p.y - pmin.y, nfz = p.z - pmin.z;
  {
    const double dr = sqrt(nfx * nfx + nfy * nfy + nfz * nfz);
    nfx /= dr; nfy /= dr; nfz /= dr;
  }
  int mode = -100;
  int di[3] = {0, 0, 0}, dj[3] = {0, 0, 0}, dk[3] = {0, 0, 0};
  double cr1 = 0, cr2 = 0, cr3 = 0, d_i = 0;
  for (int f = 0; f < 6 && mode < 0; f++) {
    int ip[9], jp[9], kp[9];
    bool ok = true;
    for (int q = 0; q < 9; q++) {
      const int u = q % 3 - 1, v = q / 3 - 1;
      switch (f) {
        case 0: ip[q] = i - 1; jp[q] = j + u; kp[q] = k + v; break;
        case 1: ip[q] = i + 1; jp[q] = j + u; kp[q] = k + v; break;
        case 2: ip[q] = i + u; jp[q] = j - 1; kp[q] = k + v; break;
        case 3: ip[q] = i + u; jp[q] = j + 1; kp[q] = k + v; break;
        case 4: ip[q] = i + u; jp[q] = j + v; kp[q] = k - 1; break;
        default: ip[q] = i + u; jp[q] = j + v; kp[q] = k + 1; break;
      }
      if (ip[q] < 0 || ip[q] >= g.mx || jp[q] < 0 || jp[q] >= g.my || kp[q] < 0 || kp[q] >= g.mz) ok = false;
    }
    if (!ok) continue;
    // ICP, ibm_iso.c:1875-1973
    for (int t = 0; t < 8; t++) {
      const int a0 = kIcpTri[t][0], a1 = kIcpTri[t][1], a2 = kIcpTri[t][2];
      const P3 p1 = node_xyz(a, nid(g, kp[a0], jp[a0], ip[a0]));
      const P3 p2 = node_xyz(a, nid(g, kp[a1], jp[a1], ip[a1]));
      const P3 p3 = node_xyz(a, nid(g, kp[a2], jp[a2], ip[a2]));
      const double dx1 = p.x - p1.x, dy1 = p.y - p1.y, dz1 = p.z - p1.z;
      const double dx2 = p2.x - p1.x, dy2 = p2.y - p1.y, dz2 = p2.z - p1.z;
      const double dx3 = p3.x - p1.x, dy3 = p3.y - p1.y, dz3 = p3.z - p1.z;
      double d = (nfx * (dy2 * dz3 - dz2 * dy3) - nfy * (dx2 * dz3 - dz2 * dx3) + nfz * (dx2 * dy3 - dy2 * dx3));
      if (!(fabs(d) > 1.e-10)) continue;
      d = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) / d;
      if (!(d > 0)) continue;
      const P3 pint{p.x + d * nfx, p.y + d * nfy, p.z + d * nfz};
      const double rx1 = p2.x - p1.x, ry1 = p2.y - p1.y, rz1 = p2.z - p1.z;
      const double rx2 = p3.x - p1.x, ry2 = p3.y - p1.y, rz2 = p3.z - p1.z;
      const double nfxt = ry1 * rz2 - rz1 * ry2;
      const double nfyt = -rx1 * rz2 + rz1 * rx2;
      const double nfzt = rx1 * ry2 - ry1 * rx2;
      if (point_in_triangle(pint, p1, p2, p3, nfxt, nfyt, nfzt) >= 0) {
        bary_by_normal(pint, p1, p2, p3, nfxt, nfyt, nfzt, cr1, cr2, cr3);
        d_i = sqrt((pint.x - p.x) * (pint.x - p.x) + (pint.y - p.y) * (pint.y - p.y) + (pint.z - p.z) * (pint.z - p.z));
        mode = t;
        di[0] = ip[a0]; dj[0] = jp[a0]; dk[0] = kp[a0];
        di[1] = ip[a1]; dj[1] = jp[a1]; dk[1] = kp[a1];
        di[2] = ip[a2]; dj[2] = jp[a2]; dk[2] = kp[a2];
        break;
      }
    }
  }
  if (mode < 0) {
    // fallback single-donor stencils, ibm_iso.c:536-634 (iso 2), :660-745 (iso 1)
    int si = -1, sj = -1, sk = -1;
    const double nz = s.nf[2 * ne + cmin];
    if (a.iso == 2) {
      if (k == 0 || k == g.mz - 1) { si = i; sj = j + 1; sk = k; }
      else if (i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1) { si = i; sj = j; sk = nz > 0 ? k + 1 : k - 1; }
    } else {
      if (i == 1 || i == g.mx - 2 || j == 1 || j == g.my - 2) { si = i; sj = j; sk = nz > 0 ? k + 1 : k - 1; }
      else if (k == 1 || k == g.mz - 2) { si = i; sj = j + 1; sk = k; }
    }
    if (si < 0 || si >= g.mx || sj < 0 || sj >= g.my || sk < 0 || sk >= g.mz) {
      flag_node(a.flags, a.first_bad, F_STENCIL, n);
      si = i; sj = j; sk = k;
    }
    const P3 q = node_xyz(a, nid(g, sk, sj, si));
    d_i = sqrt((p.x - q.x) * (p.x - q.x) + (p.y - q.y) * (p.y - q.y) + (p.z - q.z) * (p.z - q.z));
    cr1 = 1.; cr2 = 0.; cr3 = 0.;
    for (int q3 = 0; q3 < 3; q3++) { di[q3] = si; dj[q3] = sj; dk[q3] = sk; }
  }
  a.cell[m] = cmin;
  a.imode[m] = mode;
  for (int q = 0; q < 3; q++) {
    a.don[(3 * q) * nib + m] = di[q];
    a.don[(3 * q + 1) * nib + m] = dj[q];
    a.don[(3 * q + 2) * nib + m] = dk[q];
    a.dn[q * nib + m] = nid(g, dk[q], dj[q], di[q]);
  }
  const double vals[11] = {pmin.x, pmin.y, pmin.z, dmin, d_i, cs1, cs2, cs3, cr1, cr2, cr3};
  for (int q = 0; q < 11; q++) a.info[q * nib + m] = vals[q];
  // normal of the interpolation (ibm_iso.c:2662-2689): -twoD drops one
  // component before the normalisation
  double ux = a.twoD == 1 ? 0. : p.x - pmin.x;
  double uy = a.twoD == 2 ? 0. : p.y - pmin.y;
  double uz = a.twoD == 3 ? 0. : p.z - pmin.z;
  const double dr = sqrt(ux * ux + uy * uy + uz * uz);
  if (dr < 0.000000001) {
    ux = s.nf[cmin]; uy = s.nf[ne + cmin]; uz = s.nf[2 * ne + cmin];
  } else {
    ux = ux / dr; uy = uy / dr; uz = uz / dr;
  }
  a.nrm[m] = ux;
  a.nrm[nib + m] = uy;
  a.nrm[2 * nib + m] = uz;
}

// ------------------------------------------------------------ interpolation

struct InterpArgs {
  int n_ib, iso, invicid, add_dundt, sweep, twoD;
  long long np;
  const int *node, *cell;
  const long long *dn;
  const double *info, *nrm;
  Surf s;
  const double *u, *uold;  // surface velocities, 3 planes of nv
  double *q, *p;           // state being interpolated (Q, P)
  const double *qold;
  double *w;               // kW planes of n_ib
  double *carry;           // Pa carried into the next sweep
  double gamma, Ma, Twall, dt;
};
enum { W_PINT = 0, W_TB, W_X, W_LSC, W_UBX, W_UBY, W_UBZ, W_DUN, W_RHO, W_SB, W_SC, kW };

// per-node part of one sweep: reads Q and P of the previous sweep at the donors
__global__ __launch_bounds__(256) void interp_prep_kernel(InterpArgs a) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.n_ib) return;
  const int nib = a.n_ib;
  const long long np = a.np;
  const long long n = a.node[m];
  const long long d1 = a.dn[m], d2 = a.dn[nib + m], d3 = a.dn[2 * nib + m];
  const double sb = a.info[3 * nib + m], sc = sb + a.info[4 * nib + m];
  const double cs1 = a.info[5 * nib + m], cs2 = a.info[6 * nib + m], cs3 = a.info[7 * nib + m];
  const double sk1 = a.info[8 * nib + m], sk2 = a.info[9 * nib + m], sk3 = a.info[10 * nib + m];
  const double nfx = a.nrm[m], nfy = a.nrm[nib + m], nfz = a.nrm[2 * nib + m];
  const int ni = a.cell[m];
  const int nv = a.s.nv, ne = a.s.ne;
  const int v1 = a.s.tv[ni], v2 = a.s.tv[ne + ni], v3 = a.s.tv[2 * ne + ni];
  double Uax = a.uold[v1] * cs1 + a.uold[v2] * cs2 + a.uold[v3] * cs3;
  double Uay = a.uold[nv + v1] * cs1 + a.uold[nv + v2] * cs2 + a.uold[nv + v3] * cs3;
  double Uaz = a.uold[2 * nv + v1] * cs1 + a.uold[2 * nv + v2] * cs2 + a.uold[2 * nv + v3] * cs3;
  const double Ua_nold = Uax * nfx + Uay * nfy + Uaz * nfz;
  Uax = a.u[v1] * cs1 + a.u[v2] * cs2 + a.u[v3] * cs3;
  Uay = a.u[nv + v1] * cs1 + a.u[nv + v2] * cs2 + a.u[nv + v3] * cs3;
  Uaz = a.u[2 * nv + v1] * cs1 + a.u[2 * nv + v2] * cs2 + a.u[2 * nv + v3] * cs3;
  const double Ua_n = Uax * nfx + Uay * nfy + Uaz * nfz;
  const double *q = a.q;
  const double r1 = q[d1], r2 = q[d2], r3 = q[d3];
  const double Ucx = (q[np + d1] / r1 * sk1 + q[np + d2] / r2 * sk2 + q[np + d3] / r3 * sk3);
  const double Ucy = (q[2 * np + d1] / r1 * sk1 + q[2 * np + d2] / r2 * sk2 + q[2 * np + d3] / r3 * sk3);
  const double Ucz = (q[3 * np + d1] / r1 * sk1 + q[3 * np + d2] / r2 * sk2 + q[3 * np + d3] / r3 * sk3);
  double Ubx, Uby, Ubz;
  if (a.invicid) {  // freeslip, wallfunction.c:30-46
    const double Uan = Uax * nfx + Uay * nfy + Uaz * nfz;
    const double Ucn = Ucx * nfx + Ucy * nfy + Ucz * nfz;
    const double Ubn = Uan + (Ucn - Uan) * sb / sc;
    Ubx = Ucx - Ucn * nfx; Uby = Ucy - Ucn * nfy; Ubz = Ucz - Ucn * nfz;
    Ubx += Ubn * nfx; Uby += Ubn * nfy; Ubz += Ubn * nfz;
  } else {  // noslip, wallfunction.c:12-28
    const double u_c = Ucx - Uax, v_c = Ucy - Uay, w_c = Ucz - Uaz;
    Ubx = sb / sc * u_c; Uby = sb / sc * v_c; Ubz = sb / sc * w_c;
    Ubx += Uax; Uby += Uay; Ubz += Uaz;
  }
  // 2D constraint, ibm_iso.c:2744-2746
  if (a.twoD == 1) Ubx = 0.;
  if (a.twoD == 2) Uby = 0.;
  if (a.twoD == 3) Ubz = 0.;
  const double pint = (a.p[d1] * sk1 + a.p[d2] * sk2 + a.p[d3] * sk3);
  const double gm2 = a.gamma * a.Ma * a.Ma;
  double Tb = 0, X = 0, Lsc = 0, rho = 0;
  if (a.iso == 2) {
    const double rhoc = (r1 * sk1 + r2 * sk2 + r3 * sk3);
    double Tc = pint / rhoc * gm2;
    if (Tc < 0.) Tc = 0.;
    Tb = (Tc - a.Twall) * sb / sc + a.Twall;
    const double a_sound = sqrt(a.Twall / (a.Ma * a.Ma));
    const double Lambda1 = Ua_n - a_sound;
    Lsc = Lambda1 / sc;
    X = rhoc * a_sound * ((Ucx - Uax) * nfx + (Ucy - Uay) * nfy + (Ucz - Uaz) * nfz);
    rho = a.qold[n];
  } else if (a.iso == 1) {
    const double Tb1 = 1.0;
    rho = pint * gm2 / Tb1;
  } else {
    rho = (r1 * sk1 + r2 * sk2 + r3 * sk3);
  }
  double *w = a.w;
  w[W_PINT * nib + m] = pint;
  w[W_TB * nib + m] = Tb;
  w[W_X * nib + m] = X;
  w[W_LSC * nib + m] = Lsc;
  w[W_UBX * nib + m] = Ubx;
  w[W_UBY * nib + m] = Uby;
  w[W_UBZ * nib + m] = Ubz;
  w[W_DUN * nib + m] = Ua_n - Ua_nold;
  w[W_RHO * nib + m] = rho;
  w[W_SB * nib + m] = sb;
  w[W_SC * nib + m] = sc;
}

// isothermal-2 density and pressure of one node given the carried Pa
// (ibm_iso.c:2740-2762)
__device__ __forceinline__ void iso2_node(const InterpArgs &a, int m, double Pa, double &rhob, double &pb) {
  const int nib = a.n_ib;
  const double *w = a.w;
  const double a_sound = sqrt(a.Twall / (a.Ma * a.Ma));
  const double L1 = w[W_LSC * nib + m] * ((w[W_PINT * nib + m] - Pa) - w[W_X * nib + m]);
  rhob = w[W_RHO * nib + m] - a.dt * L1 / a_sound / a_sound;
  pb = rhob * w[W_TB * nib + m] / (a.gamma * a.Ma * a.Ma);
}

__device__ __forceinline__ void write_node(const InterpArgs &a, int m, double rho, double pb) {
  const int nib = a.n_ib;
  const double *w = a.w;
  const long long n = a.node[m], np = a.np;
  if (a.add_dundt) pb += rho * w[W_DUN * nib + m] / a.dt * (w[W_SC * nib + m] - w[W_SB * nib + m]);
  const double ru = rho * w[W_UBX * nib + m];
  const double rv = rho * w[W_UBY * nib + m];
  const double rw = rho * w[W_UBZ * nib + m];
  a.q[n] = rho;
  a.q[np + n] = ru;
  a.q[2 * np + n] = rv;
  a.q[3 * np + n] = rw;
  a.q[4 * np + n] = pb / (a.gamma - 1) + 0.5 * (ru * ru + rv * rv + rw * rw) / rho;
  a.p[n] = pb;
}

// list-ordered part of one sweep, one workgroup.  In sweeps after the
// first the isothermal-2 density of node m uses Pa = p of node m-1 (the
// reference's function-scope Pa), a serial chain.  The list is cut into
// one contiguous chunk per thread; the entry value of every chunk is first
// estimated by an affine-map prefix scan, then each thread replays its chunk
// with the exact serial arithmetic and the chunk entries are iterated until
// none changes.  The serial sequence is the only fixed point of that
// iteration and each pass makes at least one more chunk entry exact, so the
// result is bit-identical to evaluating the list in order.
__global__ __launch_bounds__(kChainThreads) void interp_chain_kernel(InterpArgs a) {
  __shared__ double sA[kChainThreads], sB[kChainThreads];
  const int t = threadIdx.x, nib = a.n_ib;
  const int chunk = (nib + kChainThreads - 1) / kChainThreads;
  const int m0 = t * chunk < nib ? t * chunk : nib;
  const int m1 = (t + 1) * chunk < nib ? (t + 1) * chunk : nib;
  if (a.iso != 2) {  // no chain: nodes are independent
    for (int m = t; m < nib; m += kChainThreads) write_node(a, m, a.w[W_RHO * nib + m], a.w[W_PINT * nib + m]);
    return;
  }
  if (a.sweep == 0) {  // Pa is the node's own interpolated pressure: no chain
    for (int m = t; m < nib; m += kChainThreads) {
      double rhob, pb;
      iso2_node(a, m, a.w[W_PINT * nib + m], rhob, pb);
      write_node(a, m, rhob, pb);
      if (m == nib - 1) *a.carry = pb;
    }
    return;
  }
  // estimate: affine map of each node, Pa_next = alpha + beta * Pa
  const double a_sound = sqrt(a.Twall / (a.Ma * a.Ma));
  const double gm2 = a.gamma * a.Ma * a.Ma;
  double A = 0, B = 1;
  for (int m = m0; m < m1; m++) {
    double rhob, al;
    iso2_node(a, m, 0.0, rhob, al);
    const double be = a.dt * a.w[W_LSC * nib + m] / a_sound / a_sound * a.w[W_TB * nib + m] / gm2;
    A = al + be * A;
    B = be * B;
  }
  sA[t] = A;
  sB[t] = B;
  __syncthreads();
  for (int off = 1; off < kChainThreads; off <<= 1) {  // inclusive scan of map composition
    double pa = 0, pbv = 1;
    const bool has = t >= off;
    if (has) { pa = sA[t - off]; pbv = sB[t - off]; }
    __syncthreads();
    if (has) {
      const double nA = sA[t] + sB[t] * pa, nB = sB[t] * pbv;
      sA[t] = nA;
      sB[t] = nB;
    }
    __syncthreads();
  }
  const double carry = *a.carry;
  const double guess = t == 0 ? carry : sA[t - 1] + sB[t - 1] * carry;
  __syncthreads();
  double *entry = sA;  // entry[t]: Pa entering chunk t
  entry[t] = guess;
  __syncthreads();
  for (int pass = 0; pass <= kChainThreads; pass++) {
    double x = entry[t];
    for (int m = m0; m < m1; m++) {
      double rhob, pb;
      iso2_node(a, m, x, rhob, pb);
      x = pb;
    }
    __syncthreads();
    int changed = 0;
    if (t + 1 < kChainThreads && __double_as_longlong(entry[t + 1]) != __double_as_longlong(x)) {
      entry[t + 1] = x;
      changed = 1;
    }
    if (!__syncthreads_or(changed)) break;
  }
  double x = entry[t];
  for (int m = m0; m < m1; m++) {
    double rhob, pb;
    iso2_node(a, m, x, rhob, pb);
    write_node(a, m, rhob, pb);
    x = pb;
  }
  if (m1 == nib && m0 < m1) *a.carry = x;
}

// ----------------------------------------------------------------- helpers

__global__ void copy_kernel(const double *src, double *dst, long long n) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

template <typename T> int dalloc_t(cles_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, (n ? n : 1) * sizeof(T)) != hipSuccess) return set_error(c, CLES_ENOMEM, "hipMalloc(%zu) failed", n * sizeof(T));
  *p = (T *)q;
  c->bytes += n * sizeof(T);
  return CLES_OK;
}

}  // namespace

void ibm_free(cles_ctx *c) {
  cles_ibm_dev *b = c->ibm;
  if (!b) return;
  void *bufs[] = {b->vx, b->vy, b->vz, b->tv, b->nf, b->qv, b->rad, b->u, b->uold, b->node, b->cell, b->imode,
                  b->don, b->dn, b->info, b->nrm, b->w, b->carry, b->frec, b->fout};
  for (void *q : bufs)
    if (q) (void)hipFree(q);
  delete b;
  c->ibm = nullptr;
}

int launch_ibm_interp(cles_ctx *c, int buf, int qold_buf, int add_dundt) {
  cles_ibm_dev *b = c->ibm;
  if (!b || b->n_ib == 0) return CLES_OK;
  InterpArgs a;
  a.n_ib = b->n_ib;
  a.iso = c->p.isothermal;
  a.invicid = c->p.invicid;
  a.add_dundt = add_dundt;
  a.twoD = c->p.twoD;
  a.np = c->g.np;
  a.node = b->node; a.cell = b->cell; a.dn = b->dn; a.info = b->info; a.nrm = b->nrm;
  a.s = Surf{b->nv, b->ne, b->vx, b->vy, b->vz, b->tv, b->nf, b->qv, b->rad};
  a.u = b->u; a.uold = b->uold;
  a.q = c->Q[buf]; a.p = c->P[buf]; a.qold = c->Q[qold_buf];
  a.w = b->w; a.carry = b->carry;
  a.gamma = c->p.gamma; a.Ma = c->p.Ma; a.Twall = c->p.Twall; a.dt = c->p.dt;
  const unsigned nb = (unsigned)((b->n_ib + 255) / 256);
  for (int sweep = 0; sweep < 5; sweep++) {  // ibm_iso.c:2620 itr_tmp = 5
    a.sweep = sweep;
    hipLaunchKernelGGL(interp_prep_kernel, dim3(nb), dim3(256), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(interp_chain_kernel, dim3(1), dim3(kChainThreads), 0, c->stream, a);
    CLES_HIP(c, hipGetLastError());
  }
  return CLES_OK;
}

int launch_snapshot(cles_ctx *c, int buf) {
  const long long np = c->g.np;
  const unsigned nb = (unsigned)((5 * np + 255) / 256);
  hipLaunchKernelGGL(copy_kernel, dim3(nb), dim3(256), 0, c->stream, c->Q[buf], c->qsnap, 5 * np);
  CLES_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, c->P[buf],
                     c->qsnap + 5 * np, np);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}

}  // namespace cles

using namespace cles;

extern "C" {

int cles_ibm_search(cles_ctx *c, int n_v, const double *x_bp, const double *y_bp, const double *z_bp, int n_elmt,
                    const int *nv1, const int *nv2, const int *nv3, const double *nvert_o, int *n_ib) {
  if (!c || n_v < 3 || n_elmt < 1 || !x_bp || !y_bp || !z_bp || !nv1 || !nv2 || !nv3)
    return set_error(c, CLES_EINVAL, "cles_ibm_search: empty surface or null pointer");
  if (int g_ = member_guard(c, true)) return g_;
  for (int e = 0; e < n_elmt; e++)
    if (nv1[e] < 0 || nv1[e] >= n_v || nv2[e] < 0 || nv2[e] >= n_v || nv3[e] < 0 || nv3[e] >= n_v)
      return set_error(c, CLES_EINVAL, "cles_ibm_search: triangle %d names a vertex outside 0..%d", e, n_v - 1);
  const Grid &g = c->g;
  const long long N = (long long)g.mx * g.my * g.mz;
  const int iso = c->p.isothermal;
  ibm_free(c);
  cles_ibm_dev *b = new cles_ibm_dev();
  memset(b, 0, sizeof(*b));
  c->ibm = b;
  b->nv = n_v;
  b->ne = n_elmt;

  // bounding box and bucket columns, ibm_iso.c:142-207 (host: O(triangles))
  double xmin = 1.e23, xmax = -1.e23, ymin = 1.e23, ymax = -1.e23, zmin = 1.e23, zmax = -1.e23;
  for (int n = 0; n < n_v; n++) {
    xmin = xmin < x_bp[n] ? xmin : x_bp[n]; xmax = xmax < x_bp[n] ? x_bp[n] : xmax;
    ymin = ymin < y_bp[n] ? ymin : y_bp[n]; ymax = ymax < y_bp[n] ? y_bp[n] : ymax;
    zmin = zmin < z_bp[n] ? zmin : z_bp[n]; zmax = zmax < z_bp[n] ? z_bp[n] : zmax;
  }
  xmin -= 0.05; xmax += 0.05;
  ymin -= 0.05; ymax += 0.05;
  zmin -= 0.05; zmax += 0.05;
  const double dcx = (xmax - xmin) / (kNc - 1.), dcy = (ymax - ymin) / (kNc - 1.), dcz = (zmax - zmin) / (kNc - 1.);
  std::vector<std::vector<int>> colv((size_t)kNc * kNc);
  std::vector<int> tkmin(n_elmt), tkmax(n_elmt);
  auto mn3 = [](double p, double q, double r) { double t = p < q ? p : q; return t < r ? t : r; };
  auto mx3 = [](double p, double q, double r) { double t = p < q ? q : p; return t < r ? r : t; };
  for (int e = 0; e < n_elmt; e++) {
    const int a0 = nv1[e], a1 = nv2[e], a2 = nv3[e];
    int iv0 = (int)floor((mn3(x_bp[a0], x_bp[a1], x_bp[a2]) - xmin) / dcx);
    int iv1 = (int)floor((mx3(x_bp[a0], x_bp[a1], x_bp[a2]) - xmin) / dcx) + 1;
    int jv0 = (int)floor((mn3(y_bp[a0], y_bp[a1], y_bp[a2]) - ymin) / dcy);
    int jv1 = (int)floor((mx3(y_bp[a0], y_bp[a1], y_bp[a2]) - ymin) / dcy) + 1;
    int kv0 = (int)floor((mn3(z_bp[a0], z_bp[a1], z_bp[a2]) - zmin) / dcz);
    int kv1 = (int)floor((mx3(z_bp[a0], z_bp[a1], z_bp[a2]) - zmin) / dcz) + 1;
    iv0 = iv0 < 0 ? 0 : iv0; iv1 = iv1 > kNc ? kNc : iv1;
    jv0 = jv0 < 0 ? 0 : jv0; jv1 = jv1 > kNc ? kNc : jv1;
    kv0 = kv0 < 0 ? 0 : kv0; kv1 = kv1 > kNc ? kNc : kv1;
    tkmin[e] = kv0;
    tkmax[e] = kv1;
    if (kv0 >= kv1) continue;  // in no bucket
    for (int j = jv0; j < jv1; j++)
      for (int i = iv0; i < iv1; i++) colv[(size_t)j * kNc + i].push_back(e);
  }
  std::vector<int> col_start(kNc * kNc + 1, 0), col_tri, col_kmin, col_kmax;
  for (int q = 0; q < kNc * kNc; q++) {
    col_start[q] = (int)col_tri.size();
    for (int e : colv[q]) {
      col_tri.push_back(e);
      col_kmin.push_back(tkmin[e]);
      col_kmax.push_back(tkmax[e]);
    }
  }
  col_start[kNc * kNc] = (int)col_tri.size();
  std::vector<double> sdraw(kMaxSeed);
  for (int sd = 1; sd <= kMaxSeed; sd++) sdraw[sd - 1] = glibc_first_rand((uint32_t)sd) / ((double)2147483647 + 1) - 0.5;

  // device copies of the surface
  int r = 0;
  std::vector<int> tv3(3 * (size_t)n_elmt);
  for (int e = 0; e < n_elmt; e++) {
    tv3[e] = nv1[e];
    tv3[n_elmt + e] = nv2[e];
    tv3[2 * (size_t)n_elmt + e] = nv3[e];
  }
  if ((r = dalloc_t(c, &b->vx, n_v)) || (r = dalloc_t(c, &b->vy, n_v)) || (r = dalloc_t(c, &b->vz, n_v)) ||
      (r = dalloc_t(c, &b->tv, 3 * (size_t)n_elmt)) || (r = dalloc_t(c, &b->nf, 3 * (size_t)n_elmt)) ||
      (r = dalloc_t(c, &b->qv, 3 * (size_t)n_elmt)) || (r = dalloc_t(c, &b->rad, n_elmt)) ||
      (r = dalloc_t(c, &b->u, 3 * (size_t)n_v)) || (r = dalloc_t(c, &b->uold, 3 * (size_t)n_v)) ||
      (r = dalloc_t(c, &b->carry, 1)))
    return r;
  CLES_HIP(c, hipMemcpyAsync(b->vx, x_bp, n_v * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipMemcpyAsync(b->vy, y_bp, n_v * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipMemcpyAsync(b->vz, z_bp, n_v * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipMemcpyAsync(b->tv, tv3.data(), tv3.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipMemsetAsync(b->u, 0, 3 * (size_t)n_v * sizeof(double), c->stream));
  CLES_HIP(c, hipMemsetAsync(b->uold, 0, 3 * (size_t)n_v * sizeof(double), c->stream));
  CLES_HIP(c, hipMemsetAsync(b->carry, 0, sizeof(double), c->stream));
  Surf s{n_v, n_elmt, b->vx, b->vy, b->vz, b->tv, b->nf, b->qv, b->rad};
  hipLaunchKernelGGL(surface_kernel, dim3((unsigned)((n_elmt + 255) / 256)), dim3(256), 0, c->stream, s, b->nf, b->qv, b->rad);
  CLES_HIP(c, hipGetLastError());

  // setup temporaries
  int *d_cs = nullptr, *d_ct = nullptr, *d_ck0 = nullptr, *d_ck1 = nullptr, *d_flags = nullptr, *d_counts = nullptr,
      *d_offs = nullptr;
  double *d_sd = nullptr, *d_snap = nullptr, *d_nvo = nullptr;
  unsigned long long *d_bad = nullptr;
  const long long nblk = (N + kCompactBlock - 1) / kCompactBlock;
  auto cleanup = [&]() {
    void *t[] = {d_cs, d_ct, d_ck0, d_ck1, d_flags, d_counts, d_offs, d_sd, d_snap, d_nvo, d_bad};
    for (void *q : t)
      if (q) (void)hipFree(q);
  };
#define TRY_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      cleanup();                                                                                  \
      return set_error(c, CLES_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
    }                                                                                             \
  } while (0)
  const size_t nct = col_tri.size() ? col_tri.size() : 1;
  TRY_HIP(hipMalloc((void **)&d_cs, col_start.size() * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_ct, nct * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_ck0, nct * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_ck1, nct * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_flags, F_NFLAGS * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_bad, sizeof(unsigned long long)));
  TRY_HIP(hipMalloc((void **)&d_sd, kMaxSeed * sizeof(double)));
  TRY_HIP(hipMalloc((void **)&d_snap, (size_t)N * sizeof(double)));
  TRY_HIP(hipMalloc((void **)&d_counts, (size_t)nblk * sizeof(int)));
  TRY_HIP(hipMalloc((void **)&d_offs, (size_t)nblk * sizeof(int)));
  TRY_HIP(hipMemcpyAsync(d_cs, col_start.data(), col_start.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (!col_tri.empty()) {
    TRY_HIP(hipMemcpyAsync(d_ct, col_tri.data(), col_tri.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TRY_HIP(hipMemcpyAsync(d_ck0, col_kmin.data(), col_kmin.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TRY_HIP(hipMemcpyAsync(d_ck1, col_kmax.data(), col_kmax.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  TRY_HIP(hipMemcpyAsync(d_sd, sdraw.data(), kMaxSeed * sizeof(double), hipMemcpyHostToDevice, c->stream));
  TRY_HIP(hipMemsetAsync(d_flags, 0, F_NFLAGS * sizeof(int), c->stream));
  TRY_HIP(hipMemsetAsync(d_bad, 0xff, sizeof(unsigned long long), c->stream));

  const unsigned gN = (unsigned)((N + 255) / 256);
  ClassArgs ca;
  ca.g = g; ca.coor = c->coor; ca.nvert = c->nvert; ca.s = s;
  ca.xmin = xmin; ca.xmax = xmax; ca.ymin = ymin; ca.ymax = ymax; ca.zmin = zmin; ca.zmax = zmax;
  ca.dcx = dcx; ca.dcy = dcy; ca.dcz = dcz;
  ca.col_start = d_cs; ca.col_tri = d_ct; ca.col_kmin = d_ck0; ca.col_kmax = d_ck1;
  ca.sdraw = d_sd; ca.flags = d_flags; ca.first_bad = d_bad;
  hipLaunchKernelGGL(classify_kernel, dim3(gN), dim3(256), 0, c->stream, ca);
  TRY_HIP(hipGetLastError());
  TRY_HIP(hipMemcpyAsync(d_snap, c->nvert, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(mark_kernel, dim3(gN), dim3(256), 0, c->stream, g, d_snap, c->nvert);
  TRY_HIP(hipGetLastError());
  if (iso == 1 || iso == 2) {
    TRY_HIP(hipMemcpyAsync(d_snap, c->nvert, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(island_kernel, dim3(gN), dim3(256), 0, c->stream, g, d_snap, c->nvert, iso == 2 ? 0 : 1);
    TRY_HIP(hipGetLastError());
  }
  if (nvert_o) {
    TRY_HIP(hipMalloc((void **)&d_nvo, (size_t)N * sizeof(double)));
    TRY_HIP(hipMemcpyAsync(d_nvo, nvert_o, (size_t)N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(phase_kernel, dim3(gN), dim3(256), 0, c->stream, N, d_nvo, c->nvert);
    TRY_HIP(hipGetLastError());
  }
  int hflags[F_NFLAGS];
  unsigned long long hbad = 0;
  TRY_HIP(hipMemcpyAsync(hflags, d_flags, sizeof(hflags), hipMemcpyDeviceToHost, c->stream));
  TRY_HIP(hipMemcpyAsync(&hbad, d_bad, sizeof(hbad), hipMemcpyDeviceToHost, c->stream));
  TRY_HIP(hipStreamSynchronize(c->stream));
  if (hflags[F_SEED] || hflags[F_HITS]) {
    cleanup();
    return set_error(c, CLES_EINVAL, "cles_ibm_search: ray casting %s at node %llu",
                     hflags[F_SEED] ? "found no regular ray in 1024 directions" : "crossed more than 48 triangles", hbad);
  }

  // interpolation list (ibm_iso.c:522-757), isothermal 1 or 2 only
  int nib = 0;
  if (iso == 1 || iso == 2) {
    hipLaunchKernelGGL(list_count_kernel, dim3((unsigned)nblk), dim3(256), 0, c->stream, g, c->nvert, iso, N, d_counts);
    TRY_HIP(hipGetLastError());
    std::vector<int> cnt((size_t)nblk), off((size_t)nblk);
    TRY_HIP(hipMemcpyAsync(cnt.data(), d_counts, (size_t)nblk * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TRY_HIP(hipStreamSynchronize(c->stream));
    for (long long q = 0; q < nblk; q++) {
      off[q] = nib;
      nib += cnt[q];
    }
    TRY_HIP(hipMemcpyAsync(d_offs, off.data(), (size_t)nblk * sizeof(int), hipMemcpyHostToDevice, c->stream));
  }
  b->n_ib = nib;
  if ((r = dalloc_t(c, &b->node, nib)) || (r = dalloc_t(c, &b->cell, nib)) || (r = dalloc_t(c, &b->imode, nib)) ||
      (r = dalloc_t(c, &b->don, 9 * (size_t)nib)) || (r = dalloc_t(c, &b->dn, 3 * (size_t)nib)) ||
      (r = dalloc_t(c, &b->info, 11 * (size_t)nib)) || (r = dalloc_t(c, &b->nrm, 3 * (size_t)nib)) ||
      (r = dalloc_t(c, &b->w, (size_t)kW * nib))) {
    cleanup();
    return r;
  }
  if (nib > 0) {
    hipLaunchKernelGGL(list_write_kernel, dim3((unsigned)nblk), dim3(256), 0, c->stream, g, c->nvert, iso, N, d_offs,
                       nib, b->node);
    TRY_HIP(hipGetLastError());
    StencilArgs sa;
    sa.g = g; sa.coor = c->coor; sa.s = s; sa.n_ib = nib; sa.iso = iso; sa.twoD = c->p.twoD; sa.node = b->node;
    sa.cell = b->cell; sa.imode = b->imode; sa.don = b->don; sa.info = b->info; sa.dn = b->dn; sa.nrm = b->nrm;
    sa.flags = d_flags; sa.first_bad = d_bad;
    hipLaunchKernelGGL(stencil_kernel, dim3((unsigned)((nib + 127) / 128)), dim3(128), 0, c->stream, sa);
    TRY_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(relabel_kernel, dim3(gN), dim3(256), 0, c->stream, N, c->nvert);
  TRY_HIP(hipGetLastError());
  TRY_HIP(hipMemcpyAsync(hflags, d_flags, sizeof(hflags), hipMemcpyDeviceToHost, c->stream));
  TRY_HIP(hipMemcpyAsync(&hbad, d_bad, sizeof(hbad), hipMemcpyDeviceToHost, c->stream));
  TRY_HIP(hipStreamSynchronize(c->stream));
  cleanup();
#undef TRY_HIP
  if (hflags[F_NEAREST] || hflags[F_STENCIL])
    return set_error(c, CLES_EINVAL, "cles_ibm_search: %s for boundary node %llu",
                     hflags[F_NEAREST] ? "no nearest surface triangle" : "no interpolation stencil", hbad);
  if (!c->qsnap && (r = dalloc_t(c, &c->qsnap, 6 * (size_t)g.np))) return r;
  c->has_ib = 1;
  c->immersed = 1;
  if (n_ib) *n_ib = nib;
  return CLES_OK;
}

int cles_ibm_set_velocity(cles_ctx *c, const double *u, const double *uold) {
  if (!c || !u || !uold) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  cles_ibm_dev *b = c->ibm;
  if (!b) return set_error(c, CLES_ESTATE, "no immersed surface (call cles_ibm_search first)");
  const int nv = b->nv;
  std::vector<double> su(3 * (size_t)nv), so(3 * (size_t)nv);
  for (int n = 0; n < nv; n++)
    for (int d = 0; d < 3; d++) {
      su[(size_t)d * nv + n] = u[3 * (size_t)n + d];
      so[(size_t)d * nv + n] = uold[3 * (size_t)n + d];
    }
  CLES_HIP(c, hipMemcpyAsync(b->u, su.data(), su.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipMemcpyAsync(b->uold, so.data(), so.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int cles_ibm_get_list(cles_ctx *c, cles_ibm_info *out, int n) {
  if (!c || (!out && n > 0)) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  cles_ibm_dev *b = c->ibm;
  const int nib = b ? b->n_ib : 0;
  if (n != nib) return set_error(c, CLES_EINVAL, "cles_ibm_get_list: list holds %d nodes, %d requested", nib, n);
  if (nib == 0) return CLES_OK;
  std::vector<int> node(nib), cell(nib), imode(nib), don(9 * (size_t)nib);
  std::vector<double> info(11 * (size_t)nib);
  CLES_HIP(c, hipMemcpyAsync(node.data(), b->node, nib * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipMemcpyAsync(cell.data(), b->cell, nib * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipMemcpyAsync(imode.data(), b->imode, nib * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipMemcpyAsync(don.data(), b->don, don.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipMemcpyAsync(info.data(), b->info, info.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  const Grid &g = c->g;
  for (int m = 0; m < nib; m++) {
    cles_ibm_info &o = out[m];
    const long long nn = node[m];
    o.ni = (int)(nn % g.mx);
    o.nj = (int)((nn / g.mx) % g.my);
    o.nk = (int)(nn / g.sz);
    o.cell = cell[m];
    o.imode = imode[m];
    int *d = &o.i1;
    for (int q = 0; q < 9; q++) d[q] = don[(size_t)q * nib + m];
    o.pmin[0] = info[m];
    o.pmin[1] = info[(size_t)nib + m];
    o.pmin[2] = info[2 * (size_t)nib + m];
    o.d_s = info[3 * (size_t)nib + m];
    o.d_i = info[4 * (size_t)nib + m];
    o.cs1 = info[5 * (size_t)nib + m];
    o.cs2 = info[6 * (size_t)nib + m];
    o.cs3 = info[7 * (size_t)nib + m];
    o.cr1 = info[8 * (size_t)nib + m];
    o.cr2 = info[9 * (size_t)nib + m];
    o.cr3 = info[10 * (size_t)nib + m];
  }
  return CLES_OK;
}

}  // extern "C"

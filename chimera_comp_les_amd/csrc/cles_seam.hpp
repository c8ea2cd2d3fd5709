// cles_seam.hpp -- the boundary-node walk and the periodic seam copy with the
// equation of state (rhs.c:1107-1227 seams, rhs.c:4853-4865 EOS), shared by
// the boundary kernels and the persistent batched step.
#pragma once

#include "cles_internal.hpp"

namespace cles {

// linear boundary index -> (k, j, i), every boundary node exactly once
__device__ __forceinline__ void bnd_decode(const Grid &g, long long b, int &k, int &j, int &i) {
  const long long mx = g.mx, my = g.my, mz = g.mz;
  const long long pk = mx * my;              // k = 0 and k = mz-1 planes
  const long long pj = mx * (mz - 2);        // j = 0 and j = my-1, 1<=k<=mz-2
  if (b < 2 * pk) {
    k = b < pk ? 0 : (int)(mz - 1);
    long long r = b % pk;
    j = (int)(r / mx);
    i = (int)(r % mx);
    return;
  }
  b -= 2 * pk;
  if (b < 2 * pj) {
    j = b < pj ? 0 : (int)(my - 1);
    long long r = b % pj;
    k = 1 + (int)(r / mx);
    i = (int)(r % mx);
    return;
  }
  b -= 2 * pj;
  // i = 0 and i = mx-1 with interior j, k: (my-2)(mz-2) nodes each.  The two
  // nodes of one (j, k) go to neighbouring threads:
  // each shares its 128 B line with the other's seam source (i = 1, mx-2)
  i = (b & 1) ? (int)(mx - 1) : 0;
  const long long r = b >> 1;
  k = 1 + (int)(r / (my - 2));
  j = 1 + (int)(r % (my - 2));
}

// number of boundary nodes of the walk
__host__ __device__ __forceinline__ long long bnd_count(const Grid &g) {
  const long long mx = g.mx, my = g.my, mz = g.mz;
  return 2 * mx * my + 2 * mx * (mz - 2) + 2 * (my - 2) * (mz - 2);
}

// Source node of the periodic seam copy that finally owns (k,j,i), or -1.
__device__ __forceinline__ long long seam_src(const Grid &g, int k, int j, int i) {
  if (g.pz && (k == 0 || k == g.mz - 1)) return nid(g, k == 0 ? g.mz - 2 : 1, j, i);
  if (g.py) {
    if (j == 0 && k >= 1) return nid(g, k, g.my - 2, i);
    if (j == g.my - 1 && k >= 1 && k <= g.mz - 2) return nid(g, k, 1, i);
  }
  if (g.px && (i == 0 || i == g.mx - 1) && k >= 1 && j >= 1) return nid(g, k, j, i == 0 ? g.mx - 2 : 1);
  return -1;
}

// Boundary node b (< nb): the periodic seam copy of the stage input q into
// the stage output qo (periodic grids) and the equation of state of the
// node's final state into po.  Unfused: the same bits in every translation
// unit that inlines it.
__device__ __forceinline__ void seam_copy_eos(const Grid &g, const double *q, double *qo, double *po, double gamma,
                                              int periodic, long long b, long long nb) {
#pragma clang fp contract(off)
  const long long np = g.np;
  if (b >= nb) return;
  int k, j, i;
  bnd_decode(g, b, k, j, i);
  const long long n = nid(g, k, j, i);
  double qn[5];
  const long long s = periodic ? seam_src(g, k, j, i) : -1;
#pragma unroll
  for (int c = 0; c < 5; c++) qn[c] = s >= 0 ? q[s + c * np] : qo[n + c * np];
  if (s >= 0) {
#pragma unroll
    for (int c = 0; c < 5; c++) qo[n + c * np] = qn[c];
  }
  po[n] = (gamma - 1) * (qn[4] - 0.5 * (qn[1] * qn[1] + qn[2] * qn[2] + qn[3] * qn[3]) / qn[0]);
}

}  // namespace cles

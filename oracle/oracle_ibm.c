/* Immersed boundary of the CPU restatement (test infrastructure only):
 * node classification by ray casting, interpolation stencils and the
 * interpolation sweeps of ibm_iso.c, with the geometry helpers of
 * compgeom.c and the wall models of wallfunction.c.  Single body (ibi = 0),
 * no wall function (-wallfunction 0), no thin-body pass (-thin 0).
 * The ray-casting direction uses the C library's srand/rand exactly as the
 * reference does (ibm_iso.c:54-85).  Serial on purpose: srand/rand carry
 * global state. */
#include "oracle_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ODIST(a, b) sqrt(((a).x - (b).x) * ((a).x - (b).x) + ((a).y - (b).y) * ((a).y - (b).y) + \
                         ((a).z - (b).z) * ((a).z - (b).z))

typedef struct { double x, y; } pt2;

/* ------------------------------------------------------------------ geometry */

/* intsect_triangle, compgeom.c:17-53 (Moller-Trumbore with EPSILON 1e-8) */
int orc_intsect_triangle(const double orig[3], const double dir[3], const double v0[3],
                         const double v1[3], const double v2[3], double *t, double *u, double *v) {
  double e1[3], e2[3], tv[3], pv[3], qv[3];
  for (int a = 0; a < 3; a++) { e1[a] = v1[a] - v0[a]; e2[a] = v2[a] - v0[a]; }
  pv[0] = dir[1] * e2[2] - dir[2] * e2[1];
  pv[1] = dir[2] * e2[0] - dir[0] * e2[2];
  pv[2] = dir[0] * e2[1] - dir[1] * e2[0];
  const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if (det > -0.00000001 && det < 0.00000001) return 0;
  const double inv = 1.0 / det;
  for (int a = 0; a < 3; a++) tv[a] = orig[a] - v0[a];
  *u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
  if (*u < 0.0 || *u > 1.0) return 0;
  qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
  qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
  qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
  *v = (dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2]) * inv;
  if (*v < 0.0 || *u + *v > 1.0) return 0;
  *t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
  return 1;
}

/* ISSameSide2D, compgeom.c:55-103.  The t3 branch of the reference is dead
 * (t1, t2 are recomputed unconditionally), so only the effective test is
 * restated. */
static int same_side(pt2 p, pt2 p1, pt2 p2, pt2 p3) {
  const double t1 = (p.x - p1.x) * (p2.y - p1.y) - (p.y - p1.y) * (p2.x - p1.x);
  const double t2 = (p3.x - p1.x) * (p2.y - p1.y) - (p3.y - p1.y) * (p2.x - p1.x);
  const double lt = sqrt((p1.x - p2.x) * (p1.x - p2.x) + (p1.y - p2.y) * (p1.y - p2.y));
  if (fabs(t1 / lt) < 1.e-10) return 1;
  if (t1 > 0) return t2 > 0 ? 1 : -1;
  return t2 < 0 ? 1 : -1;
}

/* ISInsideTriangle2D, compgeom.c:105-127 */
static int inside_2d(pt2 p, pt2 a, pt2 b, pt2 c) {
  int ls = same_side(p, a, b, c);
  if (ls < 0) return ls;
  ls = same_side(p, b, c, a);
  if (ls < 0) return ls;
  return same_side(p, c, a, b);
}

/* ISPointInTriangle, compgeom.c:129-157 */
int orc_point_in_triangle(orc_v3 p, orc_v3 p1, orc_v3 p2, orc_v3 p3, double nfx, double nfy, double nfz) {
  pt2 a, b, c, d;
  if (fabs(nfz) >= fabs(nfx) && fabs(nfz) >= fabs(nfy)) {
    a = (pt2){p.x, p.y}; b = (pt2){p1.x, p1.y}; c = (pt2){p2.x, p2.y}; d = (pt2){p3.x, p3.y};
  } else if (fabs(nfx) >= fabs(nfy) && fabs(nfx) >= fabs(nfz)) {
    a = (pt2){p.z, p.y}; b = (pt2){p1.z, p1.y}; c = (pt2){p2.z, p2.y}; d = (pt2){p3.z, p3.y};
  } else {
    a = (pt2){p.x, p.z}; b = (pt2){p1.x, p1.z}; c = (pt2){p2.x, p2.z}; d = (pt2){p3.x, p3.z};
  }
  return inside_2d(a, b, c, d);
}

/* Dis_P_Line, compgeom.c:160-186 */
static void dis_p_line(orc_v3 p, orc_v3 p1, orc_v3 p2, orc_v3 *po, double *d) {
  const double dx21 = p2.x - p1.x, dy21 = p2.y - p1.y, dz21 = p2.z - p1.z;
  const double dx31 = p.x - p1.x, dy31 = p.y - p1.y, dz31 = p.z - p1.z;
  const double t = (dx31 * dx21 + dy31 * dy21 + dz31 * dz21) / (dx21 * dx21 + dy21 * dy21 + dz21 * dz21);
  if (t < 0) {
    *po = p1;
    *d = sqrt(dx31 * dx31 + dy31 * dy31 + dz31 * dz31);
  } else if (t > 1) {
    *po = p2;
    *d = sqrt((p.x - po->x) * (p.x - po->x) + (p.y - po->y) * (p.y - po->y) + (p.z - po->z) * (p.z - po->z));
  } else {
    po->x = p1.x + t * dx21; po->y = p1.y + t * dy21; po->z = p1.z + t * dz21;
    *d = sqrt((p.x - po->x) * (p.x - po->x) + (p.y - po->y) * (p.y - po->y) + (p.z - po->z) * (p.z - po->z));
  }
}

/* triangle_intp / triangle_intp2, compgeom.c:188-229 */
static void tri_bary(pt2 p, pt2 p1, pt2 p2, pt2 p3, double *c1, double *c2, double *c3) {
  const double x13 = p1.x - p3.x, y13 = p1.y - p3.y;
  const double x23 = p2.x - p3.x, y23 = p2.y - p3.y;
  const double xp3 = p.x - p3.x, yp3 = p.y - p3.y;
  const double a = x13 * y23 - x23 * y13;
  *c1 = (y23 * xp3 - x23 * yp3) / a;
  *c2 = (-y13 * xp3 + x13 * yp3) / a;
  *c3 = 1. - *c1 - *c2;
}

/* barycentric coordinates in the plane picked by the largest normal
 * component (ibm_iso.c:1696-1719 and :1935-1957) */
static void bary_by_normal(orc_v3 q, orc_v3 p1, orc_v3 p2, orc_v3 p3, double nx, double ny, double nz,
                           double *c1, double *c2, double *c3) {
  if (fabs(nx) >= fabs(ny) && fabs(nx) >= fabs(nz))
    tri_bary((pt2){q.y, q.z}, (pt2){p1.y, p1.z}, (pt2){p2.y, p2.z}, (pt2){p3.y, p3.z}, c1, c2, c3);
  else if (fabs(ny) >= fabs(nx) && fabs(ny) >= fabs(nz))
    tri_bary((pt2){q.x, q.z}, (pt2){p1.x, p1.z}, (pt2){p2.x, p2.z}, (pt2){p3.x, p3.z}, c1, c2, c3);
  else if (fabs(nz) >= fabs(ny) && fabs(nz) >= fabs(nx))
    tri_bary((pt2){q.y, q.x}, (pt2){p1.y, p1.x}, (pt2){p2.y, p2.x}, (pt2){p3.y, p3.x}, c1, c2, c3);
}

/* --------------------------------------------------------------- surface */

static orc_v3 vtx(const orc_ctx *c, int n) { return (orc_v3){c->x_bp[n], c->y_bp[n], c->z_bp[n]}; }

/* unit normals from the vertex order, ibm_io.c:1716-1731 */
static void surface_normals(orc_ctx *c) {
  for (int e = 0; e < c->n_elmt; e++) {
    const int n1 = c->nv1[e], n2 = c->nv2[e], n3 = c->nv3[e];
    const double dx12 = c->x_bp[n2] - c->x_bp[n1], dy12 = c->y_bp[n2] - c->y_bp[n1], dz12 = c->z_bp[n2] - c->z_bp[n1];
    const double dx13 = c->x_bp[n3] - c->x_bp[n1], dy13 = c->y_bp[n3] - c->y_bp[n1], dz13 = c->z_bp[n3] - c->z_bp[n1];
    double nx = dy12 * dz13 - dz12 * dy13;
    double ny = -dx12 * dz13 + dz12 * dx13;
    double nz = dx12 * dy13 - dy12 * dx13;
    const double dr = sqrt(nx * nx + ny * ny + nz * nz);
    c->nf_x[e] = nx / dr; c->nf_y[e] = ny / dr; c->nf_z[e] = nz / dr;
  }
}

/* BoundingSphere, ibm_iso.c:1491-1579 */
static void bounding_sphere(orc_ctx *c) {
  const orc_v3 p0 = {0, 0, 0};
  for (int e = 0; e < c->n_elmt; e++) {
    const orc_v3 p1 = vtx(c, c->nv1[e]), p2 = vtx(c, c->nv2[e]), p3 = vtx(c, c->nv3[e]);
    const double l12 = ODIST(p1, p2), l23 = ODIST(p2, p3), l31 = ODIST(p3, p1);
    orc_v3 pa, pb, pc;
    if (l12 > l23) {
      if (l12 > l31) { pa = p1; pb = p2; pc = p3; }
      else { pa = p3; pb = p1; pc = p2; }
    } else {
      if (l31 < l23) { pa = p2; pb = p3; pc = p1; }
      else { pa = p3; pb = p1; pc = p2; }
    }
    const orc_v3 pf = {0.5 * (pa.x + pb.x), 0.5 * (pa.y + pb.y), 0.5 * (pa.z + pb.z)};
    const orc_v3 pu = {pa.x - pf.x, pa.y - pf.y, pa.z - pf.z};
    const orc_v3 pv = {pc.x - pf.x, pc.y - pf.y, pc.z - pf.z};
    const orc_v3 pt = {pu.y * pv.z - pu.z * pv.y, pu.z * pv.x - pu.x * pv.z, pu.x * pv.y - pu.y * pv.x};
    const orc_v3 pd = {pt.y * pu.z - pt.z * pu.y, pt.z * pu.x - pt.x * pu.z, pt.x * pu.y - pt.y * pu.x};
    double gama = -(ODIST(pu, p0) * ODIST(pu, p0) - ODIST(pv, p0) * ODIST(pv, p0));
    const orc_v3 pw = {pv.x - pu.x, pv.y - pu.y, pv.z - pu.z};
    double lamda = 2 * (pd.x * pw.x + pd.y * pw.y + pd.z * pw.z);
    gama /= lamda;
    lamda = gama < 0 ? 0 : gama;
    c->qvec[e] = (orc_v3){pf.x + lamda * pd.x, pf.y + lamda * pd.y, pf.z + lamda * pd.z};
    c->radvec[e] = ODIST(c->qvec[e], pa);
  }
}

/* ------------------------------------------------------------ ray casting */

typedef struct { int *item; int n, cap; } bucket;

static void bucket_add(bucket *b, int e) {
  for (int q = 0; q < b->n; q++)
    if (b->item[q] == e) return; /* insertnode keeps one entry per element (variables.c:7-27) */
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 8;
    b->item = (int *)realloc(b->item, (size_t)b->cap * sizeof(int));
  }
  b->item[b->n++] = e;
}

/* randomdirection, ibm_iso.c:54-85 */
static void random_direction(orc_v3 p, int ip, int jp, double xbp_min, double ybp_min, double zbp_max,
                             double dcx, double dcy, double dir[3], int seed) {
  const double xpc = dcx * (ip + 0.5) + xbp_min;
  const double ypc = dcy * (jp + 0.5) + ybp_min;
  srand((unsigned)seed);
  double s = rand() / ((double)RAND_MAX + 1) - 0.5;
  const double ex = xpc + s * dcx, ey = ypc + s * dcy, ez = zbp_max + 0.2;
  dir[0] = ex - p.x; dir[1] = ey - p.y; dir[2] = ez - p.z;
  s = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  dir[0] /= s; dir[1] /= s; dir[2] /= s;
}

double orc_glibc_first_rand(int seed) {
  srand((unsigned)seed);
  return (double)rand();
}

/* point_cell_advanced, ibm_iso.c:1284-1396.  Returns 4 (inside) or 0,
 * -1 when more than ORC_MAX_HITS crossings are found on one ray, -4 when
 * ORC_MAX_SEEDS directions were all singular (the reference retries without
 * bound; on a surface whose edge lies along the line the endpoints move on,
 * it never stops). */
#define ORC_MAX_HITS 1000
#define ORC_MAX_SEEDS 1024
static int point_cell(orc_ctx *c, orc_v3 p, int ip, int jp, int kp, int ncx, int ncy, int ncz, double dcx,
                      double dcy, double xbp_min, double ybp_min, double zbp_max, const bucket *cell,
                      unsigned char *searched, int *touched) {
  double dt[ORC_MAX_HITS];
  orc_v3 dnn[ORC_MAX_HITS];
  int searchtimes = 0;
  while (searchtimes < ORC_MAX_SEEDS) {
    searchtimes++;
    int nintp = 0, ntouched = 0, singular = 0;
    double dir[3];
    random_direction(p, ip, jp, xbp_min, ybp_min, zbp_max, dcx, dcy, dir, searchtimes);
    for (int k = kp; k < ncz && !singular; k++) {
      const bucket *b = &cell[(size_t)k * ncx * ncy + (size_t)jp * ncx + ip];
      for (int q = b->n - 1; q >= 0 && !singular; q--) { /* list head = last inserted */
        const int e = b->item[q];
        if (searched[e]) continue;
        searched[e] = 1;
        touched[ntouched++] = e;
        const orc_v3 nn = {c->nf_x[e], c->nf_y[e], c->nf_z[e]};
        const double orig[3] = {p.x, p.y, p.z};
        const int n1 = c->nv1[e], n2 = c->nv2[e], n3 = c->nv3[e];
        const double v0[3] = {c->x_bp[n1], c->y_bp[n1], c->z_bp[n1]};
        const double v1[3] = {c->x_bp[n2], c->y_bp[n2], c->z_bp[n2]};
        const double v2[3] = {c->x_bp[n3], c->y_bp[n3], c->z_bp[n3]};
        double t = 0, u, v;
        const int hit = orc_intsect_triangle(orig, dir, v0, v1, v2, &t, &u, &v);
        if (hit > 0 && t > 0) {
          if (nintp == ORC_MAX_HITS) {
            for (int z = 0; z < ntouched; z++) searched[touched[z]] = 0;
            return -1;
          }
          dt[nintp] = t;
          dnn[nintp] = nn;
          nintp++;
          for (int z = 0; z < nintp - 1; z++) {
            const double ndotn = dnn[z].x * nn.x + dnn[z].y * nn.y + dnn[z].z * nn.z;
            if (fabs(t - dt[z]) < 1.e-8 && ndotn > -0.97) singular = 1;
          }
        }
      }
    }
    for (int z = 0; z < ntouched; z++) searched[touched[z]] = 0;
    if (!singular) return (nintp % 2) ? 4 : 0;
  }
  return -4;
}

/* --------------------------------------------------------------- stencils */

/* nearestcell1, ibm_iso.c:1581-1727 */
static void nearest_cell(const orc_ctx *c, orc_v3 p, orc_ibm_info *info) {
  double dmin = 1.e20;
  int cell_min = -100;
  orc_v3 pmin = {0, 0, 0};
  for (int e = 0; e < c->n_elmt; e++) {
    const double d_center = ODIST(p, c->qvec[e]);
    if (!(d_center - c->radvec[e] < dmin)) continue;
    const int n1 = c->nv1[e], n2 = c->nv2[e], n3 = c->nv3[e];
    const double nfx = c->nf_x[e], nfy = c->nf_y[e], nfz = c->nf_z[e];
    const orc_v3 p1 = vtx(c, n1), p2 = vtx(c, n2), p3 = vtx(c, n3);
    double tf = (p.x - c->x_bp[n1]) * nfx + (p.y - c->y_bp[n1]) * nfy + (p.z - c->z_bp[n1]) * nfz;
    if (fabs(tf) < 1.e-10) tf = 1.e-10;
    if (tf >= 0) {
      const orc_v3 pj = {p.x - tf * nfx, p.y - tf * nfy, p.z - tf * nfz};
      if (orc_point_in_triangle(pj, p1, p2, p3, nfx, nfy, nfz) == 1) {
        if (tf < dmin) { dmin = tf; pmin = pj; cell_min = e; }
      } else {
        orc_v3 po;
        double d;
        dis_p_line(p, p1, p2, &po, &d);
        if (d < dmin) { dmin = d; pmin = po; cell_min = e; }
        dis_p_line(p, p2, p3, &po, &d);
        if (d < dmin) { dmin = d; pmin = po; cell_min = e; }
        dis_p_line(p, p3, p1, &po, &d);
        if (d < dmin) { dmin = d; pmin = po; cell_min = e; }
      }
    }
  }
  info->cell = cell_min;
  info->pmin[0] = pmin.x; info->pmin[1] = pmin.y; info->pmin[2] = pmin.z;
  info->d_s = dmin;
  if (cell_min < 0) return;
  const orc_v3 p1 = vtx(c, c->nv1[cell_min]), p2 = vtx(c, c->nv2[cell_min]), p3 = vtx(c, c->nv3[cell_min]);
  bary_by_normal(pmin, p1, p2, p3, c->nf_x[cell_min], c->nf_y[cell_min], c->nf_z[cell_min], &info->cs1,
                 &info->cs2, &info->cs3);
}

/* ICP, ibm_iso.c:1875-1973: the ray p + d*n through the 8 triangles of a 3x3
 * face of grid nodes; returns the triangle index or -100. */
static const int icp_tri[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 4, 5}, {4, 5, 8},
                                  {4, 7, 8}, {4, 6, 7}, {3, 4, 6}, {3, 4, 0}};
static int icp(orc_v3 p, const orc_v3 pc[9], double nfx, double nfy, double nfz, orc_ibm_info *info) {
  for (int t = 0; t < 8; t++) {
    const orc_v3 p1 = pc[icp_tri[t][0]], p2 = pc[icp_tri[t][1]], p3 = pc[icp_tri[t][2]];
    const double dx1 = p.x - p1.x, dy1 = p.y - p1.y, dz1 = p.z - p1.z;
    const double dx2 = p2.x - p1.x, dy2 = p2.y - p1.y, dz2 = p2.z - p1.z;
    const double dx3 = p3.x - p1.x, dy3 = p3.y - p1.y, dz3 = p3.z - p1.z;
    double d = (nfx * (dy2 * dz3 - dz2 * dy3) - nfy * (dx2 * dz3 - dz2 * dx3) + nfz * (dx2 * dy3 - dy2 * dx3));
    if (fabs(d) > 1.e-10) {
      d = -(dx1 * (dy2 * dz3 - dz2 * dy3) - dy1 * (dx2 * dz3 - dz2 * dx3) + dz1 * (dx2 * dy3 - dy2 * dx3)) / d;
      if (d > 0) {
        const orc_v3 pint = {p.x + d * nfx, p.y + d * nfy, p.z + d * nfz};
        const double rx1 = p2.x - p1.x, ry1 = p2.y - p1.y, rz1 = p2.z - p1.z;
        const double rx2 = p3.x - p1.x, ry2 = p3.y - p1.y, rz2 = p3.z - p1.z;
        const double nfxt = ry1 * rz2 - rz1 * ry2;
        const double nfyt = -rx1 * rz2 + rz1 * rx2;
        const double nfzt = rx1 * ry2 - ry1 * rx2;
        if (orc_point_in_triangle(pint, p1, p2, p3, nfxt, nfyt, nfzt) >= 0) {
          bary_by_normal(pint, p1, p2, p3, nfxt, nfyt, nfzt, &info->cr1, &info->cr2, &info->cr3);
          info->d_i = sqrt((pint.x - p.x) * (pint.x - p.x) + (pint.y - p.y) * (pint.y - p.y) +
                           (pint.z - p.z) * (pint.z - p.z));
          return t;
        }
      }
    }
  }
  return -100;
}

/* InterceptionPoint, ibm_iso.c:1975-2469: faces i-1, i+1, j-1, j+1, k-1, k+1
 * in that order.  A face that would need a node outside the grid is skipped
 * (the reference reads outside its arrays there). */
static void interception_point(const orc_ctx *c, orc_v3 p, int i, int j, int k, orc_ibm_info *info) {
  double nfx = p.x - info->pmin[0], nfy = p.y - info->pmin[1], nfz = p.z - info->pmin[2];
  const double dr = sqrt(nfx * nfx + nfy * nfy + nfz * nfz);
  nfx /= dr; nfy /= dr; nfz /= dr;
  info->imode = -100;
  for (int f = 0; f < 6; f++) {
    int ip[9], jp[9], kp[9], ok = 1;
    for (int n = 0; n < 9; n++) {
      const int a = n % 3 - 1, b = n / 3 - 1;
      switch (f) {
        case 0: ip[n] = i - 1; jp[n] = j + a; kp[n] = k + b; break;
        case 1: ip[n] = i + 1; jp[n] = j + a; kp[n] = k + b; break;
        case 2: ip[n] = i + a; jp[n] = j - 1; kp[n] = k + b; break;
        case 3: ip[n] = i + a; jp[n] = j + 1; kp[n] = k + b; break;
        case 4: ip[n] = i + a; jp[n] = j + b; kp[n] = k - 1; break;
        default: ip[n] = i + a; jp[n] = j + b; kp[n] = k + 1; break;
      }
      if (ip[n] < 0 || ip[n] >= c->mx || jp[n] < 0 || jp[n] >= c->my || kp[n] < 0 || kp[n] >= c->mz) ok = 0;
    }
    if (!ok) continue;
    orc_v3 pc[9];
    for (int n = 0; n < 9; n++) pc[n] = c->coor[GI(c, kp[n], jp[n], ip[n])];
    const int t = icp(p, pc, nfx, nfy, nfz, info);
    if (t >= 0) {
      const int a = icp_tri[t][0], b = icp_tri[t][1], d = icp_tri[t][2];
      info->i1 = ip[a]; info->j1 = jp[a]; info->k1 = kp[a];
      info->i2 = ip[b]; info->j2 = jp[b]; info->k2 = kp[b];
      info->i3 = ip[d]; info->j3 = jp[d]; info->k3 = kp[d];
      info->imode = t;
      return;
    }
  }
}

static void single_donor(const orc_ctx *c, orc_v3 p, int i, int j, int k, orc_ibm_info *info) {
  const orc_v3 q = c->coor[GI(c, k, j, i)];
  info->d_i = sqrt((p.x - q.x) * (p.x - q.x) + (p.y - q.y) * (p.y - q.y) + (p.z - q.z) * (p.z - q.z));
  info->cr1 = 1.; info->cr2 = 0.; info->cr3 = 0.;
  info->i1 = info->i2 = info->i3 = i;
  info->j1 = info->j2 = info->j3 = j;
  info->k1 = info->k2 = info->k3 = k;
}

/* ------------------------------------------------------- ibm_search_advanced */

int orc_ibm_search(orc_ctx *c, int n_v, const double *x_bp, const double *y_bp, const double *z_bp, int n_elmt,
                   const int *nv1, const int *nv2, const int *nv3, const double *nvert_o) {
  const int mx = c->mx, my = c->my, mz = c->mz;
  const int iso = c->p.isothermal;
  const int ncx = 40, ncy = 40, ncz = 40;
  /* surface (ibm_read_ucd, ibm_io.c:1530-1760) */
  free(c->x_bp); free(c->y_bp); free(c->z_bp); free(c->nf_x); free(c->nf_y); free(c->nf_z);
  free(c->nv1); free(c->nv2); free(c->nv3); free(c->qvec); free(c->radvec); free(c->u_bp); free(c->uold_bp);
  free(c->ibl);
  c->n_v = n_v; c->n_elmt = n_elmt;
  c->x_bp = (double *)malloc(n_v * sizeof(double));
  c->y_bp = (double *)malloc(n_v * sizeof(double));
  c->z_bp = (double *)malloc(n_v * sizeof(double));
  memcpy(c->x_bp, x_bp, n_v * sizeof(double));
  memcpy(c->y_bp, y_bp, n_v * sizeof(double));
  memcpy(c->z_bp, z_bp, n_v * sizeof(double));
  c->nv1 = (int *)malloc(n_elmt * sizeof(int));
  c->nv2 = (int *)malloc(n_elmt * sizeof(int));
  c->nv3 = (int *)malloc(n_elmt * sizeof(int));
  memcpy(c->nv1, nv1, n_elmt * sizeof(int));
  memcpy(c->nv2, nv2, n_elmt * sizeof(int));
  memcpy(c->nv3, nv3, n_elmt * sizeof(int));
  c->nf_x = (double *)malloc(n_elmt * sizeof(double));
  c->nf_y = (double *)malloc(n_elmt * sizeof(double));
  c->nf_z = (double *)malloc(n_elmt * sizeof(double));
  c->qvec = (orc_v3 *)malloc(n_elmt * sizeof(orc_v3));
  c->radvec = (double *)malloc(n_elmt * sizeof(double));
  c->u_bp = (orc_v3 *)calloc(n_v, sizeof(orc_v3));
  c->uold_bp = (orc_v3 *)calloc(n_v, sizeof(orc_v3));
  surface_normals(c);

  /* bounding box and bucket grid, ibm_iso.c:142-207 */
  double xbp_min = 1.e23, xbp_max = -1.e23, ybp_min = 1.e23, ybp_max = -1.e23, zbp_min = 1.e23, zbp_max = -1.e23;
  for (int n = 0; n < n_v; n++) {
    xbp_min = pmin(xbp_min, x_bp[n]); xbp_max = pmax(xbp_max, x_bp[n]);
    ybp_min = pmin(ybp_min, y_bp[n]); ybp_max = pmax(ybp_max, y_bp[n]);
    zbp_min = pmin(zbp_min, z_bp[n]); zbp_max = pmax(zbp_max, z_bp[n]);
  }
  xbp_min -= 0.05; xbp_max += 0.05;
  ybp_min -= 0.05; ybp_max += 0.05;
  zbp_min -= 0.05; zbp_max += 0.05;
  const double dcx = (xbp_max - xbp_min) / (ncx - 1.);
  const double dcy = (ybp_max - ybp_min) / (ncy - 1.);
  const double dcz = (zbp_max - zbp_min) / (ncz - 1.);
  bucket *cell = (bucket *)calloc((size_t)ncx * ncy * ncz, sizeof(bucket));
  for (int e = 0; e < n_elmt; e++) {
    const int a = nv1[e], b = nv2[e], d = nv3[e];
    const double xv_min = pmin(pmin(x_bp[a], x_bp[b]), x_bp[d]), xv_max = pmax(pmax(x_bp[a], x_bp[b]), x_bp[d]);
    const double yv_min = pmin(pmin(y_bp[a], y_bp[b]), y_bp[d]), yv_max = pmax(pmax(y_bp[a], y_bp[b]), y_bp[d]);
    const double zv_min = pmin(pmin(z_bp[a], z_bp[b]), z_bp[d]), zv_max = pmax(pmax(z_bp[a], z_bp[b]), z_bp[d]);
    int iv_min = (int)floor((xv_min - xbp_min) / dcx), iv_max = (int)floor((xv_max - xbp_min) / dcx) + 1;
    int jv_min = (int)floor((yv_min - ybp_min) / dcy), jv_max = (int)floor((yv_max - ybp_min) / dcy) + 1;
    int kv_min = (int)floor((zv_min - zbp_min) / dcz), kv_max = (int)floor((zv_max - zbp_min) / dcz) + 1;
    iv_min = iv_min < 0 ? 0 : iv_min; iv_max = iv_max > ncx ? ncx : iv_max;
    jv_min = jv_min < 0 ? 0 : jv_min; jv_max = jv_max > ncx ? ncy : jv_max;
    kv_min = kv_min < 0 ? 0 : kv_min; kv_max = kv_max > ncz ? ncz : kv_max;
    for (int k = kv_min; k < kv_max; k++)
      for (int j = jv_min; j < jv_max; j++)
        for (int i = iv_min; i < iv_max; i++) bucket_add(&cell[(size_t)k * ncx * ncy + (size_t)j * ncx + i], e);
  }

  /* ray casting, ibm_iso.c:234-250 */
  unsigned char *searched = (unsigned char *)calloc(n_elmt, 1);
  int *touched = (int *)malloc(n_elmt * sizeof(int));
  double *nvert = c->nvert;
  int status = 0;
  for (int k = 0; k < mz && !status; k++)
    for (int j = 0; j < my && !status; j++)
      for (int i = 0; i < mx; i++) {
        const size_t g = GI(c, k, j, i);
        if (nvert[g] < 5.) nvert[g] = 0;
        const orc_v3 p = c->coor[g];
        if (p.x > xbp_min && p.x < xbp_max && p.y > ybp_min && p.y < ybp_max && p.z > zbp_min && p.z < zbp_max) {
          const int ic = (int)floor((p.x - xbp_min) / dcx);
          const int jc = (int)floor((p.y - ybp_min) / dcy);
          const int kc = (int)floor((p.z - zbp_min) / dcz);
          const int r = point_cell(c, p, ic, jc, kc, ncx, ncy, ncz, dcx, dcy, xbp_min, ybp_min, zbp_max, cell,
                                   searched, touched);
          if (r < 0) { status = r == -4 ? -4 : -1; break; }
          nvert[g] = pmax(nvert[g], (double)r);
          if (nvert[g] < 0) nvert[g] = 0;
        }
      }
  free(searched); free(touched);
  for (size_t q = 0; q < (size_t)ncx * ncy * ncz; q++) free(cell[q].item);
  free(cell);
  if (status) return status;
  orc_g2l(c, c->nvert, c->lnvert, 1);

  /* near the immersed boundary, ibm_iso.c:349-381 */
#define NVL(k, j, i) c->lnvert[LI(c, k, j, i)]
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        double *nv = &nvert[GI(c, k, j, i)];
        if (*nv < 0) *nv = 0;
        const int ip = i < mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
        const int jp = j < my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
        const int kp = k < mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
        if ((int)(*nv + 0.5) != 4)
          for (int kk = km; kk < kp + 1; kk++)
            for (int jj = jm; jj < jp + 1; jj++)
              for (int ii = im; ii < ip + 1; ii++)
                if ((int)(NVL(kk, jj, ii) + 0.5) == 4) *nv = pmax(2, *nv);
      }
  orc_g2l(c, c->nvert, c->lnvert, 1);

  /* islanding, ibm_iso.c:397-486: both passes read the same lNvert */
  for (int tmp = 0; tmp < 2; tmp++) {
    if (iso != 1 && iso != 2) break;
    const int lo = iso == 2 ? 0 : 1;
    for (int k = lo; k < mz - lo; k++)
      for (int j = lo; j < my - lo; j++)
        for (int i = lo; i < mx - lo; i++) {
          double *nv = &nvert[GI(c, k, j, i)];
          if (*nv < 0) *nv = 0;
          const int ip = i < mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
          const int jp = j < my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
          const int kp = k < mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
          int around = 0;
          if ((int)(*nv + 0.5) < 0.5) {
            for (int kk = km; kk < kp + 1; kk++)
              if ((int)(NVL(kk, j, i) + 0.5) > 1 && kk != k) around++;
            for (int jj = jm; jj < jp + 1; jj++)
              if ((int)(NVL(k, jj, i) + 0.5) > 1 && jj != j) around++;
            for (int ii = im; ii < ip + 1; ii++)
              if ((int)(NVL(k, j, ii) + 0.5) > 1 && ii != i) around++;
          }
          if (around > 3) *nv = pmax(2, *nv);
        }
  }
#undef NVL
  orc_g2l(c, c->nvert, c->lnvert, 1);

  /* phase change against the previous classification, ibm_iso.c:495-507 */
  if (nvert_o)
    for (size_t g = 0; g < c->nG; g++)
      if (nvert_o[g] > 2.5 && nvert[g] < 0.5) nvert[g] = 2;
  orc_g2l(c, c->nvert, c->lnvert, 1);

  /* interpolation list, ibm_iso.c:522-757 (AddIBMNode prepends, so the
   * traversal order is the reverse of the scan order) */
  bounding_sphere(c);
  int cap = 64, n = 0;
  orc_ibm_info *scan = (orc_ibm_info *)malloc(cap * sizeof(orc_ibm_info));
  if (iso == 1 || iso == 2) {
    const int lo = iso == 2 ? 0 : 1;
    for (int k = iso == 2 ? 0 : 1; k < mz - 1; k++)
      for (int j = lo; j < my - lo; j++)
        for (int i = lo; i < mx - lo; i++) {
          if ((int)(nvert[GI(c, k, j, i)] + 0.5) != 2) continue;
          orc_ibm_info info;
          memset(&info, 0, sizeof(info));
          info.ni = i; info.nj = j; info.nk = k;
          const orc_v3 p = c->coor[GI(c, k, j, i)];
          nearest_cell(c, p, &info);
          if (info.cell < 0) { free(scan); return -2; }
          interception_point(c, p, i, j, k, &info);
          if (info.imode < 0) {
            const int cellm = info.cell;
            if (iso == 2) {
              if (k == 0 || k == mz - 1) single_donor(c, p, i, j + 1, k, &info);
              else if (i == 0 || i == mx - 1 || j == 0 || j == my - 1)
                single_donor(c, p, i, j, c->nf_z[cellm] > 0 ? k + 1 : k - 1, &info);
              else { free(scan); return -3; }
            } else {
              if (i == 1 || i == mx - 2 || j == 1 || j == my - 2)
                single_donor(c, p, i, j, c->nf_z[cellm] > 0 ? k + 1 : k - 1, &info);
              else if (k == 1 || k == mz - 2) single_donor(c, p, i, j + 1, k, &info);
              else { free(scan); return -3; }
            }
          }
          if (n == cap) { cap *= 2; scan = (orc_ibm_info *)realloc(scan, cap * sizeof(orc_ibm_info)); }
          scan[n++] = info;
        }
  }
  c->n_ibl = n;
  c->ibl = (orc_ibm_info *)malloc((n ? n : 1) * sizeof(orc_ibm_info));
  for (int q = 0; q < n; q++) c->ibl[q] = scan[n - 1 - q];
  free(scan);

  /* back to 1 and 3, ibm_iso.c:761-769 (ibi = 0) */
  for (size_t g = 0; g < c->nG; g++) {
    if ((int)(nvert[g] + 0.5) == 2) nvert[g] = 1 + 0 / 1000.0;
    if ((int)(nvert[g] + 0.5) == 4) nvert[g] = 3 + 0 / 1000.0;
  }
  orc_g2l(c, c->nvert, c->lnvert, 1);
  c->immersed = 1;
  return n;
}

int orc_ibm_count(const orc_ctx *c) { return c->n_ibl; }

void orc_ibm_get_list(const orc_ctx *c, orc_ibm_info *out) {
  if (c->n_ibl) memcpy(out, c->ibl, (size_t)c->n_ibl * sizeof(orc_ibm_info));
}

void orc_ibm_set_velocity(orc_ctx *c, const double *u, const double *uold) {
  for (int n = 0; n < c->n_v; n++) {
    c->u_bp[n] = (orc_v3){u[3 * n], u[3 * n + 1], u[3 * n + 2]};
    c->uold_bp[n] = (orc_v3){uold[3 * n], uold[3 * n + 1], uold[3 * n + 2]};
  }
}

/* --------------------------------------------------- ibm_interpolation_advanced */

/* noslip / freeslip, wallfunction.c:12-46 */
static orc_v3 noslip(double sc, double sb, orc_v3 Ua, orc_v3 Uc) {
  const double u_c = Uc.x - Ua.x, v_c = Uc.y - Ua.y, w_c = Uc.z - Ua.z;
  orc_v3 Ub = {sb / sc * u_c, sb / sc * v_c, sb / sc * w_c};
  Ub.x += Ua.x; Ub.y += Ua.y; Ub.z += Ua.z;
  return Ub;
}
static orc_v3 freeslip(double sc, double sb, orc_v3 Ua, orc_v3 Uc, double nx, double ny, double nz) {
  const double Uan = Ua.x * nx + Ua.y * ny + Ua.z * nz;
  const double Ucn = Uc.x * nx + Uc.y * ny + Uc.z * nz;
  const double Ubn = Uan + (Ucn - Uan) * sb / sc;
  orc_v3 Ub = {Uc.x - Ucn * nx, Uc.y - Ucn * ny, Uc.z - Ucn * nz};
  Ub.x += Ubn * nx; Ub.y += Ubn * ny; Ub.z += Ubn * nz;
  return Ub;
}

/* ibm_interpolation_advanced, ibm_iso.c:2572-2880: 5 sweeps that read lQ and
 * lP of the previous sweep and write Q and P at the interpolation nodes.
 * Pa carries from one list node to the next within a sweep, as in the
 * reference (it is a function-scope variable set only in the first sweep). */
void orc_ibm_interpolation(orc_ctx *c, int add_dundt) {
  const orc_params *P = &c->p;
  const double gam = P->gamma, Ma = P->Ma, Twall = P->Twall, dt = P->dt;
  const int iso = P->isothermal;
  double Pa = 0;
  for (int tmp = 0; tmp < 5; tmp++) {
    for (int q = 0; q < c->n_ibl; q++) {
      const orc_ibm_info *b = &c->ibl[q];
      const int ni = b->cell, i = b->ni, j = b->nj, k = b->nk;
      const size_t g = GI(c, k, j, i);
      const size_t l1 = LI(c, b->k1, b->j1, b->i1), l2 = LI(c, b->k2, b->j2, b->i2), l3 = LI(c, b->k3, b->j3, b->i3);
      const double sb = b->d_s, sc = sb + b->d_i;
      const double sk1 = b->cr1, sk2 = b->cr2, sk3 = b->cr3;
      const double cs1 = b->cs1, cs2 = b->cs2, cs3 = b->cs3;
      const orc_v3 x = c->coor[g];
      /* -twoD drops one component of the normal, ibm_iso.c:2662-2664 */
      const int twoD = P->twoD;
      double nfx = twoD == 1 ? 0. : x.x - b->pmin[0];
      double nfy = twoD == 2 ? 0. : x.y - b->pmin[1];
      double nfz = twoD == 3 ? 0. : x.z - b->pmin[2];
      const double dr = sqrt(nfx * nfx + nfy * nfy + nfz * nfz);
      if (dr < 0.000000001) {
        if (ni >= 0) { nfx = c->nf_x[ni]; nfy = c->nf_y[ni]; nfz = c->nf_z[ni]; }
        else nfx = nfy = nfz = 0.;
      } else {
        nfx = nfx / dr; nfy = nfy / dr; nfz = nfz / dr;
      }
      orc_v3 Ua;
      double Ua_n, Ua_nold;
      if (ni >= 0) {
        const orc_v3 *uo = c->uold_bp, *u = c->u_bp;
        const int a1 = c->nv1[ni], a2 = c->nv2[ni], a3 = c->nv3[ni];
        Ua.x = uo[a1].x * cs1 + uo[a2].x * cs2 + uo[a3].x * cs3;
        Ua.y = uo[a1].y * cs1 + uo[a2].y * cs2 + uo[a3].y * cs3;
        Ua.z = uo[a1].z * cs1 + uo[a2].z * cs2 + uo[a3].z * cs3;
        Ua_nold = Ua.x * nfx + Ua.y * nfy + Ua.z * nfz;
        Ua.x = u[a1].x * cs1 + u[a2].x * cs2 + u[a3].x * cs3;
        Ua.y = u[a1].y * cs1 + u[a2].y * cs2 + u[a3].y * cs3;
        Ua.z = u[a1].z * cs1 + u[a2].z * cs2 + u[a3].z * cs3;
        Ua_n = Ua.x * nfx + Ua.y * nfy + Ua.z * nfz;
      } else {
        Ua.x = Ua.y = Ua.z = 0; Ua_n = 0; Ua_nold = 0.;
      }
      const orc_q *q1 = &c->lQ[l1], *q2 = &c->lQ[l2], *q3 = &c->lQ[l3];
      orc_v3 Uc;
      Uc.x = (q1->rhoU / q1->rho * sk1 + q2->rhoU / q2->rho * sk2 + q3->rhoU / q3->rho * sk3);
      Uc.y = (q1->rhoV / q1->rho * sk1 + q2->rhoV / q2->rho * sk2 + q3->rhoV / q3->rho * sk3);
      Uc.z = (q1->rhoW / q1->rho * sk1 + q2->rhoW / q2->rho * sk2 + q3->rhoW / q3->rho * sk3);
      orc_v3 Ub = P->invicid ? freeslip(sc, sb, Ua, Uc, nfx, nfy, nfz) : noslip(sc, sb, Ua, Uc);
      double cv1 = c->lP[l1], cv2 = c->lP[l2], cv3 = c->lP[l3];
      double *p = &c->P[g];
      orc_q *qq = &c->Q[g];
      *p = (cv1 * sk1 + cv2 * sk2 + cv3 * sk3);
      /* 2D constraint, ibm_iso.c:2744-2746 */
      if (twoD == 1) Ub.x = 0.;
      if (twoD == 2) Ub.y = 0.;
      if (twoD == 3) Ub.z = 0.;
      if (iso == 2) {
        cv1 = q1->rho; cv2 = q2->rho; cv3 = q3->rho;
        const double rhoc = (cv1 * sk1 + cv2 * sk2 + cv3 * sk3);
        double Tc = *p / rhoc * (gam * Ma * Ma);
        if (Tc < 0.) Tc = 0.;
        const double Tb = (Tc - Twall) * sb / sc + Twall;
        const double a_sound = sqrt(Twall / (Ma * Ma));
        const double Lambda1 = Ua_n - a_sound;
        if (tmp == 0) { Pa = *p; qq->rho = rhoc; }
        const double L1 = Lambda1 / sc *
                          ((*p - Pa) - rhoc * a_sound * ((Uc.x - Ua.x) * nfx + (Uc.y - Ua.y) * nfy + (Uc.z - Ua.z) * nfz));
        const double rhob = c->Qold[g].rho - dt * L1 / a_sound / a_sound;
        qq->rho = rhob;
        *p = qq->rho * Tb / (gam * Ma * Ma);
        Pa = *p;
      } else if (iso == 1) {
        const double Tb = 1.0; /* ibm_iso.c:2790-2796 computes Tc and Tb, then sets Tb = 1 */
        qq->rho = *p * (gam * Ma * Ma) / Tb;
      } else {
        cv1 = q1->rho; cv2 = q2->rho; cv3 = q3->rho;
        qq->rho = (cv1 * sk1 + cv2 * sk2 + cv3 * sk3);
      }
      if (add_dundt) *p += qq->rho * (Ua_n - Ua_nold) / dt * (sc - sb);
      qq->rhoU = qq->rho * Ub.x;
      qq->rhoV = qq->rho * Ub.y;
      qq->rhoW = qq->rho * Ub.z;
      qq->rhoE = *p / (gam - 1) + 0.5 * (qq->rhoU * qq->rhoU + qq->rhoV * qq->rhoV + qq->rhoW * qq->rhoW) / qq->rho;
    }
    orc_g2l(c, (double *)c->Q, (double *)c->lQ, 5);
    orc_g2l(c, c->P, c->lP, 1);
  }
}

/* exported for the known-answer test of the exactly rounded x^1.5 */
double orc_pow15(double x) { return orc_pow15_cr(x); }

/* CPU restatement of the surface forces of the immersed body:
 * Calc_forces_SI2 (fsi.c:1312-2470) for one body and one block, on a single
 * rank.  Test infrastructure only.
 *
 * Loop over the interpolation list in the reference's order (the prepended
 * list, fsi.c:1453-1457); for a node inside the bounds of fsi.c:1478 (note
 * the j lower bound ys, not lys): the viscous stress from lUcat with the
 * one-sided differences of fsi.c:1486-1640, then the six faces k+, k-, j+,
 * j- (own metrics and coordinates at j = 0), i+, i- (fsi.c:1686-2290), each
 * counted when its neighbour is fluid (nvert < 0.9), then the dp/dn moments
 * (fsi.c:2292-2298).  The running sums are updated statement by statement in
 * the reference's order; the per-face MF* sums restart at 0 for every face,
 * the MFdpdn* sums for every node.  r_x, r_y, r_z keep the values of the last
 * face assigned, across nodes, as the reference's function-scope variables
 * do (they start at 0 here).
 */
#include "oracle_internal.h"

#include <string.h>

void orc_ibm_forces(const orc_ctx *c, const double center[3], double out[ORC_FORCE_N]) {
  const orc_params *P = &c->p;
  const int mx = c->mx, my = c->my, mz = c->mz;
  const double rei = 1. / P->ren;
  const double X_c = center[0], Y_c = center[1], Z_c = center[2];
  const int lxs = 1, lxe = mx - 1, ys = 0, lye = my - 1, lzs = 1, lze = mz - 1;
#define NV(k, j, i) c->lnvert[LI(c, k, j, i)]
#define U(k, j, i) c->lUcat[LI(c, k, j, i)]
#define AJ(k, j, i) c->laj[LI(c, k, j, i)]
#define CSI(k, j, i) c->lcsi[LI(c, k, j, i)]
#define ETA(k, j, i) c->leta[LI(c, k, j, i)]
#define ZET(k, j, i) c->lzet[LI(c, k, j, i)]
#define PP(k, j, i) c->lP[LI(c, k, j, i)]
#define CO(k, j, i) c->coor[GI(c, k, j, i)]
  double Cp_px = 0, Cp_py = 0, Cp_pz = 0, Cs_px = 0, Cs_py = 0, Cs_pz = 0;
  double Cp_nx = 0, Cp_ny = 0, Cp_nz = 0, Cs_nx = 0, Cs_ny = 0, Cs_nz = 0;
  double Ap_x = 0, Ap_y = 0, Ap_z = 0, An_x = 0, An_y = 0, An_z = 0;
  double Iap_x = 0, Iap_y = 0, Iap_z = 0, Ian_x = 0, Ian_y = 0, Ian_z = 0;
  double M_px = 0, M_py = 0, M_pz = 0, M_nx = 0, M_ny = 0, M_nz = 0;
  double Mdpdn_px = 0, Mdpdn_py = 0, Mdpdn_pz = 0, Mdpdn_nx = 0, Mdpdn_ny = 0, Mdpdn_nz = 0;
  double r_x = 0, r_y = 0, r_z = 0;
  for (int m = 0; m < c->n_ibl; m++) {
    const orc_ibm_info *in = &c->ibl[m];
    const int i = in->ni, j = in->nj, k = in->nk, elmt = in->cell;
    const double sb = in->d_s;
    const double nfy = c->nf_y[elmt], nfz = c->nf_z[elmt];
    if (!(i >= lxs && i < lxe && j >= ys && j < lye && k >= lzs && k < lze)) continue;
    double MFdpdn_px = 0, MFdpdn_py = 0, MFdpdn_pz = 0, MFdpdn_nx = 0, MFdpdn_ny = 0, MFdpdn_nz = 0;
    double Txx, Tyy, Tzz, Tzy, Tzx, Tyx;
    if (!P->invicid) {
      double dwdz, dvdz, dudz, dwdy, dvdy, dudy, dwdx, dvdx, dudx;
      double g1, g2, g3;
      /* k differences, fsi.c:1486-1547 */
      if (NV(k + 1, j, i) < 2.5 && NV(k - 1, j, i) < 2.5 && k + 1 < mz - 1 && k - 1 > 0) {
        g1 = 0.25 * ((ZET(k, j, i).x + ZET(k - 1, j, i).x) * (AJ(k, j, i) + AJ(k - 1, j, i)));
        g2 = 0.25 * ((ZET(k, j, i).y + ZET(k - 1, j, i).y) * (AJ(k, j, i) + AJ(k - 1, j, i)));
        g3 = 0.25 * ((ZET(k, j, i).z + ZET(k - 1, j, i).z) * (AJ(k, j, i) + AJ(k - 1, j, i)));
        dwdz = (U(k + 1, j, i).z - U(k - 1, j, i).z) / 2. * g3;
        dvdz = (U(k + 1, j, i).y - U(k - 1, j, i).y) / 2. * g3;
        dudz = (U(k + 1, j, i).x - U(k - 1, j, i).x) / 2. * g3;
        dwdy = (U(k + 1, j, i).z - U(k - 1, j, i).z) / 2. * g2;
        dvdy = (U(k + 1, j, i).y - U(k - 1, j, i).y) / 2. * g2;
        dudy = (U(k + 1, j, i).x - U(k - 1, j, i).x) / 2. * g2;
        dwdx = (U(k + 1, j, i).z - U(k - 1, j, i).z) / 2. * g1;
        dvdx = (U(k + 1, j, i).y - U(k - 1, j, i).y) / 2. * g1;
        dudx = (U(k + 1, j, i).x - U(k - 1, j, i).x) / 2. * g1;
      } else if (NV(k + 1, j, i) < 2.5 && k + 1 < mz - 1) {
        g1 = ZET(k, j, i).x * AJ(k, j, i);
        g2 = ZET(k, j, i).y * AJ(k, j, i);
        g3 = ZET(k, j, i).z * AJ(k, j, i);
        dwdz = (U(k + 1, j, i).z - U(k, j, i).z) * g3;
        dvdz = (U(k + 1, j, i).y - U(k, j, i).y) * g3;
        dudz = (U(k + 1, j, i).x - U(k, j, i).x) * g3;
        dwdy = (U(k + 1, j, i).z - U(k, j, i).z) * g2;
        dvdy = (U(k + 1, j, i).y - U(k, j, i).y) * g2;
        dudy = (U(k + 1, j, i).x - U(k, j, i).x) * g2;
        dwdx = (U(k + 1, j, i).z - U(k, j, i).z) * g1;
        dvdx = (U(k + 1, j, i).y - U(k, j, i).y) * g1;
        dudx = (U(k + 1, j, i).x - U(k, j, i).x) * g1;
      } else if (NV(k - 1, j, i) < 2.5 && k - 1 > 0) {
        g1 = ZET(k - 1, j, i).x * AJ(k - 1, j, i);
        g2 = ZET(k - 1, j, i).y * AJ(k - 1, j, i);
        g3 = ZET(k - 1, j, i).z * AJ(k - 1, j, i);
        dwdz = (U(k, j, i).z - U(k - 1, j, i).z) * g3;
        dvdz = (U(k, j, i).y - U(k - 1, j, i).y) * g3;
        dudz = (U(k, j, i).x - U(k - 1, j, i).x) * g3;
        dwdy = (U(k, j, i).z - U(k - 1, j, i).z) * g2;
        dvdy = (U(k, j, i).y - U(k - 1, j, i).y) * g2;
        dudy = (U(k, j, i).x - U(k - 1, j, i).x) * g2;
        dwdx = (U(k, j, i).z - U(k - 1, j, i).z) * g1;
        dvdx = (U(k, j, i).y - U(k - 1, j, i).y) * g1;
        dudx = (U(k, j, i).x - U(k - 1, j, i).x) * g1;
      } else {
        dwdz = dvdz = dudz = dwdy = dvdy = dudy = dwdx = dvdx = dudx = 0.;
      }
      /* j differences added, fsi.c:1550-1593; i differences, fsi.c:1594-1637 */
      for (int dir = 1; dir >= 0; dir--) {
        const int dj = dir == 1, di = dir == 0;
        const int jp = j + dj, jm = j - dj, ip = i + di, im = i - di;
        const int idx = dir == 1 ? j : i, lim = dir == 1 ? my : mx;
        const orc_v3 *mt0, *mtm;
        mt0 = dir == 1 ? &ETA(k, j, i) : &CSI(k, j, i);
        mtm = dir == 1 ? &ETA(k, jm, im) : &CSI(k, jm, im);
        int cse = 0;
        if (NV(k, jp, ip) < 2.5 && idx + 1 < lim - 1 && NV(k, jm, im) < 2.5 && idx - 1 > 0) cse = 1;
        else if (NV(k, jp, ip) < 2.5 && idx + 1 < lim - 1) cse = 2;
        else if (NV(k, jm, im) < 2.5 && idx - 1 > 0) cse = 3;
        if (cse == 1) {
          g1 = 0.25 * (mt0->x + mtm->x) * (AJ(k, j, i) + AJ(k, jm, im));
          g2 = 0.25 * (mt0->y + mtm->y) * (AJ(k, j, i) + AJ(k, jm, im));
          g3 = 0.25 * (mt0->z + mtm->z) * (AJ(k, j, i) + AJ(k, jm, im));
          dwdz += (U(k, jp, ip).z - U(k, jm, im).z) / 2. * g3;
          dvdz += (U(k, jp, ip).y - U(k, jm, im).y) / 2. * g3;
          dudz += (U(k, jp, ip).x - U(k, jm, im).x) / 2. * g3;
          dwdy += (U(k, jp, ip).z - U(k, jm, im).z) / 2. * g2;
          dvdy += (U(k, jp, ip).y - U(k, jm, im).y) / 2. * g2;
          dudy += (U(k, jp, ip).x - U(k, jm, im).x) / 2. * g2;
          dwdx += (U(k, jp, ip).z - U(k, jm, im).z) / 2. * g1;
          dvdx += (U(k, jp, ip).y - U(k, jm, im).y) / 2. * g1;
          dudx += (U(k, jp, ip).x - U(k, jm, im).x) / 2. * g1;
        } else if (cse == 2 || cse == 3) {
          /* one-sided: forward from the node, or backward from the -1 node */
          const int ja = cse == 2 ? jp : j, ia = cse == 2 ? ip : i;  /* upper node */
          const int jb = cse == 2 ? j : jm, ib = cse == 2 ? i : im;  /* lower node */
          const orc_v3 *mt = cse == 2 ? mt0 : mtm;
          const double ajb = cse == 2 ? AJ(k, j, i) : AJ(k, jm, im);
          g1 = mt->x * ajb;
          g2 = mt->y * ajb;
          g3 = mt->z * ajb;
          dwdz += (U(k, ja, ia).z - U(k, jb, ib).z) * g3;
          dvdz += (U(k, ja, ia).y - U(k, jb, ib).y) * g3;
          dudz += (U(k, ja, ia).x - U(k, jb, ib).x) * g3;
          dwdy += (U(k, ja, ia).z - U(k, jb, ib).z) * g2;
          dvdy += (U(k, ja, ia).y - U(k, jb, ib).y) * g2;
          dudy += (U(k, ja, ia).x - U(k, jb, ib).x) * g2;
          dwdx += (U(k, ja, ia).z - U(k, jb, ib).z) * g1;
          dvdx += (U(k, ja, ia).y - U(k, jb, ib).y) * g1;
          dudx += (U(k, ja, ia).x - U(k, jb, ib).x) * g1;
        }
      }
      Tzz = rei * (dwdz + dwdz);
      Tyy = rei * (dvdy + dvdy);
      Txx = rei * (dudx + dudx);
      Tzy = rei * (dwdy + dvdz);
      Tzx = rei * (dwdx + dudz);
      Tyx = rei * (dvdx + dudy);
    } else {
      Tzz = Tyy = Txx = Tzy = Tzx = Tyx = 0.;
    }
    const double p = PP(k, j, i);
    /* faces k+, k-, j+, j-, i+, i- (fsi.c:1660-2290) */
    for (int f = 0; f < 6; f++) {
      const int d = f / 2, plus = (f & 1) == 0;
      const int dk = d == 0 ? (plus ? 1 : -1) : 0, dj = d == 1 ? (plus ? 1 : -1) : 0, di = d == 2 ? (plus ? 1 : -1) : 0;
      const orc_v3 *mt0 = d == 0 ? &ZET(k, j, i) : d == 1 ? &ETA(k, j, i) : &CSI(k, j, i);
      double A_x, A_y, A_z, x, y, z;
      if (d == 1 && !plus && j == 0) { /* fsi.c:1921-1927: the node's own metric and coordinates */
        A_x = mt0->x; A_y = mt0->y; A_z = mt0->z;
        x = CO(k, j, i).x; y = CO(k, j, i).y; z = CO(k, j, i).z;
      } else if (NV(k + dk, j + dj, i + di) < 0.9) {
        const orc_v3 *mt1 = d == 0 ? &ZET(k + dk, j, i) : d == 1 ? &ETA(k, j + dj, i) : &CSI(k, j, i + di);
        A_x = 0.5 * (mt0->x + mt1->x);
        A_y = 0.5 * (mt0->y + mt1->y);
        A_z = 0.5 * (mt0->z + mt1->z);
        x = 0.5 * (CO(k, j, i).x + CO(k + dk, j + dj, i + di).x);
        y = 0.5 * (CO(k, j, i).y + CO(k + dk, j + dj, i + di).y);
        z = 0.5 * (CO(k, j, i).z + CO(k + dk, j + dj, i + di).z);
      } else {
        continue;
      }
      r_x = x - X_c;
      r_y = y - Y_c;
      r_z = z - Z_c;
      double MF_x = 0., MF_y = 0., MF_z = 0.;
      if (plus) {
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
#undef NV
#undef U
#undef AJ
#undef CSI
#undef ETA
#undef ZET
#undef PP
#undef CO
  /* totals, fsi.c:2300-2420 (one rank: the Allreduce sums are the values) */
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

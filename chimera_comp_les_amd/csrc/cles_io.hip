// cles_io.hip -- the reference's binary field files: Q_P_Binary_Output and
// Q_P_Binary_Input (main.c:215-304) with the restart of RestartApplication
// (main.c:621-650).
//
// Each field is one PETSc binary Vec file as VecView writes it for a
// single-rank DMDA vector (PETSc 3.12, 32-bit PetscInt): a big-endian int32
// class id 1211214 (VEC_FILE_CLASSID), a big-endian int32 length, then the
// values as big-endian doubles in the natural ordering (i fastest, the dof of
// a node interleaved).  VecView also writes `<file>.info` with
// "-vecload_block_size <dof>" for dof > 1.  File names follow the reference:
// Qfield%05d_%1d.dat, pfield..., nvfield..., with LES sigma... and nu_t...,
// with -averaging (after step 0) Qfield_ave..., su2_Sq..., su_cross... and
// psum...; the number is the time step, the digit the block.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cles_internal.hpp"

namespace cles {
namespace {

constexpr int32_t kVecClassId = 1211214;

uint32_t be32(uint32_t x) { return __builtin_bswap32(x); }
uint64_t be64(uint64_t x) { return __builtin_bswap64(x); }

void field_path(char *out, size_t cap, const char *dir, const char *stem, int ti, int block) {
  if (dir && dir[0])
    snprintf(out, cap, "%s/%s%5.5d_%1.1d.dat", dir, stem, ti, block);
  else
    snprintf(out, cap, "%s%5.5d_%1.1d.dat", stem, ti, block);
}

}  // namespace
}  // namespace cles

using namespace cles;

extern "C" {

int cles_petsc_vec_write(const char *path, const double *x, long long n, int block_size) {
  if (!path || (!x && n > 0) || n < 0 || n > INT32_MAX || block_size < 1)
    return set_error(nullptr, CLES_EINVAL, "cles_petsc_vec_write: bad arguments");
  FILE *f = fopen(path, "wb");
  if (!f) return set_error(nullptr, CLES_EINVAL, "cannot create %s: %s", path, strerror(errno));
  const uint32_t head[2] = {be32((uint32_t)kVecClassId), be32((uint32_t)n)};
  bool ok = fwrite(head, sizeof(head), 1, f) == 1;
  std::vector<uint64_t> buf(1 << 16);
  for (long long o = 0; ok && o < n; o += (long long)buf.size()) {
    const long long m = n - o < (long long)buf.size() ? n - o : (long long)buf.size();
    for (long long q = 0; q < m; q++) {
      uint64_t u;
      memcpy(&u, x + o + q, 8);
      buf[q] = be64(u);
    }
    ok = fwrite(buf.data(), 8, (size_t)m, f) == (size_t)m;
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok) return set_error(nullptr, CLES_EINVAL, "write error on %s", path);
  if (block_size > 1) {
    std::string info = std::string(path) + ".info";
    FILE *g = fopen(info.c_str(), "w");
    if (!g) return set_error(nullptr, CLES_EINVAL, "cannot create %s", info.c_str());
    fprintf(g, "-vecload_block_size %d\n", block_size);
    if (fclose(g) != 0) return set_error(nullptr, CLES_EINVAL, "write error on %s", info.c_str());
  }
  return CLES_OK;
}

int cles_petsc_vec_read(const char *path, double *x, long long n) {
  if (!path || (!x && n > 0) || n < 0) return set_error(nullptr, CLES_EINVAL, "cles_petsc_vec_read: bad arguments");
  FILE *f = fopen(path, "rb");
  if (!f) return set_error(nullptr, CLES_EINVAL, "cannot open %s: %s", path, strerror(errno));
  uint32_t head[2];
  if (fread(head, sizeof(head), 1, f) != 1) {
    fclose(f);
    return set_error(nullptr, CLES_EINVAL, "%s: truncated header", path);
  }
  const int32_t id = (int32_t)be32(head[0]), len = (int32_t)be32(head[1]);
  if (id != kVecClassId) {
    fclose(f);
    return set_error(nullptr, CLES_EINVAL, "%s: not a PETSc binary Vec (class id %d)", path, id);
  }
  if ((long long)len != n) {
    fclose(f);
    return set_error(nullptr, CLES_EINVAL, "%s: holds %d values, expected %lld", path, len, n);
  }
  std::vector<uint64_t> buf(1 << 16);
  bool ok = true;
  for (long long o = 0; ok && o < n; o += (long long)buf.size()) {
    const long long m = n - o < (long long)buf.size() ? n - o : (long long)buf.size();
    ok = fread(buf.data(), 8, (size_t)m, f) == (size_t)m;
    for (long long q = 0; ok && q < m; q++) {
      const uint64_t u = be64(buf[q]);
      memcpy(x + o + q, &u, 8);
    }
  }
  fclose(f);
  if (!ok) return set_error(nullptr, CLES_EINVAL, "%s: truncated data", path);
  return CLES_OK;
}

int cles_write_fields(cles_ctx *c, const char *dir, int block) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  std::vector<double> q((size_t)N * 5), p((size_t)N), nv((size_t)N);
  int r = cles_get_state(c, q.data(), p.data(), CLES_LAYOUT_GLOBAL);
  if (r) return r;
  if ((r = cles_get_nvert(c, nv.data()))) return r;
  char path[1024];
  const int ti = c->ti;
  // the order of Q_P_Binary_Output (main.c:246-300)
  field_path(path, sizeof(path), dir, "Qfield", ti, block);
  if ((r = cles_petsc_vec_write(path, q.data(), N * 5, 5))) return set_error(c, r, "%s", cles_last_error(nullptr));
  if (c->p.les) {
    std::vector<double> sig((size_t)N * 3);
    if ((r = cles_get_les(c, nullptr, nullptr, sig.data()))) return r;
    field_path(path, sizeof(path), dir, "sigma", ti, block);
    if ((r = cles_petsc_vec_write(path, sig.data(), N * 3, 3))) return set_error(c, r, "%s", cles_last_error(nullptr));
  }
  field_path(path, sizeof(path), dir, "pfield", ti, block);
  if ((r = cles_petsc_vec_write(path, p.data(), N, 1))) return set_error(c, r, "%s", cles_last_error(nullptr));
  field_path(path, sizeof(path), dir, "nvfield", ti, block);
  if ((r = cles_petsc_vec_write(path, nv.data(), N, 1))) return set_error(c, r, "%s", cles_last_error(nullptr));
  if (c->p.averaging && ti != 0) {  // main.c:274-300
    std::vector<double> u2((size_t)N * 3), uc((size_t)N * 3);
    if ((r = cles_get_averages(c, q.data(), p.data(), u2.data(), uc.data()))) return r;
    field_path(path, sizeof(path), dir, "Qfield_ave", ti, block);
    if ((r = cles_petsc_vec_write(path, q.data(), N * 5, 5))) return set_error(c, r, "%s", cles_last_error(nullptr));
    field_path(path, sizeof(path), dir, "su2_Sq", ti, block);
    if ((r = cles_petsc_vec_write(path, u2.data(), N * 3, 3))) return set_error(c, r, "%s", cles_last_error(nullptr));
    field_path(path, sizeof(path), dir, "su_cross", ti, block);
    if ((r = cles_petsc_vec_write(path, uc.data(), N * 3, 3))) return set_error(c, r, "%s", cles_last_error(nullptr));
    field_path(path, sizeof(path), dir, "psum", ti, block);
    if ((r = cles_petsc_vec_write(path, p.data(), N, 1))) return set_error(c, r, "%s", cles_last_error(nullptr));
  }
  if (c->p.les) {
    if ((r = cles_get_les(c, nullptr, p.data(), nullptr))) return r;
    field_path(path, sizeof(path), dir, "nu_t", ti, block);
    if ((r = cles_petsc_vec_write(path, p.data(), N, 1))) return set_error(c, r, "%s", cles_last_error(nullptr));
  }
  return CLES_OK;
}

int cles_read_fields(cles_ctx *c, const char *dir, int ti, int block) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  std::vector<double> q((size_t)N * 5), p((size_t)N), nv((size_t)N);
  char path[1024];
  int r;
  // the order of Q_P_Binary_Input (main.c:215-243)
  field_path(path, sizeof(path), dir, "pfield", ti, block);
  if ((r = cles_petsc_vec_read(path, p.data(), N))) return set_error(c, r, "%s", cles_last_error(nullptr));
  field_path(path, sizeof(path), dir, "nvfield", ti, block);
  if ((r = cles_petsc_vec_read(path, nv.data(), N))) return set_error(c, r, "%s", cles_last_error(nullptr));
  field_path(path, sizeof(path), dir, "Qfield", ti, block);
  if ((r = cles_petsc_vec_read(path, q.data(), N * 5))) return set_error(c, r, "%s", cles_last_error(nullptr));
  // RestartApplication (main.c:627-650): Nvert, P and Q as read, Qold = Q
  if ((r = cles_set_nvert(c, nv.data()))) return r;
  if ((r = cles_set_state_qp(c, q.data(), p.data(), CLES_LAYOUT_GLOBAL))) return r;
  c->ti = ti;
  return CLES_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- VTK output
//
// VTKOut of the post-processor (data.c:385-790, the !Pres branch): one XML
// StructuredGrid file Result%2.2d_%6.6d.vts (block, time step) with raw
// appended little-endian Float64 arrays, each preceded by its byte count as
// a 4-byte int: Position (3), nvert, p, rho, E = rhoE/rho, u = rhoU/rho ...
// (3) and with LES Sigma (3) and nu_T.  -twoD collapses that extent to one
// node: the fields are taken at node 1 along x (twoD 1), the coordinates at
// node 0, and for twoD 2 / 3 every field at node 0 along y / z, as the
// reference's loops do.  No -skip subsampling and no block rotation (theta
// = 0 leaves the coordinates as they are).
namespace cles {
namespace {

struct VtkArray {
  const char *name;
  int nc;
};

}  // namespace
}  // namespace cles

// VTKOut (data.c:385-855) with its options.  The extents are those the
// reference writes: -twoD keeps one x/y/z node, -skip s > 1 keeps mx / s x
// nodes (node i from x node i s); the fields under -twoD 1 come from x node 1
// (Sigma from i + 1), the coordinates from node i (or i s).  Block 2's
// positions and velocities are rotated by theta about z (data.c:544-548,
// 678-681).  With pres the reference writes no file and prints the pressure
// of the first fluid nodes next to a body along the x node 20 (block 2) and
// 8 (airfoil) lines (data.c:772-844), which go to text here.  The rotation
// and the TKE sum are evaluated without fused multiply-adds, as the
// reference's C compiles on x86-64.
static int vtk_out(cles_ctx *c, const char *dir, int block, const cles_vtk_opts *o, double *tke, char *text,
                   size_t text_len, size_t *text_used) {
#pragma clang fp contract(off)
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  const int skip = o ? o->skip : 1;
  const double theta = o ? o->theta : 0.0;
  const int pres = o ? o->pres : 0, airfoil = o ? o->airfoil : 0;
  const int MX = c->g.mx, MY = c->g.my, MZ = c->g.mz;
  const long long N = (long long)MX * MY * MZ;
  const int twoD = c->p.twoD;
  int mx = twoD == 1 ? 1 : MX;
  const int my = twoD == 2 ? 1 : MY, mz = twoD == 3 ? 1 : MZ;
  if (skip > 1) mx = mx / skip;  // node (mx - 1) skip <= MX - skip stays on the grid
  const long long nn = (long long)mx * my * mz;
  if (3 * nn * 8 > INT32_MAX)
    return set_error(c, CLES_EINVAL, "VTK output: an array of %lld bytes does not fit the 4-byte size header", 3 * nn * 8);
  std::vector<double> q((size_t)N * 5), p((size_t)N), nv((size_t)N), xyz((size_t)N * 3), nut, sig;
  int r = cles_get_state(c, q.data(), p.data(), CLES_LAYOUT_GLOBAL);
  if (r) return r;
  if ((r = cles_get_nvert(c, nv.data()))) return r;
  if ((r = get_coor(c, xyz.data()))) return r;
  auto at = [&](int k, int j, int i) { return ((long long)k * MY + j) * MX + i; };
  if (pres) {
    // data.c:774-842: " PL \n " then, per line (block 2's, then airfoil's),
    // the nodes whose next node (k + 1, or j + 1 for airfoil) is solid,
    // " PU \n " and those whose previous one is; the loops run over nodes
    // 1 .. n - 3 of the written extents
    std::string txt = " PL \n ";
    char line[256];
    auto out_line = [&](const char *fmt, double a0, double a1, double a2, double a3, int n) {
      const int len = n == 4 ? snprintf(line, sizeof(line), fmt, a0, a1, a2, a3) : snprintf(line, sizeof(line), fmt, a0, a1, a2);
      txt.append(line, (size_t)len);
    };
    for (int part = 0; part < 2; part++) {
      if (part == 0 ? block != 2 : !airfoil) continue;
      const int i = part == 0 ? 20 : 8;
      if (i > MX - 1) return set_error(c, CLES_EINVAL, "pressure lines: x node %d outside the grid", i);
      for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) txt += " PU \n ";
        for (int a0 = 1; a0 < (part == 0 ? my : mz) - 2; a0++)
          for (int b0 = 1; b0 < (part == 0 ? mz : my) - 2; b0++) {
            const int j = part == 0 ? a0 : b0, k = part == 0 ? b0 : a0;
            const long long n = at(k, j, i);
            const long long nb = part == 0 ? at(pass == 0 ? k + 1 : k - 1, j, i) : at(k, pass == 0 ? j + 1 : j - 1, i);
            if (!(nv[n] == 0.0 && nv[nb] > 0.0)) continue;
            if (pass == 0)
              out_line("%f %f  %f  %f\n ", xyz[3 * n], xyz[3 * n + 2], xyz[3 * n + 1], p[n], 4);
            else
              out_line("%f  %f  %f   \n ", xyz[3 * n + 2], xyz[3 * n + 1], p[n], 0.0, 3);
          }
      }
    }
    if (text_used) *text_used = txt.size();
    if (text && text_len) {
      const size_t m = txt.size() < text_len - 1 ? txt.size() : text_len - 1;
      memcpy(text, txt.data(), m);
      text[m] = '\0';
    }
    if (tke) *tke = 0.0;
    return CLES_OK;
  }
  const bool les = c->p.les != 0;
  if (les) {
    nut.resize((size_t)N);
    sig.resize((size_t)N * 3);
    if ((r = cles_get_les(c, nullptr, nut.data(), sig.data()))) return r;
  }
  std::vector<VtkArray> arrs = {{"nvert", 1}, {"p", 1}, {"rho", 1}, {"E", 1}, {"u", 3}};
  if (les) {
    arrs.push_back({"Sigma", 3});
    arrs.push_back({"nu_T", 1});
  }
  char path[1024];
  if (dir && dir[0])
    snprintf(path, sizeof(path), "%s/Result%2.2d_%6.6d.vts", dir, block, c->ti);
  else
    snprintf(path, sizeof(path), "Result%2.2d_%6.6d.vts", block, c->ti);
  FILE *f = fopen(path, "wb");
  if (!f) return set_error(c, CLES_EINVAL, "cannot create %s: %s", path, strerror(errno));
  long long off = 0;
  fprintf(f, "<?xml version=\"1.0\"?>\n");
  fprintf(f, "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  fprintf(f, "  <StructuredGrid WholeExtent=\"%d %d %d %d %d %d\">\n", 0, mx - 1, 0, my - 1, 0, mz - 1);
  fprintf(f, "    <Piece Extent=\"%d %d %d %d %d %d\">\n", 0, mx - 1, 0, my - 1, 0, mz - 1);
  fprintf(f, "      <Points>\n");
  fprintf(f, "        <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"appended\" offset=\"%lld\" />\n", off);
  off += 3 * nn * 8 + 4;
  fprintf(f, "      </Points>\n");
  fprintf(f, "      <PointData Scalars=\"ScalarPointData\">\n");
  for (const VtkArray &a : arrs) {
    fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" offset=\"%lld\" />\n",
            a.name, a.nc, off);
    off += a.nc * nn * 8 + 4;
  }
  fprintf(f, "      </PointData>\n");
  fprintf(f, "    </Piece>\n");
  fprintf(f, "  </StructuredGrid>\n");
  fprintf(f, "  <AppendedData encoding=\"raw\">\n");
  fprintf(f, "_");
  std::vector<double> out((size_t)nn * 3);
  bool ok = true;
  auto put = [&](int nc) {
    const int bytes = (int)(nc * nn * 8);
    ok = ok && fwrite(&bytes, sizeof(int), 1, f) == 1 && fwrite(out.data(), 8, (size_t)(nc * nn), f) == (size_t)(nc * nn);
  };
  // source x node of output x node i: the coordinates i (or i skip), the
  // fields i skip, else 1 under twoD 1, else i; Sigma i skip, but i + 1
  // under twoD 1
  auto src = [&](int k, int j, int i, int kind) {
    int ii = i;
    if (kind == 0) ii = skip > 1 ? i * skip : i;
    else if (kind == 1) ii = skip > 1 ? i * skip : twoD == 1 ? 1 : i;
    else ii = twoD == 1 ? i + 1 : skip > 1 ? i * skip : i;
    return at(k, j, ii);
  };
  const bool rot = block == 2;
  const double ct = cos(theta), st = sin(theta);
  for (int k = 0; k < mz; k++)  // Position
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const long long o = ((long long)k * my + j) * mx + i, s = src(k, j, i, 0);
        double x = xyz[3 * s], y = xyz[3 * s + 1];
        if (rot) {
          x = xyz[3 * s] * ct - xyz[3 * s + 1] * st;
          y = xyz[3 * s + 1] * ct + xyz[3 * s] * st;
        }
        out[3 * o] = x;
        out[3 * o + 1] = y;
        out[3 * o + 2] = xyz[3 * s + 2];
      }
  put(3);
  double sum = 0.0;
  const double cells = (double)((long long)mz * my * mx);
  for (size_t kind = 0; kind < arrs.size(); kind++) {  // nvert p rho E u [Sigma nu_T]
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          const long long o = ((long long)k * my + j) * mx + i, s = src(k, j, i, kind == 5 ? 2 : 1);
          const double *qs = &q[5 * s];
          switch (kind) {
            case 0: out[o] = nv[s]; break;
            case 1: out[o] = p[s]; break;
            case 2: out[o] = qs[0]; break;
            case 3: out[o] = qs[4] / qs[0]; break;
            case 4: {
              const double ux = qs[1] / qs[0], uy = qs[2] / qs[0], uz = qs[3] / qs[0];
              double x = ux, y = uy;
              if (rot) {
                x = ux * ct - uy * st;
                y = uy * ct + ux * st;
              }
              out[3 * o] = x;
              out[3 * o + 1] = y;
              out[3 * o + 2] = uz;
              sum += (x * x + y * y * y + uz * uz) / cells;  // the reference's "TKE" (v cubed), data.c:688
              break;
            }
            case 5:
              for (int d = 0; d < 3; d++) out[3 * o + d] = sig[3 * s + d];
              break;
            default: out[o] = nut[s];
          }
        }
    put(arrs[kind].nc);
  }
  fprintf(f, "\n </AppendedData>\n");
  fprintf(f, "</VTKFile>\n");
  ok = (fclose(f) == 0) && ok;
  if (tke) *tke = sum;
  if (text_used) *text_used = 0;
  if (text && text_len) text[0] = '\0';
  return ok ? CLES_OK : set_error(c, CLES_EINVAL, "write error on %s", path);
}

// VTKOutQ (data.c:1562-1711): the Q criterion and lambda2 of QCriteria
// (data.c:4815-5053, cles_q_criteria) and nvert in a legacy ASCII VTK
// structured grid of (mx-1) x (my-1) x (mz-1) points, the nodes 0..m-2.  With
// vc 0 each point's value is the mean of the 8 nodes of the cell whose
// lowest corner it is (0.125 times their sum, in the reference's order);
// with vc 1 the node values of nodes 1..m-2 as CELL_DATA.  Values and
// coordinates are printed with %f, as the reference does.
extern "C" int cles_write_q_vtk(cles_ctx *c, const char *dir, int block, int vc) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  const int mx = c->g.mx, my = c->g.my, mz = c->g.mz;
  const long long N = (long long)mx * my * mz;
  std::vector<double> qc((size_t)N), l2((size_t)N), nv((size_t)N), xyz((size_t)N * 3);
  int r = cles_q_criteria(c, qc.data(), l2.data());
  if (r) return r;
  if ((r = cles_get_nvert(c, nv.data()))) return r;
  if ((r = get_coor(c, xyz.data()))) return r;
  char path[1024];
  if (dir && dir[0])
    snprintf(path, sizeof(path), "%s/QCriteria%2.2d_%5.5d.vtk", dir, block, c->ti);
  else
    snprintf(path, sizeof(path), "QCriteria%2.2d_%5.5d.vtk", block, c->ti);
  FILE *f = fopen(path, "w");
  if (!f) return set_error(c, CLES_EINVAL, "cannot create %s: %s", path, strerror(errno));
  auto at = [&](int k, int j, int i) { return ((long long)k * my + j) * mx + i; };
  fprintf(f, "# vtk DataFile Version 2.0\n");
  fprintf(f, "Structured Grid\n");
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET STRUCTURED_GRID\n");
  fprintf(f, "DIMENSIONS %d %d %d\n", mx - 1, my - 1, mz - 1);
  fprintf(f, "POINTS  %d double\n", (mx - 1) * (my - 1) * (mz - 1));
  for (int k = 0; k < mz - 1; k++)
    for (int j = 0; j < my - 1; j++)
      for (int i = 0; i < mx - 1; i++) {
        const long long n = at(k, j, i);
        fprintf(f, "%f %f %f\n", xyz[3 * n], xyz[3 * n + 1], xyz[3 * n + 2]);
      }
  if (vc)
    fprintf(f, "CELL_DATA %d\n", (mx - 2) * (my - 2) * (mz - 2));
  else
    fprintf(f, "POINT_DATA %d\n", (mx - 1) * (my - 1) * (mz - 1));
  const int lo = vc ? 1 : 0;
  const std::vector<double> *fields[3] = {&qc, &l2, &nv};
  const char *names[3] = {"q", "l2", "nvert"};
  for (int fi = 0; fi < 3; fi++) {
    const std::vector<double> &v = *fields[fi];
    fprintf(f, "SCALARS %s double\n", names[fi]);
    fprintf(f, "LOOKUP_TABLE default\n");
    for (int k = lo; k < mz - 1; k++)
      for (int j = lo; j < my - 1; j++)
        for (int i = lo; i < mx - 1; i++) {
          double x;
          if (vc)
            x = v[at(k, j, i)];
          else
            x = 0.125 * (v[at(k, j, i)] + v[at(k, j, i + 1)] + v[at(k, j + 1, i)] + v[at(k, j + 1, i + 1)] +
                         v[at(k + 1, j, i)] + v[at(k + 1, j, i + 1)] + v[at(k + 1, j + 1, i)] +
                         v[at(k + 1, j + 1, i + 1)]);
          fprintf(f, "%f\n", x);
        }
  }
  const bool ok = fclose(f) == 0;
  return ok ? CLES_OK : set_error(c, CLES_EINVAL, "write error on %s", path);
}

extern "C" int cles_write_vtk(cles_ctx *c, const char *dir, int block) {
  return vtk_out(c, dir, block, nullptr, nullptr, nullptr, 0, nullptr);
}

extern "C" int cles_write_vtk_opts(cles_ctx *c, const char *dir, int block, const cles_vtk_opts *o, double *tke,
                                   char *text, size_t text_len, size_t *text_used) {
  if (o && o->skip < 1) return c ? set_error(c, CLES_EINVAL, "skip must be at least 1") : CLES_EINVAL;
  return vtk_out(c, dir, block, o, tke, text, text_len, text_used);
}

// cles_batch.hip -- several independent periodic solves advanced together.
//
// A batch holds existing contexts (fully periodic grids, no LES model, no
// immersed body; any size, Mach and Reynolds number).  One batched step is,
// for each RK stage, one launch of the batched stage kernel per (precision,
// viscous) group, whose workgroups cover the tiles of every member, and one
// launch of the batched seam/EOS kernel over every member's boundary nodes.
// The step is captured once into a hipGraph and replayed; membership changes
// rebuild the argument tables and the graph.
//
// Buffers: a member's Q^n buffer X stays its current buffer; stage 0 writes
// Y, stage 1 writes Z and the last stage writes X in place (the stage kernel
// reads Qold only at its own node), so the pointers of a step never change.
// The arithmetic of every tile is the single-grid kernel's, so a member's
// result does not depend on the other members and equals cles_advance's.
//
// Admission is by HBM footprint: a member is accepted when its allocation
// fits into the batch's byte budget.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cles_internal.hpp"

struct cles_batch {
  hipStream_t stream = nullptr;
  size_t budget = 0, used = 0;
  int tile = 0;
  std::vector<cles_ctx *> members;
  // device argument tables: [stage][group] stage args + tile starts, [stage] seam args + block starts
  void *d_stage = nullptr, *d_seam = nullptr;
  int *d_tstart = nullptr, *d_bstart = nullptr;
  size_t cap = 0;  // members the tables hold
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
  int nstages = 0;
  char err[256] = {0};
};

namespace cles {
namespace {

constexpr int kGroups = 4;  // (fp32, viscous)

int batch_error(cles_batch *b, int code, const char *msg) {
  if (b) snprintf(b->err, sizeof(b->err), "%s", msg);
  return set_error(nullptr, code, "%s", msg);
}

void free_tables(cles_batch *b) {
  void *p[] = {b->d_stage, b->d_seam, b->d_tstart, b->d_bstart};
  for (void *q : p)
    if (q) (void)hipFree(q);
  b->d_stage = b->d_seam = nullptr;
  b->d_tstart = b->d_bstart = nullptr;
  b->cap = 0;
}

void drop_graph(cles_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int group_of(const cles_ctx *c) { return (c->p.precision == CLES_PREC_FP32BULK ? 2 : 0) + (c->p.invicid ? 0 : 1); }

// rebuild the tables and capture one batched step
int rebuild(cles_batch *b) {
  drop_graph(b);
  const size_t n = b->members.size();
  if (n == 0) {
    b->dirty = false;
    return CLES_OK;
  }
  int nst = b->members[0]->p.fluxorder;
  for (cles_ctx *c : b->members)
    if (c->p.fluxorder != nst) return batch_error(b, CLES_EINVAL, "batch members must share fluxorder");
  b->nstages = nst;
  const size_t sa = stage_args_bytes(), ba = seam_args_bytes();
  if (n > b->cap) {
    free_tables(b);
    const size_t cap = std::max<size_t>(n, 2 * b->cap);
    if (hipMalloc(&b->d_stage, 3 * kGroups * cap * sa) != hipSuccess ||
        hipMalloc(&b->d_seam, 3 * cap * ba) != hipSuccess ||
        hipMalloc((void **)&b->d_tstart, 3 * kGroups * (cap + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&b->d_bstart, 3 * (cap + 1) * sizeof(int)) != hipSuccess) {
      free_tables(b);
      return batch_error(b, CLES_ENOMEM, "batch argument tables");
    }
    b->cap = cap;
  }
  std::vector<char> hstage(3 * kGroups * b->cap * sa), hseam(3 * b->cap * ba);
  std::vector<int> htstart(3 * kGroups * (b->cap + 1), 0), hbstart(3 * (b->cap + 1), 0);
  int ngroup[kGroups] = {0, 0, 0, 0}, nblk[3][kGroups] = {}, nbnd[3] = {0, 0, 0};
  for (int s = 0; s < nst; s++) {
    int cnt[kGroups] = {0, 0, 0, 0};
    int r = 0;
    for (cles_ctx *c : b->members) {
      const int X = c->cur, Y = (X + 1) % 3, Z = (X + 2) % 3;
      int in, out;
      if (nst == 3) {
        in = s == 0 ? X : s == 1 ? Y : Z;
        out = s == 0 ? Y : s == 1 ? Z : X;
      } else {
        in = s == 0 ? X : Y;
        out = s == 0 ? Y : X;
      }
      double coef[3];
      rk_coef(c->p, s, coef);
      const int gi = group_of(c);
      const size_t slot = ((size_t)s * kGroups + gi) * b->cap + cnt[gi];
      const int tiles = stage_batch_args(c, in, out, s, coef, b->tile, hstage.data() + slot * sa);
      htstart[((size_t)s * kGroups + gi) * (b->cap + 1) + cnt[gi] + 1] =
          htstart[((size_t)s * kGroups + gi) * (b->cap + 1) + cnt[gi]] + tiles;
      cnt[gi]++;
      const int blocks = seam_batch_args(c, in, out, hseam.data() + ((size_t)s * b->cap + r) * ba);
      hbstart[(size_t)s * (b->cap + 1) + r + 1] = hbstart[(size_t)s * (b->cap + 1) + r] + blocks;
      r++;
    }
    for (int gi = 0; gi < kGroups; gi++) {
      ngroup[gi] = cnt[gi];
      nblk[s][gi] = htstart[((size_t)s * kGroups + gi) * (b->cap + 1) + cnt[gi]];
    }
    nbnd[s] = hbstart[(size_t)s * (b->cap + 1) + n];
  }
  if (hipMemcpy(b->d_stage, hstage.data(), hstage.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_seam, hseam.data(), hseam.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_tstart, htstart.data(), htstart.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_bstart, hbstart.data(), hbstart.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "batch table upload");
  // capture one step
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "hipStreamBeginCapture");
  hipError_t e = hipSuccess;
  for (int s = 0; s < nst && e == hipSuccess; s++) {
    for (int gi = 0; gi < kGroups && e == hipSuccess; gi++) {
      if (!ngroup[gi]) continue;
      const size_t off = (size_t)s * kGroups + gi;
      e = launch_stage_batch(b->tile, gi >= 2, gi & 1, (const char *)b->d_stage + off * b->cap * sa,
                             b->d_tstart + off * (b->cap + 1), ngroup[gi], nblk[s][gi], b->stream);
    }
    if (e == hipSuccess)
      e = launch_seam_batch((const char *)b->d_seam + (size_t)s * b->cap * ba, b->d_bstart + (size_t)s * (b->cap + 1),
                            (int)n, nbnd[s], b->stream);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return batch_error(b, CLES_EDEVICE, "batched step capture failed");
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph(b);
    return batch_error(b, CLES_EDEVICE, "hipGraphInstantiate");
  }
  b->dirty = false;
  return CLES_OK;
}

}  // namespace
}  // namespace cles

using namespace cles;

extern "C" {

size_t cles_footprint(const cles_params *p) { return p ? ctx_footprint(p) : 0; }

cles_batch *cles_batch_create(size_t hbm_budget, int *err) {
  cles_batch *b = new cles_batch();
  b->budget = hbm_budget;
  if (hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    if (err) *err = CLES_EDEVICE;
    set_error(nullptr, CLES_EDEVICE, "hipStreamCreate failed");
    return nullptr;
  }
  if (err) *err = CLES_OK;
  return b;
}

void cles_batch_destroy(cles_batch *b) {
  if (!b) return;
  drop_graph(b);
  free_tables(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int cles_batch_add(cles_batch *b, cles_ctx *c) {
  if (!b || !c) return CLES_EINVAL;
  if (hipStreamQuery(b->stream) == hipErrorNotReady) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (!c->have_state) return batch_error(b, CLES_ESTATE, "member has no state");
  for (int q = 0; q < 6; q++)
    if (c->p.bctype[q] != 7) return batch_error(b, CLES_EINVAL, "batch members must be periodic in every direction");
  if (c->p.les || c->has_ib || c->immersed) return batch_error(b, CLES_EINVAL, "batch members run without LES and immersed body");
  if (std::find(b->members.begin(), b->members.end(), c) != b->members.end())
    return batch_error(b, CLES_EINVAL, "context is already a member");
  if (c->p.fluxorder < 2) return batch_error(b, CLES_EINVAL, "batch members run RK2 or RK3 (fluxorder 2 or 3)");
  if (c->p.weno5 || c->p.roe) return batch_error(b, CLES_EINVAL, "batch members use the reference schemes (no weno5 / roe)");
  if (!b->members.empty() && b->members[0]->p.fluxorder != c->p.fluxorder)
    return batch_error(b, CLES_EINVAL, "batch members must share fluxorder");
  const size_t need = ctx_footprint(&c->p);
  if (b->used + need > b->budget) return batch_error(b, CLES_ENOMEM, "member does not fit the HBM budget");
  b->used += need;
  b->members.push_back(c);
  b->dirty = true;
  return CLES_OK;
}

int cles_batch_remove(cles_batch *b, cles_ctx *c) {
  if (!b || !c) return CLES_EINVAL;
  if (hipStreamQuery(b->stream) == hipErrorNotReady) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  auto it = std::find(b->members.begin(), b->members.end(), c);
  if (it == b->members.end()) return batch_error(b, CLES_EINVAL, "context is not a member");
  b->members.erase(it);
  b->used -= ctx_footprint(&c->p);
  b->dirty = true;
  return CLES_OK;
}

int cles_batch_size(const cles_batch *b) { return b ? (int)b->members.size() : CLES_EINVAL; }
size_t cles_batch_used(const cles_batch *b) { return b ? b->used : 0; }

int cles_batch_set_tile(cles_batch *b, int tile) {
  if (!b || tile < 0 || tile > 7) return CLES_EINVAL;
  if (tile != b->tile) {
    b->tile = tile;
    b->dirty = true;
  }
  return CLES_OK;
}

int cles_batch_launch(cles_batch *b, int nsteps) {
  if (!b || nsteps < 0) return CLES_EINVAL;
  if (b->members.empty() || nsteps == 0) return CLES_OK;
  if (b->dirty) {
    int r = rebuild(b);
    if (r) return r;
  }
  for (int s = 0; s < nsteps; s++)
    if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) return batch_error(b, CLES_EDEVICE, "hipGraphLaunch");
  for (cles_ctx *c : b->members) {
    c->ti += nsteps;
    c->ucat_src = -1;
  }
  return CLES_OK;
}

int cles_batch_wait(cles_batch *b) {
  if (!b) return CLES_EINVAL;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_error(b, CLES_EDEVICE, "batch step failed");
  return CLES_OK;
}

int cles_batch_advance(cles_batch *b, int nsteps) {
  int r = cles_batch_launch(b, nsteps);
  if (r) return r;
  return cles_batch_wait(b);
}

int cles_batch_time_step(cles_batch *b, int nrep, double *ms) {
  if (!b || !ms || nrep < 1) return CLES_EINVAL;
  if (b->members.empty()) return batch_error(b, CLES_ESTATE, "empty batch");
  if (b->dirty) {
    int r = rebuild(b);
    if (r) return r;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  for (int q = 0; q < nrep && e == hipSuccess; q++) e = hipGraphLaunch(b->exec, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return batch_error(b, CLES_EDEVICE, "batch timing");
  *ms = (double)t / nrep;
  for (cles_ctx *c : b->members) c->ti += nrep;
  return CLES_OK;
}

const char *cles_batch_last_error(const cles_batch *b) { return b ? b->err : ""; }

}  // extern "C"

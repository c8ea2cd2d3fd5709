// cles_batch.hip -- several independent periodic solves advanced together
// (the serving form of the time loop, solvers.c:223-294, over many grids).
//
// A batch holds existing contexts (fully periodic grids, any size, Mach and
// Reynolds number, with or without the LES model and the shock sensor; no
// immersed body, no statistics).  One batched step is captured once into a
// hipGraph and replayed:
//   [LES model of each LES member] [shock sensor of each -calc_sigma member]
//   stages: launch mode (default) -- one batched stage launch per group and
//           stage plus one batched seam launch per stage, each grid's tiles
//           numbered XCD-aware as in the single-grid launch;
//           persistent mode -- one persistent kernel per (precision, model)
//           group runs every stage of every member of the group, its
//           workgroups pulling tile and seam items from per-XCD queues with
//           per-member completion counters between stages.  Every item
//           releases its stores at agent scope (an L2 write-back on each of
//           the eight XCDs' caches) and every dependent item acquires (an L2
//           invalidation), which costs more than the kernel boundaries it
//           replaces: launch mode is the faster form on gfx950.
//   [the stale-state cells of the LES members advance]
// Membership and member changes mark the batch dirty; the next launch
// rebuilds the tables and the graph.  Without the graph (cles_batch_set_graph
// off) the same kernels are launched step by step and a rebuild only uploads
// the tables.
//
// Buffers: a member's Q^n buffer X stays its current buffer; stage 0 writes
// Y, stage 1 writes Z and the last stage writes X in place (the stage kernel
// reads Qold only at its own node), so the pointers of a step never change.
// The last stage's input (Z, or Y for RK2) is what the next step's LES model
// reads as the state of the last CalcUCat; the graph reads it through a
// device cell, so that the first step of a launch may start from any
// buffer.  The LES model swaps its two Cs buffers each step: the batch keeps
// two graphs (from the current and from the swapped pointers) and alternates
// them.  The arithmetic of every tile is the single-grid kernel's, so a
// member's result does not depend on the other members and equals
// cles_advance's.
//
// Admission is by HBM footprint: a member is accepted when its allocation
// fits into the batch's byte budget.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cles_internal.hpp"
#include "cles_seam.hpp"

// launch counts of one batched step
struct BatchPlan {
  // launch mode: members per group, workgroups per stage and group, seam blocks per stage
  int ngroup[16] = {}, nblk[3][16] = {}, nbnd[3] = {};
  // persistent mode: phases per group
  int nph[16] = {};
};

struct cles_batch {
  hipStream_t stream = nullptr;
  size_t budget = 0, used = 0;
  int tile = 0;
  int mode = CLES_BATCH_LAUNCHES;
  unsigned spin_limit = 1u << 22;
  std::vector<cles_ctx *> members;
  size_t cap = 0;  // members the device tables hold
  // launch mode: [stage][group] stage args + tile starts, [stage] seam args + block starts
  void *d_stage = nullptr, *d_seam = nullptr;
  int *d_tstart = nullptr, *d_bstart = nullptr;
  // persistent mode: [stage][member] stage and seam args, [group][queue][phase]
  void *d_pstage = nullptr;
  cles::SeamArgs *d_pseam = nullptr;
  cles::PersistPhase *d_phase = nullptr;
  int *d_nitems = nullptr;  // [group][queue]
  unsigned *d_ctr = nullptr;  // [group][queue] heads, [member] done, abort; kCtrStride apart
  // LES members: device cells with the stale-state pointer, and the values
  // the graph writes into them after a step
  const double **d_cells = nullptr, **d_cellnext = nullptr;
  std::vector<const double *> h_cells;
  std::vector<char> swaps;  // members whose LES model swaps its Cs buffers each step
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int nexec = 0, next = 0;
  bool dirty = true;
  bool use_graph = true;  // replay a captured step, or launch every step's kernels
  BatchPlan plan;
  bool queued = false;  // steps launched and not yet waited for
  int nstages = 0;
  char err[256] = {0};
};

namespace cles {
namespace {

// (fp32, LES, viscous, sigma = 1 on every face of an LES member: the central
// instance of the stage kernel, in launch mode)
constexpr int kGroups = 16;
constexpr int kQueues = 8;

int batch_error(cles_batch *b, int code, const char *msg) {
  if (b) snprintf(b->err, sizeof(b->err), "%s", msg);
  return set_error(nullptr, code, "%s", msg);
}

// launched steps not yet collected by cles_batch_wait (the batch's own
// record: a stream query can also report work that other threads' device-wide
// calls put in flight)
bool running(const cles_batch *b) { return b->queued; }

void free_tables(cles_batch *b) {
  void *p[] = {b->d_stage, b->d_seam, b->d_tstart, b->d_bstart, b->d_pstage, b->d_pseam,
               b->d_phase, b->d_nitems, b->d_ctr, (void *)b->d_cells, (void *)b->d_cellnext};
  for (void *q : p)
    if (q) (void)hipFree(q);
  b->d_stage = b->d_seam = b->d_pstage = nullptr;
  b->d_tstart = b->d_bstart = b->d_nitems = nullptr;
  b->d_pseam = nullptr;
  b->d_phase = nullptr;
  b->d_ctr = nullptr;
  b->d_cells = b->d_cellnext = nullptr;
  b->cap = 0;
}

void drop_graph(cles_batch *b) {
  for (int q = 0; q < 2; q++) {
    if (b->exec[q]) (void)hipGraphExecDestroy(b->exec[q]);
    if (b->graph[q]) (void)hipGraphDestroy(b->graph[q]);
    b->exec[q] = nullptr;
    b->graph[q] = nullptr;
  }
  b->nexec = 0;
  b->next = 0;
}

unsigned *abort_flag(cles_batch *b) { return b->d_ctr + ((size_t)kGroups * kQueues + b->cap) * kCtrStride; }

bool les_on(const cles_ctx *c) { return c->p.les != 0; }
// flow_step runs the model before the stages (solvers.c:16-30)
bool les_model(const cles_ctx *c) { return c->p.les && !c->p.invicid && c->p.les != 10; }
bool sigma_one(const cles_ctx *c) {
  if (c->sgm) return false;
  for (int d = 0; d < 3; d++)
    if (c->sig.in[d] != 1.0 || c->sig.bnd[d] != 1.0) return false;
  return true;
}
int group_of(const cles_ctx *c) {
  return (c->p.precision == CLES_PREC_FP32BULK ? 4 : 0) + (les_on(c) ? 2 : 0) + (c->p.invicid ? 0 : 1) +
         (les_on(c) && sigma_one(c) ? 8 : 0);
}
int member_tile(const cles_batch *b, const cles_ctx *c) { return les_on(c) ? 7 : b->tile; }
int last_input(const cles_batch *b, const cles_ctx *c) { return (c->cur + (b->nstages == 3 ? 2 : 1)) % 3; }

// why c cannot be a member, or null
const char *ineligible(const cles_ctx *c) {
  if (!c->have_state) return "member has no state";
  for (int q = 0; q < 6; q++)
    if (c->p.bctype[q] != 7) return "batch members must be periodic in every direction";
  if (c->has_ib || c->immersed) return "batch members run without an immersed body";
  if (c->p.fluxorder < 2) return "batch members run RK2 or RK3 (fluxorder 2 or 3)";
  if (c->p.weno5 || c->p.roe) return "batch members use the reference schemes (no weno5 / roe)";
  if (c->p.averaging) return "batch members run without statistics (averaging)";
  // les.c:139: only a fresh run with tistart = -1 takes the Cs = 0 branch on
  // its first steps, a branch a captured step cannot follow
  if (les_model(c) && c->p.tistart + 1 == 0 && !c->p.rstart)
    return "LES batch members need tistart >= 0 or a restart";
  return nullptr;
}

__global__ void zero_kernel(unsigned *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

__global__ void cells_kernel(const double **cells, const double *const *next, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) cells[t] = next[t];
}

int grow(cles_batch *b, size_t n) {
  if (n <= b->cap) return CLES_OK;
  free_tables(b);
  const size_t cap = std::max<size_t>(n, 2 * b->cap);
  const size_t sa = stage_args_bytes(), ba = seam_args_bytes();
  bool ok = hipMalloc(&b->d_stage, 3 * kGroups * cap * sa) == hipSuccess &&
            hipMalloc(&b->d_seam, 3 * cap * ba) == hipSuccess &&
            hipMalloc((void **)&b->d_tstart, 3 * kGroups * (cap + 1) * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&b->d_bstart, 3 * (cap + 1) * sizeof(int)) == hipSuccess &&
            hipMalloc(&b->d_pstage, 3 * cap * sa) == hipSuccess &&
            hipMalloc((void **)&b->d_pseam, 3 * cap * sizeof(SeamArgs)) == hipSuccess &&
            hipMalloc((void **)&b->d_phase, (size_t)kGroups * kQueues * 3 * cap * sizeof(PersistPhase)) == hipSuccess &&
            hipMalloc((void **)&b->d_nitems, kGroups * kQueues * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&b->d_ctr, (kGroups * kQueues + cap + 1) * kCtrStride * sizeof(unsigned)) == hipSuccess &&
            hipMalloc((void **)&b->d_cells, cap * sizeof(double *)) == hipSuccess &&
            hipMalloc((void **)&b->d_cellnext, cap * sizeof(double *)) == hipSuccess;
  if (!ok) {
    free_tables(b);
    return batch_error(b, CLES_ENOMEM, "batch argument tables");
  }
  b->cap = cap;
  return CLES_OK;
}

// stage s of a member whose Q^n buffer is X: (input, output) buffers
void stage_buffers(int nst, int X, int s, int &in, int &out) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  if (nst == 3) {
    in = s == 0 ? X : s == 1 ? Y : Z;
    out = s == 0 ? Y : s == 1 ? Z : X;
  } else {
    in = s == 0 ? X : Y;
    out = s == 0 ? Y : X;
  }
}


// the launch-mode tables
int build_launch_tables(cles_batch *b, BatchPlan &pl) {
  const size_t n = b->members.size(), sa = stage_args_bytes(), ba = seam_args_bytes();
  const int nst = b->nstages;
  std::vector<char> hstage(3 * kGroups * b->cap * sa), hseam(3 * b->cap * ba);
  std::vector<int> htstart(3 * kGroups * (b->cap + 1), 0), hbstart(3 * (b->cap + 1), 0);
  for (int s = 0; s < nst; s++) {
    int cnt[kGroups] = {};
    int r = 0;
    for (cles_ctx *c : b->members) {
      int in, out;
      stage_buffers(nst, c->cur, s, in, out);
      double coef[3];
      rk_coef(c->p, s, coef);
      const int gi = group_of(c);
      const size_t slot = ((size_t)s * kGroups + gi) * b->cap + cnt[gi];
      const int wgs = stage_batch_args(c, in, out, s, coef, member_tile(b, c), hstage.data() + slot * sa);
      int *ts = htstart.data() + ((size_t)s * kGroups + gi) * (b->cap + 1);
      ts[cnt[gi] + 1] = ts[cnt[gi]] + wgs;
      cnt[gi]++;
      const int blocks = seam_batch_args(c, in, out, hseam.data() + ((size_t)s * b->cap + r) * ba);
      hbstart[(size_t)s * (b->cap + 1) + r + 1] = hbstart[(size_t)s * (b->cap + 1) + r] + blocks;
      r++;
    }
    for (int gi = 0; gi < kGroups; gi++) {
      pl.ngroup[gi] = cnt[gi];
      pl.nblk[s][gi] = htstart[((size_t)s * kGroups + gi) * (b->cap + 1) + cnt[gi]];
    }
    pl.nbnd[s] = hbstart[(size_t)s * (b->cap + 1) + n];
  }
  if (hipMemcpy(b->d_stage, hstage.data(), hstage.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_seam, hseam.data(), hseam.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_tstart, htstart.data(), htstart.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_bstart, hbstart.data(), hbstart.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "batch table upload");
  return CLES_OK;
}

// tiles per item: large members take several tiles per queue item, so that
// the per-item synchronisation stays small beside the item's work
int grain_of(int tiles) { return tiles >= 16384 ? 4 : tiles >= 4096 ? 2 : 1; }

// the persistent-mode tables: per group, per queue, the phases in stage
// order; a member's items of one stage are split into eight contiguous
// shares, one per queue (the tiles an XCD takes share their halos in its L2)
int build_persist_tables(cles_batch *b, BatchPlan &pl) {
  const size_t n = b->members.size(), sa = stage_args_bytes();
  const int nst = b->nstages;
  std::vector<char> hstage(3 * b->cap * sa);
  std::vector<SeamArgs> hseam(3 * b->cap);
  std::vector<PersistPhase> hph((size_t)kGroups * kQueues * 3 * b->cap);
  std::vector<int> hn(kGroups * kQueues, 0);
  std::vector<int> tiles(n), titems(n), sitems(n), grain(n);
  for (size_t r = 0; r < n; r++) {
    cles_ctx *c = b->members[r];
    tiles[r] = stage_tiles(c, member_tile(b, c));
    grain[r] = grain_of(tiles[r]);
    titems[r] = 0;
    const long long nbn = bnd_count(c->g);
    sitems[r] = (int)((nbn + kSeamItem - 1) / kSeamItem);
    for (int s = 0; s < nst; s++) {
      int in, out;
      stage_buffers(nst, c->cur, s, in, out);
      double coef[3];
      rk_coef(c->p, s, coef);
      (void)stage_batch_args(c, in, out, s, coef, member_tile(b, c), hstage.data() + ((size_t)s * b->cap + r) * sa);
      SeamArgs &sm = hseam[(size_t)s * b->cap + r];
      sm.g = c->g;
      sm.q = c->Q[in];
      sm.qo = c->Q[out];
      sm.po = c->P[out];
      sm.gamma = c->p.gamma;
      sm.nb = nbn;
    }
  }
  for (int gi = 0; gi < kGroups; gi++) {
    int nph = 0;
    for (int s = 0; s < nst; s++)
      for (size_t r = 0; r < n; r++)
        if (group_of(b->members[r]) == gi) nph++;
    pl.nph[gi] = nph;
    if (!nph) continue;
    for (int x = 0; x < kQueues; x++) {
      PersistPhase *ph = hph.data() + ((size_t)gi * kQueues + x) * 3 * b->cap;
      int p = 0, items = 0;
      for (int s = 0; s < nst; s++)
        for (size_t r = 0; r < n; r++) {
          if (group_of(b->members[r]) != gi) continue;
          const int T = tiles[r], per = (T + kQueues - 1) / kQueues;
          const int t0 = std::min(T, x * per), t1 = std::min(T, (x + 1) * per);
          const int S = sitems[r], sper = (S + kQueues - 1) / kQueues;
          const int b0 = std::min(S, x * sper), b1 = std::min(S, (x + 1) * sper);
          // items of the member per stage, over all queues
          int per_stage = 0;
          for (int y = 0; y < kQueues; y++) {
            const int u0 = std::min(T, y * per), u1 = std::min(T, (y + 1) * per);
            per_stage += (u1 - u0 + grain[r] - 1) / grain[r];
          }
          per_stage += S;
          PersistPhase &q = ph[p++];
          q.member = (int)r;
          q.stage = s;
          q.t0 = t0;
          q.nt = t1 - t0;
          q.grain = grain[r];
          q.b0 = b0;
          q.nb = b1 - b0;
          q.need = (unsigned)(s * per_stage);
          q.istart = items;
          items += (q.nt + q.grain - 1) / q.grain + q.nb;
        }
      hn[gi * kQueues + x] = items;
    }
  }
  if (hipMemcpy(b->d_pstage, hstage.data(), hstage.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_pseam, hseam.data(), hseam.size() * sizeof(SeamArgs), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_phase, hph.data(), hph.size() * sizeof(PersistPhase), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_nitems, hn.data(), hn.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "batch table upload");
  return CLES_OK;
}

// one batched step on b->stream (under capture)
hipError_t record_step(cles_batch *b, const BatchPlan &pl) {
  const size_t n = b->members.size();
  const int nst = b->nstages;
  hipError_t e = hipSuccess;
  bool any_les = false;
  for (size_t r = 0; r < n && e == hipSuccess; r++) {
    cles_ctx *c = b->members[r];
    if (!les_model(c)) continue;
    any_les = true;
    if (launch_les_on(c, b->stream, b->d_cells + r)) e = hipErrorUnknown;
    c->ucat_les_mixed = 0;  // the stages of this step evaluate the right-hand side
  }
  for (size_t r = 0; r < n && e == hipSuccess; r++) {
    cles_ctx *c = b->members[r];
    if (c->sgm && launch_sigma_on(c, c->cur, b->stream)) e = hipErrorUnknown;  // rhs.c:3327, first stage's input
  }
  if (e != hipSuccess) return e;
  if (b->mode == CLES_BATCH_PERSISTENT) {
    const int nctr = (int)((kGroups * kQueues + n) * kCtrStride);
    hipLaunchKernelGGL(zero_kernel, dim3((nctr + 255) / 256), dim3(256), 0, b->stream, b->d_ctr, nctr);
    e = hipGetLastError();
    for (int gi = 0; gi < kGroups && e == hipSuccess; gi++) {
      if (!pl.nph[gi]) continue;
      PersistArgs a;
      a.stage = b->d_pstage;
      a.seam = b->d_pseam;
      a.phase = b->d_phase + (size_t)gi * kQueues * 3 * b->cap;
      a.nitems = b->d_nitems + gi * kQueues;
      a.nph = pl.nph[gi];
      a.pstride = (int)(3 * b->cap);
      a.cap = (int)b->cap;
      a.head = b->d_ctr + (size_t)gi * kQueues * kCtrStride;
      a.done = b->d_ctr + (size_t)kGroups * kQueues * kCtrStride;
      a.abort = abort_flag(b);
      a.spin_limit = b->spin_limit;
      e = launch_stage_persist(b->tile, (gi >> 2) & 1, gi & 1, (gi >> 1) & 1, a, b->stream);
    }
  } else {
    const size_t sa = stage_args_bytes(), ba = seam_args_bytes();
    for (int s = 0; s < nst && e == hipSuccess; s++) {
      for (int gi = 0; gi < kGroups && e == hipSuccess; gi++) {
        if (!pl.ngroup[gi]) continue;
        const size_t off = (size_t)s * kGroups + gi;
        e = launch_stage_batch(b->tile, (gi >> 2) & 1, gi & 1, (gi >> 1) & 1, gi >> 3,
                               (const char *)b->d_stage + off * b->cap * sa,
                               b->d_tstart + off * (b->cap + 1), pl.ngroup[gi], pl.nblk[s][gi], b->stream);
      }
      if (e == hipSuccess)
        e = launch_seam_batch((const char *)b->d_seam + (size_t)s * b->cap * ba, b->d_bstart + (size_t)s * (b->cap + 1),
                              (int)n, pl.nbnd[s], b->stream);
    }
  }
  if (e == hipSuccess && any_les) {
    hipLaunchKernelGGL(cells_kernel, dim3(((int)n + 255) / 256), dim3(256), 0, b->stream, b->d_cells,
                       (const double *const *)b->d_cellnext, (int)n);
    e = hipGetLastError();
  }
  return e;
}

int capture(cles_batch *b, const BatchPlan &pl, int q) {
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "hipStreamBeginCapture");
  const hipError_t e = record_step(b, pl);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return batch_error(b, CLES_EDEVICE, "batched step capture failed");
  }
  b->graph[q] = g;
  if (hipGraphInstantiate(&b->exec[q], g, nullptr, nullptr, 0) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "hipGraphInstantiate");
  return CLES_OK;
}

// rebuild the tables and capture the batched step
int rebuild(cles_batch *b) {
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_error(b, CLES_EDEVICE, "batch step failed");
  drop_graph(b);
  const size_t n = b->members.size();
  if (n == 0) {
    b->dirty = false;
    return CLES_OK;
  }
  const int nst = b->members[0]->p.fluxorder;
  for (cles_ctx *c : b->members) {
    if (const char *why = ineligible(c)) return batch_error(b, CLES_EINVAL, why);
    if (c->p.fluxorder != nst) return batch_error(b, CLES_EINVAL, "batch members must share fluxorder");
    if (les_model(c) && c->ucat_les_mixed)
      return batch_error(b, CLES_ESTATE, "LES model called twice without a right-hand-side evaluation in between");
  }
  b->nstages = nst;
  int r = grow(b, n);
  if (r) return r;
  BatchPlan &pl = b->plan;
  pl = BatchPlan();
  r = b->mode == CLES_BATCH_PERSISTENT ? build_persist_tables(b, pl) : build_launch_tables(b, pl);
  if (r) return r;
  // stale-state cells: Q of the last CalcUCat now, the last stage's input after a step
  b->h_cells.assign(n, nullptr);
  std::vector<const double *> nxt(n, nullptr);
  for (size_t q = 0; q < n; q++) {
    cles_ctx *c = b->members[q];
    b->h_cells[q] = c->ucat_src >= 0 ? c->Q[c->ucat_src] : nullptr;
    nxt[q] = c->Q[last_input(b, c)];
  }
  if (hipMemcpy((void *)b->d_cells, b->h_cells.data(), n * sizeof(double *), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy((void *)b->d_cellnext, nxt.data(), n * sizeof(double *), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(abort_flag(b), 0, sizeof(unsigned), b->stream) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "batch table upload");
  if (!b->use_graph) {  // every step's kernels are launched by replay()
    b->dirty = false;
    return CLES_OK;
  }
  // the LES model swaps each member's Cs buffers per step: capture the step
  // from the current pointers and, when any member swapped, from the swapped
  // ones (that capture swaps them back)
  std::vector<double *> cs0(n);
  for (size_t q = 0; q < n; q++) cs0[q] = b->members[q]->cs;
  r = capture(b, pl, 0);
  if (r) return r;
  bool swapped = false;
  b->swaps.assign(n, 0);
  for (size_t q = 0; q < n; q++) {
    b->swaps[q] = b->members[q]->cs != cs0[q];
    swapped |= b->swaps[q] != 0;
  }
  b->nexec = 1;
  if (swapped) {
    r = capture(b, pl, 1);
    if (r) return r;
    b->nexec = 2;
  }
  b->next = 0;
  b->dirty = false;
  return CLES_OK;
}

// host bookkeeping of n replayed steps: step counters, the state of the last
// CalcUCat, the Cs buffers
void after_steps(cles_batch *b, int n) {
  for (size_t q = 0; q < b->members.size(); q++) {
    cles_ctx *c = b->members[q];
    c->ti += n;
    if (!c->p.invicid) {
      c->ucat_src = last_input(b, c);
      c->ucat_les_mixed = 0;
      b->h_cells[q] = c->Q[c->ucat_src];
    }
    if (b->nexec == 2 && (n & 1) && b->swaps[q]) std::swap(c->cs, c->cs_alt);
  }
  if (b->nexec == 2) b->next ^= n & 1;
}

int prepare(cles_batch *b) {
  if (b->dirty) {
    int r = rebuild(b);
    if (r) return r;
  }
  return CLES_OK;
}

int replay(cles_batch *b, int n) {
  if (!b->nexec) {  // no graph: launch each step (the LES model swaps its Cs buffers as it goes)
    for (int s = 0; s < n; s++)
      if (record_step(b, b->plan) != hipSuccess) return batch_error(b, CLES_EDEVICE, "batched step launch failed");
    return CLES_OK;
  }
  for (int s = 0; s < n; s++) {
    if (hipGraphLaunch(b->exec[b->nexec == 2 ? (b->next + s) & 1 : 0], b->stream) != hipSuccess)
      return batch_error(b, CLES_EDEVICE, "hipGraphLaunch");
  }
  return CLES_OK;
}

// after the stream drained: did a dependency wait give up?
int check_abort(cles_batch *b) {
  if (b->mode != CLES_BATCH_PERSISTENT || !b->d_ctr || b->members.empty()) return CLES_OK;
  unsigned h = 0;
  if (hipMemcpy(&h, abort_flag(b), sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
    return batch_error(b, CLES_EDEVICE, "batch status");
  if (h) {
    (void)hipMemsetAsync(abort_flag(b), 0, sizeof(unsigned), b->stream);
    (void)hipStreamSynchronize(b->stream);
    return batch_error(b, CLES_EDEVICE, "batched step: a stage dependency wait gave up; member states are undefined");
  }
  return CLES_OK;
}

}  // namespace

int member_guard(cles_ctx *c, bool modifies) {
  if (!c || !c->batch) return CLES_OK;
  cles_batch *b = c->batch;
  if (running(b)) return set_error(c, CLES_ESTATE, "context is a member of a running batch (cles_batch_wait)");
  if (modifies) b->dirty = true;
  return CLES_OK;
}

void member_leave(cles_ctx *c) {
  if (!c || !c->batch) return;
  cles_batch *b = c->batch;
  (void)hipStreamSynchronize(b->stream);
  auto it = std::find(b->members.begin(), b->members.end(), c);
  if (it != b->members.end()) {
    b->members.erase(it);
    b->used -= ctx_footprint(&c->p);
  }
  b->dirty = true;
  c->batch = nullptr;
}

}  // namespace cles

using namespace cles;

extern "C" {

size_t cles_footprint(const cles_params *p) { return p ? ctx_footprint(p) : 0; }

cles_batch *cles_batch_create(size_t hbm_budget, int *err) {
  cles_batch *b = new cles_batch();
  b->budget = hbm_budget;
  // non-blocking: set-up and downloads of other contexts on the null stream
  // do not wait for the batch
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
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
  (void)hipStreamSynchronize(b->stream);
  for (cles_ctx *c : b->members) c->batch = nullptr;
  drop_graph(b);
  free_tables(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int cles_batch_add(cles_batch *b, cles_ctx *c) {
  if (!b || !c) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (c->batch) return batch_error(b, CLES_EINVAL, c->batch == b ? "context is already a member" : "context is a member of another batch");
  if (const char *why = ineligible(c)) return batch_error(b, c->have_state ? CLES_EINVAL : CLES_ESTATE, why);
  if (!b->members.empty() && b->members[0]->p.fluxorder != c->p.fluxorder)
    return batch_error(b, CLES_EINVAL, "batch members must share fluxorder");
  const size_t need = ctx_footprint(&c->p);
  if (b->used + need > b->budget) return batch_error(b, CLES_ENOMEM, "member does not fit the HBM budget");
  b->used += need;
  b->members.push_back(c);
  c->batch = b;
  b->dirty = true;
  return CLES_OK;
}

int cles_batch_remove(cles_batch *b, cles_ctx *c) {
  if (!b || !c) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (c->batch != b) return batch_error(b, CLES_EINVAL, "context is not a member");
  member_leave(c);
  return CLES_OK;
}

int cles_batch_size(const cles_batch *b) { return b ? (int)b->members.size() : CLES_EINVAL; }
size_t cles_batch_used(const cles_batch *b) { return b ? b->used : 0; }

int cles_batch_set_tile(cles_batch *b, int tile) {
  if (!b || tile < 0 || tile > 7) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (tile != b->tile) {
    b->tile = tile;
    b->dirty = true;
  }
  return CLES_OK;
}

int cles_batch_set_mode(cles_batch *b, int mode) {
  if (!b || (mode != CLES_BATCH_PERSISTENT && mode != CLES_BATCH_LAUNCHES)) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (mode != b->mode) {
    b->mode = mode;
    b->dirty = true;
  }
  return CLES_OK;
}

int cles_batch_set_graph(cles_batch *b, int on) {
  if (!b) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if ((on != 0) != b->use_graph) {
    b->use_graph = on != 0;
    b->dirty = true;
  }
  return CLES_OK;
}

int cles_batch_set_spin_limit(cles_batch *b, unsigned polls) {
  if (!b || polls == 0) return CLES_EINVAL;
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  if (polls != b->spin_limit) {
    b->spin_limit = polls;
    b->dirty = true;
  }
  return CLES_OK;
}

int cles_batch_launch(cles_batch *b, int nsteps) {
  if (!b || nsteps < 0) return CLES_EINVAL;
  if (b->members.empty() || nsteps == 0) return CLES_OK;
  int r = prepare(b);
  if (r) return r;
  b->queued = true;
  r = replay(b, nsteps);
  if (r) return r;
  after_steps(b, nsteps);
  return CLES_OK;
}

int cles_batch_wait(cles_batch *b) {
  if (!b) return CLES_EINVAL;
  b->queued = false;
  if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_error(b, CLES_EDEVICE, "batch step failed");
  return check_abort(b);
}

int cles_batch_advance(cles_batch *b, int nsteps) {
  int r = cles_batch_launch(b, nsteps);
  if (r) return r;
  return cles_batch_wait(b);
}

int cles_batch_time_step(cles_batch *b, int nrep, double *ms) {
  if (!b || !ms || nrep < 1) return CLES_EINVAL;
  if (b->members.empty()) return batch_error(b, CLES_ESTATE, "empty batch");
  if (running(b)) return batch_error(b, CLES_ESTATE, "batch is running (cles_batch_wait)");
  int r = prepare(b);
  if (r) return r;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  if (e == hipSuccess && replay(b, nrep)) e = hipErrorUnknown;
  if (e == hipSuccess) {
    after_steps(b, nrep);
    e = hipEventRecord(e1, b->stream);
  }
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return batch_error(b, CLES_EDEVICE, "batch timing");
  *ms = (double)t / nrep;
  return check_abort(b);
}

const char *cles_batch_last_error(const cles_batch *b) { return b ? b->err : ""; }

}  // extern "C"

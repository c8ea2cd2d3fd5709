// cles_api.hip -- the C ABI of include/cles.h: context lifetime, host <->
// device layout conversion, the Runge-Kutta driver and measurement hooks.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <cmath>
#include <vector>

#include "cles_internal.hpp"

namespace cles {

static char g_err[256];

int set_error(cles_ctx *c, int code, const char *fmt, ...) {
  char *dst = c ? c->err : g_err;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 256, fmt, ap);
  va_end(ap);
  return code;
}

namespace {

// ---- layout conversion: interleaved DMDA arrays <-> device planes ----
struct XferArgs {
  Grid g;
  const double *src;
  double *dst;
  int nc;         // components per node
  int local;      // 1: host array is the ghosted local Vec
  int gx, gy, gz; // ghost widths of that local Vec
};

// host interleaved (global or local) -> device planes (owned nodes)
__global__ void unpack_kernel(XferArgs a) {
  const Grid &g = a.g;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (long long)g.mx * g.my * g.mz) return;
  long long s = n;
  if (a.local) {
    const int i = (int)(n % g.mx), j = (int)((n / g.mx) % g.my), k = (int)(n / g.sz);
    const long long LX = g.mx + 2 * a.gx, LY = g.my + 2 * a.gy;
    s = ((long long)(k + a.gz) * LY + (j + a.gy)) * LX + (i + a.gx);
  }
  for (int q = 0; q < a.nc; q++) a.dst[n + q * g.np] = a.src[s * a.nc + q];
}

// device planes -> host interleaved (global, or local with periodic ghosts)
__global__ void pack_kernel(XferArgs a) {
  const Grid &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long LX = g.mx + 2 * a.gx, LY = g.my + 2 * a.gy, LZ = g.mz + 2 * a.gz;
  const long long total = a.local ? LX * LY * LZ : (long long)g.mx * g.my * g.mz;
  if (t >= total) return;
  long long n = t;
  if (a.local) {
    const int i = (int)(t % LX) - a.gx, j = (int)((t / LX) % LY) - a.gy, k = (int)(t / (LX * LY)) - a.gz;
    n = nidw(g, k, j, i);
  }
  for (int q = 0; q < a.nc; q++) a.dst[t * a.nc + q] = a.src[n + q * g.np];
}

__global__ void any_nonzero_kernel(const double *x, long long n, int *flag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && x[t] != 0.0) atomicOr(flag, 1);
}

long long nodes(const cles_ctx *c) { return (long long)c->g.mx * c->g.my * c->g.mz; }
long long local_nodes(const cles_ctx *c) {
  return (long long)(c->g.mx + 4 * c->g.px) * (c->g.my + 4 * c->g.py) * (c->g.mz + 4 * c->g.pz);
}

int upload(cles_ctx *c, const double *host, double *dplanes, int nc, int layout) {
  const long long cnt = (layout == CLES_LAYOUT_LOCAL ? local_nodes(c) : nodes(c)) * nc;
  if ((size_t)cnt > c->xfer_elems) return set_error(c, CLES_EINVAL, "transfer larger than staging buffer");
  CLES_HIP(c, hipMemcpyAsync(c->xfer, host, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  XferArgs a{c->g, c->xfer, dplanes, nc, layout == CLES_LAYOUT_LOCAL, 2 * c->g.px, 2 * c->g.py, 2 * c->g.pz};
  const int bs = 256;
  hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((nodes(c) + bs - 1) / bs)), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int download(cles_ctx *c, const double *dplanes, double *host, int nc, int layout) {
  const long long tot = layout == CLES_LAYOUT_LOCAL ? local_nodes(c) : nodes(c);
  if ((size_t)(tot * nc) > c->xfer_elems) return set_error(c, CLES_EINVAL, "transfer larger than staging buffer");
  XferArgs a{c->g, dplanes, c->xfer, nc, layout == CLES_LAYOUT_LOCAL, 2 * c->g.px, 2 * c->g.py, 2 * c->g.pz};
  const int bs = 256;
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((tot + bs - 1) / bs)), dim3(bs), 0, c->stream, a);
  CLES_HIP(c, hipGetLastError());
  CLES_HIP(c, hipMemcpyAsync(host, c->xfer, tot * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int dalloc(cles_ctx *c, double **p, size_t elems) {
  hipError_t e = hipMalloc((void **)p, elems * sizeof(double));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(c, CLES_ENOMEM, "hipMalloc of %zu bytes failed: %s", elems * sizeof(double), hipGetErrorString(e));
  }
  c->bytes += elems * sizeof(double);
  CLES_HIP(c, hipMemsetAsync(*p, 0, elems * sizeof(double), c->stream));
  return CLES_OK;
}

int free_buffer(const cles_ctx *c, int a, int b) {
  for (int q = 0; q < 3; q++)
    if (q != a && q != b) return q;
  (void)c;
  return -1;
}

int check_params(const cles_params *p) {
  if (p->mx < 5 || p->my < 5 || p->mz < 5) return set_error(nullptr, CLES_EINVAL, "grid must have at least 5 nodes per direction");
  // the stage kernel addresses a plane with 32-bit byte offsets
  if ((long long)p->mx * p->my * p->mz > (1LL << 29) - 64)
    return set_error(nullptr, CLES_EINVAL, "grid has more than 2^29 - 64 nodes (a plane must stay below 4 GB)");
  for (int d = 0; d < 3; d++)
    if ((p->bctype[2 * d] == 7) != (p->bctype[2 * d + 1] == 7))
      return set_error(nullptr, CLES_EINVAL, "periodic direction %d needs bctype 7 on both sides", d);
  for (int q = 0; q < 6; q++)
    if (!supported_bc(q, p->bctype[q]))
      return set_error(nullptr, CLES_EINVAL, "bctype[%d]=%d is not on this path", q, p->bctype[q]);
  if (p->fluxorder < 1 || p->fluxorder > 3) return set_error(nullptr, CLES_EINVAL, "fluxorder must be 1, 2 or 3");
  if (p->eno_order != 2 && p->eno_order != 3) return set_error(nullptr, CLES_EINVAL, "ENOorder must be 2 or 3");
  if (p->twoD < 0 || p->twoD > 3) return set_error(nullptr, CLES_EINVAL, "-twoD must be 0, 1, 2 or 3");
  if (p->precision != CLES_PREC_FP64 && p->precision != CLES_PREC_FP32BULK)
    return set_error(nullptr, CLES_EINVAL, "unknown precision %d", p->precision);
  if (p->les && p->invicid) return set_error(nullptr, CLES_EINVAL, "LES requires a viscous run (-inv 0)");
  if ((p->weno5 || p->roe) && p->precision != CLES_PREC_FP64)
    return set_error(nullptr, CLES_EINVAL, "the WENO5 and Roe schemes run in the FP64 form only");
  return CLES_OK;
}

// LES filter widths, formed once per grid on the host with the C library's
// pow, the function the reference calls (rhs.c:1693, les.c:396-397, les.c:1517):
// fil2 = pow(pow(1/aj, 1/3), 2) and the squared test-filter width.  The
// squares are the exactly rounded products, as in the device kernels.
int filter_widths(cles_ctx *c) {
  const long long N = (long long)c->g.mx * c->g.my * c->g.mz;
  std::vector<double> aj((size_t)N), f2((size_t)N), t2((size_t)N), iv((size_t)N);
  int r = download(c, c->aj, aj.data(), 1, CLES_LAYOUT_GLOBAL);
  if (r) return r;
  const double c5 = pow(5.0, 1. / 3.);
  for (long long n = 0; n < N; n++) {
    const double filter = pow(1. / aj[(size_t)n], 1. / 3.);
    const double test_filter = c->p.testfilter_ik ? c5 * filter : 2.0 * filter;
    f2[(size_t)n] = filter * filter;
    t2[(size_t)n] = test_filter * test_filter;
    iv[(size_t)n] = 1. / aj[(size_t)n];  // cell volume, the Cs smoothing weight (les.c:1301)
  }
  if ((r = upload(c, f2.data(), c->fil2, 1, CLES_LAYOUT_GLOBAL))) return r;
  if (c->ivol && (r = upload(c, iv.data(), c->ivol, 1, CLES_LAYOUT_GLOBAL))) return r;
  return upload(c, t2.data(), c->tfil2, 1, CLES_LAYOUT_GLOBAL);
}

void sigma_spec(cles_ctx *c) {
  const cles_params &p = c->p;
  c->sig.bnd[0] = p.s_x;
  c->sig.bnd[1] = p.s_y;
  c->sig.bnd[2] = p.s_z;
  const double m = p.s_z < p.s_y ? p.s_z : p.s_y;  // PetscMin(sigma_z, sigma_y)
  c->sig.in[0] = p.s_x;
  c->sig.in[1] = m;
  c->sig.in[2] = m;
}

int rk_coefficients(const cles_params &p, int istage, double coef[3], double *alfa_out) {
  double alfa[3] = {1., 1 / 4., 2 / 3.}, beta[3] = {1., 3 / 4., 1 / 3.};
  if (p.fluxorder < 3) {
    alfa[1] = 1 / 2.;
    beta[1] = 1 / 2.;
  }
  coef[0] = alfa[istage] * p.dt * p.st;
  coef[1] = beta[istage];
  coef[2] = alfa[istage];
  *alfa_out = alfa[istage];
  return 0;
}

__global__ void zero_norms_kernel(unsigned long long *x) { x[threadIdx.x] = 0ull; }

// the stage kernels' max slots start from zero (a fill kernel: no memset node
// when a caller captures the stream)
int reset_norms(cles_ctx *c) {
  hipLaunchKernelGGL(zero_norms_kernel, dim3(1), dim3(128), 0, c->stream, c->dnorm);
  CLES_HIP(c, hipGetLastError());
  return CLES_OK;
}

}  // namespace

int rk_coef(const cles_params &p, int istage, double coef[3]) {
  double alfa;
  return rk_coefficients(p, istage, coef, &alfa);
}

int get_coor(cles_ctx *c, double *xyz) { return download(c, c->coor, xyz, 3, CLES_LAYOUT_GLOBAL); }

// device bytes cles_create allocates for p (the allocation list of cles_create)
// doubles that hold n values of the LES model's fields: double in the FP64
// form, float in the FP32-bulk form
size_t les_field_doubles(const cles_params *p, size_t n) {
  return p->precision == CLES_PREC_FP32BULK ? (n + 1) / 2 : n;
}

// sums of the homogeneous LM/MM averages: 5 per (k, j) line, 4 per j plane, 2
size_t homo_doubles(const cles_params *p) { return 5 * (size_t)p->my * p->mz + 4 * (size_t)p->my + 2; }

size_t ctx_footprint(const cles_params *p) {
  const long long N = (long long)p->mx * p->my * p->mz;
  const size_t np = (size_t)((N + 63) / 64 * 64);
  size_t planes = 3 + 9 + 1 + 1 + 1 + 15 + 3 + 5;  // coor, metrics, aj, nvert, fil2, Q x3, P x3, rhs
  size_t extra = 0;  // doubles outside whole planes
  if (p->les) {
    planes += 1 + 3 + 10;  // tfil2, cs, cs_alt, nut, packed metrics
    if (p->calc_sigma) planes += 6;  // shock sensor and its smoothing buffer
    if (p->les != 1 && p->les != 10) {
      planes += 1;  // 1/aj
      extra += les_field_doubles(p, 8 * np) + 2 * les_field_doubles(p, np);  // strain rates, |S| and |S| of nu_t, LM, MM
      if (p->i_homo_filter) extra += homo_doubles(p);
    }
  }
  if (p->averaging) planes += 12;  // Do_averaging sums
  const size_t xfer = (size_t)((p->mx + 4) * (long long)(p->my + 4) * (p->mz + 4)) * 5;
  return (planes * np + extra + xfer) * sizeof(double);
}
}  // namespace cles

using namespace cles;

extern "C" {

void cles_default_params(cles_params *p, int mx, int my, int mz) {
  memset(p, 0, sizeof(*p));
  p->mx = mx; p->my = my; p->mz = mz;
  for (int q = 0; q < 6; q++) p->bctype[q] = 7;
  p->ren = 3000.; p->Ma = 2.81; p->Pr = 0.72; p->gamma = 1.4; p->dt = 0.01; p->st = 1.;
  p->Tref = 273.15; p->max_cs = 0.05; p->Twall = 1.;
  p->s_x = p->s_y = p->s_z = 1.0;
  p->sigma_thrsh = 0.00001;
  p->fluxorder = 3; p->eno_order = 2; p->central = 2;
  p->precision = CLES_PREC_FP64;
}

cles_ctx *cles_create(const cles_params *p, const double *coor, int *err) {
  int e = check_params(p);
  if (e) {
    if (err) *err = e;
    return nullptr;
  }
  cles_ctx *c = (cles_ctx *)calloc(1, sizeof(cles_ctx));
  if (!c) {
    if (err) *err = CLES_ENOMEM;
    return nullptr;
  }
  c->p = *p;
  Grid &g = c->g;
  g.mx = p->mx; g.my = p->my; g.mz = p->mz;
  g.px = p->bctype[0] == 7; g.py = p->bctype[2] == 7; g.pz = p->bctype[4] == 7;
  g.sy = g.mx;
  g.sz = (long long)g.mx * g.my;
  const long long N = (long long)g.mx * g.my * g.mz;
  g.np = (N + 63) / 64 * 64;
  c->ucat_src = -1;
  c->tile = 7;  // 8x4x3 in 128-thread workgroups: fastest measured on the LES and viscous forms
  c->ti = p->tistart;
  sigma_spec(c);
#define TRY(x)                          \
  do {                                  \
    int r_ = (x);                       \
    if (r_) {                           \
      if (err) *err = r_;               \
      fprintf(stderr, "cles_create: %s\n", c->err); \
      cles_destroy(c);                  \
      return nullptr;                   \
    }                                   \
  } while (0)
  TRY(hipGetDevice(&c->dev) == hipSuccess ? 0 : set_error(c, CLES_EDEVICE, "no HIP device"));
  // non-blocking: a context's work neither waits for nor holds up the null
  // stream, other contexts or a batch being captured on another thread
  TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess
          ? 0
          : set_error(c, CLES_EDEVICE, "hipStreamCreate failed"));
  const size_t np = (size_t)g.np;
  TRY(dalloc(c, &c->coor, 3 * np));
  TRY(dalloc(c, &c->csi, 3 * np));
  TRY(dalloc(c, &c->eta, 3 * np));
  TRY(dalloc(c, &c->zet, 3 * np));
  TRY(dalloc(c, &c->aj, np));
  TRY(dalloc(c, &c->nvert, np));
  TRY(dalloc(c, &c->fil2, np));
  if (p->les) TRY(dalloc(c, &c->mpack, 10 * np));
  for (int b = 0; b < 3; b++) {
    TRY(dalloc(c, &c->Q[b], 5 * np));
    TRY(dalloc(c, &c->P[b], np));
  }
  TRY(dalloc(c, &c->rhs, 5 * np));
  c->xfer_elems = (size_t)((g.mx + 4) * (long long)(g.my + 4) * (g.mz + 4)) * 5;
  TRY(dalloc(c, &c->xfer, c->xfer_elems));
  if (p->les) {
    TRY(dalloc(c, &c->tfil2, np));
    TRY(dalloc(c, &c->cs, np));
    TRY(dalloc(c, &c->cs_alt, np));
    TRY(dalloc(c, &c->nut, np));
    if (p->les != 1 && p->les != 10) {
      // the model's own fields in the precision of the form (les_field_doubles)
      TRY(dalloc(c, &c->grad, les_field_doubles(p, 8 * np)));
      TRY(dalloc(c, &c->lm, les_field_doubles(p, np)));
      TRY(dalloc(c, &c->mm, les_field_doubles(p, np)));
      TRY(dalloc(c, &c->ivol, np));
      if (p->i_homo_filter) TRY(dalloc(c, &c->homo, homo_doubles(p)));
    }
    if (p->calc_sigma) {
      TRY(dalloc(c, &c->sgm, 3 * np));
      TRY(dalloc(c, &c->sgm_tmp, 3 * np));
    }
  }
  if (p->averaging) TRY(dalloc(c, &c->avg, 12 * np));
  {
    void *dn = nullptr;
    TRY(hipMalloc(&dn, 128 * sizeof(unsigned long long)) == hipSuccess ? 0 : set_error(c, CLES_ENOMEM, "norm buffer"));
    c->dnorm = (unsigned long long *)dn;
    TRY(hipMemsetAsync(dn, 0, 128 * sizeof(unsigned long long), c->stream) == hipSuccess ? 0 : set_error(c, CLES_EDEVICE, "memset"));
  }
  TRY(upload(c, coor, c->coor, 3, CLES_LAYOUT_GLOBAL));
  TRY(launch_metrics(c));
  TRY(hipStreamSynchronize(c->stream) == hipSuccess ? 0 : set_error(c, CLES_EDEVICE, "metrics"));
  if (p->les) TRY(filter_widths(c));
#undef TRY
  if (err) *err = CLES_OK;
  return c;
}

void cles_destroy(cles_ctx *c) {
  if (!c) return;
  member_leave(c);  // waits for its batch, which then rebuilds without it
  double *bufs[] = {c->coor, c->csi, c->eta, c->zet, c->aj, c->nvert, c->Q[0], c->Q[1], c->Q[2], c->P[0],
                    c->P[1], c->P[2], c->rhs, c->xfer, c->cs, c->cs_alt, c->nut, c->grad, c->lm, c->mm,
                    c->fil2, c->qsnap, c->tfil2, c->ivol, c->avg, c->sgm, c->sgm_tmp, c->homo, c->mpack};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  ibm_free(c);
  if (c->dnorm) (void)hipFree(c->dnorm);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  free(c);
}

int cles_set_nvert(cles_ctx *c, const double *nvert) {
  if (!c || !nvert) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  int r = upload(c, nvert, c->nvert, 1, CLES_LAYOUT_GLOBAL);
  if (r) return r;
  int *flag = nullptr;
  CLES_HIP(c, hipMalloc((void **)&flag, sizeof(int)));
  CLES_HIP(c, hipMemsetAsync(flag, 0, sizeof(int), c->stream));
  const long long N = nodes(c);
  hipLaunchKernelGGL(any_nonzero_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, c->nvert, N, flag);
  int h = 0;
  CLES_HIP(c, hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  (void)hipFree(flag);
  c->has_ib = h;
  return CLES_OK;
}

int cles_get_nvert(cles_ctx *c, double *nvert) {
  if (!c || !nvert) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  return download(c, c->nvert, nvert, 1, CLES_LAYOUT_GLOBAL);
}

// Q (and P, or EqOfState when P is null) into a free state buffer, which
// becomes Q and Qold
static int set_state(cles_ctx *c, const double *Q, const double *P, int layout) {
  if (!c || !Q) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  // keep the buffer that holds the last CalcUCat input intact
  const int b = c->ucat_src == c->cur ? free_buffer(c, c->cur, -1) : c->cur;
  int r = upload(c, Q, c->Q[b], 5, layout);
  if (r) return r;
  c->cur = b;
  r = P ? upload(c, P, c->P[b], 1, layout) : launch_eos_all(c, b);
  if (r) return r;
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  if (c->p.precision == CLES_PREC_FP32BULK) {
    // reference state of the split fluxes: node (1,1,1) of this state
    const long long n = nid(c->g, 1, 1, 1);
    CLES_HIP(c, hipMemcpyAsync(&c->ref_rho, c->Q[b] + n, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CLES_HIP(c, hipMemcpyAsync(&c->ref_p, c->P[b] + n, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CLES_HIP(c, hipStreamSynchronize(c->stream));
    if (!(c->ref_rho > 0.0)) return set_error(c, CLES_EINVAL, "non-positive density at node (1,1,1)");
  }
  c->have_state = 1;
  return CLES_OK;
}

int cles_set_state(cles_ctx *c, const double *Q, int layout) { return set_state(c, Q, nullptr, layout); }

int cles_set_state_qp(cles_ctx *c, const double *Q, const double *P, int layout) {
  if (!P) return CLES_EINVAL;
  return set_state(c, Q, P, layout);
}

int cles_get_state(cles_ctx *c, double *Q, double *P, int layout) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  if (Q) {
    int r = download(c, c->Q[c->cur], Q, 5, layout);
    if (r) return r;
  }
  if (P) {
    int r = download(c, c->P[c->cur], P, 1, layout);
    if (r) return r;
  }
  return CLES_OK;
}

int cles_get_metrics(cles_ctx *c, double *csi, double *eta, double *zet, double *aj) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  int r = 0;
  if (csi && (r = download(c, c->csi, csi, 3, CLES_LAYOUT_GLOBAL))) return r;
  if (eta && (r = download(c, c->eta, eta, 3, CLES_LAYOUT_GLOBAL))) return r;
  if (zet && (r = download(c, c->zet, zet, 3, CLES_LAYOUT_GLOBAL))) return r;
  if (aj && (r = download(c, c->aj, aj, 1, CLES_LAYOUT_GLOBAL))) return r;
  return CLES_OK;
}

int cles_get_averages(cles_ctx *c, double *q_sum, double *p_sum, double *u2_sum, double *ucross_sum) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  if (!c->avg) return set_error(c, CLES_ESTATE, "averaging is off (cles_params.averaging = 0)");
  const long long np = c->g.np;
  int r;
  if (q_sum && (r = download(c, c->avg, q_sum, 5, CLES_LAYOUT_GLOBAL))) return r;
  if (p_sum && (r = download(c, c->avg + 5 * np, p_sum, 1, CLES_LAYOUT_GLOBAL))) return r;
  if (u2_sum && (r = download(c, c->avg + 6 * np, u2_sum, 3, CLES_LAYOUT_GLOBAL))) return r;
  if (ucross_sum && (r = download(c, c->avg + 9 * np, ucross_sum, 3, CLES_LAYOUT_GLOBAL))) return r;
  return CLES_OK;
}

int cles_get_les(cles_ctx *c, double *cs, double *nu_t, double *sigma) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  const long long N = nodes(c);
  int r = 0;
  if (cs) {
    if (c->cs) {
      if ((r = download(c, c->cs, cs, 1, CLES_LAYOUT_GLOBAL))) return r;
    } else {
      memset(cs, 0, N * sizeof(double));
    }
  }
  if (nu_t) {
    if (c->nut) {
      if ((r = download(c, c->nut, nu_t, 1, CLES_LAYOUT_GLOBAL))) return r;
    } else {
      memset(nu_t, 0, N * sizeof(double));
    }
  }
  if (sigma && c->sgm) {
    if ((r = download(c, c->sgm, sigma, 3, CLES_LAYOUT_GLOBAL))) return r;
  } else if (sigma) {
    for (int k = 0; k < c->g.mz; k++)
      for (int j = 0; j < c->g.my; j++)
        for (int i = 0; i < c->g.mx; i++) {
          const long long n = nid(c->g, k, j, i);
          const bool in = i >= 1 && i <= c->g.mx - 2 && j >= 1 && j <= c->g.my - 2 && k >= 1 && k <= c->g.mz - 2;
          for (int d = 0; d < 3; d++) sigma[3 * n + d] = in ? c->sig.in[d] : c->sig.bnd[d];
        }
  }
  return CLES_OK;
}

int cles_eq_of_state(cles_ctx *c) {
  if (!c || !c->have_state) return CLES_ESTATE;
  if (int g_ = member_guard(c, true)) return g_;
  int r = launch_eos_all(c, c->cur);
  if (r) return r;
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int cles_form_rhs(cles_ctx *c, int istage, double *rhs) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  const double coef[3] = {0, 0, 0};
  CLES_HIP(c, hipMemsetAsync(c->rhs, 0, 5 * c->g.np * sizeof(double), c->stream));
  int r = reset_norms(c);
  if (r) return r;
  if (c->sgm && istage == 0 && (r = launch_sigma(c, c->cur))) return r;  // formConvectionLES, rhs.c:3327
  r = launch_stage(c, c->cur, -1, istage, 0, coef);
  if (r) return r;
  if (!c->p.invicid) {
    c->ucat_src = c->cur;  // CalcUCat (rhs.c:4718)
    c->ucat_les_mixed = 0;
  }
  if (rhs) {
    r = download(c, c->rhs, rhs, 5, CLES_LAYOUT_GLOBAL);
    if (r) return r;
  }
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

static int runge_kutta_async(cles_ctx *c) {
  const int nst = c->p.fluxorder;
  int in = c->cur;
  for (int s = 0; s < nst; s++) {
    const int out = free_buffer(c, in, c->cur);
    double coef[3], alfa;
    rk_coefficients(c->p, s, coef, &alfa);
    int r = reset_norms(c);
    if (r) return r;
    if (c->sgm && s == 0 && (r = launch_sigma(c, in))) return r;  // formConvectionLES, rhs.c:3327
    r = launch_stage(c, in, out, s, 1, coef);
    if (r) return r;
    r = launch_boundary(c, in, out, s, alfa, coef);
    if (r) return r;
    if (!c->p.invicid) {
      c->ucat_src = in;
      c->ucat_les_mixed = 0;
    }
    in = out;
  }
  if (c->immersed) {  // solvers.c:279-286
    int r = launch_ibm_interp(c, in, c->cur, 1);
    if (r) return r;
    r = launch_snapshot(c, in);
    if (r) return r;
    r = launch_boundary_post(c, in);
    if (r) return r;
  }
  c->cur = in;  // Qold <- Q (main.c:971)
  return CLES_OK;
}

int cles_runge_kutta(cles_ctx *c) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  int r = runge_kutta_async(c);
  if (r) return r;
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int cles_les_update(cles_ctx *c) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  if (!c->p.les) return set_error(c, CLES_EINVAL, "LES is off (-les 0)");
  int r = launch_les(c);
  if (r) return r;
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

static int flow_step_async(cles_ctx *c) {
  const cles_params &p = c->p;
  c->ti += 1;
  if (p.les && !p.invicid && p.les != 10) {
    int r = launch_les(c);
    if (r) return r;
  }
  int r = runge_kutta_async(c);
  if (r) return r;
  // Flow_Solver (solvers.c:52-53): statistics after the Runge-Kutta stages
  return p.averaging ? launch_averaging(c) : CLES_OK;
}

int cles_flow_step(cles_ctx *c) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  int r = flow_step_async(c);
  if (r) return r;
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int cles_advance(cles_ctx *c, int nsteps) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  for (int s = 0; s < nsteps; s++) {
    int r = flow_step_async(c);
    if (r) return r;
  }
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  return CLES_OK;
}

int cles_norms(cles_ctx *c, double out[2]) {
  if (!c || !out) return CLES_EINVAL;
  if (int g_ = member_guard(c, false)) return g_;
  unsigned long long h[128];
  CLES_HIP(c, hipMemcpyAsync(h, c->dnorm, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  CLES_HIP(c, hipStreamSynchronize(c->stream));
  unsigned long long m0 = 0, m1 = 0;
  for (int q = 0; q < 64; q++) {
    if (h[q] > m0) m0 = h[q];
    if (h[64 + q] > m1) m1 = h[64 + q];
  }
  memcpy(&out[0], &m0, sizeof(double));
  memcpy(&out[1], &m1, sizeof(double));
  return CLES_OK;
}

int cles_get_time(cles_ctx *c) { return c ? c->ti : CLES_EINVAL; }
int cles_set_time(cles_ctx *c, int ti) {
  if (!c) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  c->ti = ti;
  return CLES_OK;
}

int cles_time_stage_kernel(cles_ctx *c, int nrep, double *ms) {
  if (!c || !ms || nrep < 1) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  const int out = free_buffer(c, c->cur, c->ucat_src);
  // input: the last stage's input where a step has run, so that the timed
  // launch reads Qold as stages after the first do (the first reads its own
  // input as Qold); else the current state
  const int in = c->ucat_src >= 0 && c->ucat_src != c->cur ? c->ucat_src : c->cur;
  double coef[3], alfa;
  rk_coefficients(c->p, c->p.fluxorder > 1 ? 1 : 0, coef, &alfa);
  hipEvent_t e0, e1;
  CLES_HIP(c, hipEventCreate(&e0));
  CLES_HIP(c, hipEventCreate(&e1));
  int r = launch_stage(c, in, out, 1, 1, coef);  // warm-up
  if (r) return r;
  CLES_HIP(c, hipEventRecord(e0, c->stream));
  for (int q = 0; q < nrep; q++) {
    r = launch_stage(c, in, out, 1, 1, coef);
    if (r) return r;
  }
  CLES_HIP(c, hipEventRecord(e1, c->stream));
  CLES_HIP(c, hipEventSynchronize(e1));
  float t = 0;
  CLES_HIP(c, hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / nrep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return CLES_OK;
}

int cles_time_les(cles_ctx *c, int nrep, double *ms) {
  if (!c || !ms || nrep < 1) return CLES_EINVAL;
  if (int g_ = member_guard(c, true)) return g_;
  if (!c->have_state) return set_error(c, CLES_ESTATE, "no state uploaded");
  if (!c->p.les) return set_error(c, CLES_EINVAL, "LES is off");
  hipEvent_t e0, e1;
  CLES_HIP(c, hipEventCreate(&e0));
  CLES_HIP(c, hipEventCreate(&e1));
  CLES_HIP(c, hipEventRecord(e0, c->stream));
  for (int q = 0; q < nrep; q++) {
    c->ucat_les_mixed = 0;
    int r = launch_les(c);
    if (r) return r;
  }
  CLES_HIP(c, hipEventRecord(e1, c->stream));
  CLES_HIP(c, hipEventSynchronize(e1));
  c->ucat_les_mixed = 0;
  float t = 0;
  CLES_HIP(c, hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / nrep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return CLES_OK;
}

size_t cles_device_bytes(const cles_ctx *c) { return c ? c->bytes : 0; }

int cles_set_tile(cles_ctx *c, int tile) {
  if (!c || tile < 0 || tile > 7) return CLES_EINVAL;
  c->tile = tile;
  return CLES_OK;
}

const char *cles_last_error(const cles_ctx *c) { return c ? c->err : g_err; }

}  // extern "C"

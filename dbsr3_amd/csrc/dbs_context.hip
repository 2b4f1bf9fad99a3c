// dbs_context.hip -- grid context: knots, Gauss tables and B-spline values
// built on the device (gfx950), error reporting and memory helpers.
//
// Reference: def_grid / alloc_DBS_gauss / Def_pbsp / Def_qbsp
// (DBSR_HF/dbsr_lib_dbs.f90:480-596, 911-1074), gauleg
// (DBSR_HF/dbsr_lib_zcom.f90:1990-2039), zbsplvd/zbsplvb (:4197-4468).
// The table kernels keep the reference's operation order with FP
// contraction off, so the tables match the reference bit for bit.
#include <cmath>
#include <cstring>
#include <mutex>

#include "dbs_device.h"
#include <algorithm>

#include "dbs_internal.h"

namespace dbs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int kband(int type, int ks) { return type >= DBS_PPQQ ? 2 * ks - 1 : ks; }

int ensure_scratch(dbs_ctx* c, size_t n) {
  if (c->scratch_n >= n) return DBS_OK;
  if (c->scratch) {
    DBS_HIP(hipStreamSynchronize(c->stream));
    DBS_HIP(hipFree(c->scratch));
  }
  c->scratch = nullptr;
  c->scratch_n = 0;
  DBS_HIP(hipMalloc(&c->scratch, n * sizeof(double)));
  c->scratch_n = n;
  return DBS_OK;
}

int list_stage_begin(dbs_ctx* c, size_t bytes, char** host) {
  const int hb = c->list_cur;
  if (!c->list_ev[hb]) DBS_HIP(hipEventCreateWithFlags(&c->list_ev[hb], hipEventDisableTiming));
  DBS_HIP(hipEventSynchronize(c->list_ev[hb]));
  if (c->list_host_n[hb] < bytes) {
    if (c->list_host[hb]) DBS_HIP(hipHostFree(c->list_host[hb]));
    c->list_host[hb] = nullptr;
    c->list_host_n[hb] = 0;
    const size_t cap = std::max(bytes, (size_t)1 << 20);
    DBS_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->list_host[hb]), cap, hipHostMallocDefault));
    c->list_host_n[hb] = cap;
  }
  *host = c->list_host[hb];
  return DBS_OK;
}

int list_stage_end(dbs_ctx* c) {
  const int hb = c->list_cur;
  c->list_cur ^= 1;
  DBS_HIP(hipEventRecord(c->list_ev[hb], c->stream));
  return DBS_OK;
}

int ensure_stage(dbs_ctx* c, size_t n) {
  if (c->stage_n >= n) return DBS_OK;
  if (c->stage) {
    DBS_HIP(hipStreamSynchronize(c->stream));
    DBS_HIP(hipFree(c->stage));
  }
  c->stage = nullptr;
  c->stage_n = 0;
  DBS_HIP(hipMalloc(&c->stage, n * sizeof(double)));
  c->stage_n = n;
  return DBS_OK;
}

// Gauss-Legendre nodes and weights on [x1,x2] by Newton iteration on the
// Legendre recurrence (same iteration and stopping rule as gauleg).
static void gauss_legendre(double x1, double x2, double* x, double* w, int n) {
  const double eps = 1.0e-14;
  const int half = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= half; ++i) {
    double z = std::cos(3.141592654 * (i - 0.25) / (n + 0.5));
    double z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > eps);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

// ---------------------------------------------------------------------------
// table kernels
// ---------------------------------------------------------------------------
// outer Gauss points: gr(iv,m) = h*gx(m) + t(iv+ks-1), grw = h*gw(m), grm = 1/gr
__global__ void k_gauss_points(int nv, int ks, const double* __restrict__ t,
                               const double* __restrict__ gx, const double* __restrict__ gw,
                               double* gr, double* grw, double* grm) {
#pragma clang fp contract(off)
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv * ks) return;
  int iv = q % nv, m = q / nv;
  double h = t[iv + ks] - t[iv + ks - 1];
  double r = h * gx[m] + t[iv + ks - 1];
  gr[q] = r;
  grm[q] = 1.0 / r;
  grw[q] = h * gw[m];
}

// B-spline values at the outer points, xbsp(nv+1,ks,kx): rows 0..nv-1 at
// gr(iv,m) in interval left=iv+kx (1-based) of tx; row nv, column 0 holds
// the values at the last knot tx(nsx+1) (left = nsx), the rest zero.
__global__ void k_outer_bsplines(int nv, int ks, int kx, int nsx, const double* __restrict__ tx,
                                 const double* __restrict__ gr, double* xbsp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = nv + 1;
  if (q >= nw * ks) return;
  int iv = q % nw, m = q / nw;
  double b[DBS_MAXK];
  if (iv < nv) {
    bspline_values(tx, kx, iv + kx, gr[iv + nv * m], b);
  } else if (m == 0) {
    bspline_values(tx, kx, nsx, tx[nsx], b);
  } else {
    for (int j = 0; j < kx; ++j) b[j] = 0.0;
  }
  for (int j = 0; j < kx; ++j) xbsp[iv + (size_t)nw * (m + (size_t)ks * j)] = b[j];
}

// nested rule of the diagonal cells (triang_*, dbsr_lib_dbs.f90:3012-3031):
// inner point i of outer point m in interval iv.
__global__ void k_inner_points(int nv, int ks, int ksp, int ksq, int nsp, int nsq,
                               const double* __restrict__ t, const double* __restrict__ tp,
                               const double* __restrict__ tq, const double* __restrict__ gx,
                               const double* __restrict__ gw, const double* __restrict__ gr,
                               double* xin, double* win, double* bpin, double* bqin) {
#pragma clang fp contract(off)
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv * ks * ks) return;
  int i = q % ks, m = (q / ks) % ks, iv = q / (ks * ks);
  double xbase = t[iv + ks - 1];
  double h = gr[iv + nv * m] - xbase;
  double x = h * gx[i] + xbase;
  xin[q] = x;
  win[q] = h * gw[i];
  double b[DBS_MAXK];
  bspline_values(tp, ksp, iv + ksp, x, b);
  for (int j = 0; j < ksp; ++j) bpin[(size_t)q * ksp + j] = b[j];
  bspline_values(tq, ksq, iv + ksq, x, b);
  for (int j = 0; j < ksq; ++j) bqin[(size_t)q * ksq + j] = b[j];
}

}  // namespace dbs

using namespace dbs;

extern "C" {

const char* dbs_last_error(void) { return g_last_error.c_str(); }
int dbs_version(void) { return 10000; }

int dbs_ctx_create(dbs_ctx** out, int ns, int ks, int ksp, int ksq, const double* t, int device) {
  DBS_CHECK(out != nullptr && t != nullptr, DBS_ERR_ARG, "dbs_ctx_create: null argument");
  *out = nullptr;
  DBS_CHECK(ks >= 2 && ks <= DBS_MAXK, DBS_ERR_ARG, "dbs_ctx_create: ks out of range [2,16]");
  DBS_CHECK(ns >= ks, DBS_ERR_ARG, "dbs_ctx_create: ns < ks");
  for (int i = 1; i < ns + ks; ++i)
    DBS_CHECK(t[i] >= t[i - 1], DBS_ERR_ARG, "dbs_ctx_create: knots not non-decreasing");
  for (int i = ks - 1; i < ns; ++i)
    DBS_CHECK(t[i + 1] > t[i], DBS_ERR_ARG, "dbs_ctx_create: zero-length interval");
  if (ksp <= 0 || ksp > ks) ksp = ks - 1;
  if (ksq <= 0 || ksq > ks) ksq = ks;
  DBS_CHECK(ksp >= 2, DBS_ERR_ARG, "dbs_ctx_create: ksp < 2");
  int ndev = 0;
  DBS_HIP(hipGetDeviceCount(&ndev));
  DBS_CHECK(device >= 0 && device < ndev, DBS_ERR_ARG, "dbs_ctx_create: no such device");
  DBS_HIP(hipSetDevice(device));

  dbs_ctx* c = new dbs_ctx();
  c->device = device;
  c->ns = ns; c->ks = ks; c->ksp = ksp; c->ksq = ksq;
  c->ms = 2 * ns; c->nv = ns - ks + 1;
  c->nsp = c->nv + ksp - 1; c->nsq = c->nv + ksq - 1;
  const int nv = c->nv, nsp = c->nsp, nsq = c->nsq;

  c->h_t.assign(t, t + ns + ks);
  c->h_gx.resize(ks);
  c->h_gw.resize(ks);
  gauss_legendre(0.0, 1.0, c->h_gx.data(), c->h_gw.data(), ks);
  std::vector<double> htp(nsp + ksp), htq(nsq + ksq);
  for (int i = 0; i < ksp; ++i) htp[i] = t[0];
  for (int i = ksp; i < nsp; ++i) htp[i] = t[ks + (i - ksp)];
  for (int i = nsp; i < nsp + ksp; ++i) htp[i] = t[ns + (i - nsp)];
  for (int i = 0; i < ksq; ++i) htq[i] = t[0];
  for (int i = ksq; i < nsq; ++i) htq[i] = t[ks + (i - ksq)];
  for (int i = nsq; i < nsq + ksq; ++i) htq[i] = t[ns + (i - nsq)];

  auto cleanup = [&](int rc) {
    dbs_ctx_destroy(c);
    return rc;
  };
#define ALLOC(ptr, n)                                                              \
  do {                                                                             \
    hipError_t e_ = hipMalloc(&(ptr), (size_t)(n) * sizeof(double));               \
    if (e_ != hipSuccess) return cleanup(fail(DBS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e_))); \
  } while (0)
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail(DBS_ERR_HIP, "hipStreamCreate failed"));
  c->own_stream = true;
  ALLOC(c->t, ns + ks);
  ALLOC(c->tp, nsp + ksp);
  ALLOC(c->tq, nsq + ksq);
  ALLOC(c->gx, ks);
  ALLOC(c->gw, ks);
  ALLOC(c->gr, nv * ks);
  ALLOC(c->grw, nv * ks);
  ALLOC(c->grm, nv * ks);
  ALLOC(c->pbsp, (nv + 1) * ks * ksp);
  ALLOC(c->qbsp, (nv + 1) * ks * ksq);
  ALLOC(c->xin, (size_t)nv * ks * ks);
  ALLOC(c->win, (size_t)nv * ks * ks);
  ALLOC(c->bpin, (size_t)nv * ks * ks * ksp);
  ALLOC(c->bqin, (size_t)nv * ks * ks * ksq);
  ALLOC(c->mom, (size_t)4 * ks * ks * nv);
  ALLOC(c->rkd, (size_t)ks * ks * ks * ks * nv);
#undef ALLOC
  hipStream_t s = c->stream;
  e = hipMemcpyAsync(c->t, t, sizeof(double) * (ns + ks), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->tp, htp.data(), sizeof(double) * (nsp + ksp), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->tq, htq.data(), sizeof(double) * (nsq + ksq), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->gx, c->h_gx.data(), sizeof(double) * ks, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(c->gw, c->h_gw.data(), sizeof(double) * ks, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return cleanup(fail(DBS_ERR_HIP, "dbs_ctx_create: upload failed"));
  const int B = 256;
  k_gauss_points<<<(nv * ks + B - 1) / B, B, 0, s>>>(nv, ks, c->t, c->gx, c->gw, c->gr, c->grw, c->grm);
  k_outer_bsplines<<<((nv + 1) * ks + B - 1) / B, B, 0, s>>>(nv, ks, ksp, nsp, c->tp, c->gr, c->pbsp);
  k_outer_bsplines<<<((nv + 1) * ks + B - 1) / B, B, 0, s>>>(nv, ks, ksq, nsq, c->tq, c->gr, c->qbsp);
  k_inner_points<<<(nv * ks * ks + B - 1) / B, B, 0, s>>>(nv, ks, ksp, ksq, nsp, nsq, c->t, c->tp,
                                                         c->tq, c->gx, c->gw, c->gr, c->xin,
                                                         c->win, c->bpin, c->bqin);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return cleanup(fail(DBS_ERR_HIP, std::string("dbs_ctx_create: table kernels: ") + hipGetErrorString(e)));
  *out = c;
  return DBS_OK;
}

int dbs_ctx_destroy(dbs_ctx* c) {
  if (!c) return DBS_OK;
  // teardown: failures here leave nothing to recover
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  double* bufs[] = {c->t, c->tp, c->tq, c->gx, c->gw, c->gr, c->grw, c->grm, c->pbsp, c->qbsp,
                    c->xin, c->win, c->bpin, c->bqin, c->mom, c->rkd, c->scratch, c->stage};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  for (auto& kv : c->rkb_cache) (void)hipFree(kv.second);
  for (int b = 0; b < 2; ++b) {
    if (c->list_host[b]) (void)hipHostFree(c->list_host[b]);
    if (c->list_ev[b]) (void)hipEventDestroy(c->list_ev[b]);
  }
  if (c->solver && c->solver_free) c->solver_free(c->solver);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DBS_OK;
}

int dbs_ctx_dims(const dbs_ctx* c, int* d) {
  DBS_CHECK(c && d, DBS_ERR_ARG, "dbs_ctx_dims: null argument");
  d[0] = c->ns; d[1] = c->ks; d[2] = c->ksp; d[3] = c->ksq;
  d[4] = c->nv; d[5] = c->nsp; d[6] = c->nsq; d[7] = c->ms;
  return DBS_OK;
}

int dbs_ctx_sync(dbs_ctx* c) {
  DBS_CHECK(c, DBS_ERR_ARG, "dbs_ctx_sync: null context");
  DBS_HIP(hipStreamSynchronize(c->stream));
  return DBS_OK;
}

int dbs_ctx_tables(dbs_ctx* c, double* gr, double* grw, double* pbsp, double* qbsp) {
  DBS_CHECK(c, DBS_ERR_ARG, "dbs_ctx_tables: null context");
  const size_t nvk = (size_t)c->nv * c->ks;
  DBS_HIP(hipStreamSynchronize(c->stream));
  if (gr) DBS_HIP(hipMemcpy(gr, c->gr, nvk * sizeof(double), hipMemcpyDeviceToHost));
  if (grw) DBS_HIP(hipMemcpy(grw, c->grw, nvk * sizeof(double), hipMemcpyDeviceToHost));
  if (pbsp) DBS_HIP(hipMemcpy(pbsp, c->pbsp, (size_t)(c->nv + 1) * c->ks * c->ksp * sizeof(double), hipMemcpyDeviceToHost));
  if (qbsp) DBS_HIP(hipMemcpy(qbsp, c->qbsp, (size_t)(c->nv + 1) * c->ks * c->ksq * sizeof(double), hipMemcpyDeviceToHost));
  return DBS_OK;
}

int dbs_ctx_set_cache_limit(dbs_ctx* c, size_t bytes) {
  DBS_CHECK(c, DBS_ERR_ARG, "dbs_ctx_set_cache_limit: null context");
  c->rkb_cache_limit = bytes;
  return DBS_OK;
}

int dbs_ctx_set_workspace_limit(dbs_ctx* c, size_t bytes) {
  DBS_CHECK(c, DBS_ERR_ARG, "dbs_ctx_set_workspace_limit: null context");
  DBS_CHECK(bytes >= ((size_t)64 << 10), DBS_ERR_ARG, "dbs_ctx_set_workspace_limit: below 64 KiB");
  c->work_limit = bytes;
  return DBS_OK;
}

int dbs_malloc(dbs_ctx* c, void** p, size_t bytes) {
  DBS_CHECK(c && p, DBS_ERR_ARG, "dbs_malloc: null argument");
  DBS_HIP(hipSetDevice(c->device));
  DBS_HIP(hipMalloc(p, bytes ? bytes : 8));
  return DBS_OK;
}
int dbs_free(dbs_ctx* c, void* p) {
  DBS_CHECK(c, DBS_ERR_ARG, "dbs_free: null context");
  if (p) {
    DBS_HIP(hipStreamSynchronize(c->stream));
    DBS_HIP(hipFree(p));
  }
  return DBS_OK;
}
int dbs_memcpy_h2d(dbs_ctx* c, void* dst, const void* src, size_t bytes) {
  DBS_CHECK(c && dst && src, DBS_ERR_ARG, "dbs_memcpy_h2d: null argument");
  DBS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  DBS_HIP(hipStreamSynchronize(c->stream));
  return DBS_OK;
}
int dbs_memcpy_d2h(dbs_ctx* c, void* dst, const void* src, size_t bytes) {
  DBS_CHECK(c && dst && src, DBS_ERR_ARG, "dbs_memcpy_d2h: null argument");
  DBS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  DBS_HIP(hipStreamSynchronize(c->stream));
  return DBS_OK;
}

}  // extern "C"

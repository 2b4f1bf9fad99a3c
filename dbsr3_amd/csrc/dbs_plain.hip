// dbs_plain.hip -- the plain path: the reference's routines one for one on
// the device.  Moment tiles, diagonal-cell quadrature, materialised R^k/S^k
// arrays, densities, convolutions, band dot products, UPDATE_HS, quadr, rk.
//
// This path keeps the reference's data layout (rkb(ns,ns,kb,kb) in HBM) and
// is the drop-in baseline the engine (dbs_engine.hip) is measured against.
// Kernels that build arrays run with FP contraction off and the reference's
// summation order per element, so rkb matches the reference bit for bit.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "dbs_device.h"
#include "dbs_internal.h"

namespace dbs {

// which moment tile: 0 pp, 1 qq, 2 pq, 3 qp
struct MomSpec {
  const double* b1;
  const double* b2;
  int k1, k2, sym;
};
static MomSpec mom_spec(const dbs_ctx* c, int which) {
  switch (which) {
    case 0: return {c->pbsp, c->pbsp, c->ksp, c->ksp, 1};
    case 1: return {c->qbsp, c->qbsp, c->ksq, c->ksq, 1};
    case 2: return {c->pbsp, c->qbsp, c->ksp, c->ksq, 0};
    default: return {c->qbsp, c->pbsp, c->ksq, c->ksp, 0};
  }
}

// moments_* (dbsr_lib_dbs.f90:1384-1564): rkm(ii,iv) = sum_m b1(m,i)*(b2(m,j)*gw(m))
__global__ void k_moments(int nv, int ks, int k, const double* __restrict__ gr,
                          const double* __restrict__ grw, const double* __restrict__ grm,
                          const double* __restrict__ b1t, const double* __restrict__ b2t, int k1,
                          int k2, int sym, double* rkm) {
#pragma clang fp contract(off)
  const int kk = ks * ks;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kk * nv) return;
  int ii = q % kk, iv = q / kk;
  int npair = sym ? k1 * (k1 + 1) / 2 : k1 * k2;
  if (ii >= npair) {
    rkm[q] = 0.0;
    return;
  }
  int i, j;
  if (sym) {
    i = 0;
    int rem = ii;
    while (rem >= k1 - i) {
      rem -= k1 - i;
      ++i;
    }
    j = i + rem;
  } else {
    i = ii / k2;
    j = ii % k2;
  }
  const int nw = nv + 1;
  double s = 0.0;
  for (int m = 0; m < ks; ++m) {
    double w = grw[iv + nv * m];
    if (k > 0) w = w * ipow(gr[iv + nv * m], k);
    if (k < 0) w = w * ipow(grm[iv + nv * m], -k);
    double b2 = b2t[iv + (size_t)nw * (m + (size_t)ks * j)] * w;
    s = s + b1t[iv + (size_t)nw * (m + (size_t)ks * i)] * b2;
  }
  rkm[q] = s;
}

// triang_* diagonal cells, one workgroup per interval.  type 1..6.
// rkd(ks*ks, ks*ks, nv), column-major, rows = left (r1) pair, cols = right.
__global__ void __launch_bounds__(256) k_diag_cells(
    int nv, int ks, int ksp, int ksq, int type, int k, const double* __restrict__ grw,
    const double* __restrict__ grm, const double* __restrict__ pbsp,
    const double* __restrict__ qbsp, const double* __restrict__ xin,
    const double* __restrict__ win, const double* __restrict__ bpin,
    const double* __restrict__ bqin, double* rkd) {
#pragma clang fp contract(off)
  const int iv = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int kk = ks * ks;
  const int nw = nv + 1;
  __shared__ double s_w[DBS_MAXK * DBS_MAXK];           // inner weight (m,i)
  __shared__ double s_gx2[DBS_MAXK];                    // outer weight
  __shared__ double s_intp[DBS_MAXK * DBS_MAXK * DBS_MAXK];               // INTP or INTPQ
  __shared__ double s_intq[DBS_MAXK * (DBS_MAXK + 1) / 2 * DBS_MAXK];     // INTQ
  const bool symt = (type == DBS_PPPP || type == DBS_QQQQ);
  const bool mixed = (type == DBS_PPQQ || type == DBS_PQQP);
  const bool needP = type != DBS_QQQQ;
  const bool needQ = type != DBS_PPPP;
  const int TP = ksp * (ksp + 1) / 2, TQ = ksq * (ksq + 1) / 2;
  const size_t base = (size_t)iv * ks * ks;

  for (int q = tid; q < ks * ks; q += nt) {
    int m = q / ks, i = q % ks;
    double x = xin[base + q], gw = win[base + q], v;
    if (symt) {
      if (k > 1) v = gw * ipow(x, k);
      else if (k == 1) v = gw * x;
      else v = gw;
    } else {
      v = gw * ipow(x, k);
    }
    s_w[m * ks + i] = v;
  }
  for (int m = tid; m < ks; m += nt) {
    double g = grm[iv + nv * m];
    s_gx2[m] = (symt && k == 0) ? grw[iv + nv * m] * g : grw[iv + nv * m] * ipow(g, k + 1);
  }
  __syncthreads();
  // inner integrals INT(pair,m) = sum_i (w(m,i)*B_j(x_mi)) * B_jp(x_mi)
  if (!mixed) {
    if (needP)
      for (int q = tid; q < TP * ks; q += nt) {
        int pr = q / ks, m = q % ks;
        int j = 0, rem = pr;
        while (rem >= ksp - j) { rem -= ksp - j; ++j; }
        int jp = j + rem;
        double s = 0.0;
        for (int i = 0; i < ks; ++i) {
          const double* b = bpin + (base + (size_t)m * ks + i) * ksp;
          double g = s_w[m * ks + i] * b[j];
          s = s + g * b[jp];
        }
        s_intp[pr * ks + m] = s;
      }
    if (needQ)
      for (int q = tid; q < TQ * ks; q += nt) {
        int pr = q / ks, m = q % ks;
        int j = 0, rem = pr;
        while (rem >= ksq - j) { rem -= ksq - j; ++j; }
        int jp = j + rem;
        double s = 0.0;
        for (int i = 0; i < ks; ++i) {
          const double* b = bqin + (base + (size_t)m * ks + i) * ksq;
          double g = s_w[m * ks + i] * b[j];
          s = s + g * b[jp];
        }
        s_intq[pr * ks + m] = s;
      }
  } else {
    // INTPQ(j in P, jp in Q)
    for (int q = tid; q < ksp * ksq * ks; q += nt) {
      int pr = q / ks, m = q % ks;
      int j = pr / ksq, jp = pr % ksq;
      double s = 0.0;
      for (int i = 0; i < ks; ++i) {
        const double* bp = bpin + (base + (size_t)m * ks + i) * ksp;
        const double* bq = bqin + (base + (size_t)m * ks + i) * ksq;
        double g = s_w[m * ks + i] * bp[j];
        s = s + g * bq[jp];
      }
      s_intp[pr * ks + m] = s;
    }
  }
  __syncthreads();
  double* out = rkd + (size_t)iv * kk * kk;
  auto PB = [&](int m, int j) { return pbsp[iv + (size_t)nw * (m + (size_t)ks * j)]; };
  auto QB = [&](int m, int j) { return qbsp[iv + (size_t)nw * (m + (size_t)ks * j)]; };
  for (int q = tid; q < kk * kk; q += nt) {
    int r = q % kk, cc = q / kk;  // rkd(r, cc)
    double v = 0.0;
    if (symt) {
      const int kx = (type == DBS_PPPP) ? ksp : ksq;
      const int T = kx * (kx + 1) / 2;
      const double* IN = (type == DBS_PPPP) ? s_intp : s_intq;
      if (r < T && cc < T) {
        // a(r,cc) + a(cc,r); a(ii,jj) = sum_m ((B_i B_ip)(m) * gx2(m)) * INT(jj,m)
        double a1 = 0.0, a2 = 0.0;
        int i1 = 0, rem = r;
        while (rem >= kx - i1) { rem -= kx - i1; ++i1; }
        int ip1 = i1 + rem;
        int i2 = 0; rem = cc;
        while (rem >= kx - i2) { rem -= kx - i2; ++i2; }
        int ip2 = i2 + rem;
        for (int m = 0; m < ks; ++m) {
          double bi = (type == DBS_PPPP) ? PB(m, i1) * PB(m, ip1) : QB(m, i1) * QB(m, ip1);
          bi = bi * s_gx2[m];
          a1 = a1 + bi * IN[cc * ks + m];
        }
        for (int m = 0; m < ks; ++m) {
          double bi = (type == DBS_PPPP) ? PB(m, i2) * PB(m, ip2) : QB(m, i2) * QB(m, ip2);
          bi = bi * s_gx2[m];
          a2 = a2 + bi * IN[r * ks + m];
        }
        v = a1 + a2;
      }
    } else if (!mixed) {
      // pqpq: rows P pairs, cols Q pairs; qpqp: rows Q pairs, cols P pairs
      const bool pq = (type == DBS_PQPQ);
      const int TR = pq ? TP : TQ, TC = pq ? TQ : TP;
      if (r < TR && cc < TC) {
        int pP = pq ? r : cc, pQ = pq ? cc : r;
        int i = 0, rem = pP;
        while (rem >= ksp - i) { rem -= ksp - i; ++i; }
        int ip = i + rem;
        int iq = 0; rem = pQ;
        while (rem >= ksq - iq) { rem -= ksq - iq; ++iq; }
        int ipq = iq + rem;
        // pass 1: P outer, Q inner; pass 2: Q outer, P inner.  The
        // reference's pass order is (1 then 2) for pqpq and qpqp alike.
        double s1 = 0.0, s2 = 0.0;
        for (int m = 0; m < ks; ++m) {
          double bi = PB(m, i) * PB(m, ip);
          bi = bi * s_gx2[m];
          s1 = s1 + bi * s_intq[pQ * ks + m];
        }
        for (int m = 0; m < ks; ++m) {
          double bi = QB(m, iq) * QB(m, ipq);
          bi = bi * s_gx2[m];
          s2 = s2 + bi * s_intp[pP * ks + m];
        }
        double z = 0.0;
        v = (z + s1) + s2;
      }
    } else {
      // S types: rkd(jj,ii) with jj inner (P,Q) pair (rows), ii outer pair
      // (ppqq: (P,Q); pqqp: (Q,P)) (cols)
      const int ko1 = (type == DBS_PPQQ) ? ksp : ksq, ko2 = (type == DBS_PPQQ) ? ksq : ksp;
      if (r < ksp * ksq && cc < ko1 * ko2) {
        int i = cc / ko2, ip = cc % ko2;
        double s = 0.0;
        for (int m = 0; m < ks; ++m) {
          double bv = (type == DBS_PPQQ) ? PB(m, i) * QB(m, ip) : QB(m, i) * PB(m, ip);
          double bi = bv * s_gx2[m];
          s = s + bi * s_intp[r * ks + m];
        }
        v = s;
      }
    }
    out[r + (size_t)kk * cc] = v;
  }
}

// Assemble rkb(i,j,ip,jp) by gathering every cell contribution in the
// reference's accumulation order (jv outer, iv inner; mrk_* loops).
// M1..M4 = moment tiles (kk,nv) as in mrk_*; rkd = diagonal cells.
__global__ void k_rkb_assemble(int ns, int nv, int ks, int kl, int kr, int type,
                               const double* __restrict__ M1, const double* __restrict__ M2,
                               const double* __restrict__ M3, const double* __restrict__ M4,
                               const double* __restrict__ rkd, double* rkb) {
#pragma clang fp contract(off)
  const int kk = ks * ks;
  const size_t nsz = (size_t)ns;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsz * nsz * ks * ks) return;
  int i = (int)(q % nsz);
  int j = (int)((q / nsz) % nsz);
  int ip = (int)((q / (nsz * nsz)) % ks);
  int jp = (int)(q / (nsz * nsz * ks));
  double acc = 0.0;
  // right (r2) cells jv: j = jv + jh, jhp = jh + jp < kr
  int jv_lo = j - (kr - 1 - jp);
  if (jv_lo < 0) jv_lo = 0;
  int jv_hi = j < nv - 1 ? j : nv - 1;
  int iv_lo = i - (kl - 1 - ip);
  if (iv_lo < 0) iv_lo = 0;
  int iv_hi = i < nv - 1 ? i : nv - 1;
  if (ip < kl && jp < kr) {
    for (int jv = jv_lo; jv <= jv_hi; ++jv) {
      int jh = j - jv;
      int jj = packed_index(jh, jh + jp, kr);
      for (int iv = iv_lo; iv <= iv_hi; ++iv) {
        int ih = i - iv;
        int ii = packed_index(ih, ih + ip, kl);
        double c;
        if (iv < jv) {
          c = (type == DBS_QPQP) ? M3[ii + kk * iv] * M4[jj + kk * jv]
                                 : M1[ii + kk * iv] * M2[jj + kk * jv];
        } else if (iv > jv) {
          c = (type == DBS_PQPQ) ? M3[jj + kk * jv] * M4[ii + kk * iv]
                                 : M1[jj + kk * jv] * M2[ii + kk * iv];
        } else {
          c = rkd[ii + (size_t)kk * (jj + (size_t)kk * iv)];
        }
        acc = acc + c;
      }
    }
  }
  rkb[q] = acc;
}

// S^k arrays (msk_ppqq / msk_pqqp): cells iv<=jv, non-symmetric band
// ip = ihp-ih+ks-1 (0-based), kb = 2ks-1.
__global__ void k_rsk_assemble(int ns, int nv, int ks, int ksp, int ksq, int type,
                               const double* __restrict__ M1, const double* __restrict__ M2,
                               const double* __restrict__ rkd, double* rkb) {
#pragma clang fp contract(off)
  const int kk = ks * ks, kb = 2 * ks - 1;
  const size_t nsz = (size_t)ns;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsz * nsz * kb * kb) return;
  int i = (int)(q % nsz);
  int j = (int)((q / nsz) % nsz);
  int ip = (int)((q / (nsz * nsz)) % kb);
  int jp = (int)(q / (nsz * nsz * kb));
  const int kr1 = (type == DBS_PPQQ) ? ksp : ksq, kr2 = (type == DBS_PPQQ) ? ksq : ksp;
  double acc = 0.0;
  for (int jv = 0; jv < nv; ++jv) {
    int jh = j - jv;
    if (jh < 0) break;
    if (jh >= kr1) continue;
    int jhp = jp - ks + 1 + jh;
    if (jhp < 0 || jhp >= kr2) continue;
    int jj = jh * kr2 + jhp;
    for (int iv = 0; iv <= jv; ++iv) {
      int ih = i - iv;
      if (ih < 0) break;
      if (ih >= ksp) continue;
      int ihp = ip - ks + 1 + ih;
      if (ihp < 0 || ihp >= ksq) continue;
      int ii = ih * ksq + ihp;
      double c = (iv == jv) ? rkd[ii + (size_t)kk * (jj + (size_t)kk * iv)]
                            : M1[ii + kk * iv] * M2[jj + kk * jv];
      acc = acc + c;
    }
  }
  rkb[q] = acc;
}

static int launch_moments(dbs_ctx* c, int which, int k, double* rkm) {
  MomSpec s = mom_spec(c, which);
  const int n = c->ks * c->ks * c->nv;
  k_moments<<<(n + 255) / 256, 256, 0, c->stream>>>(c->nv, c->ks, k, c->gr, c->grw, c->grm, s.b1,
                                                    s.b2, s.k1, s.k2, s.sym, rkm);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

static int launch_diag(dbs_ctx* c, int type, int k, double* rkd) {
  k_diag_cells<<<c->nv, 256, 0, c->stream>>>(c->nv, c->ks, c->ksp, c->ksq, type, k, c->grw,
                                              c->grm, c->pbsp, c->qbsp, c->xin, c->win, c->bpin,
                                              c->bqin, rkd);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int build_rkb(dbs_ctx* c, int type, int k, double** out) {
  DBS_CHECK(type >= 1 && type <= 6, DBS_ERR_ARG, "Rk_evaluate: unknown type of intgrals");
  DBS_CHECK(k >= 0 && k < 64, DBS_ERR_ARG, "mrk: multipole index out of range");
  const int key = type * 1000 + k;
  auto it = c->rkb_cache.find(key);
  if (it != c->rkb_cache.end()) {
    *out = it->second;
    auto p = std::find(c->rkb_lru.begin(), c->rkb_lru.end(), key);
    if (p != c->rkb_lru.end()) c->rkb_lru.erase(p);
    c->rkb_lru.push_back(key);
    return DBS_OK;
  }
  const int ns = c->ns, ks = c->ks, kb = kband(type, ks), kk = ks * ks, nv = c->nv;
  const size_t n = (size_t)ns * ns * kb * kb;
  const size_t bytes = n * sizeof(double);
  // evict least recently used arrays beyond the cache limit
  while (!c->rkb_lru.empty() && c->rkb_cache_bytes + bytes > c->rkb_cache_limit) {
    int old = c->rkb_lru.front();
    c->rkb_lru.erase(c->rkb_lru.begin());
    DBS_HIP(hipStreamSynchronize(c->stream));
    int okb = kband((old >= 0 ? old : -old - 1) / 1000, ks);  // negative keys: build_rkm's matrices
    DBS_HIP(hipFree(c->rkb_cache[old]));
    c->rkb_cache_bytes -= (size_t)ns * ns * okb * okb * sizeof(double);
    c->rkb_cache.erase(old);
  }
  double* rkb = nullptr;
  DBS_HIP(hipMalloc(&rkb, bytes));
  double* M1 = c->mom;
  double* M2 = M1 + (size_t)kk * nv;
  double* M3 = M2 + (size_t)kk * nv;
  double* M4 = M3 + (size_t)kk * nv;
  int rc = DBS_OK;
  switch (type) {
    case DBS_PPPP: rc |= launch_moments(c, 0, k, M1); rc |= launch_moments(c, 0, -(k + 1), M2); break;
    case DBS_QQQQ: rc |= launch_moments(c, 1, k, M1); rc |= launch_moments(c, 1, -(k + 1), M2); break;
    case DBS_PQPQ:
    case DBS_QPQP:
      rc |= launch_moments(c, 0, k, M1); rc |= launch_moments(c, 1, -(k + 1), M2);
      rc |= launch_moments(c, 1, k, M3); rc |= launch_moments(c, 0, -(k + 1), M4); break;
    case DBS_PPQQ: rc |= launch_moments(c, 2, k, M1); rc |= launch_moments(c, 2, -(k + 1), M2); break;
    default: rc |= launch_moments(c, 2, k, M1); rc |= launch_moments(c, 3, -(k + 1), M2); break;
  }
  if (rc == DBS_OK) rc = launch_diag(c, type, k, c->rkd);
  if (rc != DBS_OK) {
    (void)hipFree(rkb);
    return rc;
  }
  const int B = 256;
  if (type <= DBS_QPQP) {
    int kl = (type == DBS_PPPP || type == DBS_PQPQ) ? c->ksp : c->ksq;
    int kr = (type == DBS_PPPP || type == DBS_QPQP) ? c->ksp : c->ksq;
    k_rkb_assemble<<<(unsigned)((n + B - 1) / B), B, 0, c->stream>>>(ns, nv, ks, kl, kr, type, M1,
                                                                      M2, M3, M4, c->rkd, rkb);
  } else {
    k_rsk_assemble<<<(unsigned)((n + B - 1) / B), B, 0, c->stream>>>(ns, nv, ks, c->ksp, c->ksq,
                                                                      type, M1, M2, c->rkd, rkb);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(rkb);
    return fail(DBS_ERR_HIP, std::string("mrk: ") + hipGetErrorString(e));
  }
  c->rkb_cache[key] = rkb;
  c->rkb_lru.push_back(key);
  c->rkb_cache_bytes += bytes;
  *out = rkb;
  return DBS_OK;
}

// rkm(r, c) = rkb(i, j, ip, jp) with r = i + ns*ip, c = j + ns*jp (K x K,
// K = ns*kb, column-major): the read side coalesced along i, the write side
// along r
__global__ void k_rkb_permute(int ns, int kb, const double* __restrict__ rkb, double* __restrict__ rkm) {
  const long long K = (long long)ns * kb, n = K * K;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long r = q % K, cc = q / K;
    const int i = (int)(r % ns), ip = (int)(r / ns), j = (int)(cc % ns), jp = (int)(cc / ns);
    rkm[q] = rkb[i + (long long)ns * (j + (long long)ns * (ip + (long long)kb * jp))];
  }
}

int build_rkm(dbs_ctx* c, int type, int k, double** out) {
  const int key = -(type * 1000 + k) - 1;
  auto it = c->rkb_cache.find(key);
  if (it != c->rkb_cache.end()) {
    *out = it->second;
    auto p = std::find(c->rkb_lru.begin(), c->rkb_lru.end(), key);
    if (p != c->rkb_lru.end()) c->rkb_lru.erase(p);
    c->rkb_lru.push_back(key);
    return DBS_OK;
  }
  double* rkb = nullptr;
  int rc = build_rkb(c, type, k, &rkb);
  if (rc) return rc;
  const int ns = c->ns, kb = kband(type, c->ks);
  const size_t bytes = (size_t)ns * ns * kb * kb * sizeof(double);
  // evict least recently used arrays beyond the cache limit, never the
  // array just built or touched (the back of the list)
  while (c->rkb_lru.size() > 1 && c->rkb_cache_bytes + bytes > c->rkb_cache_limit) {
    int old = c->rkb_lru.front();
    c->rkb_lru.erase(c->rkb_lru.begin());
    DBS_HIP(hipStreamSynchronize(c->stream));
    int okb = kband((old >= 0 ? old : -old - 1) / 1000, c->ks);
    DBS_HIP(hipFree(c->rkb_cache[old]));
    c->rkb_cache_bytes -= (size_t)ns * ns * okb * okb * sizeof(double);
    c->rkb_cache.erase(old);
  }
  double* rkm = nullptr;
  DBS_HIP(hipMalloc(&rkm, bytes));
  const long long n = (long long)ns * ns * kb * kb;
  k_rkb_permute<<<(unsigned)std::min<long long>((n + 255) / 256, 1 << 20), 256, 0, c->stream>>>(ns, kb, rkb, rkm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(rkm);
    return fail(DBS_ERR_HIP, std::string("mrk: ") + hipGetErrorString(e));
  }
  c->rkb_cache[key] = rkm;
  c->rkb_lru.push_back(key);
  c->rkb_cache_bytes += bytes;
  *out = rkm;
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// densities (density, dbsr_lib_dbs.f90:2834-2891)
// ---------------------------------------------------------------------------
__global__ void k_density(int ns, int ks, int type, const double* __restrict__ p1,
                          const double* __restrict__ p2, double* d) {
#pragma clang fp contract(off)
  int ncol = (type == 's' || type == 'u' || type == 'l') ? ks : (type == 'n' ? 2 * ks - 1 : ns);
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)ns * ncol) return;
  int i = (int)(q % ns), j = (int)(q / ns);  // 0-based
  double v = 0.0;
  if (type == 's' || type == 'u') {
    if (j == 0) v = p1[i] * p2[i];
    else if (i < ns - j) v = p1[i] * p2[i + j] + p1[i + j] * p2[i];
  } else if (type == 'l') {
    int jf = j + 1, iff = i + 1;
    if (jf == ks) v = p1[i] * p2[i];
    else if (iff >= ks + 1 - jf) v = p1[i] * p2[i + jf - ks] + p1[i + jf - ks] * p2[i];
  } else if (type == 'n') {
    int jf = j + 1, iff = i + 1;
    int imin = 1 > 1 + ks - jf ? 1 : 1 + ks - jf, imax = ns < ns + ks - jf ? ns : ns + ks - jf;
    if (iff >= imin && iff <= imax) v = p1[i] * p2[i + jf - ks];
  } else {
    v = p1[i] * p2[j];
  }
  d[q] = v;
}

// ---------------------------------------------------------------------------
// convol cases 1 and 2 (band outputs)
// ---------------------------------------------------------------------------
// band ranges of the reference (1-based): 's' ip 1..ks, i 1..ns-ip+1;
// 'l' ip 1..ks, i ks+1-ip..ns; 'n' ip 1..2ks-1, i max(1,1+ks-ip)..min(ns,ns+ks-ip)
__device__ inline int band_cols(char s, int ks) { return s == 'n' ? 2 * ks - 1 : ks; }
__device__ inline void band_range(char s, int ns, int ks, int ipf, int* lo, int* hi) {
  if (s == 's') { *lo = 1; *hi = ns - ipf + 1; }
  else if (s == 'l') { *lo = ks + 1 - ipf; *hi = ns; }
  else { *lo = 1 > 1 + ks - ipf ? 1 : 1 + ks - ipf; *hi = ns < ns + ks - ipf ? ns : ns + ks - ipf; }
}

// case 1: partial(i, ip; jp, chunk) = sum_{j in chunk} d(j,jp) rkb(i,j,ip,jp)
__global__ void k_convol1_partial(int ns, int ks, int kb, char si, char sj, int jchunk,
                                  const double* __restrict__ rkb, const double* __restrict__ d,
                                  double* part) {
  const int ip0 = blockIdx.x, jp0 = blockIdx.y, ch = blockIdx.z;
  const int nci = band_cols(si, ks);
  const int ncj = band_cols(sj, ks);
  double* out = part + ((size_t)(jp0 * gridDim.z + ch) * ns) * nci + (size_t)ip0 * ns;
  if (ip0 >= nci) return;
  int jlo, jhi;
  if (jp0 < ncj) band_range(sj, ns, ks, jp0 + 1, &jlo, &jhi);
  else { jlo = 1; jhi = 0; }
  int ilo, ihi;
  band_range(si, ns, ks, ip0 + 1, &ilo, &ihi);
  int j0 = jlo + ch * jchunk, j1 = j0 + jchunk - 1;
  if (j1 > jhi) j1 = jhi;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    double c = 0.0;
    if (i + 1 >= ilo && i + 1 <= ihi)
      for (int jf = j0; jf <= j1; ++jf)
        c += d[(jf - 1) + (size_t)ns * jp0] *
             rkb[i + (size_t)ns * ((jf - 1) + (size_t)ns * (ip0 + (size_t)kb * jp0))];
    out[i] = c;
  }
}

__global__ void k_convol1_reduce(int ns, int ks, char si, int nparts, const double* __restrict__ part,
                                 double* a) {
  const int nci = band_cols(si, ks);
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ns * nci) return;
  int i = q % ns, ip0 = q / ns;
  int ilo, ihi;
  band_range(si, ns, ks, ip0 + 1, &ilo, &ihi);
  double c = 0.0;
  if (i + 1 >= ilo && i + 1 <= ihi)
    for (int p = 0; p < nparts; ++p) c += part[(size_t)p * ns * nci + q];
  a[q] = c;
}

// case 2: a(j,jp) = sum_ip sum_i d(i,ip) rkb(i,j,ip,jp); one wavefront per output
__global__ void k_convol2(int ns, int ks, int kb, char si, char sj, const double* __restrict__ rkb,
                          const double* __restrict__ d, double* a) {
  const int ncj = band_cols(sj, ks), nci = band_cols(si, ks);
  const int lane = threadIdx.x & 63;
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (w >= ns * ncj) return;
  const int j = w % ns, jp0 = w / ns;
  int jlo, jhi;
  band_range(sj, ns, ks, jp0 + 1, &jlo, &jhi);
  double c = 0.0;
  if (j + 1 >= jlo && j + 1 <= jhi) {
    for (int ip0 = 0; ip0 < nci; ++ip0) {
      int ilo, ihi;
      band_range(si, ns, ks, ip0 + 1, &ilo, &ihi);
      const double* col = rkb + (size_t)ns * (j + (size_t)ns * (ip0 + (size_t)kb * jp0));
      for (int i = ilo - 1 + lane; i <= ihi - 1; i += 64) c += d[i + (size_t)ns * ip0] * col[i];
    }
  }
  c = wave_sum(c);
  if (lane == 0) a[j + (size_t)ns * jp0] = c;
}

// ---------------------------------------------------------------------------
// convol cases 3 and 4 (full outputs), gathered per output element.
// Each reference update a(X,Y) += rkb(i,j,ip,jp)*d(U,V) has X in {i,ii},
// Y in {j,jj} and (U,V) the other pair members; terms carry conditions on
// ip/jp (convol, dbsr_lib_dbs.f90:2343-2466).
// ---------------------------------------------------------------------------
struct Term {
  int xs, ys;   // 0: first member (i or j), 1: second (ii or jj)
  int cond;     // 0 none, 1 ip-cond, 2 jp-cond, 3 both
};

__device__ inline int term_list(int icase, char si, char sj, Term* t) {
  // 's'/'l' conditions: 's' ip>1, 'l' ip!=ks
  if (si != 'n' && sj != 'n') {
    if (icase == 3) { t[0] = {0, 1, 0}; t[1] = {1, 1, 1}; t[2] = {0, 0, 2}; t[3] = {1, 0, 3}; }
    else { t[0] = {1, 0, 0}; t[1] = {0, 0, 1}; t[2] = {1, 1, 2}; t[3] = {0, 1, 3}; }
    return 4;
  }
  if (si == 'n' && sj == 'n') {
    t[0] = icase == 3 ? Term{0, 1, 0} : Term{1, 0, 0};
    return 1;
  }
  if (si == 'n') {  // ('n','l')
    if (icase == 3) { t[0] = {0, 1, 0}; t[1] = {0, 0, 2}; }
    else { t[0] = {1, 0, 0}; t[1] = {1, 1, 2}; }
    return 2;
  }
  // ('l','n')
  if (icase == 3) { t[0] = {0, 1, 0}; t[1] = {1, 1, 1}; }
  else { t[0] = {1, 0, 0}; t[1] = {0, 0, 1}; }
  return 2;
}

__device__ inline bool cond_ok(char s, int pf, int ks) { return s == 's' ? pf > 1 : pf != ks; }
__device__ inline int band_off(char s, int pf, int ks) { return s == 's' ? pf - 1 : pf - ks; }

__global__ void k_convol34(int ns, int ks, int kb, int icase, char si, char sj,
                           const double* __restrict__ rkb, const double* __restrict__ d, double* a) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)ns * ns) return;
  const int x = (int)(q % ns) + 1, y = (int)(q / ns) + 1;  // 1-based output a(x,y)
  Term terms[4];
  const int nt = term_list(icase, si, sj, terms);
  const int nci = band_cols(si, ks), ncj = band_cols(sj, ks);
  double acc = 0.0;
  for (int t = 0; t < nt; ++t) {
    const Term tm = terms[t];
    for (int jpf = 1; jpf <= ncj; ++jpf) {
      if ((tm.cond & 2) && !cond_ok(sj, jpf, ks)) continue;
      int joff = band_off(sj, jpf, ks);
      int j = tm.ys == 0 ? y : y - joff;
      int jlo, jhi;
      band_range(sj, ns, ks, jpf, &jlo, &jhi);
      if (j < jlo || j > jhi) continue;
      int jj = j + joff;
      int v = tm.ys == 0 ? jj : j;
      for (int ipf = 1; ipf <= nci; ++ipf) {
        if ((tm.cond & 1) && !cond_ok(si, ipf, ks)) continue;
        int ioff = band_off(si, ipf, ks);
        int i = tm.xs == 0 ? x : x - ioff;
        int ilo, ihi;
        band_range(si, ns, ks, ipf, &ilo, &ihi);
        if (i < ilo || i > ihi) continue;
        int ii = i + ioff;
        int u = tm.xs == 0 ? ii : i;
        double c = rkb[(i - 1) + (size_t)ns * ((j - 1) + (size_t)ns * ((ipf - 1) + (size_t)kb * (jpf - 1)))];
        acc += c * d[(u - 1) + (size_t)ns * (v - 1)];
      }
    }
  }
  a[q] = acc;
}

// ---------------------------------------------------------------------------
// SUM_AmB and quadr (single deterministic reductions)
// ---------------------------------------------------------------------------
__global__ void k_sum_amb(int ns, int ks, char sym, const double* __restrict__ a,
                          const double* __restrict__ b, double* out) {
  __shared__ double sh[1024];
  double s = 0.0;
  const int ncol = sym == 'x' ? ns : band_cols(sym, ks);
  for (int q = threadIdx.x; q < ns * ncol; q += blockDim.x) {
    int i = q % ns + 1, jf = q / ns + 1;
    bool ok;
    if (sym == 'x') ok = true;
    else {
      int lo, hi;
      band_range(sym, ns, ks, jf, &lo, &hi);
      ok = i >= lo && i <= hi;
    }
    if (ok) s += a[q] * b[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void k_quadr(int ns, int nv, int ks, int ksp, int ksq, int mpow,
                        const double* __restrict__ gr, const double* __restrict__ grw,
                        const double* __restrict__ pbsp, const double* __restrict__ qbsp,
                        const double* __restrict__ fi, const double* __restrict__ fj, double* out) {
  __shared__ double sh[1024];
  const int nw = nv + 1;
  double s = 0.0;
  for (int q = threadIdx.x; q < 2 * nv; q += blockDim.x) {
    int comp = q / nv, iv = q % nv;
    int kx = comp == 0 ? ksp : ksq;
    const double* xb = comp == 0 ? pbsp : qbsp;
    for (int m = 0; m < ks; ++m) {
      double g = ipow(gr[iv + nv * m], mpow) * grw[iv + nv * m];
      double vi = 0.0, vj = 0.0;
      for (int th = 0; th < kx; ++th) {
        double b = xb[iv + (size_t)nw * (m + (size_t)ks * th)];
        vi += b * fi[iv + th + (size_t)ns * comp];
        vj += b * fj[iv + th + (size_t)ns * comp];
      }
      s += g * vi * vj;
    }
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// UPDATE_HS (library form): pass 0 adds y/2 at (i1:i2,j1:j2); pass 1 adds
// (y/2)^T at (j1:j2,i1:i2).  Two launches keep the reference's order when
// the sub-blocks coincide.
__global__ void k_update_hs(int ns, int ks, int ms, char sym, int ii, int jj, int pass,
                            const double* __restrict__ d, double* hm) {
#pragma clang fp contract(off)
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)ns * ns) return;
  int a = (int)(q % ns), b = (int)(q / ns);
  // y(a,b) for pass 0, y(b,a) for pass 1
  int r = pass == 0 ? a : b, cidx = pass == 0 ? b : a;
  double y = 0.0;
  if (sym == 'x') y = d[r + (size_t)ns * cidx];
  else if (sym == 's') {
    int lo = r < cidx ? r : cidx, hi = r < cidx ? cidx : r;
    int jp = hi - lo;
    if (jp < ks) y = d[lo + (size_t)ns * jp];
  } else {
    int jp = cidx - r + ks - 1;  // 0-based column of 'n' band
    if (jp >= 0 && jp < 2 * ks - 1) y = d[r + (size_t)ns * jp];
  }
  y = y / 2.0;
  int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  if (pass == 0) hm[(i1 + a) + (size_t)ms * (j1 + b)] += y;
  else hm[(j1 + a) + (size_t)ms * (i1 + b)] += y;
}

__global__ void k_rk_sum(const double* __restrict__ p, double* o) {
#pragma clang fp contract(off)
  if (threadIdx.x == 0) o[0] = (((0.0 + p[0]) + p[1]) + p[2]) + p[3];
}

}  // namespace dbs

using namespace dbs;

static bool valid_sym(char s) { return s == 's' || s == 'l' || s == 'n'; }

namespace dbs {
int convol_launch(dbs_ctx* c, const double* rkb, int type, int icase, char si, char sj, const double* d,
                  double* a) {
  const int ns = c->ns, ks = c->ks, kb = kband(type, ks);
  if (icase == 1) {
    const int jchunk = 32;
    int nci = si == 'n' ? 2 * ks - 1 : ks, ncj = sj == 'n' ? 2 * ks - 1 : ks;
    int nch = (ns + jchunk - 1) / jchunk;
    size_t need = (size_t)ncj * nch * ns * nci;
    int rc = ensure_scratch(c, need);
    if (rc) return rc;
    dim3 grid(nci, ncj, nch);
    k_convol1_partial<<<grid, 256, 0, c->stream>>>(ns, ks, kb, si, sj, jchunk, rkb, d, c->scratch);
    k_convol1_reduce<<<(ns * nci + 255) / 256, 256, 0, c->stream>>>(ns, ks, si, ncj * nch, c->scratch, a);
  } else if (icase == 2) {
    int ncj = sj == 'n' ? 2 * ks - 1 : ks;
    size_t threads = (size_t)ns * ncj * 64;
    k_convol2<<<(unsigned)((threads + 255) / 256), 256, 0, c->stream>>>(ns, ks, kb, si, sj, rkb, d, a);
  } else {
    size_t n = (size_t)ns * ns;
    k_convol34<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(ns, ks, kb, icase, si, sj, rkb, d, a);
  }
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}
}  // namespace dbs

extern "C" {

int dbs_moments_dev(dbs_ctx* c, int which, int k, double* rkm) {
  DBS_CHECK(c && rkm, DBS_ERR_ARG, "dbs_moments_dev: null argument");
  DBS_CHECK(which >= 0 && which <= 3, DBS_ERR_ARG, "dbs_moments_dev: which not in 0..3");
  return launch_moments(c, which, k, rkm);
}

int dbs_diag_cells_dev(dbs_ctx* c, int type, int k, double* rkd) {
  DBS_CHECK(c && rkd, DBS_ERR_ARG, "dbs_diag_cells_dev: null argument");
  DBS_CHECK(type >= 1 && type <= 6, DBS_ERR_ARG, "dbs_diag_cells_dev: unknown type");
  DBS_CHECK(k >= 0, DBS_ERR_ARG, "dbs_diag_cells_dev: k < 0");
  return launch_diag(c, type, k, rkd);
}

int dbs_mrk(dbs_ctx* c, int type, int k, const double** rkb) {
  DBS_CHECK(c && rkb, DBS_ERR_ARG, "dbs_mrk: null argument");
  double* p = nullptr;
  int rc = build_rkb(c, type, k, &p);
  *rkb = p;
  return rc;
}

int dbs_mrk_host(dbs_ctx* c, int type, int k, double* out) {
  DBS_CHECK(c && out, DBS_ERR_ARG, "dbs_mrk_host: null argument");
  double* p = nullptr;
  int rc = build_rkb(c, type, k, &p);
  if (rc) return rc;
  int kb = kband(type, c->ks);
  size_t n = (size_t)c->ns * c->ns * kb * kb;
  DBS_HIP(hipMemcpyAsync(out, p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DBS_HIP(hipStreamSynchronize(c->stream));
  return DBS_OK;
}

int dbs_density_dev(dbs_ctx* c, const double* p1, const double* p2, char type, double* d) {
  DBS_CHECK(c && p1 && p2 && d, DBS_ERR_ARG, "dbs_density_dev: null argument");
  int ns = c->ns, ks = c->ks;
  int ncol = (type == 's' || type == 'u' || type == 'l') ? ks : (type == 'n' ? 2 * ks - 1 : ns);
  size_t n = (size_t)ns * ncol;
  k_density<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(ns, ks, type, p1, p2, d);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_convol_dev(dbs_ctx* c, int type, int k, int icase, char si, char sj, const double* d,
                   double* a) {
  DBS_CHECK(c && d && a, DBS_ERR_ARG, "dbs_convol_dev: null argument");
  DBS_CHECK(icase >= 1 && icase <= 4, DBS_ERR_ARG, "convol: unknown case");
  DBS_CHECK(valid_sym(si) && valid_sym(sj), DBS_ERR_ARG, "convol: unknown symmetry");
  DBS_CHECK(type >= DBS_PPQQ || (si != 'n' && sj != 'n'), DBS_ERR_ARG,
            "convol: 'n' storage needs an S^k array (2ks-1 band)");
  double* rkb = nullptr;
  int rc = build_rkb(c, type, k, &rkb);
  if (rc) return rc;
  return convol_launch(c, rkb, type, icase, si, sj, d, a);
}

int dbs_sum_amb_dev(dbs_ctx* c, const double* a, const double* b, char sym, double* out) {
  DBS_CHECK(c && a && b && out, DBS_ERR_ARG, "dbs_sum_amb_dev: null argument");
  DBS_CHECK(sym == 's' || sym == 'l' || sym == 'n' || sym == 'x', DBS_ERR_ARG,
            " SUM_AmB:  unknown symmetry ");
  k_sum_amb<<<1, 1024, 0, c->stream>>>(c->ns, c->ks, sym, a, b, out);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_quadr_dev(dbs_ctx* c, const double* fi, const double* fj, int m, double* out) {
  DBS_CHECK(c && fi && fj && out, DBS_ERR_ARG, "dbs_quadr_dev: null argument");
  k_quadr<<<1, 1024, 0, c->stream>>>(c->ns, c->nv, c->ks, c->ksp, c->ksq, m, c->gr, c->grw,
                                      c->pbsp, c->qbsp, fi, fj, out);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_update_hs_dev(dbs_ctx* c, double* hm, int ii, int jj, const double* d, char sym) {
  DBS_CHECK(c && hm && d, DBS_ERR_ARG, "dbs_update_hs_dev: null argument");
  DBS_CHECK((ii == 1 || ii == 2) && (jj == 1 || jj == 2), DBS_ERR_ARG, "unknown case in update_hs");
  DBS_CHECK(sym == 's' || sym == 'n' || sym == 'x', DBS_ERR_ARG, "update_hs: unknown symmetry");
  size_t n = (size_t)c->ns * c->ns;
  for (int pass = 0; pass < 2; ++pass)
    k_update_hs<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->ns, c->ks, c->ms, sym, ii, jj,
                                                                      pass, d, hm);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

// rk(): four densities, four case-2 convolutions, four band dot products
int dbs_rk_dev(dbs_ctx* c, const double* f1, const double* f2, const double* f3, const double* f4,
               int k, double* out) {
  DBS_CHECK(c && f1 && f2 && f3 && f4 && out, DBS_ERR_ARG, "dbs_rk_dev: null argument");
  const int ns = c->ns, ks = c->ks;
  const size_t nb = (size_t)ns * ks;
  // scratch layout: d1 d2 d3 d4 conv parts(4)
  int rc = ensure_scratch(c, nb * 5 + 8);
  if (rc) return rc;
  double* d1 = c->scratch;
  double* d2 = d1 + nb;
  double* d3 = d2 + nb;
  double* d4 = d3 + nb;
  double* cv = d4 + nb;
  double* parts = cv + nb;  // 4 scalars
  const int types[4] = {DBS_PPPP, DBS_QQQQ, DBS_QPQP, DBS_PQPQ};
  if ((rc = dbs_density_dev(c, f1, f3, 's', d1))) return rc;
  if ((rc = dbs_density_dev(c, f2, f4, 's', d2))) return rc;
  if ((rc = dbs_density_dev(c, f1 + ns, f3 + ns, 's', d3))) return rc;
  if ((rc = dbs_density_dev(c, f2 + ns, f4 + ns, 's', d4))) return rc;
  const double* left[4] = {d1, d3, d3, d1};
  const double* right[4] = {d2, d4, d2, d4};
  size_t threads = (size_t)ns * ks * 64;
  // each array is fetched just before its convolution: building the next
  // one may evict this one, and the cache synchronises the stream before it
  // frees an array, so no kernel can read a freed array
  for (int t = 0; t < 4; ++t) {
    double* rkb = nullptr;
    if ((rc = build_rkb(c, types[t], k, &rkb))) return rc;
    k_convol2<<<(unsigned)((threads + 255) / 256), 256, 0, c->stream>>>(ns, ks, ks, 's', 's',
                                                                         rkb, left[t], cv);
    k_sum_amb<<<1, 1024, 0, c->stream>>>(ns, ks, 's', cv, right[t], parts + t);
  }
  // rk = (((0 + S_pppp) + S_qqqq) + S_qpqp) + S_pqpq
  k_rk_sum<<<1, 64, 0, c->stream>>>(parts, out);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// rk() without any allocation: the four densities, the convolution and the
// four partial sums live in `scratch`, a fixed device array of this file
// (one copy per GPU, held until the process ends).  Its device address is
// taken once per device with hipGetSymbolAddress and kept in scratch_dev;
// host code never passes the array's own name to a kernel.  The kernels are
// those of dbs_rk_dev and overwrite every element they later read, so no
// clearing is needed and results are bit-identical to dbs_rk_dev.
// ---------------------------------------------------------------------------
static constexpr int kRkNoallocMaxBand = 32768;  // ns*ks
static constexpr size_t kRkNoallocDoubles = 5 * (size_t)kRkNoallocMaxBand + 8;
static __device__ __align__(256) double scratch[kRkNoallocDoubles];  // 1.25 MiB per GPU

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

int scratch_on(int device, double** out) {
  DBS_CHECK(device >= 0 && device < kMaxDevices, DBS_ERR_ARG, "rk (no-alloc): device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[device]) {
    int cur = 0;
    DBS_HIP(hipGetDevice(&cur));
    DBS_HIP(hipSetDevice(device));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("rk (no-alloc): hipGetSymbolAddress: ") +
                                                      hipGetErrorString(e));
    scratch_dev[device] = static_cast<double*>(p);
  }
  *out = scratch_dev[device];
  return DBS_OK;
}
}  // namespace

extern "C" {

int dbs_rk_noalloc_dev(dbs_ctx* c, const double* f1, const double* f2, const double* f3, const double* f4,
                       int k, double* out) {
  DBS_CHECK(c && f1 && f2 && f3 && f4 && out, DBS_ERR_ARG, "dbs_rk_noalloc_dev: null argument");
  const int ns = c->ns, ks = c->ks;
  const size_t nb = (size_t)ns * ks;
  DBS_CHECK(nb <= (size_t)kRkNoallocMaxBand, DBS_ERR_CAPACITY,
            "rk (no-alloc): ns*ks = " + std::to_string(nb) + " exceeds the fixed scratch (" +
                std::to_string(kRkNoallocMaxBand) + ")");
  const int types[4] = {DBS_PPPP, DBS_QQQQ, DBS_QPQP, DBS_PQPQ};
  const double* rkb[4];
  for (int t = 0; t < 4; ++t) {
    auto it = c->rkb_cache.find(types[t] * 1000 + k);
    DBS_CHECK(it != c->rkb_cache.end(), DBS_ERR_STATE,
              "rk (no-alloc): the R^k arrays of multipole " + std::to_string(k) +
                  " are not built; build types 1-4 first with dbs_mrk");
    rkb[t] = it->second;
  }
  double* s = nullptr;
  int rc = scratch_on(c->device, &s);
  if (rc) return rc;
  double *d1 = s, *d2 = d1 + nb, *d3 = d2 + nb, *d4 = d3 + nb, *cv = d4 + nb, *parts = cv + nb;
  const unsigned gd = (unsigned)((nb + 255) / 256);
  k_density<<<gd, 256, 0, c->stream>>>(ns, ks, 's', f1, f3, d1);
  k_density<<<gd, 256, 0, c->stream>>>(ns, ks, 's', f2, f4, d2);
  k_density<<<gd, 256, 0, c->stream>>>(ns, ks, 's', f1 + ns, f3 + ns, d3);
  k_density<<<gd, 256, 0, c->stream>>>(ns, ks, 's', f2 + ns, f4 + ns, d4);
  const double* left[4] = {d1, d3, d3, d1};
  const double* right[4] = {d2, d4, d2, d4};
  const size_t threads = nb * 64;
  for (int t = 0; t < 4; ++t) {
    k_convol2<<<(unsigned)((threads + 255) / 256), 256, 0, c->stream>>>(ns, ks, ks, 's', 's', rkb[t], left[t],
                                                                         cv);
    k_sum_amb<<<1, 1024, 0, c->stream>>>(ns, ks, 's', cv, right[t], parts + t);
  }
  k_rk_sum<<<1, 64, 0, c->stream>>>(parts, out);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// host-pointer forms: stage inputs into the context's device area, run the
// device routine, copy the result back and synchronise.
// ---------------------------------------------------------------------------
static int h2d(dbs_ctx* c, double* dst, const double* src, size_t n) {
  DBS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return DBS_OK;
}
static int d2h_sync(dbs_ctx* c, double* dst, const double* src, size_t n) {
  DBS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DBS_HIP(hipStreamSynchronize(c->stream));
  return DBS_OK;
}
static int band_ncol(char sym, int ns, int ks) {
  return sym == 'n' ? 2 * ks - 1 : (sym == 'x' ? ns : ks);
}

int dbs_density(dbs_ctx* c, const double* p1, const double* p2, char type, double* d) {
  DBS_CHECK(c && p1 && p2 && d, DBS_ERR_ARG, "dbs_density: null argument");
  const size_t ns = c->ns;
  const size_t nd = ns * (size_t)((type == 's' || type == 'u' || type == 'l') ? c->ks
                                  : (type == 'n' ? 2 * c->ks - 1 : c->ns));
  int rc = ensure_stage(c, 2 * ns + nd);
  if (rc) return rc;
  double *a = c->stage, *b = a + ns, *o = b + ns;
  if ((rc = h2d(c, a, p1, ns)) || (rc = h2d(c, b, p2, ns))) return rc;
  if ((rc = dbs_density_dev(c, a, b, type, o))) return rc;
  return d2h_sync(c, d, o, nd);
}

int dbs_convol(dbs_ctx* c, int type, int k, int icase, char si, char sj, const double* d, double* a) {
  DBS_CHECK(c && d && a, DBS_ERR_ARG, "dbs_convol: null argument");
  DBS_CHECK(icase >= 1 && icase <= 4, DBS_ERR_ARG, "convol: unknown case");
  const size_t ns = c->ns;
  size_t nin, nout;
  if (icase == 1) { nin = ns * band_ncol(sj, c->ns, c->ks); nout = ns * band_ncol(si, c->ns, c->ks); }
  else if (icase == 2) { nin = ns * band_ncol(si, c->ns, c->ks); nout = ns * band_ncol(sj, c->ns, c->ks); }
  else { nin = ns * ns; nout = ns * ns; }
  int rc = ensure_stage(c, nin + nout);
  if (rc) return rc;
  double *din = c->stage, *dout = din + nin;
  if ((rc = h2d(c, din, d, nin))) return rc;
  if ((rc = dbs_convol_dev(c, type, k, icase, si, sj, din, dout))) return rc;
  return d2h_sync(c, a, dout, nout);
}

int dbs_sum_amb(dbs_ctx* c, const double* a, const double* b, char sym, double* out) {
  DBS_CHECK(c && a && b && out, DBS_ERR_ARG, "dbs_sum_amb: null argument");
  const size_t n = (size_t)c->ns * band_ncol(sym, c->ns, c->ks);
  int rc = ensure_stage(c, 2 * n + 1);
  if (rc) return rc;
  double *x = c->stage, *y = x + n, *o = y + n;
  if ((rc = h2d(c, x, a, n)) || (rc = h2d(c, y, b, n))) return rc;
  if ((rc = dbs_sum_amb_dev(c, x, y, sym, o))) return rc;
  return d2h_sync(c, out, o, 1);
}

int dbs_rk(dbs_ctx* c, const double* f1, const double* f2, const double* f3, const double* f4,
           int k, double* out) {
  DBS_CHECK(c && f1 && f2 && f3 && f4 && out, DBS_ERR_ARG, "dbs_rk: null argument");
  const size_t n = 2 * (size_t)c->ns;
  int rc = ensure_stage(c, 4 * n + 1);
  if (rc) return rc;
  double* s = c->stage;
  const double* src[4] = {f1, f2, f3, f4};
  for (int q = 0; q < 4; ++q)
    if ((rc = h2d(c, s + q * n, src[q], n))) return rc;
  if ((rc = dbs_rk_dev(c, s, s + n, s + 2 * n, s + 3 * n, k, s + 4 * n))) return rc;
  return d2h_sync(c, out, s + 4 * n, 1);
}

int dbs_quadr(dbs_ctx* c, const double* fi, const double* fj, int m, double* out) {
  DBS_CHECK(c && fi && fj && out, DBS_ERR_ARG, "dbs_quadr: null argument");
  const size_t n = 2 * (size_t)c->ns;
  int rc = ensure_stage(c, 2 * n + 1);
  if (rc) return rc;
  double* s = c->stage;
  if ((rc = h2d(c, s, fi, n)) || (rc = h2d(c, s + n, fj, n))) return rc;
  if ((rc = dbs_quadr_dev(c, s, s + n, m, s + 2 * n))) return rc;
  return d2h_sync(c, out, s + 2 * n, 1);
}

}  // extern "C"

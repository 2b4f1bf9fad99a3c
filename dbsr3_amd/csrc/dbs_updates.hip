// dbs_updates.hip -- B-spline overlap matrices, band unfolding and the
// remaining matrix updates of the integral path:
//   ZINTYM overlaps fpbs/fqbs and the block matrix fppqq
//     (DBSR_HF/dbsr_lib_dbs.f90:4472-4497; Def_pbsp/Def_qbsp :1026, :1072;
//      Def_pq_arrays :1078-1102)
//   the Dirac one-electron operator with a finite nucleus: V_nuclear and
//     SKFUN (DBSR_HF/dbsr_lib_dbs.f90:220-310), Def_Vnucl (:2756-2780),
//     the derivative tables of Def_pbsp/Def_qbsp (:987-1074, zbsplvd with
//     two derivatives :4197-4351), MAT_dhl_pq with the Bloch term
//     (:1164-1206); exp() in the Fermi potential is the device library's
//   Full_mat_sym (DBSR_HF/dbsr_lib_zcom.f90:1247-1294)
//   UPDATE_HX / HW / HV / HB / ACF (DBSR_MAT3/updates.f90:79-106, 183-292)
//   O_data (DBSR_MAT3/O_data.f90:2-93) and the CI overlap SM(i,j) += CC
//     (DBSR_CI3/setup_matrix.f90:108-120)
// Loop and summation orders follow the reference with FMA contraction off,
// so every result is bit-identical to it.  The updates of one matrix element
// are applied by one thread in list order: no atomics, no reassociation.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>

#include "dbs_device.h"
#include "dbs_internal.h"


namespace dbs {

// ym(i,j) = sum_iv sum_m ygr(iv,m) b1(iv,m,ith) b2(iv,m,jth), i = iv+ith-1,
// j = iv+jth-1.  One thread per (i,j); intervals ascending, points in order
// (the reference's SUM of the elementwise product (ygr*b1)*b2).  ym is
// (mm,mm) column-major, only its leading (ns,ns) part is written.
__global__ void k_zintym(int nv, int ks, int ks1, int ks2, const double* __restrict__ b1,
                         const double* __restrict__ b2, const double* __restrict__ ygr, int ns, int mm,
                         double* __restrict__ ym) {
#pragma clang fp contract(off)
  const long long total = (long long)ns * ns;
  const long long nw = nv + 1;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % ns), j = (int)(q / ns);
    double acc = 0.0;
    const int lo = max(max(i - ks1 + 1, j - ks2 + 1), 0), hi = min(min(i, j), nv - 1);
    for (int iv = lo; iv <= hi; ++iv) {
      const int ith = i - iv, jth = j - iv;
      double s = 0.0;
      for (int m = 0; m < ks; ++m) {
        const double y = ygr[iv + (long long)nv * m];
        s = s + y * b1[iv + nw * (m + (long long)ks * ith)] * b2[iv + nw * (m + (long long)ks * jth)];
      }
      acc = acc + s;
    }
    ym[i + (long long)mm * j] = acc;
  }
}

// zero the off-diagonal (ns,ns) blocks of a (2ns,2ns) matrix
__global__ void k_zero_offdiag(int ns, int mm, double* __restrict__ y) {
  const long long total = (long long)mm * mm;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % mm), j = (int)(q / mm);
    if ((i >= ns) != (j >= ns)) y[q] = 0.0;
  }
}

// y(ns,ns) from storage sym 0 's', 1 'l', 2 'n', 3 'x'
__global__ void k_full_mat_sym(int ns, int ks, const double* __restrict__ x, int sym, double* __restrict__ y) {
  const long long total = (long long)ns * ns;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % ns), j = (int)(q / ns);
    double v = 0.0;
    if (sym == 3) {
      v = x[q];
    } else if (sym == 2) {
      const int jp = j - i + ks - 1;  // x(i, jp+1), j = i + (jp+1) - ks
      if (jp >= 0 && jp < 2 * ks - 1) v = x[i + (long long)ns * jp];
    } else {
      const int a = min(i, j), b = max(i, j), d = b - a;
      if (d < ks) v = sym == 0 ? x[a + (long long)ns * d]             // 's': x(a, d+1)
                               : x[b + (long long)ns * (ks - 1 - d)];  // 'l': x(b, ks-d)
    }
    y[q] = v;
  }
}

// block += C*d (trans 0) or C*d^T (trans 1)
__global__ void k_update_hx(int ms, double* __restrict__ blk, const double* __restrict__ d, int trans, double C) {
#pragma clang fp contract(off)
  const long long total = (long long)ms * ms;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % ms), j = (int)(q / ms);
    const double v = trans ? d[j + (long long)ms * i] : d[q];
    blk[q] = blk[q] + C * v;
  }
}

__device__ __forceinline__ double hw_term(double vi, double wj, double vj, double wi, int mode, double C) {
#pragma clang fp contract(off)
  const double x = vi * wj;
  if (mode == 0) return C * x;
  const double xt = vj * wi;  // TRANSPOSE(x)(i,j) = v(j)*w(i)
  if (mode == 1) return C * xt;
  return (C / 2.0) * (x + xt);
}

// block += C v w^T (mode 0), C (v w^T)^T (1), (C/2)(x + x^T) (2)
__global__ void k_update_hw(int ms, double* __restrict__ blk, const double* __restrict__ v,
                            const double* __restrict__ w, int mode, double C) {
#pragma clang fp contract(off)
  const long long total = (long long)ms * ms;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % ms), j = (int)(q / ms);
    blk[q] = blk[q] + hw_term(v[i], w[j], v[j], w[i], mode, C);
  }
}

// col += C v
__global__ void k_update_hv(int ms, double* __restrict__ col, const double* __restrict__ v, double C) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ms; i += gridDim.x * blockDim.x)
    col[i] = col[i] + C * v[i];
}

// out[tpos[t]] += vals[e] for the entries e of target t, in list order
__global__ void k_accumulate_list(int ntarget, const long long* __restrict__ tpos,
                                  const int* __restrict__ toff, const double* __restrict__ vals,
                                  double* __restrict__ out) {
#pragma clang fp contract(off)
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntarget; t += gridDim.x * blockDim.x) {
    double acc = out[tpos[t]];
    for (int e = toff[t]; e < toff[t + 1]; ++e) acc = acc + vals[e];
    out[tpos[t]] = acc;
  }
}

// vector of a list entry: idx >= 0 column idx of V (leading dimension ldv),
// idx < 0 column -idx-1 of W (leading dimension ms)
__device__ __forceinline__ const double* entry_vec(int idx, const double* V, int ldv, const double* W, int ms) {
  return idx >= 0 ? V + (long long)ldv * idx : W + (long long)ms * (-idx - 1);
}

// Block targets on blockIdx.y (strided), elements on x.  Each element
// applies its target's entries in list order: HX with the matrix d
// (mode = transpose) or HW with vectors v, w (mode 0/1/2).
__global__ void k_block_list(int ms, int ntarget, const int* __restrict__ tblk, const int* __restrict__ toff,
                             const MatUpd* __restrict__ ent, const double* __restrict__ d,
                             const double* __restrict__ V, int ldv, const double* __restrict__ W,
                             double* __restrict__ hch) {
#pragma clang fp contract(off)
  const long long total = (long long)ms * ms;
  for (int t = blockIdx.y; t < ntarget; t += gridDim.y) {
    double* blk = hch + total * tblk[t];
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
      const int i = (int)(q % ms), j = (int)(q / ms);
      double acc = blk[q];
      for (int e = toff[t]; e < toff[t + 1]; ++e) {
        const MatUpd u = ent[e];
        if (u.kind == MatUpd::HW) {
          const double* v = entry_vec(u.v, V, ldv, W, ms);
          const double* w = entry_vec(u.w, V, ldv, W, ms);
          acc = acc + hw_term(v[i], w[j], v[j], w[i], u.mode, u.C);
        } else {
          const double x = u.mode ? d[j + (long long)ms * i] : d[q];
          acc = acc + u.C * x;
        }
      }
      blk[q] = acc;
    }
  }
}

// column targets: col(tcol[t]) += C_e v_e in list order
__global__ void k_vec_list(int ms, int ntarget, const int* __restrict__ tcol, const int* __restrict__ toff,
                           const MatUpd* __restrict__ ent, const double* __restrict__ V, int ldv,
                           const double* __restrict__ W, double* __restrict__ hcp) {
#pragma clang fp contract(off)
  for (int t = blockIdx.y; t < ntarget; t += gridDim.y) {
    double* col = hcp + (long long)ms * tcol[t];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ms; i += gridDim.x * blockDim.x) {
      double acc = col[i];
      for (int e = toff[t]; e < toff[t + 1]; ++e) {
        const MatUpd u = ent[e];
        acc = acc + u.C * entry_vec(u.v, V, ldv, W, ms)[i];
      }
      col[i] = acc;
    }
  }
}

// Gen_Lvec for a list: W(:,q) = hd(kappa index qv.x) (P,Q)(orbital qv.y),
// columns summed in order (MATMUL)
__global__ void k_lvec(int ms, int nq, const int2* __restrict__ qv, const double* __restrict__ hd,
                       const double* __restrict__ pq, double* __restrict__ W) {
#pragma clang fp contract(off)
  const long long total = (long long)ms * nq;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int q = (int)(e / ms), r = (int)(e % ms);
    const double* A = hd + (long long)ms * ms * qv[q].x;
    const double* v = pq + (long long)ms * (qv[q].y - 1);
    double acc = 0.0;
    for (int c = 0; c < ms; ++c) acc += A[r + (long long)ms * c] * v[c];
    W[e] = acc;
  }
}

namespace {

unsigned grid_n(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (unsigned)g;
}

int sym_code(char sym) {
  switch (sym) {
    case 's': return 0;
    case 'l': return 1;
    case 'n': return 2;
    case 'x': return 3;
    default: return -1;
  }
}

// device copies of host lists, freed after the stream has drained
struct Uploads {
  std::vector<void*> p;
  hipStream_t s;
  explicit Uploads(hipStream_t st) : s(st) {}
  ~Uploads() {
    if (!p.empty()) (void)hipStreamSynchronize(s);
    for (void* x : p) (void)hipFree(x);
  }
  template <class T>
  int put(const std::vector<T>& h, T** d) {
    *d = nullptr;
    if (h.empty()) return DBS_OK;
    void* x = nullptr;
    if (hipMalloc(&x, sizeof(T) * h.size()) != hipSuccess) return fail(DBS_ERR_HIP, "list upload: hipMalloc");
    p.push_back(x);
    if (hipMemcpyAsync(x, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(DBS_ERR_HIP, "list upload: hipMemcpyAsync");
    *d = (T*)x;
    return DBS_OK;
  }
};

int finish(hipStream_t s, const char* who) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
  return DBS_OK;
}

// stable grouping of entries by target: order = permutation, tval = the
// distinct targets ascending, toff = their entry ranges in the permutation
void group_by(const std::vector<long long>& tgt, std::vector<int>& order, std::vector<long long>& tval,
              std::vector<int>& toff) {
  order.clear();
  for (size_t e = 0; e < tgt.size(); ++e)
    if (tgt[e] >= 0) order.push_back((int)e);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tgt[a] < tgt[b]; });
  tval.clear();
  toff.clear();
  for (size_t q = 0; q < order.size(); ++q) {
    if (q == 0 || tgt[order[q]] != tgt[order[q - 1]]) {
      tval.push_back(tgt[order[q]]);
      toff.push_back((int)q);
    }
  }
  toff.push_back((int)order.size());
}

// out[target[e]] += val[e] over the entries with target >= 0, list order
// kept inside each target
int accumulate_list(hipStream_t s, const std::vector<long long>& target, const double* val, double* out,
                    const char* who) {
  std::vector<int> order, toff;
  std::vector<long long> tpos;
  group_by(target, order, tpos, toff);
  if (tpos.empty()) return DBS_OK;
  std::vector<double> v(order.size());
  for (size_t q = 0; q < order.size(); ++q) v[q] = val[order[q]];
  Uploads up(s);
  long long* d_tpos;
  int* d_toff;
  double* d_v;
  int rc = up.put(tpos, &d_tpos);
  if (!rc) rc = up.put(toff, &d_toff);
  if (!rc) rc = up.put(v, &d_v);
  if (rc) return rc;
  const int nt = (int)tpos.size();
  k_accumulate_list<<<grid_n(nt, 256), 256, 0, s>>>(nt, d_tpos, d_toff, d_v, out);
  return finish(s, who);
}

}  // namespace

// icc(ich,jch) with the reference's checks (updates.f90:89-103); *b = 0
// when the block is not stored here
int chan_block(const dbs_mat_blocks* m, int ich, int jch, const char* who, int* b) {
  const std::string w(who);
  if (ich < 0) return fail(DBS_ERR_ARG, w + ": ich < 0");
  if (jch < 0) return fail(DBS_ERR_ARG, w + ": jch < 0");
  if (ich > m->nch) return fail(DBS_ERR_ARG, w + ": ich > nch");
  if (jch > m->nch) return fail(DBS_ERR_ARG, w + ": jch > nch");
  if (ich == 0 || jch == 0) return fail(DBS_ERR_ARG, w + ": channel index 0");
  const int i = m->icc[(ich - 1) + (long long)m->nch * (jch - 1)];
  if (i < 0 || i > m->nblocks) return fail(DBS_ERR_ARG, "UPDATE_HX: i out of range");
  *b = i;
  return DBS_OK;
}

int check_blocks(const dbs_mat_blocks* m) {
  DBS_CHECK(m, DBS_ERR_ARG, "matrix storage: null descriptor");
  DBS_CHECK(m->nch >= 0 && m->npert >= 0 && m->nblocks >= 0 && m->ncb >= 0 && m->nbb >= 0, DBS_ERR_ARG,
            "matrix storage: negative dimension");
  return DBS_OK;
}

// UPDATE_ACF of a list queued on the context stream without waiting: the
// tables (targets ascending, their entries in list order, values) go
// through the context's pinned list buffers into the scratch, which the
// stream order keeps behind the kernels that used the scratch before
int update_acf_queued(dbs_ctx* ctx, const dbs_mat_blocks* m, int n, const int* k, const int* ich, const int* jch,
                      const double* C) {
  DBS_CHECK(ctx && (n <= 0 || (k && ich && jch && C)), DBS_ERR_ARG, "Update_cf: null argument");
  int rc = check_blocks(m);
  if (rc || n <= 0) return rc;
  std::vector<long long> t(n);
  for (int e = 0; e < n; ++e) {
    if (k[e] > m->mk) return fail(DBS_ERR_ARG, "Update_cf: k > kmk ");
    if (k[e] < 0) return fail(DBS_ERR_ARG, "Update_cf: k < 0");
    int b = 0;
    if ((rc = chan_block(m, ich[e], jch[e], "Update_cf", &b))) return rc;
    t[e] = b == 0 ? -1 : (long long)(b - 1) + (long long)m->nblocks * k[e];
  }
  std::vector<int> order, toff;
  std::vector<long long> tpos;
  group_by(t, order, tpos, toff);
  if (tpos.empty()) return DBS_OK;
  const int nt = (int)tpos.size();
  const size_t nv = order.size();
  const size_t b_v = 8 * (size_t)nt, b_toff = b_v + 8 * nv, nbytes = (b_toff + 4 * ((size_t)nt + 1) + 15) & ~(size_t)15;
  if ((rc = ensure_scratch(ctx, nbytes / 8 + 2))) return rc;
  char* hb = nullptr;
  if ((rc = list_stage_begin(ctx, nbytes, &hb))) return rc;
  std::memcpy(hb, tpos.data(), 8 * (size_t)nt);
  double* hv = reinterpret_cast<double*>(hb + b_v);
  for (size_t q = 0; q < nv; ++q) hv[q] = C[order[q]];
  std::memcpy(hb + b_toff, toff.data(), 4 * ((size_t)nt + 1));
  char* base = reinterpret_cast<char*>(ctx->scratch);
  if (hipMemcpyAsync(base, hb, nbytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    return fail(DBS_ERR_HIP, "UPDATE_ACF: upload");
  if ((rc = list_stage_end(ctx))) return rc;
  k_accumulate_list<<<grid_n(nt, 256), 256, 0, ctx->stream>>>(nt, reinterpret_cast<const long long*>(base),
                                                               reinterpret_cast<const int*>(base + b_toff),
                                                               reinterpret_cast<const double*>(base + b_v),
                                                               m->acf_dev);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("UPDATE_ACF: ") + hipGetErrorString(e));
  return DBS_OK;
}

int hb_target(const dbs_mat_blocks* m, int ic, int jc, long long* t) {
  if (ic > m->npert || jc > m->npert || ic <= 0 || jc <= 0)
    return fail(DBS_ERR_ARG, "UPDATE_HB: index is out of range");
  const int i = m->ibb[(ic - 1) + (long long)m->npert * (jc - 1)];
  if (i < 0 || i > m->nbb) return fail(DBS_ERR_ARG, "UPDATE_HB: ic, jc out of range");
  *t = i - 1;  // -1: not stored
  return DBS_OK;
}

int hv_column(const dbs_mat_blocks* m, int ich, int ip, int* col) {
  if (ich < 1 || ich > m->nch) return fail(DBS_ERR_ARG, "UPDATE_HV: channel index out is of range");
  if (ip < 1 || ip > m->npert) return fail(DBS_ERR_ARG, "UPDATE_HV: bound index out is of range");
  const int i = m->icb[(ich - 1) + (long long)m->nch * (ip - 1)];
  if (i < 0 || i > m->ncb) return fail(DBS_ERR_ARG, "UPDATE_HV: ich, ip out of range");
  *col = i;
  return DBS_OK;
}

// Target_h (DBSR_MAT3/target_subs.f90:2-31): tgt(ij) += C for a stored
// block, ij the packed (max,min) pair of the channels' targets; tgt may be
// NULL (bookkeeping not kept)
int target_h(const dbs_mat_blocks* m, int ich, int jch, const int* iptar, double* tgt, double C) {
  if (ich < 0) return fail(DBS_ERR_ARG, "target_h: ich < 0");
  if (jch < 0) return fail(DBS_ERR_ARG, "target_h: jch < 0");
  if (ich > m->nch) return fail(DBS_ERR_ARG, "target_h: ich > nch");
  if (jch > m->nch) return fail(DBS_ERR_ARG, "target_h: jch > nch");
  DBS_CHECK(ich >= 1 && jch >= 1, DBS_ERR_ARG, "target_h: channel index 0");
  if (m->icc[(ich - 1) + (long long)m->nch * (jch - 1)] == 0) return DBS_OK;
  if (!tgt || !iptar) return DBS_OK;
  const int it = iptar[ich - 1], jt = iptar[jch - 1];
  const int hi = std::max(it, jt), lo = std::min(it, jt);
  DBS_CHECK(lo >= 1, DBS_ERR_ARG, "target_h: target index < 1");
  tgt[(long long)(hi - 1) * hi / 2 + lo - 1] += C;
  return DBS_OK;
}

// the updates of one list, grouped by matrix and target (stable), each
// target's entries applied in list order by one thread per element
int apply_mat3(dbs_ctx* ctx, const dbs_mat_blocks* m, const std::vector<MatUpd>& u, const double* V, int ldv,
               const double* W, const double* d, const char* who) {
  hipStream_t s = ctx->stream;
  const int ms = ctx->ms;
  // HB: scalar targets
  {
    std::vector<long long> t(u.size(), -1);
    std::vector<double> c(u.size());
    bool any = false;
    for (size_t e = 0; e < u.size(); ++e) {
      c[e] = u[e].C;
      if (u[e].kind == MatUpd::HB) { t[e] = u[e].target; any = any || t[e] >= 0; }
    }
    if (any) {
      int rc = accumulate_list(s, t, c.data(), m->hp_dev, who);
      if (rc) return rc;
    }
  }
  // HV: columns of hcp; HW/HX: blocks of hch
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<long long> t(u.size(), -1);
    for (size_t e = 0; e < u.size(); ++e) {
      const int k = u[e].kind;
      if (pass == 0 ? k == MatUpd::HV : (k == MatUpd::HW || k == MatUpd::HX)) t[e] = u[e].target;
    }
    std::vector<int> order, toff;
    std::vector<long long> tval;
    group_by(t, order, tval, toff);
    if (tval.empty()) continue;
    std::vector<int> tt(tval.begin(), tval.end());
    std::vector<MatUpd> es(order.size());
    for (size_t q = 0; q < order.size(); ++q) {
      es[q] = u[order[q]];
      if (pass == 1 && es[q].kind == MatUpd::HX) DBS_CHECK(d, DBS_ERR_ARG, std::string(who) + ": fppqq missing");
      if (es[q].kind != MatUpd::HX) {
        DBS_CHECK((es[q].v >= 0 ? V != nullptr : W != nullptr), DBS_ERR_ARG, std::string(who) + ": vector missing");
        if (es[q].kind == MatUpd::HW)
          DBS_CHECK((es[q].w >= 0 ? V != nullptr : W != nullptr), DBS_ERR_ARG, std::string(who) + ": vector missing");
      }
    }
    Uploads up(s);
    int *d_t, *d_o;
    MatUpd* d_e;
    int rc;
    if ((rc = up.put(tt, &d_t)) || (rc = up.put(toff, &d_o)) || (rc = up.put(es, &d_e))) return rc;
    const unsigned gy = (unsigned)std::min<size_t>(tt.size(), 65535);
    if (pass == 0) {
      k_vec_list<<<dim3(grid_n(ms, 256), gy), 256, 0, s>>>(ms, (int)tt.size(), d_t, d_o, d_e, V, ldv, W,
                                                          m->hcp_dev);
    } else {
      const unsigned gx = std::min(grid_n((long long)ms * ms, 256), std::max(1u, 262144u / gy));
      k_block_list<<<dim3(gx, gy), 256, 0, s>>>(ms, (int)tt.size(), d_t, d_o, d_e, d, V, ldv, W, m->hch_dev);
    }
    if ((rc = finish(s, who))) return rc;
  }
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// Dirac one-electron operator
// ---------------------------------------------------------------------------
// the constants of module zconst (DBSR_HF/dbsr_lib_zcom.f90:72-98)
constexpr double kCau = 137.03599976;
constexpr double kPi = 3.141592653589793238462643;

// S_k(x) = sum_{n>=1} (-1)^n e^{nx} / n^k, summed until the relative term
// falls below 10 eps (SKFUN)
__device__ inline double skfun(int k, double x) {
#pragma clang fp contract(off)
  if (x < -500.0) return 0.0;
  const double eps = 10.0 * 2.220446049250313e-16;
  const double base = -exp(x);
  double dnum = base, s = base;
  int n = 1;
  for (;;) {
    ++n;
    dnum = dnum * base;
    const double dk = ipow((double)n, k);
    const double delta = dnum / dk;
    s = s + delta;
    if (fabs(delta / s) <= eps) break;
  }
  return s;
}

// nuclear potential at r: point, uniform sphere or Fermi distribution
// (Parpia and Mohanty, Phys Rev A46, 3735 (1992)), as V_nuclear
__device__ inline double v_nuclear(const dbs_nucleus& nu, double r) {
#pragma clang fp contract(off)
  const double Z = nu.z;
  if (r <= 0.0) return 0.0;
  if (nu.model == DBS_NUCLEUS_POINT) return -Z / r;
  if (nu.model == DBS_NUCLEUS_UNIFORM) {
    const double b = nu.r_uniform;
    if (r > b) return -Z / r;
    return -(3.0 * Z) / (2.0 * b) * (1.0 - r * r / (3.0 * b * b));
  }
  const double a = nu.a_fermi, c = nu.c_fermi;
  const double ac = a / c, ac2 = ac * ac, ac3 = ac2 * ac;
  const double p = kPi * kPi * ac2, s = 6.0 * ac3 * skfun(3, -c / a);
  const double N = 1.0 + p - s;
  const double rc = r / c;
  double V;
  if (r <= c) {
    const double ra = (r - c) / a;
    const double s2 = ac2 * skfun(2, ra), s3 = ac3 * skfun(3, ra);
    V = (3.0 - rc * rc + p) / 2.0 - 3.0 * s2 - s / rc + 6.0 / rc * s3;
    V = -(Z * V) / (N * c);
  } else {
    const double ra = (c - r) / a;
    const double s2 = ac2 * skfun(2, ra), s3 = ac3 * skfun(3, ra);
    V = N + 6.0 * s3 + 3.0 * rc * s2;
    V = -Z / r * (V / N);
  }
  return V;
}

// per outer Gauss point: vr = V_nuclear(gr) * grw (VR_nucl of Def_Vnucl)
// and ygw = grm * grw (the 1/r weights of Def_pq_arrays)
__global__ void k_dirac_weights(int nv, int ks, const double* __restrict__ gr, const double* __restrict__ grw,
                                const double* __restrict__ grm, dbs_nucleus nu, double* __restrict__ vr,
                                double* __restrict__ ygw) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nv * ks) return;
  vr[q] = v_nuclear(nu, gr[q]) * grw[q];
  ygw[q] = grm[q] * grw[q];
}

// first derivatives of the kx B-splines that do not vanish on
// [T(left), T(left+1)) at x: bsplvd with two derivatives (order kx-1
// values, then the B-coefficients w31 of the derivative), tx 0-based
__device__ inline void bspline_d1(const double* __restrict__ tx, int kx, int left, double x, double* d) {
#pragma clang fp contract(off)
  double lo[DBS_MAXK];
  bspline_values(tx, kx - 1, left, x, lo);
  double w31[DBS_MAXK + 1][DBS_MAXK + 1];
  int jlow = 1;
  for (int i = 1; i <= kx; ++i) {
    for (int j = jlow; j <= kx; ++j) w31[j][i] = 0.0;
    jlow = i;
    w31[i][i] = 1.0;
  }
  const int kpimm = kx - 1;
  const double fkpimm = kpimm;
  int i = kx, il = 0;
  for (int ld = 1; ld <= kpimm; ++ld) {
    const int n = left - il;
    const double w1 = fkpimm / (tx[n + kpimm - 1] - tx[n - 1]);
    for (int j = 1; j <= i; ++j) w31[i][j] = (w31[i][j] - w31[i - 1][j]) * w1;
    ++il;
    --i;
  }
  // DB(j,2) = order kx-1 value j-1 for j = 2..kx
  for (int ii = 1; ii <= kx; ++ii) {
    double w2 = 0.0;
    for (int j = ii > 2 ? ii : 2; j <= kx; ++j) w2 += w31[j][ii] * lo[j - 2];
    d[ii - 1] = w2;
  }
}

// derivative table xbsd(nv+1,ks,kx) in the layout of pbsp/qbsp (row nv,
// the last knot, left zero: no ZINTYm term reads it)
__global__ void k_outer_bspline_d1(int nv, int ks, int kx, const double* __restrict__ tx,
                                   const double* __restrict__ gr, double* __restrict__ xbsd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = nv + 1;
  if (q >= nw * ks) return;
  const int iv = q % nw, m = q / nw;
  double d[DBS_MAXK];
  if (iv < nv) {
    bspline_d1(tx, kx, iv + kx, gr[iv + nv * m], d);
  } else {
    for (int j = 0; j < kx; ++j) d[j] = 0.0;
  }
  for (int j = 0; j < kx; ++j) xbsd[iv + (size_t)nw * (m + (size_t)ks * j)] = d[j];
}

// dhl(ms,ms): (1:1) <B|V|B>; (2:2) <B|V|B> - 2c^2 <B|B>; (2:1)
// c<B|d/dr|B> + kc<B|1/r|B>; (1:2) -c<B|d/dr|B> + kc<B|1/r|B>; then the
// Bloch term at the last P and Q splines
struct DiracMats {
  const double *fpbn, *fqbn, *fqbs, *fpqbs, *fqpbs, *fpqbsd, *fqpbsd;
};
__global__ void k_dhl_assemble(int ns, int nsp, int nsq, int kappa, DiracMats M, int bloch, double rn, double pnu,
                               double* __restrict__ dhl) {
#pragma clang fp contract(off)
  const int ms = 2 * ns;
  const long long total = (long long)ms * ms;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(q % ms), j = (int)(q / ms);
    const int ii = i % ns, jj = j % ns;
    const long long e = ii + (long long)ns * jj;
    double v;
    if (i < ns && j < ns) {
      v = M.fpbn[e];
    } else if (i >= ns && j >= ns) {
      const double sb = -2.0 * kCau * kCau;
      v = M.fqbn[e] + sb * M.fqbs[e];
    } else if (i >= ns) {
      const double sa = kCau, sb = (double)kappa * kCau;
      v = sa * M.fqpbsd[e] + sb * M.fqpbs[e];
    } else {
      const double sa = -kCau, sb = (double)kappa * kCau;
      v = sa * M.fpqbsd[e] + sb * M.fpqbs[e];
    }
    if (bloch) {
      const int bi = nsp - 1, bj = ns + nsq - 1;
      if (i == bi && j == bi) v = v - pnu * rn * kCau;
      if (i == bi && j == bj) v = v + pnu * kCau;
      if (i == bj && j == bi) v = v + (pnu - 1.0) * kCau;
      if (i == bj && j == bj) v = v - (pnu - 1.0) / rn * kCau;
    }
    dhl[q] = v;
  }
}

static int check_nucleus(const dbs_nucleus* nu) {
  DBS_CHECK(nu, DBS_ERR_ARG, "dirac: null nucleus");
  DBS_CHECK(nu->z > 0.0, DBS_ERR_ARG, "dirac: Z <= 0");
  DBS_CHECK(nu->model == DBS_NUCLEUS_POINT || nu->model == DBS_NUCLEUS_UNIFORM || nu->model == DBS_NUCLEUS_FERMI,
            DBS_ERR_ARG, "dirac: unknown charge distribution");
  if (nu->model == DBS_NUCLEUS_UNIFORM) DBS_CHECK(nu->r_uniform > 0.0, DBS_ERR_ARG, "dirac: r_uniform <= 0");
  if (nu->model == DBS_NUCLEUS_FERMI)
    DBS_CHECK(nu->a_fermi > 0.0 && nu->c_fermi > 0.0, DBS_ERR_ARG, "dirac: Fermi a, c must be > 0");
  return DBS_OK;
}

}  // namespace dbs

using namespace dbs;

extern "C" {

int dbs_bspline_overlaps_dev(dbs_ctx* ctx, double* fpbs_dev, double* fqbs_dev, double* fppqq_dev) {
  DBS_CHECK(ctx, DBS_ERR_ARG, "dbs_bspline_overlaps_dev: null context");
  const int ns = ctx->ns, ms = ctx->ms, nv = ctx->nv, ks = ctx->ks;
  hipStream_t s = ctx->stream;
  const unsigned g = grid_n((long long)ns * ns, 256);
  if (fpbs_dev) k_zintym<<<g, 256, 0, s>>>(nv, ks, ctx->ksp, ctx->ksp, ctx->pbsp, ctx->pbsp, ctx->grw, ns, ns, fpbs_dev);
  if (fqbs_dev) k_zintym<<<g, 256, 0, s>>>(nv, ks, ctx->ksq, ctx->ksq, ctx->qbsp, ctx->qbsp, ctx->grw, ns, ns, fqbs_dev);
  if (fppqq_dev) {
    // fpbs in the PP block, fqbs in the QQ block, zeros elsewhere
    k_zero_offdiag<<<grid_n((long long)ms * ms, 256), 256, 0, s>>>(ns, ms, fppqq_dev);
    k_zintym<<<g, 256, 0, s>>>(nv, ks, ctx->ksp, ctx->ksp, ctx->pbsp, ctx->pbsp, ctx->grw, ns, ms, fppqq_dev);
    k_zintym<<<g, 256, 0, s>>>(nv, ks, ctx->ksq, ctx->ksq, ctx->qbsp, ctx->qbsp, ctx->grw, ns, ms,
                               fppqq_dev + ns + (size_t)ms * ns);
  }
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_full_mat_sym_dev(dbs_ctx* ctx, const double* x_dev, char sym, double* y_dev) {
  DBS_CHECK(ctx && x_dev && y_dev, DBS_ERR_ARG, "Full_mat_sym: null argument");
  const int sc = sym_code(sym);
  DBS_CHECK(sc >= 0, DBS_ERR_ARG, "Full_mat_sym: unknown storage mode");
  const int ns = ctx->ns;
  k_full_mat_sym<<<grid_n((long long)ns * ns, 256), 256, 0, ctx->stream>>>(ns, ctx->ks, x_dev, sc, y_dev);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_update_hx_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int ich, int jch, const double* d_dev, double C) {
  DBS_CHECK(ctx && d_dev, DBS_ERR_ARG, "UPDATE_HX: null argument");
  int rc = check_blocks(m), b = 0;
  if (!rc) rc = chan_block(m, ich, jch, "UPDATE_HX", &b);
  if (rc) return rc;
  if (b == 0) return DBS_OK;
  const int ms = ctx->ms;
  k_update_hx<<<grid_n((long long)ms * ms, 256), 256, 0, ctx->stream>>>(
      ms, m->hch_dev + (size_t)ms * ms * (b - 1), d_dev, ich >= jch ? 0 : 1, C);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_update_hw_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int ich, int jch, const double* v_dev,
                      const double* w_dev, double C) {
  DBS_CHECK(ctx && v_dev && w_dev, DBS_ERR_ARG, "UPDATE_HW: null argument");
  int rc = check_blocks(m), b = 0;
  if (!rc) rc = chan_block(m, ich, jch, "UPDATE_HW", &b);
  if (rc) return rc;
  if (b == 0) return DBS_OK;
  const int ms = ctx->ms;
  const int mode = ich > jch ? 0 : (jch > ich ? 1 : 2);
  k_update_hw<<<grid_n((long long)ms * ms, 256), 256, 0, ctx->stream>>>(
      ms, m->hch_dev + (size_t)ms * ms * (b - 1), v_dev, w_dev, mode, C);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_update_hv_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int ich, int ip, const double* v_dev, double C) {
  DBS_CHECK(ctx && v_dev, DBS_ERR_ARG, "UPDATE_HV: null argument");
  int rc = check_blocks(m), col = 0;
  if (!rc) rc = hv_column(m, ich, ip, &col);
  if (rc) return rc;
  if (col == 0) return DBS_OK;
  const int ms = ctx->ms;
  k_update_hv<<<grid_n(ms, 256), 256, 0, ctx->stream>>>(ms, m->hcp_dev + (size_t)ms * (col - 1), v_dev, C);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_update_hb_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int n, const int* ic, const int* jc,
                      const double* c) {
  DBS_CHECK(ctx && (n <= 0 || (ic && jc && c)), DBS_ERR_ARG, "UPDATE_HB: null argument");
  int rc = check_blocks(m);
  if (rc || n <= 0) return rc;
  std::vector<long long> t(n);
  for (int e = 0; e < n; ++e)
    if ((rc = hb_target(m, ic[e], jc[e], &t[e]))) return rc;
  return accumulate_list(ctx->stream, t, c, m->hp_dev, "UPDATE_HB");
}

int dbs_update_acf_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int n, const int* k, const int* ich,
                       const int* jch, const double* C) {
  DBS_CHECK(ctx && (n <= 0 || (k && ich && jch && C)), DBS_ERR_ARG, "Update_cf: null argument");
  int rc = check_blocks(m);
  if (rc || n <= 0) return rc;
  std::vector<long long> t(n);
  for (int e = 0; e < n; ++e) {
    if (k[e] > m->mk) return fail(DBS_ERR_ARG, "Update_cf: k > kmk ");
    if (k[e] < 0) return fail(DBS_ERR_ARG, "Update_cf: k < 0");
    int b = 0;
    if ((rc = chan_block(m, ich[e], jch[e], "Update_cf", &b))) return rc;
    t[e] = b == 0 ? -1 : (long long)(b - 1) + (long long)m->nblocks * k[e];
  }
  return accumulate_list(ctx->stream, t, C, m->acf_dev, "UPDATE_ACF");
}

int dbs_o_data_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int itype, int n, const int* k1, const int* k2,
                   const double* c, double eps_c, int ibo, const int* ipbs, dbs_kbort_fn kbort,
                   void* kbort_user, const double* vch_dev, int nvch, const double* fppqq_dev, int iitar,
                   const int* iptar, double* otarg) {
  DBS_CHECK(ctx && (n <= 0 || (k1 && k2 && c)), DBS_ERR_ARG, "O_data: null argument");
  int rc = check_blocks(m);
  if (rc || n <= 0) return rc;
  DBS_CHECK(itype >= 1 && itype <= 4, DBS_ERR_ARG, "O_data: unknown structure (itype) ");
  // orbital io = i1*ibo + i2 -> (i1, 0-based column KBORT(i1,i2)-1 of V_ch)
  auto vcol = [&](int io, int* col, int* i1) -> int {
    DBS_CHECK(ibo > 0 && ipbs && kbort && vch_dev, DBS_ERR_ARG, "O_data: bound-orbital tables missing");
    *i1 = io / ibo;
    DBS_CHECK(*i1 >= 1, DBS_ERR_ARG, "O_data: orbital index < 1");
    const int k = kbort(kbort_user, io / ibo, io % ibo);
    if (k == 0) return fail(DBS_ERR_ARG, "O: KBORT=0");
    DBS_CHECK(k >= 1 && k <= nvch, DBS_ERR_ARG, "O_data: KBORT column out of range");
    *col = k - 1;
    return DBS_OK;
  };
  std::vector<MatUpd> u;
  u.reserve(n);
  for (int e = 0; e < n; ++e) {
    const double C = c[e];
    if (std::fabs(C) < eps_c) continue;
    if (itype == 1) {  // <ic|jc>: UPDATE_HB
      long long t = -1;
      if ((rc = hb_target(m, k1[e], k2[e], &t))) return rc;
      u.push_back({MatUpd::HB, 0, 0, 0, t, C});
    } else if (itype == 2) {  // <i|.> jc: UPDATE_HV(ipbs(i1), jc, V_ch(:,KBORT))
      int vc = 0, i1 = 0, hc = 0;
      if ((rc = vcol(k1[e], &vc, &i1))) return rc;
      if ((rc = hv_column(m, ipbs[i1 - 1], k2[e], &hc))) return rc;
      u.push_back({MatUpd::HV, 0, vc, 0, (long long)hc - 1, C});
    } else if (itype == 3) {  // <i|.><.|j>: UPDATE_HW
      int vi = 0, i1 = 0, wj = 0, j1 = 0, b = 0;
      if ((rc = vcol(k1[e], &vi, &i1)) || (rc = vcol(k2[e], &wj, &j1))) return rc;
      const int ich = ipbs[i1 - 1], jch = ipbs[j1 - 1];
      if ((rc = chan_block(m, ich, jch, "UPDATE_HW", &b))) return rc;
      u.push_back({MatUpd::HW, ich > jch ? 0 : (jch > ich ? 1 : 2), vi, wj, (long long)b - 1, C});
    } else {  // <i|j>: Target_h(ich,jch,0,C) and, for iitar > 1, UPDATE_HX(fppqq)
      DBS_CHECK(ibo > 0 && ipbs, DBS_ERR_ARG, "O_data: bound-orbital tables missing");
      const int io = k1[e], i1 = io / ibo, i2 = io % ibo;
      DBS_CHECK(i1 >= 1 && i2 >= 1, DBS_ERR_ARG, "O_data: orbital index < 1");
      const int ich = ipbs[i1 - 1], jch = ipbs[i2 - 1];
      if ((rc = target_h(m, ich, jch, iptar, otarg, C))) return rc;
      if (iitar > 1 && ich != jch) {
        int b = 0;
        if ((rc = chan_block(m, ich, jch, "UPDATE_HX", &b))) return rc;
        u.push_back({MatUpd::HX, ich >= jch ? 0 : 1, 0, 0, (long long)b - 1, C});
      }
    }
  }
  return apply_mat3(ctx, m, u, vch_dev, ctx->ms, nullptr, fppqq_dev, "O_data");
}

int dbs_ldata_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, const dbs_ldata_aux* aux, int itype, int n, const int* k1,
                  const int* k2, const int* k3, const int* k4, const double* c, double eps_c, const double* pq_dev) {
  DBS_CHECK(ctx && aux && aux->mat3 && (n <= 0 || (k1 && k2 && k3 && k4 && c)), DBS_ERR_ARG,
            "L_data: null argument");
  int rc = check_blocks(m);
  if (rc || n <= 0) return rc;
  DBS_CHECK(itype >= 1 && itype <= 3, DBS_ERR_ARG, " L_data: uknown itype ");
  const dbs_mat3_aux* X = aux->mat3;
  const int ibo = X->ibo, norb = aux->norb;
  auto ipb = [&](int o) -> int { return X->ipbs[o - 1]; };
  auto orb_ok = [&](int o) { return o >= 1 && o <= norb; };
  // io = i1*ibo + i2 -> column KBORT(i1,i2)-1 of V_ch (L_data.f90:58-60)
  auto vcol = [&](int io, int* col, int* i1) -> int {
    DBS_CHECK(ibo > 0 && X->ipbs && X->kbort && X->vch_dev, DBS_ERR_ARG, "L_data: bound-orbital tables missing");
    *i1 = io / ibo;
    DBS_CHECK(orb_ok(*i1), DBS_ERR_ARG, "L_data: orbital index out of range");
    const int k = X->kbort(X->kbort_user, io / ibo, io % ibo);
    if (k == 0) return fail(DBS_ERR_ARG, "O: KBORT=0");
    DBS_CHECK(k >= 1 && k <= X->nvch, DBS_ERR_ARG, "L_data: KBORT column out of range");
    *col = k - 1;
    return DBS_OK;
  };
  auto kap_index = [&](int o, int* q) -> int {  // Ipointer(nkap, kap_list, kbs(o))
    DBS_CHECK(aux->kbs && aux->kap_list && aux->hd_dev, DBS_ERR_ARG, "L_data: one-electron operators missing");
    for (int t = 0; t < aux->nkap; ++t)
      if (aux->kap_list[t] == aux->kbs[o - 1]) { *q = t; return DBS_OK; }
    *q = -1;
    return DBS_OK;
  };
  std::vector<MatUpd> u;
  if (itype == 1) {
    DBS_CHECK(aux->lval, DBS_ERR_ARG, "L_data: Lval table missing");
    for (int e = 0; e < n; ++e) {
      double C = c[e];
      if (std::fabs(C) < eps_c) continue;
      DBS_CHECK(orb_ok(k1[e]) && orb_ok(k2[e]), DBS_ERR_ARG, "L_data: orbital index out of range");
      C = C * aux->lval[(k1[e] - 1) + (size_t)norb * (k2[e] - 1)];
      const int ic = k3[e], jc = k4[e];
      if (ic > 0 && jc > 0) {  // L(..) ic,jc
        long long t = -1;
        if ((rc = hb_target(m, ic, jc, &t))) return rc;
        u.push_back({MatUpd::HB, 0, 0, 0, t, C});
      } else if (ic < 0 && jc > 0) {  // L(..) <i|.> jc
        int vc = 0, i1 = 0, hc = 0;
        if ((rc = vcol(-ic, &vc, &i1))) return rc;
        if ((rc = hv_column(m, ipb(i1), jc, &hc))) return rc;
        u.push_back({MatUpd::HV, 0, vc, 0, (long long)hc - 1, C});
      } else if (ic < 0 && jc < 0) {  // L(..) <i|.> <.|j>
        int vi = 0, i1 = 0, wj = 0, j1 = 0, b = 0;
        if ((rc = vcol(-ic, &vi, &i1)) || (rc = vcol(-jc, &wj, &j1))) return rc;
        const int ich = ipb(i1), jch = ipb(j1);
        if ((rc = chan_block(m, ich, jch, "UPDATE_HW", &b))) return rc;
        u.push_back({MatUpd::HW, ich > jch ? 0 : (jch > ich ? 1 : 2), vi, wj, (long long)b - 1, C});
      } else if (ic < 0 && jc == 0) {  // L(..) <i|j>
        DBS_CHECK(ibo > 0 && X->ipbs, DBS_ERR_ARG, "L_data: bound-orbital tables missing");
        const int io = -ic, i1 = io / ibo, i2 = io % ibo;
        DBS_CHECK(orb_ok(i1) && orb_ok(i2), DBS_ERR_ARG, "L_data: orbital index out of range");
        const int ich = ipb(i1), jch = ipb(i2);
        if (X->iitar > 0 && ich != jch) {
          int b = 0;
          if ((rc = chan_block(m, ich, jch, "UPDATE_HX", &b))) return rc;
          u.push_back({MatUpd::HX, ich >= jch ? 0 : 1, 0, 0, (long long)b - 1, C});
        }
        if ((rc = target_h(m, ich, jch, X->iptar, X->htarg, C))) return rc;
      } else {
        return fail(DBS_ERR_ARG, "L_data:  wrong combination of ic,jc for itype=1");
      }
    }
    return apply_mat3(ctx, m, u, X->vch_dev, ctx->ms, nullptr, X->fppqq_dev, "L_data");
  }
  if (itype == 2) {
    // Gen_Lvec(i1, v): v = hd(kappa of i1) (P,Q)(i1), one column per orbital
    DBS_CHECK(pq_dev, DBS_ERR_ARG, "L_data: orbitals missing");
    std::vector<int> orbs;
    std::vector<int2> qv;
    std::vector<int> col_of(norb + 1, -1);
    for (int e = 0; e < n; ++e) {
      if (std::fabs(c[e]) < eps_c) continue;
      const int i1 = k1[e];
      DBS_CHECK(orb_ok(i1), DBS_ERR_ARG, "L_data: orbital index out of range");
      if (col_of[i1] >= 0) continue;
      int q = -1;
      if ((rc = kap_index(i1, &q))) return rc;
      DBS_CHECK(q >= 0, DBS_ERR_ARG, "L_data: unknown kappa, itype=2");
      col_of[i1] = (int)qv.size();
      qv.push_back(make_int2(q, i1));
    }
    for (int e = 0; e < n; ++e) {
      const double C = c[e];
      if (std::fabs(C) < eps_c) continue;
      const int ich = k2[e], jc = k3[e], wq = -(col_of[k1[e]] + 1);
      if (jc > 0) {  // L(.i) jc
        int hc = 0;
        if ((rc = hv_column(m, ich, jc, &hc))) return rc;
        u.push_back({MatUpd::HV, 0, wq, 0, (long long)hc - 1, C});
      } else if (jc < 0) {  // L(.i) <.|j>
        int wj = 0, j1 = 0, b = 0;
        if ((rc = vcol(-jc, &wj, &j1))) return rc;
        const int jch = ipb(j1);
        if ((rc = chan_block(m, ich, jch, "UPDATE_HW", &b))) return rc;
        u.push_back({MatUpd::HW, ich > jch ? 0 : (jch > ich ? 1 : 2), wq, wj, (long long)b - 1, C});
      } else {
        return fail(DBS_ERR_ARG, "L_data:  wrong value jc = 0 for itype=2");
      }
    }
    if (u.empty()) return DBS_OK;
    double* W = nullptr;
    if (hipMalloc(&W, sizeof(double) * (size_t)ctx->ms * qv.size()) != hipSuccess)
      return fail(DBS_ERR_HIP, "L_data: hipMalloc");
    int2* d_qv = nullptr;
    if (hipMalloc(&d_qv, sizeof(int2) * qv.size()) != hipSuccess) {
      (void)hipFree(W);
      return fail(DBS_ERR_HIP, "L_data: hipMalloc");
    }
    hipStream_t s = ctx->stream;
    rc = DBS_OK;
    if (hipMemcpyAsync(d_qv, qv.data(), sizeof(int2) * qv.size(), hipMemcpyHostToDevice, s) != hipSuccess)
      rc = fail(DBS_ERR_HIP, "L_data: hipMemcpyAsync");
    if (!rc) {
      k_lvec<<<grid_n((long long)ctx->ms * qv.size(), 256), 256, 0, s>>>(ctx->ms, (int)qv.size(), d_qv, aux->hd_dev,
                                                                          pq_dev, W);
      rc = apply_mat3(ctx, m, u, X->vch_dev, ctx->ms, W, X->fppqq_dev, "L_data");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(W);
    (void)hipFree(d_qv);
    return rc;
  }
  // itype 3: L(ij) -> UPDATE_HX(ich,jch,hd(kappa)) or the target overlap;
  // one apply per kappa (a block's channels fix its kappa)
  std::vector<std::vector<MatUpd>> per(aux->nkap);
  for (int e = 0; e < n; ++e) {
    const double C = c[e];
    if (std::fabs(C) < eps_c) continue;
    const int i1 = k1[e], j1 = k2[e];
    DBS_CHECK(orb_ok(i1) && orb_ok(j1) && X->ipbs, DBS_ERR_ARG, "L_data: orbital index out of range");
    const int ich = ipb(i1), jch = ipb(j1);
    DBS_CHECK(aux->kbs, DBS_ERR_ARG, "L_data: one-electron operators missing");
    if (aux->kbs[i1 - 1] != aux->kbs[j1 - 1]) return fail(DBS_ERR_ARG, "L_data: different kappa, itype=3");
    int q = -1;
    if ((rc = kap_index(i1, &q))) return rc;
    if (q < 0) return fail(DBS_ERR_ARG, "L_data: unknown kappa, itype=3");
    if (ich == jch || X->iitar > 1) {
      int b = 0;
      if ((rc = chan_block(m, ich, jch, "UPDATE_HX", &b))) return rc;
      per[q].push_back({MatUpd::HX, ich >= jch ? 0 : 1, 0, 0, (long long)b - 1, C});
    } else if (aux->check_target == 1) {
      if ((rc = target_h(m, ich, jch, X->iptar, aux->otarg, C))) return rc;
    }
  }
  const size_t mm = (size_t)ctx->ms * ctx->ms;
  for (int q = 0; q < aux->nkap; ++q)
    if (!per[q].empty() && (rc = apply_mat3(ctx, m, per[q], nullptr, ctx->ms, nullptr, aux->hd_dev + mm * q, "L_data")))
      return rc;
  return DBS_OK;
}

int dbs_overlap_ci_dev(dbs_ctx* ctx, int n, const int* ic, const int* jc, const double* cc, double eps_c,
                       int nzero, double* SM_dev, int ldsm) {
  DBS_CHECK(ctx && (n <= 0 || (ic && jc && cc)) && SM_dev, DBS_ERR_ARG, "overlap: null argument");
  DBS_CHECK(ldsm >= 1, DBS_ERR_ARG, "overlap: ldsm < 1");
  if (n <= 0) return DBS_OK;
  std::vector<long long> t(n, -1);
  for (int e = 0; e < n; ++e) {
    if (std::fabs(cc[e]) < eps_c) continue;  // CC below eps_C is skipped
    DBS_CHECK(ic[e] >= 1 && ic[e] <= ldsm && jc[e] >= 1, DBS_ERR_ARG, "overlap: index out of range");
    if (jc[e] <= nzero) t[e] = (long long)(ic[e] - 1) + (long long)ldsm * (jc[e] - 1);
  }
  return accumulate_list(ctx->stream, t, cc, SM_dev, "overlap");
}

}  // extern "C"

int dbs_vnucl_dev(dbs_ctx* ctx, const dbs_nucleus* nuc, double* vr_dev) {
  DBS_CHECK(ctx && vr_dev, DBS_ERR_ARG, "dbs_vnucl_dev: null argument");
  int rc = check_nucleus(nuc);
  if (rc) return rc;
  const int nv = ctx->nv, ks = ctx->ks;
  rc = ensure_scratch(ctx, (size_t)nv * ks);
  if (rc) return rc;
  k_dirac_weights<<<grid_n((long long)nv * ks, 256), 256, 0, ctx->stream>>>(nv, ks, ctx->gr, ctx->grw, ctx->grm, *nuc,
                                                                        vr_dev, ctx->scratch);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int dbs_dhl_pq_dev(dbs_ctx* ctx, const dbs_nucleus* nuc, int kappa, int bloch, double rb, double pnu,
                   double* dhl_dev) {
  DBS_CHECK(ctx && dhl_dev, DBS_ERR_ARG, "dbs_dhl_pq_dev: null argument");
  DBS_CHECK(kappa != 0, DBS_ERR_ARG, "dbs_dhl_pq_dev: kappa = 0");
  int rc = check_nucleus(nuc);
  if (rc) return rc;
  const int ns = ctx->ns, nv = ctx->nv, ks = ctx->ks, ksp = ctx->ksp, ksq = ctx->ksq;
  const size_t nk = (size_t)nv * ks, nwk = (size_t)(nv + 1) * ks, nn = (size_t)ns * ns;
  // scratch: vr, ygw | pbsd, qbsd | 7 matrices
  rc = ensure_scratch(ctx, 2 * nk + nwk * (ksp + ksq) + 7 * nn);
  if (rc) return rc;
  double* vr = ctx->scratch;
  double* ygw = vr + nk;
  double* pbsd = ygw + nk;
  double* qbsd = pbsd + nwk * ksp;
  double* mats = qbsd + nwk * ksq;
  double* fpbn = mats;
  double* fqbn = fpbn + nn;
  double* fqbs = fqbn + nn;
  double* fpqbs = fqbs + nn;
  double* fqpbs = fpqbs + nn;
  double* fpqbsd = fqpbs + nn;
  double* fqpbsd = fpqbsd + nn;
  hipStream_t s = ctx->stream;
  k_dirac_weights<<<grid_n((long long)nk, 256), 256, 0, s>>>(nv, ks, ctx->gr, ctx->grw, ctx->grm, *nuc, vr, ygw);
  k_outer_bspline_d1<<<grid_n((long long)nwk, 256), 256, 0, s>>>(nv, ks, ksp, ctx->tp, ctx->gr, pbsd);
  k_outer_bspline_d1<<<grid_n((long long)nwk, 256), 256, 0, s>>>(nv, ks, ksq, ctx->tq, ctx->gr, qbsd);
  const unsigned g = grid_n((long long)nn, 256);
  // Def_Vnucl
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksp, ksp, ctx->pbsp, ctx->pbsp, vr, ns, ns, fpbn);
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksq, ksq, ctx->qbsp, ctx->qbsp, vr, ns, ns, fqbn);
  // Def_qbsp's overlap and Def_pq_arrays
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksq, ksq, ctx->qbsp, ctx->qbsp, ctx->grw, ns, ns, fqbs);
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksp, ksq, ctx->pbsp, ctx->qbsp, ygw, ns, ns, fpqbs);
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksq, ksp, ctx->qbsp, ctx->pbsp, ygw, ns, ns, fqpbs);
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksp, ksq, ctx->pbsp, qbsd, ctx->grw, ns, ns, fpqbsd);
  k_zintym<<<g, 256, 0, s>>>(nv, ks, ksq, ksp, ctx->qbsp, pbsd, ctx->grw, ns, ns, fqpbsd);
  // RN = (RB + k) / (2 tmax c), tmax = t(ns+1)
  const double tmax = ctx->h_t[ns];
  const double rn = (rb + (double)kappa) / (2.0 * tmax * kCau);
  DiracMats M{fpbn, fqbn, fqbs, fpqbs, fqpbs, fpqbsd, fqpbsd};
  k_dhl_assemble<<<grid_n(4LL * (long long)nn, 256), 256, 0, s>>>(ns, ctx->nsp, ctx->nsq, kappa, M, bloch ? 1 : 0, rn,
                                                              pnu, dhl_dev);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

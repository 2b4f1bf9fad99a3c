// dbs_mat3.hip -- DBSR_MAT3's structure drivers on the GPU:
//   R_data itype 1..4 (DBSR_MAT3/R_data.f90:2-290) and S_data itype 1..9,
//   atype 0/1 (DBSR_MAT3/S_data.f90:2-287), with BAV (called at
//   R_data.f90:181, S_data.f90:205; its source is in the un-vendored
//   Libraries_03: restated as c = A b ('r') / c = A^T b ('l') on the
//   unfolded band matrix).
//
// Structure 1 (integral times a bound/channel factor): the list's integrals
// come from the engine in one batch, C = S*c, and the updates go through the
// in-order list kernels.  Structures 2-5 (bound-channel vectors): one
// density and convolution per distinct pair K1 (the plain kernels), one BAV
// vector per distinct (K1, orbital), then UPDATE_HW / UPDATE_HV in list
// order.  Structures 6-9 (S, channel-channel): runs of equal K1 accumulate
// their densities (one batched kernel), one convolution per run, UPDATE_HS
// per run in list order.
#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "dbs_internal.h"

namespace dbs {

// element (i,j) (0-based) of the (ns,ns) matrix held in storage sym
// (0 's', 1 'l', 2 'n', 3 'x'), as Full_mat_sym unfolds it
__device__ __forceinline__ double band_elem(const double* __restrict__ x, int ns, int ks, int sym, int i, int j) {
  if (sym == 3) return x[i + (long long)ns * j];
  if (sym == 2) {
    const int jp = j - i + ks - 1;
    return (jp >= 0 && jp < 2 * ks - 1) ? x[i + (long long)ns * jp] : 0.0;
  }
  const int a = min(i, j), b = max(i, j), d = b - a;
  if (d >= ks) return 0.0;
  return sym == 0 ? x[a + (long long)ns * d] : x[b + (long long)ns * (ks - 1 - d)];
}

// BAV for a list of (matrix g, orbital): W(:,q) = 0 except the ns entries at
// offset woff: sum_j A_g(i,j) b(j) ('r') or A_g(j,i) b(j) ('l'), j ascending
__global__ void k_bav_list(int nvec, const int2* __restrict__ vg, const double* __restrict__ A, long long astride,
                           int sym, int left, const double* __restrict__ pq, int comp, int ns, int ks, int woff,
                           double* __restrict__ W) {
#pragma clang fp contract(off)
  const int ms = 2 * ns;
  const long long total = (long long)nvec * ns;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int vq = (int)(q / ns), i = (int)(q % ns);
    const int2 e = vg[vq];  // (matrix, 1-based orbital)
    const double* a = A + astride * e.x;
    const double* b = pq + (long long)ns * ((comp - 1) + 2LL * (e.y - 1));
    double s = 0.0;
    for (int j = 0; j < ns; ++j)
      s += (left ? band_elem(a, ns, ks, sym, j, i) : band_elem(a, ns, ks, sym, i, j)) * b[j];
    double* w = W + (long long)ms * vq;
    w[woff + i] = s;
    w[(woff == 0 ? ns : 0) + i] = 0.0;
  }
}

__global__ void k_bav(int ns, int ks, const double* __restrict__ a, const double* __restrict__ b, int sym, int left,
                      double* __restrict__ c) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < ns; ++j)
      s += (left ? band_elem(a, ns, ks, sym, j, i) : band_elem(a, ns, ks, sym, i, j)) * b[j];
    c[i] = s;
  }
}

// density of one orbital pair in storage type (0 's', 2 'n', 3 'x'), the
// element rule of density (dbsr_lib_dbs.f90:2834-2891)
__device__ __forceinline__ double dens_elem(const double* p1, const double* p2, int ns, int ks, int type, int i,
                                            int j) {
#pragma clang fp contract(off)
  if (type == 0) {
    if (j == 0) return p1[i] * p2[i];
    if (i < ns - j) return p1[i] * p2[i + j] + p1[i + j] * p2[i];
    return 0.0;
  }
  if (type == 2) {
    const int jf = j + 1, iff = i + 1;
    const int imin = max(1, 1 + ks - jf), imax = min(ns, ns + ks - jf);
    return (iff >= imin && iff <= imax) ? p1[i] * p2[i + jf - ks] : 0.0;
  }
  return p1[i] * p2[j];
}

// xx_g = sum over the run's entries (list order) of c_e * density(pq(jp1,
// j1_e), pq(jp2, j2_e)), as S_data's xx = xx + cdata*dd
__global__ void k_run_density(int ng, const int* __restrict__ goff, const int* __restrict__ ej1,
                              const int* __restrict__ ej2, const double* __restrict__ ec,
                              const double* __restrict__ pq, int ns, int ks, int jp1, int jp2, int type,
                              double* __restrict__ xx) {
#pragma clang fp contract(off)
  const int ncol = type == 0 ? ks : (type == 2 ? 2 * ks - 1 : ns);
  const long long per = (long long)ns * ncol;
  const long long total = per * ng;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int g = (int)(q / per);
    const int r = (int)(q % per);
    const int i = r % ns, j = r / ns;
    double acc = 0.0;
    for (int e = goff[g]; e < goff[g + 1]; ++e) {
      const double* p1 = pq + (long long)ns * ((jp1 - 1) + 2LL * (ej1[e] - 1));
      const double* p2 = pq + (long long)ns * ((jp2 - 1) + 2LL * (ej2[e] - 1));
      acc = acc + ec[e] * dens_elem(p1, p2, ns, ks, type, i, j);
    }
    xx[q] = acc;
  }
}

// UPDATE_HS, MAT3 form (DBSR_MAT3/updates.f90:110-179), one block: y =
// unfold(d, sym); ich>jch: h(ii,jj) += y; jch>ich: += y^T; ich==jch: pass 0
// += y/2, pass 1 h(jj,ii) += (y/2)^T
__global__ void k_update_hs_one(int ns, int ks, int sym, int ich, int jch, int ii, int jj, int pass,
                                const double* __restrict__ d, double* __restrict__ h) {
#pragma clang fp contract(off)
  const int ms = 2 * ns;
  const long long total = (long long)ns * ns;
  const int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(q % ns), b = (int)(q / ns);
    if (ich > jch) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += band_elem(d, ns, ks, sym, a, b);
    } else if (jch > ich) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += band_elem(d, ns, ks, sym, b, a);
    } else if (pass == 0) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += band_elem(d, ns, ks, sym, a, b) / 2.0;
    } else {
      h[(j1 + a) + (long long)ms * (i1 + b)] += band_elem(d, ns, ks, sym, b, a) / 2.0;
    }
  }
}

__global__ void k_div(long long n, double* __restrict__ h, double q) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    h[e] = h[e] / q;
}


// ---------------------------------------------------------------------------
// DBSR_MCHF orbital matrices: small dense pieces of mchf_matrix / core_matrix
// ---------------------------------------------------------------------------
// p = scale * src, or zero when src is null
__global__ void k_set(long long n, double* __restrict__ p, const double* __restrict__ src, double scale) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    p[e] = src ? scale * src[e] : 0.0;
}

// y(r) = y(r) + sum_c A(r,c) x(c), A (m,m) column-major, c ascending (MATMUL)
__global__ void k_gemv_add(int m, const double* __restrict__ A, const double* __restrict__ x, double* __restrict__ y) {
#pragma clang fp contract(off)
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < m; ++c) s += A[r + (long long)m * c] * x[c];
    y[r] = y[r] + s;
  }
}

// y(r) = y(r) + sum_c Full_mat_sym(dd,'s')(r,c) p(c): the band columns of row r
__global__ void k_band_mv_add(int ns, int ks, const double* __restrict__ dd, const double* __restrict__ p,
                              double* __restrict__ y) {
#pragma clang fp contract(off)
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ns; r += gridDim.x * blockDim.x) {
    double s = 0.0;
    const int c0 = max(0, r - ks + 1), c1 = min(ns - 1, r + ks - 1);
    for (int c = c0; c <= c1; ++c) s += band_elem(dd, ns, ks, 0, r, c) * p[c];
    y[r] = y[r] + s;
  }
}

// vv = sum_e v(orb_e) * coef_e / S / 2, v = (P, Q) of the orbital (Get_pv_df)
__global__ void k_rhs_vec(int ms, int n, const int* __restrict__ orb, const double* __restrict__ coef, double S,
                          const double* __restrict__ pq, double* __restrict__ vv) {
#pragma clang fp contract(off)
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ms; r += gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int e = 0; e < n; ++e) a = a + pq[r + (long long)ms * (orb[e] - 1)] * coef[e] / S / 2.0;
    vv[r] = a;
  }
}

// one workgroup of 256: sum over 256 strided partial sums
__device__ double block_sum256(double x, double* sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] = sh[t] + sh[t + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// mchf_matrix.f90:238-246: srhs = sum|rhs|, rhs dropped below 1e-8;
// scal[0] = 1 when rhs moves into hfm (irhs > 0, srhs > 0), scal[1] = SS =
// -rhs.v.  One workgroup of 256.
__global__ void k_mchf_rhs_scan(int ms, int irhs, const double* __restrict__ v, double* __restrict__ rhs,
                                double* __restrict__ scal) {
#pragma clang fp contract(off)
  __shared__ double sh[256];
  double a = 0.0;
  for (int r = threadIdx.x; r < ms; r += 256) a += fabs(rhs[r]);
  double srhs = block_sum256(a, sh);
  if (srhs < 1e-8) {
    for (int r = threadIdx.x; r < ms; r += 256) rhs[r] = 0.0;
    srhs = 0.0;
  }
  __syncthreads();
  a = 0.0;
  for (int r = threadIdx.x; r < ms; r += 256) a += rhs[r] * v[r];
  const double rv = block_sum256(a, sh);
  if (threadIdx.x == 0) {
    scal[0] = (irhs > 0 && srhs > 0.0) ? 1.0 : 0.0;
    scal[1] = -rv;
  }
}

// irhs conversion (mchf_matrix.f90:243-251): h = h + SS B, then
// h(a,b) = h(a,b) + rhs(a) vv(b) + rhs(b) vv(a), vv = B v
__global__ void k_mchf_irhs(int ms, const double* __restrict__ B, const double* __restrict__ rhs,
                            const double* __restrict__ vv, const double* __restrict__ scal, double* __restrict__ h) {
#pragma clang fp contract(off)
  if (scal[0] == 0.0) return;
  const double SS = scal[1];
  const long long n = (long long)ms * ms;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(e % ms), b = (int)(e / ms);
    double x = h[e] + SS * B[e];
    h[e] = x + rhs[a] * vv[b] + rhs[b] * vv[a];
  }
}

// hfm = dhl + hfm; rhs = -rhs (zero after the irhs conversion)
__global__ void k_mchf_final(int ms, const double* __restrict__ dhl, const double* __restrict__ scal,
                             double* __restrict__ h, double* __restrict__ rhs) {
#pragma clang fp contract(off)
  const long long n = (long long)ms * ms;
  const bool moved = scal[0] != 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    h[e] = dhl[e] + h[e];
    if (e < ms) rhs[e] = moved ? -0.0 : -rhs[e];
  }
}

// e(i) = v.(w - rhs), w = hfm v; one workgroup of 256 -> scal[2]
__global__ void k_mchf_energy(int ms, const double* __restrict__ v, const double* __restrict__ w,
                              const double* __restrict__ rhs, double* __restrict__ scal) {
#pragma clang fp contract(off)
  __shared__ double sh[256];
  double a = 0.0;
  for (int r = threadIdx.x; r < ms; r += 256) a += v[r] * (w[r] - rhs[r]);
  const double e = block_sum256(a, sh);
  if (threadIdx.x == 0) scal[2] = e;
}


// ---------------------------------------------------------------------------
// hf_eiv: orbital update of DBSR_HF (hf_solve_HF.f90:118-202)
// ---------------------------------------------------------------------------
// u(c) = sum_r v(r) h(r,c), rows in order (thread per column)
__global__ void k_vt_h(int m, const double* __restrict__ v, const double* __restrict__ h, double* __restrict__ u) {
#pragma clang fp contract(off)
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x) {
    double s = 0.0;
    const double* hc = h + (long long)m * c;
    for (int r = 0; r < m; ++r) s += v[r] * hc[r];
    u[c] = s;
  }
}

// h(r,c) = h(r,c) - x(r) y(c)
__global__ void k_rank1_sub(int m, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ h) {
#pragma clang fp contract(off)
  const long long n = (long long)m * m;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    h[e] = h[e] - x[e % m] * y[e / m];
}

// the kept rows/columns of h and B (boundary conditions): a(p,q) = h(keep(p), keep(q))
__global__ void k_compact2(int m, int kk, const int* __restrict__ keep, const double* __restrict__ h,
                           const double* __restrict__ B, double* __restrict__ a, double* __restrict__ s) {
  const long long n = (long long)kk * kk;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(e % kk), q = (int)(e / kk);
    const long long src = keep[p] + (long long)m * keep[q];
    a[e] = h[src];
    s[e] = B[src];
  }
}

// v = column col of Z expanded onto the kept B-splines, then the sign that
// makes the entry of largest magnitude (first one) positive.  One workgroup
// of 256.
__global__ void k_expand_sign(int m, int kk, const int* __restrict__ keep, const double* __restrict__ Z, int col,
                              double* __restrict__ v) {
  __shared__ double sa[256];
  __shared__ int si[256];
  for (int r = threadIdx.x; r < m; r += 256) v[r] = 0.0;
  __syncthreads();
  for (int p = threadIdx.x; p < kk; p += 256) v[keep[p]] = Z[p + (long long)kk * col];
  __syncthreads();
  double best = -1.0;
  int bi = m;
  for (int r = threadIdx.x; r < m; r += 256) {
    const double a = fabs(v[r]);
    if (a > best) { best = a; bi = r; }
  }
  sa[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = sa[threadIdx.x + w];
      const int i = si[threadIdx.x + w];
      if (a > sa[threadIdx.x] || (a == sa[threadIdx.x] && i < si[threadIdx.x])) {
        sa[threadIdx.x] = a;
        si[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  const bool flip = si[0] < m && v[si[0]] < 0.0;
  __syncthreads();
  if (flip)
    for (int r = threadIdx.x; r < m; r += 256) v[r] = -v[r];
}

namespace {

unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (unsigned)g;
}

int sym_of(char c) { return c == 's' ? 0 : c == 'l' ? 1 : c == 'n' ? 2 : c == 'x' ? 3 : -1; }

// device buffer freed after the stream drained
struct DevMem {
  hipStream_t s;
  std::vector<void*> p;
  explicit DevMem(hipStream_t st) : s(st) {}
  ~DevMem() {
    if (!p.empty()) (void)hipStreamSynchronize(s);
    for (void* x : p) (void)hipFree(x);
  }
  template <class T>
  int alloc(T** d, size_t n) {
    void* x = nullptr;
    if (hipMalloc(&x, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return fail(DBS_ERR_HIP, "hipMalloc");
    p.push_back(x);
    *d = (T*)x;
    return DBS_OK;
  }
  template <class T>
  int put(const std::vector<T>& h, T** d) {
    int rc = alloc(d, h.size());
    if (rc || h.empty()) return rc;
    if (hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(DBS_ERR_HIP, "hipMemcpyAsync");
    return DBS_OK;
  }
};

// the component/case table of one structure
struct Plan {
  int type, kb, ip1, ip2, jp1, jp2, icase;
  char sym_i, sym_j, sym_d, sym_r, sym_v;
};

// R_data.f90:65-86 (stype 0) and S_data.f90:73-115 (stype 1)
int make_plan(int stype, int itype, int atype, int ks, Plan* p) {
  static const int s_tab[2][9][6] = {
      {{1, 2, 1, 2, 2, 0}, {1, 2, 1, 2, 1, 'r'}, {1, 2, 1, 2, 2, 'r'}, {2, 1, 1, 2, 1, 'l'}, {2, 1, 1, 2, 2, 'l'},
       {1, 2, 1, 2, 1, 0}, {1, 2, 1, 2, 2, 0}, {1, 2, 2, 1, 3, 0}, {2, 1, 1, 2, 4, 0}},
      {{1, 2, 2, 1, 2, 0}, {1, 2, 2, 1, 1, 'r'}, {2, 1, 1, 2, 2, 'r'}, {2, 1, 2, 1, 1, 'l'}, {1, 2, 1, 2, 2, 'l'},
       {1, 2, 2, 1, 1, 0}, {2, 1, 1, 2, 2, 0}, {1, 1, 2, 2, 3, 0}, {2, 2, 1, 1, 4, 0}}};
  Plan q{};
  q.sym_v = 'r';
  if (stype == 0) {
    DBS_CHECK(itype >= 1 && itype <= 4, DBS_ERR_ARG, " R_data: unknown itype ");
    DBS_CHECK(atype >= 0 && atype <= 3, DBS_ERR_ARG, "R_data: atype out of range");
    q.type = atype + 1;
    q.kb = ks;
    q.sym_i = q.sym_j = 's';
    q.icase = itype == 1 ? 2 : (itype == 4 ? 3 : 1);
    q.sym_d = q.sym_r = itype == 4 ? 'x' : 's';
    const int tab[4][4] = {{1, 1, 1, 1}, {2, 2, 2, 2}, {1, 1, 2, 2}, {2, 2, 1, 1}};
    q.ip1 = tab[atype][0]; q.ip2 = tab[atype][1]; q.jp1 = tab[atype][2]; q.jp2 = tab[atype][3];
    if (itype == 4 && atype == 2) { q.ip1 = 1; q.ip2 = 2; q.jp1 = 1; q.jp2 = 2; }
    if (itype == 4 && atype == 3) { q.ip1 = 2; q.ip2 = 1; q.jp1 = 2; q.jp2 = 1; }
  } else {
    DBS_CHECK(itype >= 1 && itype <= 9, DBS_ERR_ARG, " S_data: unknown itype ");
    DBS_CHECK(atype == 0 || atype == 1, DBS_ERR_ARG, "S_data: atype out of range");
    const int* t = s_tab[atype][itype - 1];
    q.type = DBS_PPQQ + atype;
    q.kb = 2 * ks - 1;
    q.ip1 = t[0]; q.ip2 = t[1]; q.jp1 = t[2]; q.jp2 = t[3]; q.icase = t[4];
    if (t[5]) q.sym_v = (char)t[5];
    q.sym_i = q.sym_j = 'n';
    q.sym_d = q.sym_r = q.icase <= 2 ? 'n' : 'x';
  }
  *p = q;
  return DBS_OK;
}

int band_ncol(char sym, int ns, int ks) { return sym == 'x' ? ns : (sym == 'n' ? 2 * ks - 1 : ks); }

struct Aux {
  const dbs_mat3_aux* a;
  const char* kbort_msg;
  // 0-based V_ch column of the bound orbital io = i1*ibo + i2; *i1 = io/ibo
  int vcol(int io, int* col, int* i1) const {
    DBS_CHECK(a->ibo > 0 && a->ipbs && a->kbort && a->vch_dev, DBS_ERR_ARG, "bound-orbital tables missing");
    *i1 = io / a->ibo;
    DBS_CHECK(*i1 >= 1, DBS_ERR_ARG, "orbital index < 1");
    const int k = a->kbort(a->kbort_user, io / a->ibo, io % a->ibo);
    if (k == 0) return fail(DBS_ERR_ARG, kbort_msg);
    DBS_CHECK(k >= 1 && k <= a->nvch, DBS_ERR_ARG, "KBORT column out of range");
    *col = k - 1;
    return DBS_OK;
  }
  int ipbs(int orb) const { return a->ipbs[orb - 1]; }
};

// structure 1: Rk/Sk( . . ; . . ) times ic,jc / <i|.> / <i|.><.|j> / <i|j>
int structure1(dbs_ctx* ctx, const dbs_mat_blocks* m, const Aux& X, const Plan& P, int stype, int k, int nc,
               const double* c, const int* K1, const int* K2, const int* K3, const int* K4,
               const double* pq_dev, int norb) {
  const int ibo = X.a->ibo;
  DBS_CHECK(ibo > 0, DBS_ERR_ARG, "ibo < 1");
  std::vector<int> rows;
  rows.reserve(5 * (size_t)nc);
  for (int j = 0; j < nc; ++j) {
    const int a = K1[j] / ibo, cc = K1[j] % ibo, b = K2[j] / ibo, d = K2[j] % ibo;
    DBS_CHECK(a >= 1 && a <= norb && cc >= 1 && cc <= norb && b >= 1 && b <= norb && d >= 1 && d <= norb,
              DBS_ERR_ARG, "R/S_data: orbital index out of range");
    rows.insert(rows.end(), {k, a, b, cc, d});
  }
  // the integrals S of the entries, in one engine pass
  dbs_engine* eng = nullptr;
  int rc = dbs_engine_create(&eng, 1, &ctx, &norb, &nc, rows.data(), P.type);
  if (rc) return rc;
  std::vector<double> S(nc);
  double *epq = nullptr, *evals = nullptr;
  dbs_engine_buffers(eng, 0, &epq, nullptr, &evals, nullptr, nullptr);
  hipStream_t s = ctx->stream;
  do {
    // the engine runs on its own stream: the orbitals must have landed first
    if (hipMemcpyAsync(epq, pq_dev, sizeof(double) * (size_t)ctx->ns * 2 * norb, hipMemcpyDeviceToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { rc = fail(DBS_ERR_HIP, "R/S_data: copy"); break; }
    if ((rc = dbs_engine_forward(eng))) break;
    if ((rc = dbs_engine_sync(eng))) break;
    if (hipMemcpy(S.data(), evals, sizeof(double) * nc, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(DBS_ERR_HIP, "R/S_data: values");
  } while (0);
  dbs_engine_destroy(eng);
  if (rc) return rc;
  const char* unknown = stype ? " S_data: unknown structure for itype=1" : " R_data: unknown structure for itype=1";
  std::vector<MatUpd> u;
  u.reserve(nc);
  for (int j = 0; j < nc; ++j) {
    const double C = S[j] * c[j];
    const int ic = -K3[j], jc = -K4[j], io = -ic, jo = -jc;
    if (ic > 0 && jc > 0) {
      long long t = -1;
      if ((rc = hb_target(m, ic, jc, &t))) return rc;
      u.push_back({MatUpd::HB, 0, 0, 0, t, C});
    } else if (ic < 0 && jc > 0) {
      int vc = 0, i1 = 0, col = 0;
      if ((rc = X.vcol(io, &vc, &i1)) || (rc = hv_column(m, X.ipbs(i1), jc, &col))) return rc;
      u.push_back({MatUpd::HV, 0, vc, 0, (long long)col - 1, C});
    } else if (ic < 0 && jc < 0) {
      int vi = 0, i1 = 0, wj = 0, j1 = 0, b = 0;
      if ((rc = X.vcol(io, &vi, &i1)) || (rc = X.vcol(jo, &wj, &j1))) return rc;
      const int ich = X.ipbs(i1), jch = X.ipbs(j1);
      if ((rc = chan_block(m, ich, jch, "UPDATE_HW", &b))) return rc;
      u.push_back({MatUpd::HW, ich > jch ? 0 : (jch > ich ? 1 : 2), vi, wj, (long long)b - 1, C});
    } else if (ic < 0 && jc == 0) {
      DBS_CHECK(X.a->ipbs, DBS_ERR_ARG, "bound-orbital tables missing");
      DBS_CHECK(io / ibo >= 1 && io % ibo >= 1, DBS_ERR_ARG, "orbital index < 1");
      const int ich = X.ipbs(io / ibo), jch = X.ipbs(io % ibo);
      if ((rc = target_h(m, ich, jch, X.a->iptar, X.a->htarg, C))) return rc;
      if (X.a->iitar > 0 && ich != jch) {
        int b = 0;
        if ((rc = chan_block(m, ich, jch, "UPDATE_HX", &b))) return rc;
        u.push_back({MatUpd::HX, ich >= jch ? 0 : 1, 0, 0, (long long)b - 1, C});
      }
    } else {
      return fail(DBS_ERR_ARG, unknown);
    }
  }
  return apply_mat3(ctx, m, u, X.a->vch_dev, ctx->ms, nullptr, X.a->fppqq_dev, stype ? "S_data" : "R_data");
}

// structures 2-5: Rk/Sk( i . . .) <j|.> or ic -- bound-channel vectors
int structure2(dbs_ctx* ctx, const dbs_mat_blocks* m, const Aux& X, const Plan& P, int stype, int k, int nc,
               const double* c, const int* K1, const int* K2, const int* K3, const int* K4,
               const double* pq_dev, int norb) {
  const int ns = ctx->ns, ks = ctx->ks, ibo = X.a->ibo;
  DBS_CHECK(ibo > 0, DBS_ERR_ARG, "ibo < 1");
  hipStream_t s = ctx->stream;
  // distinct pairs K1 -> matrices, distinct (K1, orbital) -> vectors
  std::map<int, int> gidx;
  std::vector<int> gkey;
  std::map<std::pair<int, int>, int> vidx;
  std::vector<int2> vg;
  std::vector<int> evec(nc);
  for (int j = 0; j < nc; ++j) {
    const int jj = K1[j], ii = K2[j];
    DBS_CHECK(jj / ibo >= 1 && jj / ibo <= norb && jj % ibo >= 1 && jj % ibo <= norb && ii >= 1 && ii <= norb,
              DBS_ERR_ARG, "R/S_data: orbital index out of range");
    auto it = gidx.find(jj);
    int g;
    if (it == gidx.end()) {
      g = (int)gkey.size();
      gidx[jj] = g;
      gkey.push_back(jj);
    } else {
      g = it->second;
    }
    auto iv = vidx.find({g, ii});
    if (iv == vidx.end()) {
      evec[j] = (int)vg.size();
      vidx[{g, ii}] = evec[j];
      vg.push_back(make_int2(g, ii));
    } else {
      evec[j] = iv->second;
    }
  }
  // the updates (validated before any launch)
  const char* unknown = stype ? " S_data: uknown structure for itype 2-5 " : " R_data: uknown structure for itype=2 ";
  std::vector<MatUpd> u;
  int rc = DBS_OK;
  for (int j = 0; j < nc; ++j) {
    const int ich = K3[j], io = K4[j];
    const int wv = -evec[j] - 1;
    if (io > 0) {
      int wc = 0, j1 = 0, b = 0;
      if ((rc = X.vcol(io, &wc, &j1))) return rc;
      const int jch = X.ipbs(j1);
      if ((rc = chan_block(m, ich, jch, "UPDATE_HW", &b))) return rc;
      u.push_back({MatUpd::HW, ich > jch ? 0 : (jch > ich ? 1 : 2), wv, wc, (long long)b - 1, c[j]});
    } else if (io < 0) {
      int col = 0;
      if ((rc = hv_column(m, ich, -io, &col))) return rc;
      u.push_back({MatUpd::HV, 0, wv, 0, (long long)col - 1, c[j]});
    } else {
      return fail(DBS_ERR_ARG, unknown);
    }
  }
  double* rkb = nullptr;
  if ((rc = build_rkb(ctx, P.type, k, &rkb))) return rc;
  const int ng = (int)gkey.size(), nvec = (int)vg.size();
  const int ncd = band_ncol(P.sym_d, ns, ks);
  const int nco = band_ncol(P.icase == 1 ? P.sym_i : P.sym_j, ns, ks);
  DevMem mem(s);
  double *d = nullptr, *A = nullptr, *W = nullptr;
  int2* d_vg = nullptr;
  if ((rc = mem.alloc(&d, (size_t)ns * ncd)) || (rc = mem.alloc(&A, (size_t)ns * nco * ng)) ||
      (rc = mem.alloc(&W, (size_t)2 * ns * nvec)) || (rc = mem.put(vg, &d_vg)))
    return rc;
  for (int g = 0; g < ng; ++g) {
    const int jj = gkey[g];
    const double* p1 = pq_dev + (size_t)ns * ((P.jp1 - 1) + 2 * (size_t)(jj / ibo - 1));
    const double* p2 = pq_dev + (size_t)ns * ((P.jp2 - 1) + 2 * (size_t)(jj % ibo - 1));
    if ((rc = dbs_density_dev(ctx, p1, p2, P.sym_d, d))) return rc;
    if ((rc = convol_launch(ctx, rkb, P.type, P.icase, P.sym_i, P.sym_j, d, A + (size_t)ns * nco * g))) return rc;
  }
  // R: BAV(xx, pq(ip1,ii), 'r'); S: BAV(xx, pq(ip2,ii), sym_v); both placed by ip2
  const int comp = stype ? P.ip2 : P.ip1;
  const int left = stype ? (P.sym_v == 'l') : 0;
  k_bav_list<<<grid_for((long long)nvec * ns, 256), 256, 0, s>>>(nvec, d_vg, A, (long long)ns * nco,
                                                                  sym_of(P.sym_r), left, pq_dev, comp, ns, ks,
                                                                  P.ip2 == 1 ? 0 : ns, W);
  DBS_HIP(hipGetLastError());
  return apply_mat3(ctx, m, u, X.a->vch_dev, ctx->ms, W, nullptr, stype ? "S_data" : "R_data");
}

// structures 6-9 (S): Sk(i . j .) etc. -- channel-channel blocks
int structure6(dbs_ctx* ctx, const dbs_mat_blocks* m, const Aux& X, const Plan& P, int k, int nc, const double* c,
               const int* K1, const int* K2, const int* K3, const double* pq_dev, int norb) {
  const int ns = ctx->ns, ks = ctx->ks, ibo = X.a->ibo;
  DBS_CHECK(ibo > 0, DBS_ERR_ARG, "ibo < 1");
  hipStream_t s = ctx->stream;
  std::vector<int> goff(1, 0), ej1, ej2, gch, gjch, gblk;
  std::vector<double> ec;
  int rc = DBS_OK;
  for (int j = 0; j < nc; ++j) {
    DBS_CHECK(K2[j] >= 1 && K2[j] <= norb && K3[j] >= 1 && K3[j] <= norb, DBS_ERR_ARG,
              "S_data: orbital index out of range");
    ej1.push_back(K2[j]);
    ej2.push_back(K3[j]);
    ec.push_back(c[j]);
    if (j < nc - 1 && K1[j] == K1[j + 1]) continue;
    const int ich = K1[j] / ibo, jch = K1[j] % ibo;
    int b = 0;
    if ((rc = chan_block(m, ich, jch, "UPDATE_HS", &b))) return rc;
    if (b == 0) {  // block not stored: drop the run
      ej1.resize(goff.back()); ej2.resize(goff.back()); ec.resize(goff.back());
      continue;
    }
    gch.push_back(ich);
    gjch.push_back(jch);
    gblk.push_back(b - 1);
    goff.push_back((int)ec.size());
  }
  const int ng = (int)gblk.size();
  if (ng == 0) return DBS_OK;
  double* rkb = nullptr;
  if ((rc = build_rkb(ctx, P.type, k, &rkb))) return rc;
  const int ncd = band_ncol(P.sym_d, ns, ks);
  const int nco = P.icase >= 3 ? ns : band_ncol(P.icase == 1 ? P.sym_i : P.sym_j, ns, ks);
  DevMem mem(s);
  int *d_goff, *d_j1, *d_j2;
  double *d_c, *XX, *D;
  if ((rc = mem.put(goff, &d_goff)) || (rc = mem.put(ej1, &d_j1)) || (rc = mem.put(ej2, &d_j2)) ||
      (rc = mem.put(ec, &d_c)) || (rc = mem.alloc(&XX, (size_t)ns * ncd * ng)) ||
      (rc = mem.alloc(&D, (size_t)ns * nco * ng)))
    return rc;
  k_run_density<<<grid_for((long long)ns * ncd * ng, 256), 256, 0, s>>>(ng, d_goff, d_j1, d_j2, d_c, pq_dev, ns, ks,
                                                                         P.jp1, P.jp2, sym_of(P.sym_d), XX);
  DBS_HIP(hipGetLastError());
  const int ms = ctx->ms;
  for (int g = 0; g < ng; ++g) {
    double* dg = D + (size_t)ns * nco * g;
    if ((rc = convol_launch(ctx, rkb, P.type, P.icase, P.sym_i, P.sym_j, XX + (size_t)ns * ncd * g, dg))) return rc;
    double* h = m->hch_dev + (size_t)ms * ms * gblk[g];
    for (int pass = 0; pass < (gch[g] == gjch[g] ? 2 : 1); ++pass)
      k_update_hs_one<<<grid_for((long long)ns * ns, 256), 256, 0, s>>>(ns, ks, sym_of(P.sym_r), gch[g], gjch[g],
                                                                         P.ip1, P.ip2, pass, dg, h);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("S_data: ") + hipGetErrorString(e));
  return DBS_OK;
}

// hf_matrix, two-electron part (DBSR_HF/hf_matrix.f90:2-78, Breit block
// :97-163): per (ii,jj) the direct and exchange coefficient densities of
// all k in one launch each, then per k one convolution and one UPDATE_HS,
// in the reference's order
int hf_matrix(dbs_ctx* ctx, int i, int norb, const double* pq_dev, const double* A, const double* B, int kmax,
              double qsum, int mbreit, int nrk, const int* kr1, const int* kr2, const int* kr3, const int* kr4,
              const double* crk, double* hfm) {
  const int ns = ctx->ns, ks = ctx->ks, ms = ctx->ms, ng = kmax + 1;
  hipStream_t s = ctx->stream;
  DevMem mem(s);
  double *DD, *XX, *D, *X;
  int rc;
  if ((rc = mem.alloc(&DD, (size_t)ns * ks * ng)) || (rc = mem.alloc(&XX, (size_t)ns * ns * ng)) ||
      (rc = mem.alloc(&D, (size_t)ns * ns)) || (rc = mem.alloc(&X, (size_t)ns * ns)))
    return rc;
  for (int ii = 1; ii <= 2; ++ii)
    for (int jj = 1; jj <= 2; ++jj) {
      const int type = ii == 1 ? (jj == 1 ? DBS_PPPP : DBS_PQPQ) : (jj == 1 ? DBS_QPQP : DBS_QQQQ);
      // coefficient lists per k: direct (a, doubled for j == i) and exchange (b)
      std::vector<int> goff_d(1, 0), goff_x(1, 0), jd, jx;
      std::vector<double> cd, cx;
      for (int k = 0; k <= kmax; ++k) {
        for (int j = 1; j <= norb; ++j) {
          double c = A[(j - 1) + (size_t)norb * k];
          if (c != 0.0) {
            if (i == j) c = c + c;
            jd.push_back(j);
            cd.push_back(c);
          }
          const double b = B[(j - 1) + (size_t)norb * k];
          if (b != 0.0) {
            jx.push_back(j);
            cx.push_back(b);
          }
        }
        goff_d.push_back((int)cd.size());
        goff_x.push_back((int)cx.size());
      }
      DevMem lists(s);
      int *d_gd, *d_gx, *d_jd, *d_jx;
      double *d_cd, *d_cx;
      if ((rc = lists.put(goff_d, &d_gd)) || (rc = lists.put(goff_x, &d_gx)) || (rc = lists.put(jd, &d_jd)) ||
          (rc = lists.put(jx, &d_jx)) || (rc = lists.put(cd, &d_cd)) || (rc = lists.put(cx, &d_cx)))
        return rc;
      k_run_density<<<grid_for((long long)ns * ks * ng, 256), 256, 0, s>>>(ng, d_gd, d_jd, d_jd, d_cd, pq_dev, ns,
                                                                            ks, jj, jj, 0, DD);
      k_run_density<<<grid_for((long long)ns * ns * ng, 256), 256, 0, s>>>(ng, d_gx, d_jx, d_jx, d_cx, pq_dev, ns,
                                                                            ks, ii, jj, 3, XX);
      DBS_HIP(hipGetLastError());
      for (int k = 0; k <= kmax; ++k) {
        const bool dir = goff_d[k + 1] > goff_d[k], exch = goff_x[k + 1] > goff_x[k];
        if (!dir && !exch) continue;
        double* rkb = nullptr;
        if ((rc = build_rkb(ctx, type, k, &rkb))) return rc;
        if (dir) {
          if ((rc = convol_launch(ctx, rkb, type, 1, 's', 's', DD + (size_t)ns * ks * k, D))) return rc;
          if ((rc = dbs_update_hs_dev(ctx, hfm, ii, ii, D, 's'))) return rc;
        }
        if (exch) {
          if ((rc = convol_launch(ctx, rkb, type, 4, 's', 's', XX + (size_t)ns * ns * k, X))) return rc;
          if ((rc = dbs_update_hs_dev(ctx, hfm, ii, jj, X, 'x'))) return rc;
        }
      }
      DBS_HIP(hipStreamSynchronize(s));  // the lists are freed next
    }
  if (mbreit == 2 && nrk > 0) {
    static const int tab[2][4][4] = {{{1, 2, 1, 2}, {1, 2, 1, 2}, {1, 2, 2, 1}, {2, 1, 1, 2}},
                                     {{1, 2, 2, 1}, {2, 1, 1, 2}, {1, 1, 2, 2}, {2, 2, 1, 1}}};
    static const int kint[4] = {2, 1, 4, 3};
    for (int itype = 0; itype <= 1; ++itype)
      for (int k = 0; k <= kmax; ++k) {
        const int kk = k + 1000 * itype;
        // int 3/4 call convol_sk with sym_i = sym_j = 'x', for which it has no
        // branch (dbsr_lib_dbs.f90:2618-2743): the result stays zero and the
        // update adds nothing, so only int 1/2 contribute
        for (int in = 1; in <= 2; ++in) {
          const char sym = 'n';
          const int* t = tab[itype][in - 1];
          std::vector<int> goff(1, 0), jl;
          std::vector<double> cl;
          for (int ik = 0; ik < nrk; ++ik) {
            if (kk != kr1[ik]) continue;
            int jnt, j;
            if (i == kr3[ik]) { jnt = kr2[ik]; j = kr4[ik]; }
            else if (i == kr4[ik]) {
              DBS_CHECK(kr2[ik] >= 1 && kr2[ik] <= 4, DBS_ERR_ARG, "hf_matrix_breit: kr2 out of range");
              jnt = kint[kr2[ik] - 1]; j = kr3[ik];
            } else continue;
            if (jnt != in) continue;
            DBS_CHECK(j >= 1 && j <= norb, DBS_ERR_ARG, "hf_matrix_breit: orbital index out of range");
            jl.push_back(j);
            cl.push_back(crk[ik]);
          }
          if (cl.empty()) continue;
          goff.push_back((int)cl.size());
          DevMem lists(s);
          int *d_g, *d_j;
          double* d_c;
          if ((rc = lists.put(goff, &d_g)) || (rc = lists.put(jl, &d_j)) || (rc = lists.put(cl, &d_c))) return rc;
          const int type = DBS_PPQQ + itype;
          const long long nel = (long long)ns * (2 * ks - 1);
          k_run_density<<<grid_for(nel, 256), 256, 0, s>>>(1, d_g, d_j, d_j, d_c, pq_dev, ns, ks, t[2], t[3],
                                                            sym_of(sym), XX);
          double* rkb = nullptr;
          if ((rc = build_rkb(ctx, type, k, &rkb))) return rc;
          if ((rc = convol_launch(ctx, rkb, type, in, sym, sym, XX, X))) return rc;
          if ((rc = dbs_update_hs_dev(ctx, hfm, t[0], t[1], X, sym))) return rc;
        }
      }
  }
  k_div<<<grid_for((long long)ms * ms, 256), 256, 0, s>>>((long long)ms * ms, hfm, qsum);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("hf_matrix: ") + hipGetErrorString(e));
  return DBS_OK;
}


// ---------------------------------------------------------------------------
// mchf_matrix (DBSR_MCHF/mchf_matrix.f90:2-248).  The angular list arrives
// flattened, one entry per coefficient with c = WW(ic,jc)*coef.  The
// collection list (module id4_data, un-vendored Libraries_03) keeps entries
// in order of first insertion and sums equal keys (k, id2, id3, id4).
// ---------------------------------------------------------------------------
struct Id4 {
  int id2, id3, id4;
  double c;
};

void id4_add(std::vector<Id4>& l, int id2, int id3, int id4, double c) {
  for (Id4& e : l)
    if (e.id2 == id2 && e.id3 == id3 && e.id4 == id4) {
      e.c = e.c + c;
      return;
    }
  l.push_back({id2, id3, id4, c});
}

// mchf_matrix.f90:89-144 for one k
void id4_collect(int i, int k, int nr, const int* rk, const int* r1, const int* r2, const int* r3, const int* r4,
                 const double* rc, double S, double eps_c, std::vector<Id4>& l) {
  for (int n = 0; n < nr; ++n) {
    if (rk[n] != k) continue;
    const int i1 = r1[n], i2 = r2[n], i3 = r3[n], i4 = r4[n];
    const double C = rc[n] / S;
    if (fabs(C) < eps_c) continue;
    const int ii = (i == i1 ? 1000 : 0) + (i == i2 ? 100 : 0) + (i == i3 ? 10 : 0) + (i == i4 ? 1 : 0);
    switch (ii) {
      case 1111: id4_add(l, -1, i1, i3, 2 * C); id4_add(l, -3, i1, i3, 4 * C); break;
      case 1010: id4_add(l, -1, i2, i4, C); break;
      case 101: id4_add(l, -1, i1, i3, C); break;
      case 1001: id4_add(l, -2, i2, i3, C); break;
      case 1100: id4_add(l, -2, i3, i4, C); break;
      case 110: id4_add(l, -2, i1, i4, C); break;
      case 11: id4_add(l, -2, i1, i2, C); break;
      case 111:
      case 1101: id4_add(l, -1, i1, i3, C); id4_add(l, -2, i1, i3, C / 2); id4_add(l, -3, i1, i3, 3 * C / 2); break;
      case 1011:
      case 1110: id4_add(l, -1, i2, i4, C); id4_add(l, -2, i2, i4, C / 2); id4_add(l, -3, i2, i4, 3 * C / 2); break;
      case 1000: id4_add(l, i3, i2, i4, C / 2); break;
      case 100: id4_add(l, i4, i1, i3, C / 2); break;
      case 10: id4_add(l, i1, i2, i4, C / 2); break;
      case 1: id4_add(l, i2, i1, i3, C / 2); break;
      default: break;
    }
  }
}

// group lists over k (g = k) or over (k, orbital) for one P/Q block
struct Groups {
  std::vector<int> off{0}, j1, j2;
  std::vector<double> c;
  void close() { off.push_back((int)c.size()); }
  void add(int a, int b, double x) {
    j1.push_back(a);
    j2.push_back(b);
    c.push_back(x);
  }
  bool has(int g) const { return off[g + 1] > off[g]; }
};

int launch_groups(DevMem& mem, hipStream_t s, const Groups& G, const double* pq, int ns, int ks, int jp1, int jp2,
                  int type, double** out) {
  const int ng = (int)G.off.size() - 1;
  const int ncol = type == 0 ? ks : ns;
  int rc;
  if ((rc = mem.alloc(out, (size_t)ns * ncol * std::max(ng, 1)))) return rc;
  if (G.c.empty()) return DBS_OK;
  int *d_off, *d_j1, *d_j2;
  double* d_c;
  if ((rc = mem.put(G.off, &d_off)) || (rc = mem.put(G.j1, &d_j1)) || (rc = mem.put(G.j2, &d_j2)) ||
      (rc = mem.put(G.c, &d_c)))
    return rc;
  k_run_density<<<grid_for((long long)ns * ncol * ng, 256), 256, 0, s>>>(ng, d_off, d_j1, d_j2, d_c, pq, ns, ks, jp1,
                                                                         jp2, type, *out);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

int mchf_matrix(dbs_ctx* ctx, int i, int norb, const double* pq, const double* dhl, int nl, const int* li1,
                const int* li2, const double* lc, int nr, const int* rk, const int* r1, const int* r2, const int* r3,
                const int* r4, const double* rc_, int kmax, double eps_c, double qsum_min, int newton, int irhs,
                const double* fppqq, double* hfm, double* rhs, double* hx, double* e_out, double* qsum_out) {
  const int ns = ctx->ns, ks = ctx->ks, ms = ctx->ms;
  hipStream_t s = ctx->stream;
  const long long mm = (long long)ms * ms;
  // qsum(i) and CI (mchf_matrix.f90:59-72)
  double S = 0.0;
  std::vector<double> CI(norb, 0.0);
  for (int n = 0; n < nl; ++n) {
    const double C = lc[n];
    if (li1[n] == i && li2[n] == i) S = S + C;
    else if (li1[n] == i) CI[li2[n] - 1] = CI[li2[n] - 1] + C;
    else if (li2[n] == i) CI[li1[n] - 1] = CI[li1[n] - 1] + C;
  }
  if (fabs(S) < qsum_min) S = qsum_min;
  DevMem mem(s);
  double *Dt, *Xt, *W, *scal;
  int rc;
  if ((rc = mem.alloc(&Dt, (size_t)ns * ks)) || (rc = mem.alloc(&Xt, (size_t)ns * ns)) || (rc = mem.alloc(&W, ms)) ||
      (rc = mem.alloc(&scal, 4)))
    return rc;
  k_set<<<grid_for(mm, 256), 256, 0, s>>>(mm, hfm, nullptr, 0.0);
  k_set<<<grid_for(mm, 256), 256, 0, s>>>(mm, hx, nullptr, 0.0);
  k_set<<<1, 256, 0, s>>>(ms, rhs, nullptr, 0.0);
  // one-electron rhs (mchf_matrix.f90:74-82)
  double sci = 0.0;
  std::vector<int> io;
  std::vector<double> ci;
  for (int o = 1; o <= norb; ++o) {
    sci += fabs(CI[o - 1]);
    if (CI[o - 1] != 0.0) {
      io.push_back(o);
      ci.push_back(CI[o - 1]);
    }
  }
  if (sci != 0.0) {
    int* d_io;
    double* d_ci;
    if ((rc = mem.put(io, &d_io)) || (rc = mem.put(ci, &d_ci))) return rc;
    k_rhs_vec<<<grid_for(ms, 256), 256, 0, s>>>(ms, (int)io.size(), d_io, d_ci, S, pq, W);
    k_gemv_add<<<grid_for(ms, 64), 64, 0, s>>>(ms, dhl, W, rhs);
  }
  // collection per k, then per P/Q block the densities of all k in one
  // launch each: direct, exchange, rhs (per orbital) and Newton
  std::vector<std::vector<Id4>> lists(kmax + 1);
  for (int k = 0; k <= kmax; ++k) id4_collect(i, k, nr, rk, r1, r2, r3, r4, rc_, S, eps_c, lists[k]);
  static const int blk[4][3] = {{1, 1, DBS_PPPP}, {1, 2, DBS_PQPQ}, {2, 1, DBS_QPQP}, {2, 2, DBS_QQQQ}};
  Groups Gd[4], Gx[4], Gr[4], Gh[4];
  std::vector<std::vector<int>> rj[4];  // rhs group orbitals per k
  for (int b = 0; b < 4; ++b) rj[b].resize(kmax + 1);
  for (int k = 0; k <= kmax; ++k) {
    const std::vector<Id4>& l = lists[k];
    int last3 = -1;
    for (int n = 0; n < (int)l.size(); ++n)
      if (l[n].id2 == -3 && l[n].c != 0.0) last3 = n;
    std::vector<int> rorb;
    for (const Id4& e : l)
      if (e.id2 > 0 && std::find(rorb.begin(), rorb.end(), e.id2) == rorb.end()) rorb.push_back(e.id2);
    std::sort(rorb.begin(), rorb.end());
    for (int b = 0; b < 4; ++b) {
      for (const Id4& e : l)
        if (e.id2 == -1 && e.c != 0.0) Gd[b].add(e.id3, e.id4, e.c);
      Gd[b].close();
      for (const Id4& e : l)
        if (e.id2 == -2 && e.c != 0.0) Gx[b].add(e.id3, e.id4, e.c);
      Gx[b].close();
      for (int j : rorb) {
        for (const Id4& e : l)
          if (e.id2 == j) Gr[b].add(e.id3, e.id4, e.c);
        Gr[b].close();
        rj[b][k].push_back(j);
      }
      // the Newton term takes the last -3 density without its coefficient
      // (mchf_matrix.f90:222-231: Convol's output is xx, its input x)
      if (newton && last3 >= 0) Gh[b].add(l[last3].id3, l[last3].id4, 1.0);
      Gh[b].close();
    }
  }
  double *DD[4], *XX[4], *RA[4], *HX[4];
  for (int b = 0; b < 4; ++b) {
    const int ii = blk[b][0], jj = blk[b][1];
    if ((rc = launch_groups(mem, s, Gd[b], pq, ns, ks, jj, jj, 0, &DD[b])) ||
        (rc = launch_groups(mem, s, Gx[b], pq, ns, ks, ii, jj, 3, &XX[b])) ||
        (rc = launch_groups(mem, s, Gr[b], pq, ns, ks, jj, jj, 0, &RA[b])) ||
        (rc = launch_groups(mem, s, Gh[b], pq, ns, ks, jj, jj, 3, &HX[b])))
      return rc;
  }
  std::vector<int> rg(4, 0);  // next rhs group per block
  for (int k = 0; k <= kmax; ++k) {
    if (lists[k].empty()) continue;
    for (int b = 0; b < 4; ++b) {
      const int ii = blk[b][0], jj = blk[b][1], type = blk[b][2];
      const bool dir = Gd[b].has(k), exch = Gx[b].has(k), nwt = Gh[b].has(k);
      const int nrg = (int)rj[b][k].size();
      if (!dir && !exch && !nwt && nrg == 0) continue;
      double* rkb = nullptr;
      if ((rc = build_rkb(ctx, type, k, &rkb))) return rc;
      if (dir) {
        if ((rc = convol_launch(ctx, rkb, type, 1, 's', 's', DD[b] + (size_t)ns * ks * k, Dt))) return rc;
        if ((rc = dbs_update_hs_dev(ctx, hfm, ii, ii, Dt, 's'))) return rc;
      }
      if (exch) {
        if ((rc = convol_launch(ctx, rkb, type, 4, 's', 's', XX[b] + (size_t)ns * ns * k, Xt))) return rc;
        if ((rc = dbs_update_hs_dev(ctx, hfm, ii, jj, Xt, 'x'))) return rc;
      }
      for (int q = 0; q < nrg; ++q) {
        const int g = rg[b]++, j = rj[b][k][q];
        if ((rc = convol_launch(ctx, rkb, type, 1, 's', 's', RA[b] + (size_t)ns * ks * g, Dt))) return rc;
        k_band_mv_add<<<grid_for(ns, 64), 64, 0, s>>>(ns, ks, Dt, pq + (size_t)ns * ((ii - 1) + 2 * (size_t)(j - 1)),
                                                     rhs + (size_t)(ii - 1) * ns);
      }
      if (nwt) {
        if ((rc = convol_launch(ctx, rkb, type, 4, 's', 's', HX[b] + (size_t)ns * ns * k, Xt))) return rc;
        if ((rc = dbs_update_hs_dev(ctx, hx, ii, jj, Xt, 'x'))) return rc;
      }
    }
  }
  // rhs into hfm, final matrix and e(i) (mchf_matrix.f90:236-258)
  const double* v = pq + (size_t)ms * (i - 1);
  k_mchf_rhs_scan<<<1, 256, 0, s>>>(ms, irhs, v, rhs, scal);
  if (irhs > 0) {
    k_set<<<grid_for(ms, 256), 256, 0, s>>>(ms, W, nullptr, 0.0);
    k_gemv_add<<<grid_for(ms, 64), 64, 0, s>>>(ms, fppqq, v, W);
    k_mchf_irhs<<<grid_for(mm, 256), 256, 0, s>>>(ms, fppqq, rhs, W, scal, hfm);
  }
  k_mchf_final<<<grid_for(mm, 256), 256, 0, s>>>(ms, dhl, scal, hfm, rhs);
  k_set<<<grid_for(ms, 256), 256, 0, s>>>(ms, W, nullptr, 0.0);
  k_gemv_add<<<grid_for(ms, 64), 64, 0, s>>>(ms, hfm, v, W);
  k_mchf_energy<<<1, 256, 0, s>>>(ms, v, W, rhs, scal);
  double h_scal[4];
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h_scal, scal, sizeof h_scal, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("mchf_matrix: ") + hipGetErrorString(e));
  if (e_out) *e_out = h_scal[2];
  if (qsum_out) *qsum_out = S;
  return DBS_OK;
}

// core_matrix (DBSR_MCHF/mchf_matrix.f90:253-345): hf_matrix's loops with
// a(i,j,k) qsum(i) qsum(j) (qsum(i)(qsum(i)-1) for j = i, which hf_matrix
// doubles, so half of it goes in) and b likewise; hfm starts as qsum(i) hd
int core_matrix(dbs_ctx* ctx, int i, int norb, const double* pq, const double* hd, const double* A, const double* B,
                int kmax, const double* qsum, int newton, double* hfm, double* hx) {
  const int ns = ctx->ns, ms = ctx->ms;
  hipStream_t s = ctx->stream;
  const long long mm = (long long)ms * ms;
  const double q = qsum[i - 1];
  std::vector<double> A2((size_t)norb * (kmax + 1)), B2((size_t)norb * (kmax + 1));
  for (int k = 0; k <= kmax; ++k)
    for (int j = 1; j <= norb; ++j) {
      const size_t x = (j - 1) + (size_t)norb * k;
      if (j != i) {
        A2[x] = A[x] * q * qsum[j - 1];
        B2[x] = B[x] * q * qsum[j - 1];
      } else {
        A2[x] = A[x] * q * (q - 1) / 2.0;
        B2[x] = B[x] * q * (q - 1);
      }
    }
  k_set<<<grid_for(mm, 256), 256, 0, s>>>(mm, hfm, hd, q);
  k_set<<<grid_for(mm, 256), 256, 0, s>>>(mm, hx, nullptr, 0.0);
  DBS_HIP(hipGetLastError());
  int rc = hf_matrix(ctx, i, norb, pq, A2.data(), B2.data(), kmax, q, 0, 0, nullptr, nullptr, nullptr, nullptr,
                     nullptr, hfm);
  if (rc || !newton) return rc;
  // H_aa terms (:323-331): c = 2 a(i,i,k) q (q-1), exchange density of i
  DevMem mem(s);
  double *X, *Xt;
  if ((rc = mem.alloc(&X, (size_t)ns * ns)) || (rc = mem.alloc(&Xt, (size_t)ns * ns))) return rc;
  static const int blk[4][3] = {{1, 1, DBS_PPPP}, {1, 2, DBS_PQPQ}, {2, 1, DBS_QPQP}, {2, 2, DBS_QQQQ}};
  std::vector<int> off{0, 1}, jl{i};
  for (int b = 0; b < 4; ++b)
    for (int k = 0; k <= kmax; ++k) {
      double c = 2.0 * A[(i - 1) + (size_t)norb * k];
      if (c == 0.0) continue;
      c = c * q * (q - 1);
      const int ii = blk[b][0], jj = blk[b][1], type = blk[b][2];
      std::vector<double> cl{c};
      DevMem lists(s);
      int *d_off, *d_j;
      double* d_c;
      if ((rc = lists.put(off, &d_off)) || (rc = lists.put(jl, &d_j)) || (rc = lists.put(cl, &d_c))) return rc;
      k_run_density<<<grid_for((long long)ns * ns, 256), 256, 0, s>>>(1, d_off, d_j, d_j, d_c, pq, ns, ctx->ks, ii,
                                                                       jj, 3, X);
      double* rkb = nullptr;
      if ((rc = build_rkb(ctx, type, k, &rkb))) return rc;
      if ((rc = convol_launch(ctx, rkb, type, 4, 's', 's', X, Xt))) return rc;
      if ((rc = dbs_update_hs_dev(ctx, hx, ii, jj, Xt, 'x'))) return rc;
    }
  k_div<<<grid_for(mm, 256), 256, 0, s>>>(mm, hx, q);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("core_matrix: ") + hipGetErrorString(e));
  return DBS_OK;
}


void free_solver(void* h) { (void)rocblas_destroy_handle((rocblas_handle)h); }

}  // namespace

int solver_of(dbs_ctx* ctx, rocblas_handle* h) {
  if (!ctx->solver) {
    rocblas_handle x = nullptr;
    if (rocblas_create_handle(&x) != rocblas_status_success) return fail(DBS_ERR_HIP, "hf_eiv: solver handle");
    ctx->solver = x;
    ctx->solver_free = free_solver;
  }
  *h = (rocblas_handle)ctx->solver;
  if (rocblas_set_stream(*h, ctx->stream) != rocblas_status_success)
    return fail(DBS_ERR_HIP, "hf_eiv: solver stream");
  return DBS_OK;
}

int conv_gemm(dbs_ctx* c, int icase, int K, int n, const double* rkm, const double* dens, double* conv) {
  rocblas_handle h;
  int rc = solver_of(c, &h);
  if (rc) return rc;
  const double one = 1.0, zero = 0.0;
  const rocblas_operation op = icase == 2 ? rocblas_operation_transpose : rocblas_operation_none;
  if (rocblas_dgemm(h, op, rocblas_operation_none, K, n, K, &one, rkm, K, dens, K, &zero, conv, K) !=
      rocblas_status_success)
    return fail(DBS_ERR_HIP, "convol: convolution product");
  return DBS_OK;
}

namespace {

// hf_eiv (hf_solve_HF.f90:136-202): orthogonality to each vorth column
// (H <- (1 - B v v^T) H (1 - v v^T B), orthogonality in hf_orthogonality.f90),
// the B-splines with iprm = 0 removed, H x = e B x (lower triangles), the
// m-th eigenvalue above -1.999 c^2 and its eigenvector on the full net with
// its largest entry positive
int hf_eiv(dbs_ctx* ctx, double* hfm, int northo, const double* vorth, const int* iprm, int m, const double* B,
           double* v, double* e_out) {
  const int ms = ctx->ms;
  hipStream_t s = ctx->stream;
  DevMem mem(s);
  double *w, *u;
  int rc;
  if ((rc = mem.alloc(&w, ms)) || (rc = mem.alloc(&u, ms))) return rc;
  for (int j = 0; j < northo; ++j) {
    const double* vj = vorth + (size_t)ms * j;
    k_set<<<grid_for(ms, 256), 256, 0, s>>>(ms, w, nullptr, 0.0);
    k_gemv_add<<<grid_for(ms, 64), 64, 0, s>>>(ms, B, vj, w);  // w = B v
    k_vt_h<<<grid_for(ms, 64), 64, 0, s>>>(ms, vj, hfm, u);     // u = v^T H
    k_rank1_sub<<<grid_for((long long)ms * ms, 256), 256, 0, s>>>(ms, w, u, hfm);  // H - w u^T
    k_set<<<grid_for(ms, 256), 256, 0, s>>>(ms, u, nullptr, 0.0);
    k_gemv_add<<<grid_for(ms, 64), 64, 0, s>>>(ms, hfm, vj, u);  // u = H v
    k_rank1_sub<<<grid_for((long long)ms * ms, 256), 256, 0, s>>>(ms, u, w, hfm);  // H - u w^T
  }
  std::vector<int> keep;
  for (int r = 0; r < ms; ++r)
    if (iprm[r] != 0) keep.push_back(r);
  const int kk = (int)keep.size();
  DBS_CHECK(kk > 0, DBS_ERR_ARG, "hf_eiv: every B-spline removed");
  int* d_keep;
  double *A, *S, *D, *E;
  rocblas_int* info;
  if ((rc = mem.put(keep, &d_keep)) || (rc = mem.alloc(&A, (size_t)kk * kk)) || (rc = mem.alloc(&S, (size_t)kk * kk)) ||
      (rc = mem.alloc(&D, kk)) || (rc = mem.alloc(&E, kk)) || (rc = mem.alloc(&info, 1)))
    return rc;
  k_compact2<<<grid_for((long long)kk * kk, 256), 256, 0, s>>>(ms, kk, d_keep, hfm, B, A, S);
  DBS_HIP(hipGetLastError());
  rocblas_handle h;
  if ((rc = solver_of(ctx, &h))) return rc;
  if (rocsolver_dsygvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_lower, kk, A, kk, S, kk, D, E,
                       info) != rocblas_status_success)
    return fail(DBS_ERR_HIP, "hf_eiv: rocsolver_dsygvd");
  std::vector<double> ev(kk);
  rocblas_int hinfo = 0;
  hipError_t he = hipMemcpyAsync(ev.data(), D, sizeof(double) * kk, hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) return fail(DBS_ERR_HIP, std::string("hf_eiv: ") + hipGetErrorString(he));
  if (hinfo != 0) return fail(DBS_ERR_HIP, "Error in Eigenvalue routine, dsygv " + std::to_string(hinfo));
  // the first eigenvalue above the negative-energy continuum, then m - 1 more
  const double zz = -1.999 * 137.03599976 * 137.03599976;
  int mm = 0;
  for (int j = 1; j <= kk; ++j) {
    mm = j;
    if (ev[j - 1] > zz) break;
  }
  mm = m + mm - 1;
  DBS_CHECK(mm >= 1 && mm <= kk, DBS_ERR_ARG, "hf_eiv: eigenvalue index out of range");
  k_expand_sign<<<1, 256, 0, s>>>(ms, kk, d_keep, A, mm - 1, v);
  he = hipGetLastError();
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) return fail(DBS_ERR_HIP, std::string("hf_eiv: ") + hipGetErrorString(he));
  if (e_out) *e_out = ev[mm - 1];
  return DBS_OK;
}

int rs_data(dbs_ctx* ctx, const dbs_mat_blocks* m, const dbs_mat3_aux* aux, int stype, int itype, int atype, int k,
            int nc, const double* c, const int* K1, const int* K2, const int* K3, const int* K4,
            const double* pq_dev, int norb) {
  const char* who = stype ? "S_data" : "R_data";
  DBS_CHECK(ctx && aux && pq_dev, DBS_ERR_ARG, std::string(who) + ": null argument");
  int rc = check_blocks(m);
  if (rc) return rc;
  Plan P;
  if ((rc = make_plan(stype, itype, atype, ctx->ks, &P))) return rc;
  if (nc <= 0) return DBS_OK;
  DBS_CHECK(c && K1 && K2 && K3 && K4, DBS_ERR_ARG, std::string(who) + ": null list");
  Aux X{aux, stype ? "S: KBORT = 0" : "RK: KBORT =0"};
  if (stype == 0 && (itype == 3 || itype == 4)) {
    // asymptotic coefficients of the direct P-P / Q-Q structure
    // (R_data.f90:222-226, 240): per run of equal K1, CA = sum c*QUADR_PQ
    // in list order, then UPDATE_ACF(kpol,ich,jch,CA)
    const bool acf = itype == 3 && atype <= 1 && m->acf_dev;
    std::vector<int> ak, aich, ajch;
    std::vector<double> aC;
    if (acf) {
      DBS_CHECK(aux->quadr_pq, DBS_ERR_ARG, "R_data: QUADR_PQ callback needed for the asymptotic coefficients");
      double CA = 0.0;
      for (int j = 0; j < nc; ++j) {
        CA = CA + c[j] * aux->quadr_pq(aux->quadr_pq_user, K2[j], K3[j], k);
        if (j < nc - 1 && K1[j] == K1[j + 1]) continue;
        ak.push_back(k);
        aich.push_back(K1[j] / aux->ibo);
        ajch.push_back(K1[j] % aux->ibo);
        aC.push_back(CA);
        CA = 0.0;
      }
    }
    rc = dbs_rdata_34_dev(ctx, itype, atype, k, nc, c, K1, K2, K3, aux->ibo, pq_dev, norb, m->icc, m->nch,
                          m->hch_dev, m->nblocks);
    if (rc || !acf) return rc;
    return update_acf_queued(ctx, m, (int)aC.size(), ak.data(), aich.data(), ajch.data(), aC.data());
  }
  if (itype == 1) return structure1(ctx, m, X, P, stype, k, nc, c, K1, K2, K3, K4, pq_dev, norb);
  if (itype <= 5) return structure2(ctx, m, X, P, stype, k, nc, c, K1, K2, K3, K4, pq_dev, norb);
  return structure6(ctx, m, X, P, k, nc, c, K1, K2, K3, pq_dev, norb);
}

}  // namespace
}  // namespace dbs

using namespace dbs;

extern "C" {

double dbs_quadr_pq_table(void* user, int i1, int i2, int m) {
  const dbs_qpq_table* t = static_cast<const dbs_qpq_table*>(user);
  if (!t || !t->q || i1 < 1 || i2 < 1 || i1 > t->norb || i2 > t->norb || m < 0 || m > t->mmax) return 0.0;
  return t->q[(i1 - 1) + (size_t)t->norb * ((i2 - 1) + (size_t)t->norb * m)];
}

int dbs_rdata_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, const dbs_mat3_aux* aux, int itype, int atype, int k,
                  int nc, const double* c, const int* K1, const int* K2, const int* K3, const int* K4,
                  const double* pq_dev, int norb) {
  return rs_data(ctx, m, aux, 0, itype, atype, k, nc, c, K1, K2, K3, K4, pq_dev, norb);
}

int dbs_sdata_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, const dbs_mat3_aux* aux, int ktype, int k, int nc,
                  const double* c, const int* K1, const int* K2, const int* K3, const int* K4, const double* pq_dev,
                  int norb) {
  const int ibtype = 10;  // mod_dbsr_mat.f90:103
  return rs_data(ctx, m, aux, 1, ktype % ibtype, ktype / ibtype, k, nc, c, K1, K2, K3, K4, pq_dev, norb);
}

int dbs_hf_matrix_dev(dbs_ctx* ctx, int i, int norb, const double* pq_dev, const double* acoef,
                      const double* bcoef, int kmax, double qsum, int mbreit, int nrk, const int* kr1, const int* kr2,
                      const int* kr3, const int* kr4, const double* crk, double* hfm_dev) {
  DBS_CHECK(ctx && pq_dev && acoef && bcoef && hfm_dev, DBS_ERR_ARG, "hf_matrix: null argument");
  DBS_CHECK(i >= 1 && i <= norb && kmax >= 0, DBS_ERR_ARG, "hf_matrix: orbital or kmax out of range");
  DBS_CHECK(qsum != 0.0, DBS_ERR_ARG, "hf_matrix: qsum = 0");
  DBS_CHECK(mbreit != 2 || nrk <= 0 || (kr1 && kr2 && kr3 && kr4 && crk), DBS_ERR_ARG,
            "hf_matrix: Breit list missing");
  return hf_matrix(ctx, i, norb, pq_dev, acoef, bcoef, kmax, qsum, mbreit, nrk, kr1, kr2, kr3, kr4, crk, hfm_dev);
}

int dbs_mchf_matrix_dev(dbs_ctx* ctx, int i, int norb, const double* pq_dev, const double* dhl_dev, int nl,
                        const int* l_i1, const int* l_i2, const double* l_c, int nr, const int* r_k, const int* r_i1,
                        const int* r_i2, const int* r_i3, const int* r_i4, const double* r_c, int kmax, double eps_c,
                        double qsum_min, int newton, int irhs, const double* fppqq_dev, double* hfm_dev,
                        double* rhs_dev, double* hx_dev, double* e_i, double* qsum_i) {
  DBS_CHECK(ctx && pq_dev && dhl_dev && hfm_dev && rhs_dev && hx_dev, DBS_ERR_ARG, "mchf_matrix: null argument");
  DBS_CHECK(i >= 1 && i <= norb && kmax >= 0 && nl >= 0 && nr >= 0, DBS_ERR_ARG,
            "mchf_matrix: orbital, kmax or list length out of range");
  DBS_CHECK(nl == 0 || (l_i1 && l_i2 && l_c), DBS_ERR_ARG, "mchf_matrix: L list missing");
  DBS_CHECK(nr == 0 || (r_k && r_i1 && r_i2 && r_i3 && r_i4 && r_c), DBS_ERR_ARG, "mchf_matrix: Rk list missing");
  DBS_CHECK(irhs <= 0 || fppqq_dev, DBS_ERR_ARG, "mchf_matrix: irhs > 0 needs fppqq");
  DBS_CHECK(qsum_min > 0.0, DBS_ERR_ARG, "mchf_matrix: qsum_min <= 0");
  for (int n = 0; n < nl; ++n)
    DBS_CHECK(l_i1[n] >= 1 && l_i1[n] <= norb && l_i2[n] >= 1 && l_i2[n] <= norb, DBS_ERR_ARG,
              "mchf_matrix: L orbital index out of range");
  for (int n = 0; n < nr; ++n)
    DBS_CHECK(r_k[n] >= 0 && r_i1[n] >= 1 && r_i1[n] <= norb && r_i2[n] >= 1 && r_i2[n] <= norb && r_i3[n] >= 1 &&
                  r_i3[n] <= norb && r_i4[n] >= 1 && r_i4[n] <= norb,
              DBS_ERR_ARG, "mchf_matrix: Rk orbital index or k out of range");
  return mchf_matrix(ctx, i, norb, pq_dev, dhl_dev, nl, l_i1, l_i2, l_c, nr, r_k, r_i1, r_i2, r_i3, r_i4, r_c, kmax,
                     eps_c, qsum_min, newton, irhs, fppqq_dev, hfm_dev, rhs_dev, hx_dev, e_i, qsum_i);
}

int dbs_core_matrix_dev(dbs_ctx* ctx, int i, int norb, const double* pq_dev, const double* hd_dev,
                        const double* acoef, const double* bcoef, int kmax, const double* qsum, int newton,
                        double* hfm_dev, double* hx_dev) {
  DBS_CHECK(ctx && pq_dev && hd_dev && acoef && bcoef && qsum && hfm_dev && hx_dev, DBS_ERR_ARG,
            "core_matrix: null argument");
  DBS_CHECK(i >= 1 && i <= norb && kmax >= 0, DBS_ERR_ARG, "core_matrix: orbital or kmax out of range");
  DBS_CHECK(qsum[i - 1] != 0.0, DBS_ERR_ARG, "core_matrix: qsum = 0");
  return core_matrix(ctx, i, norb, pq_dev, hd_dev, acoef, bcoef, kmax, qsum, newton, hfm_dev, hx_dev);
}

int dbs_hf_eiv_dev(dbs_ctx* ctx, double* hfm_dev, int northo, const double* vorth_dev, const int* iprm, int m,
                   const double* fppqq_dev, double* v_dev, double* e_i) {
  DBS_CHECK(ctx && hfm_dev && iprm && fppqq_dev && v_dev && (northo <= 0 || vorth_dev), DBS_ERR_ARG,
            "hf_eiv: null argument");
  DBS_CHECK(m >= 1, DBS_ERR_ARG, "hf_eiv: m < 1");
  return hf_eiv(ctx, hfm_dev, northo, vorth_dev, iprm, m, fppqq_dev, v_dev, e_i);
}

int dbs_bav_dev(dbs_ctx* ctx, const double* a_dev, const double* b_dev, char sym, char dir, double* c_dev) {
  DBS_CHECK(ctx && a_dev && b_dev && c_dev, DBS_ERR_ARG, "BAV: null argument");
  const int sc = sym_of(sym);
  DBS_CHECK(sc >= 0, DBS_ERR_ARG, "BAV: unknown symmetry");
  DBS_CHECK(dir == 'l' || dir == 'r', DBS_ERR_ARG, "BAV: unknown direction");
  k_bav<<<grid_for(ctx->ns, 256), 256, 0, ctx->stream>>>(ctx->ns, ctx->ks, a_dev, b_dev, sc, dir == 'l', c_dev);
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

}  // extern "C"

// dbs_channels.hip -- DBSR_MAT3's channel-block post-processing on the GPU:
// the zero conditions (Zero_cond), the orthogonality conditions (DBS_ORTH),
// the diagonalisation of the diagonal channel blocks (Diag_channels) and the
// transformation of the other blocks to the channel solutions
// (Record_matrix).
//
// Layout: the MAT3 storage of dbs_mat_blocks (hch(ms,ms,nblocks) column-major
// in HBM, icc/icb/ibb host tables).  The reference builds dense ms x ms
// projectors and multiplies whole blocks by them (two ms^3 products per
// block); here a channel's conditions are its nort vector pairs, so a block
// is projected by one rank-2*nmo update, X <- X - [Q_L | V - Q_L S] [U ; Q_R^T]
// with U = P_L^T X, V = X P_R, S = P_L^T X P_R -- four passes over the block,
// HBM-bound -- as batched rocBLAS products over all selected blocks.  The
// eigen-solves of all channels run as one strided batch (smaller problems
// padded by a decoupled block): Cholesky, reduction by two triangular solves,
// rocSOLVER dsyevd, back-substitution; the transformation D_i^T H D_j is two
// batched products per block.  Selection of the eigenvalue window and the
// record offsets are host work on the eigenvalues and per-block maxima only.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "dbs_internal.h"

namespace dbs {
namespace {

unsigned grid_n(long long n, int block) {
  long long g = (n + block - 1) / block;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 1 << 16));
}

// device allocations of one call, freed after the stream drains
struct Mem {
  hipStream_t s;
  std::vector<void*> p;
  explicit Mem(hipStream_t st) : s(st) {}
  ~Mem() {
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

int blas(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    return fail(DBS_ERR_HIP, std::string(what) + ": " + rocblas_status_to_string(st));
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// per-channel condition vectors, zero-padded to nmo columns; channel nch is
// the all-zero side (a block side without conditions)
__global__ void k_pad_vectors(int ms, int nmo, int nch, const int* __restrict__ nort, const int* __restrict__ iort,
                              int ldo, const double* __restrict__ pv, const double* __restrict__ qv,
                              double* __restrict__ P, double* __restrict__ Q) {
  const long long n = (long long)ms * nmo * (nch + 1);
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(x % ms);
    const long long r = x / ms;
    const int a = (int)(r % nmo), ch = (int)(r / nmo);
    double p = 0.0, q = 0.0;
    if (ch < nch && a < nort[ch]) {
      const size_t j = (size_t)(iort[(size_t)ch * ldo + a] - 1);
      p = pv[i + (size_t)ms * j];
      q = qv[i + (size_t)ms * j];
    }
    P[x] = p;
    Q[x] = q;
  }
}

// L_e(:, 0:nmo) = Q of the left channel; R_e(nmo+a, :) = Q(:, a) of the right
__global__ void k_fill_lr(int ms, int nmo, int ne, const int* __restrict__ lch, const int* __restrict__ rch,
                          const double* __restrict__ Q, double* __restrict__ L, double* __restrict__ R) {
  const long long n = (long long)ne * ms * nmo;
  const size_t per = (size_t)ms * 2 * nmo;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(x % ms);
    const long long r = x / ms;
    const int a = (int)(r % nmo), e = (int)(r / nmo);
    L[per * e + i + (size_t)ms * a] = Q[i + (size_t)ms * (a + (size_t)nmo * lch[e])];
    R[per * e + (nmo + a) + (size_t)2 * nmo * i] = Q[i + (size_t)ms * (a + (size_t)nmo * rch[e])];
  }
}

// one workgroup per perturber column: h <- h - Q (P^T h)
// (hcp(:,jc) - matmul(hcp(:,jc), bb), dbs_orth.f90:74-78)
__global__ void __launch_bounds__(256) k_orth_vec(int ms, int nmo, const int* __restrict__ col,
                                                  const int* __restrict__ ch, const double* __restrict__ P,
                                                  const double* __restrict__ Q, double* __restrict__ hcp) {
  __shared__ double red[256];
  __shared__ double u[64];
  double* h = hcp + (size_t)ms * col[blockIdx.x];
  const double* p = P + (size_t)ms * nmo * ch[blockIdx.x];
  const double* q = Q + (size_t)ms * nmo * ch[blockIdx.x];
  for (int a0 = 0; a0 < nmo; a0 += 64) {
    const int na = min(64, nmo - a0);
    for (int a = 0; a < na; ++a) {
      double s = 0.0;
      for (int i = threadIdx.x; i < ms; i += blockDim.x) s += p[i + (size_t)ms * (a0 + a)] * h[i];
      red[threadIdx.x] = s;
      __syncthreads();
      for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) u[a] = red[0];
      __syncthreads();
    }
    for (int i = threadIdx.x; i < ms; i += blockDim.x) {
      double s = 0.0;
      for (int a = 0; a < na; ++a) s += q[i + (size_t)ms * (a0 + a)] * u[a];
      h[i] -= s;
    }
    __syncthreads();
  }
}

// compacted pairs of the eigen-problems of one batch: A = hch(keep,keep,blk),
// B = diag(keep,keep,blk) in the leading kr x kr part; a channel with fewer
// kept B-splines than the batch's order kk is padded with (sigma I, I), a
// block that decouples exactly (its zeros stay zeros through the Cholesky
// factor, the reduction and the Householder steps) and whose eigenvalue
// sigma lies outside the energy window
__global__ void k_compact_pair(int ms, int kk, int nb, const int* __restrict__ keep, const int* __restrict__ kr,
                               const int* __restrict__ blk, double sigma, const double* __restrict__ hch,
                               const double* __restrict__ dg, double* __restrict__ A, double* __restrict__ B) {
  const long long n = (long long)nb * kk * kk;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(x % kk);
    const long long r = x / kk;
    const int j = (int)(r % kk), e = (int)(r / kk);
    if (i < kr[e] && j < kr[e]) {
      const size_t src = (size_t)ms * ms * blk[e] + keep[(size_t)e * kk + i] + (size_t)ms * keep[(size_t)e * kk + j];
      A[x] = hch[src];
      B[x] = dg[src];
    } else {
      A[x] = i == j ? sigma : 0.0;
      B[x] = i == j ? 1.0 : 0.0;
    }
  }
}

// one workgroup per selected solution: expand the eigenvector to the full
// B-spline net, make its first largest entry positive (maxloc(abs(v))), store
// it in diag(:,s,blk) (diag.f90:50-56)
__global__ void __launch_bounds__(256) k_expand_solutions(int ms, int kk, const int* __restrict__ sol_e,
                                                          const int* __restrict__ sol_col, const int* __restrict__ sol_s,
                                                          const int* __restrict__ blk, const int* __restrict__ kr,
                                                          const int* __restrict__ inv, const double* __restrict__ Z,
                                                          double* __restrict__ dg) {
  __shared__ double bv[256];
  __shared__ int bi[256];
  const int e = sol_e[blockIdx.x];
  const double* z = Z + (size_t)kk * kk * e + (size_t)kk * sol_col[blockIdx.x];
  double best = -1.0;
  int idx = 0x7fffffff;
  for (int i = threadIdx.x; i < kr[e]; i += blockDim.x) {
    const double a = fabs(z[i]);
    if (a > best) {
      best = a;
      idx = i;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = idx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double o = bv[threadIdx.x + w];
      const int oi = bi[threadIdx.x + w];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const double sgn = z[bi[0]] < 0.0 ? -1.0 : 1.0;
  double* out = dg + (size_t)ms * ms * blk[e] + (size_t)ms * sol_s[blockIdx.x];
  const int* iv = inv + (size_t)ms * e;
  for (int r = threadIdx.x; r < ms; r += blockDim.x) {
    const int c = iv[r];
    out[r] = c >= 0 ? sgn * z[c] : 0.0;
  }
}

// the eigenvalues into column ns, where the reference's last write leaves
// them: all of them when fewer than ns solutions, else those of solutions
// ns.. (diag.f90:57; solution ns overwrites the column before)
__global__ void k_store_energies(int ms, int ns, int nsol, const int* __restrict__ sol_e,
                                 const int* __restrict__ sol_s, const int* __restrict__ ksol,
                                 const int* __restrict__ blk, const double* __restrict__ E, double* __restrict__ dg) {
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < nsol; x += gridDim.x * blockDim.x) {
    const int e = sol_e[x], s = sol_s[x];
    if (ksol[e] < ns || s >= ns - 1) dg[(size_t)ms * ms * blk[e] + s + (size_t)ms * (ns - 1)] = E[x];
  }
}

// largest |x| of each of n segments of length len (one workgroup each)
__global__ void __launch_bounds__(256) k_seg_maxabs(long long len, const double* __restrict__ x,
                                                    const long long* __restrict__ off, double* __restrict__ mx) {
  __shared__ double red[256];
  const double* p = x + off[blockIdx.x];
  double m = 0.0;
  for (long long i = threadIdx.x; i < len; i += blockDim.x) m = fmax(m, fabs(p[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) mx[blockIdx.x] = red[0];
}

// solutions of every channel padded to nmax columns: D(ms, nmax, nch)
__global__ void k_pad_solutions(int ms, int nmax, int nch, const int* __restrict__ nsol, const double* __restrict__ dg,
                                double* __restrict__ D) {
  const long long n = (long long)ms * nmax * nch;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < n; x += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(x % ms);
    const long long r = x / ms;
    const int a = (int)(r % nmax), ch = (int)(r / nmax);
    D[x] = a < nsol[ch] ? dg[(size_t)ms * ms * ch + (size_t)ms * a + i] : 0.0;
  }
}

// one workgroup per record: c(1:ni,1:nj) of the padded product into the
// packed output, and its largest |c|
__global__ void __launch_bounds__(256) k_pack_records(int nmax, const int* __restrict__ ni, const int* __restrict__ nj,
                                                      const long long* __restrict__ off, const double* __restrict__ C,
                                                      double* __restrict__ out, double* __restrict__ mx) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  const int a = ni[e], b = nj[e];
  const double* c = C + (size_t)nmax * nmax * e;
  double* o = out + off[e];
  double m = 0.0;
  for (long long x = threadIdx.x; x < (long long)a * b; x += blockDim.x) {
    const int i = (int)(x % a), j = (int)(x / a);
    const double v = c[i + (size_t)nmax * j];
    o[x] = v;
    m = fmax(m, fabs(v));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) mx[e] = red[0];
}

// one workgroup per perturber record: w(i) = D(:,i)^T h, i < ni
// (record_matrix.f90:41), and its largest |w|
__global__ void __launch_bounds__(256) k_dt_vec(int ms, int nmax, const int* __restrict__ ch, const int* __restrict__ col,
                                                const int* __restrict__ ni, const long long* __restrict__ off,
                                                const double* __restrict__ D, const double* __restrict__ hcp,
                                                double* __restrict__ out, double* __restrict__ mx) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  const double* d = D + (size_t)ms * nmax * ch[e];
  const double* h = hcp + (size_t)ms * col[e];
  double m = 0.0;
  for (int i = threadIdx.x; i < ni[e]; i += blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < ms; ++r) s += d[r + (size_t)ms * i] * h[r];
    out[off[e] + i] = s;
    m = fmax(m, fabs(s));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) mx[e] = red[0];
}

int hip_ok(const char* who) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
  return DBS_OK;
}

int to_host(void* h, const void* d, size_t bytes, hipStream_t s, const char* who) {
  hipError_t e = hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// DBS_ORTH
// ---------------------------------------------------------------------------
int orth(dbs_ctx* ctx, const dbs_mat_blocks* m, int idiag, const int* nort, const int* iort, int ldo,
         const double* pv, const double* qv) {
  const int ms = ctx->ms, nch = m->nch;
  hipStream_t s = ctx->stream;
  int nmo = 0;
  for (int c = 0; c < nch; ++c) nmo = std::max(nmo, nort[c]);
  if (nmo == 0) return DBS_OK;
  // the blocks each channel touches (dbs_orth.f90:53-72): block (r, c),
  // r >= c, gets the left projection of r and the right one of c
  std::vector<long long> blk;
  std::vector<int> lc, rc;
  for (int r = 1; r <= nch; ++r)
    for (int c = 1; c <= r; ++c) {
      if (idiag == 1 && r != c) continue;
      if (idiag == -1 && r == c) continue;
      const int b = m->icc[(r - 1) + (size_t)nch * (c - 1)];
      if (b == 0) continue;
      const bool L = nort[r - 1] > 0, R = nort[c - 1] > 0;
      if (!L && !R) continue;
      blk.push_back(b - 1);
      lc.push_back(L ? r - 1 : nch);
      rc.push_back(R ? c - 1 : nch);
    }
  std::vector<int> vcol, vch;
  if (idiag == 0)
    for (int ich = 1; ich <= nch; ++ich) {
      if (nort[ich - 1] == 0) continue;
      for (int ip = 1; ip <= m->npert; ++ip) {
        const int k = m->icb[(ich - 1) + (size_t)nch * (ip - 1)];
        if (k == 0) continue;
        DBS_CHECK(k <= m->ncb && m->hcp_dev, DBS_ERR_ARG, "DBS_ORTH: icb out of range or hcp missing");
        vcol.push_back(k - 1);
        vch.push_back(ich - 1);
      }
    }
  if (blk.empty() && vcol.empty()) return DBS_OK;
  Mem mem(s);
  int *d_nort, *d_iort;
  double *P, *Q;
  std::vector<int> hn(nort, nort + nch), hi(iort, iort + (size_t)nch * ldo);
  int rc0;
  if ((rc0 = mem.put(hn, &d_nort)) || (rc0 = mem.put(hi, &d_iort)) ||
      (rc0 = mem.alloc(&P, (size_t)ms * nmo * (nch + 1))) || (rc0 = mem.alloc(&Q, (size_t)ms * nmo * (nch + 1))))
    return rc0;
  k_pad_vectors<<<grid_n((long long)ms * nmo * (nch + 1), 256), 256, 0, s>>>(ms, nmo, nch, d_nort, d_iort, ldo, pv, qv,
                                                                            P, Q);
  if ((rc0 = hip_ok("DBS_ORTH"))) return rc0;
  if (!vcol.empty()) {
    int *d_col, *d_ch;
    if ((rc0 = mem.put(vcol, &d_col)) || (rc0 = mem.put(vch, &d_ch))) return rc0;
    k_orth_vec<<<(unsigned)vcol.size(), 256, 0, s>>>(ms, nmo, d_col, d_ch, P, Q, m->hcp_dev);
    if ((rc0 = hip_ok("DBS_ORTH"))) return rc0;
  }
  if (blk.empty()) return DBS_OK;
  rocblas_handle h;
  if ((rc0 = solver_of(ctx, &h))) return rc0;
  const size_t per = (size_t)ms * 2 * nmo * 2 + (size_t)nmo * nmo;
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(blk.size(), ctx->work_limit / (per * 8)));
  double *Lw, *Rw, *S;
  int *d_l, *d_r;
  const double **pp;
  if ((rc0 = mem.alloc(&Lw, (size_t)chunk * ms * 2 * nmo)) || (rc0 = mem.alloc(&Rw, (size_t)chunk * ms * 2 * nmo)) ||
      (rc0 = mem.alloc(&S, (size_t)chunk * nmo * nmo)) || (rc0 = mem.alloc(&d_l, chunk)) ||
      (rc0 = mem.alloc(&d_r, chunk)) || (rc0 = mem.alloc(&pp, (size_t)chunk * 9)))
    return rc0;
  const double one = 1.0, zero = 0.0, mone = -1.0;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  std::vector<const double*> hp((size_t)chunk * 9);
  for (size_t c0 = 0; c0 < blk.size(); c0 += chunk) {
    const int ne = (int)std::min<size_t>(chunk, blk.size() - c0);
    // pointer tables: X, P_L, P_R, Q_L, L, L2 (= V), R (= U), S
    const double** X = hp.data();
    const double** PL = X + chunk;
    const double** PR = PL + chunk;
    const double** QL = PR + chunk;
    const double** Lp = QL + chunk;
    const double** L2 = Lp + chunk;
    const double** Rp = L2 + chunk;
    const double** Sp = Rp + chunk;
    for (int e = 0; e < ne; ++e) {
      X[e] = m->hch_dev + (size_t)ms * ms * blk[c0 + e];
      PL[e] = P + (size_t)ms * nmo * lc[c0 + e];
      PR[e] = P + (size_t)ms * nmo * rc[c0 + e];
      QL[e] = Q + (size_t)ms * nmo * lc[c0 + e];
      Lp[e] = Lw + (size_t)ms * 2 * nmo * e;
      L2[e] = Lp[e] + (size_t)ms * nmo;
      Rp[e] = Rw + (size_t)ms * 2 * nmo * e;
      Sp[e] = S + (size_t)nmo * nmo * e;
    }
    std::vector<int> hl(lc.begin() + c0, lc.begin() + c0 + ne), hr(rc.begin() + c0, rc.begin() + c0 + ne);
    if (hipMemcpyAsync(pp, hp.data(), sizeof(double*) * chunk * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_l, hl.data(), sizeof(int) * ne, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_r, hr.data(), sizeof(int) * ne, hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(DBS_ERR_HIP, "DBS_ORTH: hipMemcpyAsync");
    const double* const* dX = pp;
    const double* const* dPL = pp + chunk;
    const double* const* dPR = pp + 2 * chunk;
    const double* const* dQL = pp + 3 * chunk;
    double* const* dL = (double* const*)(pp + 4 * chunk);
    double* const* dL2 = (double* const*)(pp + 5 * chunk);
    double* const* dR = (double* const*)(pp + 6 * chunk);
    double* const* dS = (double* const*)(pp + 7 * chunk);
    // V = X P_R into the second half of L; U = P_L^T X into rows 0..nmo-1 of R
    if ((rc0 = blas(rocblas_dgemm_batched(h, N, N, ms, nmo, ms, &one, dX, ms, dPR, ms, &zero, dL2, ms, ne),
                    "DBS_ORTH: X P")) ||
        (rc0 = blas(rocblas_dgemm_batched(h, T, N, nmo, ms, ms, &one, dPL, ms, dX, ms, &zero, dR, 2 * nmo, ne),
                    "DBS_ORTH: P^T X")) ||
        // S = P_L^T V; V <- V - Q_L S
        (rc0 = blas(rocblas_dgemm_batched(h, T, N, nmo, nmo, ms, &one, dPL, ms, (const double* const*)dL2, ms, &zero,
                                          dS, nmo, ne),
                    "DBS_ORTH: P^T X P")) ||
        (rc0 = blas(rocblas_dgemm_batched(h, N, N, ms, nmo, nmo, &mone, dQL, ms, (const double* const*)dS, nmo, &one,
                                          dL2, ms, ne),
                    "DBS_ORTH: V - Q S")))
      return rc0;
    k_fill_lr<<<grid_n((long long)ne * ms * nmo, 256), 256, 0, s>>>(ms, nmo, ne, d_l, d_r, Q, Lw, Rw);
    if ((rc0 = hip_ok("DBS_ORTH"))) return rc0;
    // X <- X - L R
    if ((rc0 = blas(rocblas_dgemm_batched(h, N, N, ms, ms, 2 * nmo, &mone, (const double* const*)dL, ms,
                                          (const double* const*)dR, 2 * nmo, &one, (double* const*)dX, ms, ne),
                    "DBS_ORTH: X - L R")))
      return rc0;
    // the next batch rewrites the pointer and channel tables
    if (c0 + chunk < blk.size() && hipStreamSynchronize(s) != hipSuccess)
      return fail(DBS_ERR_HIP, "DBS_ORTH: synchronize");
  }
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// Diag_channels
// ---------------------------------------------------------------------------
int diag_channels(dbs_ctx* ctx, const dbs_mat_blocks* m, double* dg, int nd, const int* iprm, double edmin,
                  double edmax, double egap, int* ipsol) {
  const int ms = ctx->ms, ns = ctx->ns, nch = m->nch;
  hipStream_t s = ctx->stream;
  ipsol[0] = 0;
  // the channels' kept sizes; one strided batch of the largest order with
  // the smaller problems padded by a decoupled (sigma I, I) block when a
  // sigma outside the window of modest size exists, else one batch per size
  std::vector<int> chans_all, kr_all;
  for (int ich = 1; ich <= nch; ++ich) {
    ipsol[ich] = 0;
    const int k = m->icc[(ich - 1) + (size_t)nch * (ich - 1)];
    if (k == 0) continue;
    DBS_CHECK(k <= m->nblocks && k <= nd, DBS_ERR_ARG, "Diag_channels: icc(ich,ich) beyond the stored blocks");
    int kk = 0;
    for (int i = 0; i < ms; ++i) kk += iprm[(size_t)(ich - 1) * ms + i] != 0;
    DBS_CHECK(kk > 0, DBS_ERR_ARG, "Diag_channels: every B-spline of a channel removed");
    chans_all.push_back(ich);
    kr_all.push_back(kk);
  }
  double sigma = 0.0;
  bool pad = false;
  {
    const double lim = 1e12;
    double best = HUGE_VAL;
    if (std::fabs(edmin) < lim) {
      const double c = edmin - std::max(1.0, std::fabs(edmin));
      if (std::fabs(c) < best) best = std::fabs(c), sigma = c, pad = true;
    }
    if (std::fabs(edmax) < lim) {
      const double c = edmax + std::max(1.0, std::fabs(edmax));
      if (std::fabs(c) < best) best = std::fabs(c), sigma = c, pad = true;
    }
  }
  std::map<int, std::vector<int>> groups;  // order -> positions in chans_all
  const int kmax = kr_all.empty() ? 0 : *std::max_element(kr_all.begin(), kr_all.end());
  for (size_t x = 0; x < chans_all.size(); ++x) groups[pad ? kmax : kr_all[x]].push_back((int)x);
  rocblas_handle h;
  int rc;
  if (!groups.empty() && (rc = solver_of(ctx, &h))) return rc;
  for (auto& g : groups) {
    const int kk = g.first;
    const std::vector<int>& pos = g.second;
    const size_t per = (size_t)kk * kk * 2 + (size_t)kk * 2 + ms;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(pos.size(), ctx->work_limit / (per * 8)));
    for (size_t c0 = 0; c0 < pos.size(); c0 += chunk) {
      const int nb = (int)std::min<size_t>(chunk, pos.size() - c0);
      std::vector<int> keep((size_t)nb * kk, -1), inv((size_t)nb * ms, -1), blk(nb), kr(nb), chans(nb);
      for (int e = 0; e < nb; ++e) {
        const int ich = chans_all[pos[c0 + e]];
        chans[e] = ich;
        kr[e] = kr_all[pos[c0 + e]];
        blk[e] = m->icc[(ich - 1) + (size_t)nch * (ich - 1)] - 1;
        int n = 0;
        for (int i = 0; i < ms; ++i)
          if (iprm[(size_t)(ich - 1) * ms + i] != 0) {
            inv[(size_t)e * ms + i] = n;
            keep[(size_t)e * kk + n++] = i;
          }
      }
      Mem mem(s);
      int *d_keep, *d_inv, *d_blk, *d_kr;
      double *A, *B, *D, *E;
      rocblas_int* info;
      if ((rc = mem.put(keep, &d_keep)) || (rc = mem.put(inv, &d_inv)) || (rc = mem.put(blk, &d_blk)) ||
          (rc = mem.put(kr, &d_kr)) ||
          (rc = mem.alloc(&A, (size_t)nb * kk * kk)) || (rc = mem.alloc(&B, (size_t)nb * kk * kk)) ||
          (rc = mem.alloc(&D, (size_t)nb * kk)) || (rc = mem.alloc(&E, (size_t)nb * kk)) ||
          (rc = mem.alloc(&info, nb)))
        return rc;
      k_compact_pair<<<grid_n((long long)nb * kk * kk, 256), 256, 0, s>>>(ms, kk, nb, d_keep, d_kr, d_blk, sigma,
                                                                          m->hch_dev, dg, A, B);
      if ((rc = hip_ok("Diag_channels"))) return rc;
      // DSYGV's steps (itype 1): B = L L^T, C = L^-1 A L^-T, C z = E z,
      // x = L^-T z; the reduction as two batched triangular solves (level 3)
      // rather than rocSOLVER's column-by-column sygst
      const rocblas_stride st2 = (rocblas_stride)kk * kk;
      const double one = 1.0;
      std::vector<rocblas_int> hinfo(nb), hinfo2(nb);
      if ((rc = blas(rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, kk, B, kk, st2, info, nb),
                     "Diag_channels: rocsolver_dpotrf")) ||
          (rc = to_host(hinfo.data(), info, sizeof(rocblas_int) * nb, s, "Diag_channels")))
        return rc;
      for (int e = 0; e < nb; ++e)
        DBS_CHECK(hinfo[e] == 0, DBS_ERR_ARG, "channel diagonalization failed");
      if ((rc = blas(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                                                   rocblas_diagonal_non_unit, kk, kk, &one, B, kk, st2, A, kk, st2, nb),
                     "Diag_channels: L^-1 A")) ||
          (rc = blas(rocblas_dtrsm_strided_batched(h, rocblas_side_right, rocblas_fill_lower,
                                                   rocblas_operation_transpose, rocblas_diagonal_non_unit, kk, kk, &one,
                                                   B, kk, st2, A, kk, st2, nb),
                     "Diag_channels: A L^-T")) ||
          (rc = blas(rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, kk, A, kk, st2, D,
                                                      kk, E, kk, info, nb),
                     "Diag_channels: rocsolver_dsyevd")) ||
          (rc = blas(rocblas_dtrsm_strided_batched(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                                                   rocblas_diagonal_non_unit, kk, kk, &one, B, kk, st2, A, kk, st2, nb),
                     "Diag_channels: L^-T z")))
        return rc;
      std::vector<double> ev((size_t)nb * kk);
      if ((rc = to_host(hinfo2.data(), info, sizeof(rocblas_int) * nb, s, "Diag_channels")) ||
          (rc = to_host(ev.data(), D, sizeof(double) * ev.size(), s, "Diag_channels")))
        return rc;
      for (int e = 0; e < nb; ++e)
        DBS_CHECK(hinfo2[e] == 0, DBS_ERR_ARG, "channel diagonalization failed");
      // the window (diag.f90:44-48), solutions in ascending order
      std::vector<int> se, scol, ss, ksol(nb, 0);
      std::vector<double> se_E;
      for (int e = 0; e < nb; ++e)
        for (int i = 0; i < kk; ++i) {
          const double w = ev[(size_t)e * kk + i];
          if (w < edmin || w > edmax || std::fabs(w) < egap) continue;
          se.push_back(e);
          scol.push_back(i);
          ss.push_back(ksol[e]++);
          se_E.push_back(w);
        }
      for (int e = 0; e < nb; ++e) ipsol[chans[e]] = ksol[e];
      if (se.empty()) continue;
      DBS_CHECK(*std::max_element(ksol.begin(), ksol.end()) <= ms, DBS_ERR_ARG, "Diag_channels: ksol > ms");
      int *d_se, *d_scol, *d_ss, *d_ksol;
      double* d_E;
      if ((rc = mem.put(se, &d_se)) || (rc = mem.put(scol, &d_scol)) || (rc = mem.put(ss, &d_ss)) ||
          (rc = mem.put(ksol, &d_ksol)) || (rc = mem.put(se_E, &d_E)))
        return rc;
      k_expand_solutions<<<(unsigned)se.size(), 256, 0, s>>>(ms, kk, d_se, d_scol, d_ss, d_blk, d_kr, d_inv, A, dg);
      k_store_energies<<<grid_n((long long)se.size(), 256), 256, 0, s>>>(ms, ns, (int)se.size(), d_se, d_ss, d_ksol,
                                                                         d_blk, d_E, dg);
      if ((rc = hip_ok("Diag_channels"))) return rc;
    }
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("Diag_channels: ") + hipGetErrorString(e));
  return DBS_OK;
}

// ---------------------------------------------------------------------------
// Record_matrix
// ---------------------------------------------------------------------------
int record_matrix(dbs_ctx* ctx, const dbs_mat_blocks* m, const double* dg, const int* ipsol, double* c_dev,
                  long long c_cap, long long* coff, double* w_dev, long long w_cap, long long* woff, double* ov) {
  const int ms = ctx->ms, nch = m->nch, np = m->npert;
  const long long nt = (long long)nch + np;
  hipStream_t s = ctx->stream;
  for (long long x = 0; x < (long long)nch * nch; ++x) coff[x] = -1;
  if (woff)
    for (long long x = 0; x < (long long)nch * np; ++x) woff[x] = -1;
  int nmax = 0;
  for (int c = 1; c <= nch; ++c) nmax = std::max(nmax, ipsol[c]);
  Mem mem(s);
  int rc;
  // channel-channel records (record_matrix.f90:12-26)
  std::vector<int> ri, rj, rk;
  for (int ich = 2; ich <= nch; ++ich)
    for (int jch = 1; jch < ich; ++jch) {
      const int k = m->icc[(ich - 1) + (size_t)nch * (jch - 1)];
      if (k == 0) continue;
      DBS_CHECK(k <= m->nblocks, DBS_ERR_ARG, "Record_matrix: icc out of range");
      ri.push_back(ich);
      rj.push_back(jch);
      rk.push_back(k - 1);
    }
  double* D = nullptr;
  int* d_nsol = nullptr;
  if (nmax > 0) {
    std::vector<int> nsol(ipsol + 1, ipsol + 1 + nch);
    if ((rc = mem.put(nsol, &d_nsol)) || (rc = mem.alloc(&D, (size_t)ms * nmax * nch))) return rc;
    k_pad_solutions<<<grid_n((long long)ms * nmax * nch, 256), 256, 0, s>>>(ms, nmax, nch, d_nsol, dg, D);
    if ((rc = hip_ok("Record_matrix"))) return rc;
  }
  if (!ri.empty()) {
    std::vector<long long> boff(ri.size());
    for (size_t e = 0; e < ri.size(); ++e) boff[e] = (long long)ms * ms * rk[e];
    long long* d_boff;
    double* d_mx;
    if ((rc = mem.put(boff, &d_boff)) || (rc = mem.alloc(&d_mx, ri.size()))) return rc;
    k_seg_maxabs<<<(unsigned)ri.size(), 256, 0, s>>>((long long)ms * ms, m->hch_dev, d_boff, d_mx);
    std::vector<double> mx(ri.size());
    if ((rc = hip_ok("Record_matrix")) || (rc = to_host(mx.data(), d_mx, sizeof(double) * mx.size(), s, "Record_matrix")))
      return rc;
    std::vector<int> wr;  // written records, in record order
    long long off = 0;
    for (size_t e = 0; e < ri.size(); ++e) {
      if (mx[e] < 1e-20) continue;
      wr.push_back((int)e);
      coff[(ri[e] - 1) + (size_t)nch * (rj[e] - 1)] = off;
      off += (long long)ipsol[ri[e]] * ipsol[rj[e]];
    }
    DBS_CHECK(off <= c_cap, DBS_ERR_CAPACITY, "Record_matrix: c_dev too small");
    DBS_CHECK(off == 0 || c_dev, DBS_ERR_ARG, "Record_matrix: c_dev missing");
    if (nmax > 0 && !wr.empty()) {
      rocblas_handle h;
      if ((rc = solver_of(ctx, &h))) return rc;
      const size_t per = (size_t)nmax * ms + (size_t)nmax * nmax;
      const int chunk = (int)std::max<size_t>(1, std::min<size_t>(wr.size(), ctx->work_limit / (per * 8)));
      double *W, *Cp, *d_cm;
      const double** pp;
      int *d_ni, *d_nj;
      long long* d_off;
      if ((rc = mem.alloc(&W, (size_t)chunk * nmax * ms)) || (rc = mem.alloc(&Cp, (size_t)chunk * nmax * nmax)) ||
          (rc = mem.alloc(&pp, (size_t)chunk * 5)) || (rc = mem.alloc(&d_ni, chunk)) || (rc = mem.alloc(&d_nj, chunk)) ||
          (rc = mem.alloc(&d_off, chunk)) || (rc = mem.alloc(&d_cm, wr.size())))
        return rc;
      const double one = 1.0, zero = 0.0;
      std::vector<const double*> hp((size_t)chunk * 5);
      std::vector<int> hni(chunk), hnj(chunk);
      std::vector<long long> hoff(chunk);
      for (size_t c0 = 0; c0 < wr.size(); c0 += chunk) {
        const int ne = (int)std::min<size_t>(chunk, wr.size() - c0);
        for (int e = 0; e < ne; ++e) {
          const int r = wr[c0 + e];
          hp[e] = D + (size_t)ms * nmax * (ri[r] - 1);                     // D_ich
          hp[chunk + e] = m->hch_dev + (size_t)ms * ms * rk[r];           // H
          hp[2 * chunk + e] = W + (size_t)nmax * ms * e;                  // W = D_ich^T H
          hp[3 * chunk + e] = D + (size_t)ms * nmax * (rj[r] - 1);        // D_jch
          hp[4 * chunk + e] = Cp + (size_t)nmax * nmax * e;               // C = W D_jch
          hni[e] = ipsol[ri[r]];
          hnj[e] = ipsol[rj[r]];
          hoff[e] = coff[(ri[r] - 1) + (size_t)nch * (rj[r] - 1)];
        }
        if (hipMemcpyAsync(pp, hp.data(), sizeof(double*) * chunk * 5, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d_ni, hni.data(), sizeof(int) * ne, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d_nj, hnj.data(), sizeof(int) * ne, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d_off, hoff.data(), sizeof(long long) * ne, hipMemcpyHostToDevice, s) != hipSuccess)
          return fail(DBS_ERR_HIP, "Record_matrix: hipMemcpyAsync");
        const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
        if ((rc = blas(rocblas_dgemm_batched(h, T, N, nmax, ms, ms, &one, pp, ms, pp + chunk, ms, &zero,
                                             (double* const*)(pp + 2 * chunk), nmax, ne),
                       "Record_matrix: D^T H")) ||
            (rc = blas(rocblas_dgemm_batched(h, N, N, nmax, nmax, ms, &one, pp + 2 * chunk, nmax, pp + 3 * chunk, ms,
                                             &zero, (double* const*)(pp + 4 * chunk), nmax, ne),
                       "Record_matrix: W D")))
          return rc;
        k_pack_records<<<(unsigned)ne, 256, 0, s>>>(nmax, d_ni, d_nj, d_off, Cp, c_dev, d_cm + c0);
        if ((rc = hip_ok("Record_matrix"))) return rc;
        // the next chunk rewrites the host tables: drain this one first
        if (c0 + chunk < wr.size() && hipStreamSynchronize(s) != hipSuccess)
          return fail(DBS_ERR_HIP, "Record_matrix: synchronize");
      }
      if (ov) {
        std::vector<double> cm(wr.size());
        if ((rc = to_host(cm.data(), d_cm, sizeof(double) * cm.size(), s, "Record_matrix"))) return rc;
        for (size_t e = 0; e < wr.size(); ++e) ov[(ri[wr[e]] - 1) + nt * (rj[wr[e]] - 1)] = cm[e];
      }
    } else if (ov) {
      for (int r : wr) ov[(ri[r] - 1) + nt * (rj[r] - 1)] = 0.0;
    }
  }
  if (np <= 0) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("Record_matrix: ") + hipGetErrorString(e));
    return DBS_OK;
  }
  // channel-perturber records (record_matrix.f90:34-45)
  std::vector<int> vi, vp, vk;
  for (int ich = 1; ich <= nch; ++ich)
    for (int ip = 1; ip <= np; ++ip) {
      const int k = m->icb[(ich - 1) + (size_t)nch * (ip - 1)];
      if (k == 0) continue;
      DBS_CHECK(k <= m->ncb && m->hcp_dev, DBS_ERR_ARG, "Record_matrix: icb out of range or hcp missing");
      vi.push_back(ich);
      vp.push_back(ip);
      vk.push_back(k - 1);
    }
  if (!vi.empty()) {
    DBS_CHECK(woff, DBS_ERR_ARG, "Record_matrix: woff missing");
    std::vector<long long> boff(vi.size());
    for (size_t e = 0; e < vi.size(); ++e) boff[e] = (long long)ms * vk[e];
    long long* d_boff;
    double* d_mx;
    if ((rc = mem.put(boff, &d_boff)) || (rc = mem.alloc(&d_mx, vi.size()))) return rc;
    k_seg_maxabs<<<(unsigned)vi.size(), 256, 0, s>>>((long long)ms, m->hcp_dev, d_boff, d_mx);
    std::vector<double> mx(vi.size());
    if ((rc = hip_ok("Record_matrix")) || (rc = to_host(mx.data(), d_mx, sizeof(double) * mx.size(), s, "Record_matrix")))
      return rc;
    std::vector<int> wch, wcol, wni;
    std::vector<long long> wo;
    std::vector<size_t> wsrc;
    long long off = 0;
    for (size_t e = 0; e < vi.size(); ++e) {
      if (mx[e] < 1e-20) continue;
      woff[(vi[e] - 1) + (size_t)nch * (vp[e] - 1)] = off;
      wch.push_back(vi[e] - 1);
      wcol.push_back(vk[e]);
      wni.push_back(ipsol[vi[e]]);
      wo.push_back(off);
      wsrc.push_back(e);
      off += ipsol[vi[e]];
    }
    DBS_CHECK(off <= w_cap, DBS_ERR_CAPACITY, "Record_matrix: w_dev too small");
    DBS_CHECK(off == 0 || w_dev, DBS_ERR_ARG, "Record_matrix: w_dev missing");
    if (off > 0) {
      int *d_ch, *d_col, *d_ni;
      long long* d_wo;
      double* d_wm;
      if ((rc = mem.put(wch, &d_ch)) || (rc = mem.put(wcol, &d_col)) || (rc = mem.put(wni, &d_ni)) ||
          (rc = mem.put(wo, &d_wo)) || (rc = mem.alloc(&d_wm, wch.size())))
        return rc;
      k_dt_vec<<<(unsigned)wch.size(), 256, 0, s>>>(ms, nmax, d_ch, d_col, d_ni, d_wo, D, m->hcp_dev, w_dev, d_wm);
      if ((rc = hip_ok("Record_matrix"))) return rc;
      if (ov) {
        std::vector<double> wm(wch.size());
        if ((rc = to_host(wm.data(), d_wm, sizeof(double) * wm.size(), s, "Record_matrix"))) return rc;
        for (size_t e = 0; e < wch.size(); ++e) ov[(nch + vp[wsrc[e]] - 1) + nt * (vi[wsrc[e]] - 1)] = wm[e];
      }
    } else if (ov) {
      for (size_t e = 0; e < wch.size(); ++e) ov[(nch + vp[wsrc[e]] - 1) + nt * (vi[wsrc[e]] - 1)] = 0.0;
    }
  }
  // perturber-perturber records (record_matrix.f90:49-55): hp itself
  if (ov && m->nbb > 0) {
    DBS_CHECK(m->hp_dev, DBS_ERR_ARG, "Record_matrix: hp missing");
    std::vector<double> hpv(m->nbb);
    if ((rc = to_host(hpv.data(), m->hp_dev, sizeof(double) * m->nbb, s, "Record_matrix"))) return rc;
    for (int ip = 1; ip <= np; ++ip)
      for (int jp = 1; jp <= ip; ++jp) {
        const int k = m->ibb[(ip - 1) + (size_t)np * (jp - 1)];
        if (k == 0) continue;
        DBS_CHECK(k <= m->nbb, DBS_ERR_ARG, "Record_matrix: ibb out of range");
        if (hpv[k - 1] == 0.0) continue;
        ov[(nch + ip - 1) + nt * (nch + jp - 1)] = std::fabs(hpv[k - 1]);
      }
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("Record_matrix: ") + hipGetErrorString(e));
  return DBS_OK;
}

}  // namespace
}  // namespace dbs

using namespace dbs;

extern "C" {

int dbs_zero_cond(int nch, int ns, int nsp, int nsq, int ksp, int ksq, const int* lch, const int* kch, int ilzero,
                  int jlzero, int ibzero, int jbzero, int npert, int* iprm, int* nhm) {
  DBS_CHECK(nch >= 0 && ns > 0 && npert >= 0 && iprm && nhm && (nch == 0 || (lch && kch)), DBS_ERR_ARG,
            "Zero_cond: bad argument");
  const int ms = 2 * ns;
  for (long long x = 0; x < (long long)nch * ms + npert; ++x) iprm[x] = 1;
  int ip = 0, jp = 0;
  for (int ich = 0; ich < nch; ++ich) {
    int il, jl = 0, mi = 0, mj = 0;
    if (ilzero >= 0) {
      il = ilzero;
    } else {
      il = lch[ich] + 1;
      if (il > ksp - 1) {
        il = 1;
        mi = 1;
      }
    }
    if (jlzero >= 0) {
      jl = jlzero;
    } else {
      DBS_CHECK(kch[ich] != 0, DBS_ERR_ARG, "Zero_cond: kappa = 0");
      jl = kch[ich] < 0 ? lch[ich] + 2 : lch[ich];
      if (jl > ksq - 1) {
        jl = 1;
        mj = 1;
      }
    }
    if (mi == 1 || mj == 1) il = jl = 1;
    const int lo[2] = {il, jl}, hi[2] = {nsp + 1 - ibzero, nsq + 1 - jbzero};
    for (int c = 0; c < 2; ++c)
      for (int i = 1; i <= ns; ++i) {
        if (i <= lo[c] || i >= hi[c]) iprm[ip] = 0;
        jp += iprm[ip++];
      }
  }
  *nhm = jp + npert;
  return DBS_OK;
}

int dbs_orth_conditions(int nch, const int* ipch, int nbf, const int* ipbs, const int* kbs, dbs_ibort_fn ibort,
                        void* ibort_user, int ldo, int* nort, int* iort) {
  DBS_CHECK(nch >= 0 && nbf >= 0 && ldo >= 0 && (nch == 0 || (ipch && nort && (ldo == 0 || iort))) &&
                (nbf == 0 || (ipbs && kbs)) && ibort,
            DBS_ERR_ARG, "DBS_ORTH: bad argument");
  for (int ich = 0; ich < nch; ++ich) {
    const int i = ipch[ich];
    DBS_CHECK(i >= 1 && i <= nbf, DBS_ERR_ARG, "DBS_ORTH: channel orbital out of range");
    int n = 0;
    for (int j = 1; j <= nbf; ++j) {
      if (ipbs[j - 1] > 0) continue;
      if (kbs[j - 1] != kbs[i - 1]) continue;
      if (ibort(ibort_user, i, j) != 0) continue;
      DBS_CHECK(n < ldo, DBS_ERR_CAPACITY, "DBS_ORTH: more conditions than ldo");
      iort[(size_t)ich * ldo + n++] = j;
    }
    nort[ich] = n;
  }
  return DBS_OK;
}

int dbs_orth_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, int idiag, const int* nort, const int* iort, int ldo,
                 const double* pv_dev, const double* qv_dev, int nbf) {
  DBS_CHECK(ctx && nort && (iort || ldo == 0), DBS_ERR_ARG, "DBS_ORTH: null argument");
  int rc = check_blocks(m);
  if (rc) return rc;
  DBS_CHECK(idiag >= -1 && idiag <= 1, DBS_ERR_ARG, "DBS_ORTH: idiag not in -1..1");
  for (int c = 0; c < m->nch; ++c) {
    DBS_CHECK(nort[c] >= 0 && nort[c] <= ldo, DBS_ERR_ARG, "DBS_ORTH: nort out of range");
    for (int a = 0; a < nort[c]; ++a)
      DBS_CHECK(iort[(size_t)c * ldo + a] >= 1 && iort[(size_t)c * ldo + a] <= nbf, DBS_ERR_ARG,
                "DBS_ORTH: orbital index out of range");
    DBS_CHECK(nort[c] == 0 || (pv_dev && qv_dev), DBS_ERR_ARG, "DBS_ORTH: orbital vectors missing");
  }
  DBS_CHECK(m->nblocks == 0 || m->hch_dev, DBS_ERR_ARG, "DBS_ORTH: hch missing");
  return orth(ctx, m, idiag, nort, iort, ldo, pv_dev, qv_dev);
}

int dbs_diag_channels_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, double* diag_dev, int nd, const int* iprm,
                          double edmin, double edmax, double egap, int* ipsol) {
  DBS_CHECK(ctx && iprm && ipsol && (diag_dev || nd == 0), DBS_ERR_ARG, "Diag_channels: null argument");
  int rc = check_blocks(m);
  if (rc) return rc;
  DBS_CHECK(m->nblocks == 0 || m->hch_dev, DBS_ERR_ARG, "Diag_channels: hch missing");
  return diag_channels(ctx, m, diag_dev, nd, iprm, edmin, edmax, egap, ipsol);
}

int dbs_record_matrix_dev(dbs_ctx* ctx, const dbs_mat_blocks* m, const double* diag_dev, int nd, const int* ipsol,
                          double* c_dev, long long c_cap, long long* coff, double* w_dev, long long w_cap,
                          long long* woff, double* overlaps) {
  DBS_CHECK(ctx && ipsol && coff, DBS_ERR_ARG, "Record_matrix: null argument");
  int rc = check_blocks(m);
  if (rc) return rc;
  DBS_CHECK(m->nblocks == 0 || m->hch_dev, DBS_ERR_ARG, "Record_matrix: hch missing");
  DBS_CHECK(nd >= m->nch && (diag_dev || m->nch == 0), DBS_ERR_ARG, "Record_matrix: diag holds fewer than nch blocks");
  for (int c = 1; c <= m->nch; ++c)
    DBS_CHECK(ipsol[c] >= 0 && ipsol[c] <= ctx->ms, DBS_ERR_ARG, "Record_matrix: ipsol out of range");
  return record_matrix(ctx, m, diag_dev, ipsol, c_dev, c_cap, coff, w_dev, w_cap, woff, overlaps);
}

}  // extern "C"

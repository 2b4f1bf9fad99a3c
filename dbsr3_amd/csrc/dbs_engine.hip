// dbs_engine.hip -- batched evaluation of B-spline Slater-integral sets and
// their adjoint on MI355X.
//
// For one multipole k the reference's cell algorithm (mrk_*, triang_*,
// dbsr_lib_dbs.f90:2898-3597) makes R^k(a,b;c,d) a symmetric bilinear form
// in the pair densities rho_L = f_a f_c (r1) and rho_R = f_b f_d (r2):
//   off-diagonal cells iv<jv:  M_L^k(iv) * M_R^{-(k+1)}(jv)   (and mirrored)
//   diagonal cells iv=jv:      the nested ks x ks Gauss rule of triang_*.
// Contracting the reference's per-pair moment tiles with the densities
// first gives, per interval, the moments M^k, M^{-(k+1)} and the nested
// inner integrals Y(iv,m) = int_{t_iv}^{r_m} rho r^k dr at the ks outer
// points -- the same nodes, weights and cells, summed in another order.
// With S(iv) = sum_{iv'<iv} M^k(iv') (a prefix scan) every density has a
// feature vector
//   F = [ S(iv) | M^{-(k+1)}(iv) | W(iv,m) rho(iv,m) | Y(iv,m) ]   (2nv(1+ks))
// with W = grw*gr^{-(k+1)}, and
//   R^k(L,R) = S_L.Mm_R + Mm_L.S_R + Ro_L.Y_R + Y_L.Ro_R .
// The four P/Q types of rk() add up to one form in the total density
// P_aP_c + Q_aQ_c.  S^k (msk_ppqq/pqqp, cells iv<=jv only) keeps the two
// one-sided products S_L.Mm_R + Y_L.Ro_R.
//
// Phases (one persistent launch each, all instances together):
//   orbital values at the outer and nested Gauss points
//   -> per-interval moments and inner integrals of every needed density
//   -> wavefront prefix scans -> integral contraction per row (sparse) or
//   per density tile on the FP64 matrix cores (dense sets)
//   -> adjoint: gather partner features, suffix scans, chain rule to the
//   orbital point values, projection onto the B-spline coefficients.
// No floating-point atomics: every output is owned by one thread/wave and
// summed in a fixed order, so results are bitwise reproducible.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>

#include "dbs_device.h"
#include "dbs_internal.h"

namespace dbs {

enum DensKind { DK_TOT = 0, DK_PP = 1, DK_QQ = 2, DK_PQ = 3 };

// per-instance geometry on the device
struct InstDev {
  int ns, ks, ksp, ksq, nv, norb, np;  // np = ks + ks*ks points per interval
  int orb_base;                        // global index of the instance's first orbital
  int kg0, nkg;                        // the instance's (instance,k) groups kg0..kg0+nkg-1
  long long off_pq;                    // into pq/grad buffers
  long long off_orbv;                  // orbital point values
  const double* grw;
  const double* gr;
  const double* grm;
  const double* pbsp;
  const double* qbsp;
  const double* xin;
  const double* win;
  const double* bpin;
  const double* bqin;
};

// one (instance, k) group of densities
struct KGroupDev {
  int inst, k, ndens, dens0;   // densities dens0 .. dens0+ndens-1
  long long off_w;             // weight tables: Wk(nv*ks), Wm(nv*ks), wkin(nv*ks*ks)
};

struct DensDev {
  int kg, a, c, kind;          // orbitals 0-based
  long long off_f;             // feature vector offset (F and G buffers)
};

struct RowDev {
  int L, R, kg, pad;           // global density indices
};

// adjacency for the adjoint: partner density, row, role (0: this is L)
struct AdjDev {
  int partner, row, role, pad;
};

// orbital -> densities adjacency: density, partner orbital, factor, comp-mask
struct OrbAdjDev {
  int dens, partner, factor, kind;
};

__host__ __device__ inline int feat_dim(int nv, int ks) { return 2 * nv + 2 * nv * ks; }

// ---------------------------------------------------------------------------
// phase 1: orbital values at the points of every interval
// value (inst, orb, comp, iv, p):  p < ks outer point m=p; p >= ks inner (m,i)
// ---------------------------------------------------------------------------
__device__ inline int find_seg(const long long* __restrict__ seg, int n, long long q) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (seg[mid] <= q) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_orb_points(int ninst, const long long* __restrict__ seg,
                             const InstDev* __restrict__ inst, const double* __restrict__ pq,
                             double* __restrict__ orbv, long long total) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int s = find_seg(seg, ninst, q);
    const InstDev I = inst[s];
    long long r = q - seg[s];
    const int p = (int)(r % I.np);
    r /= I.np;
    const int iv = (int)(r % I.nv);
    r /= I.nv;
    const int comp = (int)(r % 2);
    const int orb = (int)(r / 2);
    const int kx = comp == 0 ? I.ksp : I.ksq;
    const double* coef = pq + I.off_pq + (long long)I.ns * (comp + 2LL * orb) + iv;
    double v = 0.0;
    if (p < I.ks) {
      const double* xb = comp == 0 ? I.pbsp : I.qbsp;
      const int nw = I.nv + 1;
      for (int j = 0; j < kx; ++j) v += xb[iv + (long long)nw * (p + (long long)I.ks * j)] * coef[j];
    } else {
      const int mi = p - I.ks;  // m*ks + i
      const double* b = (comp == 0 ? I.bpin : I.bqin) + ((long long)iv * I.ks * I.ks + mi) * kx;
      for (int j = 0; j < kx; ++j) v += b[j] * coef[j];
    }
    orbv[I.off_orbv + q - seg[s]] = v;
  }
}

// weight tables of a (instance,k) group: Wk = grw*gr^k, Wm = grw*gr^-(k+1),
// wkin = win*xin^k
__global__ void k_weights(int nkg, const long long* __restrict__ seg,
                          const KGroupDev* __restrict__ kgs, const InstDev* __restrict__ inst,
                          double* __restrict__ wt, long long total) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int g = find_seg(seg, nkg, q);
    const KGroupDev G = kgs[g];
    const InstDev I = inst[G.inst];
    const long long r = q - seg[g];
    const int nvk = I.nv * I.ks;
    double v;
    if (r < nvk) {  // Wk(iv,m), stored [iv*ks + m]
      int iv = (int)(r / I.ks), m = (int)(r % I.ks);
      v = I.grw[iv + I.nv * m] * ipow(I.gr[iv + I.nv * m], G.k);
    } else if (r < 2 * nvk) {
      int rr = (int)(r - nvk);
      int iv = rr / I.ks, m = rr % I.ks;
      v = I.grw[iv + I.nv * m] * ipow(I.grm[iv + I.nv * m], G.k + 1);
    } else {
      long long rr = r - 2 * nvk;  // [iv][m][i]
      v = I.win[rr] * ipow(I.xin[rr], G.k);
    }
    wt[G.off_w + r] = v;
  }
}

// density value at point p of interval iv
__device__ inline double dens_point(const double* __restrict__ va, const double* __restrict__ vc,
                                    long long compstride, int kind, int p) {
  // va/vc point at comp 0 of the orbital's interval block; comp 1 is +compstride
  switch (kind) {
    case DK_TOT: return va[p] * vc[p] + va[compstride + p] * vc[compstride + p];
    case DK_PP: return va[p] * vc[p];
    case DK_QQ: return va[compstride + p] * vc[compstride + p];
    default: return va[p] * vc[compstride + p];  // P_a Q_c
  }
}

// ---------------------------------------------------------------------------
// phase 2: per-interval features.  One workgroup per (group, iv) item,
// threads over the group's densities.  Writes Mk(iv) into the S slot (the
// scan turns it into the exclusive prefix), Mm(iv), Ro(iv,m), Y(iv,m).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_features(
    int nitems, const int2* __restrict__ items, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const DensDev* __restrict__ dens,
    const double* __restrict__ orbv, const double* __restrict__ wt, double* __restrict__ F,
    int stage) {
  // sw: Wk(ks) Wm(ks) wkin(ks*ks), then (stage) the interval's orbital
  // values [orb][comp][np]
  extern __shared__ double sw[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int g = items[it].x, iv = items[it].y;
    const KGroupDev G = kgs[g];
    const InstDev I = inst[G.inst];
    const int ks = I.ks, nv = I.nv, np = I.np;
    const int nw = 2 * ks + ks * ks;
    __syncthreads();
    for (int q = threadIdx.x; q < nw; q += blockDim.x) {
      double v;
      if (q < ks) v = wt[G.off_w + (long long)iv * ks + q];
      else if (q < 2 * ks) v = wt[G.off_w + (long long)nv * ks + (long long)iv * ks + (q - ks)];
      else v = wt[G.off_w + 2LL * nv * ks + (long long)iv * ks * ks + (q - 2 * ks)];
      sw[q] = v;
    }
    double* sv = sw + nw;
    if (stage) {
      const int per = 2 * I.norb * np;
      for (int q = threadIdx.x; q < per; q += blockDim.x) {
        const int p = q % np, oc = q / np;  // oc = 2*orb + comp
        sv[q] = orbv[I.off_orbv + ((long long)oc * nv + iv) * np + p];
      }
    }
    __syncthreads();
    const double* Wk = sw;
    const double* Wm = sw + ks;
    const double* wk = sw + 2 * ks;
    const long long cstride = stage ? (long long)np : (long long)nv * np;
    for (int d = threadIdx.x; d < G.ndens; d += blockDim.x) {
      const DensDev D = dens[G.dens0 + d];
      const double* va = stage ? sv + (2 * D.a) * np
                               : orbv + I.off_orbv + ((long long)(2 * D.a) * nv + iv) * np;
      const double* vc = stage ? sv + (2 * D.c) * np
                               : orbv + I.off_orbv + ((long long)(2 * D.c) * nv + iv) * np;
      double* f = F + D.off_f;
      double mk = 0.0, mm = 0.0;
      for (int m = 0; m < ks; ++m) {
        double ro = dens_point(va, vc, cstride, D.kind, m);
        mk += Wk[m] * ro;
        mm += Wm[m] * ro;
        f[2 * nv + iv * ks + m] = Wm[m] * ro;
        double y = 0.0;
        for (int i = 0; i < ks; ++i) y += wk[m * ks + i] * dens_point(va, vc, cstride, D.kind, ks + m * ks + i);
        f[2 * nv + nv * ks + iv * ks + m] = y;
      }
      f[iv] = mk;
      f[nv + iv] = mm;
    }
  }
}

// exclusive prefix (dir=0) or suffix (dir=1) scan of nv values, one wave
// per density; each lane owns a contiguous chunk.
__device__ inline void wave_scan_excl(double* x, int nv, int dir) {
  const int lane = threadIdx.x & 63;
  const int chunk = (nv + 63) / 64;
  const int lo = lane * chunk, hi = min(lo + chunk, nv);
  double sum = 0.0;
  if (dir == 0) { for (int i = lo; i < hi; ++i) sum += x[i]; }
  else { for (int i = hi - 1; i >= lo; --i) sum += x[i]; }
  // exclusive scan of lane sums in lane order (dir 0) or reverse (dir 1)
  double incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    double v = dir == 0 ? __shfl_up(incl, off, 64) : __shfl_down(incl, off, 64);
    bool take = dir == 0 ? lane >= off : lane + off < 64;
    if (take) incl += v;
  }
  double run = incl - sum;
  if (dir == 0) {
    for (int i = lo; i < hi; ++i) { double v = x[i]; x[i] = run; run += v; }
  } else {
    for (int i = hi - 1; i >= lo; --i) { double v = x[i]; x[i] = run; run += v; }
  }
}

__global__ void k_scan(int ndens, const DensDev* __restrict__ dens, const KGroupDev* __restrict__ kgs,
                       const InstDev* __restrict__ inst, double* __restrict__ F, int dir) {
  const int waves = blockDim.x >> 6;
  for (int d = blockIdx.x * waves + (threadIdx.x >> 6); d < ndens; d += gridDim.x * waves) {
    const DensDev D = dens[d];
    const int nv = inst[kgs[D.kg].inst].nv;
    wave_scan_excl(F + D.off_f, nv, dir);
  }
}

// ---------------------------------------------------------------------------
// phase 3: contraction, one wavefront per row
// ---------------------------------------------------------------------------
__global__ void k_contract_rows(int nrows, const RowDev* __restrict__ rows,
                                const DensDev* __restrict__ dens, const KGroupDev* __restrict__ kgs,
                                const InstDev* __restrict__ inst, const double* __restrict__ F,
                                const long long* __restrict__ row_out, int onesided,
                                double* __restrict__ vals) {
  const int waves = blockDim.x >> 6;
  const int lane = threadIdx.x & 63;
  for (int r = blockIdx.x * waves + (threadIdx.x >> 6); r < nrows; r += gridDim.x * waves) {
    const RowDev R = rows[r];
    const DensDev DL = dens[R.L], DR = dens[R.R];
    const InstDev I = inst[kgs[R.kg].inst];
    const int nv = I.nv, nk = nv * I.ks;
    const double* fl = F + DL.off_f;
    const double* fr = F + DR.off_f;
    double acc = 0.0;
    if (!onesided) {
      for (int x = lane; x < nv; x += 64) acc += fl[x] * fr[nv + x] + fl[nv + x] * fr[x];
      const double* rl = fl + 2 * nv;
      const double* rr = fr + 2 * nv;
      for (int x = lane; x < nk; x += 64) acc += rl[x] * rr[nk + x] + rl[nk + x] * rr[x];
    } else {
      for (int x = lane; x < nv; x += 64) acc += fl[x] * fr[nv + x];
      const double* rl = fl + 2 * nv;
      const double* rr = fr + 2 * nv;
      for (int x = lane; x < nk; x += 64) acc += rl[nk + x] * rr[x];
    }
    acc = wave_sum(acc);
    if (lane == 0) vals[row_out[r]] = acc;
  }
}

// ---------------------------------------------------------------------------
// dense contraction on the FP64 matrix cores.  For a group whose rows cover
// a large part of all density pairs, G = F * F_sigma^T over the feature
// dimension is one GEMM; v_mfma_f64_16x16x4f64 tiles of 16x16 outputs.
// One wavefront computes a 16x16 tile of G and scatters the requested pairs.
// ---------------------------------------------------------------------------
typedef double __attribute__((ext_vector_type(4))) dbl4;

__device__ inline double fsig(const double* f, int nv, int nk, int x, int onesided) {
  // sigma-permuted feature x of F_R (the partner vector of an L feature)
  // x in [0,nv): S_L partners Mm_R; [nv,2nv): Mm_L partners S_R (0 if onesided)
  // [2nv, 2nv+nk): Ro_L partners Y_R (0 if onesided); [2nv+nk, ...): Y_L partners Ro_R
  if (x < nv) return f[nv + x];
  if (x < 2 * nv) return onesided ? 0.0 : f[x - nv];
  int y = x - 2 * nv;
  if (y < nk) return onesided ? 0.0 : f[2 * nv + nk + y];
  return f[2 * nv + (y - nk)];
}

__global__ void __launch_bounds__(256) k_contract_dense(
    int ntiles, const int4* __restrict__ tiles, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const DensDev* __restrict__ dens,
    const double* __restrict__ F, const int* __restrict__ pairrow, const long long* __restrict__ pairoff,
    const int* __restrict__ nextrow, const long long* __restrict__ row_out, int onesided,
    double* __restrict__ vals) {
  // tile = (group, Ltile, Rtile, 0); pairrow[pairoff[g] + l*ndens + r] = row or -1
  const int waves = blockDim.x >> 6;
  const int lane = threadIdx.x & 63;
  for (int t = blockIdx.x * waves + (threadIdx.x >> 6); t < ntiles; t += gridDim.x * waves) {
    const int4 T = tiles[t];
    const KGroupDev G = kgs[T.x];
    const InstDev I = inst[G.inst];
    const int nv = I.nv, nk = nv * I.ks, D = feat_dim(nv, I.ks);
    const int l0 = T.y * 16, r0 = T.z * 16;
    // A: 16 x 4 slice (rows l0+lane%16, cols x0 + lane/16); B: 4 x 16 slice
    const int li = l0 + (lane & 15), ri = r0 + (lane & 15), kq = lane >> 4;
    const bool lv = li < G.ndens, rv = ri < G.ndens;
    const double* fl = F + dens[G.dens0 + (lv ? li : 0)].off_f;
    const double* fr = F + dens[G.dens0 + (rv ? ri : 0)].off_f;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int x0 = 0; x0 < D; x0 += 4) {
      const int x = x0 + kq;
      double a = (lv && x < D) ? fl[x] : 0.0;
      double b = (rv && x < D) ? fsig(fr, nv, nk, x, onesided) : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    // output layout of the 16x16x4 f64 MFMA (one output group per
    // register): lane holds column lane%16 and rows 4*j + lane/16, j = 0..3
    const int col = lane & 15;
    for (int j = 0; j < 4; ++j) {
      const int row = 4 * j + (lane >> 4);
      const int L = l0 + row, R = r0 + col;
      if (L < G.ndens && R < G.ndens) {
        for (int rr = pairrow[pairoff[T.x] + (long long)L * G.ndens + R]; rr >= 0; rr = nextrow[rr])
          vals[row_out[rr]] = acc[j];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// energy: E_inst = sum_rows coef*value.  Blocks reduce fixed chunks of one
// instance's rows (tree in a fixed shape); one thread per instance then adds
// the chunk partials in order.
// ---------------------------------------------------------------------------
#define DBS_ECHUNK 2048
__global__ void k_energy_partial(int nchunks, const int2* __restrict__ chunks,
                                 const long long* __restrict__ roff, const double* __restrict__ coef,
                                 const double* __restrict__ vals, double* __restrict__ part) {
  __shared__ double sh[256];
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int2 C = chunks[ch];  // (instance, chunk index within it)
    const long long r0 = roff[C.x] + (long long)C.y * DBS_ECHUNK;
    const long long r1 = min(r0 + DBS_ECHUNK, roff[C.x + 1]);
    double acc = 0.0;
    for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) acc += coef[r] * vals[r];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[ch] = sh[0];
    __syncthreads();
  }
}

__global__ void k_energy_final(int ninst, const int* __restrict__ choff, const double* __restrict__ part,
                               double* __restrict__ energy) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ninst; s += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int c = choff[s]; c < choff[s + 1]; ++c) acc += part[c];
    energy[s] = acc;
  }
}

// ---------------------------------------------------------------------------
// adjoint phase A: G_d = sum over rows touching density d of coef * partner
// vector.  Each workgroup owns one density (all feature elements), walks
// its adjacency list in row order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_adj_gather(
    int ndens, const DensDev* __restrict__ dens, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const int* __restrict__ adjoff, const AdjDev* __restrict__ adj,
    const double* __restrict__ coef, const long long* __restrict__ row_out,
    const double* __restrict__ F, int onesided, double* __restrict__ Gb) {
  for (int d = blockIdx.x; d < ndens; d += gridDim.x) {
    const DensDev Dd = dens[d];
    const InstDev I = inst[kgs[Dd.kg].inst];
    const int nv = I.nv, nk = nv * I.ks, D = feat_dim(nv, I.ks);
    double* g = Gb + Dd.off_f;
    const int a0 = adjoff[d], a1 = adjoff[d + 1];
    for (int x = threadIdx.x; x < D; x += blockDim.x) {
      double acc = 0.0;
      for (int e = a0; e < a1; ++e) {
        const AdjDev A = adj[e];
        const double c = coef[row_out[A.row]];
        const double* fp = F + dens[A.partner].off_f;
        double v;
        if (!onesided) {
          v = fsig(fp, nv, nk, x, 0);
        } else if (A.role == 0) {
          // d is L: dS_L <- Mm_R, dY_L <- Ro_R
          v = (x < nv) ? fp[nv + x] : (x >= 2 * nv + nk ? fp[2 * nv + (x - 2 * nv - nk)] : 0.0);
        } else {
          // d is R: dMm_R <- S_L, dRo_R <- Y_L
          v = (x >= nv && x < 2 * nv) ? fp[x - nv]
                                      : ((x >= 2 * nv && x < 2 * nv + nk) ? fp[x + nk] : 0.0);
        }
        acc += c * v;
      }
      g[x] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// panel forms.  The feature dimension of a group is cut into panels whose
// features for all the group's densities fit one XCD's L2; every work item
// on a panel is dispatched to the same XCD (workgroup i runs on XCD i % 8),
// so partner features are read from HBM/MALL once per panel and then served
// from that XCD's L2.  Items with d < 0 are padding.
// ---------------------------------------------------------------------------
struct PanelItem {
  int d;       // density (gather) or first row (contraction)
  int n;       // rows in this item (contraction)
  int x0, x1;  // feature range
  int tile;    // panel index within the group (contraction partials)
  int nv, nk;  // intervals and nv*ks of the group's grid
  int pad;
};

// per adjacency entry: partner feature offset, coefficient index, role
struct AdjPre {
  long long off;
  long long cidx;
  int role, pad;
};

__global__ void __launch_bounds__(256) k_adj_gather_panel(
    int nitems, const PanelItem* __restrict__ items, const long long* __restrict__ doff,
    const int* __restrict__ adjoff, const AdjPre* __restrict__ adj, const double* __restrict__ coef,
    const double* __restrict__ F, int onesided, double* __restrict__ Gb) {
  __shared__ long long s_off[256];
  __shared__ double s_c[256];
  __shared__ int s_role[256];
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const PanelItem P = items[it];
  if (P.d < 0) return;
  const int nv = P.nv, nk = P.nk;
  double* g = Gb + doff[P.d];
  const int a0 = adjoff[P.d], a1 = adjoff[P.d + 1];
  // thread-owned accumulators for x = x0 + tid + k*blockDim
  constexpr int XMAX = 8;
  double acc[XMAX];
  for (int q = 0; q < XMAX; ++q) acc[q] = 0.0;
  for (int c0 = a0; c0 < a1; c0 += 256) {
    const int cn = min(256, a1 - c0);
    __syncthreads();
    if ((int)threadIdx.x < cn) {
      const AdjPre A = adj[c0 + threadIdx.x];
      s_off[threadIdx.x] = A.off;
      s_c[threadIdx.x] = coef[A.cidx];
      s_role[threadIdx.x] = A.role;
    }
    __syncthreads();
    for (int e = 0; e < cn; ++e) {
      const double* fp = F + s_off[e];
      const double c = s_c[e];
      const int role = s_role[e];
#pragma unroll
      for (int q = 0; q < XMAX; ++q) {
        const int x = P.x0 + threadIdx.x + q * blockDim.x;
        if (x >= P.x1) break;
        double v;
        if (!onesided) {
          v = fsig(fp, nv, nk, x, 0);
        } else if (role == 0) {
          v = (x < nv) ? fp[nv + x] : (x >= 2 * nv + nk ? fp[2 * nv + (x - 2 * nv - nk)] : 0.0);
        } else {
          v = (x >= nv && x < 2 * nv) ? fp[x - nv] : ((x >= 2 * nv && x < 2 * nv + nk) ? fp[x + nk] : 0.0);
        }
        acc[q] += c * v;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < XMAX; ++q) {
    const int x = P.x0 + threadIdx.x + q * blockDim.x;
    if (x < P.x1) g[x] = acc[q];
  }
}

// contraction partials: one wavefront per row of the item, lanes over the
// panel's feature range; part[row * maxtile + tile].  roff[r] = feature
// offsets of the row's two densities.
__global__ void __launch_bounds__(256) k_contract_panel(
    int nitems, const PanelItem* __restrict__ items, const longlong2* __restrict__ roff,
    const double* __restrict__ F, int onesided, int maxtile, double* __restrict__ part) {
  const int it = blockIdx.x;
  if (it >= nitems) return;
  const PanelItem P = items[it];
  if (P.d < 0) return;
  const int nv = P.nv, nk = P.nk;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int q = wv; q < P.n; q += nw) {
    const int r = P.d + q;
    const longlong2 o = roff[r];
    const double* fl = F + o.x;
    const double* fr = F + o.y;
    double acc = 0.0;
    for (int x = P.x0 + lane; x < P.x1; x += 64) acc += fl[x] * fsig(fr, nv, nk, x, onesided);
    acc = wave_sum(acc);
    if (lane == 0) part[(long long)r * maxtile + P.tile] = acc;
  }
}

__global__ void k_contract_reduce(int nrows, const int* __restrict__ ntile, int maxtile,
                                  const double* __restrict__ part, const long long* __restrict__ row_out,
                                  double* __restrict__ vals) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int t = 0; t < ntile[r]; ++t) acc += part[(long long)r * maxtile + t];
    vals[row_out[r]] = acc;
  }
}

// adjoint phase C: gradient with respect to the orbital values at every
// point.  For density d = (a,c): rho(p) = sum over comps of f_a f_c, and
//   g_rho(outer m) = G_Mk(iv)*Wk(m) + (G_Mm(iv) + G_Ro(iv,m))*Wm(m)
//   g_rho(inner m,i) = G_Y(iv,m)*wkin(m,i)
// where G_Mk(iv) = sum_{iv'>iv} G_S(iv') (suffix scan done by k_scan).
__global__ void k_adj_orb_points(int ninst, const long long* __restrict__ seg,
                                 const InstDev* __restrict__ inst, const KGroupDev* __restrict__ kgs,
                                 const DensDev* __restrict__ dens, const int* __restrict__ oadjoff,
                                 const OrbAdjDev* __restrict__ oadj, const double* __restrict__ orbv,
                                 const double* __restrict__ wt, const double* __restrict__ Gb,
                                 double* __restrict__ gorb, long long total) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int s = find_seg(seg, ninst, q);
    const InstDev I = inst[s];
    long long r = q - seg[s];
    const int p = (int)(r % I.np);
    r /= I.np;
    const int iv = (int)(r % I.nv);
    r /= I.nv;
    const int comp = (int)(r % 2);
    const int orb = (int)(r / 2);
    const int ks = I.ks, nv = I.nv, nk = nv * ks;
    double acc = 0.0;
    const int gslot = I.orb_base + orb;
    const int e0 = oadjoff[gslot], e1 = oadjoff[gslot + 1];
    const int m = p < ks ? p : (p - ks) / ks;
    const int i = p < ks ? -1 : (p - ks) % ks;
    for (int e = e0; e < e1; ++e) {
      const OrbAdjDev A = oadj[e];
      const DensDev D = dens[A.dens];
      // component selection: TOT both comps, PP only P, QQ only Q, PQ: P for a, Q for c
      double partner;
      const double* vp = orbv + I.off_orbv + ((long long)(2 * A.partner) * nv + iv) * I.np;
      const long long cst = (long long)nv * I.np;
      if (D.kind == DK_TOT) partner = comp == 0 ? vp[p] : vp[cst + p];
      else if (D.kind == DK_PP) { if (comp != 0) continue; partner = vp[p]; }
      else if (D.kind == DK_QQ) { if (comp != 1) continue; partner = vp[cst + p]; }
      else {
        // PQ(a,c) = P_a Q_c; A.kind: 0 -> this orbital is a (needs comp P,
        // partner Q_c); 1 -> this orbital is c (comp Q, partner P_a)
        if (A.kind == 0) { if (comp != 0) continue; partner = vp[cst + p]; }
        else { if (comp != 1) continue; partner = vp[p]; }
      }
      const KGroupDev G = kgs[D.kg];
      const double* g = Gb + D.off_f;
      double grho;
      if (i < 0) {
        const double wkv = wt[G.off_w + (long long)iv * ks + m];
        const double wmv = wt[G.off_w + (long long)nk + (long long)iv * ks + m];
        grho = g[iv] * wkv + (g[nv + iv] + g[2 * nv + iv * ks + m]) * wmv;
      } else {
        const double wki = wt[G.off_w + 2LL * nk + ((long long)iv * ks + m) * ks + i];
        grho = g[2 * nv + nk + iv * ks + m] * wki;
      }
      acc += (double)A.factor * grho * partner;
    }
    gorb[I.off_orbv + q - seg[s]] = acc;
  }
}

// adjoint phase D: dE/dcoef(j) = sum over intervals and points of B_j * g_f;
// one wavefront per coefficient, lanes over (interval, point)
__global__ void k_adj_coef(int ninst, const long long* __restrict__ seg, const InstDev* __restrict__ inst,
                           const double* __restrict__ gorb, double* __restrict__ grad, long long total) {
  const int lane = threadIdx.x & 63;
  for (long long q = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < total;
       q += ((long long)gridDim.x * blockDim.x) >> 6) {
    const int s = find_seg(seg, ninst, q);
    const InstDev I = inst[s];
    long long r = q - seg[s];
    const int j = (int)(r % I.ns);
    r /= I.ns;
    const int comp = (int)(r % 2);
    const int orb = (int)(r / 2);
    const int kx = comp == 0 ? I.ksp : I.ksq;
    const int ks = I.ks, nv = I.nv, np = I.np, nw = nv + 1;
    const double* xb = comp == 0 ? I.pbsp : I.qbsp;
    const double* bin = comp == 0 ? I.bpin : I.bqin;
    double acc = 0.0;
    const int nsx = nv + kx - 1;
    if (j < nsx) {
      const int iv0 = max(0, j - kx + 1), iv1 = min(j, nv - 1);
      const int nterm = (iv1 - iv0 + 1) * np;
      for (int t = lane; t < nterm; t += 64) {
        const int iv = iv0 + t / np, p = t % np;
        const int jj = j - iv;
        const double gf = gorb[I.off_orbv + ((long long)(2 * orb + comp) * nv + iv) * np + p];
        const double b = p < ks ? xb[iv + (long long)nw * (p + (long long)ks * jj)]
                                : bin[((long long)iv * ks * ks + (p - ks)) * kx + jj];
        acc += b * gf;
      }
    }
    acc = wave_sum(acc);
    if (lane == 0) grad[I.off_pq + q - seg[s]] = acc;
  }
}

// adjoint phase C, LDS form: one workgroup per (instance, interval iv,
// outer point m).  For every (instance,k) group the density gradients at
// the point are scattered into norb x norb matrices in LDS (symmetric for
// P_aP_c / Q_aQ_c / total densities, the diagonal counted twice; plain for
// P_aQ_c), then the orbital gradients are matrix-vector products with the
// orbital values at the outer point m and at its ks nested points.
// mode 0: total density (one matrix, both components); 1: P_aP_c and
// Q_aQ_c densities (two matrices); 2: P_aQ_c densities (one matrix).
__global__ void __launch_bounds__(256) k_adj_orb_lds(
    int ninst, const long long* __restrict__ seg, const InstDev* __restrict__ inst,
    const KGroupDev* __restrict__ kgs, const DensDev* __restrict__ dens,
    const double* __restrict__ orbv, const double* __restrict__ wt, const double* __restrict__ Gb,
    int mode, double* __restrict__ gorb, long long total) {
  extern __shared__ double sm[];
  for (long long item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = find_seg(seg, ninst, item);
    const InstDev I = inst[s];
    const int r = (int)(item - seg[s]);
    const int iv = r / I.ks, m = r % I.ks;
    const int norb = I.norb, ks = I.ks, np = I.np, nv = I.nv, nk = nv * ks;
    const int ld = norb + 1;
    const int nmat = mode == 1 ? 2 : 1;
    double* M1 = sm;
    double* M2 = M1 + norb * ld;
    double* fo = M1 + nmat * norb * ld;  // [comp][orb]
    double* fi = fo + 2 * norb;          // [comp][orb][i]
    double* oo = fi + 2 * norb * ks;     // outputs, outer point
    double* oi = oo + 2 * norb;          // outputs, nested points
    double* wv = oi + 2 * norb * ks;     // Wk, Wm, wkin(ks)
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * norb; q += blockDim.x) {
      const int comp = q / norb, a = q % norb;
      fo[q] = orbv[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + m];
      oo[q] = 0.0;
    }
    for (int q = threadIdx.x; q < 2 * norb * ks; q += blockDim.x) {
      const int i = q % ks, a = (q / ks) % norb, comp = q / (ks * norb);
      fi[q] = orbv[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i];
      oi[q] = 0.0;
    }
    for (int gi = 0; gi < I.nkg; ++gi) {
      const KGroupDev G = kgs[I.kg0 + gi];
      __syncthreads();
      if (threadIdx.x < ks + 2) {
        const int q = threadIdx.x;
        wv[q] = q == 0 ? wt[G.off_w + (long long)iv * ks + m]
              : q == 1 ? wt[G.off_w + (long long)nk + (long long)iv * ks + m]
                       : wt[G.off_w + 2LL * nk + ((long long)iv * ks + m) * ks + (q - 2)];
      }
      for (int pass = 0; pass < 2; ++pass) {
        for (int q = threadIdx.x; q < nmat * norb * ld; q += blockDim.x) M1[q] = 0.0;
        __syncthreads();
        for (int d = threadIdx.x; d < G.ndens; d += blockDim.x) {
          const DensDev D = dens[G.dens0 + d];
          const double* g = Gb + D.off_f;
          double v;
          if (pass == 0) v = g[iv] * wv[0] + (g[nv + iv] + g[2 * nv + iv * ks + m]) * wv[1];
          else v = g[2 * nv + nk + iv * ks + m];
          double* M = (mode == 1 && D.kind == DK_QQ) ? M2 : M1;
          if (mode == 2) {
            M[D.a * ld + D.c] = v;
          } else if (D.a == D.c) {
            M[D.a * ld + D.a] = 2.0 * v;
          } else {
            M[D.a * ld + D.c] = v;
            M[D.c * ld + D.a] = v;
          }
        }
        __syncthreads();
        if (pass == 0) {
          for (int q = threadIdx.x; q < 2 * norb; q += blockDim.x) {
            const int comp = q / norb, a = q % norb;
            double acc = 0.0;
            if (mode == 2) {
              if (comp == 0) for (int c = 0; c < norb; ++c) acc += M1[a * ld + c] * fo[norb + c];
              else for (int c = 0; c < norb; ++c) acc += M1[c * ld + a] * fo[c];
            } else {
              const double* M = (mode == 1 && comp == 1) ? M2 : M1;
              for (int c = 0; c < norb; ++c) acc += M[a * ld + c] * fo[comp * norb + c];
            }
            oo[q] += acc;
          }
        } else {
          for (int q = threadIdx.x; q < 2 * norb * ks; q += blockDim.x) {
            const int i = q % ks, a = (q / ks) % norb, comp = q / (ks * norb);
            double acc = 0.0;
            if (mode == 2) {
              if (comp == 0) for (int c = 0; c < norb; ++c) acc += M1[a * ld + c] * fi[(norb + c) * ks + i];
              else for (int c = 0; c < norb; ++c) acc += M1[c * ld + a] * fi[c * ks + i];
            } else {
              const double* M = (mode == 1 && comp == 1) ? M2 : M1;
              for (int c = 0; c < norb; ++c) acc += M[a * ld + c] * fi[(comp * norb + c) * ks + i];
            }
            oi[q] += wv[2 + i] * acc;
          }
        }
        __syncthreads();
      }
    }
    for (int q = threadIdx.x; q < 2 * norb; q += blockDim.x) {
      const int comp = q / norb, a = q % norb;
      gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + m] = oo[q];
    }
    for (int q = threadIdx.x; q < 2 * norb * ks; q += blockDim.x) {
      const int i = q % ks, a = (q / ks) % norb, comp = q / (ks * norb);
      gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i] = oi[q];
    }
  }
}

// ---------------------------------------------------------------------------
// plain path for integral sets: materialised arrays + convol(case 2) per
// distinct left density + SUM_AmB per row (the reference's Rk_evaluate
// organisation, DBSR_CI3/rk_evaluate.f90:31-56), all on the device.
// ---------------------------------------------------------------------------
// band density of orbitals (a,c) in the reference's storage: sym 's'
// (density type 's', components comp1 = comp2) or 'n' (type 'n', P/Q
// components given separately).  spec = (a, c, comp1 | comp2<<1, slot)
__device__ inline bool band_ok(int sym, int ns, int ks, int i, int jp) {
  // 0-based i, jp; 's': i < ns-jp; 'n': 1-based imin..imax of the reference
  if (sym == 0) return i < ns - jp;
  const int jf = jp + 1;
  const int imin = 1 > 1 + ks - jf ? 1 : 1 + ks - jf, imax = ns < ns + ks - jf ? ns : ns + ks - jf;
  return i + 1 >= imin && i + 1 <= imax;
}

__global__ void k_plain_dens(int n, const int4* __restrict__ spec, int ns, int ks, int sym,
                             const double* __restrict__ pq, long long off_pq, double* __restrict__ out) {
  const int nb = sym == 0 ? ks : 2 * ks - 1;
  long long total = (long long)n * ns * nb;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int d = (int)(q / ((long long)ns * nb));
    const int rr = (int)(q % ((long long)ns * nb));
    const int i = rr % ns, j = rr / ns;
    const int4 S = spec[d];
    const int c1 = S.z & 1, c2 = (S.z >> 1) & 1;
    const double* p1 = pq + off_pq + (long long)ns * (c1 + 2LL * S.x);
    const double* p2 = pq + off_pq + (long long)ns * (c2 + 2LL * S.y);
    double v = 0.0;
    if (sym == 0) {
      if (j == 0) v = p1[i] * p2[i];
      else if (i < ns - j) v = p1[i] * p2[i + j] + p1[i + j] * p2[i];
    } else if (band_ok(1, ns, ks, i, j)) {
      v = p1[i] * p2[i + j + 1 - ks];
    }
    out[(long long)S.w * ns * nb + rr] = v;
  }
}

// conv(j,jp) = sum_ip sum_i d(i,ip) rkb(i,j,ip,jp) for many densities
// (convol case 2); one wave per (density, j, jp) with the density index
// fastest so consecutive waves share the rkb column in cache.
__global__ void k_plain_conv2(int ndl, const int* __restrict__ dslot, int ns, int ks, int sym,
                              const double* __restrict__ rkb, const double* __restrict__ dbuf,
                              double* __restrict__ conv) {
  const int lane = threadIdx.x & 63;
  const int nb = sym == 0 ? ks : 2 * ks - 1;
  const long long nw = (long long)ndl * ns * nb;
  for (long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nw;
       w += ((long long)gridDim.x * blockDim.x) >> 6) {
    const int dl = (int)(w % ndl);
    const int jj = (int)(w / ndl);
    const int j = jj % ns, jp = jj / ns;
    double c = 0.0;
    if (band_ok(sym, ns, ks, j, jp)) {
      const double* d = dbuf + (long long)dslot[dl] * ns * nb;
      for (int ip = 0; ip < nb; ++ip) {
        const double* col = rkb + (long long)ns * (j + (long long)ns * (ip + (long long)nb * jp));
        for (int i = lane; i < ns; i += 64)
          if (band_ok(sym, ns, ks, i, ip)) c += d[i + (long long)ns * ip] * col[i];
      }
    }
    c = wave_sum(c);
    if (lane == 0) conv[(long long)dl * ns * nb + j + (long long)ns * jp] = c;
  }
}

// value(row) (+)= SUM_AmB(conv_L, d_R, sym); one wave per row
__global__ void k_plain_rows(int nrows, const int2* __restrict__ lr, int ns, int ks, int sym,
                             const double* __restrict__ conv, const double* __restrict__ dbuf,
                             const long long* __restrict__ row_out, int accumulate,
                             double* __restrict__ vals) {
  const int lane = threadIdx.x & 63;
  const int nb = sym == 0 ? ks : 2 * ks - 1;
  for (long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < nrows;
       w += ((long long)gridDim.x * blockDim.x) >> 6) {
    const int2 P = lr[w];
    const double* a = conv + (long long)P.x * ns * nb;
    const double* b = dbuf + (long long)P.y * ns * nb;
    double s = 0.0;
    for (int jp = 0; jp < nb; ++jp)
      for (int i = lane; i < ns; i += 64)
        if (band_ok(sym, ns, ks, i, jp)) s += a[i + (long long)ns * jp] * b[i + (long long)ns * jp];
    s = wave_sum(s);
    if (lane == 0) {
      double* o = vals + row_out[w];
      *o = accumulate ? *o + s : s;
    }
  }
}

__global__ void k_fill(double* p, long long n, double v) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    p[q] = v;
}

}  // namespace dbs

using namespace dbs;

// ===========================================================================
// host side
// ===========================================================================
struct dbs_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int kind = 0;
  int onesided = 0;
  int ninst = 0;
  std::vector<dbs_ctx*> ctx;
  std::vector<int> norb, nrows;
  std::vector<long long> off_pq, off_rows, off_orbv;  // host copies
  long long total_pq = 0, total_rows = 0, total_orbv = 0, total_w = 0, total_feat = 0;
  // device buffers
  double *pq = nullptr, *coef = nullptr, *vals = nullptr, *grad = nullptr, *energy = nullptr;
  double *orbv = nullptr, *gorb = nullptr, *wt = nullptr, *F = nullptr, *G = nullptr;
  InstDev* d_inst = nullptr;
  KGroupDev* d_kg = nullptr;
  DensDev* d_dens = nullptr;
  RowDev* d_rows = nullptr;
  long long* d_row_out = nullptr;
  long long* d_seg_orb = nullptr;   // per instance start in flat orbital-point index
  long long* d_seg_coef = nullptr;  // per instance start in flat coefficient index
  long long* d_seg_w = nullptr;     // per group start in flat weight index
  long long* d_roff = nullptr;      // per instance row offsets (ninst+1)
  int2* d_items = nullptr;
  int* d_adjoff = nullptr;
  AdjDev* d_adj = nullptr;
  int* d_oadjoff = nullptr;
  OrbAdjDev* d_oadj = nullptr;
  // dense-contraction tables
  int4* d_tiles = nullptr;
  int* d_pairrow = nullptr;
  long long* d_pairoff = nullptr;
  int* d_nextrow = nullptr;  // next row with the same (L,R) pair, -1 at the end
  // energy reduction chunks
  int2* d_echunks = nullptr;
  int* d_echoff = nullptr;
  double* d_epart = nullptr;
  int nechunks = 0;
  // LDS sizes (doubles) of the staged kernels; 0 = not staged
  int feat_stage = 0;
  size_t feat_shm = 0;
  int adj_mode = 0;
  size_t adj_shm = 0;
  long long adj_items = 0;
  long long* d_seg_adj = nullptr;
  // panel work lists (XCD-ordered) for gather and contraction
  PanelItem* d_gitems = nullptr;
  int ngitems = 0;
  PanelItem* d_citems = nullptr;
  int ncitems = 0;
  int* d_cntile = nullptr;   // panels per sparse row
  longlong2* d_croff = nullptr;  // feature offsets (L,R) per sparse row
  long long* d_doff = nullptr;   // feature offset per density
  AdjPre* d_adjpre = nullptr;
  double* d_cpart = nullptr;
  int cmaxtile = 1;
  int ntiles = 0;
  // sizes
  int nkg = 0, ndens = 0, nitems = 0;
  long long total_coef_pts = 0;  // flat orbital-point count (== total_orbv)
  long long total_grad = 0;
  std::vector<KGroupDev> h_kg;
  std::vector<DensDev> h_dens;
  std::vector<RowDev> h_rows;
  std::vector<long long> h_row_out;
  std::vector<InstDev> h_inst;
  // rows per group (plain path)
  std::vector<std::vector<int>> group_rows;
  // tuning: [0] contraction path (0 sparse, 1 dense MFMA, 2 auto),
  //         [1] features block size, [2] rows per block (waves),
  //         [3] persistent grid cap (workgroups), [4] adjoint gather block
  //         [5] panel budget in KiB (features of one panel, per XCD L2)
  int tune[6] = {2, 128, 4, 2048, 256, 2048};
  std::vector<char> group_dense;  // per group: use the dense path
  double flops_fwd = 0, flops_adj = 0, bytes_min = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_adj = -1;
  // per-phase timing (dbs_engine_profile): events recorded after each launch
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<std::string> prof_names;
  size_t prof_pos = 0;
  // per-phase algorithmic work of one launch: flops, minimum bytes
  std::map<std::string, std::pair<double, double>> phase_work;
  // plain path scratch
  double* plain_d = nullptr;
  double* plain_c = nullptr;
  size_t plain_d_n = 0, plain_c_n = 0;
};

namespace {

template <class T>
int upload(hipStream_t s, T** dst, const std::vector<T>& v) {
  size_t n = v.size() ? v.size() : 1;
  DBS_HIP(hipMalloc(dst, n * sizeof(T)));
  if (!v.empty()) DBS_HIP(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return DBS_OK;
}

unsigned grid_for(long long n, int block, int cap) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

int choose_dense(const dbs_engine* e, const KGroupDev& G, size_t nrow) {
  if (e->tune[0] == 0) return 0;
  if (e->tune[0] == 1) return 1;
  // auto: dense when the requested pairs fill enough of the Gram matrix that
  // the MFMA GEMM (N^2 D flops) beats streaming two feature rows per row
  double n = G.ndens;
  return (double)nrow >= 0.08 * n * n ? 1 : 0;
}

int build_dense_tables(dbs_engine* e) {
  std::vector<int4> tiles;
  std::vector<long long> pairoff(e->nkg + 1, 0);
  long long tot = 0;
  for (int g = 0; g < e->nkg; ++g) {
    pairoff[g] = tot;
    if (e->group_dense[g]) tot += (long long)e->h_kg[g].ndens * e->h_kg[g].ndens;
  }
  pairoff[e->nkg] = tot;
  std::vector<int> pairrow((size_t)std::max<long long>(tot, 1), -1);
  std::vector<int> nextrow(std::max<size_t>(e->h_rows.size(), 1), -1);
  for (int g = 0; g < e->nkg; ++g) {
    if (!e->group_dense[g]) continue;
    const int n = e->h_kg[g].ndens, nt = (n + 15) / 16;
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) tiles.push_back(make_int4(g, a, b, 0));
    // rows requesting the same (L,R) pair form a chain in row order
    for (auto it = e->group_rows[g].rbegin(); it != e->group_rows[g].rend(); ++it) {
      const int r = *it;
      const RowDev& R = e->h_rows[r];
      const int l = R.L - e->h_kg[g].dens0, rr = R.R - e->h_kg[g].dens0;
      const long long slot = pairoff[g] + (long long)l * n + rr;
      nextrow[r] = pairrow[slot];
      pairrow[slot] = r;
    }
  }
  e->ntiles = (int)tiles.size();
  if (e->d_tiles) hipFree(e->d_tiles);
  if (e->d_pairrow) hipFree(e->d_pairrow);
  if (e->d_pairoff) hipFree(e->d_pairoff);
  if (e->d_nextrow) hipFree(e->d_nextrow);
  e->d_tiles = nullptr; e->d_pairrow = nullptr; e->d_pairoff = nullptr; e->d_nextrow = nullptr;
  int rc = upload(e->stream, &e->d_tiles, tiles);
  if (!rc) rc = upload(e->stream, &e->d_pairrow, pairrow);
  if (!rc) rc = upload(e->stream, &e->d_pairoff, pairoff);
  if (!rc) rc = upload(e->stream, &e->d_nextrow, nextrow);
  return rc;
}

}  // namespace

// sparse rows of groups on the sparse path, in group order
static std::vector<RowDev> sparse_rows(const dbs_engine* e, std::vector<long long>* outmap) {
  std::vector<RowDev> v;
  outmap->clear();
  for (int g = 0; g < e->nkg; ++g) {
    if (e->group_dense[g]) continue;
    std::vector<int> rs(e->group_rows[g]);
    std::stable_sort(rs.begin(), rs.end(), [&](int x, int y) {
      const RowDev &A = e->h_rows[x], &B = e->h_rows[y];
      return A.L != B.L ? A.L < B.L : A.R < B.R;
    });
    for (int r : rs) {
      v.push_back(e->h_rows[r]);
      outmap->push_back(e->h_row_out[r]);
    }
  }
  return v;
}

struct SparseCache {
  RowDev* rows = nullptr;
  long long* out = nullptr;
  int n = 0;
};
static std::map<const dbs_engine*, SparseCache> g_sparse;

static int panel_width(const dbs_engine* e, int ndens, int D) {
  const double budget = (double)e->tune[5] * 1024.0;
  long long w = (long long)(budget / (8.0 * std::max(ndens, 1)));
  w = (w / 64) * 64;
  if (w < 64) w = 64;
  if (w > 2048) w = 2048;  // 8 features per thread of a 256-thread item
  if (w > D) w = D;
  return (int)w;
}

// interleave per-XCD lists so that item i lands on XCD i % 8
static std::vector<PanelItem> xcd_interleave(const std::vector<std::vector<PanelItem>>& lists) {
  size_t mx = 0;
  for (auto& l : lists) mx = std::max(mx, l.size());
  PanelItem pad;
  std::memset(&pad, 0, sizeof(pad));
  pad.d = -1;
  std::vector<PanelItem> out(mx * lists.size(), pad);
  for (size_t x = 0; x < lists.size(); ++x)
    for (size_t j = 0; j < lists[x].size(); ++j) out[j * lists.size() + x] = lists[x][j];
  return out;
}

static int build_gather_panels(dbs_engine* e) {
  std::vector<std::vector<PanelItem>> lists(8);
  int panel = 0;
  for (int g = 0; g < e->nkg; ++g) {
    const KGroupDev& G = e->h_kg[g];
    const InstDev& I = e->h_inst[G.inst];
    const int D = feat_dim(I.nv, I.ks);
    const int W = panel_width(e, G.ndens, D);
    for (int x0 = 0; x0 < D; x0 += W, ++panel) {
      for (int d = 0; d < G.ndens; ++d) {
        PanelItem it;
        std::memset(&it, 0, sizeof(it));
        it.d = G.dens0 + d; it.n = 1; it.x0 = x0; it.x1 = std::min(D, x0 + W);
        it.nv = I.nv; it.nk = I.nv * I.ks;
        lists[panel % 8].push_back(it);
      }
    }
  }
  std::vector<PanelItem> items = xcd_interleave(lists);
  if (e->d_gitems) hipFree(e->d_gitems);
  e->d_gitems = nullptr;
  e->ngitems = (int)items.size();
  return upload(e->stream, &e->d_gitems, items);
}

static int refresh_sparse(dbs_engine* e) {
  SparseCache& sc = g_sparse[e];
  if (sc.rows) hipFree(sc.rows);
  if (sc.out) hipFree(sc.out);
  sc.rows = nullptr; sc.out = nullptr;
  std::vector<long long> om;
  std::vector<RowDev> rv = sparse_rows(e, &om);
  sc.n = (int)rv.size();
  int rc = upload(e->stream, &sc.rows, rv);
  if (!rc) rc = upload(e->stream, &sc.out, om);
  if (rc) return rc;
  // contraction panels over the sparse rows (row chunks per group panel)
  const int RCH = 16;
  std::vector<std::vector<PanelItem>> lists(8);
  std::vector<int> ntile(std::max(sc.n, 1), 0);
  int panel = 0, maxtile = 1;
  for (int r0 = 0; r0 < sc.n;) {
    const int g = rv[r0].kg;
    int r1 = r0;
    while (r1 < sc.n && rv[r1].kg == g) ++r1;
    const KGroupDev& G = e->h_kg[g];
    const InstDev& I = e->h_inst[G.inst];
    const int D = feat_dim(I.nv, I.ks);
    const int W = panel_width(e, G.ndens, D);
    int t = 0;
    for (int x0 = 0; x0 < D; x0 += W, ++t, ++panel)
      for (int a = r0; a < r1; a += RCH) {
        PanelItem it;
        std::memset(&it, 0, sizeof(it));
        it.d = a; it.n = std::min(RCH, r1 - a); it.x0 = x0; it.x1 = std::min(D, x0 + W); it.tile = t;
        it.nv = I.nv; it.nk = I.nv * I.ks;
        lists[panel % 8].push_back(it);
      }
    for (int r = r0; r < r1; ++r) ntile[r] = t;
    maxtile = std::max(maxtile, t);
    r0 = r1;
  }
  std::vector<PanelItem> items = xcd_interleave(lists);
  if (e->d_citems) hipFree(e->d_citems);
  if (e->d_cntile) hipFree(e->d_cntile);
  if (e->d_cpart) hipFree(e->d_cpart);
  e->d_citems = nullptr; e->d_cntile = nullptr; e->d_cpart = nullptr;
  e->ncitems = (int)items.size();
  e->cmaxtile = maxtile;
  std::vector<longlong2> croff(std::max(sc.n, 1));
  for (int r = 0; r < sc.n; ++r) {
    croff[r].x = e->h_dens[rv[r].L].off_f;
    croff[r].y = e->h_dens[rv[r].R].off_f;
  }
  if (e->d_croff) hipFree(e->d_croff);
  e->d_croff = nullptr;
  rc = upload(e->stream, &e->d_citems, items);
  if (!rc) rc = upload(e->stream, &e->d_croff, croff);
  if (!rc) rc = upload(e->stream, &e->d_cntile, ntile);
  if (!rc) {
    hipError_t er = hipMalloc(&e->d_cpart, sizeof(double) * (size_t)std::max(sc.n, 1) * maxtile);
    if (er != hipSuccess) rc = fail(DBS_ERR_HIP, "contraction partials: hipMalloc");
  }
  return rc;
}

static void mark(dbs_engine* e, const char* name) {
  if (!e->profiling) return;
  if (e->prof_pos >= e->prof_ev.size()) {
    hipEvent_t ev;
    hipEventCreate(&ev);
    e->prof_ev.push_back(ev);
    e->prof_names.emplace_back();
  }
  hipEventRecord(e->prof_ev[e->prof_pos], e->stream);
  e->prof_names[e->prof_pos] = name;
  e->prof_pos++;
}

static int run_forward(dbs_engine* e) {
  hipStream_t s = e->stream;
  mark(e, "start");
  const int cap = e->tune[3];
  k_orb_points<<<grid_for(e->total_orbv, 256, cap * 4), 256, 0, s>>>(
      e->ninst, e->d_seg_orb, e->d_inst, e->pq, e->orbv, e->total_orbv);
  mark(e, "orb_points");
  k_weights<<<grid_for(e->total_w, 256, cap * 4), 256, 0, s>>>(e->nkg, e->d_seg_w, e->d_kg, e->d_inst,
                                                                e->wt, e->total_w);
  mark(e, "weights");
  k_features<<<grid_for(e->nitems, 1, cap), e->tune[1], e->feat_shm, s>>>(
      e->nitems, e->d_items, e->d_kg, e->d_inst, e->d_dens, e->orbv, e->wt, e->F, e->feat_stage);
  mark(e, "features");
  k_scan<<<grid_for(e->ndens, 4, cap), 256, 0, s>>>(e->ndens, e->d_dens, e->d_kg, e->d_inst, e->F, 0);
  mark(e, "prefix_scan");
  SparseCache& sc = g_sparse[e];
  if (sc.n > 0) {
    const int bw = e->tune[2] * 64;
    k_contract_panel<<<(unsigned)std::max(e->ncitems, 1), bw, 0, s>>>(
        e->ncitems, e->d_citems, e->d_croff, e->F, e->onesided, e->cmaxtile, e->d_cpart);
    k_contract_reduce<<<grid_for(sc.n, 256, cap), 256, 0, s>>>(sc.n, e->d_cntile, e->cmaxtile, e->d_cpart,
                                                               sc.out, e->vals);
    mark(e, "contract_rows");
  }
  if (e->ntiles > 0) {
    k_contract_dense<<<grid_for(e->ntiles, 4, cap * 4), 256, 0, s>>>(
        e->ntiles, e->d_tiles, e->d_kg, e->d_inst, e->d_dens, e->F, e->d_pairrow, e->d_pairoff,
        e->d_nextrow, e->d_row_out, e->onesided, e->vals);
    mark(e, "contract_dense");
  }
  if (e->nechunks > 0)
    k_energy_partial<<<grid_for(e->nechunks, 1, cap), 256, 0, s>>>(e->nechunks, e->d_echunks, e->d_roff,
                                                                   e->coef, e->vals, e->d_epart);
  k_energy_final<<<grid_for(e->ninst, 64, 64), 64, 0, s>>>(e->ninst, e->d_echoff, e->d_epart, e->energy);
  mark(e, "energy");
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

static int run_adjoint(dbs_engine* e) {
  hipStream_t s = e->stream;
  const int cap = e->tune[3];
  k_adj_gather_panel<<<(unsigned)std::max(e->ngitems, 1), 256, 0, s>>>(
      e->ngitems, e->d_gitems, e->d_doff, e->d_adjoff, e->d_adjpre, e->coef, e->F, e->onesided, e->G);
  mark(e, "adj_gather");
  k_scan<<<grid_for(e->ndens, 4, cap), 256, 0, s>>>(e->ndens, e->d_dens, e->d_kg, e->d_inst, e->G, 1);
  mark(e, "adj_suffix_scan");
  if (e->adj_shm > 0) {
    k_adj_orb_lds<<<grid_for(e->adj_items, 1, cap * 4), 256, e->adj_shm, s>>>(
        e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->d_dens, e->orbv, e->wt, e->G, e->adj_mode, e->gorb,
        e->adj_items);
  } else {
    k_adj_orb_points<<<grid_for(e->total_orbv, 256, cap * 4), 256, 0, s>>>(
        e->ninst, e->d_seg_orb, e->d_inst, e->d_kg, e->d_dens, e->d_oadjoff, e->d_oadj, e->orbv, e->wt,
        e->G, e->gorb, e->total_orbv);
  }
  mark(e, "adj_orb_points");
  k_adj_coef<<<grid_for(e->total_grad, 4, cap * 4), 256, 0, s>>>(e->ninst, e->d_seg_coef, e->d_inst,
                                                                  e->gorb, e->grad, e->total_grad);
  mark(e, "adj_coef");
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

static void compute_phase_work(dbs_engine* e) {
  const int ninst = e->ninst;
  const std::vector<int>& norb = e->norb;
  // per phase (one launch): flops and minimum HBM bytes
    double pts = 0, pts_fl = 0, pq_b = 0, feat_fl = 0, feat_b = 0, scan_fl = 0, scan_b = 0;
    double rows_fl = 0, rows_b = 0, rows_dense_fl = 0, gath_fl = 0, orbp_fl = 0, coef_fl = 0;
    for (int i = 0; i < ninst; ++i) {
      const InstDev& I = e->h_inst[i];
      const double p = (double)norb[i] * 2 * I.nv * I.np;
      pts += p;
      pts_fl += p * 2 * I.ks;
      pq_b += (double)norb[i] * 2 * I.ns * 8;
      coef_fl += (double)norb[i] * 2 * I.ns * 2.0 * I.ks * I.np;
    }
    for (int g = 0; g < e->nkg; ++g) {
      const InstDev& I = e->h_inst[e->h_kg[g].inst];
      const double nv = I.nv, ks = I.ks, nd = e->h_kg[g].ndens, nr = (double)e->group_rows[g].size();
      const double D = feat_dim(I.nv, I.ks);
      const double rho = e->kind == 0 ? 3.0 : 1.0;  // flops per density point value
      feat_fl += nd * nv * (I.np * rho + 4 * ks + ks + 2 * ks * ks);
      feat_b += nd * D * 8;
      scan_fl += nd * nv;
      scan_b += nd * nv * 16;
      const double prow = (e->onesided ? 2.0 : 4.0) * nv * (1 + ks);
      if (e->group_dense[g]) rows_dense_fl += 2.0 * nd * nd * D;
      else { rows_fl += nr * prow; rows_b += nd * D * 8 + nr * 8; }
      gath_fl += 2.0 * nr * 2.0 * D;
      orbp_fl += 2.0 * nd * nv * I.np * 3.0;
    }
    e->phase_work["orb_points"] = {pts_fl, pq_b + pts * 8};
    e->phase_work["features"] = {feat_fl, feat_b};
    e->phase_work["prefix_scan"] = {scan_fl, scan_b};
    e->phase_work["contract_rows"] = {rows_fl, rows_b};
    e->phase_work["contract_dense"] = {rows_dense_fl, 0.0};
    e->phase_work["adj_gather"] = {gath_fl, feat_b * 2};
    e->phase_work["adj_suffix_scan"] = {scan_fl, scan_b};
    e->phase_work["adj_orb_points"] = {orbp_fl, pts * 8 * 2};
    e->phase_work["adj_coef"] = {coef_fl, pts * 8 + pq_b};
}

extern "C" {

int dbs_engine_create(dbs_engine** out, int ninst, dbs_ctx* const* ctxs, const int* norb,
                      const int* nrows, const int* rows, int kind) {
  DBS_CHECK(out && ctxs && norb && nrows && rows, DBS_ERR_ARG, "dbs_engine_create: null argument");
  DBS_CHECK(ninst >= 1, DBS_ERR_ARG, "dbs_engine_create: ninst < 1");
  DBS_CHECK(kind >= 0 && kind <= 6, DBS_ERR_ARG, "dbs_engine_create: unknown kind");
  *out = nullptr;
  dbs_engine* e = new dbs_engine();
  e->kind = kind;
  e->onesided = kind >= DBS_PPQQ;
  e->ninst = ninst;
  e->device = ctxs[0]->device;
  auto bail = [&](int rc) {
    dbs_engine_destroy(e);
    return rc;
  };
  for (int i = 0; i < ninst; ++i) {
    if (!ctxs[i] || ctxs[i]->device != e->device)
      return bail(fail(DBS_ERR_ARG, "dbs_engine_create: contexts must exist and share one device"));
    if (norb[i] < 1 || nrows[i] < 0) return bail(fail(DBS_ERR_ARG, "dbs_engine_create: bad sizes"));
  }
  hipSetDevice(e->device);
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(DBS_ERR_HIP, "dbs_engine_create: stream"));
  e->ctx.assign(ctxs, ctxs + ninst);
  e->norb.assign(norb, norb + ninst);
  e->nrows.assign(nrows, nrows + ninst);

  // ---- instances ----
  std::vector<long long> seg_orb(ninst), seg_coef(ninst), roff(ninst + 1, 0);
  long long po = 0, pq = 0, pr = 0;
  for (int i = 0; i < ninst; ++i) {
    const dbs_ctx* c = ctxs[i];
    InstDev I;
    I.ns = c->ns; I.ks = c->ks; I.ksp = c->ksp; I.ksq = c->ksq; I.nv = c->nv;
    I.norb = norb[i]; I.np = c->ks + c->ks * c->ks;
    I.orb_base = 0;  // set below
    I.off_pq = pq;
    I.off_orbv = po;
    I.grw = c->grw; I.gr = c->gr; I.grm = c->grm; I.pbsp = c->pbsp; I.qbsp = c->qbsp;
    I.xin = c->xin; I.win = c->win; I.bpin = c->bpin; I.bqin = c->bqin;
    e->h_inst.push_back(I);
    seg_orb[i] = po;
    seg_coef[i] = pq;
    e->off_pq.push_back(pq);
    e->off_orbv.push_back(po);
    e->off_rows.push_back(pr);
    po += (long long)norb[i] * 2 * c->nv * I.np;
    pq += (long long)norb[i] * 2 * c->ns;
    roff[i] = pr;
    pr += nrows[i];
  }
  roff[ninst] = pr;
  e->total_orbv = po;
  e->total_pq = pq;
  e->total_grad = pq;
  e->total_rows = pr;

  std::vector<long long> orb_base(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    orb_base[i + 1] = orb_base[i] + norb[i];
    e->h_inst[i].orb_base = (int)orb_base[i];
  }

  // ---- densities, groups, rows ----
  const int* rp = rows;
  long long foff = 0, woff = 0;
  std::vector<long long> seg_w;
  std::vector<AdjDev> adj_all;
  std::vector<std::vector<AdjDev>> adj_per_dens;
  std::vector<std::vector<OrbAdjDev>> oadj((size_t)orb_base[ninst]);
  e->h_row_out.resize((size_t)pr);
  for (int i = 0; i < ninst; ++i) {
    const InstDev& I = e->h_inst[i];
    const int D = feat_dim(I.nv, I.ks);
    std::map<int, std::vector<long long>> by_k;  // k -> local row indices
    for (long long r = 0; r < nrows[i]; ++r) {
      const int* q = rp + 5 * r;
      int k = q[0];
      for (int t = 1; t <= 4; ++t)
        if (q[t] < 1 || q[t] > norb[i])
          return bail(fail(DBS_ERR_ARG, "dbs_engine_create: orbital index out of range"));
      if (k < 0 || k > 60) return bail(fail(DBS_ERR_ARG, "dbs_engine_create: k out of range"));
      by_k[k].push_back(r);
    }
    for (auto& kv : by_k) {
      KGroupDev G;
      G.inst = i;
      G.k = kv.first;
      G.dens0 = (int)e->h_dens.size();
      G.off_w = woff;
      std::map<std::tuple<int, int, int>, int> dmap;
      auto dens_of = [&](int a, int c, int dk) {
        if (dk == DK_TOT || dk == DK_PP || dk == DK_QQ) {
          if (a > c) std::swap(a, c);
        }
        auto key = std::make_tuple(a, c, dk);
        auto it = dmap.find(key);
        if (it != dmap.end()) return it->second;
        DensDev dd;
        dd.kg = e->nkg;
        dd.a = a; dd.c = c; dd.kind = dk;
        dd.off_f = foff;
        foff += D;
        int id = (int)e->h_dens.size();
        e->h_dens.push_back(dd);
        dmap[key] = id;
        return id;
      };
      std::vector<int> grows;
      for (long long r : kv.second) {
        const int* q = rp + 5 * r;
        int a = q[1] - 1, b = q[2] - 1, c = q[3] - 1, d = q[4] - 1;
        int kl, kr;
        int la = a, lc = c, rb = b, rd = d;
        switch (kind) {
          case 0: kl = DK_TOT; kr = DK_TOT; break;
          case DBS_PPPP: kl = DK_PP; kr = DK_PP; break;
          case DBS_QQQQ: kl = DK_QQ; kr = DK_QQ; break;
          case DBS_PQPQ: kl = DK_PP; kr = DK_QQ; break;
          case DBS_QPQP: kl = DK_QQ; kr = DK_PP; break;
          case DBS_PPQQ: kl = DK_PQ; kr = DK_PQ; break;
          default: kl = DK_PQ; kr = DK_PQ; rb = d; rd = b; break;  // Q_b P_d = PQ(d,b)
        }
        RowDev R;
        R.L = dens_of(la, lc, kl);
        R.R = dens_of(rb, rd, kr);
        R.kg = e->nkg;
        R.pad = 0;
        int gid = (int)e->h_rows.size();
        e->h_rows.push_back(R);
        e->h_row_out[gid] = roff[i] + r;
        grows.push_back(gid);
      }
      G.ndens = (int)e->h_dens.size() - G.dens0;
      seg_w.push_back(woff);
      woff += 2LL * I.nv * I.ks + (long long)I.nv * I.ks * I.ks;
      e->h_kg.push_back(G);
      e->group_rows.push_back(grows);
      e->nkg++;
    }
    rp += 5LL * nrows[i];
  }
  e->ndens = (int)e->h_dens.size();
  for (int i = 0, g = 0; i < ninst; ++i) {
    e->h_inst[i].kg0 = g;
    int n = 0;
    while (g < e->nkg && e->h_kg[g].inst == i) { ++g; ++n; }
    e->h_inst[i].nkg = n;
  }
  e->total_feat = foff;
  e->total_w = woff;
  // work items (group, iv)
  std::vector<int2> items;
  for (int g = 0; g < e->nkg; ++g)
    for (int iv = 0; iv < e->h_inst[e->h_kg[g].inst].nv; ++iv) items.push_back(make_int2(g, iv));
  e->nitems = (int)items.size();
  // adjacency per density (row order)
  adj_per_dens.resize(e->ndens);
  for (size_t r = 0; r < e->h_rows.size(); ++r) {
    const RowDev& R = e->h_rows[r];
    adj_per_dens[R.L].push_back({R.R, (int)r, 0, 0});
    adj_per_dens[R.R].push_back({R.L, (int)r, 1, 0});
  }
  std::vector<int> adjoff(e->ndens + 1, 0);
  for (int d = 0; d < e->ndens; ++d) {
    adjoff[d + 1] = adjoff[d] + (int)adj_per_dens[d].size();
    for (auto& a : adj_per_dens[d]) adj_all.push_back(a);
  }
  // orbital adjacency (global orbital index = orb_base[inst] + orb)
  for (int d = 0; d < e->ndens; ++d) {
    const DensDev& D = e->h_dens[d];
    const int inst = e->h_kg[D.kg].inst;
    const long long ga = orb_base[inst] + D.a, gc = orb_base[inst] + D.c;
    if (D.kind == DK_PQ) {
      oadj[ga].push_back({d, D.c, 1, 0});
      oadj[gc].push_back({d, D.a, 1, 1});
    } else if (D.a == D.c) {
      oadj[ga].push_back({d, D.c, 2, 0});
    } else {
      oadj[ga].push_back({d, D.c, 1, 0});
      oadj[gc].push_back({d, D.a, 1, 0});
    }
  }
  std::vector<int> oadjoff((size_t)orb_base[ninst] + 1, 0);
  std::vector<OrbAdjDev> oadj_all;
  for (size_t o = 0; o < oadj.size(); ++o) {
    oadjoff[o + 1] = oadjoff[o] + (int)oadj[o].size();
    for (auto& a : oadj[o]) oadj_all.push_back(a);
  }
  // ---- work estimate ----
  {
    double fl = 0, fa = 0, by = 0;
    for (int g = 0; g < e->nkg; ++g) {
      const InstDev& I = e->h_inst[e->h_kg[g].inst];
      double nv = I.nv, ks = I.ks, nd = e->h_kg[g].ndens, nr = (double)e->group_rows[g].size();
      double per_dens = nv * (ks * (e->kind == 0 ? 3 : 1) + 2 * 2 * ks + ks * (2 * ks + (e->kind == 0 ? 2 * ks : 0)) + 1);
      double per_row = (e->onesided ? 2.0 : 4.0) * nv * (1 + ks);
      fl += nd * per_dens + nr * per_row;
      fa += nr * 2.0 * per_row + nd * nv * (2 + 4 * ks + 3 * ks * ks);
    }
    for (int i = 0; i < ninst; ++i) {
      const InstDev& I = e->h_inst[i];
      double pts = (double)norb[i] * 2 * I.nv * I.np;
      fl += pts * 2 * I.ks;            // orbital point values
      fa += pts * 2 * I.ks;            // projection back onto coefficients
      by += (double)norb[i] * 2 * I.ns * 8 * 2 + (double)nrows[i] * (8 + 5 * 4 + 8);
    }
    e->flops_fwd = fl;
    e->flops_adj = fa;
    e->bytes_min = by;
  }

  // ---- device allocations ----
  hipStream_t s = e->stream;
  auto dalloc = [&](double** p, long long n) -> int {
    DBS_HIP(hipMalloc(p, (size_t)std::max<long long>(n, 1) * sizeof(double)));
    return DBS_OK;
  };
  int rc = 0;
  rc |= dalloc(&e->pq, e->total_pq);
  rc |= dalloc(&e->grad, e->total_pq);
  rc |= dalloc(&e->coef, e->total_rows);
  rc |= dalloc(&e->vals, e->total_rows);
  rc |= dalloc(&e->energy, ninst);
  rc |= dalloc(&e->orbv, e->total_orbv);
  rc |= dalloc(&e->gorb, e->total_orbv);
  rc |= dalloc(&e->wt, e->total_w);
  rc |= dalloc(&e->F, e->total_feat);
  rc |= dalloc(&e->G, e->total_feat);
  if (rc) return bail(rc);
  // energy chunks: (instance, chunk) with DBS_ECHUNK rows each
  std::vector<int2> echunks;
  std::vector<int> echoff(ninst + 1, 0);
  for (int i = 0; i < ninst; ++i) {
    const int nch = (nrows[i] + DBS_ECHUNK - 1) / DBS_ECHUNK;
    for (int c = 0; c < nch; ++c) echunks.push_back(make_int2(i, c));
    echoff[i + 1] = (int)echunks.size();
  }
  e->nechunks = (int)echunks.size();
  rc |= upload(s, &e->d_echunks, echunks);
  rc |= upload(s, &e->d_echoff, echoff);
  rc |= dalloc(&e->d_epart, std::max(e->nechunks, 1));
  // LDS plans: staged features when the interval's orbital values fit,
  // the LDS form of the adjoint orbital phase when its matrices fit
  {
    const size_t lds_max = 64 * 1024 / sizeof(double);
    size_t fneed = 0, aneed = 0, wneed = 0;
    for (auto& I : e->h_inst) {
      wneed = std::max(wneed, (size_t)(2 * I.ks + I.ks * I.ks));
      fneed = std::max(fneed, (size_t)(2 * I.ks + I.ks * I.ks) + (size_t)2 * I.norb * I.np);
      const int nmat = (kind >= 1 && kind <= 4) ? 2 : 1;
      aneed = std::max(aneed, (size_t)nmat * I.norb * (I.norb + 1) + 4 * I.norb + 4 * I.norb * I.ks + I.ks + 2);
    }
    e->feat_stage = fneed <= lds_max ? 1 : 0;
    e->feat_shm = sizeof(double) * (e->feat_stage ? fneed : wneed);
    e->adj_mode = kind == 0 ? 0 : (kind <= 4 ? 1 : 2);
    e->adj_shm = aneed <= lds_max ? sizeof(double) * aneed : 0;
    std::vector<long long> seg_adj(ninst);
    long long tot = 0;
    for (int i = 0; i < ninst; ++i) {
      seg_adj[i] = tot;
      tot += (long long)e->h_inst[i].nv * e->h_inst[i].ks;
    }
    e->adj_items = tot;
    rc |= upload(s, &e->d_seg_adj, seg_adj);
  }
  if (rc) return bail(rc);
  std::vector<long long> row_out(e->h_row_out);
  rc |= upload(s, &e->d_inst, e->h_inst);
  rc |= upload(s, &e->d_kg, e->h_kg);
  rc |= upload(s, &e->d_dens, e->h_dens);
  rc |= upload(s, &e->d_rows, e->h_rows);
  rc |= upload(s, &e->d_row_out, row_out);
  rc |= upload(s, &e->d_seg_orb, seg_orb);
  rc |= upload(s, &e->d_seg_coef, seg_coef);
  rc |= upload(s, &e->d_seg_w, seg_w);
  rc |= upload(s, &e->d_roff, roff);
  rc |= upload(s, &e->d_items, items);
  rc |= upload(s, &e->d_adjoff, adjoff);
  rc |= upload(s, &e->d_adj, adj_all);
  {
    std::vector<long long> doff(std::max(e->ndens, 1));
    for (int d = 0; d < e->ndens; ++d) doff[d] = e->h_dens[d].off_f;
    std::vector<AdjPre> pre(std::max(adj_all.size(), (size_t)1));
    for (size_t q = 0; q < adj_all.size(); ++q) {
      pre[q].off = e->h_dens[adj_all[q].partner].off_f;
      pre[q].cidx = e->h_row_out[adj_all[q].row];
      pre[q].role = adj_all[q].role;
      pre[q].pad = 0;
    }
    rc |= upload(s, &e->d_doff, doff);
    rc |= upload(s, &e->d_adjpre, pre);
  }
  rc |= upload(s, &e->d_oadjoff, oadjoff);
  rc |= upload(s, &e->d_oadj, oadj_all);
  if (rc) return bail(rc);
  // zero inputs so a forward before the caller fills them is well defined
  k_fill<<<grid_for(e->total_pq, 256, 4096), 256, 0, s>>>(e->pq, e->total_pq, 0.0);
  k_fill<<<grid_for(e->total_rows, 256, 4096), 256, 0, s>>>(e->coef, e->total_rows, 0.0);
  k_fill<<<grid_for(e->total_rows, 256, 4096), 256, 0, s>>>(e->vals, e->total_rows, 0.0);
  e->group_dense.assign(e->nkg, 0);
  for (int g = 0; g < e->nkg; ++g) e->group_dense[g] = (char)choose_dense(e, e->h_kg[g], e->group_rows[g].size());
  rc = build_dense_tables(e);
  compute_phase_work(e);
  if (!rc) rc = refresh_sparse(e);
  if (!rc) rc = build_gather_panels(e);
  if (rc) return bail(rc);
  if (hipStreamSynchronize(s) != hipSuccess) return bail(fail(DBS_ERR_HIP, "dbs_engine_create: setup"));
  *out = e;
  return DBS_OK;
}

int dbs_engine_destroy(dbs_engine* e) {
  if (!e) return DBS_OK;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  if (e->gexec) hipGraphExecDestroy(e->gexec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (hipEvent_t ev : e->prof_ev) hipEventDestroy(ev);
  void* bufs[] = {e->pq, e->coef, e->vals, e->grad, e->energy, e->orbv, e->gorb, e->wt, e->F, e->G,
                  e->d_inst, e->d_kg, e->d_dens, e->d_rows, e->d_row_out, e->d_seg_orb, e->d_seg_coef,
                  e->d_seg_w, e->d_roff, e->d_items, e->d_adjoff, e->d_adj, e->d_oadjoff, e->d_oadj,
                  e->d_tiles, e->d_pairrow, e->d_pairoff, e->d_nextrow, e->plain_d, e->plain_c,
                  e->d_echunks, e->d_echoff, e->d_epart, e->d_seg_adj, e->d_gitems, e->d_citems,
                  e->d_cntile, e->d_cpart, e->d_croff, e->d_doff, e->d_adjpre};
  for (void* p : bufs)
    if (p) hipFree(p);
  auto it = g_sparse.find(e);
  if (it != g_sparse.end()) {
    if (it->second.rows) hipFree(it->second.rows);
    if (it->second.out) hipFree(it->second.out);
    g_sparse.erase(it);
  }
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return DBS_OK;
}

int dbs_engine_buffers(dbs_engine* e, int inst, double** pq, double** coef, double** values,
                       double** grad, double** energy) {
  DBS_CHECK(e && inst >= 0 && inst < e->ninst, DBS_ERR_ARG, "dbs_engine_buffers: bad instance");
  if (pq) *pq = e->pq + e->off_pq[inst];
  if (grad) *grad = e->grad + e->off_pq[inst];
  if (coef) *coef = e->coef + e->off_rows[inst];
  if (values) *values = e->vals + e->off_rows[inst];
  if (energy) *energy = e->energy + inst;
  return DBS_OK;
}

int dbs_engine_forward(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_forward: null engine");
  return run_forward(e);
}

int dbs_engine_forward_adjoint(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_forward_adjoint: null engine");
  int rc = run_forward(e);
  if (rc) return rc;
  return run_adjoint(e);
}

int dbs_engine_sync(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_sync: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  return DBS_OK;
}

int dbs_engine_graph_capture(dbs_engine* e, int with_adjoint) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_graph_capture: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  if (e->gexec) { hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
  if (e->graph) { hipGraphDestroy(e->graph); e->graph = nullptr; }
  DBS_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = run_forward(e);
  if (!rc && with_adjoint) rc = run_adjoint(e);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return fail(DBS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ee));
  e->graph = g;
  DBS_HIP(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
  e->graph_adj = with_adjoint;
  return DBS_OK;
}

int dbs_engine_graph_launch(dbs_engine* e) {
  DBS_CHECK(e && e->gexec, DBS_ERR_STATE, "dbs_engine_graph_launch: no captured graph");
  DBS_HIP(hipGraphLaunch(e->gexec, e->stream));
  return DBS_OK;
}

int dbs_engine_set_tuning(dbs_engine* e, int n, const int* p) {
  DBS_CHECK(e && p && n >= 1 && n <= 6, DBS_ERR_ARG, "dbs_engine_set_tuning: bad arguments");
  int t[6];
  std::memcpy(t, e->tune, sizeof(t));
  for (int i = 0; i < n; ++i) t[i] = p[i];
  DBS_CHECK(t[0] >= 0 && t[0] <= 2, DBS_ERR_ARG, "tuning[0]: path 0 sparse, 1 dense, 2 auto");
  DBS_CHECK(t[1] >= 64 && t[1] <= 256 && t[1] % 64 == 0, DBS_ERR_ARG, "tuning[1]: block size 64..256");
  DBS_CHECK(t[2] >= 1 && t[2] <= 16, DBS_ERR_ARG, "tuning[2]: waves per block 1..16");
  DBS_CHECK(t[3] >= 64 && t[3] <= 65536, DBS_ERR_ARG, "tuning[3]: grid cap");
  DBS_CHECK(t[4] >= 64 && t[4] <= 256 && t[4] % 64 == 0, DBS_ERR_ARG, "tuning[4]: block size 64..256");
  DBS_CHECK(t[5] >= 64 && t[5] <= 65536, DBS_ERR_ARG, "tuning[5]: panel budget 64..65536 KiB");
  const bool path_changed = t[0] != e->tune[0];
  const bool panels_changed = t[5] != e->tune[5];
  std::memcpy(e->tune, t, sizeof(t));
  if (path_changed) {
    for (int g = 0; g < e->nkg; ++g) e->group_dense[g] = (char)choose_dense(e, e->h_kg[g], e->group_rows[g].size());
    int rc = build_dense_tables(e);
    if (!rc) rc = refresh_sparse(e);
    if (rc) return rc;
    compute_phase_work(e);
    DBS_HIP(hipStreamSynchronize(e->stream));
  }
  if (panels_changed && !path_changed) {
    int rc = refresh_sparse(e);
    if (rc) return rc;
  }
  if (panels_changed) {
    int rc = build_gather_panels(e);
    if (rc) return rc;
    DBS_HIP(hipStreamSynchronize(e->stream));
  }
  // a captured graph holds the old launch shapes
  if (e->gexec) { hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
  if (e->graph) { hipGraphDestroy(e->graph); e->graph = nullptr; }
  return DBS_OK;
}

int dbs_engine_get_tuning(dbs_engine* e, int n, int* p) {
  DBS_CHECK(e && p && n >= 1 && n <= 6, DBS_ERR_ARG, "dbs_engine_get_tuning: bad arguments");
  for (int i = 0; i < n; ++i) p[i] = e->tune[i];
  return DBS_OK;
}

int dbs_engine_work(dbs_engine* e, int with_adjoint, double* flops, double* bytes) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_work: null engine");
  if (flops) *flops = e->flops_fwd + (with_adjoint ? e->flops_adj : 0.0);
  if (bytes) *bytes = e->bytes_min;
  return DBS_OK;
}

// Autotuner: times each candidate setting on the engine's own data (event
// timing, median of reps) and keeps the fastest.  Results are identical for
// every candidate up to summation order within 1e-13.
int dbs_engine_autotune(dbs_engine* e, int with_adjoint, int reps) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_autotune: null engine");
  if (reps < 1) reps = 3;
  const int paths[] = {0, 1};
  const int fblocks[] = {128, 256};
  const int waves[] = {4};
  const int caps[] = {4096};
  const int budgets[] = {1024, 2048, 4096};
  int best[6];
  std::memcpy(best, e->tune, sizeof(best));
  float best_ms = 1e30f;
  hipEvent_t a, b;
  DBS_HIP(hipEventCreate(&a));
  DBS_HIP(hipEventCreate(&b));
  for (int p : paths)
    for (int fb : fblocks)
      for (int w : waves)
        for (int cap : caps)
          for (int bud : budgets) {
          int t[6] = {p, fb, w, cap, 256, bud};
          if (dbs_engine_set_tuning(e, 6, t)) continue;
          // warm
          int rc = with_adjoint ? dbs_engine_forward_adjoint(e) : dbs_engine_forward(e);
          if (rc) continue;
          std::vector<float> ms;
          for (int r = 0; r < reps; ++r) {
            hipEventRecord(a, e->stream);
            rc = with_adjoint ? dbs_engine_forward_adjoint(e) : dbs_engine_forward(e);
            hipEventRecord(b, e->stream);
            hipEventSynchronize(b);
            float x = 0;
            hipEventElapsedTime(&x, a, b);
            ms.push_back(x);
          }
          std::sort(ms.begin(), ms.end());
          float med = ms[ms.size() / 2];
          if (med < best_ms) {
            best_ms = med;
            std::memcpy(best, t, sizeof(best));
          }
        }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return dbs_engine_set_tuning(e, 6, best);
}

// ---------------------------------------------------------------------------
// plain evaluation of the engine's rows: per (instance, k, type) the
// materialised array, one case-2 convolution per distinct left density and
// one band dot product per row; types of rk() summed in its order.
// ---------------------------------------------------------------------------
static int plain_group(dbs_engine* e, int g, int type, int accumulate) {
  const KGroupDev& G = e->h_kg[g];
  dbs_ctx* c = e->ctx[G.inst];
  const int ns = c->ns, ks = c->ks;
  const int sym = type >= DBS_PPQQ ? 1 : 0;
  const int nb = sym ? 2 * ks - 1 : ks;
  // component choice for the left and right densities of the type
  int lc1, lc2, rc1, rc2;
  switch (type) {
    case DBS_PPPP: lc1 = lc2 = 0; rc1 = rc2 = 0; break;
    case DBS_QQQQ: lc1 = lc2 = 1; rc1 = rc2 = 1; break;
    case DBS_PQPQ: lc1 = lc2 = 0; rc1 = rc2 = 1; break;
    case DBS_QPQP: lc1 = lc2 = 1; rc1 = rc2 = 0; break;
    case DBS_PPQQ: lc1 = 0; lc2 = 1; rc1 = 0; rc2 = 1; break;
    default: lc1 = 0; lc2 = 1; rc1 = 1; rc2 = 0; break;
  }
  const std::vector<int>& rows = e->group_rows[g];
  // original (a,b,c,d) of each row are not kept; recover them from the
  // engine's densities: L = (a,c), R = (b,d) up to the symmetric swap,
  // which the 's' band density does not see.
  std::map<std::tuple<int, int, int>, int> slot;
  std::vector<int4> spec;
  std::vector<int> lslots;
  std::map<int, int> lindex;
  std::vector<int2> lr;
  std::vector<long long> outmap;
  auto get_slot = [&](int a, int cc, int cm) {
    auto key = std::make_tuple(a, cc, cm);
    auto it = slot.find(key);
    if (it != slot.end()) return it->second;
    int id = (int)spec.size();
    spec.push_back(make_int4(a, cc, cm, id));
    slot[key] = id;
    return id;
  };
  for (int r : rows) {
    const RowDev& R = e->h_rows[r];
    const DensDev& DL = e->h_dens[R.L];
    const DensDev& DR = e->h_dens[R.R];
    int la = DL.a, lcc = DL.c, rb = DR.a, rd = DR.c;
    if (type == DBS_PQQP) std::swap(rb, rd);  // right density Q_b P_d was stored as PQ(d,b)
    int ls = get_slot(la, lcc, lc1 | (lc2 << 1));
    int rs = get_slot(rb, rd, rc1 | (rc2 << 1));
    auto it = lindex.find(ls);
    int li;
    if (it == lindex.end()) {
      li = (int)lslots.size();
      lslots.push_back(ls);
      lindex[ls] = li;
    } else {
      li = it->second;
    }
    lr.push_back(make_int2(li, rs));
    outmap.push_back(e->h_row_out[r]);
  }
  double* rkb = nullptr;
  int rc = build_rkb(c, type, G.k, &rkb);
  if (rc) return rc;
  size_t need_d = spec.size() * (size_t)ns * nb, need_c = lslots.size() * (size_t)ns * nb;
  if (need_d > e->plain_d_n) {
    if (e->plain_d) hipFree(e->plain_d);
    DBS_HIP(hipMalloc(&e->plain_d, need_d * sizeof(double)));
    e->plain_d_n = need_d;
  }
  if (need_c > e->plain_c_n) {
    if (e->plain_c) hipFree(e->plain_c);
    DBS_HIP(hipMalloc(&e->plain_c, need_c * sizeof(double)));
    e->plain_c_n = need_c;
  }
  int4* d_spec = nullptr;
  int* d_ls = nullptr;
  int2* d_lr = nullptr;
  long long* d_out = nullptr;
  hipStream_t s = c->stream;
  rc = upload(s, &d_spec, spec);
  if (!rc) rc = upload(s, &d_ls, lslots);
  if (!rc) rc = upload(s, &d_lr, lr);
  if (!rc) rc = upload(s, &d_out, outmap);
  if (!rc) {
    const long long nd = (long long)spec.size() * ns * nb;
    k_plain_dens<<<grid_for(nd, 256, 8192), 256, 0, s>>>((int)spec.size(), d_spec, ns, ks, sym, e->pq,
                                                         e->off_pq[G.inst], e->plain_d);
    const long long nwv = (long long)lslots.size() * ns * nb;
    k_plain_conv2<<<grid_for(nwv, 4, 65535), 256, 0, s>>>((int)lslots.size(), d_ls, ns, ks, sym, rkb,
                                                         e->plain_d, e->plain_c);
    k_plain_rows<<<grid_for((long long)lr.size(), 4, 65535), 256, 0, s>>>(
        (int)lr.size(), d_lr, ns, ks, sym, e->plain_c, e->plain_d, d_out, accumulate, e->vals);
    hipError_t ee = hipGetLastError();
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) rc = fail(DBS_ERR_HIP, std::string("plain path: ") + hipGetErrorString(ee));
  }
  hipFree(d_spec);
  hipFree(d_ls);
  hipFree(d_lr);
  hipFree(d_out);
  return rc;
}

int dbs_engine_forward_plain(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_forward_plain: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  for (int g = 0; g < e->nkg; ++g) {
    int rc;
    if (e->kind == 0) {
      const int order[4] = {DBS_PPPP, DBS_QQQQ, DBS_QPQP, DBS_PQPQ};
      for (int t = 0; t < 4; ++t)
        if ((rc = plain_group(e, g, order[t], t > 0))) return rc;
    } else if ((rc = plain_group(e, g, e->kind, 0))) {
      return rc;
    }
  }
  return DBS_OK;
}

// Per-phase device time: events recorded on the engine's stream after each
// launch, averaged over reps.  names is nmax*32 chars.
int dbs_engine_profile(dbs_engine* e, int with_adjoint, int reps, int nmax, char* names, double* ms,
                       double* flops, double* bytes, int* nphase) {
  DBS_CHECK(e && names && ms && nphase && nmax > 0, DBS_ERR_ARG, "dbs_engine_profile: bad arguments");
  if (reps < 1) reps = 1;
  std::vector<std::string> order;
  std::map<std::string, double> acc;
  for (int r = 0; r < reps; ++r) {
    e->profiling = true;
    e->prof_pos = 0;
    int rc = run_forward(e);
    if (!rc && with_adjoint) rc = run_adjoint(e);
    e->profiling = false;
    if (rc) return rc;
    DBS_HIP(hipStreamSynchronize(e->stream));
    for (size_t i = 1; i < e->prof_pos; ++i) {
      float x = 0.f;
      DBS_HIP(hipEventElapsedTime(&x, e->prof_ev[i - 1], e->prof_ev[i]));
      const std::string& n = e->prof_names[i];
      if (!acc.count(n)) order.push_back(n);
      acc[n] += x;
    }
  }
  int n = 0;
  for (const auto& name : order) {
    if (n >= nmax) break;
    std::strncpy(names + 32 * n, name.c_str(), 31);
    names[32 * n + 31] = 0;
    ms[n] = acc[name] / reps;
    auto w = e->phase_work.find(name);
    if (flops) flops[n] = w == e->phase_work.end() ? 0.0 : w->second.first;
    if (bytes) bytes[n] = w == e->phase_work.end() ? 0.0 : w->second.second;
    ++n;
  }
  *nphase = n;
  return DBS_OK;
}

}  // extern "C"

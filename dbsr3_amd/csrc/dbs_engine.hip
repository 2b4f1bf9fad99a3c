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

// one (instance, k) group of densities.  Its features (and their
// gradients) form a (2J) x ndens row-major matrix at off_F, J = nv(1+ks)
// pair slots j with rows x = 2j+h:
//   j < nv (interval iv = j):     h=0 S(iv) (M^k(iv) before the prefix scan), h=1 M^{-(k+1)}(iv)
//   j = nv + iv*ks + m:           h=0 Ro(iv,m) = Wm(iv,m) rho(iv,m),          h=1 Y(iv,m)
// so R^k(L,R) = sum_j u_L(j) v_R(j) + v_L(j) u_R(j) with u = row 2j, v = row 2j+1.
struct KGroupDev {
  int inst, k, ndens, dens0;   // densities dens0 .. dens0+ndens-1
  long long off_w;             // weight tables: Wk(nv*ks), Wm(nv*ks), wkin(nv*ks*ks)
  long long off_F;             // feature / gradient matrix of the group
  long long off_part;          // contraction partials: npr x nrows
  int nrows, row0;             // the group's rows row0 .. row0+nrows-1 (group order)
  int npr, dense;              // partial slots per row; rows contracted on the matrix cores
  int chunk0, nchunk;          // adjacency segments: nchunk blocks of 64 densities x nch
  int gmask, full;             // density kinds present (bit DK_TOT..DK_PQ); 1: every
                               // orbital pair a <= c is a total density of the group
  long long off_pidx;          // orbital pair -> density tables, one NO x NO per kind present
  int nch, pcs;                // partner chunks of the pair panels, densities per chunk
};

struct DensDev {
  int kg, a, c, kind;          // orbitals 0-based
};

struct RowDev {
  int L, R, kg, pad;           // global density indices
};

// orbital -> densities adjacency: density, partner orbital, factor, comp-mask
struct OrbAdjDev {
  int dens, partner, factor, kind;
};

// adjacency of the densities of a group in segments of 64-wide column
// blocks: the entries of density X = 64b+t whose partner lies in partner
// chunk c are the first entries of column t of segment chunk0 + b*nch + c,
// entry i at base + 64 i + t; the entries where X is the row's left density
// come first.  A segment has len rows (its longest column, rounded up to
// DBS_PAIR_PF) plus DBS_PAIR_PF more, the rest of every column filled with
// inert entries (the chunk's first partner, no slot, coefficient 0), so the
// 64 lanes run the same trip count: len with the adjoint, lenl (the longest
// run of left entries, rounded up) without.  An entry holds the partner
// density (group-local) and,
// when X is the row's left density and the row's value is wanted, the
// row's partial slot >= 0: the group's left entries numbered in position
// order, so the lanes of a wavefront that update partials touch
// consecutive doubles.  -1: X is the row's right density; -2: left one,
// value produced elsewhere.
struct AdjChunk {
  long long base;
  int len, lenl;
};
struct AdjEnt {
  int P, slot;
  double c;  // the row's coefficient, gathered at every launch
};

// pair-panel work: slots [j0, j1) of group g, partial column pr, owner
// densities [o0, o1) (the chunked kernel)
struct PairWork {
  int g, j0, j1, pr, o0, o1;
};

// sparse-row reduction: value = sum_p part[pbase + p*stride], p < npr
struct RowRed {
  long long pbase, out;
  int stride, npr;
};

typedef double __attribute__((ext_vector_type(4))) dbl4;

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
    // index within the instance (fits 32 bits: norb*2*nv*np values)
    unsigned r = (unsigned)(q - seg[s]);
    const int p = (int)(r % (unsigned)I.np);
    r /= (unsigned)I.np;
    const int iv = (int)(r % (unsigned)I.nv);
    r /= (unsigned)I.nv;
    const int comp = (int)(r & 1u);
    const int orb = (int)(r >> 1);
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
      // column d of the group's feature matrix
      double* f = F + G.off_F + d;
      const long long ld = G.ndens;
      double mk = 0.0, mm = 0.0;
      for (int m = 0; m < ks; ++m) {
        double ro = dens_point(va, vc, cstride, D.kind, m);
        mk += Wk[m] * ro;
        mm += Wm[m] * ro;
        double y = 0.0;
        for (int i = 0; i < ks; ++i) y += wk[m * ks + i] * dens_point(va, vc, cstride, D.kind, ks + m * ks + i);
        const long long x = 2LL * (nv + iv * ks + m);
        f[x * ld] = Wm[m] * ro;
        f[(x + 1) * ld] = y;
      }
      f[(2LL * iv) * ld] = mk;
      f[(2LL * iv + 1) * ld] = mm;
    }
  }
}

// ---------------------------------------------------------------------------
// phase 2, matrix-core form.  Every feature of an interval iv is an entry
// of a Gram matrix over the orbitals, one per density kind:
//   Y(m)[a][c]     = sum_i wkin(m,i) sum_comp f_a(x_mi) f_c(x_mi)  (ks nested points)
//   M^k[a][c]      = sum_m Wk(m)     sum_comp f_a(x_m) f_c(x_m)    (ks outer points)
//   M^-(k+1)[a][c] = sum_m Wm(m)     sum_comp f_a(x_m) f_c(x_m)
//   Ro(m)[a][c]    = Wm(m)           sum_comp f_a(x_m) f_c(x_m)
// (sum_comp over P and Q for total densities, P or Q alone for P_aP_c /
// Q_aQ_c, P of a with Q of c for P_aQ_c).  Each wave takes its own items
// (group, interval, kind, quantity): it stages the orbital values of the
// quantity's points in its slice of LDS, V[comp][point][orbital] with the
// orbital index XOR-swizzled by the point (conflict-free staging writes and
// MFMA operand reads), builds the 16x16 tiles with v_mfma_f64_16x16x4f64
// and stores every tile entry that is a density of the group straight from
// the accumulators: the pair -> density table maps (a,c) to the column, and
// with densities sorted by pair the 16 lanes of a tile row store 16
// consecutive columns.  No workgroup barrier: waves run independently.
// item = (first group g, interval, kind | ng << 8, quantity q): q < ks Y(q)
// and Ro(q); ks M^k; ks+1 M^-(k+1).  The item covers the ng consecutive
// groups g.. of one instance (all with the kind): the orbital values of the
// points are staged once and each group's tiles use its own weights.
// ---------------------------------------------------------------------------
__device__ inline int vsw(int comp, int p, int o, int npt, int NO) {
  return (comp * npt + p) * NO + (o ^ (p & (NO - 1)));
}

__device__ inline void wave_lds_sync() {
  // LDS accesses of one wave are performed in order; keep the compiler from
  // moving accesses across this point
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(256) k_features_mfma(
    int nitems, const int4* __restrict__ items, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const int* __restrict__ pidx,
    const double* __restrict__ orbv, const double* __restrict__ wt, double* __restrict__ F, int wave_lds) {
  extern __shared__ double sm[];
  // the wave index is uniform across the wave: readfirstlane tells the
  // compiler, so item fields, loop bounds and branches stay scalar
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
  double* V = sm + (long long)wv * wave_lds;  // this wave's slice
  for (int it = blockIdx.x * nw + wv; it < nitems; it += gridDim.x * nw) {
    const int4 T = items[it];
    const int g0 = T.x, iv = T.y, kind = T.z & 255, ng = T.z >> 8, q = T.w;
    const InstDev I = inst[kgs[g0].inst];
    const int ks = I.ks, nv = I.nv, np = I.np, norb = I.norb;
    const int NO = norb <= 16 ? 16 : (norb <= 32 ? 32 : 64);
    const int npt = q < ks ? ks + 1 : ks;  // Y(q): nested points of q, then the outer point q
    double* Wa = V + 2 * npt * NO;         // [ng][npt] weight of each point; W[ks] = Wm(q)
    wave_lds_sync();                        // the previous item's reads are done
    // e = oc*npt + lp split with a float reciprocal: exact for e < 2^12,
    // npt <= 17 ((e + 0.5)/npt is at least 0.5/npt away from an integer)
    const float rnpt = 1.0f / (float)npt;
    for (int e = lane; e < 2 * NO * npt; e += 64) {
      const int oc = (int)(((float)e + 0.5f) * rnpt), lp = e - oc * npt;  // oc = 2*orb + comp
      const int o = oc >> 1, comp = oc & 1;
      const int p = q < ks ? (lp < ks ? ks + q * ks + lp : q) : lp;
      V[vsw(comp, lp, o, npt, NO)] = o < norb ? orbv[I.off_orbv + ((long long)oc * nv + iv) * np + p] : 0.0;
    }
    for (int e = lane; e < ng * npt; e += 64) {
      const int lp = e % npt;
      const long long ow = kgs[g0 + e / npt].off_w;
      double w;
      if (q < ks) w = lp < ks ? wt[ow + 2LL * nv * ks + ((long long)iv * ks + q) * ks + lp]
                              : wt[ow + (long long)nv * ks + (long long)iv * ks + q];
      else w = wt[ow + (q == ks ? 0LL : (long long)nv * ks) + (long long)iv * ks + lp];
      Wa[e] = w;
    }
    wave_lds_sync();
    const int ntl = NO / 16;
    const bool sym = kind != DK_PQ;
    const int npass = q < ks ? 2 : 1;
    for (int tt = 0; tt < ntl * ntl; ++tt) {
      const int ta = tt / ntl, tc = tt % ntl;
      if (sym && ta > tc) continue;  // densities of symmetric kinds have a <= c
      const int ra = ta * 16 + (lane & 15), cb = tc * 16 + (lane & 15);
      // this lane's operand offsets for every k-step, the same for all
      // groups of the item: A = V[ao]*W[wl], B = V[bo]; pass 0 the Gram over
      // the ks points (up to 2ks = 32 entries: 8 steps), pass 1 (Y items)
      // Ro over the outer point (1 step); lanes past K contribute zero
      constexpr int MAXS = (2 * DBS_MAXK + 3) / 4;
      int ao[MAXS + 1], bo[MAXS + 1], wl[MAXS + 1];
      bool okk[MAXS + 1];
      const int K0 = kind == DK_TOT ? 2 * ks : ks, ns0 = (K0 + 3) / 4;
#pragma unroll
      for (int st = 0; st <= MAXS; ++st) {
        const bool p1 = st == MAXS;  // the pass-1 step
        const int K = p1 ? (kind == DK_TOT ? 2 : 1) : K0;
        const int kk = (p1 ? 0 : 4 * st) + (lane >> 4);
        const bool ok = kk < K;
        const int kc = ok ? kk : 0;
        const int per = p1 ? 1 : ks;
        int ca, cbp, lp;
        if (kind == DK_TOT) { ca = cbp = kc >= per ? 1 : 0; lp = kc - ca * per; }
        else if (kind == DK_PP) { ca = cbp = 0; lp = kc; }
        else if (kind == DK_QQ) { ca = cbp = 1; lp = kc; }
        else { ca = 0; cbp = 1; lp = kc; }
        if (p1) lp = q < ks ? ks : 0;  // the outer point (only Y items run pass 1)
        ao[st] = vsw(ca, lp, ra, npt, NO);
        bo[st] = vsw(cbp, lp, cb, npt, NO);
        wl[st] = lp;
        okk[st] = ok;
      }
      for (int gi = 0; gi < ng; ++gi) {
        const KGroupDev G = kgs[g0 + gi];
        const int nd = G.ndens;
        const double* W = Wa + gi * npt;
        const int slot = __popc(G.gmask & ((1 << kind) - 1));
        const int* pt = pidx + G.off_pidx + (long long)slot * NO * NO;
        double* Fg = F + G.off_F;
        long long xrow;
        if (q < ks) xrow = 2LL * (nv + iv * ks + q) + 1;  // Y(iv,q); Ro(iv,q) is the row before
        else xrow = 2LL * iv + (q - ks);                   // M^k (the S slot), M^-(k+1)
        for (int pass = 0; pass < npass; ++pass) {
          dbl4 acc = {0.0, 0.0, 0.0, 0.0};
          if (pass == 0) {
#pragma unroll
            for (int st = 0; st < MAXS; ++st) {
              if (st >= ns0) break;
              const double a = W[wl[st]] * V[ao[st]];
              const double b = V[bo[st]];
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(okk[st] ? a : 0.0, okk[st] ? b : 0.0, acc, 0, 0, 0);
            }
          } else {
            const double a = W[wl[MAXS]] * V[ao[MAXS]];
            const double b = V[bo[MAXS]];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(okk[MAXS] ? a : 0.0, okk[MAXS] ? b : 0.0, acc, 0, 0, 0);
          }
          // lane holds column cb and rows ta*16 + 4j + lane/16
          const long long x = pass == 0 ? xrow : xrow - 1;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int row = ta * 16 + 4 * j + (lane >> 4);
            const int d = pt[row * NO + cb];
            if (d >= 0) Fg[x * nd + d] = acc[j];
          }
        }
      }  // groups of the item
    }    // tiles
  }
}

// exclusive prefix (dir=0) or suffix (dir=1) scan over the intervals of the
// S rows (x = 2 iv) of every density column.  Workgroup = 16 columns x 16
// interval segments; items (group, first column).  Each thread sums its
// segment, the segment sums are combined in LDS in a fixed order, then the
// thread rewrites its segment with the exclusive running sums.
#define DBS_SCAN_COLS 16
#define DBS_SCAN_SEGS 16
__global__ void __launch_bounds__(256) k_scan_cols(int nitems, const int2* __restrict__ items,
                                                   const KGroupDev* __restrict__ kgs,
                                                   const InstDev* __restrict__ inst, double* __restrict__ F,
                                                   int dir) {
  __shared__ double ssum[DBS_SCAN_SEGS][DBS_SCAN_COLS];
  const int col = threadIdx.x % DBS_SCAN_COLS, seg = threadIdx.x / DBS_SCAN_COLS;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const KGroupDev G = kgs[items[it].x];
    const int nv = inst[G.inst].nv, nd = G.ndens;
    const int d = items[it].y + col;
    const int per = (nv + DBS_SCAN_SEGS - 1) / DBS_SCAN_SEGS;
    const int lo = min(nv, seg * per), hi = min(nv, lo + per);
    double* f = F + G.off_F + d;
    const long long st = 2LL * nd;  // S(iv) at row 2 iv
    double sum = 0.0;
    if (d < nd) {
      if (dir == 0) {
#pragma unroll 4
        for (int iv = lo; iv < hi; ++iv) sum += f[iv * st];
      } else {
#pragma unroll 4
        for (int iv = hi - 1; iv >= lo; --iv) sum += f[iv * st];
      }
    }
    __syncthreads();
    ssum[seg][col] = sum;
    __syncthreads();
    if (d < nd) {
      double run = 0.0;
      if (dir == 0) { for (int q = 0; q < seg; ++q) run += ssum[q][col]; }
      else { for (int q = DBS_SCAN_SEGS - 1; q > seg; --q) run += ssum[q][col]; }
      if (dir == 0) {
        for (int iv = lo; iv < hi; ++iv) { const double v = f[iv * st]; f[iv * st] = run; run += v; }
      } else {
        for (int iv = hi - 1; iv >= lo; --iv) { const double v = f[iv * st]; f[iv * st] = run; run += v; }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// dense contraction on the FP64 matrix cores.  For a group whose rows cover
// a large part of all density pairs, the values are entries of
// U V^T + V U^T over the pair slots, one GEMM over the feature rows;
// v_mfma_f64_16x16x4f64 tiles of 16x16 outputs.  One wavefront computes a
// 16x16 tile and scatters the requested pairs.
// ---------------------------------------------------------------------------
// partner row of feature row x of the left density (two-sided: u<->v;
// one-sided S^k: S_L with Mm_R for j < nv, Y_L with Ro_R beyond); -1: none
__device__ inline int partner_row(int x, int nv, int onesided) {
  if (!onesided) return x ^ 1;
  if (x < 2 * nv) return (x & 1) ? -1 : x + 1;
  return (x & 1) ? x - 1 : -1;
}

// one workgroup per 16x16 tile; its DBS_DENSE_SPLIT wavefronts take every
// DBS_DENSE_SPLIT-th 4-row step of the feature rows (split K, so a tile's
// long feature sum runs on many SIMDs), loads issued in batches of 4 steps
// ahead of their MFMAs; the partial tiles are added in wave order through
// LDS.
#define DBS_DENSE_SPLIT 16
__global__ void __launch_bounds__(64 * DBS_DENSE_SPLIT) k_contract_dense(
    int ntiles, const int4* __restrict__ tiles, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const double* __restrict__ F, const int* __restrict__ pairrow, const long long* __restrict__ pairoff,
    const int* __restrict__ nextrow, const long long* __restrict__ row_out, int onesided,
    double* __restrict__ vals) {
  // tile = (group, Ltile, Rtile, 0); pairrow[pairoff[g] + l*ndens + r] = row or -1
  __shared__ dbl4 red[DBS_DENSE_SPLIT][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // scalar wave index
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 T = tiles[t];
    const KGroupDev G = kgs[T.x];
    const InstDev I = inst[G.inst];
    const int nv = I.nv, D = feat_dim(nv, I.ks), nd = G.ndens;
    const double* Fg = F + G.off_F;
    const int l0 = T.y * 16, r0 = T.z * 16;
    // A: 16 x 4 slice (rows l0+lane%16, cols x0 + lane/16); B: 4 x 16 slice
    const int li = l0 + (lane & 15), ri = r0 + (lane & 15), kq = lane >> 4;
    const bool lv = li < nd, rv = ri < nd;
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    constexpr int STEP = 4 * DBS_DENSE_SPLIT;
    for (int xb = 4 * w; xb < D; xb += 4 * STEP) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int x = xb + u * STEP + kq;
        const int xp = x < D ? partner_row(x, nv, onesided) : -1;
        a[u] = (lv && xp >= 0) ? Fg[(long long)x * nd + li] : 0.0;
        b[u] = (rv && xp >= 0) ? Fg[(long long)xp * nd + ri] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0) {
      dbl4 sum = red[0][lane];
      for (int q = 1; q < DBS_DENSE_SPLIT; ++q) sum += red[q][lane];
      // output layout of the 16x16x4 f64 MFMA (one output group per
      // register): lane holds column lane%16 and rows 4*j + lane/16, j = 0..3
      const int col = lane & 15;
      for (int j = 0; j < 4; ++j) {
        const int row = 4 * j + (lane >> 4);
        const int L = l0 + row, R = r0 + col;
        if (L < nd && R < nd) {
          for (int rr = pairrow[pairoff[T.x] + (long long)L * nd + R]; rr >= 0; rr = nextrow[rr])
            vals[row_out[rr]] = sum[j];
        }
      }
    }
    __syncthreads();  // red is rewritten by the next tile
  }
}

// ---------------------------------------------------------------------------
// energy: E_inst = sum_rows coef*value.  Blocks reduce fixed chunks of one
// instance's rows (tree in a fixed shape); the last block then adds each
// instance's chunk partials in order.
// ---------------------------------------------------------------------------
#define DBS_ECHUNK 2048
__global__ void k_energy(int nchunks, const int2* __restrict__ chunks, const long long* __restrict__ roff,
                         const double* __restrict__ coef, const double* __restrict__ vals,
                         double* __restrict__ part, int ninst, const int* __restrict__ choff,
                         double* __restrict__ energy, unsigned* __restrict__ done) {
  __shared__ double sh[256];
  __shared__ bool last;
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
  // the last workgroup to finish adds each instance's chunk partials in
  // order (integer counter; it resets the counter for the next launch)
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int s = threadIdx.x; s < ninst; s += blockDim.x) {
    double acc = 0.0;
    for (int c = choff[s]; c < choff[s + 1]; ++c) acc += part[c];
    energy[s] = acc;
  }
  if (threadIdx.x == 0) *done = 0u;
}

// ---------------------------------------------------------------------------
// pair panels: contraction of the rows and (with the adjoint) the gather
//   G_X = sum over rows touching X of coef * (partner features, u<->v)
// in one pass over the features.  A workgroup takes pair slots [j0, j1) of
// one group W slots at a time: the W-slot panels of the densities are
// staged in LDS, each thread owns a density X, keeps its own panel and
// gradient in registers and streams its adjacency: per entry the partner's
// 2W values come from LDS, the row's partial value (X = left density) and
// the gradient are accumulated in a fixed order.  Partial values of a row
// over the workgroup's panels go to the row's partial column pr; the row
// value is their sum in pr order (k_rows_reduce).
// ---------------------------------------------------------------------------
constexpr int kFeatGroups = 16;  // most groups per item of k_features_mfma (LDS sizing)

// Lane order of the pair panels.  The panel of a density is E = 2W doubles
// [u0 v0 u1 v1 ...] (element e = 2 jl + h) and sits in LDS as a row of E
// doubles, so the bank of an element does not depend on the row.  Lane l
// takes the E/2 element pairs in the order k ^ r, r = l mod min(E/2, 16),
// and moves a pair with one 16-byte access: the 16 lanes that share an LDS
// cycle (64 banks) touch 16 different pairs, which for rows of 32 doubles
// (W = 16) are the 64 banks exactly, wherever their rows are.  Registers
// t = 2k + h hold element 2 (k ^ r) + h; the byte address of that pair of
// a row at byte offset R (a multiple of E*8) is (R | r << 4) ^ (k << 4): one
// XOR per access.
template <int E>
__device__ inline int pair_lane(int lane) {
  return lane & ((E / 2 < 16 ? E / 2 : 16) - 1);
}

// v[2k + h] <- v[2 (k ^ r) + h] for r < E/2: per bit b of r, pairs k and
// k | b swap where the bit is set (an involution: the same call maps back).
// Written as explicit swaps, one select per register and bit.
template <int E>
__device__ inline void permute_pairs(double (&v)[E], int r) {
#pragma unroll
  for (int b = 1; b < E / 2; b <<= 1) {
    const bool sw = (r & b) != 0;
#pragma unroll
    for (int k = 0; k < E / 2; ++k) {
      if (k & b) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double x = v[2 * k + h], y = v[2 * (k | b) + h];
        v[2 * k + h] = sw ? y : x;
        v[2 * (k | b) + h] = sw ? x : y;
      }
    }
  }
}

// LDS accesses by 32-bit byte address: the address of an element pair is
// one XOR away from its row's, with no base to add.  The pair-panel kernels
// use dynamic LDS only, which starts at address 0, so row offsets keep the
// zero low bits the XOR relies on.
typedef double pair_v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) pair_v lds_pair_v;
__device__ inline unsigned lds_addr(const double* p) {
  return (unsigned)(size_t)((const __attribute__((address_space(3))) double*)p);
}
__device__ inline pair_v lds_get(unsigned a) { return *(const lds_pair_v*)(size_t)a; }
__device__ inline void lds_put(unsigned a, double x, double y) {
  pair_v v;
  v.x = x;
  v.y = y;
  *(lds_pair_v*)(size_t)a = v;
}

// S^k pair panels: bit t set where the element of register t of a panel at
// slot j0 pairs as the row's left density (S_L Mm_R below nv, Y_L Ro_R above)
template <int E>
__device__ inline unsigned left_mask(int j0, int nv, int r) {
  unsigned mL = 0;
#pragma unroll
  for (int t = 0; t < E; ++t) {
    const int e = (((t >> 1) ^ r) << 1) | (t & 1);
    const bool lower = j0 + (e >> 1) < nv;
    if (lower ? (e & 1) == 0 : (e & 1) == 1) mL |= 1u << t;
  }
  return mL;
}

// one adjacency entry against the partner row at LDS byte address a (the
// row's address | r << 4): the row's partial dot and, with the adjoint, the
// gradient update; the pairing swaps u and v.  The dot runs in four
// accumulators (pairs k mod 4), summed as (d0 + d1) + (d2 + d3).
template <int E, bool ONE, bool ADJ>
__device__ inline double pair_entry(unsigned a, const double (&fx)[E], double (&gx)[E], double c, bool left,
                                    unsigned mL) {
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  if (!ONE) {
#pragma unroll
    for (int t = 0; t < E; t += 2) {
      const pair_v pp = lds_get(a ^ (unsigned)(t << 3));
      d[(t >> 1) & 3] += fx[t] * pp.y;
      d[(t >> 1) & 3] += fx[t + 1] * pp.x;
      if (ADJ) {
        gx[t] += c * pp.y;
        gx[t + 1] += c * pp.x;
      }
    }
  } else {
    const unsigned mg = left ? mL : ~mL;
#pragma unroll
    for (int t = 0; t < E; t += 2) {
      const pair_v pp = lds_get(a ^ (unsigned)(t << 3));
      if ((mL >> t) & 1u) d[(t >> 1) & 3] += fx[t] * pp.y;
      if ((mL >> (t + 1)) & 1u) d[(t >> 1) & 3] += fx[t + 1] * pp.x;
      if (ADJ && ((mg >> t) & 1u)) gx[t] += c * pp.y;
      if (ADJ && ((mg >> (t + 1)) & 1u)) gx[t + 1] += c * pp.x;
    }
  }
  return (d[0] + d[1]) + (d[2] + d[3]);
}

// adjacency entries in flight per thread of the pair-panel kernels
#ifndef DBS_PAIR_PF
#define DBS_PAIR_PF 4
#endif

// one column of an adjacency segment, len entries (a multiple of
// DBS_PAIR_PF) DBS_PAIR_PF at a time, with the same trip count on every
// lane.  Each entry is loaded one group ahead of its use, into the
// registers it is used from.  A row's partial over the workgroup's panels
// is stored at its first panel and added to with a no-return atomic add at
// the others: nothing waits on a partial, and since only this thread adds
// to the slot, in program order, the sum is formed in panel order.
template <int E, bool ONE, bool ADJ>
__device__ inline void pair_entries(const AdjEnt* __restrict__ ent, long long eb, int len, unsigned rowbase, int P0,
                                    int rsh, const double (&fx)[E], double (&gx)[E], double* __restrict__ pcol,
                                    unsigned mL, bool first) {
  constexpr int PF = DBS_PAIR_PF;
  AdjEnt a[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) a[u] = ent[eb + 64LL * u];
  for (int i = 0; i < len; i += PF) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (i + u >= len) break;
      const unsigned ad = (rowbase + (unsigned)((a[u].P - P0) * E * 8)) | (unsigned)rsh;
      const double dot = pair_entry<E, ONE, ADJ>(ad, fx, gx, ADJ ? a[u].c : 0.0, a[u].slot != -1, mL);
      if (a[u].slot >= 0) {
        if (first) pcol[a[u].slot] = dot;
        else unsafeAtomicAdd(pcol + a[u].slot, dot);
      }
      a[u] = ent[eb + 64LL * (i + u + PF)];
      // one entry's partner reads in flight at a time (registers)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Pair panels for groups whose panels all fit in LDS (E*ndens doubles).  A
// workgroup owns the densities [o0, o1) of the group (thread t owns o0 + t
// + q blockDim, q < DPT; T = 512 threads, 1024 at W = 4; DPT = 2 for the
// narrow panels W = 8, 4 of owner ranges up to 1024 or 2048 densities) and
// stages the panels of all the group's densities: owner ranges split a
// group's rows between workgroups that read the same panels (from L2 when
// they run on one XCD), so a workgroup can take more panels, and a row
// fewer partial values, at the same number of workgroups.  The gradient is
// permuted back to element order before its coalesced store.
// workgroup size limit of the kernel: W = 4 panels leave room for 1024
// threads (more waves to hide the adjacency loads of narrow panels)
template <int W>
constexpr int rot_threads_max() { return W == 4 ? 1024 : 512; }

template <int W, bool ONE, bool ADJ, int DPT>
__global__ void __launch_bounds__(rot_threads_max<W>()) k_pair_rot(
    int nwork, const PairWork* __restrict__ work, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const AdjChunk* __restrict__ chunks,
    const AdjEnt* __restrict__ ent,
    const double* __restrict__ F, double* __restrict__ Gout, double* __restrict__ part) {
  extern __shared__ double sm[];
  constexpr int E = 2 * W;
  const int r = pair_lane<E>(threadIdx.x);
  const unsigned sbase = lds_addr(sm);
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const PairWork Wk = work[w];
    const KGroupDev Gp = kgs[Wk.g];
    const int nv = inst[Gp.inst].nv, nd = Gp.ndens;
    const double* Fg = F + Gp.off_F;
    // the thread's densities X = o0 + threadIdx.x + q blockDim.x < o1, q <
    // DPT (one partner chunk: segment chunk0 + X/64)
    bool act[DPT];
    int n[DPT];
    long long eb[DPT];
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int X = Wk.o0 + threadIdx.x + q * blockDim.x;
      act[q] = X < Wk.o1;
      n[q] = 0;
      eb[q] = 0;
      if (act[q]) {
        const AdjChunk C = chunks[Gp.chunk0 + (X >> 6)];
        eb[q] = C.base + (X & 63);
        // the forward stops after the entries where X is the row's left one
        n[q] = ADJ ? C.len : C.lenl;
      }
    }
    bool first = true;
    for (int j0 = Wk.j0; j0 < Wk.j1; j0 += W) {
      const int nj = min(W, Wk.j1 - j0);
      // the previous panel is no longer read
      __syncthreads();
      // every density of the group is a partner: the workgroup stages all
      // their panels (rows read coalesced in element order, permuted into
      // the staging lane's order, stored as pairs: a row's layout in LDS
      // is element order whichever lane stores it)
      for (int X = threadIdx.x; X < nd; X += blockDim.x) {
        double st[E];
#pragma unroll
        for (int e = 0; e < E; ++e) st[e] = (e >> 1) < nj ? Fg[(2LL * j0 + e) * nd + X] : 0.0;
        permute_pairs<E>(st, r);
        const unsigned a = (sbase + (unsigned)(X * E * 8)) | (unsigned)(r << 4);
#pragma unroll
        for (int t = 0; t < E; t += 2) lds_put(a ^ (unsigned)(t << 3), st[t], st[t + 1]);
      }
      __syncthreads();
      // the owned densities' panels, in this lane's order, from LDS
      double fx[DPT][E], gx[DPT][E];
      const unsigned mL = ONE ? left_mask<E>(j0, nv, r) : 0u;
#pragma unroll
      for (int q = 0; q < DPT; ++q) {
        const int X = Wk.o0 + threadIdx.x + q * blockDim.x;
#pragma unroll
        for (int t = 0; t < E; ++t) gx[q][t] = 0.0;
        if (act[q]) {
          const unsigned a = (sbase + (unsigned)(X * E * 8)) | (unsigned)(r << 4);
#pragma unroll
          for (int t = 0; t < E; t += 2) {
            const pair_v pp = lds_get(a ^ (unsigned)(t << 3));
            fx[q][t] = pp.x;
            fx[q][t + 1] = pp.y;
          }
        }
      }
      double* const pcol = part + Gp.off_part + (long long)Wk.pr * Gp.nrows;
#pragma unroll
      for (int q = 0; q < DPT; ++q) {
        if (!act[q]) continue;
        pair_entries<E, ONE, ADJ>(ent, eb[q], n[q], sbase, 0, r << 4, fx[q], gx[q], pcol, mL, first);
      }
      if (ADJ) {
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
          if (!act[q]) continue;
          // back to element order, then coalesced rows of G
          permute_pairs<E>(gx[q], r);
          double* g = Gout + Gp.off_F + Wk.o0 + threadIdx.x + q * blockDim.x;
#pragma unroll
          for (int e = 0; e < E; ++e)
            if ((e >> 1) < nj) g[(2LL * j0 + e) * nd] = gx[q][e];
        }
      }
      first = false;
    }
  }
}

// Pair panels of groups too large for k_pair_rot.  The group's densities
// form nch partner chunks of pcs; per panel each chunk is staged in LDS in
// turn, in the same row form and lane order, so the panel width does not
// depend on the group size.  A workgroup owns the densities [o0, o1) of the
// group (one per thread, o0 a multiple of 64); per chunk each thread
// streams the entries of its density whose partner lies in that chunk, its
// panel and gradient staying in registers across the chunks.  Staging
// moves 8 element pairs (128 bytes) per step, lane l taking pair k ^ (l mod
// 8): for W = 16 lanes with l mod 16 >= 8 take the other half of the row
// first, for W = 8 they stage partners of the other parity, so the 16 lanes
// of an LDS cycle write all 64 banks.
template <int W, bool ONE, bool ADJ>
__global__ void __launch_bounds__(512) k_pair_chunk(
    int nwork, const PairWork* __restrict__ work, const KGroupDev* __restrict__ kgs,
    const InstDev* __restrict__ inst, const AdjChunk* __restrict__ chunks,
    const AdjEnt* __restrict__ ent,
    const double* __restrict__ F, double* __restrict__ Gout, double* __restrict__ part) {
  static_assert(W == 8 || W == 16, "chunked pair panels: W = 8 or 16");
  extern __shared__ double sm[];
  constexpr int E = 2 * W;
  const int r = pair_lane<E>(threadIdx.x);
  const int r8 = threadIdx.x & 7;
  const int hsw = (threadIdx.x >> 3) & 1;
  const unsigned sbase = lds_addr(sm);
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const PairWork Wk = work[w];
    const KGroupDev Gp = kgs[Wk.g];
    const int nv = inst[Gp.inst].nv, nd = Gp.ndens, nch = Gp.nch, pcs = Gp.pcs;
    const double* Fg = F + Gp.off_F;
    const int X = Wk.o0 + threadIdx.x;
    const bool act = X < Wk.o1;
    double* const pcol = part + Gp.off_part + (long long)Wk.pr * Gp.nrows;
    bool first = true;
    for (int j0 = Wk.j0; j0 < Wk.j1; j0 += W) {
      const int nj = min(W, Wk.j1 - j0);
      const unsigned mL = ONE ? left_mask<E>(j0, nv, r) : 0u;
      double fx[E], gx[E];
#pragma unroll
      for (int t = 0; t < E; ++t) {
        fx[t] = 0.0;
        gx[t] = 0.0;
      }
      if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) fx[e] = (e >> 1) < nj ? Fg[(2LL * j0 + e) * nd + X] : 0.0;
        permute_pairs<E>(fx, r);
      }
      for (int c = 0; c < nch; ++c) {
        const int c0 = c * pcs, cn = min(pcs, nd - c0);
        // the previous chunk is no longer read
        __syncthreads();
        const int cn2 = (cn + 1) & ~1;
        for (int s = threadIdx.x; s < cn2; s += blockDim.x) {
          const int p = W == 8 ? s ^ hsw : s;
          if (p >= cn) continue;
#pragma unroll 1
          for (int hh = 0; hh < E / 16; ++hh) {
            const int h0 = 16 * (W == 16 ? hh ^ hsw : hh);
            // one half at a time: its element rows from one base
            const double* fr = Fg + (size_t)(2 * j0 + h0) * nd + (c0 + p);
            double v[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = ((h0 + e) >> 1) < nj ? fr[(size_t)e * nd] : 0.0;
            permute_pairs<16>(v, r8);
            const unsigned a = (sbase + (unsigned)(p * E * 8) + (unsigned)(h0 * 8)) | (unsigned)(r8 << 4);
#pragma unroll
            for (int t = 0; t < 16; t += 2) lds_put(a ^ (unsigned)(t << 3), v[t], v[t + 1]);
          }
        }
        __syncthreads();
        if (!act) continue;
        const AdjChunk S = chunks[Gp.chunk0 + (X >> 6) * nch + c];
        // the forward stops after the entries where X is the row's left one
        const int n = ADJ ? S.len : S.lenl;
        pair_entries<E, ONE, ADJ>(ent, S.base + (X & 63), n, sbase, c0, r << 4, fx, gx, pcol, mL, first);
      }
      if (ADJ && act) {
        permute_pairs<E>(gx, r);
        double* g = Gout + Gp.off_F + X;
#pragma unroll
        for (int e = 0; e < E; ++e)
          if ((e >> 1) < nj) g[(2LL * j0 + e) * nd] = gx[e];
      }
      first = false;
    }
  }
}

// row values of the pair-panel path: DBS_RED_LANES lanes per row, lane s
// sums slots p = s, s+L, ... in order, the lane sums are combined by a
// fixed butterfly
#define DBS_RED_LANES 8
__global__ void k_rows_reduce(int n, const RowRed* __restrict__ red, const double* __restrict__ part,
                              double* __restrict__ vals) {
  // the DBS_RED_LANES consecutive lanes of a row share r, so they enter and
  // leave the loop together and all take part in the shuffles
  const int sub = threadIdx.x % DBS_RED_LANES;
  const long long step = (long long)gridDim.x * blockDim.x / DBS_RED_LANES;
  for (long long r = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / DBS_RED_LANES; r < n; r += step) {
    const RowRed R = red[r];
    double acc = 0.0;
    for (int p = sub; p < R.npr; p += DBS_RED_LANES) acc += part[R.pbase + (long long)p * R.stride];
#pragma unroll
    for (int off = 1; off < DBS_RED_LANES; off <<= 1) acc += __shfl_xor(acc, off, DBS_RED_LANES);
    if (sub == 0) vals[R.out] = acc;
  }
}

// coefficients of the adjacency entries (0 on padding)
__global__ void k_adj_cgather(long long n, const long long* __restrict__ cidx, const double* __restrict__ coef,
                              AdjEnt* __restrict__ ent) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long c = cidx[q];
    ent[q].c = c >= 0 ? coef[c] : 0.0;
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
                                 double* __restrict__ gorb, long long total, int gb0, int gb1) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int s = find_seg(seg, ninst, q);
    const InstDev I = inst[s];
    // groups [gb0, gb1) only; earlier groups' sums continue from gorb
    if (max(I.kg0, gb0) >= min(I.kg0 + I.nkg, gb1)) continue;
    const bool cont = I.kg0 < gb0;
    long long r = q - seg[s];
    const int p = (int)(r % I.np);
    r /= I.np;
    const int iv = (int)(r % I.nv);
    r /= I.nv;
    const int comp = (int)(r % 2);
    const int orb = (int)(r / 2);
    const int ks = I.ks, nv = I.nv, nk = nv * ks;
    double acc = cont ? gorb[I.off_orbv + q - seg[s]] : 0.0;
    const int gslot = I.orb_base + orb;
    const int e0 = oadjoff[gslot], e1 = oadjoff[gslot + 1];
    const int m = p < ks ? p : (p - ks) / ks;
    const int i = p < ks ? -1 : (p - ks) % ks;
    for (int e = e0; e < e1; ++e) {
      const OrbAdjDev A = oadj[e];
      const DensDev D = dens[A.dens];
      if (D.kg < gb0 || D.kg >= gb1) continue;
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
      const long long nd = G.ndens;
      const double* g = Gb + G.off_F + (A.dens - G.dens0);
      const long long xr = 2LL * (nv + iv * ks + m);
      double grho;
      if (i < 0) {
        const double wkv = wt[G.off_w + (long long)iv * ks + m];
        const double wmv = wt[G.off_w + (long long)nk + (long long)iv * ks + m];
        grho = g[2LL * iv * nd] * wkv + (g[(2LL * iv + 1) * nd] + g[xr * nd]) * wmv;
      } else {
        const double wki = wt[G.off_w + 2LL * nk + ((long long)iv * ks + m) * ks + i];
        grho = g[(xr + 1) * nd] * wki;
      }
      acc += (double)A.factor * grho * partner;
    }
    gorb[I.off_orbv + q - seg[s]] = acc;
  }
}

// Orbital values at the points of interval iv, one workgroup per
// (instance, interval, component): the kx B-spline values at the np points
// and the kx coefficients of every orbital are staged in LDS, then
// f(o,p) = sum_j B_j(p) c_o(iv+j) with j ascending.  Workgroups past the
// first nbo gather the adjacency coefficients (k_adj_cgather's work) in the
// same launch.
__global__ void __launch_bounds__(256) k_orb_values(int ninst, const long long* __restrict__ seg,
                                                   const InstDev* __restrict__ inst, const double* __restrict__ pq,
                                                   double* __restrict__ orbv, long long total, int nbo,
                                                   long long nent, const long long* __restrict__ cidx,
                                                   const double* __restrict__ coef, AdjEnt* __restrict__ ent) {
  extern __shared__ double sm[];
  if ((int)blockIdx.x >= nbo) {
    // workgroups past the first nbo gather the adjacency coefficients
    const long long nb = gridDim.x - nbo;
    for (long long q = (long long)(blockIdx.x - nbo) * blockDim.x + threadIdx.x; q < nent; q += nb * blockDim.x) {
      const long long c = cidx[q];
      ent[q].c = c >= 0 ? coef[c] : 0.0;
    }
    return;
  }
  for (long long item = blockIdx.x; item < total; item += nbo) {
    const int s = find_seg(seg, ninst, item);
    const InstDev I = inst[s];
    const int r = (int)(item - seg[s]);
    const int iv = r >> 1, comp = r & 1;
    const int kx = comp == 0 ? I.ksp : I.ksq;
    const int ks = I.ks, np = I.np, nv = I.nv, ns = I.ns, norb = I.norb, nw = nv + 1;
    double* Bv = sm;              // [np][kx]
    double* Cf = Bv + np * kx;    // [norb][kx]
    const double* xb = comp == 0 ? I.pbsp : I.qbsp;
    const double* bin = comp == 0 ? I.bpin : I.bqin;
    __syncthreads();
    for (int q = threadIdx.x; q < np * kx; q += blockDim.x) {
      const int p = q / kx, j = q % kx;
      Bv[q] = p < ks ? xb[iv + (long long)nw * (p + (long long)ks * j)]
                     : bin[((long long)iv * ks * ks + (p - ks)) * kx + j];
    }
    for (int q = threadIdx.x; q < norb * kx; q += blockDim.x) {
      const int o = q / kx, j = q % kx;
      Cf[q] = pq[I.off_pq + (long long)ns * (comp + 2LL * o) + iv + j];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < norb * np; q += blockDim.x) {
      const int o = q / np, p = q % np;
      double v = 0.0;
      for (int j = 0; j < kx; ++j) v += Bv[p * kx + j] * Cf[o * kx + j];
      orbv[I.off_orbv + ((long long)(2 * o + comp) * nv + iv) * np + p] = v;
    }
  }
}

// Projection of the point gradients onto the B-spline coefficients, first
// half: per (instance, interval, component) workgroup the partial sums
//   part(o, iv, jj) = sum_p B_jj(x_p) g_f(o, iv, p),  jj < kx
// stored in rows of ks entries per (orbital, component, interval)
// (point gradients and B-spline values staged in LDS).
__global__ void __launch_bounds__(256) k_adj_proj(int ninst, const long long* __restrict__ seg,
                                                 const InstDev* __restrict__ inst, const double* __restrict__ gorb,
                                                 const long long* __restrict__ poff, double* __restrict__ part,
                                                 long long total) {
  extern __shared__ double sm[];
  for (long long item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = find_seg(seg, ninst, item);
    const InstDev I = inst[s];
    const int r = (int)(item - seg[s]);
    const int iv = r >> 1, comp = r & 1;
    const int kx = comp == 0 ? I.ksp : I.ksq;
    const int ks = I.ks, np = I.np, nv = I.nv, norb = I.norb, nw = nv + 1;
    double* Bv = sm;              // [np][kx]
    double* Gs = Bv + np * kx;    // [norb][np]
    const double* xb = comp == 0 ? I.pbsp : I.qbsp;
    const double* bin = comp == 0 ? I.bpin : I.bqin;
    __syncthreads();
    for (int q = threadIdx.x; q < np * kx; q += blockDim.x) {
      const int p = q / kx, j = q % kx;
      Bv[q] = p < ks ? xb[iv + (long long)nw * (p + (long long)ks * j)]
                     : bin[((long long)iv * ks * ks + (p - ks)) * kx + j];
    }
    for (int q = threadIdx.x; q < norb * np; q += blockDim.x) {
      const int o = q / np, p = q % np;
      Gs[q] = gorb[I.off_orbv + ((long long)(2 * o + comp) * nv + iv) * np + p];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < norb * kx; q += blockDim.x) {
      const int o = q / kx, jj = q % kx;
      double acc = 0.0;
      for (int p = 0; p < np; ++p) acc += Bv[p * kx + jj] * Gs[o * np + p];
      part[poff[s] + ((long long)(2 * o + comp) * nv + iv) * ks + jj] = acc;  // rows of ks >= kx
    }
  }
}

// second half: dE/dc(o, comp, j) = sum_jj part(o, iv = j - jj, jj), jj ascending
__global__ void k_adj_coef_sum(int ninst, const long long* __restrict__ seg, const InstDev* __restrict__ inst,
                               const long long* __restrict__ poff, const double* __restrict__ part,
                               double* __restrict__ grad, long long total) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int s = find_seg(seg, ninst, q);
    const InstDev I = inst[s];
    unsigned r = (unsigned)(q - seg[s]);
    const int j = (int)(r % (unsigned)I.ns);
    r /= (unsigned)I.ns;
    const int comp = (int)(r & 1u);
    const int o = (int)(r >> 1);
    const int kx = comp == 0 ? I.ksp : I.ksq, nv = I.nv;
    const int nsx = nv + kx - 1;
    double acc = 0.0;
    if (j < nsx) {
      for (int jj = 0; jj < kx; ++jj) {
        const int iv = j - jj;
        if (iv >= 0 && iv < nv) acc += part[poff[s] + ((long long)(2 * o + comp) * nv + iv) * I.ks + jj];
      }
    }
    grad[I.off_pq + q - seg[s]] = acc;
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
    unsigned r = (unsigned)(q - seg[s]);
    const int j = (int)(r % (unsigned)I.ns);
    r /= (unsigned)I.ns;
    const int comp = (int)(r & 1u);
    const int orb = (int)(r >> 1);
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
    int mode, double* __restrict__ gorb, long long total, int gb0, int gb1) {
  extern __shared__ double sm[];
  for (long long item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = find_seg(seg, ninst, item);
    const InstDev I = inst[s];
    // groups [gb0, gb1) only; earlier groups' sums continue from gorb
    const int ga = max(I.kg0, gb0) - I.kg0, gz = min(I.kg0 + I.nkg, gb1) - I.kg0;
    if (ga >= gz) continue;
    const bool cont = ga > 0;
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
      oo[q] = cont ? gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + m] : 0.0;
    }
    for (int q = threadIdx.x; q < 2 * norb * ks; q += blockDim.x) {
      const int i = q % ks, a = (q / ks) % norb, comp = q / (ks * norb);
      fi[q] = orbv[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i];
      oi[q] = cont ? gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i] : 0.0;
    }
    for (int gi = ga; gi < gz; ++gi) {
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
        const long long nd = G.ndens;
        const double* gb = Gb + G.off_F;
        const long long xr = 2LL * (nv + iv * ks + m);
        for (int d = threadIdx.x; d < G.ndens; d += blockDim.x) {
          const DensDev D = dens[G.dens0 + d];
          const double* g = gb + d;
          double v;
          if (pass == 0) v = g[2LL * iv * nd] * wv[0] + (g[(2LL * iv + 1) * nd] + g[xr * nd]) * wv[1];
          else v = g[(xr + 1) * nd];
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

// adjoint phase C, matrix-core form for total densities (engine kind
// rk sum).  One workgroup per (instance, interval iv, outer point m).  For
// every (instance,k) group the density gradients at the point are scattered
// into two symmetric orbital matrices in LDS (diagonal doubled):
//   MO[a][c] = G_Mk(iv) Wk(m) + (G_Mm(iv) + G_Ro(iv,m)) Wm(m)   (outer point)
//   MY[a][c] = G_Y(iv,m)                                       (nested points)
// and the orbital gradients at the eleven points are the products
//   gf(x_m)  += MO f(x_m),   gf(x_mi) += MY (wkin(m,i) f(x_mi))
// on v_mfma_f64_16x16x4f64 tiles, accumulated over the groups in the
// waves' registers (tasks: row tile x {nested P/Q columns 0-15, 16-31,
// outer columns}).  The matrices alternate between two LDS buffers so one
// group's scatter overlaps the previous group's products.
// dpk[d] = a | c << 8 of every density.
#define DBS_ADJ_PF 2  // densities per thread whose gradients are fetched one group ahead
__global__ void __launch_bounds__(256) k_adj_orb_mfma(
    int ninst, const long long* __restrict__ seg, const InstDev* __restrict__ inst,
    const KGroupDev* __restrict__ kgs, const int* __restrict__ dpk, const double* __restrict__ orbv,
    const double* __restrict__ wt, const double* __restrict__ Gb, double* __restrict__ gorb, long long total,
    const int* __restrict__ accf, int gb0, int gb1) {
  extern __shared__ double sm[];
  // wave index uniform across the wave (scalar for the compiler)
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
  for (long long item = blockIdx.x; item < total; item += gridDim.x) {
    const int s = find_seg(seg, ninst, item);
    if (accf[s]) continue;  // instance done by k_adj_orb_acc
    const InstDev I = inst[s];
    // groups [gb0, gb1) only; the sums of earlier groups continue from gorb
    // (the accumulators are reloaded with the values they were stored as)
    const int ga = max(I.kg0, gb0) - I.kg0, gz = min(I.kg0 + I.nkg, gb1) - I.kg0;
    if (ga >= gz) continue;
    const bool cont = ga > 0;
    const int r = (int)(item - seg[s]);
    const int iv = r / I.ks, m = r % I.ks;
    const int norb = I.norb, ks = I.ks, np = I.np, nv = I.nv, nk = nv * ks;
    const int NO = norb <= 16 ? 16 : (norb <= 32 ? 32 : 64);
    const int LM = NO + 1;
    double* FI = sm;              // [NO][32]: columns comp*ks + i
    double* FO = FI + NO * 32;    // [NO][16]: columns comp
    double* WV = FO + NO * 16;    // 2 x 16: wkin(m, i) of the group, per buffer
    double* MB = WV + 32;         // 2 buffers x (MO, MY) of [NO][LM]
    __syncthreads();
    for (int q = threadIdx.x; q < NO * 32; q += blockDim.x) {
      const int c = q / 32, col = q % 32;
      double v = 0.0;
      if (c < norb && col < 2 * ks) {
        const int comp = col / ks, i = col % ks;
        v = orbv[I.off_orbv + ((long long)(2 * c + comp) * nv + iv) * np + ks + m * ks + i];
      }
      FI[q] = v;
    }
    for (int q = threadIdx.x; q < NO * 16; q += blockDim.x) {
      const int c = q / 16, col = q % 16;
      FO[q] = (c < norb && col < 2) ? orbv[I.off_orbv + ((long long)(2 * c + col) * nv + iv) * np + m] : 0.0;
    }
    // both buffers start at zero (entries that no density writes stay zero;
    // rows and columns beyond norb meet zero operands)
    for (int q = threadIdx.x; q < 4 * NO * LM; q += blockDim.x) MB[q] = 0.0;
    __syncthreads();
    const int ntl = NO / 16, ntask = ntl * 2;  // nested-point tiles; the outer point on VALU
    // thread t < 2 norb: orbital t/2, component t%2
    double oacc = cont && (int)threadIdx.x < 2 * norb
                      ? gorb[I.off_orbv + ((long long)threadIdx.x * nv + iv) * np + m] : 0.0;
    dbl4 accs[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      accs[n] = dbl4{0.0, 0.0, 0.0, 0.0};
      const int t = wv + n * nw;
      if (!cont || t >= ntask) continue;
      const int rt = t / 2, ct = t % 2;
      const int col = ct * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int a = rt * 16 + 4 * j + (lane >> 4);
        if (a >= norb || col >= 2 * ks) continue;
        const int comp = col / ks, i = col % ks;
        accs[n][j] = gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i];
      }
    }
    const long long xr = 2LL * (nv + iv * ks + m);
    // gradients of the first DBS_ADJ_PF densities per thread, one group ahead
    double pgo[DBS_ADJ_PF], pgy[DBS_ADJ_PF];
    int ppk[DBS_ADJ_PF];
    auto fetch = [&](int gi) {
      const KGroupDev Gn = kgs[I.kg0 + gi];
      const double wkv = wt[Gn.off_w + (long long)iv * ks + m];
      const double wmv = wt[Gn.off_w + (long long)nk + (long long)iv * ks + m];
      const double* gb = Gb + Gn.off_F;
      const int ndn = Gn.ndens;
#pragma unroll
      for (int u = 0; u < DBS_ADJ_PF; ++u) {
        const int d = threadIdx.x + u * blockDim.x;
        ppk[u] = -1;
        pgo[u] = pgy[u] = 0.0;
        if (d < ndn) {
          ppk[u] = dpk[Gn.dens0 + d];
          pgo[u] = gb[2LL * iv * ndn + d] * wkv + (gb[(2LL * iv + 1) * ndn + d] + gb[xr * ndn + d]) * wmv;
          pgy[u] = gb[(xr + 1) * ndn + d];
        }
      }
    };
    fetch(ga);
    for (int gi = ga; gi < gz; ++gi) {
      const KGroupDev G = kgs[I.kg0 + gi];
      const int nd = G.ndens;
      const int b = gi & 1;
      double* MO = MB + (long long)b * 2 * NO * LM;
      double* MY = MO + NO * LM;
      // this buffer's previous products finished before the last barrier
      if (!G.full && gi >= 2)
        for (int q = threadIdx.x; q < 2 * NO * LM; q += blockDim.x) MO[q] = 0.0;
      if (threadIdx.x < 16) WV[b * 16 + threadIdx.x] =
          threadIdx.x < ks ? wt[G.off_w + 2LL * nk + ((long long)iv * ks + m) * ks + threadIdx.x] : 0.0;
      if (!G.full && gi >= 2) __syncthreads();
      double go[DBS_ADJ_PF], gy[DBS_ADJ_PF];
      int pk[DBS_ADJ_PF];
#pragma unroll
      for (int u = 0; u < DBS_ADJ_PF; ++u) {
        go[u] = pgo[u];
        gy[u] = pgy[u];
        pk[u] = ppk[u];
      }
      if (gi + 1 < gz) fetch(gi + 1);
      auto put = [&](int p, double vo, double vy) {
        const int a = p & 255, c = (p >> 8) & 255;
        if (a == c) {
          MO[a * LM + a] = 2.0 * vo;
          MY[a * LM + a] = 2.0 * vy;
        } else {
          MO[a * LM + c] = vo;
          MO[c * LM + a] = vo;
          MY[a * LM + c] = vy;
          MY[c * LM + a] = vy;
        }
      };
#pragma unroll
      for (int u = 0; u < DBS_ADJ_PF; ++u)
        if (pk[u] >= 0) put(pk[u], go[u], gy[u]);
      if (nd > DBS_ADJ_PF * (int)blockDim.x) {
        const double wkv = wt[G.off_w + (long long)iv * ks + m];
        const double wmv = wt[G.off_w + (long long)nk + (long long)iv * ks + m];
        const double* gb = Gb + G.off_F;
        for (int d = threadIdx.x + DBS_ADJ_PF * blockDim.x; d < nd; d += blockDim.x)
          put(dpk[G.dens0 + d], gb[2LL * iv * nd + d] * wkv + (gb[(2LL * iv + 1) * nd + d] + gb[xr * nd + d]) * wmv,
              gb[(xr + 1) * nd + d]);
      }
      __syncthreads();
      if ((int)threadIdx.x < 2 * norb) {
        const int a = threadIdx.x >> 1, comp = threadIdx.x & 1;
        double x = 0.0;
        for (int c = 0; c < norb; ++c) x += MO[a * LM + c] * FO[c * 16 + comp];
        oacc += x;
      }
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const int t = wv + n * nw;
        if (t >= ntask) continue;
        const int rt = t / 2, ct = t % 2;
        const int row = rt * 16 + (lane & 15), col = ct * 16 + (lane & 15);
        dbl4 acc = accs[n];
        for (int k0 = 0; k0 < NO; k0 += 4) {
          const int c = k0 + (lane >> 4);
          const double av = MY[row * LM + c];
          const double bv = WV[b * 16 + (col % ks)] * FI[c * 32 + col];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
        accs[n] = acc;
      }
    }
    // results: lane holds column (lane & 15) of its tile, rows rt*16 + 4j + lane/16
    if ((int)threadIdx.x < 2 * norb)
      gorb[I.off_orbv + ((long long)threadIdx.x * nv + iv) * np + m] = oacc;  // 2a+comp = t
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const int t = wv + n * nw;
      if (t >= ntask) continue;
      const int rt = t / 2, ct = t % 2;
      const dbl4 acc = accs[n];
      const int col = ct * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int a = rt * 16 + 4 * j + (lane >> 4);
        if (a >= norb || col >= 2 * ks) continue;
        const int comp = col / ks, i = col % ks;
        gorb[I.off_orbv + ((long long)(2 * a + comp) * nv + iv) * np + ks + m * ks + i] = acc[j];
      }
    }
  }
}

// adjoint phase C when every (instance,k) group of the instance holds all
// orbital pairs of the total density (the densities are then listed in the
// same (a,c) order, a <= c, in every group).  One workgroup per (instance,
// iv, m); thread t owns densities t, t+256, ...: it sums their gradients over
// the groups in registers -- the outer-point combination G_Mk Wk + (G_Mm +
// G_Ro) Wm and the nested-point G_Y weighted by each group's wkin(m,i) --
// with no barrier inside the group loop, and stores the ks+1 sums per density
// to LDS (V[q][d], q = 0 the outer point, 1..ks the nested points).  After
// one barrier every (point, orbital, component) output is a row of the
// symmetric orbital matrix (diagonal doubled) applied to the orbital values,
// read straight from V through the pair index d(a,c).  Items are laid out
// XCD-major (workgroup b runs item (b&7)*per + (b>>3)) so the ks items of an
// interval share one L2 for the interval's S/Mm rows.
template <int U, int SG>
__global__ void __launch_bounds__(256) k_adj_orb_acc(
    int ninst, const long long* __restrict__ seg, const InstDev* __restrict__ inst,
    const KGroupDev* __restrict__ kgs, const double* __restrict__ orbv, const double* __restrict__ wt,
    const double* __restrict__ Gb, double* __restrict__ gorb, long long total, long long per,
    const int* __restrict__ accf, int slice) {
  // SG threads per item: 256 (the workgroup, U*256 densities) or 64 (one
  // wavefront per item, four items per workgroup, up to U*64 densities; the
  // item's LDS is the wave's slice and its barriers are wave-level)
  constexpr int KSM = 10, NSG = 256 / SG;
  extern __shared__ double sm[];
  const int sg = SG == 256 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x / SG);
  const int tid = (int)threadIdx.x % SG;
  const auto sync = [] {
    if (SG == 256) __syncthreads();
    else wave_lds_sync();
  };
  const int xcd = blockIdx.x & 7;
  for (long long j = (long long)(blockIdx.x >> 3) * NSG + sg; j < per; j += (long long)(gridDim.x >> 3) * NSG) {
    const long long item = xcd * per + j;
    if (item >= total) break;
    const int s = find_seg(seg, ninst, item);
    if (!accf[s]) continue;  // instance done by k_adj_orb_mfma
    const InstDev I = inst[s];
    const int r = (int)(item - seg[s]);
    const int iv = r / I.ks, m = r % I.ks;
    const int norb = I.norb, ks = I.ks, np = I.np, nv = I.nv, nk = nv * ks;
    const int nd = norb * (norb + 1) / 2, FW = 2 * (ks + 1);
    const int QH = (ks + 2) / 2;             // points per half of the table
    double* V = sm + (size_t)sg * slice;     // [QH][nd]: the points of one half
    double* FV = V + (size_t)QH * nd;        // [norb][FW]: column 2q + comp
    sync();                                  // the previous item's readers are done
    for (int q = tid; q < norb * FW; q += SG) {
      const int c = q / FW, col = q % FW, pt = col >> 1, comp = col & 1;
      const long long base = I.off_orbv + ((long long)(2 * c + comp) * nv + iv) * np;
      FV[q] = orbv[base + (pt == 0 ? m : ks + m * ks + pt - 1)];
    }
    double vo[U], vy[U][KSM];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      vo[u] = 0.0;
#pragma unroll
      for (int i = 0; i < KSM; ++i) vy[u][i] = 0.0;
    }
    const long long xr = 2LL * (nv + iv * ks + m);
    for (int gi = 0; gi < I.nkg; ++gi) {
      const KGroupDev G = kgs[I.kg0 + gi];
      const double* w = wt + G.off_w;
      const double wkv = w[(long long)iv * ks + m], wmv = w[(long long)nk + (long long)iv * ks + m];
      double wi[KSM];
#pragma unroll
      for (int i = 0; i < KSM; ++i) wi[i] = i < ks ? w[2LL * nk + ((long long)iv * ks + m) * ks + i] : 0.0;
      const double* gb = Gb + G.off_F;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int d = tid + u * SG;
        if (d < nd) {
          vo[u] += gb[2LL * iv * nd + d] * wkv + (gb[(2LL * iv + 1) * nd + d] + gb[xr * nd + d]) * wmv;
          const double gy = gb[(xr + 1) * nd + d];
#pragma unroll
          for (int i = 0; i < KSM; ++i) vy[u][i] += gy * wi[i];
        }
      }
    }
    // the points in two halves (q0 = 0: points 0..QH-1, q0 = QH: the rest)
    for (int q0 = 0; q0 <= ks; q0 += QH) {
      const int nq = min(QH, ks + 1 - q0);
      if (q0 > 0) sync();  // the first half's readers are done
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int d = tid + u * SG;
        if (d < nd) {
          if (q0 == 0) V[d] = vo[u];
#pragma unroll
          for (int i = 0; i < KSM; ++i) {
            const int q = 1 + i - q0;  // slot of nested point i in this half
            if (i < ks && q >= 0 && q < nq) V[(size_t)q * nd + d] = vy[u][i];
          }
        }
      }
      sync();
      // outputs (orbital a, component comp, point q0 + q): consecutive
      // threads write the consecutive points of one orbital component
      for (int t = tid; t < nq * 2 * norb; t += SG) {
        const int o = t / nq, q = t % nq, a = o >> 1, comp = o & 1, qq = q0 + q;
        const double* Vq = V + (size_t)q * nd;
        const double* f = FV + 2 * qq + comp;
        double x = 0.0;
        for (int c = 0; c < norb; ++c) {
          const int lo = a < c ? a : c, hi = a < c ? c : a;
          const double v = Vq[lo * norb - lo * (lo - 1) / 2 + (hi - lo)];
          x += (a == c ? 2.0 * v : v) * f[c * FW];
        }
        const long long base = I.off_orbv + ((long long)o * nv + iv) * np;
        gorb[base + (qq == 0 ? m : ks + m * ks + qq - 1)] = x;
      }
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
      double* o = vals + (row_out ? row_out[w] : w);  // no table: row w to vals[w]
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
  long long* d_row_out = nullptr;
  long long* d_seg_orb = nullptr;   // per instance start in flat orbital-point index
  long long* d_seg_coef = nullptr;  // per instance start in flat coefficient index
  long long* d_seg_w = nullptr;     // per group start in flat weight index
  long long* d_roff = nullptr;      // per instance row offsets (ninst+1)
  int2* d_items = nullptr;          // features: (group, interval)
  int2* d_scan_items = nullptr;     // scans: (group, first column)
  int nscan_items = 0;
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
  unsigned* d_edone = nullptr;  // workgroups of k_energy done (zero between launches)
  int nechunks = 0;
  // LDS sizes (doubles) of the staged kernels; 0 = not staged
  int feat_stage = 0;
  size_t feat_shm = 0;
  size_t featm_shm = 0;   // LDS of the matrix-core features kernel (0: not usable)
  int4* d_fitems = nullptr;  // its work items (group, interval, kind, quantity)
  int nfitems = 0;
  int* d_pidx = nullptr;     // orbital pair -> density per group and kind
  int* d_dpk = nullptr;      // a | c << 8 per density
  long long* d_seg_ivc = nullptr;  // per instance start of its (interval, component) items
  long long n_ivc = 0;
  size_t ov_shm = 0, proj_shm = 0; // LDS of k_orb_values / k_adj_proj
  long long* d_ppoff = nullptr;    // per instance offset of the projection partials
  double* d_ppart = nullptr;
  size_t adjm_shm = 0;       // LDS of the matrix-core adjoint kernel (0: not usable)
  int adjacc_u = 0;          // densities per thread of k_adj_orb_acc (0: no instance uses it)
  size_t adjacc_shm = 0;
  int adjacc_sg = 256;       // threads per item of k_adj_orb_acc: 256, or 64 when every group has <= 64 densities
  bool adjacc_all = false;   // every instance uses k_adj_orb_acc
  int* d_accf = nullptr;     // per instance: 1 = adjoint by k_adj_orb_acc
  int featm_wave = 0;        // LDS doubles per wave of the matrix-core kernel
  int max_nd = 0;         // largest group
  int adj_mode = 0;
  size_t adj_shm = 0;
  long long adj_items = 0;
  long long* d_seg_adj = nullptr;
  // pair panels: adjacency (64-wide column blocks), work lists, partials
  AdjChunk* d_chunks = nullptr;
  AdjEnt* d_ent = nullptr;
  long long* d_ecidx = nullptr;
  long long nent = 0;
  // work lists [adjoint][variant]: 0 partner panels read from global
  // memory, 1 LDS panels, 2 LDS panels with the conflict-free rotation
  // pair-panel work lists [forward / with adjoint][rotated / chunked kernel]
  PairWork* d_work[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int nwork[2][2] = {{0, 0}, {0, 0}};
  int pair_w = 16;        // panel width (pair slots) of the rotated kernel
  size_t rot_shm = 0;     // LDS bytes of the rotated kernel
  int rot_block = 64;     // its workgroup size (rot_block * rot_dpt >= ndens of its groups)
  int rot_dpt = 1;        // densities per thread of the rotated kernel
  int chunk_w = 16;       // panel width of the chunked kernel
  size_t chunk_shm = 0;   // its LDS bytes (one partner chunk)
  int chunk_block = 64;   // its workgroup size (owner densities per workgroup)
  size_t lds_block = 0;   // device limit of LDS per workgroup (bytes)
  RowRed* d_red = nullptr;
  int nred = 0;
  double* d_part = nullptr;
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
  // adjacency of each density (host): partner (global), row (global)
  std::vector<std::vector<int2>> h_adj;
  // rows per group (group order)
  std::vector<std::vector<int>> group_rows;
  // tuning: [0] contraction path (0 pair panels, 1 dense MFMA, 2 auto),
  //         [1] features block size, [2] pair-panel width W (4, 8, 16),
  //         [3] persistent grid cap (workgroups), [4] owner densities per
  //         workgroup of the chunked pair-panel kernel (its block size),
  //         [5] W-panels per pair-panel workgroup, [6] owner ranges per
  //         group of the rotated pair-panel kernel
  int tune[7] = {2, 256, 16, 4096, 512, 1, 1};
  // memory budget of the features and their gradients (bytes of F and G
  // together); 0: every group's are held at once.  Otherwise the groups run
  // in batches of consecutive groups that fit, F and G reused per batch
  // (dbs_engine_set_memory_budget).
  long long mem_budget = 0;
  std::vector<int> bat{0, 0};  // batch b: groups [bat[b], bat[b+1])
  // per batch: first item of the feature, scan, matrix-core feature,
  // dense-tile, row-reduction and pair-work lists (one entry past the end)
  std::vector<int> bat_items, bat_scan, bat_fit, bat_tiles, bat_red;
  std::vector<int> bat_work[2][2];
  std::vector<int> h_accf;         // adjoint by k_adj_orb_acc per instance, as chosen at creation
  int adjacc_u0 = 0;
  bool adjacc_all0 = false;
  int feat_fg = 16;                // groups per matrix-core feature item
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

template <class T>
void dfree(T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
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

// Batches of consecutive groups whose features and gradients fit the
// memory budget (at least one group each), the groups' feature offsets
// within their batch, and the first feature / scan item of every batch
// (both lists are in group order).  Returns the largest batch's feature
// count.
long long layout_batches(dbs_engine* e) {
  e->bat.assign(1, 0);
  long long cur = 0, big = 0;
  for (int g = 0; g < e->nkg; ++g) {
    KGroupDev& G = e->h_kg[g];
    const InstDev& I = e->h_inst[G.inst];
    const long long nf = (long long)feat_dim(I.nv, I.ks) * G.ndens;
    if (e->mem_budget > 0 && cur > 0 && 2 * (cur + nf) * (long long)sizeof(double) > e->mem_budget) {
      e->bat.push_back(g);
      big = std::max(big, cur);
      cur = 0;
    }
    G.off_F = cur;
    cur += nf;
  }
  e->bat.push_back(e->nkg);
  big = std::max(big, cur);
  const int nb = (int)e->bat.size() - 1;
  e->bat_items.assign(nb + 1, 0);
  e->bat_scan.assign(nb + 1, 0);
  int ni = 0, nsc = 0;
  for (int b = 0; b < nb; ++b) {
    e->bat_items[b] = ni;
    e->bat_scan[b] = nsc;
    for (int g = e->bat[b]; g < e->bat[b + 1]; ++g) {
      ni += e->h_inst[e->h_kg[g].inst].nv;
      nsc += (e->h_kg[g].ndens + DBS_SCAN_COLS - 1) / DBS_SCAN_COLS;
    }
  }
  e->bat_items[nb] = ni;
  e->bat_scan[nb] = nsc;
  return big;
}

// the batch of group g
int batch_of(const dbs_engine* e, int g) {
  return (int)(std::upper_bound(e->bat.begin(), e->bat.end(), g) - e->bat.begin()) - 1;
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
  const int nbat = (int)e->bat.size() - 1;
  e->bat_tiles.assign(nbat + 1, 0);
  for (int g = 0; g < e->nkg; ++g) {
    if (g == e->bat[batch_of(e, g)]) e->bat_tiles[batch_of(e, g)] = (int)tiles.size();
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
  e->bat_tiles[nbat] = e->ntiles;
  dfree(&e->d_tiles);
  dfree(&e->d_pairrow);
  dfree(&e->d_pairoff);
  dfree(&e->d_nextrow);
  int rc = upload(e->stream, &e->d_tiles, tiles);
  if (!rc) rc = upload(e->stream, &e->d_pairrow, pairrow);
  if (!rc) rc = upload(e->stream, &e->d_pairoff, pairoff);
  if (!rc) rc = upload(e->stream, &e->d_nextrow, nextrow);
  return rc;
}

}  // namespace

// dynamic LDS above 64 KB for every variant of a kernel; the first failure
template <int W, int DPT>
static hipError_t rot_lds_attr(size_t bytes) {
  const void* fs[] = {(const void*)k_pair_rot<W, false, true, DPT>, (const void*)k_pair_rot<W, false, false, DPT>,
                      (const void*)k_pair_rot<W, true, true, DPT>, (const void*)k_pair_rot<W, true, false, DPT>};
  for (const void* f : fs) {
    const hipError_t err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

static size_t rot_lds_bytes(int W, int ndens) {
  return (size_t)2 * W * (size_t)ndens * sizeof(double);
}

template <int W>
static hipError_t chunk_lds_attr(size_t bytes) {
  const void* fs[] = {(const void*)k_pair_chunk<W, false, true>, (const void*)k_pair_chunk<W, false, false>,
                      (const void*)k_pair_chunk<W, true, true>, (const void*)k_pair_chunk<W, true, false>};
  for (const void* f : fs) {
    const hipError_t err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

// LDS of one partner chunk of the chunked kernel (KiB, capped by the
// device's limit per workgroup): the whole of a CU's 160 KiB.  The kernel's
// registers allow one 512-thread workgroup per CU either way, and fewer,
// larger chunks balance the entry counts of the 64 lanes better.
#ifndef DBS_CHUNK_LDS_KB
#define DBS_CHUNK_LDS_KB 160
#endif

// Pair-panel tables for the current tuning: the kernel of each group (the
// rotated one when the group's panel fits in LDS, the chunked one
// otherwise), adjacency segments (the slot code depends on which groups
// are dense and on the kernel), work lists, partial layout and the
// sparse-row reduction list.
static int build_pair_tables(dbs_engine* e) {
  const int W = e->tune[2];
  const int NP = e->tune[5];
  if (e->lds_block == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) != hipSuccess || v <= 0)
      v = 64 * 1024;
    e->lds_block = (size_t)v;
  }
  const size_t kPairLdsMax = e->lds_block;
  // rotated kernel: up to T threads (512, 1024 at W = 4) with DPT densities
  // each (DPT = 1 at W = 16, up to 2 at W = 8 and 4)
  const int tmax = W == 4 ? 1024 : 512;
  const int rot_max = W >= 16 ? tmax : 2 * tmax;
  std::vector<char> rot(std::max(e->nkg, 1), 0);
  int rot_nd = 0;
  for (int g = 0; g < e->nkg; ++g) {
    const int nd = e->h_kg[g].ndens;
    if (nd <= rot_max && rot_lds_bytes(W, nd) <= kPairLdsMax) {
      rot[g] = 1;
      rot_nd = std::max(rot_nd, nd);
    }
  }
  e->pair_w = W;
  // owner ranges of the rotated groups: tuning[6] ranges of whole
  // 64-density blocks
  const int osplit = e->tune[6];
  std::vector<int> rot_os(std::max(e->nkg, 1), 0);
  int rot_own = 0;
  for (int g = 0; g < e->nkg; ++g) {
    if (!rot[g]) continue;
    const int nd = std::max(e->h_kg[g].ndens, 1);
    rot_os[g] = std::min(nd, ((nd + osplit - 1) / osplit + 63) / 64 * 64);
    rot_own = std::max(rot_own, rot_os[g]);
  }
  e->rot_dpt = rot_own <= tmax ? 1 : 2;
  e->rot_block = std::max(64, ((rot_own + e->rot_dpt - 1) / e->rot_dpt + 63) / 64 * 64);
  e->rot_shm = rot_lds_bytes(W, std::max(rot_nd, 1));
  if (e->rot_shm > 64 * 1024) {
    hipError_t err;
    if (W == 16) err = rot_lds_attr<16, 1>(e->rot_shm);
    else if (W == 8) err = e->rot_dpt == 1 ? rot_lds_attr<8, 1>(e->rot_shm) : rot_lds_attr<8, 2>(e->rot_shm);
    else err = e->rot_dpt == 1 ? rot_lds_attr<4, 1>(e->rot_shm) : rot_lds_attr<4, 2>(e->rot_shm);
    if (err != hipSuccess) return fail(DBS_ERR_HIP, std::string("pair tables: LDS size: ") + hipGetErrorString(err));
  }
  // chunked kernel: panel width at least 8 (128-byte LDS rows), partner
  // chunks within the LDS budget, owner ranges of at most tuning[4]
  // densities in whole 64-density blocks
  const int CW = W < 8 ? 8 : W;
  const size_t rowb = 2 * (size_t)CW * sizeof(double);
  const size_t cbudget = std::min(kPairLdsMax, (size_t)DBS_CHUNK_LDS_KB * 1024);
  const int omax = e->tune[4];
  e->chunk_w = CW;
  e->chunk_shm = 0;
  e->chunk_block = 64;
  std::vector<int> osz(std::max(e->nkg, 1), 0);
  for (int g = 0; g < e->nkg; ++g) {
    KGroupDev& G = e->h_kg[g];
    G.nch = 1;
    G.pcs = std::max(G.ndens, 1);
    if (rot[g]) continue;
    G.nch = (int)std::max<size_t>(1, ((size_t)G.ndens * rowb + cbudget - 1) / cbudget);
    G.pcs = (G.ndens + G.nch - 1) / G.nch;
    const int nor = (G.ndens + omax - 1) / omax;
    osz[g] = ((G.ndens + nor - 1) / nor + 63) / 64 * 64;
    e->chunk_shm = std::max(e->chunk_shm, (size_t)G.pcs * rowb);
    e->chunk_block = std::max(e->chunk_block, osz[g]);
  }
  if (e->chunk_shm > 64 * 1024) {
    const hipError_t err = CW == 8 ? chunk_lds_attr<8>(e->chunk_shm) : chunk_lds_attr<16>(e->chunk_shm);
    if (err != hipSuccess) return fail(DBS_ERR_HIP, std::string("pair tables: LDS size: ") + hipGetErrorString(err));
  }
  // adjacency segments: per block of 64 densities and partner chunk
  const int PF = DBS_PAIR_PF;
  std::vector<AdjChunk> chunks;
  std::vector<AdjEnt> ent;
  std::vector<long long> ecidx;
  // per group: the row of each partial slot (left entries in position order)
  std::vector<std::vector<int>> slot_row(std::max(e->nkg, 1));
  long long base = 0;
  for (int g = 0; g < e->nkg; ++g) {
    KGroupDev& G = e->h_kg[g];
    const int nd = G.ndens, nch = G.nch, pcs = G.pcs;
    G.chunk0 = (int)chunks.size();
    G.nchunk = (nd + 63) / 64;
    const long long ebase = base;
    std::vector<std::vector<int2>> lst((size_t)64 * nch);
    for (int b = 0; b < G.nchunk; ++b) {
      // per density and partner chunk: the entries where the density is the
      // row's left one (they also form the row value), then the others, each
      // part in row order
      for (auto& v : lst) v.clear();
      for (int t = 0; t < 64 && 64 * b + t < nd; ++t) {
        const int X = G.dens0 + 64 * b + t;
        for (int pass = 0; pass < 2; ++pass)
          for (const int2& a : e->h_adj[X]) {
            const bool left = a.x >= 0;
            if (left != (pass == 0)) continue;
            const RowDev& R = e->h_rows[a.y];
            const int partner = (left ? R.R : R.L) - G.dens0;
            lst[(size_t)t * nch + partner / pcs].push_back(a);
          }
      }
      for (int c = 0; c < nch; ++c) {
        int len = 0, lenl = 0;
        for (int t = 0; t < 64; ++t) {
          const std::vector<int2>& L = lst[(size_t)t * nch + c];
          int nl = 0;
          for (const int2& a : L) nl += a.x >= 0 ? 1 : 0;
          len = std::max(len, (int)L.size());
          lenl = std::max(lenl, nl);
        }
        len = (len + PF - 1) / PF * PF;
        lenl = (lenl + PF - 1) / PF * PF;
        chunks.push_back(AdjChunk{base, len, lenl});
        // inert entries: the chunk's first partner, no slot, coefficient 0
        const long long rows_alloc = 64LL * (len + PF);
        ent.resize((size_t)(base + rows_alloc), AdjEnt{c * pcs, -1, 0.0});
        ecidx.resize((size_t)(base + rows_alloc), -1);
        for (int t = 0; t < 64 && 64 * b + t < nd; ++t) {
          const std::vector<int2>& L = lst[(size_t)t * nch + c];
          for (size_t i = 0; i < L.size(); ++i) {
            const int r = L[i].y;
            const RowDev& R = e->h_rows[r];
            const bool left = L[i].x >= 0;
            AdjEnt A;
            A.P = (left ? R.R : R.L) - G.dens0;
            // slot: the row itself for now, numbered below
            A.slot = left ? (e->group_dense[g] ? -2 : r) : -1;
            A.c = 0.0;
            const long long pos = base + 64LL * (long long)i + t;
            ent[(size_t)pos] = A;
            ecidx[(size_t)pos] = e->h_row_out[r];
          }
        }
        base += rows_alloc;
      }
    }
    // partial slots of the left entries in position order
    for (long long pos = ebase; pos < base; ++pos) {
      AdjEnt& A = ent[(size_t)pos];
      if (A.slot < 0) continue;
      slot_row[g].push_back(A.slot);
      A.slot = (int)slot_row[g].size() - 1;
    }
  }
  e->nent = base;
  // work lists and partial layout, batch by batch: a batch's partials and
  // work items start afresh (the partials of a batch are reduced before the
  // next batch's pair panels run)
  std::vector<PairWork> work[2][2];
  std::vector<RowRed> red;
  long long pmax = 0;
  const int nbat = (int)e->bat.size() - 1;
  e->bat_red.assign(nbat + 1, 0);
  for (int a = 0; a < 2; ++a)
    for (int l = 0; l < 2; ++l) e->bat_work[a][l].assign(nbat + 1, 0);
  for (int bt = 0; bt < nbat; ++bt) {
    std::vector<PairWork> wb[2][2];
    long long poff = 0;
    e->bat_red[bt] = (int)red.size();
    for (int g = e->bat[bt]; g < e->bat[bt + 1]; ++g) {
      KGroupDev& G = e->h_kg[g];
      const InstDev& I = e->h_inst[G.inst];
      const int J = I.nv * (1 + I.ks);
      const int Wg = rot[g] ? W : CW;
      // partial values of the group's rows are bounded to kPartMax doubles:
      // large row sets take more panels per workgroup
      const long long kPartMax = 32LL << 20;
      int np_g = NP;
      if (!e->group_dense[g] && G.nrows > 0) {
        const long long maxpr = std::max(1LL, kPartMax / G.nrows);
        const long long need = (J + (long long)Wg * np_g - 1) / ((long long)Wg * np_g);
        if (need > maxpr) np_g = (int)((J + Wg * maxpr - 1) / (Wg * maxpr));
      }
      const int span = Wg * np_g;
      const int npr = (J + span - 1) / span;
      G.npr = e->group_dense[g] ? 0 : npr;
      G.dense = e->group_dense[g];
      G.off_part = poff;
      if (!G.dense) poff += (long long)npr * G.nrows;
      if (G.nrows == 0) continue;
      const int var = rot[g] ? 0 : 1;
      const int ostep = rot[g] ? rot_os[g] : osz[g];
      auto push = [&](const PairWork& Wk) {
        wb[1][var].push_back(Wk);
        if (!G.dense) wb[0][var].push_back(Wk);
      };
      if (rot[g] && ostep < G.ndens) {
        // the owner ranges of one slot range 8 items apart, so that they
        // run on one XCD (workgroups go to the XCDs in turn) and read its
        // panels from one L2; blocks of 8 slot ranges, padded with empty
        // items so every group keeps the alignment
        for (int p0 = 0; p0 < npr; p0 += 8)
          for (int o0 = 0; o0 < G.ndens; o0 += ostep)
            for (int p = p0; p < p0 + 8; ++p) {
              if (p < npr)
                push(PairWork{g, p * span, std::min(J, (p + 1) * span), p, o0, std::min(G.ndens, o0 + ostep)});
              else
                push(PairWork{g, 0, 0, 0, 0, 0});
            }
      } else {
        for (int p = 0; p < npr; ++p)
          for (int o0 = 0; o0 < G.ndens; o0 += ostep)
            push(PairWork{g, p * span, std::min(J, (p + 1) * span), p, o0, std::min(G.ndens, o0 + ostep)});
      }
      // in slot order: consecutive rows read consecutive partials
      if (!G.dense)
        for (size_t q = 0; q < slot_row[g].size(); ++q) {
          RowRed R;
          R.pbase = G.off_part + (long long)q;
          R.out = e->h_row_out[slot_row[g][q]];
          R.stride = G.nrows;
          R.npr = npr;
          red.push_back(R);
        }
    }
    pmax = std::max(pmax, poff);
    // longest groups first so the tail of the launch is short
    for (int a = 0; a < 2; ++a)
      for (int l = 0; l < 2; ++l) {
        std::stable_sort(wb[a][l].begin(), wb[a][l].end(), [&](const PairWork& x, const PairWork& y) {
          return e->h_kg[x.g].nrows > e->h_kg[y.g].nrows;
        });
        // a group with owner ranges starts at a multiple of 8
        e->bat_work[a][l][bt] = (int)work[a][l].size();
        std::vector<PairWork>& al = work[a][l];
        const size_t base = al.size();
        for (size_t i = 0; i < wb[a][l].size(); ++i) {
          const PairWork& x = wb[a][l][i];
          const bool start = i == 0 || wb[a][l][i - 1].g != x.g;
          if (start && rot[x.g] && rot_os[x.g] < e->h_kg[x.g].ndens)
            while ((al.size() - base) % 8) al.push_back(PairWork{x.g, 0, 0, 0, 0, 0});
          al.push_back(x);
        }
      }
  }
  e->bat_red[nbat] = (int)red.size();
  for (int a = 0; a < 2; ++a)
    for (int l = 0; l < 2; ++l) e->bat_work[a][l][nbat] = (int)work[a][l].size();
  dfree(&e->d_chunks);
  dfree(&e->d_ent);
  dfree(&e->d_ecidx);
  dfree(&e->d_red);
  dfree(&e->d_part);
  dfree(&e->d_kg);
  for (int a = 0; a < 2; ++a)
    for (int l = 0; l < 2; ++l) {
      dfree(&e->d_work[a][l]);
      e->nwork[a][l] = (int)work[a][l].size();
    }
  e->nred = (int)red.size();
  int rc = upload(e->stream, &e->d_chunks, chunks);
  if (!rc) rc = upload(e->stream, &e->d_ent, ent);
  if (!rc) rc = upload(e->stream, &e->d_ecidx, ecidx);
  if (!rc) rc = upload(e->stream, &e->d_red, red);
  if (!rc) rc = upload(e->stream, &e->d_kg, e->h_kg);
  for (int a = 0; a < 2 && !rc; ++a)
    for (int l = 0; l < 2 && !rc; ++l) rc = upload(e->stream, &e->d_work[a][l], work[a][l]);
  if (!rc && hipMalloc(&e->d_part, sizeof(double) * (size_t)std::max<long long>(pmax, 1)) != hipSuccess)
    rc = fail(DBS_ERR_HIP, "pair tables: hipMalloc");
  return rc;
}

static void mark(dbs_engine* e, const char* name) {
  if (!e->profiling) return;
  if (e->prof_pos >= e->prof_ev.size()) {
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return;  // profiling only: the phase is skipped
    e->prof_ev.push_back(ev);
    e->prof_names.emplace_back();
  }
  (void)hipEventRecord(e->prof_ev[e->prof_pos], e->stream);  // profiling only
  e->prof_names[e->prof_pos] = name;
  e->prof_pos++;
}

template <int W, int DPT>
static void launch_rot(dbs_engine* e, int adj, unsigned grid, const PairWork* wk, int n, hipStream_t s) {
  const int b = e->rot_block;
  const size_t shm = e->rot_shm;
  if (e->onesided) {
    if (adj)
      k_pair_rot<W, true, true, DPT><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                          e->F, e->G, e->d_part);
    else
      k_pair_rot<W, true, false, DPT><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                           e->F, e->G, e->d_part);
  } else {
    if (adj)
      k_pair_rot<W, false, true, DPT><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                           e->F, e->G, e->d_part);
    else
      k_pair_rot<W, false, false, DPT><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                            e->F, e->G, e->d_part);
  }
}

template <int W>
static void launch_chunk(dbs_engine* e, int adj, unsigned grid, const PairWork* wk, int n, hipStream_t s) {
  const int b = e->chunk_block;
  const size_t shm = e->chunk_shm;
  if (e->onesided) {
    if (adj)
      k_pair_chunk<W, true, true><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                       e->F, e->G, e->d_part);
    else
      k_pair_chunk<W, true, false><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                        e->F, e->G, e->d_part);
  } else {
    if (adj)
      k_pair_chunk<W, false, true><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                        e->F, e->G, e->d_part);
    else
      k_pair_chunk<W, false, false><<<grid, b, shm, s>>>(n, wk, e->d_kg, e->d_inst, e->d_chunks, e->d_ent,
                                                         e->F, e->G, e->d_part);
  }
}

static void launch_pair(dbs_engine* e, int adj, int bt, hipStream_t s) {
  const int cap = e->tune[3] * 4;
  const int r0 = e->bat_work[adj][0][bt], nr = e->bat_work[adj][0][bt + 1] - r0;
  const int c0 = e->bat_work[adj][1][bt], nc = e->bat_work[adj][1][bt + 1] - c0;
  if (nr > 0) {
    const unsigned grid = (unsigned)std::min(nr, cap);
    const PairWork* wk = e->d_work[adj][0] + r0;
    switch (e->pair_w) {
      case 4:
        if (e->rot_dpt == 1) launch_rot<4, 1>(e, adj, grid, wk, nr, s);
        else launch_rot<4, 2>(e, adj, grid, wk, nr, s);
        break;
      case 8:
        if (e->rot_dpt == 1) launch_rot<8, 1>(e, adj, grid, wk, nr, s);
        else launch_rot<8, 2>(e, adj, grid, wk, nr, s);
        break;
      default: launch_rot<16, 1>(e, adj, grid, wk, nr, s); break;
    }
  }
  if (nc > 0) {
    const unsigned grid = (unsigned)std::min(nc, cap);
    const PairWork* wk = e->d_work[adj][1] + c0;
    if (e->chunk_w == 8) launch_chunk<8>(e, adj, grid, wk, nc, s);
    else launch_chunk<16>(e, adj, grid, wk, nc, s);
  }
}

// orbital values at the points; with the adjoint the coefficient gather
// into the adjacency entries rides in the same launch
static void run_points(dbs_engine* e, int with_adj) {
  hipStream_t s = e->stream;
  mark(e, "start");
  const int cap = e->tune[3];
  const long long ng = with_adj ? e->nent : 0;
  if (e->ov_shm <= 64 * 1024) {
    const int nbo = (int)grid_for(e->n_ivc, 1, cap * 4);
    const int nbg = ng > 0 ? (int)grid_for(ng, 256, cap * 4) : 0;
    k_orb_values<<<nbo + nbg, 256, e->ov_shm, s>>>(e->ninst, e->d_seg_ivc, e->d_inst, e->pq, e->orbv, e->n_ivc,
                                                   nbo, ng, e->d_ecidx, e->coef, e->d_ent);
    mark(e, "orb_points");
  } else {
    k_orb_points<<<grid_for(e->total_orbv, 256, cap * 4), 256, 0, s>>>(
        e->ninst, e->d_seg_orb, e->d_inst, e->pq, e->orbv, e->total_orbv);
    mark(e, "orb_points");
    if (ng > 0) k_adj_cgather<<<grid_for(ng, 256, cap * 4), 256, 0, s>>>(ng, e->d_ecidx, e->coef, e->d_ent);
    mark(e, "coef_gather");
  }
}

// the orbital point gradients of groups [g0, g1) (all groups, or one batch:
// a batch continues the sums of the earlier batches from gorb)
static void run_adj_orb(dbs_engine* e, int g0, int g1) {
  hipStream_t s = e->stream;
  const int cap = e->tune[3];
  if (e->adjacc_u > 0) {
    const long long per = (e->adj_items + 7) / 8;
    const int nsg = e->adjacc_sg == 64 ? 4 : 1;  // items per workgroup
    const unsigned grid = 8u * (unsigned)std::min<long long>((per + nsg - 1) / nsg, std::max(1, cap * 4 / 8));
    const int sl = (int)(e->adjacc_shm / sizeof(double) / nsg);
    const size_t shm = e->adjacc_shm;
    if (e->adjacc_sg == 64)
      k_adj_orb_acc<1, 64><<<grid, 256, shm, s>>>(e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->orbv, e->wt,
                                                  e->G, e->gorb, e->adj_items, per, e->d_accf, sl);
    else if (e->adjacc_u == 1)
      k_adj_orb_acc<1, 256><<<grid, 256, shm, s>>>(e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->orbv, e->wt,
                                                   e->G, e->gorb, e->adj_items, per, e->d_accf, sl);
    else if (e->adjacc_u == 2)
      k_adj_orb_acc<2, 256><<<grid, 256, shm, s>>>(e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->orbv, e->wt,
                                                   e->G, e->gorb, e->adj_items, per, e->d_accf, sl);
    else
      k_adj_orb_acc<4, 256><<<grid, 256, shm, s>>>(e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->orbv, e->wt,
                                                   e->G, e->gorb, e->adj_items, per, e->d_accf, sl);
  }
  if (e->adjm_shm > 0) {
    if (!e->adjacc_all)
      k_adj_orb_mfma<<<grid_for(e->adj_items, 1, cap * 4), 256, e->adjm_shm, s>>>(
          e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->d_dpk, e->orbv, e->wt, e->G, e->gorb, e->adj_items,
          e->d_accf, g0, g1);
  } else if (e->adj_shm > 0) {
    k_adj_orb_lds<<<grid_for(e->adj_items, 1, cap * 4), 256, e->adj_shm, s>>>(
        e->ninst, e->d_seg_adj, e->d_inst, e->d_kg, e->d_dens, e->orbv, e->wt, e->G, e->adj_mode, e->gorb,
        e->adj_items, g0, g1);
  } else {
    k_adj_orb_points<<<grid_for(e->total_orbv, 256, cap * 4), 256, 0, s>>>(
        e->ninst, e->d_seg_orb, e->d_inst, e->d_kg, e->d_dens, e->d_oadjoff, e->d_oadj, e->orbv, e->wt,
        e->G, e->gorb, e->total_orbv, g0, g1);
  }
  mark(e, "adj_orb_points");
}

// one batch of groups: features, prefix scans, pair panels (fused with the
// gradient gather of the adjoint) and row values; with the adjoint the
// suffix scans and the batch's share of the orbital point gradients
static void run_batch(dbs_engine* e, int bt, int with_adj) {
  hipStream_t s = e->stream;
  const int cap = e->tune[3];
  if (e->featm_shm > 0) {
    const int f0 = e->bat_fit[bt], nf = e->bat_fit[bt + 1] - f0;
    if (nf > 0)
      k_features_mfma<<<grid_for(nf, 4, cap * 4), 256, e->featm_shm, s>>>(
          nf, e->d_fitems + f0, e->d_kg, e->d_inst, e->d_pidx, e->orbv, e->wt, e->F, e->featm_wave);
  } else {
    const int i0 = e->bat_items[bt], ni = e->bat_items[bt + 1] - i0;
    if (ni > 0)
      k_features<<<grid_for(ni, 1, cap), e->tune[1], e->feat_shm, s>>>(
          ni, e->d_items + i0, e->d_kg, e->d_inst, e->d_dens, e->orbv, e->wt, e->F, e->feat_stage);
  }
  mark(e, "features");
  const int s0 = e->bat_scan[bt], nsc = e->bat_scan[bt + 1] - s0;
  if (nsc > 0)
    k_scan_cols<<<grid_for(nsc, 1, cap), 256, 0, s>>>(nsc, e->d_scan_items + s0, e->d_kg, e->d_inst, e->F, 0);
  mark(e, "prefix_scan");
  const int wa = with_adj ? 1 : 0;
  if (e->bat_work[wa][0][bt + 1] - e->bat_work[wa][0][bt] + e->bat_work[wa][1][bt + 1] - e->bat_work[wa][1][bt] > 0) {
    launch_pair(e, wa, bt, s);
    mark(e, with_adj ? "pair_panel" : "contract_rows");
  }
  const int r0 = e->bat_red[bt], nred = e->bat_red[bt + 1] - r0;
  if (nred > 0) {
    k_rows_reduce<<<grid_for((long long)nred * DBS_RED_LANES, 256, cap * 4), 256, 0, s>>>(nred, e->d_red + r0,
                                                                                          e->d_part, e->vals);
    mark(e, "rows_reduce");
  }
  const int t0 = e->bat_tiles[bt], nt = e->bat_tiles[bt + 1] - t0;
  if (nt > 0) {
    k_contract_dense<<<grid_for(nt, 1, cap * 4), 64 * DBS_DENSE_SPLIT, 0, s>>>(
        nt, e->d_tiles + t0, e->d_kg, e->d_inst, e->F, e->d_pairrow, e->d_pairoff, e->d_nextrow, e->d_row_out,
        e->onesided, e->vals);
    mark(e, "contract_dense");
  }
  if (with_adj) {
    if (nsc > 0)
      k_scan_cols<<<grid_for(nsc, 1, cap), 256, 0, s>>>(nsc, e->d_scan_items + s0, e->d_kg, e->d_inst, e->G, 1);
    mark(e, "adj_suffix_scan");
    run_adj_orb(e, e->bat[bt], e->bat[bt + 1]);
  }
}

static int run_all(dbs_engine* e, int with_adj) {
  hipStream_t s = e->stream;
  const int cap = e->tune[3];
  run_points(e, with_adj);
  for (int bt = 0; bt + 1 < (int)e->bat.size(); ++bt) run_batch(e, bt, with_adj);
  k_energy<<<grid_for(e->nechunks, 1, cap), 256, 0, s>>>(e->nechunks, e->d_echunks, e->d_roff, e->coef, e->vals,
                                                         e->d_epart, e->ninst, e->d_echoff, e->energy,
                                                         e->d_edone);
  mark(e, "energy");
  if (with_adj) {
    if (e->proj_shm <= 64 * 1024) {
      k_adj_proj<<<grid_for(e->n_ivc, 1, cap * 4), 256, e->proj_shm, s>>>(e->ninst, e->d_seg_ivc, e->d_inst,
                                                                         e->gorb, e->d_ppoff, e->d_ppart, e->n_ivc);
      k_adj_coef_sum<<<grid_for(e->total_grad, 256, cap * 4), 256, 0, s>>>(e->ninst, e->d_seg_coef, e->d_inst,
                                                                           e->d_ppoff, e->d_ppart, e->grad,
                                                                           e->total_grad);
    } else {
      k_adj_coef<<<grid_for(e->total_grad, 4, cap * 4), 256, 0, s>>>(e->ninst, e->d_seg_coef, e->d_inst,
                                                                      e->gorb, e->grad, e->total_grad);
    }
    mark(e, "adj_coef");
  }
  DBS_HIP(hipGetLastError());
  return DBS_OK;
}

static void compute_phase_work(dbs_engine* e) {
  const int ninst = e->ninst;
  const std::vector<int>& norb = e->norb;
  // per phase (one launch): algorithmic flops and bytes (each input read
  // once, each output written once)
  double pts = 0, pts_fl = 0, pq_b = 0, feat_fl = 0, feat_b = 0, scan_fl = 0, scan_b = 0;
  double rows_fl = 0, rows_b = 0, rows_dense_fl = 0, gath_fl = 0, orbp_fl = 0, coef_fl = 0, red_b = 0;
  for (int i = 0; i < ninst; ++i) {
    const InstDev& I = e->h_inst[i];
    const double p = (double)norb[i] * 2 * I.nv * I.np;
    pts += p;
    pts_fl += p * 2 * I.ks;
    pq_b += (double)norb[i] * 2 * I.ns * 8;
    coef_fl += (double)norb[i] * 2 * I.ns * 2.0 * I.ks * I.np;
  }
  for (int g = 0; g < e->nkg; ++g) {
    const KGroupDev& G = e->h_kg[g];
    const InstDev& I = e->h_inst[G.inst];
    const double nv = I.nv, ks = I.ks, nd = G.ndens, nr = (double)e->group_rows[g].size();
    const double D = feat_dim(I.nv, I.ks);
    const double rho = e->kind == 0 ? 3.0 : 1.0;  // flops per density point value
    feat_fl += nd * nv * (I.np * rho + 4 * ks + ks + 2 * ks * ks);
    feat_b += nd * D * 8;
    scan_fl += nd * nv;
    scan_b += nd * nv * 16;
    const double prow = (e->onesided ? 1.0 : 2.0) * D;  // flops of one row value
    if (e->group_dense[g]) {
      rows_dense_fl += 2.0 * nd * nd * D;
    } else {
      rows_fl += nr * prow;
      rows_b += nd * D * 8;
      red_b += nr * 8 * (1.0 + G.npr);
    }
    gath_fl += 2.0 * nr * (e->onesided ? 1.0 : 2.0) * D;
    orbp_fl += 2.0 * nd * nv * I.np * 3.0;
  }
  const double coef_b = (double)e->total_rows * 8;
  e->phase_work["orb_points"] = {pts_fl, pq_b + pts * 8};
  e->phase_work["features"] = {feat_fl, pts * 8 + feat_b};
  e->phase_work["prefix_scan"] = {scan_fl, scan_b};
  // pair panels: features read once, (adjoint) gradients written once
  e->phase_work["contract_rows"] = {rows_fl, rows_b};
  e->phase_work["pair_panel"] = {rows_fl + gath_fl, feat_b * 2 + coef_b};
  e->phase_work["rows_reduce"] = {0.0, red_b};
  e->phase_work["contract_dense"] = {rows_dense_fl, 0.0};
  e->phase_work["adj_suffix_scan"] = {scan_fl, scan_b};
  e->phase_work["adj_orb_points"] = {orbp_fl, feat_b + pts * 8 * 2};
  e->phase_work["adj_coef"] = {coef_fl, pts * 8 + pq_b};
}

extern "C" {

// Items of the matrix-core feature kernel, batch by batch: runs of
// consecutive groups of one instance (within one batch) with the kind, at
// most fg per item, spread evenly.  Returns the item count.
static long long feature_items(dbs_engine* e, int fg, std::vector<int4>* out) {
  long long n = 0;
  const int nbat = (int)e->bat.size() - 1;
  if (out) e->bat_fit.assign(nbat + 1, 0);
  for (int bt = 0; bt < nbat; ++bt) {
    if (out) e->bat_fit[bt] = (int)out->size();
    for (int i = 0; i < e->ninst; ++i) {
      const InstDev& I = e->h_inst[i];
      const int g0 = std::max(I.kg0, e->bat[bt]) - I.kg0, g1 = std::min(I.kg0 + I.nkg, e->bat[bt + 1]) - I.kg0;
      for (int kd = 0; kd < 4; ++kd) {
        int gi = g0;
        while (gi < g1) {
          if (!((e->h_kg[I.kg0 + gi].gmask >> kd) & 1)) { ++gi; continue; }
          int ge = gi;
          while (ge < g1 && ((e->h_kg[I.kg0 + ge].gmask >> kd) & 1)) ++ge;
          const int run = ge - gi, nch = (run + fg - 1) / fg;
          n += (long long)nch * I.nv * (I.ks + 2);
          for (int c = 0; out && c < nch; ++c) {
            const int a = gi + run * c / nch, b = gi + run * (c + 1) / nch;
            for (int iv = 0; iv < I.nv; ++iv)
              for (int q = 0; q < I.ks + 2; ++q) out->push_back(make_int4(I.kg0 + a, iv, kd | ((b - a) << 8), q));
          }
          gi = ge;
        }
      }
    }
  }
  if (out) e->bat_fit[nbat] = (int)out->size();
  return n;
}

static int build_feature_items(dbs_engine* e) {
  std::vector<int4> fit;
  feature_items(e, e->feat_fg, &fit);
  e->nfitems = (int)fit.size();
  dfree(&e->d_fitems);
  return upload(e->stream, &e->d_fitems, fit);
}

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
  if (hipSetDevice(e->device) != hipSuccess) return bail(fail(DBS_ERR_HIP, "dbs_engine_create: device"));
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
      // densities of the group in (kind, a, c) order: neighbouring threads
      // then read neighbouring orbital-pair entries
      {
        std::vector<int> perm(G.ndens);
        std::iota(perm.begin(), perm.end(), 0);
        auto key = [&](int x) {
          const DensDev& D = e->h_dens[G.dens0 + x];
          return std::make_tuple(D.kind, D.a, D.c);
        };
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return key(x) < key(y); });
        std::vector<int> newpos(G.ndens);
        std::vector<DensDev> sorted(G.ndens);
        for (int q = 0; q < G.ndens; ++q) {
          newpos[perm[q]] = q;
          sorted[q] = e->h_dens[G.dens0 + perm[q]];
        }
        for (int q = 0; q < G.ndens; ++q) e->h_dens[G.dens0 + q] = sorted[q];
        for (int r : grows) {
          e->h_rows[r].L = G.dens0 + newpos[e->h_rows[r].L - G.dens0];
          e->h_rows[r].R = G.dens0 + newpos[e->h_rows[r].R - G.dens0];
        }
        G.gmask = 0;
        for (int q = 0; q < G.ndens; ++q) G.gmask |= 1 << e->h_dens[G.dens0 + q].kind;
        G.full = (G.gmask == (1 << DK_TOT) && G.ndens == norb[i] * (norb[i] + 1) / 2) ? 1 : 0;
      }
      G.off_F = foff;
      foff += (long long)D * G.ndens;
      G.nrows = (int)grows.size();
      G.row0 = grows.empty() ? (int)e->h_rows.size() : grows[0];
      G.off_part = 0; G.npr = 0; G.dense = 0; G.chunk0 = 0; G.nchunk = 0; G.off_pidx = 0;
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
  e->total_feat = layout_batches(e);  // one batch: off_F as above
  (void)foff;
  e->total_w = woff;
  // work items (group, iv)
  std::vector<int2> items;
  for (int g = 0; g < e->nkg; ++g)
    for (int iv = 0; iv < e->h_inst[e->h_kg[g].inst].nv; ++iv) items.push_back(make_int2(g, iv));
  e->nitems = (int)items.size();
  // scan items (group, first of 32 columns)
  std::vector<int2> scan_items;
  for (int g = 0; g < e->nkg; ++g)
    for (int c0 = 0; c0 < e->h_kg[g].ndens; c0 += DBS_SCAN_COLS) scan_items.push_back(make_int2(g, c0));
  e->nscan_items = (int)scan_items.size();
  // adjacency per density in row order: (0: left density of the row, -1:
  // right density, row)
  e->h_adj.assign(e->ndens, {});
  for (size_t r = 0; r < e->h_rows.size(); ++r) {
    const RowDev& R = e->h_rows[r];
    e->h_adj[R.L].push_back(make_int2(0, (int)r));
    e->h_adj[R.R].push_back(make_int2(-1, (int)r));
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
  if (hipMalloc(&e->d_edone, sizeof(unsigned)) != hipSuccess ||
      hipMemsetAsync(e->d_edone, 0, sizeof(unsigned), s) != hipSuccess)
    rc |= fail(DBS_ERR_HIP, "dbs_engine_create: energy counter");
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
    // matrix-core features: per wave V[2][ks+1][NO] + weights
    size_t wneed2 = 0;
    bool fits = true;
    for (int g = 0; g < e->nkg; ++g) e->max_nd = std::max(e->max_nd, e->h_kg[g].ndens);
    for (auto& I : e->h_inst) {
      if (I.norb > 64) fits = false;
      const size_t NO = I.norb <= 16 ? 16 : (I.norb <= 32 ? 32 : 64);
      wneed2 = std::max(wneed2, 2 * (size_t)(I.ks + 1) * NO + kFeatGroups * (I.ks + 1));
    }
    wneed2 = (wneed2 + 1) & ~(size_t)1;  // 16-byte aligned slices
    e->featm_wave = (int)wneed2;
    e->featm_shm = fits ? 4 * wneed2 * sizeof(double) : 0;
    // a kernel whose LDS size cannot be granted falls back to the next form
    if (e->featm_shm > 64 * 1024 &&
        hipFuncSetAttribute((const void*)k_features_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)e->featm_shm) != hipSuccess)
      e->featm_shm = 0;
    if (e->featm_shm > 0) {
      std::vector<int> pidx;
      for (int g = 0; g < e->nkg; ++g) {
        KGroupDev& G = e->h_kg[g];
        const InstDev& I = e->h_inst[G.inst];
        const int NO = I.norb <= 16 ? 16 : (I.norb <= 32 ? 32 : 64);
        G.off_pidx = (long long)pidx.size();
        for (int kd = 0; kd < 4; ++kd) {
          if (!((G.gmask >> kd) & 1)) continue;
          const size_t base = pidx.size();
          pidx.resize(base + (size_t)NO * NO, -1);
          for (int d = 0; d < G.ndens; ++d) {
            const DensDev& D = e->h_dens[G.dens0 + d];
            if (D.kind == kd) pidx[base + (size_t)D.a * NO + D.c] = d;
          }
        }
      }
      e->feat_fg = feature_items(e, kFeatGroups, nullptr) >= 4096 ? kFeatGroups : kFeatGroups / 2;
      rc |= build_feature_items(e);
      rc |= upload(s, &e->d_pidx, pidx);
    }

    e->adj_mode = kind == 0 ? 0 : (kind <= 4 ? 1 : 2);
    e->adj_shm = aneed <= lds_max ? sizeof(double) * aneed : 0;
    // matrix-core adjoint (total densities, up to 64 orbitals)
    {
      size_t need = 0;
      bool ok = kind == 0;
      for (auto& I : e->h_inst) {
        if (I.norb > 64 || 2 * I.ks > 32 || I.ks > 16) ok = false;
        const size_t NO = I.norb <= 16 ? 16 : (I.norb <= 32 ? 32 : 64);
        need = std::max(need, NO * 32 + NO * 16 + 32 + 4 * NO * (NO + 1));
      }
      int lmx = 0;
      if (hipDeviceGetAttribute(&lmx, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) != hipSuccess)
        lmx = 64 * 1024;
      e->adjm_shm = ok && need * sizeof(double) <= (size_t)lmx ? need * sizeof(double) : 0;
      if (e->adjm_shm > 64 * 1024 &&
          hipFuncSetAttribute((const void*)k_adj_orb_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)e->adjm_shm) != hipSuccess)
        e->adjm_shm = 0;
      std::vector<int> pk(std::max(e->ndens, 1), 0);
      for (int d = 0; d < e->ndens; ++d) pk[d] = (e->h_dens[d].a & 255) | ((e->h_dens[d].c & 255) << 8);
      rc |= upload(s, &e->d_dpk, pk);
      // register-accumulating form, per instance: every group complete,
      // ks <= 10, up to 1024 densities (the choice depends on the instance
      // alone, so batched and single-instance results are identical)
      std::vector<int> accf(ninst, 0);
      int maxnd = 0, nacc = 0;
      size_t aneed2 = 0;
      for (int i = 0; i < ninst && e->adjm_shm > 0; ++i) {
        const InstDev& I = e->h_inst[i];
        bool a = I.ks <= 10 && I.nkg > 0;
        int nd = 0;
        for (int gi = 0; gi < I.nkg; ++gi) {
          const KGroupDev& G = e->h_kg[I.kg0 + gi];
          if (!G.full) a = false;
          nd = std::max(nd, G.ndens);
          // the kernel finds pair (a,c), a <= c, at a*norb - a(a-1)/2 + c - a
          for (int d = 0; a && d < G.ndens; ++d) {
            const DensDev& D = e->h_dens[G.dens0 + d];
            if (D.a > D.c || D.a * I.norb - D.a * (D.a - 1) / 2 + (D.c - D.a) != d) a = false;
          }
        }
        if (nd > 1024) a = false;
        if (!a) continue;
        accf[i] = 1;
        ++nacc;
        maxnd = std::max(maxnd, nd);
        aneed2 = std::max(aneed2, (size_t)((I.ks + 2) / 2) * nd + (size_t)I.norb * 2 * (I.ks + 1));
      }
      const int U = maxnd <= 256 ? 1 : (maxnd <= 512 ? 2 : 4);
      if (aneed2 * sizeof(double) > (size_t)lmx) { std::fill(accf.begin(), accf.end(), 0); nacc = 0; }
      e->adjacc_u = nacc > 0 ? U : 0;
      e->adjacc_sg = maxnd <= 64 && 4 * aneed2 * sizeof(double) <= (size_t)lmx ? 64 : 256;
      e->adjacc_all = nacc == ninst;
      e->adjacc_shm = e->adjacc_u ? aneed2 * sizeof(double) * (e->adjacc_sg == 64 ? 4 : 1) : 0;
      rc |= upload(s, &e->d_accf, accf);
      e->h_accf = accf;
      e->adjacc_u0 = e->adjacc_u;
      e->adjacc_all0 = e->adjacc_all;
      if (e->adjacc_shm > 64 * 1024) {
        const void* fs[] = {(const void*)k_adj_orb_acc<1, 64>, (const void*)k_adj_orb_acc<1, 256>,
                            (const void*)k_adj_orb_acc<2, 256>, (const void*)k_adj_orb_acc<4, 256>};
        for (const void* f : fs)
          if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->adjacc_shm) != hipSuccess)
            rc |= fail(DBS_ERR_HIP, "dbs_engine_create: LDS size of the register-summed adjoint");
      }
    }
    std::vector<long long> seg_adj(ninst);
    long long tot = 0;
    for (int i = 0; i < ninst; ++i) {
      seg_adj[i] = tot;
      tot += (long long)e->h_inst[i].nv * e->h_inst[i].ks;
    }
    e->adj_items = tot;
    rc |= upload(s, &e->d_seg_adj, seg_adj);
    // (interval, component) items of the orbital-value and projection kernels
    std::vector<long long> seg_ivc(ninst), ppoff(ninst);
    long long ti = 0, tp = 0;
    size_t ovn = 0, prn = 0;
    for (int i = 0; i < ninst; ++i) {
      const InstDev& I = e->h_inst[i];
      seg_ivc[i] = ti;
      ti += 2LL * I.nv;
      ppoff[i] = tp;
      tp += (long long)I.norb * 2 * I.nv * I.ks;
      ovn = std::max(ovn, (size_t)I.np * I.ks + (size_t)I.norb * I.ks);
      prn = std::max(prn, (size_t)I.np * I.ks + (size_t)I.norb * I.np);
    }
    e->n_ivc = ti;
    e->ov_shm = ovn * sizeof(double);
    e->proj_shm = prn * sizeof(double);
    rc |= upload(s, &e->d_seg_ivc, seg_ivc);
    rc |= upload(s, &e->d_ppoff, ppoff);
    rc |= dalloc(&e->d_ppart, tp);
  }
  if (rc) return bail(rc);
  std::vector<long long> row_out(e->h_row_out);
  rc |= upload(s, &e->d_inst, e->h_inst);
  rc |= upload(s, &e->d_dens, e->h_dens);
  rc |= upload(s, &e->d_row_out, row_out);
  rc |= upload(s, &e->d_seg_orb, seg_orb);
  rc |= upload(s, &e->d_seg_coef, seg_coef);
  rc |= upload(s, &e->d_seg_w, seg_w);
  rc |= upload(s, &e->d_roff, roff);
  rc |= upload(s, &e->d_items, items);
  rc |= upload(s, &e->d_scan_items, scan_items);
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
  if (!rc) rc = build_pair_tables(e);
  if (rc) return bail(rc);
  // the quadrature weights of every (instance,k) group depend on the grid
  // and k only: computed once here
  k_weights<<<grid_for(e->total_w, 256, 4096), 256, 0, s>>>(e->nkg, e->d_seg_w, e->d_kg, e->d_inst, e->wt,
                                                             e->total_w);
  compute_phase_work(e);
  if (hipStreamSynchronize(s) != hipSuccess) return bail(fail(DBS_ERR_HIP, "dbs_engine_create: setup"));
  *out = e;
  return DBS_OK;
}

int dbs_engine_destroy(dbs_engine* e) {
  if (!e) return DBS_OK;
  // teardown: failures here leave nothing to recover
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (hipEvent_t ev : e->prof_ev) (void)hipEventDestroy(ev);
  void* bufs[] = {e->pq, e->coef, e->vals, e->grad, e->energy, e->orbv, e->gorb, e->wt, e->F, e->G,
                  e->d_inst, e->d_kg, e->d_dens, e->d_row_out, e->d_seg_orb, e->d_seg_coef,
                  e->d_seg_w, e->d_roff, e->d_items, e->d_scan_items, e->d_oadjoff, e->d_oadj,
                  e->d_tiles, e->d_pairrow, e->d_pairoff, e->d_nextrow, e->plain_d, e->plain_c,
                  e->d_echunks, e->d_echoff, e->d_epart, e->d_seg_adj, e->d_chunks, e->d_ent,
                  e->d_ecidx, e->d_red, e->d_part, e->d_fitems, e->d_pidx, e->d_dpk, e->d_accf, e->d_seg_ivc, e->d_ppoff, e->d_ppart, e->d_edone,
                  e->d_work[0][0], e->d_work[0][1], e->d_work[1][0], e->d_work[1][1]};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
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
  return run_all(e, 0);
}

int dbs_engine_forward_adjoint(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_forward_adjoint: null engine");
  return run_all(e, 1);
}

int dbs_engine_sync(dbs_engine* e) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_sync: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  return DBS_OK;
}

int dbs_engine_graph_capture(dbs_engine* e, int with_adjoint) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_graph_capture: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
  if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
  DBS_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = run_all(e, with_adjoint ? 1 : 0);
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
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
  DBS_CHECK(e && p && n >= 1 && n <= 7, DBS_ERR_ARG, "dbs_engine_set_tuning: bad arguments");
  int t[7];
  std::memcpy(t, e->tune, sizeof(t));
  for (int i = 0; i < n; ++i) t[i] = p[i];
  DBS_CHECK(t[0] >= 0 && t[0] <= 2, DBS_ERR_ARG, "tuning[0]: path 0 pair panels, 1 dense, 2 auto");
  DBS_CHECK(t[1] >= 64 && t[1] <= 256 && t[1] % 64 == 0, DBS_ERR_ARG, "tuning[1]: block size 64..256");
  DBS_CHECK(t[2] == 4 || t[2] == 8 || t[2] == 16, DBS_ERR_ARG, "tuning[2]: panel width 4, 8 or 16");
  DBS_CHECK(t[3] >= 64 && t[3] <= 65536, DBS_ERR_ARG, "tuning[3]: grid cap");
  DBS_CHECK(t[4] >= 64 && t[4] <= 512 && t[4] % 64 == 0, DBS_ERR_ARG,
            "tuning[4]: owner densities per chunked pair-panel workgroup, 64..512");
  DBS_CHECK(t[5] >= 1 && t[5] <= 256, DBS_ERR_ARG, "tuning[5]: panels per workgroup 1..256");
  DBS_CHECK(t[6] == 1 || t[6] == 2 || t[6] == 4 || t[6] == 8, DBS_ERR_ARG,
            "tuning[6]: owner ranges per rotated group 1, 2, 4 or 8");
  DBS_HIP(hipStreamSynchronize(e->stream));
  const bool path_changed = t[0] != e->tune[0];
  const bool pair_changed = path_changed || t[2] != e->tune[2] || t[4] != e->tune[4] || t[5] != e->tune[5] ||
                            t[6] != e->tune[6];
  std::memcpy(e->tune, t, sizeof(t));
  if (path_changed) {
    for (int g = 0; g < e->nkg; ++g) e->group_dense[g] = (char)choose_dense(e, e->h_kg[g], e->group_rows[g].size());
    int rc = build_dense_tables(e);
    if (rc) return rc;
  }
  if (pair_changed) {
    int rc = build_pair_tables(e);
    if (rc) return rc;
    compute_phase_work(e);
    DBS_HIP(hipStreamSynchronize(e->stream));
  }
  // a captured graph holds the old launch shapes
  if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
  if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
  return DBS_OK;
}

int dbs_engine_set_memory_budget(dbs_engine* e, size_t bytes) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_set_memory_budget: null engine");
  DBS_HIP(hipStreamSynchronize(e->stream));
  e->mem_budget = (long long)bytes;
  const long long nf = layout_batches(e);
  dfree(&e->F);
  dfree(&e->G);
  DBS_HIP(hipMalloc(&e->F, (size_t)std::max(nf, 1LL) * sizeof(double)));
  DBS_HIP(hipMalloc(&e->G, (size_t)std::max(nf, 1LL) * sizeof(double)));
  e->total_feat = nf;
  // the register-summed adjoint (k_adj_orb_acc) adds all of an instance's
  // groups in registers before the orbital products, which a batch cannot
  // continue: under a budget the other adjoint kernels run
  std::vector<int> accf(e->h_accf);
  if (bytes > 0) {
    std::fill(accf.begin(), accf.end(), 0);
    e->adjacc_u = 0;
    e->adjacc_all = false;
  } else {
    e->adjacc_u = e->adjacc_u0;
    e->adjacc_all = e->adjacc_all0;
  }
  int rc = DBS_OK;
  if (!accf.empty()) {
    dfree(&e->d_accf);
    rc = upload(e->stream, &e->d_accf, accf);
  }
  if (!rc) rc = build_dense_tables(e);
  if (!rc) rc = build_pair_tables(e);
  if (!rc && e->featm_shm > 0) rc = build_feature_items(e);
  if (rc) return rc;
  compute_phase_work(e);
  DBS_HIP(hipStreamSynchronize(e->stream));
  if (e->gexec) { (void)hipGraphExecDestroy(e->gexec); e->gexec = nullptr; }
  if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
  return DBS_OK;
}

int dbs_engine_memory(dbs_engine* e, int* nbatch, size_t* feature_bytes) {
  DBS_CHECK(e, DBS_ERR_ARG, "dbs_engine_memory: null engine");
  if (nbatch) *nbatch = (int)e->bat.size() - 1;
  if (feature_bytes) *feature_bytes = (size_t)(2 * e->total_feat * (long long)sizeof(double));
  return DBS_OK;
}

int dbs_engine_get_tuning(dbs_engine* e, int n, int* p) {
  DBS_CHECK(e && p && n >= 1 && n <= 7, DBS_ERR_ARG, "dbs_engine_get_tuning: bad arguments");
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
  // the dense path only where some group is dense under "auto"
  bool any_dense = false;
  {
    int save = e->tune[0];
    e->tune[0] = 2;
    for (int g = 0; g < e->nkg; ++g) any_dense |= choose_dense(e, e->h_kg[g], e->group_rows[g].size()) != 0;
    e->tune[0] = save;
  }
  std::vector<int> paths = {0};
  if (any_dense) paths.push_back(2);
  const int fblocks[] = {128, 256};
  const int widths[] = {8, 16};
  const int pblocks[] = {256, 512};
  const int nps[] = {1, 2, 4};
  const int splits[] = {1, 2, 4};
  int best[7];
  std::memcpy(best, e->tune, sizeof(best));
  float best_ms = 1e30f;
  hipEvent_t a, b;
  DBS_HIP(hipEventCreate(&a));
  DBS_HIP(hipEventCreate(&b));
  for (int p : paths)
    for (int fb : fblocks)
      for (int w : widths)
        for (int pb : pblocks)
          for (int np : nps)
          for (int sp : splits) {
            int t[7] = {p, fb, w, e->tune[3], pb, np, sp};
            if (dbs_engine_set_tuning(e, 7, t)) continue;
            int rc = run_all(e, with_adjoint ? 1 : 0);  // warm
            if (rc) continue;
            std::vector<float> ms;
            for (int r = 0; r < reps && !rc; ++r) {
              if (hipEventRecord(a, e->stream) != hipSuccess) { rc = DBS_ERR_HIP; break; }
              rc = run_all(e, with_adjoint ? 1 : 0);
              float x = 0;
              if (!rc && (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                          hipEventElapsedTime(&x, a, b) != hipSuccess))
                rc = DBS_ERR_HIP;
              if (!rc) ms.push_back(x);
            }
            if (rc || ms.empty()) continue;
            std::sort(ms.begin(), ms.end());
            const float med = ms[ms.size() / 2];
            if (med < best_ms) {
              best_ms = med;
              std::memcpy(best, t, sizeof(best));
            }
          }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return dbs_engine_set_tuning(e, 7, best);
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
    if (e->plain_d) (void)hipFree(e->plain_d);
    DBS_HIP(hipMalloc(&e->plain_d, need_d * sizeof(double)));
    e->plain_d_n = need_d;
  }
  if (need_c > e->plain_c_n) {
    if (e->plain_c) (void)hipFree(e->plain_c);
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
  (void)hipFree(d_spec);
  (void)hipFree(d_ls);
  (void)hipFree(d_lr);
  (void)hipFree(d_out);
  return rc;
}

// HM/DM += S*c per target, the entries of one target in list order
__global__ void k_list_accumulate(int ntarget, const long long* __restrict__ tpos, const int* __restrict__ toff,
                                  const int* __restrict__ tent, const double* __restrict__ S,
                                  const double* __restrict__ c, double* __restrict__ HM, double* __restrict__ DM) {
#pragma clang fp contract(off)
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntarget; t += gridDim.x * blockDim.x) {
    const long long p = tpos[t];
    double* dst = p >= 0 ? HM + p : DM + (-p - 1);
    double acc = *dst;
    for (int e = toff[t]; e < toff[t + 1]; ++e) {
      const int q = tent[e];
      const double ss = S[q] * c[q];
      acc = acc + ss;
    }
    *dst = acc;
  }
}

// open-addressing table (linear probing) from 64-bit keys >= 0 to dense
// ids 0, 1, ... in order of first insertion
struct KeyIds {
  std::vector<long long> key;
  std::vector<int> id;
  size_t mask = 0;
  int n = 0;
  void reset(size_t expect) {
    size_t cap = 64;
    while (cap < 2 * expect) cap <<= 1;
    key.assign(cap, -1);
    id.resize(cap);
    mask = cap - 1;
    n = 0;
  }
  // the id of k; a new id (== previous n) when k is new
  int get(long long k) {
    unsigned long long h = (unsigned long long)k * 0x9E3779B97F4A7C15ull;
    size_t q = (size_t)(h ^ (h >> 29)) & mask;
    while (key[q] >= 0) {
      if (key[q] == k) return id[q];
      q = (q + 1) & mask;
    }
    key[q] = k;
    id[q] = n;
    return n++;
  }
};

// Rk_evaluate / Sk_evaluate (DBSR_CI3/rk_evaluate.f90:2-63,
// sk_evaluate.f90:2-63) in the reference's organisation: the list's
// distinct densities (one batched launch; the left densities K1 numbered
// first), one case-2 convolution per distinct left density K1 (one matrix
// product against the context's cached array of (itype, k), permuted to
// K x K), one SUM_AmB per entry, then S*c added to HM(ic,jc) / DM(jc) per
// target in list order.  Every buffer lives in the context's grow-only
// scratch and the host tables go up in one copy, so a flush allocates
// nothing once the scratch and the array cache have grown.
int dbs_rk_evaluate_dev(dbs_ctx* ctx, int itype, int k, int nc, const double* c, const int* K1, const int* K2,
                        const int* K3, const int* K4, double eps_c, int ibi, const double* pq_dev, int norb,
                        double* HM_dev, int ldhm, int nzero, double* DM_dev) {
  DBS_CHECK(ctx && pq_dev && HM_dev, DBS_ERR_ARG, "Rk_evaluate: null argument");
  DBS_CHECK(itype >= 1 && itype <= 6, DBS_ERR_ARG, "Rk_evaluate: unknown type of intgrals");
  DBS_CHECK(nc <= 0 || (c && K1 && K2 && K3 && K4), DBS_ERR_ARG, "Rk_evaluate: null list");
  DBS_CHECK(ibi > 0, DBS_ERR_ARG, "Rk_evaluate: ibi < 1");
  if (nc <= 0) return DBS_OK;
  const int ns = ctx->ns, ks = ctx->ks;
  const int sym = itype >= DBS_PPQQ ? 1 : 0;
  const int nb = sym ? 2 * ks - 1 : ks;
  int lc1, lc2, rc1, rc2;  // components of the left (K1) and right (K2) densities
  switch (itype) {
    case DBS_PPPP: lc1 = lc2 = 0; rc1 = rc2 = 0; break;
    case DBS_QQQQ: lc1 = lc2 = 1; rc1 = rc2 = 1; break;
    case DBS_PQPQ: lc1 = lc2 = 0; rc1 = rc2 = 1; break;
    case DBS_QPQP: lc1 = lc2 = 1; rc1 = rc2 = 0; break;
    case DBS_PPQQ: lc1 = 0; lc2 = 1; rc1 = 0; rc2 = 1; break;
    default: lc1 = 0; lc2 = 1; rc1 = 1; rc2 = 0; break;
  }
  const int cmL = lc1 | (lc2 << 1), cmR = rc1 | (rc2 << 1);
  // the entries kept (|c| >= eps_c, a target in HM or DM) and their targets
  std::vector<int> keep;
  std::vector<long long> pos;
  keep.reserve(nc);
  pos.reserve(nc);
  for (int j = 0; j < nc; ++j) {
    if (std::fabs(c[j]) < eps_c) continue;
    const int a = K1[j] / ibi, cc = K1[j] % ibi, b = K2[j] / ibi, d = K2[j] % ibi;
    DBS_CHECK(a >= 1 && a <= norb && cc >= 1 && cc <= norb && b >= 1 && b <= norb && d >= 1 && d <= norb,
              DBS_ERR_ARG, "Rk_evaluate: orbital index out of range");
    const int ic = K3[j], jc = K4[j];
    long long p;
    if (jc <= nzero) {
      DBS_CHECK(ic >= 1 && jc >= 1 && ic <= ldhm, DBS_ERR_ARG, "Rk_evaluate: HM index out of range");
      p = (long long)(ic - 1) + (long long)ldhm * (jc - 1);
    } else if (ic == jc) {
      DBS_CHECK(DM_dev, DBS_ERR_ARG, "Rk_evaluate: DM needed");
      p = -(long long)jc;  // DM(jc) at -(jc-1)-1
    } else {
      continue;
    }
    keep.push_back(j);
    pos.push_back(p);
  }
  const int n = (int)keep.size();
  if (n == 0) return DBS_OK;
  // densities: key (component pattern, a, c); the left ones take ids
  // 0..nl-1, the right ones follow (sharing a left id when equal)
  const long long nob = (long long)norb + 1;
  auto dkey = [&](int cm, int a, int b) { return ((long long)cm * nob + a) * nob + b; };
  KeyIds dens_ids, targ_ids;
  dens_ids.reset((size_t)n);
  std::vector<int4> spec;
  std::vector<int2> lr(n);
  std::vector<double> ck(n);
  for (int q = 0; q < n; ++q) {
    const int j = keep[q];
    const int l = dens_ids.get(dkey(cmL, K1[j] / ibi, K1[j] % ibi));
    if (l == (int)spec.size()) spec.push_back(make_int4(K1[j] / ibi - 1, K1[j] % ibi - 1, cmL, l));
    lr[q].x = l;
    ck[q] = c[j];
  }
  const int nl = (int)spec.size();
  for (int q = 0; q < n; ++q) {
    const int j = keep[q];
    const int r = dens_ids.get(dkey(cmR, K2[j] / ibi, K2[j] % ibi));
    if (r == (int)spec.size()) spec.push_back(make_int4(K2[j] / ibi - 1, K2[j] % ibi - 1, cmR, r));
    lr[q].y = r;
  }
  // targets: HM/DM positions in order of first appearance, each with its
  // entries in list order (CSR)
  targ_ids.reset((size_t)n);
  std::vector<int> tid(n);
  std::vector<long long> tpos;
  std::vector<int> toff(1, 0);
  for (int q = 0; q < n; ++q) {
    const long long key = pos[q] >= 0 ? 2 * pos[q] : 2 * (-pos[q] - 1) + 1;
    tid[q] = targ_ids.get(key);
    if (tid[q] == (int)tpos.size()) {
      tpos.push_back(pos[q]);
      toff.push_back(0);
    }
    ++toff[tid[q] + 1];
  }
  const int nt = (int)tpos.size();
  for (int t = 0; t < nt; ++t) toff[t + 1] += toff[t];
  std::vector<int> tflat(n), fill(toff.begin(), toff.end() - 1);
  for (int q = 0; q < n; ++q) tflat[fill[tid[q]]++] = q;
  double* rkm = nullptr;
  int rc = build_rkm(ctx, itype, k, &rkm);
  if (rc) return rc;
  // scratch (doubles): densities | convolutions | values | then the host
  // tables (coefficients, spec, row pairs, targets), 16-byte aligned, staged
  // in one of the context's two pinned buffers and uploaded in one copy
  const size_t per = (size_t)ns * nb;
  const size_t o_conv = spec.size() * per, o_vals = o_conv + (size_t)nl * per, o_tab = o_vals + n;
  size_t nblob = 0;
  auto slot = [&](size_t bytes) {
    const size_t off = (nblob + 15) & ~(size_t)15;
    nblob = off + bytes;
    return off;
  };
  const size_t b_c = slot(ck.size() * sizeof(double));
  const size_t b_spec = slot(spec.size() * sizeof(int4));
  const size_t b_lr = slot(lr.size() * sizeof(int2));
  const size_t b_tpos = slot(tpos.size() * sizeof(long long));
  const size_t b_toff = slot(toff.size() * sizeof(int));
  const size_t b_tent = slot(tflat.size() * sizeof(int));
  nblob = (nblob + 15) & ~(size_t)15;
  if ((rc = ensure_scratch(ctx, o_tab + nblob / 8 + 2))) return rc;
  char* blob = nullptr;
  if ((rc = list_stage_begin(ctx, nblob, &blob))) return rc;
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) std::memcpy(blob + off, src, bytes);
  };
  put(b_c, ck.data(), ck.size() * sizeof(double));
  put(b_spec, spec.data(), spec.size() * sizeof(int4));
  put(b_lr, lr.data(), lr.size() * sizeof(int2));
  put(b_tpos, tpos.data(), tpos.size() * sizeof(long long));
  put(b_toff, toff.data(), toff.size() * sizeof(int));
  put(b_tent, tflat.data(), tflat.size() * sizeof(int));
  char* base = reinterpret_cast<char*>(ctx->scratch + o_tab);
  base = reinterpret_cast<char*>(((uintptr_t)base + 15) & ~(uintptr_t)15);
  hipStream_t s = ctx->stream;
  double* dens = ctx->scratch;
  double* conv = ctx->scratch + o_conv;
  double* vals = ctx->scratch + o_vals;
  hipError_t e = hipMemcpyAsync(base, blob, nblob, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && (rc = list_stage_end(ctx))) return rc;
  if (e == hipSuccess) {
    const int nsp = (int)spec.size();
    k_plain_dens<<<grid_for((long long)nsp * per, 256, 8192), 256, 0, s>>>(
        nsp, reinterpret_cast<const int4*>(base + b_spec), ns, ks, sym, pq_dev, 0, dens);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    // conv(:, l) = rkm^T dens(:, l) for the left densities 0..nl-1
    rc = conv_gemm(ctx, 2, (int)per, nl, rkm, dens, conv);
    if (rc) return rc;
    k_plain_rows<<<grid_for(n, 4, 65535), 256, 0, s>>>(n, reinterpret_cast<const int2*>(base + b_lr), ns, ks,
                                                         sym, conv, dens, nullptr, 0, vals);
    k_list_accumulate<<<grid_for(nt, 256, 65535), 256, 0, s>>>(
        nt, reinterpret_cast<const long long*>(base + b_tpos), reinterpret_cast<const int*>(base + b_toff),
        reinterpret_cast<const int*>(base + b_tent), vals, reinterpret_cast<const double*>(base + b_c), HM_dev,
        DM_dev);
    e = hipGetLastError();
  }
  // queued, not waited for: the pinned tables belong to the context
  if (e != hipSuccess) return fail(DBS_ERR_HIP, std::string("Rk_evaluate: ") + hipGetErrorString(e));
  return DBS_OK;
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
    int rc = run_all(e, with_adjoint ? 1 : 0);
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

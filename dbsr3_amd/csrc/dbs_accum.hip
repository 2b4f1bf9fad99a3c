// dbs_accum.hip -- coefficient lists and their accumulation into Hamiltonian
// matrices on MI355X.
//
//  * c_data: blocked per-(kpol,itype) sorted lists (DBSR_MAT3/MOD_c_data.f90).
//    Host integer work with the reference's exact semantics: binary-search
//    insert with equal-key summation, block overflow into new blocks or an
//    early Add_matrix of the fullest type, k-way merge summing equal keys in
//    block order, eps drop, IPT order.
//  * Add_integral (CI3 and MAT3 channel-channel forms), Idef_Rtype, SMU,
//    Decode_int: host integer work.
//  * Rk_evaluate / Sk_evaluate: every list entry's integral on the device
//    (engine), then deterministic accumulation into HM/DM -- contributions
//    to one element are added in list order, as the reference does.
//  * R_data itype 3/4 + UPDATE_HS: per channel pair the c-weighted density
//    (list order), a batched convolution over all channel pairs of the list,
//    and the channel-block update, all on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dbs_device.h"
#include "dbs_internal.h"

// ===========================================================================
// c_data
// ===========================================================================
struct dbs_cdata {
  int ntype = 0, kpol1 = 0, kpol2 = 0, mb = 0, nb = 0, kb = 0;
  double eps = 0.0;
  // storage, 1-based positions (index 0 unused)
  std::vector<double> C;
  std::vector<int> K1, K2, K3, K4;
  std::vector<int> ipblk, jpblk;
  std::vector<int> iblk, nblk;            // (k,type)
  std::vector<std::vector<int>> kblk;     // (k,type) -> block chain
  dbs_gen_matrix_fn gen = nullptr;
  void* user = nullptr;
  int slot(int k, int t) const { return (k - kpol1) + (kpol2 - kpol1 + 1) * (t - 1); }
};

namespace {

inline int cmp_key(const dbs_cdata* cd, int p, int j1, int j2, int j3, int j4) {
  if (j1 != cd->K1[p]) return j1 < cd->K1[p] ? -1 : 1;
  if (j2 != cd->K2[p]) return j2 < cd->K2[p] ? -1 : 1;
  if (j3 != cd->K3[p]) return j3 < cd->K3[p] ? -1 : 1;
  if (j4 != cd->K4[p]) return j4 < cd->K4[p] ? -1 : 1;
  return 0;
}

// insert into sub-list ip..jp (sorted); returns new jp or -1 on overflow
int cdata_insert(dbs_cdata* cd, int ip, int jp, double c, int j1, int j2, int j3, int j4) {
  auto put = [&](int p) {
    cd->C[p] = c; cd->K1[p] = j1; cd->K2[p] = j2; cd->K3[p] = j3; cd->K4[p] = j4;
  };
  if (jp < ip) {
    put(ip);
    return ip;
  }
  int lo = ip, hi = jp;
  while (lo <= hi) {
    const int m = (lo + hi) / 2;
    const int r = cmp_key(cd, m, j1, j2, j3, j4);
    if (r < 0) hi = m - 1;
    else if (r > 0) lo = m + 1;
    else {
      cd->C[m] = cd->C[m] + c;  // the same integral
      return jp;
    }
  }
  if (jp - ip + 1 >= cd->mb) return -1;
  for (int i = jp; i >= lo; --i) {
    cd->C[i + 1] = cd->C[i];
    cd->K1[i + 1] = cd->K1[i]; cd->K2[i + 1] = cd->K2[i];
    cd->K3[i + 1] = cd->K3[i]; cd->K4[i + 1] = cd->K4[i];
  }
  put(lo);
  return jp + 1;
}

// k-way merge of the blocks' heads: equal keys are summed in block order
// (pairs ii<jj), the smallest key is taken next, |C| <= eps dropped.
void cdata_merge(dbs_cdata* cd, std::vector<int>& IP, std::vector<int>& JP, std::vector<int>& ipt) {
  const int nn = (int)IP.size() - 1;
  int mm = 0;
  for (int m = 1; m <= nn; ++m)
    if (JP[m] > 0) { mm = m; break; }
  if (mm == 0) return;
  for (;;) {
    bool merged = false;
    for (int ii = 1; ii <= nn - 1 && !merged; ++ii) {
      if (JP[ii] <= 0) continue;
      const int i = IP[ii];
      for (int jj = ii + 1; jj <= nn; ++jj) {
        if (JP[jj] <= 0) continue;
        const int j = IP[jj];
        if (cmp_key(cd, i, cd->K1[j], cd->K2[j], cd->K3[j], cd->K4[j]) != 0) continue;
        cd->C[i] = cd->C[i] + cd->C[j];
        cd->C[j] = 0.0;
        mm = jj;
        merged = true;
        break;
      }
    }
    if (!merged) {
      int j = IP[mm];
      for (int m = 1; m <= nn; ++m) {
        if (JP[m] == 0) continue;
        const int i = IP[m];
        if (cmp_key(cd, j, cd->K1[i], cd->K2[i], cd->K3[i], cd->K4[i]) < 0) { mm = m; j = i; }
      }
      if (std::fabs(cd->C[j]) > cd->eps) ipt.push_back(IP[mm]);
    }
    IP[mm] += 1;
    if (IP[mm] <= JP[mm]) continue;
    JP[mm] = 0;
    int nxt = 0;
    for (int m = 1; m <= nn; ++m)
      if (JP[m] > 0) { nxt = m; break; }
    if (nxt == 0) break;
    mm = nxt;
  }
}

int cdata_add_matrix(dbs_cdata* cd, int itype, int kpol) {
  const int s = cd->slot(kpol, itype);
  int n = cd->nblk[s];
  const int first = cd->kblk[s][0];
  if (n == 1 && cd->jpblk[first] <= 0) n = 0;
  if (n <= 0) return DBS_OK;
  std::vector<int> ipt;
  if (n == 1) {
    for (int p = cd->ipblk[first]; p <= cd->jpblk[first]; ++p) ipt.push_back(p);
  } else {
    std::vector<int> IP(n + 1, 0), JP(n + 1, 0);
    for (int i = 1; i <= n; ++i) {
      const int b = cd->kblk[s][i - 1];
      IP[i] = cd->ipblk[b];
      JP[i] = cd->jpblk[b];
    }
    cdata_merge(cd, IP, JP, ipt);
    for (int b : cd->kblk[s]) cd->jpblk[b] = -1;
  }
  // re-arm the first block of the type
  cd->iblk[s] = first;
  cd->nblk[s] = 1;
  cd->kblk[s].assign(1, first);
  cd->jpblk[first] = 0;
  const int nc = (int)ipt.size();
  std::vector<double> c(nc);
  std::vector<int> k1(nc), k2(nc), k3(nc), k4(nc);
  for (int q = 0; q < nc; ++q) {
    const int p = ipt[q];
    c[q] = cd->C[p]; k1[q] = cd->K1[p]; k2[q] = cd->K2[p]; k3[q] = cd->K3[p]; k4[q] = cd->K4[p];
  }
  if (cd->gen) cd->gen(cd->user, itype, kpol, nc, c.data(), k1.data(), k2.data(), k3.data(), k4.data());
  return DBS_OK;
}

int cdata_new_block(dbs_cdata* cd, int s) {
  for (int i = 1; i <= cd->nb; ++i) {
    if (cd->jpblk[i] >= 0) continue;
    cd->iblk[s] = i;
    cd->jpblk[i] = 0;
    cd->nblk[s] += 1;
    cd->kblk[s].push_back(i);
    return 1;
  }
  return 0;
}

inline int l_kappa(int kappa) { return kappa > 0 ? kappa : -kappa - 1; }

// Breit factors (Grant and Pyper), the arithmetic of SMU
double smu_factors(int ka, int kb, int kc, int kd, int L, int v, double* S) {
  double smu = 0.0;
  for (int i = 0; i < 8; ++i) S[i] = 0.0;
  if (L < 0 || v < 0) return 0.0;
  if ((l_kappa(ka) + l_kappa(kc) + v) % 2 == 0) return 0.0;
  if ((l_kappa(kb) + l_kappa(kd) + v) % 2 == 0) return 0.0;
  if (v == L) {
    if (L == 0) return 0.0;
    smu = (double)(-(ka + kc) * (kb + kd));
    smu = smu / L / (L + 1);
    for (int i = 0; i < 8; ++i) S[i] = smu;
    return smu;
  }
  const int K = kc - ka, Kp = kd - kb;
  if (v == L - 1) {
    double bp = L + 1;
    bp = bp / 2 / (L + L - 1);
    double cp = -(L - 2);
    cp = cp / 2 / L / (L + L - 1);
    S[0] = (L + K) * (bp + cp * Kp);
    S[1] = (L + Kp) * (bp + cp * K);
    S[2] = (L - K) * (bp - cp * Kp);
    S[3] = (L - Kp) * (bp - cp * K);
    S[4] = -(L + K) * (bp - cp * Kp);
    S[5] = -(L - Kp) * (bp + cp * K);
    S[6] = -(L - K) * (bp + cp * Kp);
    S[7] = -(L + Kp) * (bp - cp * K);
  } else if (v == L + 1) {
    const int i = -L - 1;
    double b = L;
    b = b / 2 / (L + L + 3);
    double c = (L + 3);
    c = c / (L + L + 2) / (L + L + 3);
    S[0] = (i + Kp) * (b + c * K);
    S[1] = (i + K) * (b + c * Kp);
    S[2] = (i - Kp) * (b - c * K);
    S[3] = (i - K) * (b - c * Kp);
    S[4] = -(i - Kp) * (b + c * K);
    S[5] = -(i + K) * (b - c * Kp);
    S[6] = -(i + Kp) * (b - c * K);
    S[7] = -(i - K) * (b + c * Kp);
  }
  for (int q = 0; q < 8; ++q) smu = smu + std::fabs(S[q]);
  return smu;
}

}  // namespace

// ===========================================================================
// device kernels for the accumulation
// ===========================================================================
namespace dbs {

// xx_g = sum_e c_e * density(pq(jp1,j1), pq(jp2,j2)) in list order;
// sym 0: band 's' (ns,ks); 1: full 'x' (ns,ns); one group after another
__global__ void k_group_density(int ng, const int* __restrict__ goff, const int* __restrict__ ej1,
                                const int* __restrict__ ej2, const double* __restrict__ ec,
                                const double* __restrict__ pq, int ns, int ks, int jp1, int jp2,
                                int sym, double* __restrict__ xx) {
#pragma clang fp contract(off)
  const int ncol = sym == 0 ? ks : ns;
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
      double d;
      if (sym == 0) {
        if (j == 0) d = p1[i] * p2[i];
        else if (i < ns - j) d = p1[i] * p2[i + j] + p1[i + j] * p2[i];
        else d = 0.0;
      } else {
        d = p1[i] * p2[j];
      }
      acc = acc + ec[e] * d;
    }
    xx[q] = acc;
  }
}

// the full 'x' densities of the groups, xx_g = sum_e (c_e p1_e) p2_e^T,
// as register-tiled rank-k updates: one workgroup per 64x64 tile of one
// group, the group's entries staged DX_E at a time in LDS (c_e p1_e over
// the tile's rows, p2_e over its columns), each thread 4x4 outputs (rows
// tx + 16a, columns ty + 16b) summed over the entries in list order
// (fused multiply-adds); grid = ng * nt * nt, nt = ceil(ns / 64)
constexpr int DX_T = 64, DX_E = 16;

__global__ __launch_bounds__(256) void k_group_density_x(int ng, int nt, const int* __restrict__ goff,
                                                         const int* __restrict__ ej1, const int* __restrict__ ej2,
                                                         const double* __restrict__ ec,
                                                         const double* __restrict__ pq, int ns, int jp1, int jp2,
                                                         double* __restrict__ xx) {
  __shared__ double A[DX_E][DX_T];
  __shared__ double B[DX_E][DX_T];
  const int nb = nt * nt;
  const int g = (int)blockIdx.x / nb, tile = (int)blockIdx.x % nb;
  if (g >= ng) return;
  const int i0 = (tile % nt) * DX_T, j0 = (tile / nt) * DX_T;
  const int tid = (int)threadIdx.x, tx = tid % 16, ty = tid / 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  const int e0 = goff[g], e1 = goff[g + 1];
  for (int eb = e0; eb < e1; eb += DX_E) {
    const int ne = min(DX_E, e1 - eb);
    __syncthreads();
    for (int q = tid; q < DX_E * DX_T; q += 256) {
      const int e = q / DX_T, r = q % DX_T;
      double av = 0.0, bv = 0.0;
      if (e < ne) {
        const int ee = eb + e;
        if (i0 + r < ns) av = ec[ee] * pq[(long long)ns * ((jp1 - 1) + 2LL * (ej1[ee] - 1)) + i0 + r];
        if (j0 + r < ns) bv = pq[(long long)ns * ((jp2 - 1) + 2LL * (ej2[ee] - 1)) + j0 + r];
      }
      A[e][r] = av;
      B[e][r] = bv;
    }
    __syncthreads();
    for (int e = 0; e < ne; ++e) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = A[e][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = B[e][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
  }
  double* out = xx + (long long)g * ns * ns;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = j0 + ty + 16 * b;
    if (j >= ns) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = i0 + tx + 16 * a;
      if (i < ns) out[i + (long long)ns * j] = acc[a][b];
    }
  }
}

// the outputs of convol case 1 with sym 's' outside the band (i >= ns-ip)
// set to zero, as the reference's convol leaves them
__global__ void k_band_zero(int ng, int ns, int ks, double* __restrict__ a) {
  const long long per = (long long)ns * ks, total = per * ng;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(q % per), i = r % ns, ip = r / ns;
    if (i >= ns - ip) a[q] = 0.0;
  }
}

// tiled convol case 3: one workgroup computes a CV_TX x CV_TY block of
// outputs a(x,y) for GB groups, with those groups' densities over the block
// and its ks-1 halo staged in LDS, so the 4*ks*ks terms of each output read
// LDS instead of global memory and each rkb value feeds GB products. Block
// order is XCD-aware: the dispatcher deals workgroups to the 8 XCDs round
// robin, so logical block (xcd, w) runs one tile's group chunks back to back
// on one XCD and that tile's rkb slice stays in the XCD's L2. Terms are
// summed in the same order as k_group_convol3. Needs ks <= CV_KS and a grid
// that is a multiple of 8.
constexpr int CV_TX = 32, CV_TY = 8, CV_KS = 9, CV_H = CV_KS - 1;
constexpr int CV_SX = CV_TX + 2 * CV_H, CV_SY = CV_TY + 2 * CV_H;

template <int GB>
__global__ __launch_bounds__(CV_TX* CV_TY) void k_group_convol3_tile(int ng, int ns, int ks, int ntx, int nty,
                                                                     const double* __restrict__ rkb,
                                                                     const double* __restrict__ xx,
                                                                     double* __restrict__ out) {
  __shared__ double L[GB][CV_SY][CV_SX];
  const int nchunk = (ng + GB - 1) / GB;
  const int lid = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (lid >= ntx * nty * nchunk) return;  // whole workgroup: before any barrier
  const int tile = lid / nchunk, chunk = lid % nchunk;
  const int x0 = (tile % ntx) * CV_TX, y0 = (tile / ntx) * CV_TY;
  const int g0 = chunk * GB, nb = min(GB, ng - g0);
  const long long per = (long long)ns * ns;
  for (int e = threadIdx.x; e < GB * CV_SY * CV_SX; e += CV_TX * CV_TY) {
    const int b = e / (CV_SY * CV_SX), rem = e % (CV_SY * CV_SX);
    const int u = x0 - CV_H + rem % CV_SX, v = y0 - CV_H + rem / CV_SX;
    double val = 0.0;
    if (b < nb && u >= 0 && u < ns && v >= 0 && v < ns) val = xx[(long long)(g0 + b) * per + u + (long long)ns * v];
    (&L[0][0][0])[e] = val;
  }
  __syncthreads();
  const int tx = (int)threadIdx.x % CV_TX, ty = (int)threadIdx.x / CV_TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= ns || y >= ns) return;  // no barrier follows
  double acc[GB];
#pragma unroll
  for (int b = 0; b < GB; ++b) acc[b] = 0.0;
  // terms: (xs,ys,cond): (0,1,-),(1,1,ip>0),(0,0,jp>0),(1,0,ip>0&&jp>0)
  for (int t = 0; t < 4; ++t) {
    const int xs = (t == 1 || t == 3), ys = (t == 0 || t == 1);
    for (int jp = 0; jp < ks; ++jp) {
      if (t >= 2 && jp == 0) continue;
      const int j = ys == 0 ? y : y - jp;
      if (j < 0 || j >= ns - jp) continue;
      const int vl = CV_H + ty + (ys == 0 ? jp : -jp);  // LDS row of v
      const double* rr = rkb + (long long)ns * (j + (long long)ns * ks * jp);
      double rv[CV_KS];
      bool ok[CV_KS];
#pragma unroll
      for (int ip = 0; ip < CV_KS; ++ip) {
        const int i = xs == 0 ? x : x - ip;
        ok[ip] = ip < ks && !(xs == 1 && ip == 0) && i >= 0 && i < ns - ip;
        rv[ip] = ok[ip] ? rr[i + per * ip] : 0.0;
      }
#pragma unroll
      for (int ip = 0; ip < CV_KS; ++ip) {
        if (!ok[ip]) continue;
        const int ul = CV_H + tx + (xs == 0 ? ip : -ip);  // LDS column of u
#pragma unroll
        for (int b = 0; b < GB; ++b) acc[b] += rv[ip] * L[b][vl][ul];
      }
    }
  }
#pragma unroll
  for (int b = 0; b < GB; ++b)
    if (b < nb) out[(long long)(g0 + b) * per + x + (long long)ns * y] = acc[b];
}

// convol case 3 ('s','s') for many groups, gathered per output element:
// a(x,y) collects rkb(i,j,ip,jp)*d(u,v) over the four symmetric-storage
// terms of the reference (convol, dbsr_lib_dbs.f90:2344-2357); GB groups
// per thread so each rkb value and its index arithmetic feed GB products
template <int GB>
__global__ void k_group_convol3(int ng, int ns, int ks, const double* __restrict__ rkb,
                                const double* __restrict__ xx, double* __restrict__ out) {
  const long long per = (long long)ns * ns;
  const int ngc = (ng + GB - 1) / GB;
  const long long total = per * ngc;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int g0 = (int)(q / per) * GB;
    const int r = (int)(q % per);
    const int x = r % ns, y = r / ns;  // 0-based output a(x,y)
    const double* d = xx + (long long)g0 * per;
    const int nb = min(GB, ng - g0);    // groups of this thread
    double acc[GB];
#pragma unroll
    for (int b = 0; b < GB; ++b) acc[b] = 0.0;
    // terms: (xs,ys,cond): (0,1,-),(1,1,ip>0),(0,0,jp>0),(1,0,ip>0&&jp>0)
    for (int t = 0; t < 4; ++t) {
      const int xs = (t == 1 || t == 3), ys = (t == 0 || t == 1);
      for (int jp = 0; jp < ks; ++jp) {
        if ((t == 2 || t == 3) && jp == 0) continue;
        const int j = ys == 0 ? y : y - jp;
        if (j < 0 || j >= ns - jp) continue;
        const int jj = j + jp;
        const int v = ys == 0 ? jj : j;
        for (int ip = 0; ip < ks; ++ip) {
          if ((t == 1 || t == 3) && ip == 0) continue;
          const int i = xs == 0 ? x : x - ip;
          if (i < 0 || i >= ns - ip) continue;
          const int ii = i + ip;
          const int u = xs == 0 ? ii : i;
          // one rkb value feeds the GB groups' products
          const double rv = rkb[i + (long long)ns * (j + (long long)ns * (ip + (long long)ks * jp))];
          const double* du = d + u + (long long)ns * v;
#pragma unroll
          for (int b = 0; b < GB; ++b)
            if (b < nb) acc[b] += rv * du[(long long)b * per];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < GB; ++b)
      if (b < nb) out[(long long)(g0 + b) * per + r] = acc[b];
  }
}

// UPDATE_HS, MAT3 form (DBSR_MAT3/updates.f90:110-179): y = unfold(d);
// ich>jch: h += y; jch>ich: h += y^T; ich==jch: pass 0 h(i1..,j1..) += y/2,
// pass 1 h(j1..,i1..) += (y/2)^T.
__global__ void k_update_hs_mat(int ng, const int4* __restrict__ gi, int ns, int ks, int sym,
                                int ii, int jj, int pass, const double* __restrict__ d,
                                double* __restrict__ hch) {
#pragma clang fp contract(off)
  const long long per = (long long)ns * ns;
  const long long total = per * ng;
  const int ms = 2 * ns;
  const int ncol = sym == 0 ? ks : ns;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int g = (int)(q / per);
    const int r = (int)(q % per);
    const int a = r % ns, b = r / ns;
    const int4 G = gi[g];  // (ich, jch, block0, 0)
    const int ich = G.x, jch = G.y;
    if (pass == 1 && ich != jch) continue;
    const double* dg = d + (long long)g * ns * ncol;
    auto yv = [&](int u, int v) -> double {
      if (sym == 1) return dg[u + (long long)ns * v];
      const int lo = u < v ? u : v, hi = u < v ? v : u;
      return hi - lo < ks ? dg[lo + (long long)ns * (hi - lo)] : 0.0;
    };
    double* h = hch + (long long)G.z * ms * ms;
    const int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
    if (ich > jch) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += yv(a, b);
    } else if (jch > ich) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += yv(b, a);
    } else if (pass == 0) {
      h[(i1 + a) + (long long)ms * (j1 + b)] += yv(a, b) / 2.0;
    } else {
      h[(j1 + a) + (long long)ms * (i1 + b)] += yv(b, a) / 2.0;
    }
  }
}

// UPDATE_HS for full 'x' convolutions (sym 1), both passes in one launch:
// one workgroup per 32x32 tile (a, b) of one group's output block, y over
// the tile staged in LDS as read (Td: y(a,b)) and, when the update needs
// it, transposed (Tt: y(b,a)), so global reads and writes are coalesced.
// Per element the reference's order: ich > jch h += y(a,b); jch > ich
// h += y(b,a); ich == jch pass 0 h(i1+a,j1+b) += y(a,b)/2 then pass 1
// h(j1+a,i1+b) += y(b,a)/2 (one element when i1 == j1, added in that
// order by one thread).  grid = ng * nt * nt, nt = ceil(ns / 32).
__global__ __launch_bounds__(256) void k_update_hs_x(int ng, int nt, const int4* __restrict__ gi, int ns, int ii,
                                                     int jj, const double* __restrict__ d,
                                                     double* __restrict__ hch) {
#pragma clang fp contract(off)
  __shared__ double Td[32][33];
  __shared__ double Tt[32][33];
  const int nb = nt * nt;
  const int g = (int)blockIdx.x / nb, tile = (int)blockIdx.x % nb;
  if (g >= ng) return;
  const int a0 = (tile % nt) * 32, b0 = (tile / nt) * 32;
  const int tx = (int)threadIdx.x % 32, ty = (int)threadIdx.x / 32;
  const int4 G = gi[g];  // (ich, jch, block0, 0)
  const int ich = G.x, jch = G.y;
  const double* dg = d + (long long)g * ns * ns;
  for (int m = ty; m < 32; m += 8) {
    // Td[bl][al] = y(a0+al, b0+bl); Tt[al][bl] = y(b0+bl, a0+al)
    const int a = a0 + tx, b = b0 + m;
    Td[m][tx] = (a < ns && b < ns) ? dg[a + (long long)ns * b] : 0.0;
    const int bt = b0 + tx, at = a0 + m;
    Tt[m][tx] = (ich <= jch && bt < ns && at < ns) ? dg[bt + (long long)ns * at] : 0.0;
  }
  __syncthreads();
  const int ms = 2 * ns;
  double* h = hch + (long long)G.z * ms * ms;
  const int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  const int a = a0 + tx;
  if (a >= ns) return;
  for (int m = ty; m < 32; m += 8) {
    const int b = b0 + m;
    if (b >= ns) break;
    const double yab = Td[m][tx], yba = Tt[tx][m];
    double* p = h + (i1 + a) + (long long)ms * (j1 + b);
    if (ich > jch) {
      *p += yab;
    } else if (jch > ich) {
      *p += yba;
    } else if (i1 == j1) {
      double v = *p;
      v = v + yab / 2.0;
      v = v + yba / 2.0;
      *p = v;
    } else {
      *p += yab / 2.0;
      h[(j1 + a) + (long long)ms * (i1 + b)] += yba / 2.0;
    }
  }
}

}  // namespace dbs

using namespace dbs;

static unsigned grid_of(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (unsigned)g;
}

extern "C" {

int dbs_cdata_create(dbs_cdata** out, int ntype, int kpol1, int kpol2, int mblock, int nblock,
                     int kblock, double eps_c, dbs_gen_matrix_fn gen, void* user) {
  DBS_CHECK(out, DBS_ERR_ARG, "dbs_cdata_create: null argument");
  DBS_CHECK(ntype >= 1 && kpol2 >= kpol1 && mblock >= 1 && kblock >= 1, DBS_ERR_ARG,
            "dbs_cdata_create: bad dimensions");
  dbs_cdata* cd = new dbs_cdata();
  cd->ntype = ntype; cd->kpol1 = kpol1; cd->kpol2 = kpol2;
  cd->mb = mblock; cd->nb = nblock; cd->kb = kblock;
  const int need = ntype * (kpol2 - kpol1 + 1) * 2 + 1;
  if (cd->nb < need) cd->nb = need;
  cd->eps = eps_c;
  cd->gen = gen;
  cd->user = user;
  const size_t m = (size_t)cd->mb * cd->nb;
  cd->C.assign(m + 2, 0.0);
  cd->K1.assign(m + 2, 0); cd->K2.assign(m + 2, 0); cd->K3.assign(m + 2, 0); cd->K4.assign(m + 2, 0);
  cd->ipblk.assign(cd->nb + 1, 0);
  cd->jpblk.assign(cd->nb + 1, -1);
  for (int i = 1; i <= cd->nb; ++i) cd->ipblk[i] = (i - 1) * cd->mb + 1;
  const int ns = (kpol2 - kpol1 + 1) * ntype;
  cd->iblk.assign(ns, 0);
  cd->nblk.assign(ns, 0);
  cd->kblk.assign(ns, {});
  int i = 0;
  for (int k = kpol1; k <= kpol2; ++k)
    for (int t = 1; t <= ntype; ++t) {
      ++i;
      const int s = cd->slot(k, t);
      cd->iblk[s] = i;
      cd->nblk[s] = 1;
      cd->kblk[s].assign(1, i);
      cd->jpblk[i] = 0;
    }
  *out = cd;
  return DBS_OK;
}

int dbs_cdata_destroy(dbs_cdata* cd) {
  delete cd;
  return DBS_OK;
}

int dbs_cdata_add_coef(dbs_cdata* cd, double c, int kpol, int j1, int j2, int j3, int j4, int itype) {
  DBS_CHECK(cd, DBS_ERR_ARG, "dbs_cdata_add_coef: null list");
  DBS_CHECK(kpol >= cd->kpol1 && kpol <= cd->kpol2 && itype >= 1 && itype <= cd->ntype, DBS_ERR_ARG,
            "Add_coef: kpol or itype out of range");
  const int s = cd->slot(kpol, itype);
  int i = cd->iblk[s];
  const int ip = cd->ipblk[i];
  const int jp = cdata_insert(cd, ip, cd->jpblk[i], c, j1, j2, j3, j4);
  DBS_CHECK(jp >= 0, DBS_ERR_CAPACITY, "Promlems with jp in Add_cdata");
  cd->jpblk[i] = jp;
  if (jp - ip + 1 < cd->mb) return DBS_OK;
  if (cd->nblk[s] < cd->kb && cdata_new_block(cd, s)) return DBS_OK;
  // everything is full: generate the matrix of the fullest type
  int jtype = itype, jpol = kpol, n = cd->nblk[s];
  for (int t = 1; t <= cd->ntype; ++t)
    for (int k = cd->kpol1; k <= cd->kpol2; ++k) {
      const int q = cd->slot(k, t);
      if (cd->nblk[q] <= n) continue;
      jtype = t; jpol = k; n = cd->nblk[q];
    }
  int rc = cdata_add_matrix(cd, jtype, jpol);
  if (rc) return rc;
  if (jtype == itype && jpol == kpol) return DBS_OK;
  DBS_CHECK(cdata_new_block(cd, s), DBS_ERR_CAPACITY, " Add_coef: problems with new block ");
  return DBS_OK;
}

int dbs_cdata_add_matrix(dbs_cdata* cd, int itype, int kpol) {
  DBS_CHECK(cd, DBS_ERR_ARG, "dbs_cdata_add_matrix: null list");
  DBS_CHECK(kpol >= cd->kpol1 && kpol <= cd->kpol2 && itype >= 1 && itype <= cd->ntype, DBS_ERR_ARG,
            "Add_matrix: kpol or itype out of range");
  return cdata_add_matrix(cd, itype, kpol);
}

int dbs_smu(int ka, int kb, int kc, int kd, int L, int v, double* S, double* smu) {
  DBS_CHECK(S && smu, DBS_ERR_ARG, "dbs_smu: null argument");
  DBS_CHECK(ka != 0 && kb != 0 && kc != 0 && kd != 0, DBS_ERR_ARG, "l_kappa: kappa=0");
  *smu = smu_factors(ka, kb, kc, kd, L, v, S);
  return DBS_OK;
}

int dbs_add_integral_ci(dbs_cdata* cd, int kpol, int ic, int jc, int i1, int i2, int i3, int i4,
                        double c, const int* kappa, int ibi, int mbreit) {
  DBS_CHECK(cd && kappa, DBS_ERR_ARG, "Add_integral: null argument");
  int rc;
  const int m = (l_kappa(kappa[i1]) + l_kappa(kappa[i3]) + kpol) % 2 +
                (l_kappa(kappa[i2]) + l_kappa(kappa[i4]) + kpol) % 2;
  if (m == 0)
    for (int t = 1; t <= 4; ++t)
      if ((rc = dbs_cdata_add_coef(cd, c, kpol, i1 * ibi + i3, i2 * ibi + i4, ic, jc, t))) return rc;
  if (mbreit == 0) return DBS_OK;
  double S[8];
  for (int v = kpol - 1; v <= kpol + 1; ++v) {
    if (smu_factors(kappa[i1], kappa[i2], kappa[i3], kappa[i4], kpol, v, S) == 0.0) continue;
    const int a = i1 * ibi + i3, b = i2 * ibi + i4, a2 = i3 * ibi + i1, b2 = i4 * ibi + i2;
    const int keys[8][3] = {{a, b, 5}, {b, a, 5}, {a2, b2, 5}, {b2, a2, 5},
                            {a, b, 6}, {b2, a2, 6}, {a2, b2, 6}, {b, a, 6}};
    for (int q = 0; q < 8; ++q)
      if ((rc = dbs_cdata_add_coef(cd, c * S[q], v, keys[q][0], keys[q][1], ic, jc, keys[q][2]))) return rc;
  }
  return DBS_OK;
}

// MAT3 coefficient routing (DBSR_MAT3/idef_type.f90).  ipbs(o) is read as
// ipb(o): the channel of orbital o (1-based), 0 for a bound orbital.
}  // extern "C"

namespace {

// Idef_Rtype (idef_type.f90:141-302): canonical orbital order, structure
// 1..4 (integral x bound/channel factor, channel vectors, direct and
// exchange channel-channel) and the c_data key.  Returns the itype, 0 when
// the entry is dropped (|C| < eps_C, or a target structure without
// check_target), or a negative status after fail().
template <class IPB>
int idef_rtype(const IPB& ipb, int ibo, double eps_c, int check_target, int j1, int j2, int j3, int j4, double C,
               int ic, int jc, int io, int jo, int* key) {
  if (std::fabs(C) < eps_c) return 0;
  if (io != 0 && std::abs(io) <= ibo) return -fail(DBS_ERR_ARG, "Idef_Rtype: io<ibo");
  if (jo != 0 && std::abs(jo) <= ibo) return -fail(DBS_ERR_ARG, "Idef_Rtype: jo<ibo");
  if (jc > 0 && ic <= 0) return -fail(DBS_ERR_ARG, "jc>0, ic<=0");
  int i1, i2, i3, i4;
  if (ipb(j1) + ipb(j2) == 0 || ipb(j3) + ipb(j4) == 0) {
    i1 = j1; i3 = j3;
    if (j1 > j3) { i1 = j3; i3 = j1; }
    i2 = j2; i4 = j4;
    if (j2 > j4) { i2 = j4; i4 = j2; }
    if (i1 > i2) { std::swap(i1, i2); std::swap(i3, i4); }
  } else {
    i1 = j1; i2 = j2; i3 = j3; i4 = j4;
  }
  int k1 = 0, k2 = 0, k3 = 0, k4 = 0, itype = 0;
  auto chan_vec = [&](int last) {  // Rk with one channel orbital: vectors (2.x)
    if (ipb(i1) > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipb(i1); k4 = last; itype = 2; }
    else if (ipb(i2) > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipb(i2); k4 = last; itype = 2; }
    else if (ipb(i3) > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipb(i3); k4 = last; itype = 2; }
    else if (ipb(i4) > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipb(i4); k4 = last; itype = 2; }
  };
  if (jc > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = -ic; k4 = -jc; itype = 1;
  } else if (ic > 0) {
    if (io == 0) chan_vec(-ic);
    else if (io > 0) { k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = -ic; itype = 1; }
  } else if (io == 0) {
    if (ipb(i1) > 0 && ipb(i3) > 0) {
      itype = 3;
      if (ipb(i1) >= ipb(i3)) { k1 = ipb(i1) * ibo + ipb(i3); k2 = i2; k3 = i4; if (k2 > k3) { k2 = i4; k3 = i2; } }
      else { k1 = ipb(i3) * ibo + ipb(i1); k2 = i4; k3 = i2; if (k2 > k3) { k2 = i2; k3 = i4; } }
    } else if (ipb(i2) > 0 && ipb(i4) > 0) {
      itype = 3;
      if (ipb(i2) >= ipb(i4)) { k1 = ipb(i2) * ibo + ipb(i4); k2 = i1; k3 = i3; if (k2 > k3) { k2 = i3; k3 = i1; } }
      else { k1 = ipb(i4) * ibo + ipb(i2); k2 = i3; k3 = i1; if (k2 > k3) { k2 = i1; k3 = i3; } }
    } else if (ipb(i1) > 0 && ipb(i4) > 0) {
      itype = 4;
      if (ipb(i1) >= ipb(i4)) { k1 = ipb(i1) * ibo + ipb(i4); k2 = i3; k3 = i2; }
      else { k1 = ipb(i4) * ibo + ipb(i1); k2 = i2; k3 = i3; }
    } else if (ipb(i2) > 0 && ipb(i3) > 0) {
      itype = 4;
      if (ipb(i3) >= ipb(i2)) { k1 = ipb(i3) * ibo + ipb(i2); k2 = i1; k3 = i4; }
      else { k1 = ipb(i2) * ibo + ipb(i3); k2 = i4; k3 = i1; }
    }
  } else if (io > 0 && jo == 0) {
    chan_vec(io);
    if (itype == 0) {
      k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = 0; itype = 1;
      if (check_target == 0) return 0;
    }
  } else if (io > 0 && jo > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = jo; itype = 1;
  }
  if (itype == 0) return -fail(DBS_ERR_ARG, "Idef_Rtype: itype=0");
  key[0] = k1; key[1] = k2; key[2] = k3; key[3] = k4;
  return itype;
}

// Idef_itype (idef_type.f90:306-437): the S^k structures 1..9, no orbital
// canonicalisation; itype + atype*ibtype
template <class IPB>
int idef_itype(const IPB& ipb, int ibo, int atype, int i1, int i2, int i3, int i4, double C, int ic, int jc,
               int io, int jo, int* key) {
  if (C == 0.0) return 0;
  if (io != 0 && std::abs(io) <= ibo) return -fail(DBS_ERR_ARG, "Idef_type: io,ibo = " + std::to_string(io) + " " + std::to_string(ibo));
  if (jo != 0 && std::abs(jo) <= ibo) return -fail(DBS_ERR_ARG, "Idef_type: jo,ibo = " + std::to_string(jo) + " " + std::to_string(ibo));
  if (jc > 0 && ic <= 0) return -fail(DBS_ERR_ARG, "jc>0,ic<=0");
  int k1 = 0, k2 = 0, k3 = 0, k4 = 0, itype = 0;
  auto chan_vec = [&](int last) {
    if (ipb(i1) > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipb(i1); k4 = last; itype = 2; }
    else if (ipb(i2) > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipb(i2); k4 = last; itype = 3; }
    else if (ipb(i3) > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipb(i3); k4 = last; itype = 4; }
    else if (ipb(i4) > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipb(i4); k4 = last; itype = 5; }
  };
  if (jc > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = -ic; k4 = -jc; itype = 1;
  } else if (ic > 0) {
    if (io == 0) chan_vec(-ic);
    else if (io > 0) { k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = -ic; itype = 1; }
  } else if (io == 0) {
    if (ipb(i1) > 0 && ipb(i3) > 0) { k1 = ipb(i1) * ibo + ipb(i3); k2 = i2; k3 = i4; itype = 6; }
    else if (ipb(i2) > 0 && ipb(i4) > 0) { k1 = ipb(i2) * ibo + ipb(i4); k2 = i1; k3 = i3; itype = 7; }
    else if (ipb(i1) > 0 && ipb(i4) > 0) { k1 = ipb(i1) * ibo + ipb(i4); k2 = i3; k3 = i2; itype = 8; }
    else if (ipb(i2) > 0 && ipb(i3) > 0) { k1 = ipb(i3) * ibo + ipb(i2); k2 = i1; k3 = i4; itype = 9; }
  } else if (io > 0 && jo == 0) {
    chan_vec(io);
    if (itype == 0) { k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = 0; itype = 1; }
  } else if (io > 0 && jo > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = jo; itype = 1;
  }
  if (itype == 0) return -fail(DBS_ERR_ARG, "IDEF_TYPE: itype=0");
  key[0] = k1; key[1] = k2; key[2] = k3; key[3] = k4;
  return atype * 10 + itype;  // ibtype = 10 (mod_dbsr_mat.f90:103)
}

// Idef_Otype (idef_type.f90:2-50): overlap structures 1..4
int idef_otype(double eps_c, int check_target, double C, int ic, int jc, int io, int jo, int* key) {
  if (std::fabs(C) < eps_c) return 0;
  if (jc > 0 && ic <= 0) return -fail(DBS_ERR_ARG, "jc > 0, ic <= 0");
  int itype = 0;
  key[2] = key[3] = 0;
  if (jc > 0) { key[0] = ic; key[1] = jc; itype = 1; }
  else if (ic > 0 && io > 0) { key[0] = io; key[1] = ic; itype = 2; }
  else if (io > 0 && jo > 0) { key[0] = io; key[1] = jo; itype = 3; }
  else if (io > 0 && jo == 0) { key[0] = io; key[1] = 0; itype = 4; }
  if (itype == 0) return -fail(DBS_ERR_ARG, "Idef_Otype: itype=0");
  if (check_target == 0 && itype == 4) return 0;
  return itype;
}

// Idef_Ltype (idef_type.f90:54-138): one-electron structures 1..3 with
// key (k1,k2,k3,k4); 0 = dropped (|C| < eps_C, or a target term with
// check_target = 0)
template <class IPB>
int idef_ltype(const IPB& ipb, double eps_c, int check_target, int ii, int jj, double C, int ic, int jc, int io,
               int jo, int* key) {
  if (std::fabs(C) < eps_c) return 0;
  if (jc > 0 && ic <= 0) return -fail(DBS_ERR_ARG, "jc>0, ic<=0");
  if (jo > 0 && io <= 0) return -fail(DBS_ERR_ARG, "jo>0, io<=0");
  if (io > 0 && jc != 0) return -fail(DBS_ERR_ARG, "jc>0, io>0");
  if (jo > 0 && ic != 0) return -fail(DBS_ERR_ARG, "jc>0, jo>0");
  int itype = 0, k1 = 0, k2 = 0, k3 = 0, k4 = 0;
  const int i = std::min(ii, jj), j = std::max(ii, jj), ich = ipb(i), jch = ipb(j);
  if (jc > 0) {
    k1 = i; k2 = j; k3 = ic; k4 = jc; itype = 1;
  } else if (ic > 0) {
    if (io == 0) {
      if (ich != 0 && jch == 0) { k1 = j; k2 = ich; k3 = ic; k4 = 0; itype = 2; }
      else if (ich == 0 && jch != 0) { k1 = i; k2 = jch; k3 = ic; k4 = 0; itype = 2; }
    } else if (io > 0) {
      k1 = i; k2 = j; k3 = -io; k4 = ic; itype = 1;
    }
  } else if (io == 0) {
    k1 = j; k2 = i; itype = 3;
  } else if (io > 0 && jo == 0) {
    if (ich != 0 && jch == 0) { k1 = j; k2 = ich; k3 = -io; itype = 2; }
    else if (jch != 0 && ich == 0) { k1 = i; k2 = jch; k3 = -io; itype = 2; }
    else if (ich == 0 && jch == 0) {
      k1 = i; k2 = j; k3 = -io; itype = 1;
      if (check_target == 0) return 0;
    }
  } else if (io > 0 && jo > 0) {
    k1 = i; k2 = j; k3 = -io; k4 = -jo; itype = 1;
  }
  if (itype == 0) return -fail(DBS_ERR_ARG, "Idef_Ltype: itype=0");
  key[0] = k1; key[1] = k2; key[2] = k3; key[3] = k4;
  return itype;
}

}  // namespace

extern "C" {

int dbs_add_integral_mat(dbs_cdata* cd, int kpol, int j1, int j2, int j3, int j4, double c,
                         const int* lorb, const int* ipbs, int ibo, int mk, double eps_c) {
  DBS_CHECK(cd && lorb && ipbs, DBS_ERR_ARG, "Add_integral: null argument");
  const int m = (lorb[j1] + lorb[j3] + kpol) % 2 + (lorb[j2] + lorb[j4] + kpol) % 2;
  if (m != 0 || kpol > mk) return DBS_OK;
  auto ipb = [&](int o) { return ipbs[o]; };
  int key[4];
  const int t = idef_rtype(ipb, ibo, eps_c, 1, j1, j2, j3, j4, c, 0, 0, 0, 0, key);
  if (t < 0) return -t;
  if (t == 0) return DBS_OK;
  return dbs_cdata_add_coef(cd, c, kpol, key[0], key[1], key[2], key[3], t);
}

int dbs_add_integral_mat3(dbs_cdata* cd, const dbs_mat3_orbs* o, int icase, int kpol, int i1, int i2, int i3,
                          int i4, double C, int ic, int jc, int io, int jo) {
  DBS_CHECK(cd && o && o->ipbs, DBS_ERR_ARG, "Add_integral: null argument");
  auto ipb = [&](int orb) { return (orb >= 1 && orb <= o->norb) ? o->ipbs[orb - 1] : 0; };
  int key[4], t, rc;
  switch (icase) {
    case 0:  // overlaps
      t = idef_otype(o->eps_c, o->check_target, C, ic, jc, io, jo, key);
      if (t <= 0) return -t;
      return dbs_cdata_add_coef(cd, C, 0, key[0], key[1], key[2], key[3], t);
    case 1:  // L-integrals (add_integral.f90:24-26): Idef_Ltype of (i1, i3)
      for (int q : {i1, i3}) DBS_CHECK(q >= 1 && q <= o->norb, DBS_ERR_ARG, "Add_integral: orbital index out of range");
      t = idef_ltype(ipb, o->eps_c, o->check_target, i1, i3, C, ic, jc, io, jo, key);
      if (t <= 0) return -t;
      return dbs_cdata_add_coef(cd, C, 0, key[0], key[1], key[2], key[3], t);
    case 2: {  // R-integrals: the parity rule, then Idef_Rtype
      DBS_CHECK(o->lbs, DBS_ERR_ARG, "Add_integral: lbs missing");
      for (int q : {i1, i2, i3, i4}) DBS_CHECK(q >= 1 && q <= o->norb, DBS_ERR_ARG, "Add_integral: orbital index out of range");
      const int m = (o->lbs[i1 - 1] + o->lbs[i3 - 1] + kpol) % 2 + (o->lbs[i2 - 1] + o->lbs[i4 - 1] + kpol) % 2;
      if (m != 0 || kpol > o->mk) return DBS_OK;
      t = idef_rtype(ipb, o->ibo, o->eps_c, o->check_target, i1, i2, i3, i4, C, ic, jc, io, jo, key);
      if (t <= 0) return -t;
      return dbs_cdata_add_coef(cd, C, kpol, key[0], key[1], key[2], key[3], t);
    }
    case 3: {  // S-integrals: SMU factors, eight permuted terms per multipole
      DBS_CHECK(o->kbs, DBS_ERR_ARG, "Add_integral: kbs missing");
      for (int q : {i1, i2, i3, i4}) DBS_CHECK(q >= 1 && q <= o->norb, DBS_ERR_ARG, "Add_integral: orbital index out of range");
      double S[8];
      for (int v = kpol - 1; v <= kpol + 1; ++v) {
        if (v > o->mk) continue;
        if (smu_factors(o->kbs[i1 - 1], o->kbs[i2 - 1], o->kbs[i3 - 1], o->kbs[i4 - 1], kpol, v, S) == 0.0) continue;
        const int perm[8][4] = {{i1, i2, i3, i4}, {i2, i1, i4, i3}, {i3, i4, i1, i2}, {i4, i3, i2, i1},
                                {i1, i2, i3, i4}, {i4, i3, i2, i1}, {i3, i4, i1, i2}, {i2, i1, i4, i3}};
        for (int q = 0; q < 8; ++q) {
          const double Cq = C * S[q];
          t = idef_itype(ipb, o->ibo, q < 4 ? 0 : 1, perm[q][0], perm[q][1], perm[q][2], perm[q][3], Cq, ic, jc, io,
                         jo, key);
          if (t < 0) return -t;
          if (t == 0) continue;
          if ((rc = dbs_cdata_add_coef(cd, Cq, v, key[0], key[1], key[2], key[3], t))) return rc;
        }
      }
      return DBS_OK;
    }
    default:
      return fail(DBS_ERR_ARG, "Add_integral: unknown icase");
  }
}

int dbs_decode_int(int n, const int* code, int* met, int* k, int* i1, int* i2, int* i3, int* i4) {
  DBS_CHECK(n >= 0 && code && met && k && i1 && i2 && i3 && i4, DBS_ERR_ARG, "Decode_int: null argument");
  for (int q = 0; q < n; ++q) {
    int ii = code[q];
    met[q] = ii % 4; ii /= 4;
    k[q] = ii % 1024; ii /= 1024;
    i4[q] = ii % 32; ii /= 32;
    i3[q] = ii % 32; ii /= 32;
    i2[q] = ii % 32;
    i1[q] = ii / 32;
  }
  return DBS_OK;
}

int dbs_rdata_34_dev(dbs_ctx* ctx, int itype, int atype, int k, int nc, const double* c,
                     const int* K1, const int* K2, const int* K3, int ibo, const double* pq_dev,
                     int norb, const int* icc, int nch, double* hch_dev, int nblocks) {
  DBS_CHECK(ctx && c && K1 && K2 && K3 && pq_dev && icc && hch_dev, DBS_ERR_ARG, "R_data: null argument");
  DBS_CHECK(itype == 3 || itype == 4, DBS_ERR_ARG, " R_data: unknown itype ");
  DBS_CHECK(atype >= 0 && atype <= 3, DBS_ERR_ARG, "R_data: atype out of range");
  if (nc <= 0) return DBS_OK;
  const int ns = ctx->ns, ks = ctx->ks;
  int ip1, ip2, jp1, jp2;
  switch (atype) {
    case 0: ip1 = 1; ip2 = 1; jp1 = 1; jp2 = 1; break;
    case 1: ip1 = 2; ip2 = 2; jp1 = 2; jp2 = 2; break;
    case 2: ip1 = 1; ip2 = 1; jp1 = 2; jp2 = 2; if (itype == 4) { ip1 = 1; ip2 = 2; jp1 = 1; jp2 = 2; } break;
    default: ip1 = 2; ip2 = 2; jp1 = 1; jp2 = 1; if (itype == 4) { ip1 = 2; ip2 = 1; jp1 = 2; jp2 = 1; } break;
  }
  const int type = atype + 1;  // mrk_pppp, qqqq, pqpq, qpqp
  // runs of equal K1 -> groups
  std::vector<int> goff(1, 0), ej1, ej2;
  std::vector<double> ec;
  std::vector<int4> ginfo;
  for (int j = 0; j < nc; ++j) {
    DBS_CHECK(K2[j] >= 1 && K2[j] <= norb && K3[j] >= 1 && K3[j] <= norb, DBS_ERR_ARG,
              "R_data: orbital index out of range");
    ej1.push_back(K2[j]);
    ej2.push_back(K3[j]);
    ec.push_back(c[j]);
    const bool last = (j == nc - 1) || (K1[j] != K1[j + 1]);
    if (!last) continue;
    const int ich = K1[j] / ibo, jch = K1[j] % ibo;
    DBS_CHECK(ich >= 0, DBS_ERR_ARG, "UPDATE_HS: ich < 0");
    DBS_CHECK(jch >= 0, DBS_ERR_ARG, "UPDATE_HS: jch < 0");
    DBS_CHECK(ich <= nch, DBS_ERR_ARG, "UPDATE_HS: ich > nch");
    DBS_CHECK(jch <= nch, DBS_ERR_ARG, "UPDATE_HS: jch > nch");
    const int blk = (ich >= 1 && jch >= 1) ? icc[(ich - 1) + (size_t)nch * (jch - 1)] : 0;
    DBS_CHECK(blk >= 0 && blk <= nblocks, DBS_ERR_ARG, "UPDATE_HX: i out of range");
    if (blk == 0) {  // block not stored: drop the run
      ej1.resize(goff.back()); ej2.resize(goff.back()); ec.resize(goff.back());
      continue;
    }
    ginfo.push_back(make_int4(ich, jch, blk - 1, 0));
    goff.push_back((int)ec.size());
  }
  const int ng = (int)ginfo.size();
  if (ng == 0) return DBS_OK;
  const int sym = itype == 3 ? 0 : 1;
  // launch sizes, checked before anything is queued on the stream
  const int ntx = (ns + CV_TX - 1) / CV_TX, nty = (ns + CV_TY - 1) / CV_TY;
  if (sym == 1 && ks <= CV_KS && (long long)ntx * nty * ((ng + 3) / 4) > (1ll << 30))
    return fail(DBS_ERR_ARG, "R_data: too many convolution tiles");
  if (sym == 1 && (long long)ng * ((ns + 31) / 32) * ((ns + 31) / 32) > (1ll << 31) - 1)
    return fail(DBS_ERR_ARG, "R_data: too many density tiles");
  // case 1 (itype 3): one matrix product against the permuted array;
  // case 3 (itype 4): the array itself
  double* rkb = nullptr;
  int rc = sym == 0 ? build_rkm(ctx, type, k, &rkb) : build_rkb(ctx, type, k, &rkb);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const long long per = (long long)ns * (sym == 0 ? ks : ns);
  // every buffer of the call in the context's grow-only scratch (no device
  // allocation per call): densities, convolutions, coefficients, then the
  // integer tables (run offsets, orbitals, run info), 16-byte aligned;
  // densities and convolutions one group after another
  const size_t nxx = (size_t)per * ng, nce = ec.size();
  const size_t n_int = goff.size() + ej1.size() + ej2.size();
  const size_t off_c = 2 * nxx, off_i = off_c + nce, off_g = off_i + (n_int + 1) / 2 + 1;
  const size_t need = (off_g + 2 * (size_t)ng + 2) & ~(size_t)1;
  if ((rc = ensure_scratch(ctx, need))) return rc;
  double* d_xx = ctx->scratch;
  double* d_a = d_xx + nxx;
  double* d_c = ctx->scratch + off_c;
  int* d_goff = reinterpret_cast<int*>(ctx->scratch + off_i);
  int* d_j1 = d_goff + goff.size();
  int* d_j2 = d_j1 + ej1.size();
  int4* d_gi = reinterpret_cast<int4*>(ctx->scratch + ((off_g + 1) & ~(size_t)1));
  // the tables in the scratch layout from off_c on, one upload from the
  // context's pinned list buffers (the call does not wait for it)
  const size_t nblob = (need - off_c) * sizeof(double);
  char* hb = nullptr;
  if ((rc = list_stage_begin(ctx, nblob, &hb))) return rc;
  auto stage = [&](const void* dev, const void* src, size_t bytes) {
    if (bytes) std::memcpy(hb + (reinterpret_cast<const char*>(dev) - reinterpret_cast<const char*>(d_c)), src, bytes);
  };
  stage(d_c, ec.data(), nce * sizeof(double));
  stage(d_goff, goff.data(), goff.size() * sizeof(int));
  stage(d_j1, ej1.data(), ej1.size() * sizeof(int));
  stage(d_j2, ej2.data(), ej2.size() * sizeof(int));
  stage(d_gi, ginfo.data(), ginfo.size() * sizeof(int4));
  if (hipMemcpyAsync(d_c, hb, nblob, hipMemcpyHostToDevice, s) != hipSuccess) return fail(DBS_ERR_HIP, "R_data: upload");
  if ((rc = list_stage_end(ctx))) return rc;
  if (sym == 0) {
    k_group_density<<<grid_of(per * ng, 256), 256, 0, s>>>(ng, d_goff, d_j1, d_j2, d_c, pq_dev, ns, ks, jp1, jp2,
                                                            sym, d_xx);
  } else {
    const int nt = (ns + DX_T - 1) / DX_T;
    k_group_density_x<<<(unsigned)((long long)ng * nt * nt), 256, 0, s>>>(ng, nt, d_goff, d_j1, d_j2, d_c, pq_dev,
                                                                          ns, jp1, jp2, d_xx);
  }
  if (sym == 0) {
    if ((rc = conv_gemm(ctx, 1, ns * ks, ng, rkb, d_xx, d_a))) return rc;
    k_band_zero<<<grid_of(per * ng, 256), 256, 0, s>>>(ng, ns, ks, d_a);
  } else if (ks <= CV_KS) {
    constexpr int GB = 4;  // 37 KB of LDS, four workgroups per CU (measured best of 2, 4, 8, 16)
    const long long nblk = (long long)ntx * nty * ((ng + GB - 1) / GB);
    const unsigned grid = (unsigned)((nblk + 7) / 8 * 8);
    k_group_convol3_tile<GB><<<grid, CV_TX * CV_TY, 0, s>>>(ng, ns, ks, ntx, nty, rkb, d_xx, d_a);
  } else {
    constexpr int GB = 8;
    k_group_convol3<GB><<<grid_of(per * ((ng + GB - 1) / GB), 256), 256, 0, s>>>(ng, ns, ks, rkb, d_xx, d_a);
  }
  // one launch for all runs when their blocks differ; runs sharing a block
  // (K1 and its transpose, or a K1 repeated out of order) one launch each,
  // in list order
  std::vector<int> blks;
  for (const int4& G : ginfo) blks.push_back(G.z);
  std::sort(blks.begin(), blks.end());
  const bool shared = std::adjacent_find(blks.begin(), blks.end()) != blks.end();
  const long long nupd = (long long)ns * ns * ng;
  bool any_diag = false;
  for (const int4& G : ginfo) any_diag = any_diag || G.x == G.y;
  const int npass = any_diag ? 2 : 1;  // pass 1 only adds for ich == jch
  const int nt32 = (ns + 31) / 32;
  if (!shared) {
    if (sym == 1) {
      k_update_hs_x<<<(unsigned)((long long)ng * nt32 * nt32), 256, 0, s>>>(ng, nt32, d_gi, ns, ip1, ip2, d_a,
                                                                              hch_dev);
    } else {
      for (int pass = 0; pass < npass; ++pass)
        k_update_hs_mat<<<grid_of(nupd, 256), 256, 0, s>>>(ng, d_gi, ns, ks, sym, ip1, ip2, pass, d_a, hch_dev);
    }
  } else {
    for (int g = 0; g < ng; ++g) {
      if (sym == 1) {
        k_update_hs_x<<<(unsigned)(nt32 * nt32), 256, 0, s>>>(1, nt32, d_gi + g, ns, ip1, ip2, d_a + per * g,
                                                               hch_dev);
        continue;
      }
      for (int pass = 0; pass < (ginfo[g].x == ginfo[g].y ? 2 : 1); ++pass)
        k_update_hs_mat<<<grid_of((long long)ns * ns, 256), 256, 0, s>>>(1, d_gi + g, ns, ks, sym, ip1, ip2, pass,
                                                                         d_a + per * g, hch_dev);
    }
  }
  // queued, not waited for: the pinned tables belong to the context
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) rc = fail(DBS_ERR_HIP, std::string("R_data: ") + hipGetErrorString(e));
  return rc;
}

}  // extern "C"

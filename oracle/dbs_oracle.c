/*
 * dbs_oracle.c -- CPU restatement of DBSR's B-spline Slater-integral path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed by bench.py.  The product never links or calls this file.
 *
 * Restates, loop for loop, the reference algorithms in
 *   DBSR_HF/dbsr_lib_dbs.f90   (grid, Gauss tables, moments, diagonal cells,
 *                               mrk_ and msk_ arrays, density, convol, SUM_AmB,
 *                               quadr, rk, UPDATE_HS)
 *   DBSR_HF/dbsr_lib_zcom.f90  (gauleg)
 *   DBSR_MAT3/MOD_c_data.f90   (blocked sorted coefficient lists)
 *   DBSR_CI3/rk_evaluate.f90   (list evaluation into HM/DM)
 *   DBSR_MAT3/R_data.f90, updates.f90, idef_type.f90 (channel blocks)
 * Loop orders follow the reference so that summation orders agree; arrays
 * are column-major with the reference's dimensions, indices 0-based here.
 */
#include "dbs_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXK 16  /* largest B-spline order handled by the fixed-size scratch */

/* x**n for integer n, by binary powering (the lowering of Fortran real**int) */
static double ipow(double a, int b) {
  int recip = b < 0;
  unsigned int n = recip ? (unsigned int)(-b) : (unsigned int)b;
  double r = 1.0;
  for (;;) {
    if (n & 1u) r *= a;
    n >>= 1;
    if (n == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

/* ===================================================================== */
/* gauleg: DBSR_HF/dbsr_lib_zcom.f90:1990-2039                           */
/* ===================================================================== */
void orc_gauleg(double x1, double x2, double *x, double *w, int n) {
  const double EPS = 1.0e-14;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double z = cos(3.141592654 * (i - 0.25) / (n + 0.5));
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
    } while (fabs(z - z1) > EPS);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* ===================================================================== */
/* zbsplvd / zbsplvb: DBSR_HF/dbsr_lib_dbs.f90:4197-4468                 */
/* de Boor's bsplvd, one point at a time.  t is 1-based in the reference; */
/* here T(i) = t[i-1].  dbiatx is (nv,ks,ks) column-major: entry          */
/* (ip,j,m) = value of the (m-1)st derivative of B-spline left-ks+j at    */
/* point ip, left = kg+ip-1 (1-based).                                    */
/* ===================================================================== */
#define T1(i) t[(i) - 1]
void orc_zbsplvd(int ns, int ks, int nv, const double *t, int kg, int ni,
                 const double *x, int nderiv, double *dbiatx) {
  (void)ns;
#define DB(ip, j, m) dbiatx[(ip) + (size_t)nv * (((j) - 1) + (size_t)ks * ((m) - 1))]
  for (int ip = 0; ip < ni; ++ip) {
    const double xx = x[ip];
    const int left = kg + ip;  /* 1-based knot index of the interval start */
    double deltar[MAXK + 1], deltal[MAXK + 1];
    int jstate = 1;
    /* bsplvb(jhigh, index): raises the order held in DB(ip,1..j,1) */
#define BSPLVB(JHIGH, INDEX)                                                   \
    do {                                                                       \
      if ((INDEX) == 1) {                                                      \
        jstate = 1;                                                            \
        DB(ip, 1, 1) = 1.0;                                                    \
        if (jstate >= (JHIGH)) break;                                          \
      }                                                                        \
      for (;;) {                                                               \
        int jp1 = jstate + 1;                                                  \
        double saved = 0.0;                                                    \
        deltar[jstate] = T1(left + jstate) - xx;                               \
        deltal[jstate] = xx - T1(left + 1 - jstate);                           \
        for (int m_ = 1; m_ <= jstate; ++m_) {                                 \
          double term = DB(ip, m_, 1) / (deltar[m_] + deltal[jp1 - m_]);       \
          DB(ip, m_, 1) = saved + deltar[m_] * term;                           \
          saved = deltal[jp1 - m_] * term;                                     \
        }                                                                      \
        DB(ip, jp1, 1) = saved;                                                \
        jstate = jp1;                                                          \
        if (jstate >= (JHIGH)) break;                                          \
      }                                                                        \
    } while (0)

    int mhigh = nderiv < ks ? nderiv : ks;
    if (mhigh < 1) mhigh = 1;
    int kp1 = ks + 1;
    int jhigh = kp1 - mhigh;
    BSPLVB(jhigh, 1);
    if (mhigh == 1) continue;

    int ideriv = mhigh;
    for (int m = 2; m <= mhigh; ++m) {
      int jpimid = 1;
      for (int j = ideriv; j <= ks; ++j) {
        DB(ip, j, ideriv) = DB(ip, jpimid, 1);
        jpimid++;
      }
      ideriv--;
      jhigh = kp1 - ideriv;
      BSPLVB(jhigh, 2);
    }
    /* w31(i,j): B-coefficients of the derivatives */
    double w31[MAXK + 1][MAXK + 1];
    int jlow = 1;
    for (int i = 1; i <= ks; ++i) {
      for (int j = jlow; j <= ks; ++j) w31[j][i] = 0.0;
      jlow = i;
      w31[i][i] = 1.0;
    }
    for (int m = 2; m <= mhigh; ++m) {
      int kpimm = kp1 - m;
      double fkpimm = kpimm;
      int i = ks, il = 0;
      for (int ld = 1; ld <= kpimm; ++ld) {
        int n = left - il;
        double w1 = fkpimm / (T1(n + kpimm) - T1(n));
        for (int j = 1; j <= i; ++j) w31[i][j] = (w31[i][j] - w31[i - 1][j]) * w1;
        il++;
        i--;
      }
      for (int ii = 1; ii <= ks; ++ii) {
        double w2 = 0.0;
        int jl = ii > m ? ii : m;
        for (int j = jl; j <= ks; ++j) w2 += w31[j][ii] * DB(ip, j, m);
        DB(ip, ii, m) = w2;
      }
    }
#undef BSPLVB
  }
#undef DB
}

/* ===================================================================== */
/* first knot step of the grid recipes by nucleus (mkgrid_01 :737-745,    */
/* mkgrid_02 :800-807): nucl 0 point hi/z, 1 Fermi 4 ln3 a_fermi hi,      */
/* 2 uniform r_uniform hi                                                 */
/* ===================================================================== */
double orc_grid_start(int nucl, double hi, double z, double a_fermi, double r_uniform) {
  if (nucl == 0) return hi / z;
  if (nucl == 1) {
    double ts = 4.0 * log(3.0) * a_fermi;
    return ts * hi;
  }
  return r_uniform * hi;
}

/* ===================================================================== */
/* mkgrid_01: DBSR_HF/dbsr_lib_dbs.f90:724-783, from the first step hs    */
/* returns ns; t must hold ns+ks entries (tcap checked)                   */
/* ===================================================================== */
int orc_mkgrid_01s(int ks, int ml, double hs, double he, double hmax, double rmax, double *t, int tcap) {
  double ti = hs * ml, tj, ht = 1.0 + he;
  int ns = ks + ml, ne = ns;
  for (;;) {
    tj = ti * ht;
    if (tj - ti > hmax) break;
    ns = ns + 1; ti = tj; ne = ns;
    if (ti > rmax) { ns = ns - 1; break; }
  }
  if (ti < rmax) {
    for (;;) {
      ti = ti + hmax;
      if (ti >= rmax) break;
      ns = ns + 1;
    }
  }
  if (t == NULL || tcap < ns + ks) return ns;
  for (int i = 1; i <= ns + ks; ++i) T1(i) = 0.0;
  for (int i = ks + 1; i <= ks + ml; ++i) T1(i) = hs * (i - ks);
  for (int i = ks + ml + 1; i <= ne; ++i) T1(i) = T1(i - 1) * ht;
  if (T1(ne) < rmax)
    for (int i = ne + 1; i <= ns + 1; ++i) T1(i) = T1(i - 1) + hmax;
  for (int i = ns + 2; i <= ns + ks; ++i) T1(i) = T1(ns + 1);
  return ns;
}

/* mkgrid_01, point nucleus */
int orc_mkgrid_01(int ks, int ml, double hi, double he, double hmax, double rmax,
                  double z, double *t, int tcap) {
  return orc_mkgrid_01s(ks, ml, orc_grid_start(0, hi, z, 0.0, 0.0), he, hmax, rmax, t, tcap);
}

/* mkgrid_02: DBSR_HF/dbsr_lib_dbs.f90:787-823, from the first knot ti */
void orc_mkgrid_02s(int ns, int ks, double ti, double rmax, double *t) {
  double tmax = rmax;
  double he = exp((log(tmax) - log(ti)) / (ns - ks));
  for (int i = 1; i <= ks; ++i) T1(i) = 0.0;
  T1(ks + 1) = ti;
  for (int i = ks + 2; i <= ns + 1; ++i) T1(i) = T1(i - 1) * he;
  for (int i = ns + 1; i <= ns + ks; ++i) T1(i) = tmax;
}

/* mkgrid_02, point nucleus */
void orc_mkgrid_02(int ns, int ks, double hi, double rmax, double z, double *t) {
  orc_mkgrid_02s(ns, ks, orc_grid_start(0, hi, z, 0.0, 0.0), rmax, t);
}
#undef T1

/* ===================================================================== */
/* Gauss tables: alloc_DBS_gauss, Def_pbsp, Def_qbsp                      */
/* DBSR_HF/dbsr_lib_dbs.f90:911-1074, 565-589                             */
/* ===================================================================== */
static void def_xbsp(const orc_grid *g, int kx, int nsx, const double *tx, double *xbsp) {
  const int nv = g->nv, ks = g->ks;
  const int nw = nv + 1;
  double *dbi = (double *)calloc((size_t)nv * kx * kx, sizeof(double));
  double *xm = (double *)malloc(sizeof(double) * nv);
#define XB(iv, m, j) xbsp[(iv) + (size_t)nw * ((m) + (size_t)ks * (j))]
  for (size_t q = 0; q < (size_t)nw * ks * kx; ++q) xbsp[q] = 0.0;
  for (int m = 0; m < ks; ++m) {
    for (int iv = 0; iv < nv; ++iv) xm[iv] = g->gr[iv + (size_t)nv * m];
    orc_zbsplvd(nsx, kx, nv, tx, kx, nv, xm, 1, dbi);
    for (int j = 0; j < kx; ++j)
      for (int iv = 0; iv < nv; ++iv) XB(iv, m, j) = dbi[iv + (size_t)nv * j];
  }
  /* values at the last knot, row nv+1 (1-based) */
  double xe = tx[nsx];  /* tx(nsx+1) */
  orc_zbsplvd(nsx, kx, nv, tx, nsx, 1, &xe, 1, dbi);
  for (int j = 0; j < kx; ++j) XB(nv, 0, j) = dbi[0 + (size_t)nv * j];
#undef XB
  free(dbi);
  free(xm);
}

int orc_grid_init(orc_grid *g, int ns, int ks, int ksp, int ksq, const double *t) {
  memset(g, 0, sizeof(*g));
  if (ks > MAXK || ks < 2) return -1;
  if (ksp <= 0 || ksp > ks) ksp = ks - 1;
  if (ksq <= 0 || ksq > ks) ksq = ks;
  g->ns = ns; g->ks = ks; g->ksp = ksp; g->ksq = ksq;
  g->ms = ns + ns; g->nv = ns - ks + 1;
  g->nsp = g->nv + ksp - 1; g->nsq = g->nv + ksq - 1;
  const int nv = g->nv, nsp = g->nsp, nsq = g->nsq;
  g->t = (double *)malloc(sizeof(double) * (ns + ks));
  memcpy(g->t, t, sizeof(double) * (ns + ks));
  g->gx = (double *)malloc(sizeof(double) * ks);
  g->gw = (double *)malloc(sizeof(double) * ks);
  g->gr = (double *)malloc(sizeof(double) * nv * ks);
  g->grm = (double *)malloc(sizeof(double) * nv * ks);
  g->grw = (double *)malloc(sizeof(double) * nv * ks);
  orc_gauleg(0.0, 1.0, g->gx, g->gw, ks);
  for (int m = 0; m < ks; ++m)
    for (int iv = 0; iv < nv; ++iv) {
      /* t(iv+ks) - t(iv+ks-1) with iv 1-based => t[iv+ks] - t[iv+ks-1] 0-based iv */
      double h = g->t[iv + ks] - g->t[iv + ks - 1];
      g->gr[iv + (size_t)nv * m] = h * g->gx[m] + g->t[iv + ks - 1];
      g->grm[iv + (size_t)nv * m] = 1.0 / g->gr[iv + (size_t)nv * m];
      g->grw[iv + (size_t)nv * m] = h * g->gw[m];
    }
  /* tp(1:ksp)=t(1); tp(ksp+1:nsp)=t(ks+1:ns); tp(nsp+1:nsp+ksp)=t(ns+1:ns+ksp) */
  g->tp = (double *)malloc(sizeof(double) * (nsp + ksp));
  g->tq = (double *)malloc(sizeof(double) * (nsq + ksq));
  for (int i = 0; i < ksp; ++i) g->tp[i] = g->t[0];
  for (int i = ksp; i < nsp; ++i) g->tp[i] = g->t[ks + (i - ksp)];
  for (int i = nsp; i < nsp + ksp; ++i) g->tp[i] = g->t[ns + (i - nsp)];
  for (int i = 0; i < ksq; ++i) g->tq[i] = g->t[0];
  for (int i = ksq; i < nsq; ++i) g->tq[i] = g->t[ks + (i - ksq)];
  for (int i = nsq; i < nsq + ksq; ++i) g->tq[i] = g->t[ns + (i - nsq)];
  g->pbsp = (double *)malloc(sizeof(double) * (nv + 1) * ks * ksp);
  g->qbsp = (double *)malloc(sizeof(double) * (nv + 1) * ks * ksq);
  def_xbsp(g, ksp, nsp, g->tp, g->pbsp);
  def_xbsp(g, ksq, nsq, g->tq, g->qbsp);
  return 0;
}

void orc_grid_free(orc_grid *g) {
  free(g->t); free(g->tp); free(g->tq); free(g->gx); free(g->gw);
  free(g->gr); free(g->grm); free(g->grw); free(g->pbsp); free(g->qbsp);
  memset(g, 0, sizeof(*g));
}

orc_grid *orc_grid_new(int ns, int ks, int ksp, int ksq, const double *t) {
  orc_grid *g = (orc_grid *)malloc(sizeof(orc_grid));
  if (orc_grid_init(g, ns, ks, ksp, ksq, t) != 0) { free(g); return NULL; }
  return g;
}
void orc_grid_delete(orc_grid *g) { if (g) { orc_grid_free(g); free(g); } }
void orc_grid_dims(const orc_grid *g, int *d) {
  d[0] = g->ns; d[1] = g->ks; d[2] = g->ksp; d[3] = g->ksq;
  d[4] = g->nv; d[5] = g->nsp; d[6] = g->nsq; d[7] = g->ms;
}
void orc_grid_tables(const orc_grid *g, double *gr, double *grw, double *pbsp, double *qbsp) {
  size_t nvk = (size_t)g->nv * g->ks;
  if (gr) memcpy(gr, g->gr, sizeof(double) * nvk);
  if (grw) memcpy(grw, g->grw, sizeof(double) * nvk);
  if (pbsp) memcpy(pbsp, g->pbsp, sizeof(double) * (g->nv + 1) * g->ks * g->ksp);
  if (qbsp) memcpy(qbsp, g->qbsp, sizeof(double) * (g->nv + 1) * g->ks * g->ksq);
}

/* ===================================================================== */
/* moments_pp/qq/pq/qp + moment_*: DBSR_HF/dbsr_lib_dbs.f90:1384-1564     */
/* rkm is (ks*ks, nv); only the packed leading part is written.           */
/* ===================================================================== */
void orc_moments(const orc_grid *g, int which, int k, double *rkm) {
  const int nv = g->nv, ks = g->ks, kk = ks * ks;
  const int nw = nv + 1;
  memset(rkm, 0, sizeof(double) * (size_t)kk * nv);
  const double *b1t, *b2t;
  int k1, k2, sym;
  switch (which) {
    case ORC_MOM_PP: b1t = g->pbsp; b2t = g->pbsp; k1 = g->ksp; k2 = g->ksp; sym = 1; break;
    case ORC_MOM_QQ: b1t = g->qbsp; b2t = g->qbsp; k1 = g->ksq; k2 = g->ksq; sym = 1; break;
    case ORC_MOM_PQ: b1t = g->pbsp; b2t = g->qbsp; k1 = g->ksp; k2 = g->ksq; sym = 0; break;
    default:         b1t = g->qbsp; b2t = g->pbsp; k1 = g->ksq; k2 = g->ksp; sym = 0; break;
  }
  double gw[MAXK], b2[MAXK][MAXK];
  for (int iv = 0; iv < nv; ++iv) {
    for (int m = 0; m < ks; ++m) {
      double w = g->grw[iv + (size_t)nv * m];
      if (k > 0) w = w * ipow(g->gr[iv + (size_t)nv * m], k);
      if (k < 0) w = w * ipow(g->grm[iv + (size_t)nv * m], -k);
      gw[m] = w;
    }
    for (int j = 0; j < k2; ++j)
      for (int m = 0; m < ks; ++m)
        b2[j][m] = b2t[iv + (size_t)nw * (m + (size_t)ks * j)] * gw[m];
    double *rv = rkm + (size_t)kk * iv;
    int ii = 0;
    for (int i = 0; i < k1; ++i)
      for (int j = (sym ? i : 0); j < k2; ++j) {
        double s = 0.0;
        for (int m = 0; m < ks; ++m) s += b1t[iv + (size_t)nw * (m + (size_t)ks * i)] * b2[j][m];
        rv[ii++] = s;
      }
  }
}

/* ===================================================================== */
/* triang_*: DBSR_HF/dbsr_lib_dbs.f90:2979-3063 (pppp), 3151-3245 (pqpq), */
/* 3333-3427 (qpqp), 3512-3597 (qqqq), 3682-3767 (ppqq), 3852-3936 (pqqp) */
/* rkd is (kk,kk) for one interval iv (0-based), zero-filled.             */
/* ===================================================================== */
/* inner-rule B-spline values: zbsplvd(nsx,kx,nv,tx,left,1,x,1,...) */
static void bvals(const double *tx, int nsx, int kx, int left, double x, double *out) {
  double dbi[MAXK * MAXK];
  orc_zbsplvd(nsx, kx, 1, tx, left, 1, &x, 1, dbi);
  for (int j = 0; j < kx; ++j) out[j] = dbi[j];  /* dbiatx(1,j,1) with nv=1 */
}

void orc_triang(const orc_grid *g, int type, int k, int iv, double *rkd) {
  const int nv = g->nv, ks = g->ks, kk = ks * ks, ksp = g->ksp, ksq = g->ksq;
  const int nw = nv + 1;
  const int ivf = iv + 1;  /* 1-based interval */
  memset(rkd, 0, sizeof(double) * (size_t)kk * kk);
#define RKD(a, b) rkd[(a) + (size_t)kk * (b)]
#define GR(iv_, m) g->gr[(iv_) + (size_t)nv * (m)]
#define GRW(iv_, m) g->grw[(iv_) + (size_t)nv * (m)]
#define GRM(iv_, m) g->grm[(iv_) + (size_t)nv * (m)]
#define PB(iv_, m, j) g->pbsp[(iv_) + (size_t)nw * ((m) + (size_t)ks * (j))]
#define QB(iv_, m, j) g->qbsp[(iv_) + (size_t)nw * ((m) + (size_t)ks * (j))]
  double x[MAXK], w[MAXK], gxl[MAXK], gwl[MAXK], gx2[MAXK], bi[MAXK];
  double INTP[MAXK][MAXK][MAXK], INTQ[MAXK][MAXK][MAXK];  /* (j,jp,m) */
  double bTp[MAXK][MAXK], bTq[MAXK][MAXK];                        /* (i,j)    */
  orc_gauleg(0.0, 1.0, x, w, ks);

  int needP = (type == ORC_PPPP || type == ORC_PQPQ || type == ORC_QPQP ||
               type == ORC_PPQQ || type == ORC_PQQP);
  int needQ = (type != ORC_PPPP);
  int mixed = (type == ORC_PPQQ || type == ORC_PQQP);
  double xbase;
  if (type == ORC_PPPP) xbase = g->tp[ivf + ksp - 1 - 1];
  else if (type == ORC_QQQQ) xbase = g->tq[ivf + ksq - 1 - 1];
  else xbase = g->t[ivf + ks - 1 - 1];
  int leftp = ivf + ksp - 1, leftq = ivf + ksq - 1;

  for (int m = 0; m < ks; ++m) {
    for (int i = 0; i < ks; ++i) gxl[i] = (GR(iv, m) - xbase) * x[i] + xbase;
    if (needP) for (int i = 0; i < ks; ++i) bvals(g->tp, g->nsp, ksp, leftp, gxl[i], bTp[i]);
    if (needQ) for (int i = 0; i < ks; ++i) bvals(g->tq, g->nsq, ksq, leftq, gxl[i], bTq[i]);
    for (int i = 0; i < ks; ++i) gwl[i] = (GR(iv, m) - xbase) * w[i];
    if (type == ORC_PPPP || type == ORC_QQQQ) {
      if (k > 1) for (int i = 0; i < ks; ++i) gxl[i] = gwl[i] * ipow(gxl[i], k);
      else if (k == 1) for (int i = 0; i < ks; ++i) gxl[i] = gwl[i] * gxl[i];
      else if (k == 0) for (int i = 0; i < ks; ++i) gxl[i] = gwl[i];
    } else {
      for (int i = 0; i < ks; ++i) gxl[i] = gwl[i] * ipow(gxl[i], k);
    }
    if (!mixed) {
      if (needP)
        for (int j = 0; j < ksp; ++j) {
          double gwj[MAXK];
          for (int i = 0; i < ks; ++i) gwj[i] = gxl[i] * bTp[i][j];
          for (int jp = j; jp < ksp; ++jp) {
            double s = 0.0;
            for (int i = 0; i < ks; ++i) s += gwj[i] * bTp[i][jp];
            INTP[j][jp][m] = s;
          }
        }
      if (needQ)
        for (int j = 0; j < ksq; ++j) {
          double gwj[MAXK];
          for (int i = 0; i < ks; ++i) gwj[i] = gxl[i] * bTq[i][j];
          for (int jp = j; jp < ksq; ++jp) {
            double s = 0.0;
            for (int i = 0; i < ks; ++i) s += gwj[i] * bTq[i][jp];
            INTQ[j][jp][m] = s;
          }
        }
    } else {
      /* INTPQ(j,jp,m): j over P, jp over Q (stored in INTP) */
      for (int j = 0; j < ksp; ++j) {
        double gwj[MAXK];
        for (int i = 0; i < ks; ++i) gwj[i] = gxl[i] * bTp[i][j];
        for (int jp = 0; jp < ksq; ++jp) {
          double s = 0.0;
          for (int i = 0; i < ks; ++i) s += gwj[i] * bTq[i][jp];
          INTP[j][jp][m] = s;
        }
      }
    }
  }
  /* second integration weights */
  for (int m = 0; m < ks; ++m) {
    if ((type == ORC_PPPP || type == ORC_QQQQ) && k == 0) gx2[m] = GRW(iv, m) * GRM(iv, m);
    else gx2[m] = GRW(iv, m) * ipow(GRM(iv, m), k + 1);
  }

  if (type == ORC_PPPP || type == ORC_QQQQ) {
    int kx = (type == ORC_PPPP) ? ksp : ksq;
    double (*IN)[MAXK][MAXK] = (type == ORC_PPPP) ? INTP : INTQ;
    int ik = kx * (kx + 1) / 2;
    double *a = (double *)malloc(sizeof(double) * ik * ik);
    int ii = 0;
    for (int i = 0; i < kx; ++i)
      for (int ip = i; ip < kx; ++ip) {
        for (int m = 0; m < ks; ++m) {
          double bv = (type == ORC_PPPP) ? PB(iv, m, i) * PB(iv, m, ip) : QB(iv, m, i) * QB(iv, m, ip);
          bi[m] = bv * gx2[m];
        }
        int jj = 0;
        for (int j = 0; j < kx; ++j)
          for (int jp = j; jp < kx; ++jp) {
            double s = 0.0;
            for (int m = 0; m < ks; ++m) s += bi[m] * IN[j][jp][m];
            a[ii + (size_t)ik * jj] = s;
            jj++;
          }
        ii++;
      }
    for (int b = 0; b < ik; ++b)
      for (int c = 0; c < ik; ++c) RKD(c, b) = a[c + (size_t)ik * b] + a[b + (size_t)ik * c];
    free(a);
  } else if (type == ORC_PQPQ || type == ORC_QPQP) {
    /* pass 1: P pairs outer, Q pairs inner; pass 2: Q outer, P inner */
    int ii = 0;
    for (int i = 0; i < ksp; ++i)
      for (int ip = i; ip < ksp; ++ip) {
        for (int m = 0; m < ks; ++m) bi[m] = PB(iv, m, i) * PB(iv, m, ip) * gx2[m];
        int jj = 0;
        for (int j = 0; j < ksq; ++j)
          for (int jp = j; jp < ksq; ++jp) {
            double s = 0.0;
            for (int m = 0; m < ks; ++m) s += bi[m] * INTQ[j][jp][m];
            if (type == ORC_PQPQ) RKD(ii, jj) += s; else RKD(jj, ii) += s;
            jj++;
          }
        ii++;
      }
    ii = 0;
    for (int i = 0; i < ksq; ++i)
      for (int ip = i; ip < ksq; ++ip) {
        for (int m = 0; m < ks; ++m) bi[m] = QB(iv, m, i) * QB(iv, m, ip) * gx2[m];
        int jj = 0;
        for (int j = 0; j < ksp; ++j)
          for (int jp = j; jp < ksp; ++jp) {
            double s = 0.0;
            for (int m = 0; m < ks; ++m) s += bi[m] * INTP[j][jp][m];
            if (type == ORC_PQPQ) RKD(jj, ii) += s; else RKD(ii, jj) += s;
            jj++;
          }
        ii++;
      }
  } else {
    /* ppqq: outer (i in P, ip in Q); pqqp: outer (i in Q, ip in P);
       inner INTPQ(j in P, jp in Q); rkd(jj,ii) */
    int ii = 0;
    int ko1 = (type == ORC_PPQQ) ? ksp : ksq, ko2 = (type == ORC_PPQQ) ? ksq : ksp;
    for (int i = 0; i < ko1; ++i)
      for (int ip = 0; ip < ko2; ++ip) {
        for (int m = 0; m < ks; ++m) {
          double bv = (type == ORC_PPQQ) ? PB(iv, m, i) * QB(iv, m, ip) : QB(iv, m, i) * PB(iv, m, ip);
          bi[m] = bv * gx2[m];
        }
        int jj = 0;
        for (int j = 0; j < ksp; ++j)
          for (int jp = 0; jp < ksq; ++jp) {
            double s = 0.0;
            for (int m = 0; m < ks; ++m) s += bi[m] * INTP[j][jp][m];
            RKD(jj, ii) = s;
            jj++;
          }
        ii++;
      }
  }
#undef RKD
#undef GR
#undef GRW
#undef GRM
#undef PB
#undef QB
}

/* ===================================================================== */
/* mrk_pppp/qqqq/pqpq/qpqp: DBSR_HF/dbsr_lib_dbs.f90:2898-2959, 3431-3492,*/
/* 3067-3131, 3249-3313;  msk_ppqq/pqqp: 3601-3662, 3771-3833             */
/* ===================================================================== */
size_t orc_rkb_size(const orc_grid *g, int type) {
  size_t ns = (size_t)g->ns;
  size_t kb = (type >= ORC_PPQQ) ? (size_t)(2 * g->ks - 1) : (size_t)g->ks;
  return ns * ns * kb * kb;
}

void orc_mrk(const orc_grid *g, int type, int k, double *rkb) {
  const int nv = g->nv, ns = g->ns, ks = g->ks, kk = ks * ks;
  const int ksp = g->ksp, ksq = g->ksq;
  double *r1 = (double *)malloc(sizeof(double) * kk * nv);
  double *r2 = (double *)malloc(sizeof(double) * kk * nv);
  double *r3 = (double *)malloc(sizeof(double) * kk * nv);
  double *r4 = (double *)malloc(sizeof(double) * kk * nv);
  double *rkd = (double *)malloc(sizeof(double) * (size_t)kk * kk * nv);
  memset(rkb, 0, sizeof(double) * orc_rkb_size(g, type));
  switch (type) {
    case ORC_PPPP: orc_moments(g, ORC_MOM_PP, k, r1); orc_moments(g, ORC_MOM_PP, -(k + 1), r2); break;
    case ORC_QQQQ: orc_moments(g, ORC_MOM_QQ, k, r1); orc_moments(g, ORC_MOM_QQ, -(k + 1), r2); break;
    case ORC_PQPQ:
    case ORC_QPQP:
      orc_moments(g, ORC_MOM_PP, k, r1); orc_moments(g, ORC_MOM_QQ, -(k + 1), r2);
      orc_moments(g, ORC_MOM_QQ, k, r3); orc_moments(g, ORC_MOM_PP, -(k + 1), r4); break;
    case ORC_PPQQ: orc_moments(g, ORC_MOM_PQ, k, r1); orc_moments(g, ORC_MOM_PQ, -(k + 1), r2); break;
    default:       orc_moments(g, ORC_MOM_PQ, k, r1); orc_moments(g, ORC_MOM_QP, -(k + 1), r2); break;
  }
  for (int iv = 0; iv < nv; ++iv) orc_triang(g, type, k, iv, rkd + (size_t)kk * kk * iv);
#define M(a, ii, iv) a[(ii) + (size_t)kk * (iv)]
#define RKDV(ii, jj, iv) rkd[(ii) + (size_t)kk * ((jj) + (size_t)kk * (iv))]
  if (type <= ORC_QPQP) {
    int kl = (type == ORC_PPPP || type == ORC_PQPQ) ? ksp : ksq;  /* left (r1) order */
    int kr = (type == ORC_PPPP || type == ORC_QPQP) ? ksp : ksq;  /* right (r2) order */
#define RKB(i, j, ip, jp) rkb[(i) + (size_t)ns * ((j) + (size_t)ns * ((ip) + (size_t)ks * (jp)))]
    for (int jv = 0; jv < nv; ++jv) {
      int jj = 0;
      for (int jh = 0; jh < kr; ++jh)
        for (int jhp = jh; jhp < kr; ++jhp) {
          int j = jv + jh, jp = jhp - jh;
          for (int iv = 0; iv < nv; ++iv) {
            int ii = 0;
            for (int ih = 0; ih < kl; ++ih)
              for (int ihp = ih; ihp < kl; ++ihp) {
                int i = iv + ih, ip = ihp - ih;
                double c;
                if (iv < jv) {
                  if (type == ORC_QPQP) c = M(r3, ii, iv) * M(r4, jj, jv);
                  else c = M(r1, ii, iv) * M(r2, jj, jv);
                } else if (iv > jv) {
                  if (type == ORC_PQPQ) c = M(r3, jj, jv) * M(r4, ii, iv);
                  else c = M(r1, jj, jv) * M(r2, ii, iv);
                } else {
                  c = RKDV(ii, jj, iv);
                }
                RKB(i, j, ip, jp) += c;
                ii++;
              }
          }
          jj++;
        }
    }
#undef RKB
  } else {
    const int kb = 2 * ks - 1;
    int kr1 = (type == ORC_PPQQ) ? ksp : ksq, kr2 = (type == ORC_PPQQ) ? ksq : ksp;
#define RKB(i, j, ip, jp) rkb[(i) + (size_t)ns * ((j) + (size_t)ns * ((ip) + (size_t)kb * (jp)))]
    for (int jv = 0; jv < nv; ++jv) {
      int jj = 0;
      for (int jh = 0; jh < kr1; ++jh)
        for (int jhp = 0; jhp < kr2; ++jhp) {
          int j = jv + jh, jp = jhp - jh + ks - 1;
          for (int iv = 0; iv <= jv; ++iv) {
            int ii = 0;
            for (int ih = 0; ih < ksp; ++ih)
              for (int ihp = 0; ihp < ksq; ++ihp) {
                int i = iv + ih, ip = ihp - ih + ks - 1;
                double c = (iv == jv) ? RKDV(ii, jj, iv) : M(r1, ii, iv) * M(r2, jj, jv);
                RKB(i, j, ip, jp) += c;
                ii++;
              }
          }
          jj++;
        }
    }
#undef RKB
  }
#undef M
#undef RKDV
  free(r1); free(r2); free(r3); free(r4); free(rkd);
}

/* ===================================================================== */
/* density: DBSR_HF/dbsr_lib_dbs.f90:2834-2891                            */
/* ===================================================================== */
void orc_density(int ns, int ks, double *d, const double *p1, const double *p2, char type) {
#define D(i, j) d[(i) + (size_t)ns * (j)]
  if (type == 's' || type == 'u') {
    memset(d, 0, sizeof(double) * (size_t)ns * ks);
    for (int i = 0; i < ns; ++i) D(i, 0) = p1[i] * p2[i];
    for (int j = 1; j < ks; ++j)
      for (int i = 0; i < ns - j; ++i) D(i, j) = p1[i] * p2[i + j] + p1[i + j] * p2[i];
  } else if (type == 'l') {
    memset(d, 0, sizeof(double) * (size_t)ns * ks);
    for (int i = 0; i < ns; ++i) D(i, ks - 1) = p1[i] * p2[i];
    for (int j = 1; j <= ks - 1; ++j)   /* 1-based j */
      for (int i = ks + 1 - j; i <= ns; ++i)
        D(i - 1, j - 1) = p1[i - 1] * p2[i + j - ks - 1] + p1[i + j - ks - 1] * p2[i - 1];
  } else if (type == 'n') {
    memset(d, 0, sizeof(double) * (size_t)ns * (2 * ks - 1));
    for (int j = 1; j <= 2 * ks - 1; ++j) {
      int imin = 1 > 1 + ks - j ? 1 : 1 + ks - j;
      int imax = ns < ns + ks - j ? ns : ns + ks - j;
      for (int i = imin; i <= imax; ++i) D(i - 1, j - 1) = p1[i - 1] * p2[i + j - ks - 1];
    }
  } else {
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) D(i, j) = p1[i] * p2[j];
  }
#undef D
}

/* ===================================================================== */
/* convol / convol_sk: DBSR_HF/dbsr_lib_dbs.f90:2208-2478, 2482-2752      */
/* kb = third/fourth dimension of rkb (ks for R^k, 2ks-1 for S^k)         */
/* ===================================================================== */
void orc_convol(int ns, int ks, int kb, const double *rkb, double *a, const double *d,
                int icase, char si, char sj) {
#define R(i, j, ip, jp) rkb[((i) - 1) + (size_t)ns * (((j) - 1) + (size_t)ns * (((ip) - 1) + (size_t)kb * ((jp) - 1)))]
#define A(i, j) a[((i) - 1) + (size_t)ns * ((j) - 1)]
#define DD(i, j) d[((i) - 1) + (size_t)ns * ((j) - 1)]
#define IMIN(ip) ((1 > 1 + ks - (ip)) ? 1 : 1 + ks - (ip))
#define IMAX(ip) ((ns < ns + ks - (ip)) ? ns : ns + ks - (ip))
  int i, j, ip, jp, ii, jj;
  double c;
  if (icase <= 2) {
    int ncol = (si == 'n' || sj == 'n') ? 2 * ks - 1 : ks;
    (void)ncol;
    for (j = 1; j <= ks; ++j) for (i = 1; i <= ns; ++i) A(i, j) = 0.0;
  } else {
    memset(a, 0, sizeof(double) * (size_t)ns * ns);
  }
  if (icase == 1) {
    if (si == 's' && sj == 's') {
      for (ip = 1; ip <= ks; ++ip) for (i = 1; i <= ns - ip + 1; ++i) {
        c = 0.0;
        for (jp = 1; jp <= ks; ++jp) for (j = 1; j <= ns - jp + 1; ++j) c += DD(j, jp) * R(i, j, ip, jp);
        A(i, ip) = c;
      }
    } else if (si == 'n' && sj == 'n') {
      for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) {
        c = 0.0;
        for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) c += DD(j, jp) * R(i, j, ip, jp);
        A(i, ip) = c;
      }
    } else if (si == 'l' && sj == 'l') {
      for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) {
        c = 0.0;
        for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) c += DD(j, jp) * R(i, j, ip, jp);
        A(i, ip) = c;
      }
    } else if (si == 'l' && sj == 'n') {
      for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) {
        c = 0.0;
        for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) c += DD(j, jp) * R(i, j, ip, jp);
        A(i, ip) = c;
      }
    } else if (si == 'n' && sj == 'l') {
      for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) {
        c = 0.0;
        for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) c += DD(j, jp) * R(i, j, ip, jp);
        A(i, ip) = c;
      }
    }
  } else if (icase == 2) {
    if (si == 's' && sj == 's') {
      for (jp = 1; jp <= ks; ++jp) for (j = 1; j <= ns - jp + 1; ++j) {
        c = 0.0;
        for (ip = 1; ip <= ks; ++ip) for (i = 1; i <= ns - ip + 1; ++i) c += DD(i, ip) * R(i, j, ip, jp);
        A(j, jp) = c;
      }
    } else if (si == 'n' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) {
        c = 0.0;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) c += DD(i, ip) * R(i, j, ip, jp);
        A(j, jp) = c;
      }
    } else if (si == 'l' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) {
        c = 0.0;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) c += DD(i, ip) * R(i, j, ip, jp);
        A(j, jp) = c;
      }
    } else if (si == 'l' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) {
        c = 0.0;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) c += DD(i, ip) * R(i, j, ip, jp);
        A(j, jp) = c;
      }
    } else if (si == 'n' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) {
        c = 0.0;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) c += DD(i, ip) * R(i, j, ip, jp);
        A(j, jp) = c;
      }
    }
  } else if (icase == 3) {
    if (si == 's' && sj == 's') {
      for (jp = 1; jp <= ks; ++jp) for (j = 1; j <= ns - jp + 1; ++j) { jj = j + jp - 1;
        for (ip = 1; ip <= ks; ++ip) for (i = 1; i <= ns - ip + 1; ++i) { ii = i + ip - 1;
          c = R(i, j, ip, jp);               A(i, jj) += c * DD(ii, j);
          if (ip > 1)                        A(ii, jj) += c * DD(i, j);
          if (jp > 1)                        A(i, j) += c * DD(ii, jj);
          if (ip > 1 && jp > 1)              A(ii, j) += c * DD(i, jj);
        } }
    } else if (si == 'l' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(i, jj) += c * DD(ii, j);
          if (ip != ks)                      A(ii, jj) += c * DD(i, j);
          if (jp != ks)                      A(i, j) += c * DD(ii, jj);
          if (ip != ks && jp != ks)          A(ii, j) += c * DD(i, jj);
        } }
    } else if (si == 'n' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(i, jj) += c * DD(ii, j);
        } }
    } else if (si == 'n' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(i, jj) += c * DD(ii, j);
          if (jp != ks)                      A(i, j) += c * DD(ii, jj);
        } }
    } else if (si == 'l' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(i, jj) += c * DD(ii, j);
          if (ip != ks)                      A(ii, jj) += c * DD(i, j);
        } }
    }
  } else if (icase == 4) {
    if (si == 's' && sj == 's') {
      for (jp = 1; jp <= ks; ++jp) for (j = 1; j <= ns - jp + 1; ++j) { jj = j + jp - 1;
        for (ip = 1; ip <= ks; ++ip) for (i = 1; i <= ns - ip + 1; ++i) { ii = i + ip - 1;
          c = R(i, j, ip, jp);               A(ii, j) += c * DD(i, jj);
          if (ip > 1)                        A(i, j) += c * DD(ii, jj);
          if (jp > 1)                        A(ii, jj) += c * DD(i, j);
          if (ip > 1 && jp > 1)              A(i, jj) += c * DD(ii, j);
        } }
    } else if (si == 'n' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(ii, j) += c * DD(i, jj);
        } }
    } else if (si == 'l' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(ii, j) += c * DD(i, jj);
          if (ip != ks)                      A(i, j) += c * DD(ii, jj);
          if (jp != ks)                      A(ii, jj) += c * DD(i, j);
          if (ip != ks && jp != ks)          A(i, jj) += c * DD(ii, j);
        } }
    } else if (si == 'n' && sj == 'l') {
      for (jp = 1; jp <= ks; ++jp) for (j = ks + 1 - jp; j <= ns; ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= 2 * ks - 1; ++ip) for (i = IMIN(ip); i <= IMAX(ip); ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(ii, j) += c * DD(i, jj);
          if (jp != ks)                      A(ii, jj) += c * DD(i, j);
        } }
    } else if (si == 'l' && sj == 'n') {
      for (jp = 1; jp <= 2 * ks - 1; ++jp) for (j = IMIN(jp); j <= IMAX(jp); ++j) { jj = j + jp - ks;
        for (ip = 1; ip <= ks; ++ip) for (i = ks + 1 - ip; i <= ns; ++i) { ii = i + ip - ks;
          c = R(i, j, ip, jp);               A(ii, j) += c * DD(i, jj);
          if (ip != ks)                      A(i, j) += c * DD(ii, jj);
        } }
    }
  }
#undef R
#undef A
#undef DD
#undef IMIN
#undef IMAX
}

/* SUM_AmB: DBSR_HF/dbsr_lib_dbs.f90:4027-4080 */
double orc_sum_amb(int ns, int ks, const double *a, const double *b, char sym) {
  double x = 0.0;
#define AB(i, j) (a[((i) - 1) + (size_t)ns * ((j) - 1)] * b[((i) - 1) + (size_t)ns * ((j) - 1)])
  if (sym == 'x') {
    for (int i = 1; i <= ns; ++i) for (int j = 1; j <= ns; ++j) x += AB(i, j);
  } else if (sym == 's') {
    for (int j = 1; j <= ks; ++j) for (int i = 1; i <= ns - j + 1; ++i) x += AB(i, j);
  } else if (sym == 'l') {
    for (int j = 1; j <= ks; ++j) for (int i = ks + 1 - j; i <= ns; ++i) x += AB(i, j);
  } else if (sym == 'n') {
    for (int j = 1; j <= 2 * ks - 1; ++j) {
      int imin = 1 > 1 + ks - j ? 1 : 1 + ks - j, imax = ns < ns + ks - j ? ns : ns + ks - j;
      for (int i = imin; i <= imax; ++i) x += AB(i, j);
    }
  }
#undef AB
  return x;
}

/* quadr: DBSR_HF/dbsr_lib_dbs.f90:3940-3976;  fi, fj are (ns,2) */
double orc_quadr(const orc_grid *g, const double *fi, const double *fj, int m) {
  const int nv = g->nv, ks = g->ks, ns = g->ns, nw = nv + 1;
  double q = 0.0, gx[MAXK], gw[MAXK];
  for (int comp = 0; comp < 2; ++comp) {
    int kx = comp == 0 ? g->ksp : g->ksq;
    const double *xb = comp == 0 ? g->pbsp : g->qbsp;
    for (int iv = 0; iv < nv; ++iv) {
      for (int mm = 0; mm < ks; ++mm) gx[mm] = ipow(g->gr[iv + (size_t)nv * mm], m) * g->grw[iv + (size_t)nv * mm];
      for (int ith = 0; ith < kx; ++ith) {
        double pi = fi[iv + ith + (size_t)ns * comp];
        for (int mm = 0; mm < ks; ++mm) gw[mm] = gx[mm] * xb[iv + (size_t)nw * (mm + (size_t)ks * ith)];
        for (int jth = 0; jth < kx; ++jth) {
          double pj = fj[iv + jth + (size_t)ns * comp];
          double s = 0.0;
          for (int mm = 0; mm < ks; ++mm) s += xb[iv + (size_t)nw * (mm + (size_t)ks * jth)] * gw[mm];
          q = q + s * pi * pj;
        }
      }
    }
  }
  return q;
}

/* unfold band/full storage into y(ns,ns) (UPDATE_HS, Full_mat_sym) */
static void unfold(int ns, int ks, const double *d, char sym, double *y) {
  memset(y, 0, sizeof(double) * (size_t)ns * ns);
#define Y(i, j) y[((i) - 1) + (size_t)ns * ((j) - 1)]
#define DV(i, j) d[((i) - 1) + (size_t)ns * ((j) - 1)]
  if (sym == 's') {
    for (int jp = 1; jp <= ks; ++jp) for (int i = 1; i <= ns - jp + 1; ++i) {
      int j = i + jp - 1; Y(i, j) = DV(i, jp); Y(j, i) = DV(i, jp);
    }
  } else if (sym == 'l') {
    for (int jp = 1; jp <= ks; ++jp) for (int i = ks + 1 - jp; i <= ns; ++i) {
      int j = i + jp - ks; Y(i, j) = DV(i, jp); Y(j, i) = DV(i, jp);
    }
  } else if (sym == 'n') {
    for (int jp = 1; jp <= 2 * ks - 1; ++jp) {
      int imin = 1 > 1 + ks - jp ? 1 : 1 + ks - jp, imax = ns < ns + ks - jp ? ns : ns + ks - jp;
      for (int i = imin; i <= imax; ++i) { int j = i + jp - ks; Y(i, j) = DV(i, jp); }
    }
  } else {
    memcpy(y, d, sizeof(double) * (size_t)ns * ns);
  }
#undef Y
#undef DV
}

/* UPDATE_HS (library version): DBSR_HF/dbsr_lib_dbs.f90:4137-4193 */
void orc_update_hs_lib(int ms, double *hm, int ii, int jj, int ns, int ks, const double *d, char sym) {
  double *y = (double *)malloc(sizeof(double) * (size_t)ns * ns);
  unfold(ns, ks, d, sym, y);
  int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  for (size_t q = 0; q < (size_t)ns * ns; ++q) y[q] = y[q] / 2.0;
  for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
    hm[(i1 + a) + (size_t)ms * (j1 + b)] += y[a + (size_t)ns * b];
  for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
    hm[(j1 + a) + (size_t)ms * (i1 + b)] += y[b + (size_t)ns * a];
  free(y);
}

/* UPDATE_HS (MAT3 channel-block version): DBSR_MAT3/updates.f90:110-179.
 * blk is the (ms,ms) block icc(ich,jch); ii,jj select the P/Q sub-block. */
void orc_update_hs_mat(int ns, int ks, double *blk, int ich, int jch, const double *d,
                       char sym, int ii, int jj) {
  const int ms = 2 * ns;
  double *y = (double *)malloc(sizeof(double) * (size_t)ns * ns);
  unfold(ns, ks, d, sym, y);
  int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  if (ich > jch) {
    for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
      blk[(i1 + a) + (size_t)ms * (j1 + b)] += y[a + (size_t)ns * b];
  } else if (jch > ich) {
    for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
      blk[(i1 + a) + (size_t)ms * (j1 + b)] += y[b + (size_t)ns * a];
  } else {
    for (size_t q = 0; q < (size_t)ns * ns; ++q) y[q] = y[q] / 2.0;
    for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
      blk[(i1 + a) + (size_t)ms * (j1 + b)] += y[a + (size_t)ns * b];
    for (int b = 0; b < ns; ++b) for (int a = 0; a < ns; ++a)
      blk[(j1 + a) + (size_t)ms * (i1 + b)] += y[b + (size_t)ns * a];
  }
  free(y);
}

/* rk: DBSR_HF/dbsr_lib_dbs.f90:3981-4022;  f* are (ns,2) */
double orc_rk(const orc_grid *g, const double *rpppp, const double *rqqqq,
              const double *rpqpq, const double *rqpqp,
              const double *f1, const double *f2, const double *f3, const double *f4) {
  const int ns = g->ns, ks = g->ks;
  size_t n = (size_t)ns * ks;
  double *d1 = (double *)malloc(sizeof(double) * n * 5);
  double *d2 = d1 + n, *d3 = d2 + n, *d4 = d3 + n, *cv = d4 + n;
  orc_density(ns, ks, d1, f1, f3, 's');
  orc_density(ns, ks, d2, f2, f4, 's');
  orc_density(ns, ks, d3, f1 + ns, f3 + ns, 's');
  orc_density(ns, ks, d4, f2 + ns, f4 + ns, 's');
  double r = 0.0;
  orc_convol(ns, ks, ks, rpppp, cv, d1, 2, 's', 's'); r += orc_sum_amb(ns, ks, cv, d2, 's');
  orc_convol(ns, ks, ks, rqqqq, cv, d3, 2, 's', 's'); r += orc_sum_amb(ns, ks, cv, d4, 's');
  orc_convol(ns, ks, ks, rqpqp, cv, d3, 2, 's', 's'); r += orc_sum_amb(ns, ks, cv, d2, 's');
  orc_convol(ns, ks, ks, rpqpq, cv, d1, 2, 's', 's'); r += orc_sum_amb(ns, ks, cv, d4, 's');
  free(d1);
  return r;
}

/* ===================================================================== */
/* c_data: DBSR_MAT3/MOD_c_data.f90:2-368.  Arrays are 1-based inside.    */
/* ===================================================================== */
struct orc_cdata {
  int ntype, kpol1, kpol2, mb, nb, kb, ncdata;
  double eps_cdata;
  double *CDATA;
  int *K1, *K2, *K3, *K4, *IPT;
  int *ipblk, *jpblk, *ipi, *ipj;
  int *iblk, *nblk, *kblk;  /* iblk(k,type), nblk(k,type), kblk(k,type,nb) */
  orc_gen_matrix_fn gen;
  void *user;
  int error;
};
#define IBLK(cd, k, t) (cd)->iblk[((k) - (cd)->kpol1) + ((cd)->kpol2 - (cd)->kpol1 + 1) * ((t) - 1)]
#define NBLK(cd, k, t) (cd)->nblk[((k) - (cd)->kpol1) + ((cd)->kpol2 - (cd)->kpol1 + 1) * ((t) - 1)]
#define KBLK(cd, k, t, n) (cd)->kblk[((k) - (cd)->kpol1) + ((cd)->kpol2 - (cd)->kpol1 + 1) * (((t) - 1) + (cd)->ntype * ((n) - 1))]

orc_cdata *orc_cdata_alloc(int nt, int kp1, int kp2, int mblock, int nblock, int kblock,
                           double eps_c, orc_gen_matrix_fn gen, void *user) {
  orc_cdata *cd = (orc_cdata *)calloc(1, sizeof(orc_cdata));
  cd->ntype = nt; cd->kpol1 = kp1; cd->kpol2 = kp2;
  cd->mb = mblock; cd->nb = nblock; cd->kb = kblock;
  int i = nt * (kp2 - kp1 + 1) * 2 + 1;
  if (cd->nb < i) cd->nb = i;
  cd->eps_cdata = eps_c;
  cd->gen = gen; cd->user = user;
  size_t m = (size_t)cd->mb * cd->nb;
  cd->CDATA = (double *)calloc(m + 1, sizeof(double));
  cd->K1 = (int *)calloc(m + 1, sizeof(int)); cd->K2 = (int *)calloc(m + 1, sizeof(int));
  cd->K3 = (int *)calloc(m + 1, sizeof(int)); cd->K4 = (int *)calloc(m + 1, sizeof(int));
  cd->IPT = (int *)calloc(m + 1, sizeof(int));
  cd->ipblk = (int *)calloc(cd->nb + 1, sizeof(int)); cd->jpblk = (int *)calloc(cd->nb + 1, sizeof(int));
  cd->ipi = (int *)calloc(cd->nb + 1, sizeof(int)); cd->ipj = (int *)calloc(cd->nb + 1, sizeof(int));
  int nk = kp2 - kp1 + 1;
  cd->iblk = (int *)calloc((size_t)nk * nt, sizeof(int));
  cd->nblk = (int *)calloc((size_t)nk * nt, sizeof(int));
  cd->kblk = (int *)calloc((size_t)nk * nt * cd->nb, sizeof(int));
  for (i = 1; i <= cd->nb; ++i) { cd->ipblk[i] = (i - 1) * cd->mb + 1; cd->jpblk[i] = -1; }
  i = 0;
  for (int k = kp1; k <= kp2; ++k)
    for (int j = 1; j <= nt; ++j) {
      i++;
      IBLK(cd, k, j) = i; NBLK(cd, k, j) = 1; KBLK(cd, k, j, 1) = i; cd->jpblk[i] = 0;
    }
  return cd;
}

void orc_cdata_free(orc_cdata *cd) {
  if (!cd) return;
  free(cd->CDATA); free(cd->K1); free(cd->K2); free(cd->K3); free(cd->K4); free(cd->IPT);
  free(cd->ipblk); free(cd->jpblk); free(cd->ipi); free(cd->ipj);
  free(cd->iblk); free(cd->nblk); free(cd->kblk); free(cd);
}

/* Add_cdata: binary-search insert into the sorted sub-list ip..jp */
static int add_cdata(orc_cdata *cd, int ip, int *jp, double C, int j1, int j2, int j3, int j4) {
  if (*jp < ip) {
    cd->CDATA[ip] = C; cd->K1[ip] = j1; cd->K2[ip] = j2; cd->K3[ip] = j3; cd->K4[ip] = j4;
    *jp = ip;
    return 0;
  }
  int k = ip, l = *jp, m;
  while (k <= l) {
    m = (k + l) / 2;
    if (j1 < cd->K1[m]) l = m - 1;
    else if (j1 > cd->K1[m]) k = m + 1;
    else if (j2 < cd->K2[m]) l = m - 1;
    else if (j2 > cd->K2[m]) k = m + 1;
    else if (j3 < cd->K3[m]) l = m - 1;
    else if (j3 > cd->K3[m]) k = m + 1;
    else if (j4 < cd->K4[m]) l = m - 1;
    else if (j4 > cd->K4[m]) k = m + 1;
    else { cd->CDATA[m] = cd->CDATA[m] + C; return 0; }
  }
  if (*jp - ip + 1 >= cd->mb) return -1;
  for (int i = *jp; i >= k; --i) {
    int ii = i + 1;
    cd->CDATA[ii] = cd->CDATA[i];
    cd->K1[ii] = cd->K1[i]; cd->K2[ii] = cd->K2[i]; cd->K3[ii] = cd->K3[i]; cd->K4[ii] = cd->K4[i];
  }
  cd->CDATA[k] = C; cd->K1[k] = j1; cd->K2[k] = j2; cd->K3[k] = j3; cd->K4[k] = j4;
  *jp = *jp + 1;
  return 0;
}

/* Merge_cdata: k-way merge of sorted blocks, summing equal keys in block order */
static int merge_cdata(orc_cdata *cd, int nn, int *IP, int *JP) {
  int nc = 0, mm = 0, i, j, m;
  for (m = 1; m <= nn; ++m) { if (JP[m] <= 0) continue; mm = m; break; }
  if (mm == 0) return 0;
  for (;;) {
    int merged = 0;
    for (int ii = 1; ii <= nn - 1 && !merged; ++ii) {
      if (JP[ii] <= 0) continue;
      i = IP[ii];
      for (int jj = ii + 1; jj <= nn; ++jj) {
        if (JP[jj] <= 0) continue;
        j = IP[jj];
        if (cd->K1[i] != cd->K1[j] || cd->K2[i] != cd->K2[j] || cd->K3[i] != cd->K3[j] ||
            cd->K4[i] != cd->K4[j]) continue;
        cd->CDATA[i] = cd->CDATA[i] + cd->CDATA[j];
        cd->CDATA[j] = 0.0;
        mm = jj; merged = 1;
        break;
      }
    }
    if (!merged) {
      j = IP[mm];
      for (m = 1; m <= nn; ++m) {
        if (JP[m] == 0) continue;
        i = IP[m];
        if (cd->K1[i] < cd->K1[j]) { mm = m; j = i; }
        else if (cd->K1[i] > cd->K1[j]) continue;
        else if (cd->K2[i] < cd->K2[j]) { mm = m; j = i; }
        else if (cd->K2[i] > cd->K2[j]) continue;
        else if (cd->K3[i] < cd->K3[j]) { mm = m; j = i; }
        else if (cd->K3[i] > cd->K3[j]) continue;
        else if (cd->K4[i] < cd->K4[j]) { mm = m; j = i; }
      }
      if (fabs(cd->CDATA[j]) > cd->eps_cdata) { nc++; cd->IPT[nc] = IP[mm]; }
    }
    IP[mm] = IP[mm] + 1;
    if (IP[mm] <= JP[mm]) continue;
    JP[mm] = 0;
    int found = 0;
    for (m = 1; m <= nn; ++m) if (JP[m] > 0) { mm = m; found = 1; break; }
    if (!found) break;
  }
  return nc;
}

int orc_cdata_add_matrix(orc_cdata *cd, int itype, int kpol) {
  int n = NBLK(cd, kpol, itype);
  int i = KBLK(cd, kpol, itype, 1), j, nc;
  if (n == 1 && cd->jpblk[i] <= 0) n = 0;
  if (n <= 0) {
    nc = 0;
    cd->ncdata = 0;
    return 0;  /* reference returns without calling Gen_matrix */
  } else if (n == 1) {
    nc = cd->jpblk[i] - cd->ipblk[i] + 1;
    j = cd->ipblk[i] - 1;
    for (i = 1; i <= nc; ++i) cd->IPT[i] = j + i;
  } else {
    for (i = 1; i <= n; ++i) {
      j = KBLK(cd, kpol, itype, i);
      cd->ipi[i] = cd->ipblk[j];
      cd->ipj[i] = cd->jpblk[j];
    }
    nc = merge_cdata(cd, n, cd->ipi, cd->ipj);
    for (i = 1; i <= n; ++i) cd->jpblk[KBLK(cd, kpol, itype, i)] = -1;
  }
  cd->ncdata = nc;
  i = KBLK(cd, kpol, itype, 1);
  IBLK(cd, kpol, itype) = i; NBLK(cd, kpol, itype) = 1; cd->jpblk[i] = 0;
  if (cd->gen && nc > 0) {
    double *c = (double *)malloc(sizeof(double) * nc);
    int *k1 = (int *)malloc(sizeof(int) * nc * 4);
    int *k2 = k1 + nc, *k3 = k2 + nc, *k4 = k3 + nc;
    for (int q = 0; q < nc; ++q) {
      int p = cd->IPT[q + 1];
      c[q] = cd->CDATA[p]; k1[q] = cd->K1[p]; k2[q] = cd->K2[p]; k3[q] = cd->K3[p]; k4[q] = cd->K4[p];
    }
    cd->gen(cd->user, itype, kpol, nc, c, k1, k2, k3, k4);
    free(c); free(k1);
  } else if (cd->gen) {
    cd->gen(cd->user, itype, kpol, 0, NULL, NULL, NULL, NULL, NULL);
  }
  return nc;
}

int orc_cdata_add_coef(orc_cdata *cd, double C, int kpol, int j1, int j2, int j3, int j4, int itype) {
  int i = IBLK(cd, kpol, itype), ip = cd->ipblk[i], jp = cd->jpblk[i];
  if (add_cdata(cd, ip, &jp, C, j1, j2, j3, j4) != 0) return -1;
  cd->jpblk[i] = jp;
  if (jp - ip + 1 < cd->mb) return 0;
  int m = 0;
  if (NBLK(cd, kpol, itype) < cd->kb) {
    for (i = 1; i <= cd->nb; ++i) {
      if (cd->jpblk[i] >= 0) continue;
      IBLK(cd, kpol, itype) = i; cd->jpblk[i] = 0;
      NBLK(cd, kpol, itype) = NBLK(cd, kpol, itype) + 1;
      KBLK(cd, kpol, itype, NBLK(cd, kpol, itype)) = i;
      m = 1; break;
    }
    if (m != 0) return 0;
  }
  int jtype = itype, jpol = kpol, n = NBLK(cd, kpol, itype);
  for (i = 1; i <= cd->ntype; ++i)
    for (int k = cd->kpol1; k <= cd->kpol2; ++k) {
      if (NBLK(cd, k, i) <= n) continue;
      jtype = i; jpol = k; n = NBLK(cd, jpol, jtype);
    }
  orc_cdata_add_matrix(cd, jtype, jpol);
  if (jtype == itype && jpol == kpol) return 1;
  m = 0;
  for (i = 1; i <= cd->nb; ++i) {
    if (cd->jpblk[i] >= 0) continue;
    IBLK(cd, kpol, itype) = i; cd->jpblk[i] = 0;
    NBLK(cd, kpol, itype) = NBLK(cd, kpol, itype) + 1;
    KBLK(cd, kpol, itype, NBLK(cd, kpol, itype)) = i;
    m = 1; break;
  }
  if (m == 0) return -2;  /* 'Add_coef: problems with new block' */
  return 1;
}

/* ===================================================================== */
/* Rk_evaluate / Sk_evaluate: DBSR_CI3/rk_evaluate.f90:2-63,              */
/* sk_evaluate.f90:2-63.  List in IPT order; pq is (ns,2,norb) with       */
/* 1-based orbital indices inside K1/K2.  R^k types 1..4 use 's' band,    */
/* S^k types 5..6 use 'n' band.  HM(ldhm,*) and DM are 1-based by ic/jc.  */
/* ===================================================================== */
void orc_rk_evaluate(int ns, int ks, int kb, const double *rkb, int itype,
                     int nc, const double *cdata, const int *K1, const int *K2,
                     const int *K3, const int *K4, double eps_c, const double *pq, int ibi,
                     double *HM, int ldhm, int nzero, double *DM) {
  int ip1, ip2, jp1, jp2;
  char sym;
  switch (itype) {
    case 1: ip1 = 1; ip2 = 1; jp1 = 1; jp2 = 1; sym = 's'; break;
    case 2: ip1 = 2; ip2 = 2; jp1 = 2; jp2 = 2; sym = 's'; break;
    case 3: ip1 = 1; ip2 = 1; jp1 = 2; jp2 = 2; sym = 's'; break;
    case 4: ip1 = 2; ip2 = 2; jp1 = 1; jp2 = 1; sym = 's'; break;
    case 5: ip1 = 1; ip2 = 2; jp1 = 1; jp2 = 2; sym = 'n'; break;
    default: ip1 = 1; ip2 = 2; jp1 = 2; jp2 = 1; sym = 'n'; break;
  }
  int nb = (sym == 'n') ? 2 * ks - 1 : ks;
  double *deni = (double *)malloc(sizeof(double) * (size_t)ns * nb * 3);
  double *denj = deni + (size_t)ns * nb, *conv = denj + (size_t)ns * nb;
#define PQ(ip, orb) (pq + (size_t)ns * (((ip) - 1) + 2 * (size_t)((orb) - 1)))
  int ii = 0, jj = 0;
  double S = 0.0;
  for (int j = 0; j < nc; ++j) {
    if (fabs(cdata[j]) < eps_c) continue;
    if (K1[j] != ii) {
      int i1 = K1[j] / ibi, i2 = K1[j] % ibi;
      orc_density(ns, ks, deni, PQ(ip1, i1), PQ(ip2, i2), sym);
      orc_convol(ns, ks, kb, rkb, conv, deni, 2, sym, sym);
    }
    if (K2[j] != jj) {
      int j1 = K2[j] / ibi, j2 = K2[j] % ibi;
      orc_density(ns, ks, denj, PQ(jp1, j1), PQ(jp2, j2), sym);
    }
    if (K1[j] != ii || K2[j] != jj) S = orc_sum_amb(ns, ks, conv, denj, sym);
    ii = K1[j]; jj = K2[j];
    double SS = S * cdata[j];
    int ic = K3[j], jc = K4[j];
    if (jc <= nzero) HM[(ic - 1) + (size_t)ldhm * (jc - 1)] += SS;
    else if (ic == jc) DM[jc - 1] += SS;
  }
#undef PQ
  free(deni);
}

/* ===================================================================== */
/* R_data itype 3 (direct) and 4 (exchange): DBSR_MAT3/R_data.f90:214-277 */
/* atype 0..3 selects mrk_* and the P/Q components (R_data.f90:76-86).    */
/* hch is (ms,ms,nblocks); icc(ich,jch) (nch x nch, column-major, 1-based */
/* values, 0 = absent) gives the block.                                    */
/* ===================================================================== */
void orc_rdata_34(int ns, int ks, const double *rkb, int itype, int atype,
                  int nc, const double *cdata, const int *K1, const int *K2,
                  const int *K3, const double *pq, int ibo, const int *icc,
                  int nch, double *hch) {
  int ip1, ip2, jp1, jp2, icase;
  char sym_d, sym_r;
  if (itype == 3) { icase = 1; sym_d = 's'; sym_r = 's'; }
  else { icase = 3; sym_d = 'x'; sym_r = 'x'; }
  switch (atype) {
    case 0: ip1 = 1; ip2 = 1; jp1 = 1; jp2 = 1; break;
    case 1: ip1 = 2; ip2 = 2; jp1 = 2; jp2 = 2; break;
    case 2: ip1 = 1; ip2 = 1; jp1 = 2; jp2 = 2;
            if (itype == 4) { ip1 = 1; ip2 = 2; jp1 = 1; jp2 = 2; } break;
    default: ip1 = 2; ip2 = 2; jp1 = 1; jp2 = 1;
             if (itype == 4) { ip1 = 2; ip2 = 1; jp1 = 2; jp2 = 1; } break;
  }
  const int ms = 2 * ns;
  size_t nn = (size_t)ns * ns;
  double *xx = (double *)calloc(nn * 3, sizeof(double));
  double *dd = xx + nn, *cv = dd + nn;
#define PQ(ip, orb) (pq + (size_t)ns * (((ip) - 1) + 2 * (size_t)((orb) - 1)))
  for (int j = 0; j < nc; ++j) {
    int j1 = K2[j], j2 = K3[j];
    orc_density(ns, ks, dd, PQ(jp1, j1), PQ(jp2, j2), sym_d);
    size_t cols = (sym_d == 's') ? (size_t)ns * ks : nn;
    for (size_t q = 0; q < cols; ++q) xx[q] = xx[q] + cdata[j] * dd[q];
    int last = 1;
    if (j < nc - 1 && K1[j] == K1[j + 1]) last = 0;
    if (!last) continue;
    orc_convol(ns, ks, ks, rkb, cv, xx, icase, 's', 's');
    int ich = K1[j] / ibo, jch = K1[j] % ibo;
    int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
    if (b > 0) orc_update_hs_mat(ns, ks, hch + (size_t)ms * ms * (b - 1), ich, jch, cv, sym_r, ip1, ip2);
    memset(xx, 0, sizeof(double) * nn);
  }
#undef PQ
  free(xx);
}

/* Asymptotic coefficients of R_data itype 3 with atype 0/1
 * (DBSR_MAT3/R_data.f90:214-240): walking the list in IPT order, CA =
 * CA + c*QUADR_PQ(K2,K3,kpol); at the end of each run of equal K1,
 * UPDATE_ACF(kpol,ich,jch,CA) (DBSR_MAT3/updates.f90:270-292: acf(i,k) += C
 * with i = icc(ich,jch); blocks not stored here are skipped) and CA = 0.
 * qpq(norb,norb) holds QUADR_PQ(i,j,kpol) of the host program.  acf is
 * (nblocks, 0:mk) column-major.  Returns 0, or 1 for k > mk ('Update_cf:
 * k > kmk'). */
int orc_rdata_acf(int nc, const double *cdata, const int *K1, const int *K2, const int *K3, int ibo,
                  int kpol, int mk, const double *qpq, int norb, const int *icc, int nch, int nblocks,
                  double *acf) {
  if (kpol > mk) return 1;
  double CA = 0.0;
  for (int j = 0; j < nc; ++j) {
    const double S = qpq[(K2[j] - 1) + (size_t)norb * (K3[j] - 1)];
    CA = CA + cdata[j] * S;
    if (j < nc - 1 && K1[j] == K1[j + 1]) continue;
    const int ich = K1[j] / ibo, jch = K1[j] % ibo;
    const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
    if (b > 0) acf[(b - 1) + (size_t)nblocks * kpol] += CA;
    CA = 0.0;
  }
  return 0;
}

/* Idef_Rtype for the channel-channel structures (ic=jc=io=jo=0):
 * DBSR_MAT3/idef_type.f90:142-302.  ipbs(orbital) > 0 marks a channel
 * orbital (1-based orbital index; ipbs[0] unused).  Returns itype (3 or 4),
 * or 0 when the reference would stop with 'Idef_Rtype: itype=0'. */
int orc_idef_rtype(int j1, int j2, int j3, int j4, const int *ipbs, int ibo,
                   int *k1, int *k2, int *k3, int *k4) {
  int i1, i2, i3, i4, t, itype = 0;
  if (ipbs[j1] + ipbs[j2] == 0 || ipbs[j3] + ipbs[j4] == 0) {
    i1 = j1; i3 = j3; if (j1 > j3) { i1 = j3; i3 = j1; }
    i2 = j2; i4 = j4; if (j2 > j4) { i2 = j4; i4 = j2; }
    if (i1 > i2) { t = i1; i1 = i2; i2 = t; t = i3; i3 = i4; i4 = t; }
  } else {
    i1 = j1; i2 = j2; i3 = j3; i4 = j4;
  }
  if (ipbs[i1] > 0 && ipbs[i3] > 0) {
    if (ipbs[i1] >= ipbs[i3]) {
      *k1 = ipbs[i1] * ibo + ipbs[i3]; *k2 = i2; *k3 = i4; *k4 = 0; itype = 3;
      if (*k2 > *k3) { *k2 = i4; *k3 = i2; }
    } else {
      *k1 = ipbs[i3] * ibo + ipbs[i1]; *k2 = i4; *k3 = i2; *k4 = 0; itype = 3;
      if (*k2 > *k3) { *k2 = i2; *k3 = i4; }
    }
  } else if (ipbs[i2] > 0 && ipbs[i4] > 0) {
    if (ipbs[i2] >= ipbs[i4]) {
      *k1 = ipbs[i2] * ibo + ipbs[i4]; *k2 = i1; *k3 = i3; *k4 = 0; itype = 3;
      if (*k2 > *k3) { *k2 = i3; *k3 = i1; }
    } else {
      *k1 = ipbs[i4] * ibo + ipbs[i2]; *k2 = i3; *k3 = i1; *k4 = 0; itype = 3;
      if (*k2 > *k3) { *k2 = i1; *k3 = i3; }
    }
  } else if (ipbs[i1] > 0 && ipbs[i4] > 0) {
    if (ipbs[i1] >= ipbs[i4]) { *k1 = ipbs[i1] * ibo + ipbs[i4]; *k2 = i3; *k3 = i2; *k4 = 0; itype = 4; }
    else { *k1 = ipbs[i4] * ibo + ipbs[i1]; *k2 = i2; *k3 = i3; *k4 = 0; itype = 4; }
  } else if (ipbs[i2] > 0 && ipbs[i3] > 0) {
    if (ipbs[i3] >= ipbs[i2]) { *k1 = ipbs[i3] * ibo + ipbs[i2]; *k2 = i1; *k3 = i4; *k4 = 0; itype = 4; }
    else { *k1 = ipbs[i2] * ibo + ipbs[i3]; *k2 = i4; *k3 = i1; *k4 = 0; itype = 4; }
  }
  return itype;
}

/* Decode_int: DBSR_CI3/setup_matrix.f90:206-223 */
void orc_decode_int(int code, int *met, int *k, int *i1, int *i2, int *i3, int *i4) {
  int ii = code;
  *met = ii % 4; ii /= 4;
  *k = ii % 1024; ii /= 1024;
  *i4 = ii % 32; ii /= 32;
  *i3 = ii % 32; ii /= 32;
  *i2 = ii % 32; *i1 = ii / 32;
}

/* l_kappa: DBSR_HF/dbsr_lib_zcom.f90:500-512 */
int orc_l_kappa(int kappa) { return kappa > 0 ? kappa : -kappa - 1; }

/* SMU: Breit factors of Grant and Pyper, DBSR_HF/hf_energy_BR.f90:58-121 */
double orc_smu(int ka, int kb, int kc, int kd, int L, int v, double *S) {
  double smu = 0.0;
  for (int i = 0; i < 8; ++i) S[i] = 0.0;
  if (L < 0 || v < 0) return 0.0;
  int la = orc_l_kappa(ka), lc = orc_l_kappa(kc);
  if ((la + lc + v) % 2 == 0) return 0.0;
  int lb = orc_l_kappa(kb), ld = orc_l_kappa(kd);
  if ((lb + ld + v) % 2 == 0) return 0.0;
  if (v == L) {
    if (L == 0) return 0.0;
    smu = (double)(-(ka + kc) * (kb + kd));
    smu = smu / L / (L + 1);
    for (int i = 0; i < 8; ++i) S[i] = smu;
    return smu;
  }
  int K = kc - ka, Kp = kd - kb;
  if (v == L - 1) {
    double bp = L + 1; bp = bp / 2 / (L + L - 1);
    double cp = -(L - 2); cp = cp / 2 / L / (L + L - 1);
    S[0] = (L + K) * (bp + cp * Kp);
    S[1] = (L + Kp) * (bp + cp * K);
    S[2] = (L - K) * (bp - cp * Kp);
    S[3] = (L - Kp) * (bp - cp * K);
    S[4] = -(L + K) * (bp - cp * Kp);
    S[5] = -(L - Kp) * (bp + cp * K);
    S[6] = -(L - K) * (bp + cp * Kp);
    S[7] = -(L + Kp) * (bp - cp * K);
  } else if (v == L + 1) {
    int i = -L - 1;
    double b = L; b = b / 2 / (L + L + 3);
    double c = (L + 3); c = c / (L + L + 2) / (L + L + 3);
    S[0] = (i + Kp) * (b + c * K);
    S[1] = (i + K) * (b + c * Kp);
    S[2] = (i - Kp) * (b - c * K);
    S[3] = (i - K) * (b - c * Kp);
    S[4] = -(i - Kp) * (b + c * K);
    S[5] = -(i + K) * (b - c * Kp);
    S[6] = -(i + Kp) * (b - c * K);
    S[7] = -(i - K) * (b + c * Kp);
  }
  for (int i = 0; i < 8; ++i) smu = smu + fabs(S[i]);
  return smu;
}

/* Add_integral, CI3: DBSR_CI3/add_integral.f90:2-43.  kappa[i] of 1-based
 * orbital i (kappa[0] unused); returns the reference's Add_coef calls in
 * order through add_coef. */
int orc_add_integral_ci(orc_cdata *cd, int kpol, int ic, int jc, int i1, int i2, int i3, int i4,
                        double C, const int *kappa, int ibi, int mbreit) {
  int rc = 0;
  int m = (orc_l_kappa(kappa[i1]) + orc_l_kappa(kappa[i3]) + kpol) % 2 +
          (orc_l_kappa(kappa[i2]) + orc_l_kappa(kappa[i4]) + kpol) % 2;
  if (m == 0)
    for (int t = 1; t <= 4; ++t)
      if (orc_cdata_add_coef(cd, C, kpol, i1 * ibi + i3, i2 * ibi + i4, ic, jc, t) < 0) rc = -1;
  if (mbreit == 0) return rc;
  double S[8];
  for (int v = kpol - 1; v <= kpol + 1; ++v) {
    if (orc_smu(kappa[i1], kappa[i2], kappa[i3], kappa[i4], kpol, v, S) == 0.0) continue;
    int r = 0;
    r |= orc_cdata_add_coef(cd, C * S[0], v, i1 * ibi + i3, i2 * ibi + i4, ic, jc, 5);
    r |= orc_cdata_add_coef(cd, C * S[1], v, i2 * ibi + i4, i1 * ibi + i3, ic, jc, 5);
    r |= orc_cdata_add_coef(cd, C * S[2], v, i3 * ibi + i1, i4 * ibi + i2, ic, jc, 5);
    r |= orc_cdata_add_coef(cd, C * S[3], v, i4 * ibi + i2, i3 * ibi + i1, ic, jc, 5);
    r |= orc_cdata_add_coef(cd, C * S[4], v, i1 * ibi + i3, i2 * ibi + i4, ic, jc, 6);
    r |= orc_cdata_add_coef(cd, C * S[5], v, i4 * ibi + i2, i3 * ibi + i1, ic, jc, 6);
    r |= orc_cdata_add_coef(cd, C * S[6], v, i3 * ibi + i1, i4 * ibi + i2, ic, jc, 6);
    r |= orc_cdata_add_coef(cd, C * S[7], v, i2 * ibi + i4, i1 * ibi + i3, ic, jc, 6);
    if (r < 0) rc = -1;
  }
  return rc;
}

/* ===================================================================== */
/* Overlaps, unfolding and the MAT3 matrix updates                        */
/* ===================================================================== */

/* Full_mat_sym: DBSR_HF/dbsr_lib_zcom.f90:1247-1294 ('s','l','n','x') */
void orc_full_mat_sym(int ns, int ks, const double *x, double *y, char sym) {
  unfold(ns, ks, x, sym, y);
}

/* ZINTYM: DBSR_HF/dbsr_lib_dbs.f90:4472-4497.  b1(nv+1,ks,ks1),
 * b2(nv+1,ks,ks2), ygr(nv,ks), ym(mm,mm). */
void orc_zintym(int nv, int ks, int ks1, int ks2, const double *b1, const double *b2,
                const double *ygr, int mm, double *ym) {
  const size_t nw = (size_t)nv + 1;
  memset(ym, 0, sizeof(double) * (size_t)mm * mm);
  for (int iv = 1; iv <= nv; ++iv)
    for (int ith = 1; ith <= ks1; ++ith) {
      const int i = iv + ith - 1;
      for (int jth = 1; jth <= ks2; ++jth) {
        const int j = iv + jth - 1;
        double s = 0.0;
        for (int m = 1; m <= ks; ++m)
          s += ygr[(iv - 1) + (size_t)nv * (m - 1)] * b1[(iv - 1) + nw * ((m - 1) + (size_t)ks * (ith - 1))] *
               b2[(iv - 1) + nw * ((m - 1) + (size_t)ks * (jth - 1))];
        ym[(i - 1) + (size_t)mm * (j - 1)] += s;
      }
    }
}

/* Def_pbsp / Def_qbsp overlaps and Def_pq_arrays' fppqq:
 * DBSR_HF/dbsr_lib_dbs.f90:1026, 1072, 1098-1101 */
void orc_overlaps(const orc_grid *g, double *fpbs, double *fqbs, double *fppqq) {
  const int ns = g->ns, ms = g->ms;
  double *p = (double *)malloc(sizeof(double) * (size_t)ns * ns * 2), *q = p + (size_t)ns * ns;
  orc_zintym(g->nv, g->ks, g->ksp, g->ksp, g->pbsp, g->pbsp, g->grw, ns, p);
  orc_zintym(g->nv, g->ks, g->ksq, g->ksq, g->qbsp, g->qbsp, g->grw, ns, q);
  if (fpbs) memcpy(fpbs, p, sizeof(double) * (size_t)ns * ns);
  if (fqbs) memcpy(fqbs, q, sizeof(double) * (size_t)ns * ns);
  if (fppqq) {
    memset(fppqq, 0, sizeof(double) * (size_t)ms * ms);
    for (int j = 0; j < ns; ++j)
      for (int i = 0; i < ns; ++i) {
        fppqq[i + (size_t)ms * j] = p[i + (size_t)ns * j];
        fppqq[(ns + i) + (size_t)ms * (ns + j)] = q[i + (size_t)ns * j];
      }
  }
  free(p);
}

/* UPDATE_HX: DBSR_MAT3/updates.f90:79-106 on one block (ms,ms) */
void orc_update_hx(int ms, double *blk, int ich, int jch, const double *d, double C) {
  for (int j = 0; j < ms; ++j)
    for (int i = 0; i < ms; ++i) {
      const double xx = ich >= jch ? C * d[i + (size_t)ms * j] : C * d[j + (size_t)ms * i];
      blk[i + (size_t)ms * j] = blk[i + (size_t)ms * j] + xx;
    }
}

/* UPDATE_HW: DBSR_MAT3/updates.f90:183-217 on one block (ms,ms) */
void orc_update_hw(int ms, double *blk, int ich, int jch, const double *v, const double *w, double C) {
  double *x = (double *)malloc(sizeof(double) * (size_t)ms * ms * 2), *xx = x + (size_t)ms * ms;
  for (int i = 0; i < ms; ++i) for (int j = 0; j < ms; ++j) x[i + (size_t)ms * j] = v[i] * w[j];
  for (int j = 0; j < ms; ++j)
    for (int i = 0; i < ms; ++i) {
      const double xij = x[i + (size_t)ms * j], xji = x[j + (size_t)ms * i];
      if (ich > jch) xx[i + (size_t)ms * j] = C * xij;
      else if (jch > ich) xx[i + (size_t)ms * j] = C * xji;
      else xx[i + (size_t)ms * j] = (C / 2.0) * (xij + xji);
    }
  for (size_t q = 0; q < (size_t)ms * ms; ++q) blk[q] = blk[q] + xx[q];
  free(x);
}

/* UPDATE_HV: DBSR_MAT3/updates.f90:221-244 on one column */
void orc_update_hv(int ms, double *col, const double *v, double C) {
  for (int i = 0; i < ms; ++i) col[i] = col[i] + C * v[i];
}

/* O_data: DBSR_MAT3/O_data.f90:2-93, one list in IPT order.  Matrix tables
 * as the MAT3 storage: icc(nch,nch), icb(nch,npert), ibb(npert,npert)
 * (1-based block numbers, 0 = not stored), hch(ms,ms,*), hcp(ms,*), hp(*).
 * kbort(i1,i2) gives the 1-based column of vch(ms,*).  ipbs, iptar are
 * 1-based tables (element 0 unused).  Returns 0, or a nonzero code where
 * the reference stops. */
int orc_o_data(int ms, int itype, int n, const int *k1, const int *k2, const double *c, double eps_c,
               int ibo, const int *ipbs, orc_kbort_fn kbort, void *user, const double *vch,
               const double *fppqq, int iitar, const int *iptar, double *otarg,
               int nch, int npert, const int *icc, double *hch, const int *icb, double *hcp,
               const int *ibb, double *hp) {
  for (int e = 0; e < n; ++e) {
    const double C = c[e];
    if (fabs(C) < eps_c) continue;
    if (itype == 1) {
      const int ic = k1[e], jc = k2[e];
      if (ic > npert || jc > npert || ic <= 0 || jc <= 0) return 1;
      const int i = ibb[(ic - 1) + (size_t)npert * (jc - 1)];
      if (i > 0) hp[i - 1] = hp[i - 1] + C;
    } else if (itype == 2) {
      const int io = k1[e], jc = k2[e], i1 = io / ibo, i2 = io % ibo;
      const int k = kbort(user, i1, i2);
      if (k == 0) return 2;
      const int ich = ipbs[i1];
      if (ich < 1 || ich > nch || jc < 1 || jc > npert) return 3;
      const int i = icb[(ich - 1) + (size_t)nch * (jc - 1)];
      if (i > 0) orc_update_hv(ms, hcp + (size_t)ms * (i - 1), vch + (size_t)ms * (k - 1), C);
    } else if (itype == 3) {
      const int io = k1[e], jo = k2[e];
      const int i1 = io / ibo, i2 = io % ibo, j1 = jo / ibo, j2 = jo % ibo;
      const int ki = kbort(user, i1, i2);
      if (ki == 0) return 2;
      const int kj = kbort(user, j1, j2);
      if (kj == 0) return 2;
      const int ich = ipbs[i1], jch = ipbs[j1];
      if (ich < 1 || jch < 1 || ich > nch || jch > nch) return 4;
      const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
      if (b > 0)
        orc_update_hw(ms, hch + (size_t)ms * ms * (b - 1), ich, jch, vch + (size_t)ms * (ki - 1),
                      vch + (size_t)ms * (kj - 1), C);
    } else if (itype == 4) {
      const int io = k1[e], i1 = io / ibo, i2 = io % ibo;
      const int ich = ipbs[i1], jch = ipbs[i2];
      if (ich < 1 || jch < 1 || ich > nch || jch > nch) return 4;
      const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
      if (b > 0 && otarg) {  /* Target_h(ich,jch,0,C): DBSR_MAT3/target_subs.f90:2-31 */
        const int it = iptar[ich], jt = iptar[jch];
        const int hi = it > jt ? it : jt, lo = it > jt ? jt : it;
        otarg[(size_t)(hi - 1) * hi / 2 + lo - 1] += C;
      }
      if (iitar > 1 && ich != jch && b > 0) orc_update_hx(ms, hch + (size_t)ms * ms * (b - 1), ich, jch, fppqq, C);
    } else {
      return 5;
    }
  }
  return 0;
}

/* BAV: called at DBSR_MAT3/R_data.f90:181 and S_data.f90:205; its source
 * lives in the un-vendored Libraries_03.  Restated from its name and call
 * sites: A(ns,ns) unfolded from storage sym as Full_mat_sym does, then
 * c = A b (dir 'r') or c = A^T b (dir 'l'), each element summed over the
 * full index range in ascending order.  Pinned through S_data structures
 * 2-5 ('n' storage, both directions) and R_data structure 2 ('s'): their
 * vectors contracted with a channel function equal the compiled
 * reference's sk_ppqq / sk_pqqp and rk() with it in the channel slot. */
void orc_bav(int ns, int ks, const double *a, const double *b, double *c, char sym, char dir) {
  double *y = (double *)malloc(sizeof(double) * (size_t)ns * ns);
  unfold(ns, ks, a, sym, y);
  for (int i = 0; i < ns; ++i) {
    double s = 0.0;
    for (int j = 0; j < ns; ++j) s += (dir == 'l' ? y[j + (size_t)ns * i] : y[i + (size_t)ns * j]) * b[j];
    c[i] = s;
  }
  free(y);
}

/* component and case table of S_data (DBSR_MAT3/S_data.f90:73-106):
 * [atype][itype-1] = {ip1, ip2, jp1, jp2, int, sym_v} */
static const int s_tab[2][9][6] = {
  {{1,2,1,2,2,0}, {1,2,1,2,1,'r'}, {1,2,1,2,2,'r'}, {2,1,1,2,1,'l'}, {2,1,1,2,2,'l'},
   {1,2,1,2,1,0}, {1,2,1,2,2,0}, {1,2,2,1,3,0}, {2,1,1,2,4,0}},
  {{1,2,2,1,2,0}, {1,2,2,1,1,'r'}, {2,1,1,2,2,'r'}, {2,1,2,1,1,'l'}, {1,2,1,2,2,'l'},
   {1,2,2,1,1,0}, {2,1,1,2,2,0}, {1,1,2,2,3,0}, {2,2,1,1,4,0}}};

/* R_data (DBSR_MAT3/R_data.f90:2-290, stype 0) and S_data (S_data.f90:2-287,
 * stype 1) for one merged list in IPT order.  rkb is the array of the
 * (type,k): R types atype+1 (ks band), S types 5+atype (2ks-1 band).
 * MAT3 storage as orc_o_data; htarg: packed target interactions
 * (Target_h(ich,jch,C,0)), may be NULL.  Returns 0, or a nonzero code where
 * the reference stops: 1 HB range, 2 KBORT=0, 3 HV range, 4 channel range,
 * 6/7 unknown structure. */
int orc_rs_data(int ns, int ks, const double *rkb, int stype, int itype, int atype, int nc,
                const double *cdata, const int *K1, const int *K2, const int *K3, const int *K4,
                const double *pq, int ibo, const int *ipbs, orc_kbort_fn kbort, void *user,
                const double *vch, const double *fppqq, int iitar, const int *iptar, double *htarg,
                int nch, int npert, const int *icc, double *hch, const int *icb, double *hcp,
                const int *ibb, double *hp) {
  const int ms = 2 * ns;
  int ip1 = 1, ip2 = 1, jp1 = 1, jp2 = 1, icase = 1, kb;
  char sym_i, sym_j, sym_d, sym_r, sym_v = 'r';
  if (stype == 0) {
    if (itype == 3 || itype == 4) {
      orc_rdata_34(ns, ks, rkb, itype, atype, nc, cdata, K1, K2, K3, pq, ibo, icc, nch, hch);
      return 0;
    }
    kb = ks; sym_i = sym_j = 's';
    if (itype == 1) { icase = 2; sym_d = 's'; sym_r = 's'; }
    else if (itype == 2) { icase = 1; sym_d = 's'; sym_r = 's'; }
    else return 7;
    switch (atype) {
      case 0: ip1 = 1; ip2 = 1; jp1 = 1; jp2 = 1; break;
      case 1: ip1 = 2; ip2 = 2; jp1 = 2; jp2 = 2; break;
      case 2: ip1 = 1; ip2 = 1; jp1 = 2; jp2 = 2; break;
      default: ip1 = 2; ip2 = 2; jp1 = 1; jp2 = 1; break;
    }
  } else {
    if (itype < 1 || itype > 9 || atype < 0 || atype > 1) return 7;
    const int *t = s_tab[atype][itype - 1];
    ip1 = t[0]; ip2 = t[1]; jp1 = t[2]; jp2 = t[3]; icase = t[4];
    if (t[5]) sym_v = (char)t[5];
    kb = 2 * ks - 1; sym_i = sym_j = 'n';
    if (icase <= 2) { sym_d = 'n'; sym_r = 'n'; } else { sym_d = 'x'; sym_r = 'x'; }
  }
  size_t nn = (size_t)ns * ns;
  double *xx = (double *)calloc(nn * 3, sizeof(double));
  double *dd = xx + nn, *d = dd + nn;
  double *v = (double *)malloc(sizeof(double) * (size_t)ns);
  double *vv = (double *)malloc(sizeof(double) * (size_t)ms);
  int rc = 0;
#define PQ(ip, orb) (pq + (size_t)ns * (((ip) - 1) + 2 * (size_t)((orb) - 1)))
#define VCH(k) (vch + (size_t)ms * ((k) - 1))
  if (itype == 1) {
    int iii = 0, jjj = 0;
    double S = 0.0;
    for (int j = 0; j < nc && !rc; ++j) {
      if (K1[j] != iii) {
        orc_density(ns, ks, d, PQ(ip1, K1[j] / ibo), PQ(ip2, K1[j] % ibo), sym_d);
        orc_convol(ns, ks, kb, rkb, xx, d, icase, sym_i, sym_j);
      }
      if (K2[j] != jjj) orc_density(ns, ks, dd, PQ(jp1, K2[j] / ibo), PQ(jp2, K2[j] % ibo), sym_r);
      if (K1[j] != iii || K2[j] != jjj) S = orc_sum_amb(ns, ks, xx, dd, sym_r);
      iii = K1[j]; jjj = K2[j];
      const double C = S * cdata[j];
      const int ic = -K3[j], jc = -K4[j], io = -ic, jo = -jc;
      if (ic > 0 && jc > 0) {                                   /* UPDATE_HB */
        if (ic > npert || jc > npert) { rc = 1; break; }
        const int i = ibb[(ic - 1) + (size_t)npert * (jc - 1)];
        if (i > 0) hp[i - 1] = hp[i - 1] + C;
      } else if (ic < 0 && jc > 0) {                            /* UPDATE_HV */
        const int k = kbort(user, io / ibo, io % ibo);
        if (k == 0) { rc = 2; break; }
        const int ich = ipbs[io / ibo];
        if (ich < 1 || ich > nch || jc > npert) { rc = 3; break; }
        const int i = icb[(ich - 1) + (size_t)nch * (jc - 1)];
        if (i > 0) orc_update_hv(ms, hcp + (size_t)ms * (i - 1), VCH(k), C);
      } else if (ic < 0 && jc < 0) {                            /* UPDATE_HW */
        const int ki = kbort(user, io / ibo, io % ibo);
        if (ki == 0) { rc = 2; break; }
        const int kj = kbort(user, jo / ibo, jo % ibo);
        if (kj == 0) { rc = 2; break; }
        const int ich = ipbs[io / ibo], jch = ipbs[jo / ibo];
        if (ich < 1 || jch < 1 || ich > nch || jch > nch) { rc = 4; break; }
        const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
        if (b > 0) orc_update_hw(ms, hch + nn * 4 * (b - 1), ich, jch, VCH(ki), VCH(kj), C);
      } else if (ic < 0 && jc == 0) {                           /* Target_h + UPDATE_HX */
        const int ich = ipbs[io / ibo], jch = ipbs[io % ibo];
        if (ich < 1 || jch < 1 || ich > nch || jch > nch) { rc = 4; break; }
        const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
        if (b > 0 && htarg) {
          const int it = iptar[ich], jt = iptar[jch];
          const int hi = it > jt ? it : jt, lo = it > jt ? jt : it;
          htarg[(size_t)(hi - 1) * hi / 2 + lo - 1] += C;
        }
        if (iitar > 0 && ich != jch && b > 0) orc_update_hx(ms, hch + nn * 4 * (b - 1), ich, jch, fppqq, C);
      } else {
        rc = 6;
      }
    }
  } else if (itype <= 5) {
    int iii = 0, jjj = 0;
    const int ipb = stype == 0 ? ip1 : ip2;
    const char dir = stype == 0 ? 'r' : sym_v;
    for (int j = 0; j < nc && !rc; ++j) {
      const int jj = K1[j], ii = K2[j], ich = K3[j], io = K4[j];
      if (jj != jjj || ii != iii) {
        if (jj != jjj) {
          orc_density(ns, ks, dd, PQ(jp1, jj / ibo), PQ(jp2, jj % ibo), sym_d);
          orc_convol(ns, ks, kb, rkb, xx, dd, icase, sym_i, sym_j);
          jjj = jj;
        }
        orc_bav(ns, ks, xx, PQ(ipb, ii), v, sym_r, dir);
        iii = ii;
      }
      memset(vv, 0, sizeof(double) * (size_t)ms);
      memcpy(vv + (ip2 == 1 ? 0 : ns), v, sizeof(double) * (size_t)ns);
      if (io > 0) {                                             /* UPDATE_HW */
        const int k = kbort(user, io / ibo, io % ibo);
        if (k == 0) { rc = 2; break; }
        const int jch = ipbs[io / ibo];
        if (ich < 1 || jch < 1 || ich > nch || jch > nch) { rc = 4; break; }
        const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
        if (b > 0) orc_update_hw(ms, hch + nn * 4 * (b - 1), ich, jch, vv, VCH(k), cdata[j]);
      } else if (io < 0) {                                      /* UPDATE_HV */
        const int jc = -io;
        if (ich < 1 || ich > nch || jc > npert) { rc = 3; break; }
        const int i = icb[(ich - 1) + (size_t)nch * (jc - 1)];
        if (i > 0) orc_update_hv(ms, hcp + (size_t)ms * (i - 1), vv, cdata[j]);
      } else {
        rc = 7;
      }
    }
  } else {                                                      /* S 6..9 */
    const size_t cols = sym_d == 'x' ? nn : (size_t)ns * (2 * ks - 1);
    for (int j = 0; j < nc; ++j) {
      orc_density(ns, ks, dd, PQ(jp1, K2[j]), PQ(jp2, K3[j]), sym_d);
      for (size_t q = 0; q < cols; ++q) xx[q] = xx[q] + cdata[j] * dd[q];
      if (j < nc - 1 && K1[j] == K1[j + 1]) continue;
      orc_convol(ns, ks, kb, rkb, dd, xx, icase, sym_i, sym_j);
      const int ich = K1[j] / ibo, jch = K1[j] % ibo;
      if (ich < 1 || jch < 1 || ich > nch || jch > nch) { rc = 4; break; }
      const int b = icc[(ich - 1) + (size_t)nch * (jch - 1)];
      if (b > 0) orc_update_hs_mat(ns, ks, hch + nn * 4 * (b - 1), ich, jch, dd, sym_r, ip1, ip2);
      memset(xx, 0, sizeof(double) * nn);
    }
  }
#undef PQ
#undef VCH
  free(xx); free(v); free(vv);
  return rc;
}

/* ===================================================================== */
/* MAT3 coefficient routing: DBSR_MAT3/idef_type.f90 and add_integral.f90 */
/* ipbs[o] for orbital o (1-based, element 0 unused).  Each function      */
/* returns the c_data itype (> 0), 0 for a dropped entry, or -1 where the  */
/* reference stops.                                                        */
/* ===================================================================== */

/* Idef_Rtype: idef_type.f90:141-302 */
int orc_idef_rtype_full(int j1, int j2, int j3, int j4, double C, int ic, int jc, int io, int jo,
                        const int *ipbs, int ibo, double eps_c, int check_target, int *key) {
  if (fabs(C) < eps_c) return 0;
  if (io != 0 && abs(io) <= ibo) return -1;
  if (jo != 0 && abs(jo) <= ibo) return -1;
  if (jc > 0 && ic <= 0) return -1;
  int i1, i2, i3, i4, t;
  if (ipbs[j1] + ipbs[j2] == 0 || ipbs[j3] + ipbs[j4] == 0) {
    i1 = j1; i3 = j3; if (j1 > j3) { i1 = j3; i3 = j1; }
    i2 = j2; i4 = j4; if (j2 > j4) { i2 = j4; i4 = j2; }
    if (i1 > i2) { t = i1; i1 = i2; i2 = t; t = i3; i3 = i4; i4 = t; }
  } else {
    i1 = j1; i2 = j2; i3 = j3; i4 = j4;
  }
  int k1 = 0, k2 = 0, k3 = 0, k4 = 0, itype = 0;
  if (jc > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = -ic; k4 = -jc; itype = 1;
  } else if (ic > 0) {
    if (io == 0) {
      if (ipbs[i1] > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipbs[i1]; k4 = -ic; itype = 2; }
      else if (ipbs[i2] > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipbs[i2]; k4 = -ic; itype = 2; }
      else if (ipbs[i3] > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipbs[i3]; k4 = -ic; itype = 2; }
      else if (ipbs[i4] > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipbs[i4]; k4 = -ic; itype = 2; }
    } else if (io > 0) {
      k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = -ic; itype = 1;
    }
  } else if (io == 0) {
    int kk[4];
    itype = orc_idef_rtype(j1, j2, j3, j4, ipbs, ibo, &kk[0], &kk[1], &kk[2], &kk[3]);
    k1 = kk[0]; k2 = kk[1]; k3 = kk[2]; k4 = kk[3];
  } else if (io > 0 && jo == 0) {
    if (ipbs[i1] > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipbs[i1]; k4 = io; itype = 2; }
    else if (ipbs[i2] > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipbs[i2]; k4 = io; itype = 2; }
    else if (ipbs[i3] > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipbs[i3]; k4 = io; itype = 2; }
    else if (ipbs[i4] > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipbs[i4]; k4 = io; itype = 2; }
    else {
      k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = 0; itype = 1;
      if (check_target == 0) return 0;
    }
  } else if (io > 0 && jo > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = jo; itype = 1;
  }
  if (itype == 0) return -1;
  key[0] = k1; key[1] = k2; key[2] = k3; key[3] = k4;
  return itype;
}

/* Idef_itype: idef_type.f90:306-437 (S^k structures; itype + 10*atype) */
int orc_idef_itype(int atype, int i1, int i2, int i3, int i4, double C, int ic, int jc, int io, int jo,
                   const int *ipbs, int ibo, int *key) {
  if (C == 0.0) return 0;
  if (io != 0 && abs(io) <= ibo) return -1;
  if (jo != 0 && abs(jo) <= ibo) return -1;
  if (jc > 0 && ic <= 0) return -1;
  int k1 = 0, k2 = 0, k3 = 0, k4 = 0, itype = 0;
  if (jc > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = -ic; k4 = -jc; itype = 1;
  } else if (ic > 0) {
    if (io == 0) {
      if (ipbs[i1] > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipbs[i1]; k4 = -ic; itype = 2; }
      else if (ipbs[i2] > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipbs[i2]; k4 = -ic; itype = 3; }
      else if (ipbs[i3] > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipbs[i3]; k4 = -ic; itype = 4; }
      else if (ipbs[i4] > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipbs[i4]; k4 = -ic; itype = 5; }
    } else if (io > 0) {
      k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = -ic; itype = 1;
    }
  } else if (io == 0) {
    if (ipbs[i1] > 0 && ipbs[i3] > 0) { k1 = ipbs[i1] * ibo + ipbs[i3]; k2 = i2; k3 = i4; itype = 6; }
    else if (ipbs[i2] > 0 && ipbs[i4] > 0) { k1 = ipbs[i2] * ibo + ipbs[i4]; k2 = i1; k3 = i3; itype = 7; }
    else if (ipbs[i1] > 0 && ipbs[i4] > 0) { k1 = ipbs[i1] * ibo + ipbs[i4]; k2 = i3; k3 = i2; itype = 8; }
    else if (ipbs[i2] > 0 && ipbs[i3] > 0) { k1 = ipbs[i3] * ibo + ipbs[i2]; k2 = i1; k3 = i4; itype = 9; }
  } else if (io > 0 && jo == 0) {
    if (ipbs[i1] > 0) { k1 = i2 * ibo + i4; k2 = i3; k3 = ipbs[i1]; k4 = io; itype = 2; }
    else if (ipbs[i2] > 0) { k1 = i1 * ibo + i3; k2 = i4; k3 = ipbs[i2]; k4 = io; itype = 3; }
    else if (ipbs[i3] > 0) { k1 = i2 * ibo + i4; k2 = i1; k3 = ipbs[i3]; k4 = io; itype = 4; }
    else if (ipbs[i4] > 0) { k1 = i1 * ibo + i3; k2 = i2; k3 = ipbs[i4]; k4 = io; itype = 5; }
    else { k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = 0; itype = 1; }
  } else if (io > 0 && jo > 0) {
    k1 = i1 * ibo + i3; k2 = i2 * ibo + i4; k3 = io; k4 = jo; itype = 1;
  }
  if (itype == 0) return -1;
  key[0] = k1; key[1] = k2; key[2] = k3; key[3] = k4;
  return atype * 10 + itype;
}

/* Idef_Otype: idef_type.f90:2-50 */
int orc_idef_otype(double C, int ic, int jc, int io, int jo, double eps_c, int check_target, int *key) {
  if (fabs(C) < eps_c) return 0;
  if (jc > 0 && ic <= 0) return -1;
  int itype = 0;
  key[2] = 0; key[3] = 0;
  if (jc > 0) { key[0] = ic; key[1] = jc; itype = 1; }
  else if (ic > 0 && io > 0) { key[0] = io; key[1] = ic; itype = 2; }
  else if (io > 0 && jo > 0) { key[0] = io; key[1] = jo; itype = 3; }
  else if (io > 0 && jo == 0) { key[0] = io; key[1] = 0; itype = 4; }
  if (itype == 0) return -1;
  if (check_target == 0 && itype == 4) return 0;
  return itype;
}

/* Add_integral, MAT3 form: DBSR_MAT3/add_integral.f90:2-57.  lbs, kbs,
 * ipbs indexed by orbital (element 0 unused).  Returns 0 or -1. */
int orc_add_integral_mat3(orc_cdata *cd, int icase, int kpol, int i1, int i2, int i3, int i4, double C,
                          int ic, int jc, int io, int jo, const int *lbs, const int *kbs, const int *ipbs,
                          int ibo, int mk, double eps_c, int check_target) {
  int key[4], t;
  if (icase == 0) {
    t = orc_idef_otype(C, ic, jc, io, jo, eps_c, check_target, key);
    if (t <= 0) return t;
    return orc_cdata_add_coef(cd, C, 0, key[0], key[1], key[2], key[3], t) < 0 ? -1 : 0;
  }
  if (icase == 2) {
    int m = (lbs[i1] + lbs[i3] + kpol) % 2 + (lbs[i2] + lbs[i4] + kpol) % 2;
    if (m != 0 || kpol > mk) return 0;
    t = orc_idef_rtype_full(i1, i2, i3, i4, C, ic, jc, io, jo, ipbs, ibo, eps_c, check_target, key);
    if (t <= 0) return t;
    return orc_cdata_add_coef(cd, C, kpol, key[0], key[1], key[2], key[3], t) < 0 ? -1 : 0;
  }
  if (icase == 3) {
    double S[8];
    for (int v = kpol - 1; v <= kpol + 1; ++v) {
      if (v > mk) continue;
      if (orc_smu(kbs[i1], kbs[i2], kbs[i3], kbs[i4], kpol, v, S) == 0.0) continue;
      const int perm[8][4] = {{i1, i2, i3, i4}, {i2, i1, i4, i3}, {i3, i4, i1, i2}, {i4, i3, i2, i1},
                              {i1, i2, i3, i4}, {i4, i3, i2, i1}, {i3, i4, i1, i2}, {i2, i1, i4, i3}};
      for (int q = 0; q < 8; ++q) {
        const double Cq = C * S[q];
        t = orc_idef_itype(q < 4 ? 0 : 1, perm[q][0], perm[q][1], perm[q][2], perm[q][3], Cq, ic, jc, io, jo,
                           ipbs, ibo, key);
        if (t < 0) return -1;
        if (t == 0) continue;
        if (orc_cdata_add_coef(cd, Cq, v, key[0], key[1], key[2], key[3], t) < 0) return -1;
      }
    }
    return 0;
  }
  return -1;
}

/* ===================================================================== */
/* hf_matrix, two-electron part: DBSR_HF/hf_matrix.f90:2-78 (direct and   */
/* exchange per (ii,jj) block and k) and hf_matrix_breit :97-163.         */
/* a(j,k), b(j,k): the coefficients a(i,j,k), b(i,j,k) of orbital i       */
/* ((norb, kmax+1) column-major).  hfm(ms,ms) holds on entry what the     */
/* caller put there (the reference's qsum(i)*dhl); on return it has the    */
/* two-electron terms added and is divided by qsum.  The Breit list       */
/* (kr1..kr4, crk) is used when mbreit == 2.                              */
/* ===================================================================== */
void orc_hf_matrix(const orc_grid *g, int i, int norb, const double *pq, const double *acoef,
                   const double *bcoef, int kmax, double qsum, int mbreit, int nrk, const int *kr1,
                   const int *kr2, const int *kr3, const int *kr4, const double *crk, double *hfm) {
  const int ns = g->ns, ks = g->ks, ms = g->ms;
  const size_t nn = (size_t)ns * ns;
  double *d = (double *)malloc(sizeof(double) * nn * 4);
  double *dd = d + nn, *x = dd + nn, *xx = x + nn;
  double *rkb = (double *)malloc(sizeof(double) * orc_rkb_size(g, ORC_PPQQ));
#define P(comp, orb) (pq + (size_t)ns * (((comp) - 1) + 2 * (size_t)((orb) - 1)))
  for (int ii = 1; ii <= 2; ++ii)
    for (int jj = 1; jj <= 2; ++jj)
      for (int k = 0; k <= kmax; ++k) {
        const int type = ii == 1 ? (jj == 1 ? ORC_PPPP : ORC_PQPQ) : (jj == 1 ? ORC_QPQP : ORC_QQQQ);
        orc_mrk(g, type, k, rkb);
        memset(dd, 0, sizeof(double) * nn);
        int met = 0;
        for (int j = 1; j <= norb; ++j) {
          double c = acoef[(j - 1) + (size_t)norb * k];
          if (c == 0.0) continue;
          if (i == j) c = c + c;
          orc_density(ns, ks, d, P(jj, j), P(jj, j), 's');
          for (size_t q = 0; q < (size_t)ns * ks; ++q) dd[q] = dd[q] + c * d[q];
          ++met;
        }
        if (met > 0) {
          orc_convol(ns, ks, ks, rkb, d, dd, 1, 's', 's');
          orc_update_hs_lib(ms, hfm, ii, ii, ns, ks, d, 's');
        }
        memset(xx, 0, sizeof(double) * nn);
        met = 0;
        for (int j = 1; j <= norb; ++j) {
          const double c = bcoef[(j - 1) + (size_t)norb * k];
          if (c == 0.0) continue;
          orc_density(ns, ks, x, P(ii, j), P(jj, j), 'x');
          for (size_t q = 0; q < nn; ++q) xx[q] = xx[q] + x[q] * c;
          ++met;
        }
        if (met > 0) {
          orc_convol(ns, ks, ks, rkb, x, xx, 4, 's', 's');
          orc_update_hs_lib(ms, hfm, ii, jj, ns, ks, x, 'x');
        }
      }
  if (mbreit == 2) {
    /* (itype, int) -> ip1, ip2, jp1, jp2 (hf_matrix.f90:122-137) */
    static const int tab[2][4][4] = {{{1, 2, 1, 2}, {1, 2, 1, 2}, {1, 2, 2, 1}, {2, 1, 1, 2}},
                                     {{1, 2, 2, 1}, {2, 1, 1, 2}, {1, 1, 2, 2}, {2, 2, 1, 1}}};
    static const int kint[4] = {2, 1, 4, 3};
    const size_t sk = (size_t)ns * (2 * ks - 1);
    for (int itype = 0; itype <= 1; ++itype)
      for (int k = 0; k <= kmax; ++k) {
        const int kk = k + 1000 * itype;
        orc_mrk(g, itype == 0 ? ORC_PPQQ : ORC_PQQP, k, rkb);
        for (int in = 1; in <= 4; ++in) {
          const char sym = in <= 2 ? 'n' : 'x';
          const int *t = tab[itype][in - 1];
          memset(xx, 0, sizeof(double) * nn);
          int met = 0;
          for (int ik = 0; ik < nrk; ++ik) {
            if (kk != kr1[ik]) continue;
            int jnt, j;
            if (i == kr3[ik]) { jnt = kr2[ik]; j = kr4[ik]; }
            else if (i == kr4[ik]) { jnt = kint[kr2[ik] - 1]; j = kr3[ik]; }
            else continue;
            if (jnt != in) continue;
            orc_density(ns, ks, x, P(t[2], j), P(t[3], j), sym);
            const size_t cols = sym == 'x' ? nn : sk;
            for (size_t q = 0; q < cols; ++q) xx[q] = xx[q] + crk[ik] * x[q];
            ++met;
          }
          if (met > 0) {
            orc_convol(ns, ks, 2 * ks - 1, rkb, x, xx, in, sym, sym);
            orc_update_hs_lib(ms, hfm, t[0], t[1], ns, ks, x, sym);
          }
        }
      }
  }
#undef P
  for (size_t q = 0; q < (size_t)ms * ms; ++q) hfm[q] = hfm[q] / qsum;
  free(d);
  free(rkb);
}

/* ===================================================================== */
/* R^k of a row set the way Rk_evaluate organises the work               */
/* (DBSR_CI3/rk_evaluate.f90:31-56) for rk()'s type sum                  */
/* (DBSR_HF/dbsr_lib_dbs.f90:3981-4022): rows sorted by k and left pair;  */
/* per multipole the four arrays are built (mrk_*), then per (left pair,  */
/* type) one density + case-2 convolution and, per row, one density +     */
/* SUM_AmB.  OpenMP over the four array builds and over the (left pair,   */
/* type) tasks on nthreads threads; the four type values of a row are     */
/* added in rk()'s order, so results do not depend on the thread count.   */
/* Work stops at the first multipole boundary or task after `budget`      */
/* seconds (budget <= 0: no limit); returns the number of leading rows     */
/* whose values are complete.  rows (n,5) row-major (k,a,b,c,d), 1-based;  */
/* pq(ns,2,norb).  CPU baseline only.                                     */
/* ===================================================================== */
int orc_rk_rows_par(const orc_grid *g, int n, const int *rows, const double *pq, int nthreads, double budget,
                    double *out) {
  const int ns = g->ns, ks = g->ks;
  static const int types[4] = {ORC_PPPP, ORC_QQQQ, ORC_QPQP, ORC_PQPQ};
  static const int compl_[4] = {0, 1, 1, 0}, compr_[4] = {0, 1, 0, 1};
  const double t0 = omp_get_wtime();
  double *part = (double *)calloc((size_t)4 * (n > 0 ? n : 1), sizeof(double));
  int done = 0, lo = 0;
  while (lo < n) {
    const int k = rows[5 * lo];
    int hi = lo;
    while (hi < n && rows[5 * hi] == k) ++hi;
    /* left-pair groups of this multipole */
    int ng = 0;
    int *gs = (int *)malloc(sizeof(int) * (size_t)(hi - lo + 1));
    for (int i = lo; i < hi; ++i)
      if (i == lo || rows[5 * i + 1] != rows[5 * (i - 1) + 1] || rows[5 * i + 3] != rows[5 * (i - 1) + 3])
        gs[ng++] = i;
    gs[ng] = hi;
    double *rkb[4];
#pragma omp parallel for num_threads(nthreads < 4 ? nthreads : 4)
    for (int t = 0; t < 4; ++t) {
      rkb[t] = (double *)malloc(sizeof(double) * orc_rkb_size(g, types[t]));
      orc_mrk(g, types[t], k, rkb[t]);
    }
    char *ok = (char *)calloc((size_t)ng * 4, 1);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int task = 0; task < 4 * ng; ++task) {
      if (budget > 0 && omp_get_wtime() - t0 > budget) continue;
      const int gi = task / 4, t = task % 4;
      double *d1 = (double *)malloc(sizeof(double) * (size_t)ns * ks);
      double *d2 = (double *)malloc(sizeof(double) * (size_t)ns * ks);
      double *cv = (double *)malloc(sizeof(double) * (size_t)ns * ks);
      const int a = rows[5 * gs[gi] + 1], c = rows[5 * gs[gi] + 3];
      orc_density(ns, ks, d1, pq + (size_t)ns * (compl_[t] + 2 * (size_t)(a - 1)),
                  pq + (size_t)ns * (compl_[t] + 2 * (size_t)(c - 1)), 's');
      orc_convol(ns, ks, ks, rkb[t], cv, d1, 2, 's', 's');
      for (int i = gs[gi]; i < gs[gi + 1]; ++i) {
        const int b = rows[5 * i + 2], d = rows[5 * i + 4];
        orc_density(ns, ks, d2, pq + (size_t)ns * (compr_[t] + 2 * (size_t)(b - 1)),
                    pq + (size_t)ns * (compr_[t] + 2 * (size_t)(d - 1)), 's');
        part[4 * (size_t)i + t] = orc_sum_amb(ns, ks, cv, d2, 's');
      }
      ok[task] = 1;
      free(d1); free(d2); free(cv);
    }
    for (int t = 0; t < 4; ++t) free(rkb[t]);
    int gdone = 0;
    while (gdone < ng && ok[4 * gdone] && ok[4 * gdone + 1] && ok[4 * gdone + 2] && ok[4 * gdone + 3]) ++gdone;
    done = gs[gdone];
    const int all = gdone == ng;
    free(ok); free(gs);
    if (!all || (budget > 0 && omp_get_wtime() - t0 > budget)) break;
    lo = hi;
  }
  for (int i = 0; i < done; ++i) {
    double r = 0.0;
    r = r + part[4 * (size_t)i];
    r = r + part[4 * (size_t)i + 1];
    r = r + part[4 * (size_t)i + 2];
    r = r + part[4 * (size_t)i + 3];
    out[i] = r;
  }
  free(part);
  return done;
}

/* ===================================================================== */
/* Dirac one-electron operator with a finite nucleus                     */
/* V_nuclear / SKFUN: DBSR_HF/dbsr_lib_dbs.f90:220-310                   */
/* Def_Vnucl :2756-2780; derivative tables of Def_pbsp/Def_qbsp          */
/* :987-1074 (zbsplvd with two derivatives); Def_pq_arrays :1078-1102;   */
/* MAT_dhl_pq :1164-1206 (Bloch term with RB, pnu); constants of zconst  */
/* (DBSR_HF/dbsr_lib_zcom.f90:72-98).                                    */
/* ===================================================================== */
#define ORC_CAU 137.03599976
#define ORC_PI 3.141592653589793238462643

double orc_skfun(int k, double x) {
  if (x < -500.0) return 0.0;
  const double eps = 10.0 * DBL_EPSILON;
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

/* model 0 point, 1 uniform (radius b), 2 Fermi (a, c) */
double orc_vnuclear(int model, double z, double b, double a, double c, double r) {
  if (r <= 0.0) return 0.0;
  if (model == 0) return -z / r;
  if (model == 1) {
    if (r > b) return -z / r;
    return -(3.0 * z) / (2.0 * b) * (1.0 - r * r / (3.0 * b * b));
  }
  const double ac = a / c, ac2 = ac * ac, ac3 = ac2 * ac;
  const double p = ORC_PI * ORC_PI * ac2, s = 6.0 * ac3 * orc_skfun(3, -c / a);
  const double N = 1.0 + p - s, rc = r / c;
  double V;
  if (r <= c) {
    const double ra = (r - c) / a;
    const double s2 = ac2 * orc_skfun(2, ra), s3 = ac3 * orc_skfun(3, ra);
    V = (3.0 - rc * rc + p) / 2.0 - 3.0 * s2 - s / rc + 6.0 / rc * s3;
    V = -(z * V) / (N * c);
  } else {
    const double ra = (c - r) / a;
    const double s2 = ac2 * orc_skfun(2, ra), s3 = ac3 * orc_skfun(3, ra);
    V = N + 6.0 * s3 + 3.0 * rc * s2;
    V = -z / r * (V / N);
  }
  return V;
}

/* first-derivative table xbsd(nv+1,ks,kx) at the outer Gauss points */
static void def_xbsd(const orc_grid *g, int kx, int nsx, const double *tx, double *xbsd) {
  const int nv = g->nv, ks = g->ks, nw = nv + 1;
  double *dbi = (double *)calloc((size_t)nv * kx * kx, sizeof(double));
  double *xm = (double *)malloc(sizeof(double) * nv);
  for (size_t q = 0; q < (size_t)nw * ks * kx; ++q) xbsd[q] = 0.0;
  for (int m = 0; m < ks; ++m) {
    for (int iv = 0; iv < nv; ++iv) xm[iv] = g->gr[iv + (size_t)nv * m];
    orc_zbsplvd(nsx, kx, nv, tx, kx, nv, xm, 2, dbi);
    for (int j = 0; j < kx; ++j)
      for (int iv = 0; iv < nv; ++iv)
        xbsd[iv + (size_t)nw * (m + (size_t)ks * j)] = dbi[iv + (size_t)nv * (j + (size_t)kx * 1)];
  }
  free(dbi);
  free(xm);
}

void orc_vnucl(const orc_grid *g, int model, double z, double b, double a, double c, double *vr) {
  const int nv = g->nv, ks = g->ks;
  for (int j = 0; j < ks; ++j)
    for (int i = 0; i < nv; ++i) {
      const size_t q = i + (size_t)nv * j;
      vr[q] = orc_vnuclear(model, z, b, a, c, g->gr[q]) * g->grw[q];
    }
}

void orc_dhl_pq(const orc_grid *g, int model, double z, double b, double a, double c, int kappa, int bloch,
                double rb, double pnu, double *dhl) {
  const int ns = g->ns, ms = g->ms, nv = g->nv, ks = g->ks, ksp = g->ksp, ksq = g->ksq;
  const size_t nn = (size_t)ns * ns, nk = (size_t)nv * ks, nwk = (size_t)(nv + 1) * ks;
  double *vr = (double *)malloc(sizeof(double) * 2 * nk), *ygw = vr + nk;
  double *pbsd = (double *)malloc(sizeof(double) * nwk * (ksp + ksq)), *qbsd = pbsd + nwk * ksp;
  double *m7 = (double *)malloc(sizeof(double) * 7 * nn);
  double *fpbn = m7, *fqbn = m7 + nn, *fqbs = m7 + 2 * nn, *fpqbs = m7 + 3 * nn, *fqpbs = m7 + 4 * nn,
         *fpqbsd = m7 + 5 * nn, *fqpbsd = m7 + 6 * nn;
  orc_vnucl(g, model, z, b, a, c, vr);
  for (size_t q = 0; q < nk; ++q) ygw[q] = g->grm[q] * g->grw[q];
  def_xbsd(g, ksp, g->nsp, g->tp, pbsd);
  def_xbsd(g, ksq, g->nsq, g->tq, qbsd);
  orc_zintym(nv, ks, ksp, ksp, g->pbsp, g->pbsp, vr, ns, fpbn);
  orc_zintym(nv, ks, ksq, ksq, g->qbsp, g->qbsp, vr, ns, fqbn);
  orc_zintym(nv, ks, ksq, ksq, g->qbsp, g->qbsp, g->grw, ns, fqbs);
  orc_zintym(nv, ks, ksp, ksq, g->pbsp, g->qbsp, ygw, ns, fpqbs);
  orc_zintym(nv, ks, ksq, ksp, g->qbsp, g->pbsp, ygw, ns, fqpbs);
  orc_zintym(nv, ks, ksp, ksq, g->pbsp, qbsd, g->grw, ns, fpqbsd);
  orc_zintym(nv, ks, ksq, ksp, g->qbsp, pbsd, g->grw, ns, fqpbsd);
#define DH(i, j) dhl[(i) + (size_t)ms * (j)]
  for (int j = 0; j < ns; ++j)
    for (int i = 0; i < ns; ++i) {
      const size_t e = i + (size_t)ns * j;
      DH(i, j) = fpbn[e];
      const double sq = -2.0 * ORC_CAU * ORC_CAU;
      DH(ns + i, ns + j) = fqbn[e] + sq * fqbs[e];
      DH(ns + i, j) = ORC_CAU * fqpbsd[e] + ((double)kappa * ORC_CAU) * fqpbs[e];
      DH(i, ns + j) = (-ORC_CAU) * fpqbsd[e] + ((double)kappa * ORC_CAU) * fpqbs[e];
    }
  if (bloch) {
    const double tmax = g->t[ns];
    const double rn = (rb + kappa) / (2.0 * tmax * ORC_CAU);
    const int i = g->nsp - 1, j = ns + g->nsq - 1;
    DH(i, i) = DH(i, i) - pnu * rn * ORC_CAU;
    DH(i, j) = DH(i, j) + pnu * ORC_CAU;
    DH(j, i) = DH(j, i) + (pnu - 1.0) * ORC_CAU;
    DH(j, j) = DH(j, j) - (pnu - 1.0) / rn * ORC_CAU;
  }
#undef DH
  free(vr);
  free(pbsd);
  free(m7);
}

/*
 * dbs_oracle.h -- CPU restatement of DBSR's B-spline Slater-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against and the CPU baseline that bench.py times.  Nothing in the product
 * (dbsr3_amd/) links, loads or calls it.
 *
 * All arrays follow the reference's Fortran conventions: column-major, with
 * the reference's dimensions (e.g. rkb(ns,ns,ks,ks)); indices in this C code
 * are 0-based.  Each function cites the reference routine it restates (paths
 * relative to the reference root).
 */
#ifndef DBS_ORACLE_H
#define DBS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Integral types, numbered as the reference's rka(...,type) slots
 * (DBSR_HF/dbsr_lib_dbs.f90:1598-1601) with the two S^k types after them. */
enum { ORC_PPPP = 1, ORC_QQQQ = 2, ORC_PQPQ = 3, ORC_QPQP = 4,
       ORC_PPQQ = 5, ORC_PQQP = 6 };
/* moment kinds for orc_moments */
enum { ORC_MOM_PP = 0, ORC_MOM_QQ = 1, ORC_MOM_PQ = 2, ORC_MOM_QP = 3 };

typedef struct {
  int ns, ks, ksp, ksq, nv, nsp, nsq, ms;
  double *t;      /* t(ns+ks)                      */
  double *tp;     /* tp(nsp+ksp)                   */
  double *tq;     /* tq(nsq+ksq)                   */
  double *gx, *gw;            /* Gauss-Legendre on [0,1], ks points */
  double *gr, *grm, *grw;     /* (nv,ks)                            */
  double *pbsp;               /* (nv+1,ks,ksp)                      */
  double *qbsp;               /* (nv+1,ks,ksq)                      */
} orc_grid;

/* --- grid and Gauss tables ------------------------------------------- */
void orc_gauleg(double x1, double x2, double *x, double *w, int n);
void orc_zbsplvd(int ns, int ks, int nv, const double *t, int kg, int ni,
                 const double *x, int nderiv, double *dbiatx);
int  orc_mkgrid_01(int ks, int ml, double hi, double he, double hmax,
                   double rmax, double z, double *t, int tcap);
void orc_mkgrid_02(int ns, int ks, double hi, double rmax, double z, double *t);
int  orc_grid_init(orc_grid *g, int ns, int ks, int ksp, int ksq, const double *t);
void orc_grid_free(orc_grid *g);
orc_grid *orc_grid_new(int ns, int ks, int ksp, int ksq, const double *t);
void orc_grid_delete(orc_grid *g);
void orc_grid_dims(const orc_grid *g, int *dims8);
void orc_grid_tables(const orc_grid *g, double *gr, double *grw, double *pbsp, double *qbsp);

/* --- moments, diagonal cells, arrays --------------------------------- */
void orc_moments(const orc_grid *g, int which, int k, double *rkm);
void orc_triang(const orc_grid *g, int type, int k, int iv, double *rkd);
void orc_mrk(const orc_grid *g, int type, int k, double *rkb);
size_t orc_rkb_size(const orc_grid *g, int type);

/* --- contraction primitives ------------------------------------------ */
void orc_density(int ns, int ks, double *d, const double *p1, const double *p2, char type);
void orc_convol(int ns, int ks, int kb, const double *rkb, double *a, const double *d,
                int icase, char sym_i, char sym_j);
double orc_sum_amb(int ns, int ks, const double *a, const double *b, char sym);
double orc_quadr(const orc_grid *g, const double *fi, const double *fj, int m);
void orc_update_hs_lib(int ms, double *hm, int ii, int jj, int ns, int ks,
                       const double *d, char sym);
void orc_update_hs_mat(int ns, int ks, double *blk, int ich, int jch,
                       const double *d, char sym, int ii, int jj);

/* --- scalar integrals -------------------------------------------------- */
double orc_rk(const orc_grid *g, const double *rkb_pppp, const double *rkb_qqqq,
              const double *rkb_pqpq, const double *rkb_qpqp,
              const double *f1, const double *f2, const double *f3, const double *f4);

/* --- c_data: blocked sorted coefficient lists ------------------------- */
typedef struct orc_cdata orc_cdata;
/* called by Add_matrix with the merged list in IPT order */
typedef void (*orc_gen_matrix_fn)(void *user, int itype, int kpol, int nc,
                                  const double *c, const int *k1, const int *k2,
                                  const int *k3, const int *k4);
orc_cdata *orc_cdata_alloc(int ntype, int kp1, int kp2, int mblock, int nblock,
                           int kblock, double eps_c, orc_gen_matrix_fn gen, void *user);
void orc_cdata_free(orc_cdata *cd);
int  orc_cdata_add_coef(orc_cdata *cd, double c, int kpol, int j1, int j2, int j3, int j4,
                        int itype);
int  orc_cdata_add_matrix(orc_cdata *cd, int itype, int kpol);

/* --- accumulation drivers (lists already in IPT order) ----------------- */
void orc_rk_evaluate(int ns, int ks, int kb, const double *rkb, int itype,
                     int nc, const double *cdata, const int *K1, const int *K2,
                     const int *K3, const int *K4, double eps_c,
                     const double *pq, int ibi, double *HM, int ldhm, int nzero,
                     double *DM);
void orc_rdata_34(int ns, int ks, const double *rkb, int itype, int atype,
                  int nc, const double *cdata, const int *K1, const int *K2,
                  const int *K3, const double *pq, int ibo, const int *icc,
                  int nch, double *hch);
int  orc_idef_rtype(int j1, int j2, int j3, int j4, const int *ipbs, int ibo,
                    int *k1, int *k2, int *k3, int *k4);
void orc_decode_int(int code, int *met, int *k, int *i1, int *i2, int *i3, int *i4);
int  orc_l_kappa(int kappa);
double orc_smu(int ka, int kb, int kc, int kd, int L, int v, double *S);
int  orc_add_integral_ci(orc_cdata *cd, int kpol, int ic, int jc, int i1, int i2, int i3, int i4,
                         double C, const int *kappa, int ibi, int mbreit);

/* --- overlaps, unfolding, MAT3 updates ------------------------------- */
void orc_full_mat_sym(int ns, int ks, const double *x, double *y, char sym);
void orc_zintym(int nv, int ks, int ks1, int ks2, const double *b1, const double *b2,
                const double *ygr, int mm, double *ym);
void orc_overlaps(const orc_grid *g, double *fpbs, double *fqbs, double *fppqq);
/* Dirac operator with a finite nucleus (model 0 point, 1 uniform sphere of
 * radius b, 2 Fermi a, c): V_nuclear, Def_Vnucl's VR_nucl(nv,ks),
 * MAT_dhl_pq(kappa) into dhl(ms,ms) */
double orc_skfun(int k, double x);
double orc_vnuclear(int model, double z, double b, double a, double c, double r);
void orc_vnucl(const orc_grid *g, int model, double z, double b, double a, double c, double *vr);
void orc_dhl_pq(const orc_grid *g, int model, double z, double b, double a, double c, int kappa, int bloch,
                double rb, double pnu, double *dhl);
void orc_update_hx(int ms, double *blk, int ich, int jch, const double *d, double C);
void orc_update_hw(int ms, double *blk, int ich, int jch, const double *v, const double *w, double C);
void orc_update_hv(int ms, double *col, const double *v, double C);
typedef int (*orc_kbort_fn)(void *user, int i1, int i2);
int  orc_o_data(int ms, int itype, int n, const int *k1, const int *k2, const double *c, double eps_c,
                int ibo, const int *ipbs, orc_kbort_fn kbort, void *user, const double *vch,
                const double *fppqq, int iitar, const int *iptar, double *otarg,
                int nch, int npert, const int *icc, double *hch, const int *icb, double *hcp,
                const int *ibb, double *hp);
int  orc_idef_rtype_full(int j1, int j2, int j3, int j4, double C, int ic, int jc, int io, int jo,
                         const int *ipbs, int ibo, double eps_c, int check_target, int *key);
int  orc_idef_itype(int atype, int i1, int i2, int i3, int i4, double C, int ic, int jc, int io, int jo,
                    const int *ipbs, int ibo, int *key);
int  orc_idef_otype(double C, int ic, int jc, int io, int jo, double eps_c, int check_target, int *key);
int  orc_add_integral_mat3(orc_cdata *cd, int icase, int kpol, int i1, int i2, int i3, int i4, double C,
                           int ic, int jc, int io, int jo, const int *lbs, const int *kbs, const int *ipbs,
                           int ibo, int mk, double eps_c, int check_target);
void orc_hf_matrix(const orc_grid *g, int i, int norb, const double *pq, const double *acoef,
                   const double *bcoef, int kmax, double qsum, int mbreit, int nrk, const int *kr1,
                   const int *kr2, const int *kr3, const int *kr4, const double *crk, double *hfm);
void orc_bav(int ns, int ks, const double *a, const double *b, double *c, char sym, char dir);
int  orc_rs_data(int ns, int ks, const double *rkb, int stype, int itype, int atype, int nc,
                 const double *cdata, const int *K1, const int *K2, const int *K3, const int *K4,
                 const double *pq, int ibo, const int *ipbs, orc_kbort_fn kbort, void *user,
                 const double *vch, const double *fppqq, int iitar, const int *iptar, double *htarg,
                 int nch, int npert, const int *icc, double *hch, const int *icb, double *hcp,
                 const int *ibb, double *hp);
double orc_grid_start(int nucl, double hi, double z, double a_fermi, double r_uniform);
int  orc_mkgrid_01s(int ks, int ml, double hs, double he, double hmax, double rmax, double *t, int tcap);
void orc_mkgrid_02s(int ns, int ks, double ti, double rmax, double *t);
int  orc_rk_rows_par(const orc_grid *g, int n, const int *rows, const double *pq, int nthreads, double budget,
                     double *out);
int  orc_rdata_acf(int nc, const double *cdata, const int *K1, const int *K2, const int *K3, int ibo,
                   int kpol, int mk, const double *qpq, int norb, const int *icc, int nch, int nblocks,
                   double *acf);

#ifdef __cplusplus
}
#endif
#endif

/*
 * dbsr3_amd.h -- C-ABI of libdbsr3_amd.so, the MI355X (gfx950) engine for
 * DBSR's two-electron B-spline Slater integrals.
 *
 * Conventions
 *  - Every function returns an int status (DBS_OK = 0); dbs_last_error()
 *    gives the message of the last failure on the calling thread.  The
 *    messages of the reference's Stop statements are kept where the same
 *    condition is detected (e.g. "convol: unknown case").
 *  - Arrays are column-major with the reference's dimensions (Fortran
 *    layout), e.g. rkb(ns,ns,ks,ks), d(ns,ks), pq(ns,2,norb).  Orbital and
 *    channel indices inside coefficient keys are 1-based, as in the reference.
 *  - Pointers named *_dev are device (HBM) pointers; all others are host.
 *  - Work runs on the context's HIP stream; calls returning results into
 *    host memory synchronise that stream, *_dev calls do not.
 *  - Types: 1 pppp, 2 qqqq, 3 pqpq, 4 qpqp (R^k), 5 ppqq, 6 pqqp (S^k).
 * Reference paths are relative to the zatsaroi/DBSR3 tree.
 */
#ifndef DBSR3_AMD_H
#define DBSR3_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DBS_OK 0
#define DBS_ERR_ARG 1      /* invalid argument                          */
#define DBS_ERR_HIP 2      /* HIP runtime failure                       */
#define DBS_ERR_STATE 3    /* call out of sequence / missing object     */
#define DBS_ERR_CAPACITY 4 /* a capacity limit was exceeded             */

#define DBS_PPPP 1
#define DBS_QQQQ 2
#define DBS_PQPQ 3
#define DBS_QPQP 4
#define DBS_PPQQ 5
#define DBS_PQQP 6
#define DBS_RK_TOTAL 0 /* engine kind: the rk() sum over pppp+qqqq+qpqp+pqpq */

typedef struct dbs_ctx dbs_ctx;
typedef struct dbs_cdata dbs_cdata;
typedef struct dbs_engine dbs_engine;

const char *dbs_last_error(void);
int dbs_version(void);

/* ---------------------------------------------------------------------
 * Grid context.  Replaces the module state set by def_grid +
 * alloc_DBS_gauss + Def_pbsp/Def_qbsp (DBSR_HF/dbsr_lib_dbs.f90:480-596,
 * 911-1074): knots t(1:ns+ks), orders ks/ksp/ksq (ksp<=0 -> ks-1,
 * ksq<=0 -> ks, as :565-566).  Gauss tables and B-spline values are built
 * on the device and stay resident.  A context has one stream, one scratch
 * and two pinned list buffers: one host thread at a time may call into it.
 * ------------------------------------------------------------------- */
int dbs_ctx_create(dbs_ctx **ctx, int ns, int ks, int ksp, int ksq, const double *t, int device);
int dbs_ctx_destroy(dbs_ctx *ctx);
/* dims8 = {ns, ks, ksp, ksq, nv, nsp, nsq, ms} */
int dbs_ctx_dims(const dbs_ctx *ctx, int *dims8);
int dbs_ctx_sync(dbs_ctx *ctx);
/* copies gr(nv,ks), grw(nv,ks), pbsp(nv+1,ks,ksp), qbsp(nv+1,ks,ksq) to host
 * (module DBS_gauss, dbsr_lib_dbs.f90:830-891); any pointer may be NULL */
int dbs_ctx_tables(dbs_ctx *ctx, double *gr, double *grw, double *pbsp, double *qbsp);
/* bytes of materialised arrays the context may cache (default 8 GiB): the
 * arrays rkb(ns,ns,kb,kb) of (type, k) and, for Rk_evaluate / Sk_evaluate
 * and R_data's band convolutions, their K x K permutations (K = ns*kb, the
 * same size), least recently used evicted first */
int dbs_ctx_set_cache_limit(dbs_ctx *ctx, size_t bytes);
/* bytes of device workspace the channel post-processing (DBS_ORTH,
 * Diag_channels, Record_matrix) takes per batch of blocks (default 256 MiB);
 * larger partial waves run in more batches with the same results; at least 64 KiB */
int dbs_ctx_set_workspace_limit(dbs_ctx *ctx, size_t bytes);

/* device memory helpers for callers without HIP headers (Fortran shim) */
int dbs_malloc(dbs_ctx *ctx, void **dptr, size_t bytes);
int dbs_free(dbs_ctx *ctx, void *dptr);
int dbs_memcpy_h2d(dbs_ctx *ctx, void *dst_dev, const void *src, size_t bytes);
int dbs_memcpy_d2h(dbs_ctx *ctx, void *dst, const void *src_dev, size_t bytes);

/* ---------------------------------------------------------------------
 * Plain path: the reference's routines one for one.
 * ------------------------------------------------------------------- */
/* moments_pp/qq/pq/qp (dbsr_lib_dbs.f90:1384-1564): which 0 pp, 1 qq,
 * 2 pq, 3 qp; rkm(ks*ks,nv), packed as the reference */
int dbs_moments_dev(dbs_ctx *ctx, int which, int k, double *rkm_dev);
/* triang_* for every interval (dbsr_lib_dbs.f90:2979-3063, 3151-3245,
 * 3333-3427, 3512-3597, 3682-3767, 3852-3936): rkd(ks*ks,ks*ks,nv) */
int dbs_diag_cells_dev(dbs_ctx *ctx, int type, int k, double *rkd_dev);
/* mrk_pppp/qqqq/pqpq/qpqp and msk_ppqq/pqqp (dbsr_lib_dbs.f90:2898-2959,
 * 3431-3492, 3067-3131, 3249-3313, 3601-3662, 3771-3833): returns the
 * context's cached device array rkb(ns,ns,kb,kb), kb = ks (R) or 2ks-1 (S).
 * The pointer stays valid only until the next array build on this context
 * (any entry point that needs another (type,k) may evict it when the cache
 * limit is reached; eviction synchronises the context stream first) or
 * until the context is destroyed. */
int dbs_mrk(dbs_ctx *ctx, int type, int k, const double **rkb_dev);
/* same, copied into a host array of ns*ns*kb*kb doubles */
int dbs_mrk_host(dbs_ctx *ctx, int type, int k, double *rkb);
/* density (dbsr_lib_dbs.f90:2834-2891): type 's','l','n','x' */
int dbs_density_dev(dbs_ctx *ctx, const double *p1_dev, const double *p2_dev, char type, double *d_dev);
/* convol / convol_sk (dbsr_lib_dbs.f90:2208-2478, 2482-2752) applied to the
 * array of (type,k); icase 1..4, sym 's','l','n' */
int dbs_convol_dev(dbs_ctx *ctx, int type, int k, int icase, char sym_i, char sym_j,
                   const double *d_dev, double *a_dev);
/* SUM_AmB (dbsr_lib_dbs.f90:4027-4080); result written to out_dev[0] */
int dbs_sum_amb_dev(dbs_ctx *ctx, const double *a_dev, const double *b_dev, char sym, double *out_dev);
/* rk (dbsr_lib_dbs.f90:3981-4022): f*(ns,2) device; out_dev[0] */
int dbs_rk_dev(dbs_ctx *ctx, const double *f1, const double *f2, const double *f3,
               const double *f4, int k, double *out_dev);
/* rk() that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  Same kernels and results (bit for bit) as
 * dbs_rk_dev, but the densities, the convolution and the partial sums live
 * in a fixed device array of the library (1.25 MiB, one copy per GPU, kept
 * until the process ends) instead of the context's scratch.  Limits: ns*ks
 * <= 32768 (DBS_ERR_CAPACITY otherwise); the four R^k arrays of multipole
 * k (types 1-4) must already be in the context's cache, e.g. built by
 * dbs_mrk (DBS_ERR_STATE otherwise; the call neither builds nor evicts
 * arrays).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams or threads) race and must be serialised by the
 * caller. */
int dbs_rk_noalloc_dev(dbs_ctx *ctx, const double *f1, const double *f2, const double *f3,
                       const double *f4, int k, double *out_dev);
/* quadr (dbsr_lib_dbs.f90:3940-3976): <fi|r^m|fj>, out_dev[0] */
int dbs_quadr_dev(dbs_ctx *ctx, const double *fi_dev, const double *fj_dev, int m, double *out_dev);
/* UPDATE_HS, library form (dbsr_lib_dbs.f90:4137-4193): hm(ms,ms) += y/2 and
 * its transpose at sub-block (ii,jj); sym 's','n','x' */
int dbs_update_hs_dev(dbs_ctx *ctx, double *hm_dev, int ii, int jj, const double *d_dev, char sym);

/* host-pointer forms of the per-call routines, for drop-in callers that
 * keep their arrays in host memory (each stages through the device) */
int dbs_density(dbs_ctx *ctx, const double *p1, const double *p2, char type, double *d);
int dbs_convol(dbs_ctx *ctx, int type, int k, int icase, char sym_i, char sym_j,
               const double *d, double *a);
int dbs_sum_amb(dbs_ctx *ctx, const double *a, const double *b, char sym, double *out);
int dbs_rk(dbs_ctx *ctx, const double *f1, const double *f2, const double *f3,
           const double *f4, int k, double *out);
int dbs_quadr(dbs_ctx *ctx, const double *fi, const double *fj, int m, double *out);

/* ---------------------------------------------------------------------
 * Coefficient lists: module c_data (DBSR_MAT3/MOD_c_data.f90:2-368).
 * Blocked per-(kpol,itype) sorted lists with equal-key summation at insert
 * (Add_cdata :177-236), block overflow handling (Add_coef :119-173), k-way
 * merge with eps drop (Merge_cdata :240-314) and Add_matrix (:318-368).
 * Add_matrix calls gen(user, itype, kpol, nc, c, K1..K4) with the merged
 * list in IPT order -- the reference's Gen_matrix hook.
 * ------------------------------------------------------------------- */
typedef void (*dbs_gen_matrix_fn)(void *user, int itype, int kpol, int nc, const double *c,
                                  const int *K1, const int *K2, const int *K3, const int *K4);
int dbs_cdata_create(dbs_cdata **cd, int ntype, int kpol1, int kpol2, int mblock, int nblock,
                     int kblock, double eps_c, dbs_gen_matrix_fn gen, void *user);
int dbs_cdata_destroy(dbs_cdata *cd);
int dbs_cdata_add_coef(dbs_cdata *cd, double c, int kpol, int j1, int j2, int j3, int j4, int itype);
int dbs_cdata_add_matrix(dbs_cdata *cd, int itype, int kpol);

/* Add_integral, CI3 form (DBSR_CI3/add_integral.f90:2-43): the parity
 * rule on l(i1)+l(i3)+k, l(i2)+l(i4)+k, then Add_coef for types 1..4 with
 * K1=i1*ibi+i3, K2=i2*ibi+i4, K3=ic, K4=jc; with mbreit != 0 also the eight
 * S^k terms (types 5/6, multipoles k-1..k+1) weighted by the Breit factors
 * SMU (DBSR_HF/hf_energy_BR.f90:58-121).  kappa[i] = kappa of 1-based
 * orbital i (kappa[0] unused); l = l_kappa(kappa). */
int dbs_add_integral_ci(dbs_cdata *cd, int kpol, int ic, int jc, int i1, int i2, int i3, int i4,
                        double c, const int *kappa, int ibi, int mbreit);
/* SMU itself: returns the factor sum and the eight factors in S[8] */
int dbs_smu(int ka, int kb, int kc, int kd, int L, int v, double *S, double *smu);
/* Add_integral, MAT3 R-branch for channel-channel terms
 * (DBSR_MAT3/add_integral.f90:28-32 + Idef_Rtype idef_type.f90:142-302,
 * structures 6.0-9.0 with ic=jc=io=jo=0).  ipbs[i] > 0 marks orbital i as
 * channel ipbs[i].  Terms with |c| < eps_c or k > mk are dropped as there. */
int dbs_add_integral_mat(dbs_cdata *cd, int kpol, int i1, int i2, int i3, int i4, double c,
                         const int *lorb, const int *ipbs, int ibo, int mk, double eps_c);
/* MAT3 orbital tables for Add_integral (module dbsr_mat / orb_jj): for
 * orbital o (1-based) element o-1 of lbs (l), kbs (kappa), ipbs (channel,
 * 0 = bound); ibo = 2**15 (mod_dbsr_mat.f90:61), mk, eps_C, check_target. */
typedef struct dbs_mat3_orbs {
  int norb;
  const int *lbs, *kbs, *ipbs;
  int ibo, mk;
  double eps_c;
  int check_target;
} dbs_mat3_orbs;
/* Add_integral, MAT3 form (DBSR_MAT3/add_integral.f90:2-57): icase 0
 * overlaps -> Idef_Otype; 1 L-integrals -> Idef_Ltype of (i1, i3)
 * (idef_type.f90:54-138, structures 1..3 with multipole 0); 2 R-integrals
 * -> parity rule, Idef_Rtype (all 16 structures with ic, jc, io, jo); 3
 * S-integrals -> SMU and Idef_itype for the eight permuted terms of
 * multipoles kpol-1..kpol+1 (atype 0 / 1, itype + 10*atype)
 * (DBSR_MAT3/idef_type.f90:2-437). */
int dbs_add_integral_mat3(dbs_cdata *cd, const dbs_mat3_orbs *o, int icase, int kpol, int i1, int i2,
                          int i3, int i4, double C, int ic, int jc, int io, int jo);
/* Decode_int (DBSR_CI3/setup_matrix.f90:206-223) for n packed int_bnk codes */
int dbs_decode_int(int n, const int *code, int *met, int *k, int *i1, int *i2, int *i3, int *i4);

/* Rk_evaluate / Sk_evaluate (DBSR_CI3/rk_evaluate.f90:2-63,
 * sk_evaluate.f90:2-63): evaluate every entry of a merged list (IPT order,
 * host arrays) of type itype (1..6) and multipole k on the device and add
 * S*c to HM(ic,jc) (jc <= nzero) or DM(jc) (ic == jc > nzero).  Entries
 * with |c| < eps_c are skipped.  Contributions to one element are added in
 * list order, as the reference does.  pq_dev(ns,2,norb); HM_dev(ldhm,*).
 * The work is queued on the context stream and the call returns without
 * waiting (the host arrays may be reused at once: the tables go through
 * the context's pinned buffers); read HM/DM after dbs_ctx_sync or with
 * the context's own copies (dbs_memcpy_d2h). */
int dbs_rk_evaluate_dev(dbs_ctx *ctx, int itype, int k, int nc, const double *c, const int *K1,
                        const int *K2, const int *K3, const int *K4, double eps_c, int ibi,
                        const double *pq_dev, int norb, double *HM_dev, int ldhm, int nzero,
                        double *DM_dev);
/* R_data itype 3 (direct) / 4 (exchange) (DBSR_MAT3/R_data.f90:214-277)
 * for one atype 0..3 (R_data.f90:76-86): per run of equal K1=(ich,jch), the
 * c-weighted density of orbitals (K2,K3) is convolved (case 1 's' or case
 * 3 'x') and added by UPDATE_HS (DBSR_MAT3/updates.f90:110-179) to block
 * icc(ich,jch) of hch_dev(ms,ms,nblocks); icc is (nch,nch) host, 1-based
 * block numbers, 0 = not stored.  UPDATE_ACF terms are not produced.
 * Queued on the context stream, returns without waiting (tables through
 * the context's pinned buffers): read hch after dbs_ctx_sync or through
 * the context's copies. */
int dbs_rdata_34_dev(dbs_ctx *ctx, int itype, int atype, int k, int nc, const double *c,
                     const int *K1, const int *K2, const int *K3, int ibo, const double *pq_dev,
                     int norb, const int *icc, int nch, double *hch_dev, int nblocks);

/* ---------------------------------------------------------------------
 * B-spline overlaps, unfolding and the remaining matrix updates
 * (dbs_updates.hip).  Loop and summation orders are the reference's, so the
 * results are bit-identical to it.
 * --------------------------------------------------------------------- */
/* ZINTYM (dbsr_lib_dbs.f90:4472-4497) as called by Def_pbsp / Def_qbsp
 * (:1026, :1072) and Def_pq_arrays (:1078-1102): fpbs(ns,ns) = <B^P_i|B^P_j>,
 * fqbs(ns,ns) = <B^Q_i|B^Q_j>, fppqq(ms,ms) = diag(fpbs, fqbs).  Any output
 * pointer may be NULL. */
int dbs_bspline_overlaps_dev(dbs_ctx *ctx, double *fpbs_dev, double *fqbs_dev, double *fppqq_dev);
/* the nucleus of module DBS_nuclear (DBSR_HF/dbsr_lib_dbs.f90:47-98): its
 * charge distribution, atomic number z and, in atomic units, the radius of
 * the uniform sphere or the Fermi parameters a, c (the a_fermi, c_fermi of
 * a knot.dat or .bsw file). */
#define DBS_NUCLEUS_POINT 0
#define DBS_NUCLEUS_UNIFORM 1
#define DBS_NUCLEUS_FERMI 2
typedef struct dbs_nucleus {
  int model;
  double z, r_uniform, a_fermi, c_fermi;
} dbs_nucleus;
/* VR_nucl of Def_Vnucl (dbsr_lib_dbs.f90:2756-2780): vr(nv,ks) =
 * V_nuclear(gr) * grw at the outer Gauss points (V_nuclear and SKFUN
 * :220-310), device output. */
int dbs_vnucl_dev(dbs_ctx *ctx, const dbs_nucleus *nuc, double *vr_dev);
/* MAT_dhl_pq(kappa) (dbsr_lib_dbs.f90:1164-1206): the Dirac operator
 * dhl(ms,ms) of the grid, column-major on the device: <P|V|P>,
 * c<P|-d/dr+k/r|Q>, c<Q|d/dr+k/r|P>, <Q|V-2c^2|Q> from Def_Vnucl and the
 * mixed arrays of Def_pq_arrays (:1078-1102).  bloch = 1 adds the Bloch
 * boundary term with parameters rb, pnu (module DBS_dhl_pq: mbloch = 1,
 * RB = 0, pnu = 1 by default).  <v1|dhl|v2> (Vp_dhl, :1210-1227) is then a
 * matrix product on the caller's side. */
int dbs_dhl_pq_dev(dbs_ctx *ctx, const dbs_nucleus *nuc, int kappa, int bloch, double rb, double pnu,
                   double *dhl_dev);
/* Full_mat_sym (DBSR_HF/dbsr_lib_zcom.f90:1247-1294): y(ns,ns)
 * from x in storage sym 's' (upper band x(ns,ks)), 'l' (lower band), 'n'
 * (non-symmetric band x(ns,2ks-1)) or 'x' (full). */
int dbs_full_mat_sym_dev(dbs_ctx *ctx, const double *x_dev, char sym, double *y_dev);

/* MAT3 matrix storage (DBSR_MAT3/mod_dbsr_mat.f90:134-147, allocated as in
 * alloc_dbsr_matrix, DBSR_MAT3/updates.f90:2-75).  Index tables are host,
 * column-major, 1-based block numbers (0 = block not stored: updates to it
 * are dropped, as for blocks another process owns).  Matrices are device. */
typedef struct dbs_mat_blocks {
  int nch, npert, mk;
  const int *icc; int nblocks; double *hch_dev; /* icc(nch,nch); hch(ms,ms,nblocks) */
  double *acf_dev;                              /* acf(nblocks,0:mk) */
  const int *icb; int ncb; double *hcp_dev;     /* icb(nch,npert); hcp(ms,ncb) */
  const int *ibb; int nbb; double *hp_dev;      /* ibb(npert,npert); hp(nbb) */
} dbs_mat_blocks;

/* UPDATE_HX (DBSR_MAT3/updates.f90:79-106): block(ich,jch) += C*d
 * (ich >= jch) or C*d^T; d_dev(ms,ms) */
int dbs_update_hx_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int ich, int jch, const double *d_dev,
                      double C);
/* UPDATE_HW (updates.f90:183-217): block(ich,jch) += C*v*w^T (ich > jch),
 * C*(v*w^T)^T (ich < jch), (C/2)*(x + x^T) (ich == jch); v,w (ms) device */
int dbs_update_hw_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int ich, int jch, const double *v_dev,
                      const double *w_dev, double C);
/* UPDATE_HV (updates.f90:221-244): hcp(:,icb(ich,ip)) += C*v */
int dbs_update_hv_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int ich, int ip, const double *v_dev,
                      double C);
/* UPDATE_HB (updates.f90:248-266) for n calls in order (host lists):
 * hp(ibb(ic,jc)) += c */
int dbs_update_hb_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int n, const int *ic, const int *jc,
                      const double *c);
/* UPDATE_ACF (updates.f90:270-292) for n calls in order (host lists):
 * acf(icc(ich,jch),k) += C */
int dbs_update_acf_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int n, const int *k, const int *ich,
                       const int *jch, const double *C);

/* KBORT(i1,i2) of the host program: column of V_ch for the bound orbital
 * (i1,i2), 0 if none (the reference's external KBORT). */
typedef int (*dbs_kbort_fn)(void *user, int i1, int i2);
/* O_data (DBSR_MAT3/O_data.f90:2-93): one merged overlap list (IPT order,
 * host arrays k1,k2,c) of structure itype 1..4 into the MAT3 matrices.
 * 1 UPDATE_HB(k1,k2); 2 UPDATE_HV(ipbs(io/ibo), k2, V_ch(:,KBORT)); 3
 * UPDATE_HW(ipbs(i1), ipbs(j1), V_ch(:,KBORT(i)), V_ch(:,KBORT(j))); 4 the
 * target overlap otarg(ij) += C (Target_h, DBSR_MAT3/target_subs.f90:2-31;
 * skipped when otarg is NULL) and, for iitar > 1 and ich != jch,
 * UPDATE_HX(ich,jch,fppqq,C).  Entries with |c| < eps_c are skipped.  The
 * updates of one matrix element are applied in list order.  vch_dev(ms,nvch),
 * fppqq_dev(ms,ms) device; ipbs(orbital) and iptar(channel) host, 1-based. */
int dbs_o_data_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int itype, int n, const int *k1,
                   const int *k2, const double *c, double eps_c, int ibo, const int *ipbs,
                   dbs_kbort_fn kbort, void *kbort_user, const double *vch_dev, int nvch,
                   const double *fppqq_dev, int iitar, const int *iptar, double *otarg);
/* CI overlap, met = 0 (DBSR_CI3/setup_matrix.f90:108-120): for each entry
 * with |cc| >= eps_c and jc <= nzero, SM(ic,jc) += cc, in list order;
 * SM_dev(ldsm,nzero). */
int dbs_overlap_ci_dev(dbs_ctx *ctx, int n, const int *ic, const int *jc, const double *cc,
                       double eps_c, int nzero, double *SM_dev, int ldsm);

/* QUADR_PQ(i1,i2,m) of the host program: <i1|r^m|i2> over the P and Q
 * components of orbitals i1, i2 (1-based), the reference's external
 * QUADR_PQ (called at DBSR_MAT3/R_data.f90:224; defined in the un-vendored
 * Libraries_03). */
typedef double (*dbs_quadr_pq_fn)(void *user, int i1, int i2, int m);
/* A ready QUADR_PQ for callers that tabulate it: pass dbs_quadr_pq_table as
 * the callback and a dbs_qpq_table as its user pointer; returns
 * q(i1,i2,m), column-major (norb,norb,0:mmax), 1-based orbitals (0.0 for
 * indices or m outside the table). */
typedef struct dbs_qpq_table {
  int norb, mmax;
  const double *q;
} dbs_qpq_table;
double dbs_quadr_pq_table(void *user, int i1, int i2, int m);
/* Bound-orbital and target data of a MAT3 run (module dbsr_mat,
 * DBSR_MAT3/mod_dbsr_mat.f90): pair packing base ibo (:61), ipbs(orbital)
 * (channel of a channel orbital, 0 for bound; 1-based orbitals), the host
 * program's KBORT, V_ch(ms,nvch) and fppqq(ms,ms) on the device, iitar,
 * iptar(channel) and the packed target interactions htarg (host, may be
 * NULL; Target_h, DBSR_MAT3/target_subs.f90:2-31). */
typedef struct dbs_mat3_aux {
  int ibo;
  const int *ipbs;
  dbs_kbort_fn kbort;
  void *kbort_user;
  const double *vch_dev;
  int nvch;
  const double *fppqq_dev;
  int iitar;
  const int *iptar;
  double *htarg;
  dbs_quadr_pq_fn quadr_pq; /* QUADR_PQ for the asymptotic coefficients */
  void *quadr_pq_user;
} dbs_mat3_aux;
/* R_data(itype, kpol, btype) (DBSR_MAT3/R_data.f90:2-290): one merged list
 * (IPT order, host arrays) of structure itype 1..4 and atype 0..3 into the
 * MAT3 matrices: 1 integral x {UPDATE_HB, UPDATE_HV, UPDATE_HW, Target_h +
 * UPDATE_HX(fppqq)}; 2 the bound-channel vectors BAV(convol(case 1,
 * density(K1)), pq(K2)) into UPDATE_HW / UPDATE_HV; 3/4 as dbs_rdata_34_dev,
 * and for itype 3 with atype 0/1 also the asymptotic coefficients
 * (R_data.f90:222-226, 240): per run of equal K1, CA = sum of
 * c*QUADR_PQ(K2,K3,k) in list order, then UPDATE_ACF(k,ich,jch,CA).  That
 * needs aux->quadr_pq when m->acf_dev is set (error otherwise); runs are
 * applied in list order.  pq_dev(ns,2,norb). */
int dbs_rdata_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, const dbs_mat3_aux *aux, int itype, int atype,
                  int k, int nc, const double *c, const int *K1, const int *K2, const int *K3,
                  const int *K4, const double *pq_dev, int norb);
/* S_data(ktype, kpol) (DBSR_MAT3/S_data.f90:2-287): ktype = atype*10 +
 * itype, atype 0 ppqq / 1 pqqp, itype 1..9 (1 as R_data 1; 2..5 vectors by
 * BAV; 6..9 runs of equal K1 into UPDATE_HS). */
int dbs_sdata_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, const dbs_mat3_aux *aux, int ktype, int k,
                  int nc, const double *c, const int *K1, const int *K2, const int *K3, const int *K4,
                  const double *pq_dev, int norb);
/* L_data (DBSR_MAT3/L_data.f90:2-145) inputs beyond the MAT3 tables: the
 * target overlaps otarg (host, may be NULL; Target_h's CC), check_target,
 * Lval(i,j) of bound orbitals as a host table lval(norb,norb) (Lval is in
 * the un-vendored Libraries_03), the orbitals' kappas kbs(norb), and the
 * one-electron operators hd_dev(ms,ms,nkap) of kap_list(nkap) (module
 * dhl_core; Gen_dhl_core, DBSR_MAT3/sub1.f90:23). */
typedef struct dbs_ldata_aux {
  const dbs_mat3_aux *mat3;
  double *otarg;
  int check_target;
  int norb;
  const double *lval;
  const int *kbs;
  int nkap;
  const int *kap_list;
  const double *hd_dev;
} dbs_ldata_aux;
/* L_data(itype) (DBSR_MAT3/L_data.f90:2-145): one merged list (IPT order,
 * host arrays) of L-integral structure itype into the MAT3 matrices.
 * 1: C*Lval(k1,k2) into UPDATE_HB (k3,k4 > 0), UPDATE_HV (k3 < 0, k4 > 0),
 *    UPDATE_HW (k3, k4 < 0), or htarg and, for iitar > 0 and ich != jch,
 *    UPDATE_HX(fppqq) (k3 < 0, k4 = 0);
 * 2: Gen_Lvec(k1) = hd(kappa of k1) (P,Q)(k1) into UPDATE_HV(k2, k3) (k3 >
 *    0) or UPDATE_HW(k2, ipbs(j1), ., V_ch(KBORT)) (k3 < 0);
 * 3: UPDATE_HX(ipbs(k1), ipbs(k2), hd(kappa)) when the channels are equal or
 *    iitar > 1, else the target overlap (check_target = 1).
 * Entries with |c| < eps_c are skipped; updates of one element in list
 * order.  pq_dev(ns,2,norb) (itype 2). */
int dbs_ldata_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, const dbs_ldata_aux *aux, int itype, int n,
                  const int *k1, const int *k2, const int *k3, const int *k4, const double *c, double eps_c,
                  const double *pq_dev);
/* hf_matrix, two-electron part (DBSR_HF/hf_matrix.f90:2-78 with
 * mrk_aaaa :82-92; Breit block hf_matrix_breit :97-163 when mbreit == 2):
 * for orbital i (1-based) of pq_dev(ns,2,norb), with acoef/bcoef(norb,
 * kmax+1) the reference's a(i,j,k) / b(i,j,k) (host), adds the direct and
 * exchange potentials of every (P/Q block, k) to hfm_dev(ms,ms) by UPDATE_HS
 * in the reference's order, then divides hfm by qsum(i).  The caller puts
 * qsum(i)*dhl (one-electron operator) or zeros into hfm first.  Breit list
 * (rk4_data, DBSR_HF/dbsr_lib_zcom.f90:2984-3003): kr1 = k + 1000*itype,
 * kr2 = int (1..4), kr3, kr4 orbitals, crk coefficients (host). */
int dbs_hf_matrix_dev(dbs_ctx *ctx, int i, int norb, const double *pq_dev, const double *acoef,
                      const double *bcoef, int kmax, double qsum, int mbreit, int nrk, const int *kr1,
                      const int *kr2, const int *kr3, const int *kr4, const double *crk,
                      double *hfm_dev);
/* mchf_matrix (DBSR_MCHF/mchf_matrix.f90:2-248) for the non-core orbital i
 * (1-based) of pq_dev(ns,2,norb): hfm_dev(ms,ms), rhs_dev(ms), hx_dev(ms,ms)
 * (device, overwritten) and, on the host, e(i) = v.(hfm v - rhs) and
 * qsum(i).  dhl_dev(ms,ms) is the one-electron operator Get_hd_core(kbs(i))
 * (MOD_hd_core.f90:337-356).  The angular lists are flattened, one entry
 * per coefficient with c = WW(ic,jc)*coef (mchf_matrix.f90:64, 94): L
 * integrals (l_i1, l_i2, l_c) and R^k integrals (r_k, r_i1..r_i4, r_c);
 * entries with |c/qsum| < eps_c are dropped, |qsum| < qsum_min takes
 * qsum_min (DBSR_MCHF/MOD_dbsr_mchf.f90:52-55: 1e-8, 1e-8).  newton != 0
 * builds hx; irhs > 0 moves rhs into hfm with fppqq_dev (ms,ms), the B-spline
 * overlap matrix.  Equal keys (k, flag, id3, id4) of the collection list sum
 * their coefficients, in order of first insertion (module id4_data is in the
 * un-vendored Libraries_03); hx is the case-4 convolution of the LAST Newton
 * density of each (k, P/Q block) without its coefficient, as the reference
 * computes it (:222-231). */
int dbs_mchf_matrix_dev(dbs_ctx *ctx, int i, int norb, const double *pq_dev, const double *dhl_dev, int nl,
                        const int *l_i1, const int *l_i2, const double *l_c, int nr, const int *r_k,
                        const int *r_i1, const int *r_i2, const int *r_i3, const int *r_i4, const double *r_c,
                        int kmax, double eps_c, double qsum_min, int newton, int irhs, const double *fppqq_dev,
                        double *hfm_dev, double *rhs_dev, double *hx_dev, double *e_i, double *qsum_i);
/* core_matrix (DBSR_MCHF/mchf_matrix.f90:253-345) for the core orbital i:
 * hfm_dev = (qsum(i) hd + direct + exchange)/qsum(i) with the coefficients
 * a(i,j,k) qsum(i) qsum(j) (qsum(i)(qsum(i)-1) for j = i) and b likewise;
 * hd_dev(ms,ms) = Get_hd(kbs(i)); acoef/bcoef(norb, kmax+1) host; qsum(norb)
 * host; hx_dev the Newton term (zero unless newton != 0). */
int dbs_core_matrix_dev(dbs_ctx *ctx, int i, int norb, const double *pq_dev, const double *hd_dev,
                        const double *acoef, const double *bcoef, int kmax, const double *qsum, int newton,
                        double *hfm_dev, double *hx_dev);
/* hf_eiv (DBSR_HF/hf_solve_HF.f90:118-202): the orbital update of DBSR_HF.
 * hfm_dev(ms,ms) (overwritten) is first projected orthogonal to each of the
 * northo vectors vorth_dev(ms,northo) (orthogonality,
 * DBSR_HF/hf_orthogonality.f90:2-29: H <- (1 - B v v^T) H (1 - v v^T B),
 * B = fppqq_dev), the B-splines with iprm[r] = 0 (host, ms entries;
 * Boundary_conditions, DBSR_HF/hf_boundary.f90:2-32) are removed, H x = e B x
 * is solved on the device (rocSOLVER dsygvd, lower triangles, as dsygv there),
 * and the m-th eigenvalue above -1.999 c^2 (m = n - l minus the orthogonality
 * conditions to lower orbitals) is returned in *e_i with its eigenvector,
 * B-normalised, on the full B-spline net in v_dev(ms), its largest entry
 * positive.  Stop texts: "Error in Eigenvalue routine, dsygv". */
int dbs_hf_eiv_dev(dbs_ctx *ctx, double *hfm_dev, int northo, const double *vorth_dev, const int *iprm, int m,
                   const double *fppqq_dev, double *v_dev, double *e_i);
/* BAV (called at DBSR_MAT3/R_data.f90:181, S_data.f90:205; defined in the
 * un-vendored Libraries_03): c(ns) = A b (dir 'r') or A^T b (dir 'l'), A the
 * (ns,ns) matrix held in storage sym 's','l','n','x'. */
int dbs_bav_dev(dbs_ctx *ctx, const double *a_dev, const double *b_dev, char sym, char dir,
                double *c_dev);

/* ---------------------------------------------------------------------
 * Channel-block post-processing of DBSR_MAT3 (dbs_channels.hip): zero
 * conditions, orthogonality conditions, diagonalisation of the diagonal
 * blocks and the transformation of the other blocks to the channel
 * solutions (DBSR_MAT3/sub1.f90:133-170, 258-265).
 * --------------------------------------------------------------------- */
/* Zero_cond (DBSR_MAT3/zero_cond.f90:2-56): iprm(nch*2*ns + npert) of 0/1
 * (0 = B-spline removed) and nhm = the number kept + npert.  lch, kch: l and
 * kappa of channel ich at element ich-1; ilzero/jlzero < 0 take the
 * l-dependent rule at r = 0, ibzero/jbzero B-splines are removed at r = R;
 * the perturber entries are 1.  Host only. */
int dbs_zero_cond(int nch, int ns, int nsp, int nsq, int ksp, int ksq, const int *lch, const int *kch,
                  int ilzero, int jlzero, int ibzero, int jbzero, int npert, int *iprm, int *nhm);
/* IBORT(i,j) of the host program: 0 when orbital i is to be kept orthogonal
 * to orbital j (the reference's external IBORT) */
typedef int (*dbs_ibort_fn)(void *user, int i, int j);
/* the conditions DBS_ORTH imposes (dbs_orth.f90:36-42): for channel ich,
 * whose orbital is ipch[ich-1], the bound orbitals j = 1..nbf (ascending)
 * with ipbs[j-1] <= 0, kbs[j-1] = kbs[ipch[ich-1]-1] and IBORT(i,j) = 0, into
 * iort[(ich-1)*ldo + n], n < nort[ich-1].  Error when a channel has more
 * than ldo.  Host only. */
int dbs_orth_conditions(int nch, const int *ipch, int nbf, const int *ipbs, const int *kbs,
                        dbs_ibort_fn ibort, void *ibort_user, int ldo, int *nort, int *iort);
/* DBS_ORTH (DBSR_MAT3/dbs_orth.f90:2-84): with P_i, Q_i the columns
 * pv_dev(:,j), qv_dev(:,j) (ms each: Get_pv / Get_qv of orbital j, nbf
 * columns) of channel i's conditions iort (dbs_orth_conditions), every
 * block idiag selects (+1 diagonal blocks, -1 the others, 0 all) becomes
 * (1 - Q_ich P_ich^T) hch(:,:,icc(ich,jch)) (1 - P_jch Q_jch^T), ich >= jch
 * (a side without conditions is left as it is), and for idiag = 0 every
 * perturber column hcp(:,icb(ich,ip)) becomes (1 - Q_ich P_ich^T) hcp.
 * Same products as the reference's matmul sequence, summed in another
 * order. */
int dbs_orth_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, int idiag, const int *nort, const int *iort, int ldo,
                 const double *pv_dev, const double *qv_dev, int nbf);
/* Diag_channels (DBSR_MAT3/diag.f90:2-79) after Zero_cond (iprm from
 * dbs_zero_cond, host): per diagonal block k = icc(ich,ich) the
 * generalised problem hch(:,:,k) x = E diag(:,:,k) x on the B-splines with
 * iprm = 1 (lower triangles), the solutions with edmin <= E <= edmax and
 * |E| >= egap, in ascending E, written to diag(:,ksol,k) (B-normalised,
 * largest entry positive, zero on removed B-splines) with E in
 * diag(ksol,ns,k), in the reference's write order (a solution ns overwrites
 * the eigenvalues stored in its column before it).  ipsol[0..nch]:
 * ipsol[0] = 0, ipsol[ich] = ksol.  diag_dev(ms,ms,nd) device; error
 * "channel diagonalization failed" when the overlap is not positive
 * definite. */
int dbs_diag_channels_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, double *diag_dev, int nd, const int *iprm,
                          double edmin, double edmax, double egap, int *ipsol);
/* Record_matrix (DBSR_MAT3/record_matrix.f90:2-60) on the device, without
 * the file: D_ich = diag_dev(:,1:ipsol[ich],ich) (channel solutions by
 * channel, nd >= nch).  For ich > jch with k = icc(ich,jch) != 0 and
 * max|hch(:,:,k)| >= 1e-20 the record c = D_ich^T hch(:,:,k) D_jch
 * (ipsol[ich] x ipsol[jch], column-major) goes to c_dev + coff[(ich-1) +
 * nch*(jch-1)]; for k = icb(ich,ip) != 0 and max|hcp(:,k)| >= 1e-20 the
 * record w = D_ich^T hcp(:,k) goes to w_dev + woff[(ich-1) + nch*(ip-1)].
 * Offsets (host, in doubles) run in record order, -1 = record not
 * written; c_cap / w_cap: capacities of c_dev / w_dev in doubles.
 * overlaps (host, (nch+npert)^2, column-major) or NULL: set as icase = 0
 * sets it (max|c|, max|w|, |hp(k)| of the written records).  The
 * perturber-perturber records are hp itself. */
int dbs_record_matrix_dev(dbs_ctx *ctx, const dbs_mat_blocks *m, const double *diag_dev, int nd,
                          const int *ipsol, double *c_dev, long long c_cap, long long *coff, double *w_dev,
                          long long w_cap, long long *woff, double *overlaps);

/* ---------------------------------------------------------------------
 * Engine: batched evaluation of integral sets and their adjoint.
 * An instance = (context, norb orbitals, rows (k,a,b,c,d)).  Row n is
 *   kind DBS_RK_TOTAL: R^k(a,b;c,d) = rk(f_a,f_b,f_c,f_d,k)
 *                      (dbsr_lib_dbs.f90:3981-4022; hf_rk DBSR_HF/hf_rk.f90:2-31)
 *   kind 1..4: the single-type term of that sum (Rk_evaluate types),
 *   kind 5/6: S^k ppqq / pqqp (Sk_evaluate, DBSR_CI3/sk_evaluate.f90).
 * The adjoint returns E = sum_n coef_n*value_n and dE/d(pq), i.e. the
 * direct+exchange potentials applied to the orbitals that hf_matrix builds
 * (DBSR_HF/hf_matrix.f90:2-71).  Many instances of different ns/ks/kmax
 * run in one persistent launch per phase; the sequence can be captured in
 * a hipGraph and replayed.
 * ------------------------------------------------------------------- */
int dbs_engine_create(dbs_engine **eng, int ninst, dbs_ctx *const *ctx, const int *norb,
                      const int *nrows, const int *rows, int kind);
int dbs_engine_destroy(dbs_engine *eng);
/* device pointers into the engine's buffers for instance inst:
 * pq (ns,2,norb) input, coef (nrows) input, values (nrows) output,
 * grad (ns,2,norb) output, energy (1) output; any may be NULL */
int dbs_engine_buffers(dbs_engine *eng, int inst, double **pq, double **coef, double **values,
                       double **grad, double **energy);
int dbs_engine_forward(dbs_engine *eng);
int dbs_engine_forward_adjoint(dbs_engine *eng);
/* the same values by the plain path -- materialised arrays (mrk_ and msk_),
 * one convol case 2 per distinct left density and one SUM_AmB per row, the
 * organisation of Rk_evaluate (DBSR_CI3/rk_evaluate.f90:31-56) and rk()
 * (dbsr_lib_dbs.f90:3981-4022).  Runs on the instances' context streams and
 * returns after they finish. */
int dbs_engine_forward_plain(dbs_engine *eng);
/* capture forward(+adjoint) into a hipGraph; launch replays it */
int dbs_engine_graph_capture(dbs_engine *eng, int with_adjoint);
int dbs_engine_graph_launch(dbs_engine *eng);
/* tuning parameters (see dbsr3_amd/csrc/dbs_engine.hip); autotune times
 * the candidates on the engine's own data and keeps the fastest */
int dbs_engine_set_tuning(dbs_engine *eng, int nparam, const int *params);
int dbs_engine_get_tuning(dbs_engine *eng, int nparam, int *params);
int dbs_engine_autotune(dbs_engine *eng, int with_adjoint, int reps);
/* Memory budget of the features and their gradients (bytes of both; 0 =
 * no budget, the default): the groups then run in batches of consecutive
 * (instance, k) groups whose features and gradients fit, the two buffers
 * reused by every batch: each batch's features are computed, contracted and
 * turned into its share of the orbital gradient before the next batch's
 * overwrite them.  A group larger than the budget is a batch of its own.
 * The orbital gradient of later batches continues the sums of earlier ones,
 * so the results do not depend on the batching; under a budget the
 * register-summed adjoint kernel is not used.  memory: the batch count and
 * the feature + gradient bytes held. */
int dbs_engine_set_memory_budget(dbs_engine *eng, size_t bytes);
int dbs_engine_memory(dbs_engine *eng, int *nbatch, size_t *feature_bytes);
/* algorithmic flop count and minimum bytes of one forward (+adjoint) */
int dbs_engine_work(dbs_engine *eng, int with_adjoint, double *flops, double *bytes);
int dbs_engine_sync(dbs_engine *eng);
/* device time of each phase of forward(+adjoint), from events recorded on
 * the engine's stream after every launch, averaged over reps; names holds
 * nmax entries of 32 chars; flops/bytes = that phase's algorithmic work
 * (minimum HBM bytes) per launch; *nphase = number of phases returned */
int dbs_engine_profile(dbs_engine *eng, int with_adjoint, int reps, int nmax, char *names,
                       double *ms, double *flops, double *bytes, int *nphase);

#ifdef __cplusplus
}
#endif
#endif

// dbs_internal.h -- shared definitions for the MI355X B-spline Slater-integral
// library (libdbsr3_amd.so).  Host runtime + HIP kernels for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "dbsr3_amd.h"

#define DBS_MAXK 16  // largest B-spline order supported by fixed-size device scratch

namespace dbs {

// ---------------------------------------------------------------------------
// error handling: every C entry point returns a status; the last message is
// kept per thread and returned by dbs_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define DBS_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::dbs::fail(DBS_ERR_HIP, std::string(#call) + ": " +              \
                                          hipGetErrorString(e_));              \
  } while (0)

#define DBS_CHECK(cond, code, msg)                                             \
  do {                                                                         \
    if (!(cond)) return ::dbs::fail((code), (msg));                            \
  } while (0)

// device buffer owned by the library
struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
};

}  // namespace dbs

// ---------------------------------------------------------------------------
// the context: one B-spline grid, its Gauss tables resident in HBM, a stream,
// and a cache of materialised R^k/S^k arrays keyed by (type, k).
// ---------------------------------------------------------------------------
struct dbs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int ns = 0, ks = 0, ksp = 0, ksq = 0, nv = 0, nsp = 0, nsq = 0, ms = 0;

  // dense-solver handle of the context (created on first use by the entry
  // point that needs it) and the routine that releases it
  void* solver = nullptr;
  void (*solver_free)(void*) = nullptr;
  // knots (1-based in the reference; here t[i-1])
  double* t = nullptr;   // ns+ks
  double* tp = nullptr;  // nsp+ksp
  double* tq = nullptr;  // nsq+ksq
  double* gx = nullptr;  // ks Gauss-Legendre nodes on [0,1]
  double* gw = nullptr;  // ks weights
  // outer Gauss points, column-major (nv,ks) as in the reference
  double* gr = nullptr;
  double* grw = nullptr;
  double* grm = nullptr;
  // B-spline values at the outer points: (nv+1,ks,ksp) / (nv+1,ks,ksq)
  double* pbsp = nullptr;
  double* qbsp = nullptr;
  // nested (inner) rule of the diagonal cells: for interval iv, outer point m,
  // inner point i: x = t_iv + (gr(iv,m)-t_iv)*gx(i), weight (gr-t_iv)*gw(i).
  // Layout [iv][m][i] (row-major), B values [iv][m][i][j].
  double* xin = nullptr;
  double* win = nullptr;
  double* bpin = nullptr;
  double* bqin = nullptr;

  // host copies of small tables
  std::vector<double> h_t, h_gx, h_gw;

  // materialised arrays (plain path), key = type*1000 + k
  std::map<int, double*> rkb_cache;
  std::vector<int> rkb_lru;
  size_t rkb_cache_limit = (size_t)8 << 30;  // bytes
  size_t rkb_cache_bytes = 0;
  // device workspace per batch of blocks of the channel post-processing
  size_t work_limit = (size_t)256 << 20;
  // scratch for the array build
  double* mom = nullptr;  // 4 x (ks*ks, nv)
  double* rkd = nullptr;  // (Tmax, Tmax, nv)
  // general scratch for contractions
  double* scratch = nullptr;
  size_t scratch_n = 0;
  // staging area of the host-pointer entry points
  double* stage = nullptr;
  size_t stage_n = 0;
  // pinned host tables of the list calls (Rk_evaluate): two buffers used in
  // turn, each reusable once the event recorded after its upload has fired,
  // so a call returns without waiting and the host builds the next list's
  // tables while the device runs this one
  char* list_host[2] = {nullptr, nullptr};
  size_t list_host_n[2] = {0, 0};
  hipEvent_t list_ev[2] = {nullptr, nullptr};
  int list_cur = 0;
};

namespace dbs {
int ensure_scratch(dbs_ctx* c, size_t n);
int ensure_stage(dbs_ctx* c, size_t n);
int kband(int type, int ks);  // third/fourth dimension of rkb for the type
int build_rkb(dbs_ctx* c, int type, int k, double** out);
// rkb permuted to the K x K matrix rkm(i + ns*ip, j + ns*jp), K = ns*kb,
// cached with the arrays (negative keys in rkb_cache, one LRU and limit)
int build_rkm(dbs_ctx* c, int type, int k, double** out);
// pinned host tables of the queued list calls (Rk_evaluate, R_data): the
// context's next buffer, grown to at least `bytes`, once the upload that
// last read it has finished; after queueing the upload from it on the
// context stream, list_stage_end records that and moves to the other one
int list_stage_begin(dbs_ctx* c, size_t bytes, char** host);
int list_stage_end(dbs_ctx* c);
// UPDATE_ACF (updates.f90) of a list, queued on the context stream without
// waiting (dbs_update_acf_dev waits)
int update_acf_queued(dbs_ctx* ctx, const dbs_mat_blocks* m, int n, const int* k, const int* ich, const int* jch,
                      const double* C);
// convol of n densities (K x n, one after another) against one array on the
// context stream (rocBLAS dgemm): case 1 conv = rkm dens, case 2
// conv = rkm^T dens (every output of the K x n result written; the entries
// outside the band are the products' and not zero)
int conv_gemm(dbs_ctx* c, int icase, int K, int n, const double* rkm, const double* dens, double* conv);
}  // namespace dbs

namespace dbs {
// convol on the context stream with a materialised array rkb of the type
// (the kernels of dbs_convol_dev, without its checks)
int convol_launch(dbs_ctx* c, const double* rkb, int type, int icase, char si, char sj, const double* d,
                  double* a);

// ---------------------------------------------------------------------------
// MAT3 matrix updates shared by O_data, R_data and S_data (dbs_updates.hip)
// ---------------------------------------------------------------------------
// One update of a list.  target: hp element (HB), hcp column (HV) or hch
// block (HW, HX), 0-based; -1 = not stored here (dropped).  Vectors: index
// >= 0 is a column of V (V_ch), < 0 column -idx-1 of W (computed, length ms).
struct MatUpd {
  enum { HB = 0, HV = 1, HW = 2, HX = 3 };
  int kind, mode, v, w;
  long long target;
  double C;
};
// applies the updates, in list order per target (stable grouping, one
// thread per matrix element); synchronises the context stream
int apply_mat3(dbs_ctx* ctx, const dbs_mat_blocks* m, const std::vector<MatUpd>& u, const double* V, int ldv,
               const double* W, const double* d, const char* who);
// icc(ich,jch) with the reference's range checks; *b = 0: block not stored
int chan_block(const dbs_mat_blocks* m, int ich, int jch, const char* who, int* b);
int hb_target(const dbs_mat_blocks* m, int ic, int jc, long long* t);
int hv_column(const dbs_mat_blocks* m, int ich, int ip, int* col);
int check_blocks(const dbs_mat_blocks* m);
int target_h(const dbs_mat_blocks* m, int ich, int jch, const int* iptar, double* tgt, double C);
}  // namespace dbs

// the context's dense-solver handle (rocBLAS/rocSOLVER), created on first
// use and bound to the context stream (dbs_mat3.hip)
struct _rocblas_handle;
namespace dbs {
int solver_of(dbs_ctx* ctx, _rocblas_handle** h);
}  // namespace dbs

// dbs_device.h -- small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#ifndef DBS_MAXK
#define DBS_MAXK 16
#endif

namespace dbs {

// x**n for integer n by binary powering: the rounding sequence of the
// Fortran real**integer the reference uses for r**k.
__host__ __device__ inline double ipow(double a, int b) {
#pragma clang fp contract(off)
  bool recip = b < 0;
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

// Values of the kx B-splines that do not vanish on interval [T(left),
// T(left+1)) at x (de Boor's bsplvb recurrence, the order of operations of
// zbsplvb, DBSR_HF/dbsr_lib_dbs.f90:4355-4466).  tx is 0-based: T(i)=tx[i-1].
__device__ inline void bspline_values(const double* __restrict__ tx, int kx, int left, double x,
                                      double* b) {
#pragma clang fp contract(off)
  double dr[DBS_MAXK + 1], dl[DBS_MAXK + 1];
  b[0] = 1.0;
  for (int j = 1; j < kx; ++j) {
    const int jp1 = j + 1;
    dr[j] = tx[left + j - 1] - x;
    dl[j] = x - tx[left - j];
    double saved = 0.0;
    for (int m = 1; m <= j; ++m) {
      double term = b[m - 1] / (dr[m] + dl[jp1 - m]);
      b[m - 1] = saved + dr[m] * term;
      saved = dl[jp1 - m] * term;
    }
    b[jp1 - 1] = saved;
  }
}

// index of pair (a,b), a<=b<k, in the packed upper-triangle order of the
// reference (do i=1,k; do j=i,k)
__host__ __device__ inline int packed_index(int a, int b, int k) {
  return a * k - (a * (a - 1)) / 2 + (b - a);
}

// wavefront (64-lane) sum
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace dbs

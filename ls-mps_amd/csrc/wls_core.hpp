// Device pieces of the 9x9 WLS solve shared by the WLS operator kernels
// and the WLS-Laplacian PPE rows: packed lower-triangle indexing, an
// in-register Cholesky and the two triangular solves, in the operation
// order of the reference's cholesky() and solve() (src/lsmps.cpp:24-25).
#pragma once
#include "common.hpp"

namespace lsmps_hip {

__device__ __forceinline__ constexpr int tri(int a, int c) { return a * (a + 1) / 2 + c; }

// In-register lower Cholesky of a packed 9x9 (45 entries), fully unrolled.
// Returns false at the first pivot that is not > thr or not finite (the
// remaining entries are then garbage, and unused).
template <typename T>
__device__ __forceinline__ bool chol9(T (&m)[45], T thr_abs, T thr_rel_scale, const T (&diag0)[9]) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      T s = m[tri(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= m[tri(i, k)] * m[tri(j, k)];
      if (i == j) {
        const T lim = thr_abs > thr_rel_scale * diag0[i] ? thr_abs : thr_rel_scale * diag0[i];
        if (!(s > lim) || !isfinite(s)) ok = false;
        m[tri(i, i)] = sqrt(s);
      } else {
        m[tri(i, j)] = s / m[tri(j, j)];
      }
    }
  }
  return ok;
}

template <typename T>
__device__ __forceinline__ void chol9_solve(const T (&L)[45], const T (&b)[9], T (&x)[9]) {
  T y[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    T s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[tri(i, k)] * y[k];
    y[i] = s / L[tri(i, i)];
  }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
    T s = y[i];
#pragma unroll
    for (int k = i + 1; k < 9; ++k) s -= L[tri(k, i)] * x[k];
    x[i] = s / L[tri(i, i)];
  }
}

}  // namespace lsmps_hip

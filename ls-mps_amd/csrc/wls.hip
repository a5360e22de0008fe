// Second-order WLS gradient/Laplacian operator on gfx950.
//
// Semantics of the reference's wls_test per particle (src/lsmps.cpp:22-25,
// 28): weight (re/r - 1)^2 on 1e-12 < r < re, 9 scaled monomials, moment
// matrix sum w q q^T and right-hand side sum w q df / r, both divided by
// the weight sum, lower Cholesky with the diagonal-shift ladder 0, reg,
// 10 reg, 100 reg, 1000 reg (pivot > 1e-14 and finite), forward/back
// substitution; gradient = x[0..2], Laplacian = (x3 + x4 + x5) / l0.
//
// Two kernels:
//  * k_wls_exact  FP64, no contraction, same operation order as the
//    reference, so the results are bit-identical to the reference's. Only
//    the lower triangle is accumulated: it is the part the factorisation
//    reads, and each entry is formed as (w q_a) q_c with a >= c, exactly as
//    the reference forms it. One thread per particle; the ladder re-runs
//    the (deterministic) accumulation instead of keeping a copy of M, which
//    keeps the thread at ~70 live doubles.
//  * k_wls_mixed  FP32 moments on offsets normalised by re, FMA
//    accumulation, FP32 Cholesky. A row whose FP32 pivots lose more than
//    three digits (pivot <= 1e-3 * diagonal), or that would need the
//    reference's regularisation, is appended to a work list and redone by
//    k_wls_exact_list in FP64 with the reference's ladder.
#include "kernels.hpp"

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

// FP64 solve of particle i with the reference's operation order. Returns
// the status; writes grad/lap.
__device__ __noinline__ uint8_t wls_exact_one(int i, const double* __restrict__ x, const double* __restrict__ y,
                                              const double* __restrict__ z, const double* __restrict__ f,
                                              const int* __restrict__ nbr, const int* __restrict__ cnt, int maxn,
                                              double re, double l0, double reg, double* g3, double* lap) {
  const double xi = x[i], yi = y[i], zi = z[i], fi = f[i];
  const int* col = nbr + (size_t)i * maxn;
  const int rb = 0, rend = cnt[i];
  const double h2 = 2 * l0;
  for (int attempt = 0; attempt < 5; ++attempt) {
    double M[45], B[9], ws = 0.0;
#pragma unroll
    for (int k = 0; k < 45; ++k) M[k] = 0.0;
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] = 0.0;
    for (int k = rb; k < rend; ++k) {
      const int j = col[k];
      const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r <= 1e-12 || r >= re) continue;
      const double tq = re / r - 1.0;
      const double w = tq * tq;
      if (w == 0.0) continue;
      const double inv = 1.0 / r, h = h2 * r, g = l0 * r;
      double q[9];
      q[0] = dx * inv;
      q[1] = dy * inv;
      q[2] = dz * inv;
      q[3] = dx * dx / h;
      q[4] = dy * dy / h;
      q[5] = dz * dz / h;
      q[6] = dx * dy / g;
      q[7] = dx * dz / g;
      q[8] = dy * dz / g;
      const double df = f[j] - fi;
      ws += w;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const double wq = w * q[a];
        B[a] += wq * df / r;
#pragma unroll
        for (int c = 0; c <= a; ++c) M[tri(a, c)] += wq * q[c];
      }
    }
    if (ws <= 0) return WLS_NO_WEIGHT;
#pragma unroll
    for (int k = 0; k < 45; ++k) M[k] /= ws;
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] /= ws;
    const double add = attempt == 0 ? 0.0 : attempt == 1 ? reg * 1.0 : attempt == 2 ? reg * 10.0
                                                            : attempt == 3 ? reg * 100.0 : reg * 1000.0;
    if (attempt > 0 && add == 0.0) return WLS_FAILED;  // zero reg: every retry equals attempt 0
    if (add != 0.0) {
#pragma unroll
      for (int a = 0; a < 9; ++a) M[tri(a, a)] += add;
    }
    const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!chol9<double>(M, 1e-14, 0.0, zero9)) continue;
    double sol[9];
    chol9_solve<double>(M, B, sol);
    g3[0] = sol[0];
    g3[1] = sol[1];
    g3[2] = sol[2];
    *lap = (sol[3] + sol[4] + sol[5]) / l0;
    return attempt == 0 ? WLS_OK : WLS_REGULARIZED;
  }
  return WLS_FAILED;
}

// Block-aggregated status histogram (statuses 0..4 and the redo slot);
// every thread of the block must call it, with slot < 0 for "nothing".
__device__ __forceinline__ void count_status(int slot, unsigned long long* __restrict__ counts) {
  __shared__ unsigned c[WLS_REDO_SLOT + 1];
  if (threadIdx.x <= WLS_REDO_SLOT) c[threadIdx.x] = 0u;
  __syncthreads();
  if (slot >= 0) atomicAdd(&c[slot], 1u);
  __syncthreads();
  if (threadIdx.x <= WLS_REDO_SLOT && c[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)c[threadIdx.x]);
}

__device__ __forceinline__ void store_result(int i, uint8_t st, const double* g3, double lap, double* __restrict__ grad,
                                             double* __restrict__ lapo, uint8_t* __restrict__ status) {
  grad[3 * (size_t)i] = g3[0];
  grad[3 * (size_t)i + 1] = g3[1];
  grad[3 * (size_t)i + 2] = g3[2];
  lapo[i] = lap;
  status[i] = st;
}

__global__ __launch_bounds__(128) void k_wls_exact(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const double* __restrict__ f,
                                                   const uint8_t* __restrict__ type, const int* __restrict__ nbr,
                                                   const int* __restrict__ cnt, int maxn, int n, double re, double l0,
                                                   double reg, double* __restrict__ grad, double* __restrict__ lap,
                                                   uint8_t* __restrict__ status,
                                                   unsigned long long* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int slot = -1;
  if (i < n) {
    const uint8_t t = type[i];
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    uint8_t st = WLS_SKIPPED;
    if (t != T_WALL && t != T_SPLASH) st = wls_exact_one(i, x, y, z, f, nbr, cnt, maxn, re, l0, reg, g3, &L);
    store_result(i, st, g3, L, grad, lap, status);
    slot = st;
  }
  count_status(slot, counts);
}

// Redo in FP64 the particles the FP32 kernel could not trust.
__global__ __launch_bounds__(128) void k_wls_exact_list(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const double* __restrict__ f,
                                                        const int* __restrict__ nbr, const int* __restrict__ cnt,
                                                        int maxn, const int* __restrict__ list,
                                                        const int* __restrict__ nlist,
                                                        double re, double l0, double reg, double* __restrict__ grad,
                                                        double* __restrict__ lap, uint8_t* __restrict__ status,
                                                        unsigned long long* __restrict__ counts) {
  const int nl = *nlist;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    const int i = list[k];
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    const uint8_t st = wls_exact_one(i, x, y, z, f, nbr, cnt, maxn, re, l0, reg, g3, &L);
    store_result(i, st, g3, L, grad, lap, status);
    atomicAdd(&counts[st], 1ull);
    atomicAdd(&counts[WLS_REDO_SLOT], 1ull);
  }
}

// FP32 moments on offsets scaled by 1/re (so re -> 1, l0 -> l0/re);
// the solution vector is unchanged by the scaling because the monomials are
// dimensionless and df/r = (df/re)/(r/re).
// Returns the status slot to count (-1 when the row went to the FP64 list).
__device__ __forceinline__ int wls_mixed_one(int i, const double* __restrict__ x, const double* __restrict__ y,
                                             const double* __restrict__ z, const double* __restrict__ f,
                                             const uint8_t* __restrict__ type, const int* __restrict__ nbr,
                                             const int* __restrict__ cnt, int maxn, double re, double l0,
                                             double* __restrict__ grad, double* __restrict__ lap,
                                             uint8_t* __restrict__ status, int* __restrict__ redo,
                                             int* __restrict__ nredo) {
  const uint8_t t = type[i];
  if (t == T_WALL || t == T_SPLASH) {
    const double g0[3] = {0.0, 0.0, 0.0};
    store_result(i, WLS_SKIPPED, g0, 0.0, grad, lap, status);
    return WLS_SKIPPED;
  }
  const double inv_re = 1.0 / re;
  const float l0s = (float)(l0 * inv_re);
  const float inv_h2 = 1.0f / (2.0f * l0s), inv_g = 1.0f / l0s;
  const float rmin = (float)(1e-12 * inv_re);
  const double xi = x[i], yi = y[i], zi = z[i], fi = f[i];
  const int* col = nbr + (size_t)i * maxn;
  const int rb = 0, rend = cnt[i];
  float M[45], B[9], ws = 0.0f;
#pragma unroll
  for (int k = 0; k < 45; ++k) M[k] = 0.0f;
#pragma unroll
  for (int a = 0; a < 9; ++a) B[a] = 0.0f;
  for (int k = rb; k < rend; ++k) {
    const int j = col[k];
    const float dx = (float)((x[j] - xi) * inv_re), dy = (float)((y[j] - yi) * inv_re),
                dz = (float)((z[j] - zi) * inv_re);
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float r = sqrtf(r2);
    if (r <= rmin || r >= 1.0f) continue;
    const float inv = 1.0f / r;
    const float tq = inv - 1.0f;
    const float w = tq * tq;
    float q[9];
    q[0] = dx * inv;
    q[1] = dy * inv;
    q[2] = dz * inv;
    const float ih = inv_h2 * inv, ig = inv_g * inv;
    q[3] = dx * dx * ih;
    q[4] = dy * dy * ih;
    q[5] = dz * dz * ih;
    q[6] = dx * dy * ig;
    q[7] = dx * dz * ig;
    q[8] = dy * dz * ig;
    const float dfr = (float)((f[j] - fi) * inv_re) * inv;
    ws += w;
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      const float wq = w * q[a];
      B[a] = fmaf(wq, dfr, B[a]);
#pragma unroll
      for (int c = 0; c <= a; ++c) M[tri(a, c)] = fmaf(wq, q[c], M[tri(a, c)]);
    }
  }
  bool ok = ws > 0.0f;
  float sol[9];
  if (ok) {
    const float iw = 1.0f / ws;
#pragma unroll
    for (int k = 0; k < 45; ++k) M[k] *= iw;
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] *= iw;
    float d0[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) d0[a] = M[tri(a, a)];
    ok = chol9<float>(M, 1e-14f, 1e-3f, d0);
    if (ok) chol9_solve<float>(M, B, sol);
  }
  if (!ok) {
    const int at = atomicAdd(nredo, 1);
    redo[at] = i;
    return -1;
  }
  const double g3[3] = {(double)sol[0], (double)sol[1], (double)sol[2]};
  store_result(i, WLS_OK, g3, (double)(sol[3] + sol[4] + sol[5]) / l0, grad, lap, status);
  return WLS_OK;
}

__global__ __launch_bounds__(128) void k_wls_mixed(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, const double* __restrict__ f,
                                                   const uint8_t* __restrict__ type, const int* __restrict__ nbr,
                                                   const int* __restrict__ cnt, int maxn, int n, double re, double l0,
                                                   double* __restrict__ grad, double* __restrict__ lap,
                                                   uint8_t* __restrict__ status, int* __restrict__ redo,
                                                   int* __restrict__ nredo, unsigned long long* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int slot = -1;
  if (i < n) slot = wls_mixed_one(i, x, y, z, f, type, nbr, cnt, maxn, re, l0, grad, lap, status, redo, nredo);
  count_status(slot, counts);
}

// Quadratic field f = c0 + cx x + cy y + cz z + cxx x^2 + cyy y^2 + czz z^2
// + cxy x y + cxz x z + cyz y z, summed left to right with (c*u)*v
// products, which reproduces the reference's test-field expression bit for
// bit (src/lsmps.cpp:28).
__global__ __launch_bounds__(256) void k_field_quadratic(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z, int n, QuadField q,
                                                         double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i], b = y[i], c = z[i];
  f[i] = q.c[0] + q.c[1] * a + q.c[2] * b + q.c[3] * c + q.c[4] * a * a + q.c[5] * b * b + q.c[6] * c * c +
         q.c[7] * a * b + q.c[8] * a * c + q.c[9] * b * c;
}

// Errors of the sampled particles i = k*stride (original index) against the
// closed-form gradient and Laplacian of the quadratic field
// (src/lsmps.cpp:28); positions and results are read at the sorted slot.
__global__ __launch_bounds__(256) void k_wls_sample_errors(const double* __restrict__ x, const double* __restrict__ y,
                                                           const double* __restrict__ z, const uint8_t* __restrict__ type,
                                                           const double* __restrict__ grad,
                                                           const double* __restrict__ lap,
                                                           const uint8_t* __restrict__ status,
                                                           const int* __restrict__ iperm, int64_t stride, int ns,
                                                           QuadField q, double* __restrict__ ge,
                                                           double* __restrict__ le, uint8_t* __restrict__ st_out,
                                                           uint8_t* __restrict__ ty_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int64_t io = (int64_t)k * stride;  // original index; arrays below are in sorted order
  const int64_t i = iperm[io];
  const double a = x[i], b = y[i], c = z[i];
  const double ex = q.c[1] + 2 * q.c[4] * a + q.c[7] * b + q.c[8] * c;
  const double ey = q.c[2] + 2 * q.c[5] * b + q.c[7] * a + q.c[9] * c;
  const double ez = q.c[3] + 2 * q.c[6] * c + q.c[8] * a + q.c[9] * b;
  const double d0 = grad[3 * i] - ex, d1 = grad[3 * i + 1] - ey, d2 = grad[3 * i + 2] - ez;
  ge[k] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  le[k] = fabs(lap[i] - q.lap_exact);
  st_out[k] = status[i];
  ty_out[k] = type[io];
}

// WLS results from sorted order back to original particle order.
__global__ __launch_bounds__(256) void k_wls_unsort(const int* __restrict__ perm, const double* __restrict__ gs,
                                                    const double* __restrict__ ls, const uint8_t* __restrict__ ss, int n,
                                                    double* __restrict__ g, double* __restrict__ l,
                                                    uint8_t* __restrict__ st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  g[3 * (size_t)i] = gs[3 * (size_t)s];
  g[3 * (size_t)i + 1] = gs[3 * (size_t)s + 1];
  g[3 * (size_t)i + 2] = gs[3 * (size_t)s + 2];
  l[i] = ls[s];
  st[i] = ss[s];
}

__global__ __launch_bounds__(256) void k_gather_double(const int* __restrict__ perm, const double* __restrict__ v, int n,
                                                       double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) out[s] = v[perm[s]];
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_wls_unsort(const int* perm, const double* gs, const double* ls, const uint8_t* ss, int n, double* g,
                       double* l, uint8_t* st, hipStream_t stream) {
  if (n > 0) hipLaunchKernelGGL(k_wls_unsort, dim3(blocks(n, 256)), dim3(256), 0, stream, perm, gs, ls, ss, n, g, l, st);
}

void launch_gather_double(const int* perm, const double* v, int n, double* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_gather_double, dim3(blocks(n, 256)), dim3(256), 0, st, perm, v, n, out);
}

void launch_field_quadratic(const double* x, const double* y, const double* z, int n, const QuadField& q, double* f,
                            hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_field_quadratic, dim3(blocks(n, 256)), dim3(256), 0, st, x, y, z, n, q, f);
}

void launch_wls_exact(const double* x, const double* y, const double* z, const double* f, const uint8_t* type,
                      const int* nbr, const int* cnt, int maxn, int n, double re, double l0, double reg, double* grad,
                      double* lap, uint8_t* status, unsigned long long* counts, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_wls_exact, dim3(blocks(n, 128)), dim3(128), 0, st, x, y, z, f, type, nbr, cnt, maxn, n, re,
                       l0, reg, grad, lap, status, counts);
}

void launch_wls_mixed(const double* x, const double* y, const double* z, const double* f, const uint8_t* type,
                      const int* nbr, const int* cnt, int maxn, int n, double re, double l0, double reg, double* grad,
                      double* lap, uint8_t* status, int* redo, int* nredo, unsigned long long* counts,
                      hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_wls_mixed, dim3(blocks(n, 128)), dim3(128), 0, st, x, y, z, f, type, nbr, cnt, maxn, n, re, l0,
                     grad, lap, status, redo, nredo, counts);
  // The redo list length is only known on the device: a bounded grid
  // strides over it (typical lists are empty or a few hundred entries).
  const unsigned g = blocks(n, 128) < 512u ? blocks(n, 128) : 512u;
  hipLaunchKernelGGL(k_wls_exact_list, dim3(g), dim3(128), 0, st, x, y, z, f, nbr, cnt, maxn, redo, nredo,
                     re, l0, reg, grad, lap, status, counts);
}

void launch_wls_sample_errors(const double* x, const double* y, const double* z, const uint8_t* type,
                              const double* grad, const double* lap, const uint8_t* status, const int* iperm,
                              int64_t stride, int ns, const QuadField& q, double* ge, double* le, uint8_t* st_out,
                              uint8_t* ty_out, hipStream_t st) {
  if (ns > 0)
    hipLaunchKernelGGL(k_wls_sample_errors, dim3(blocks(ns, 256)), dim3(256), 0, st, x, y, z, type, grad, lap, status,
                       iperm, stride, ns, q, ge, le, st_out, ty_out);
}

}  // namespace lsmps_hip

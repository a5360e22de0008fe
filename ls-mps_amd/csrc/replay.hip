// Synthetic CSR replay on gfx950.
//
// The reference's calibration tool writes a customer-profile proxy as a
// CSR neighbor graph with particle types and WLS ill-conditioned flags
// (python/workload_calibration.py:236-264), meant for GPU kernel
// performance tests without geometry: loop lengths, divergence, indirect
// access, SpMV and fallback (WORKLOAD_CALIBRATION.md). The graph has no
// positions, so the replay gives every edge a synthetic offset from a
// counter-based hash and runs the step's kernels on it:
//
//  * k_replay_ell     CSR rows -> the ELL neighbor table, each row sorted
//                     ascending (the PPE layout needs ascending columns);
//  * k_replay_wls     per row the reference's WLS (weight, basis, moments,
//                     src/lsmps.cpp:22-28) on the edges' synthetic offsets
//                     with the manufactured quadratic, FP32 moments and
//                     Cholesky; rows whose FP32 pivots are not trusted go
//                     to the FP64 redo list, taken by k_wls_exact-style
//                     code with the reference's ladder. Rows flagged
//                     ill-conditioned get near-planar offsets (z scaled by
//                     1e-4) and always take that fallback, so the redo
//                     rate is at least the profile's;
//  * the PPE fill of ppe.hip with synthetic symmetric edge weights.
//
// Offsets: edge q of row i has direction and length from splitmix64 hashes
// of (seed, i, q), the length in [0.35, 0.95) re scaled by the neighbor's
// per-node record (a gathered value, so the access pattern is the real
// kernel's: one 32-byte record per neighbor).
#include "kernels.hpp"
#include "wls_core.hpp"

namespace lsmps_hip {

// Per-node record: {0, 0, 0, u} with u in [0, 1) from the node's hash.
__global__ __launch_bounds__(256) void k_replay_nodes(int n, uint32_t seed, double4* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rec[i] = make_double4(0.0, 0.0, 0.0, unit53(mix64(((uint64_t)seed << 32) ^ (uint64_t)i ^ 0x5DEECE66Dull)));
}

__global__ __launch_bounds__(256) void k_replay_ell(const int64_t* __restrict__ row_ptr, const int* __restrict__ col,
                                                    int n, int maxn, int* __restrict__ nbr, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = row_ptr[i];
  const int k = (int)(row_ptr[i + 1] - b);
  int* out = nbr + (size_t)i * maxn;
  for (int q = 0; q < k; ++q) {
    const int v = col[b + q];
    int m = q - 1;
    while (m >= 0 && out[m] > v) {
      out[m + 1] = out[m];
      --m;
    }
    out[m + 1] = v;
  }
  cnt[i] = k;
}

// Offset of edge q of row i (FP64); uj is the neighbor's record value.
__device__ __forceinline__ void replay_offset(uint32_t seed, int i, int q, double uj, bool flat, double re,
                                              double* dx, double* dy, double* dz) {
  const uint64_t key = ((uint64_t)seed << 40) ^ ((uint64_t)(unsigned)i << 8) ^ (uint64_t)(unsigned)q;
  const uint64_t h1 = mix64(key * 0x9E3779B97F4A7C15ull + 1), h2 = mix64(h1 + 0x632BE59BD9B4E019ull),
                 h3 = mix64(h2 + 0x8CB92BA72F3D8DD7ull);
  // Direction: z uniform in [-1, 1), azimuth uniform (uniform on the sphere).
  const double z = 2.0 * unit53(h1) - 1.0, phi = 6.283185307179586 * unit53(h2);
  const double s = sqrt(fmax(0.0, 1.0 - z * z));
  double t = unit53(h3) + uj;
  t -= floor(t);
  const double len = re * (0.35 + 0.6 * t);
  *dx = len * s * cos(phi);
  *dy = len * s * sin(phi);
  *dz = len * z * (flat ? 1e-4 : 1.0);
}

// The reference's manufactured quadratic (src/lsmps.cpp:28) at offset d
// minus its value at the origin.
__device__ __forceinline__ double quad_df(double x, double y, double z) {
  return 2 * x - 3 * y + .5 * z + .7 * x * x - .2 * y * y + .4 * z * z + .3 * x * y - .1 * x * z + .25 * y * z;
}

// FP64 solve of replay row i with the reference's operation order and
// ladder (as wls_exact_one, on synthetic offsets).
__device__ __noinline__ uint8_t replay_exact_one(int i, const int* __restrict__ row, int k,
                                                 const double4* __restrict__ rec, bool flat, double re, double l0,
                                                 double reg, uint32_t seed, double* g3, double* lap) {
  const double h2 = 2 * l0;
  for (int attempt = 0; attempt < 5; ++attempt) {
    double M[45], B[9], ws = 0.0;
#pragma unroll
    for (int e = 0; e < 45; ++e) M[e] = 0.0;
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] = 0.0;
    for (int q = 0; q < k; ++q) {
      const int j = row[q];
      if (j == i) continue;
      double dx, dy, dz;
      replay_offset(seed, i, q, rec[j].w, flat, re, &dx, &dy, &dz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r <= 1e-12 || r >= re) continue;
      const double tq = re / r - 1.0;
      const double w = tq * tq;
      if (w == 0.0) continue;
      const double inv = 1.0 / r, h = h2 * r, g = l0 * r;
      const double qv[9] = {dx * inv,    dy * inv,    dz * inv,    dx * dx / h, dy * dy / h,
                            dz * dz / h, dx * dy / g, dx * dz / g, dy * dz / g};
      const double df = quad_df(dx, dy, dz);
      ws += w;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const double wq = w * qv[a];
        B[a] += wq * df / r;
#pragma unroll
        for (int c = 0; c <= a; ++c) M[tri(a, c)] += wq * qv[c];
      }
    }
    if (ws <= 0) return WLS_NO_WEIGHT;
#pragma unroll
    for (int e = 0; e < 45; ++e) M[e] /= ws;
#pragma unroll
    for (int a = 0; a < 9; ++a) B[a] /= ws;
    const double add = attempt == 0 ? 0.0 : attempt == 1 ? reg * 1.0 : attempt == 2 ? reg * 10.0
                                                            : attempt == 3 ? reg * 100.0 : reg * 1000.0;
    if (attempt > 0 && add == 0.0) return WLS_FAILED;
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

// One thread per row: FP32 moments on offsets normalised by re and an FP32
// Cholesky; flagged rows and rows with an untrusted pivot (<= 1e-3 of its
// diagonal) go to the redo list. Wall and Splash rows are skipped, as in
// the WLS stage.
// Counts per block into bpart: [0] rows solved here, [1] rows listed,
// [2] rows skipped. The redo list takes one atomic per wavefront.
__global__ __launch_bounds__(128) void k_replay_wls(const int* __restrict__ nbr, const int* __restrict__ cnt, int maxn,
                                                    const uint8_t* __restrict__ type, const uint8_t* __restrict__ bad,
                                                    const double4* __restrict__ rec, int n, double re, double l0,
                                                    uint32_t seed, double* __restrict__ grad, double* __restrict__ lap,
                                                    uint8_t* __restrict__ status, int* __restrict__ redo,
                                                    int* __restrict__ nredo, unsigned* __restrict__ bpart) {
  __shared__ unsigned lds[3 * 128 / WAVE];
  const int i = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  unsigned solved = 0, listed = 0, skipped = 0;
  if (i < n) {
    const uint8_t t = type[i];
    if (t == T_WALL || t == T_SPLASH) {
      grad[3 * (size_t)i] = grad[3 * (size_t)i + 1] = grad[3 * (size_t)i + 2] = 0.0;
      lap[i] = 0.0;
      status[i] = WLS_SKIPPED;
      skipped = 1;
    } else {
      const int* row = nbr + (size_t)i * maxn;
      const int k = cnt[i];
      const bool flat = bad[i] != 0;
      const float ire = (float)(1.0 / re), lr = (float)(l0 / re);
      float M[45], B[9], ws = 0.f;
#pragma unroll
      for (int e = 0; e < 45; ++e) M[e] = 0.f;
#pragma unroll
      for (int a = 0; a < 9; ++a) B[a] = 0.f;
      for (int q = 0; q < k; ++q) {
        const int j = row[q];
        if (j == i) continue;
        double dxd, dyd, dzd;
        replay_offset(seed, i, q, rec[j].w, flat, re, &dxd, &dyd, &dzd);
        const float dx = (float)dxd * ire, dy = (float)dyd * ire, dz = (float)dzd * ire;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        if (r <= 1e-12f || r >= 1.f) continue;
        const float tq = 1.f / r - 1.f;
        const float w = tq * tq;
        const float inv = 1.f / r, h = 2.f * lr * r, g = lr * r;
        const float qv[9] = {dx * inv,    dy * inv,    dz * inv,    dx * dx / h, dy * dy / h,
                             dz * dz / h, dx * dy / g, dx * dz / g, dy * dz / g};
        const float df = (float)quad_df(dxd, dyd, dzd) * ire;
        ws += w;
#pragma unroll
        for (int a = 0; a < 9; ++a) {
          const float wq = w * qv[a];
          B[a] = fmaf(wq, df / r, B[a]);
#pragma unroll
          for (int c = 0; c <= a; ++c) M[tri(a, c)] = fmaf(wq, qv[c], M[tri(a, c)]);
        }
      }
      bool ok = ws > 0.f;
      if (ok) {
        const float iw = 1.f / ws;
        float d0[9];
#pragma unroll
        for (int e = 0; e < 45; ++e) M[e] *= iw;
#pragma unroll
        for (int a = 0; a < 9; ++a) {
          B[a] *= iw;
          d0[a] = M[tri(a, a)];
        }
        ok = chol9<float>(M, 0.f, 1e-3f, d0) && !flat;
        if (ok) {
          float x[9];
          chol9_solve<float>(M, B, x);
          // Back to the reference's scaling: offsets were divided by re.
          grad[3 * (size_t)i] = (double)x[0];
          grad[3 * (size_t)i + 1] = (double)x[1];
          grad[3 * (size_t)i + 2] = (double)x[2];
          lap[i] = (double)(x[3] + x[4] + x[5]) / l0;
          status[i] = WLS_OK;
          solved = 1;
        }
      }
      listed = ok ? 0u : 1u;
    }
  }
  // Wavefront-aggregated append to the redo list.
  const unsigned long long mask = __ballot(listed != 0);
  if (mask) {
    const int lane = threadIdx.x & (WAVE - 1), first = __builtin_ctzll(mask);
    int at = 0;
    if (lane == first) at = atomicAdd(nredo, __popcll(mask));
    at = __shfl(at, first);
    if (listed) redo[at + __popcll(mask & ((1ull << lane) - 1ull))] = i;
  }
  unsigned v[3] = {solved, listed, skipped};
  block_count_partials<3>(v, bpart, lds);
}

__global__ __launch_bounds__(128) void k_replay_wls_redo(const int* __restrict__ nbr, const int* __restrict__ cnt,
                                                         int maxn, const uint8_t* __restrict__ bad,
                                                         const double4* __restrict__ rec, const int* __restrict__ list,
                                                         const int* __restrict__ nlist, double re, double l0,
                                                         double reg, uint32_t seed, double* __restrict__ grad,
                                                         double* __restrict__ lap, uint8_t* __restrict__ status,
                                                         unsigned long long* __restrict__ counts) {
  const int nl = *nlist;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
    const int i = list[k];
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    const uint8_t st = replay_exact_one(i, nbr + (size_t)i * maxn, cnt[i], rec, bad[i] != 0, re, l0, reg, seed, g3, &L);
    grad[3 * (size_t)i] = g3[0];
    grad[3 * (size_t)i + 1] = g3[1];
    grad[3 * (size_t)i + 2] = g3[2];
    lap[i] = L;
    status[i] = st;
    atomicAdd(&counts[st], 1ull);
  }
}

// Types of the calibration tool (0 interior, 1 near wall, 2 free surface,
// 3 splash, 4 wall/other) to ParticleType numbering; 255 for anything else.
__global__ __launch_bounds__(256) void k_replay_types(const uint8_t* __restrict__ in, int n, uint8_t* __restrict__ out,
                                                      unsigned* __restrict__ bad_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t t = in[i];
  const uint8_t map[5] = {T_INTERIOR, T_NEARWALL, T_SURFACE, T_SPLASH, T_WALL};
  if (t < 5) {
    out[i] = map[t];
  } else {
    out[i] = T_WALL;
    atomicAdd(bad_count, 1u);
  }
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_replay_load(const int64_t* row_ptr, const int* col, const uint8_t* type_in, int n, int maxn,
                        uint32_t seed, int* nbr, int* cnt, uint8_t* type_s, double4* rec, unsigned* bad_types,
                        hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_replay_ell, dim3(blocks(n, 256)), dim3(256), 0, st, row_ptr, col, n, maxn, nbr, cnt);
  hipLaunchKernelGGL(k_replay_types, dim3(blocks(n, 256)), dim3(256), 0, st, type_in, n, type_s, bad_types);
  hipLaunchKernelGGL(k_replay_nodes, dim3(blocks(n, 256)), dim3(256), 0, st, n, seed, rec);
}

void launch_replay_wls(const int* nbr, const int* cnt, int maxn, const uint8_t* type, const uint8_t* bad,
                       const double4* rec, int n, double re, double l0, double reg, uint32_t seed, double* grad,
                       double* lap, uint8_t* status, int* redo, int* nredo, unsigned long long* counts,
                       unsigned* bpart, hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
  if (n <= 0) return;
  const unsigned g = blocks(n, 128);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  hipLaunchKernelGGL(k_replay_wls, dim3(g), dim3(128), 0, st, nbr, cnt, maxn, type, bad, rec, n, re, l0, seed, grad,
                     lap, status, redo, nredo, bpart);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  launch_sum_counts(bpart, (int)g, CountSlots{3, {WLS_OK, WLS_REDO_SLOT, WLS_SKIPPED, 0, 0, 0, 0, 0}}, counts, st);
  hipLaunchKernelGGL(k_replay_wls_redo, dim3(256), dim3(128), 0, st, nbr, cnt, maxn, bad, rec, redo, nredo, re, l0, reg,
                     seed, grad, lap, status, counts);
}

}  // namespace lsmps_hip

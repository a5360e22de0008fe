// CSR pressure-Poisson (PPE) assembly on gfx950.
//
// Semantics of the reference's assemble_ppe (src/lsmps.cpp:29-64):
// unknowns are the first `cap` particles, in index order, of type Interior
// or NearWall; row 0 is the identity pin; every other row holds, in
// neighbor-list order, -a_ij = -w/(r*r) for each unknown neighbor with
// w > 0, then the diagonal sum(a_ij) last (1 with a failed-row count when
// it is not positive and finite); b = A x* summed in row order. Optional
// extension: the diagonal is scaled by (1 + shift) after the check
// (shift = 0 is the reference).
//
// Two passes (count, scan, fill), one thread per row, each row's
// diagonal summed sequentially in list order with no contraction (this file
// is built with -ffp-contract=off), so values, b and structure are
// bit-identical to the reference's.
#include "kernels.hpp"

namespace lsmps_hip {

__global__ __launch_bounds__(256) void k_unknown_flags(const uint8_t* __restrict__ type, int n, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (type[i] == T_INTERIOR || type[i] == T_NEARWALL) ? 1 : 0;
}

// Unknown numbering in particle index order; smap is the same map indexed
// by sorted slot (what the row kernels look up for each neighbor).
__global__ __launch_bounds__(256) void k_unknown_map(const int* __restrict__ flag, const int* __restrict__ pos, int n,
                                                     int cap, const int* __restrict__ iperm, int* __restrict__ smap,
                                                     int* __restrict__ rev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = pos[i];
  const bool u = flag[i] && p < cap;
  smap[iperm[i]] = u ? p : -1;
  if (u) rev[p] = i;
}

// Weight test of one pair, as the reference computes it (src/lsmps.cpp:22,45).
__device__ __forceinline__ bool pair_coef(double dx, double dy, double dz, double re, double* a) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r <= 1e-12 || r >= re) return false;
  const double t = re / r - 1.0;
  const double w = t * t;
  if (w <= 0.0) return false;
  *a = w / (r * r);
  return true;
}

// Positions (x, y, z), neighbor rows and smap are in sorted-slot order.
__global__ __launch_bounds__(256) void k_ppe_row_count(const double4* __restrict__ spos, const int* __restrict__ nbr,
                                                       const int* __restrict__ ncnt, int maxn,
                                                       const int* __restrict__ smap, const int* __restrict__ rev,
                                                       const int* __restrict__ iperm, int m, double re,
                                                       int* __restrict__ cnt) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= m) return;
  if (rr == 0) {
    cnt[0] = 1;
    return;
  }
  const int s = iperm[rev[rr]];
  const double4 pi = spos[s];
  const int* row = nbr + (size_t)s * maxn;
  const int k = ncnt[s];
  int c = 1;
  for (int q = 0; q < k; ++q) {
    const int s2 = row[q];
    if (smap[s2] < 0) continue;
    const double4 pj = spos[s2];
    double a;
    if (pair_coef(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, re, &a)) ++c;
  }
  cnt[rr] = c;
}

// Seeded manufactured solution x*_k ~ U(-1, 1): the top 53 bits of the
// splitmix64 finaliser of seed * 2^64/phi + k, scaled to [0, 1) and mapped
// to [-1, 1). Counter-based, so every row draws its value independently.
__device__ __forceinline__ double xstar_random(uint32_t seed, int k) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)k;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return -1.0 + 2.0 * ((double)(z >> 11) * 0x1.0p-53);
}

// counters: [0] failed rows, [1] NearWall rows except row 0.
__global__ __launch_bounds__(256) void k_ppe_row_fill(const double4* __restrict__ spos, const uint8_t* __restrict__ type,
                                                      const int* __restrict__ nbr, const int* __restrict__ ncnt,
                                                      int maxn, const int* __restrict__ smap,
                                                      const int* __restrict__ rev, const int* __restrict__ iperm, int m,
                                                      double re, double shift, int xmode, uint32_t xseed,
                                                      const int* __restrict__ prow,
                                                      int* __restrict__ pcol, double* __restrict__ pval,
                                                      double* __restrict__ xe, double* __restrict__ diag_out,
                                                      unsigned long long* __restrict__ counters) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  int failed = 0, neumann = 0;
  if (rr < m) {
    const int i = rev[rr];
    const int s = iperm[i];
    const double4 pi = spos[s];
    xe[rr] = xmode == 0 ? pi.x * pi.x + pi.y * pi.y  // x* = x^2 + y^2 (src/lsmps.cpp:29)
                        : xstar_random(xseed, rr);
    int at = prow[rr];
    if (rr == 0) {
      pcol[at] = 0;
      pval[at] = 1.0;
      diag_out[0] = 1.0;
    } else {
      double d = 0.0;
      const int* row = nbr + (size_t)s * maxn;
      const int kn = ncnt[s];
      for (int q = 0; q < kn; ++q) {
        const int s2 = row[q];
        const int cj = smap[s2];
        if (cj < 0) continue;
        const double4 pj = spos[s2];
        double a;
        if (!pair_coef(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, re, &a)) continue;
        d += a;
        pcol[at] = cj;
        pval[at] = -a;
        ++at;
      }
      if (d <= 0 || !isfinite(d)) {
        failed = 1;
        d = 1.0;
      }
      if (shift != 0.0) d += shift * d;
      pcol[at] = rr;
      pval[at] = d;
      diag_out[rr] = d;
      neumann = type[i] == T_NEARWALL ? 1 : 0;
    }
  }
  const unsigned long long fs = wave_sum((unsigned long long)failed), ns = wave_sum((unsigned long long)neumann);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (fs) atomicAdd(&counters[0], fs);
    if (ns) atomicAdd(&counters[1], ns);
  }
}

// b = A x* with each row summed left to right (src/lsmps.cpp:59-61).
__global__ __launch_bounds__(256) void k_ppe_rhs(const int* __restrict__ prow, const int* __restrict__ pcol,
                                                 const double* __restrict__ pval, const double* __restrict__ xe, int m,
                                                 double* __restrict__ b) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= m) return;
  double s = 0.0;
  for (int k = prow[rr]; k < prow[rr + 1]; ++k) s += pval[k] * xe[pcol[k]];
  b[rr] = s;
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_unknown_flags(const uint8_t* type, int n, int* flag, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unknown_flags, dim3(blocks(n, 256)), dim3(256), 0, st, type, n, flag);
}
void launch_unknown_map(const int* flag, const int* pos, int n, int cap, const int* iperm, int* smap, int* rev,
                        hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_unknown_map, dim3(blocks(n, 256)), dim3(256), 0, st, flag, pos, n, cap, iperm, smap, rev);
}
void launch_ppe_row_count(const double4* spos, const int* nbr, const int* ncnt, int maxn, const int* smap,
                          const int* rev, const int* iperm, int m, double re, int* cnt, hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_ppe_row_count, dim3(blocks(m, 256)), dim3(256), 0, st, spos, nbr, ncnt, maxn, smap, rev, iperm,
                       m, re, cnt);
}
void launch_ppe_row_fill(const double4* spos, const uint8_t* type, const int* nbr, const int* ncnt, int maxn,
                         const int* smap, const int* rev, const int* iperm, int m, double re, double shift, int xmode,
                         uint32_t xseed, const int* prow, int* pcol, double* pval, double* xe, double* diag,
                         unsigned long long* counters, hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_ppe_row_fill, dim3(blocks(m, 256)), dim3(256), 0, st, spos, type, nbr, ncnt, maxn, smap, rev,
                       iperm, m, re, shift, xmode, xseed, prow, pcol, pval, xe, diag, counters);
}
void launch_ppe_rhs(const int* prow, const int* pcol, const double* pval, const double* xe, int m, double* b,
                    hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_ppe_rhs, dim3(blocks(m, 256)), dim3(256), 0, st, prow, pcol, pval, xe, m, b);
}

}  // namespace lsmps_hip

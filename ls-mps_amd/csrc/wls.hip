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
//    the reference forms it. One thread per particle and the first
//    factorisation only; a row that needs the regularisation ladder goes to
//    the work list of k_wls_redo (below), which runs the whole ladder with
//    the same results.
//  * k_wls_mixed  FP32 moments on scale-free monomials (wls_accumulate), FMA
//    accumulation, FP32 Cholesky. A row whose FP32 pivots are not all above
//    tau times their diagonal entries (an ill-conditioned moment matrix,
//    whose FP32 solution carries an error of about eps32 / pivot ratio
//    times the linear part of the fit), or that would need the reference's
//    regularisation, is flagged; k_wls_pending computes the flagged rows in
//    FP64 (the reference's bits) and k_wls_redo runs the reference's
//    ladder for those whose first FP64 factorisation fails.
#include "kernels.hpp"
#include "wls_core.hpp"

namespace lsmps_hip {

// Block-aggregated status histogram (statuses 0..4 and the redo slot);
// every thread of the block must call it, with slot < 0 for "nothing".
// Block counts of each status slot into bpart[slot * gridDim.x + block].
__device__ __forceinline__ void count_status(int slot, unsigned* __restrict__ bpart) {
  __shared__ unsigned c[WLS_REDO_SLOT + 1];
  if (threadIdx.x <= WLS_REDO_SLOT) c[threadIdx.x] = 0u;
  __syncthreads();
  if (slot >= 0) atomicAdd(&c[slot], 1u);
  __syncthreads();
  if (threadIdx.x <= WLS_REDO_SLOT) bpart[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = c[threadIdx.x];
}

__device__ __forceinline__ void store_result(int i, uint8_t st, const double* g3, double lap, double* __restrict__ grad,
                                             double* __restrict__ lapo, uint8_t* __restrict__ status) {
  grad[3 * (size_t)i] = g3[0];
  grad[3 * (size_t)i + 1] = g3[1];
  grad[3 * (size_t)i + 2] = g3[2];
  lapo[i] = lap;
  status[i] = st;
}

// RN(a / b) from y = RN(1 / b) (Markstein's theorem): q0 = RN(a y) is
// within one ulp of a/b, the remainder b q0 - a is exact in one FMA, and
// RN(q0 - (b q0 - a) y) is the correctly rounded quotient; negating the
// remainder keeps the sign of a zero quotient. It holds while the remainder
// is clear of underflow and the quotient of overflow, so the caller tracks
// the binary exponents of the q0 (emin, emax; frexp gives 0 for a zero) and
// does not use the results when one leaves [2^-900, 2^900] or a sum is not
// finite (an infinite or NaN quotient).
__device__ __forceinline__ double div_rn(double a, double b, double y, int& emin, int& emax) {
  const double q0 = a * y;
  const double e = fma(b, q0, -a);
  const int ex = __builtin_amdgcn_frexp_exp(q0);  // 0 for a zero quotient
  emin = min(emin, ex);
  emax = max(emax, ex);
  return fma(-e, y, q0);
}

// FP64 WLS of one particle per lane, in the reference's operation order:
// the moments and right-hand sides are accumulated neighbor by neighbor
// (each neighbor's record is loaded one neighbor ahead and its index two
// ahead, so the loads overlap the previous neighbor's arithmetic), then the
// first Cholesky (no diagonal shift) and the solve. A row whose first
// factorisation fails goes to the work list of k_wls_redo, which runs the
// reference's whole ladder with the same results bit for bit; keeping the
// ladder out of this kernel takes its register count from 248 to 200.
// The reference's divisions by r, 2 l0 r, l0 r and the weight sum are
// correctly rounded quotients from one reciprocal each (div_rn): 14 of the
// 17 divisions per neighbor pair and the 54 per particle become three
// instructions each, 1.70 -> 1.41 ms at 3M (industrial_mix), the same
// results bit for bit; a row with a quotient div_rn cannot vouch for goes
// to k_wls_redo, which divides. Tracking the quotients' exponents (frexp,
// min, max) instead of testing each against the range: 207 VGPRs and 4%
// less time. Forcing three wavefronts per SIMD (168 VGPRs) measured 1.75 ms
// against 1.67, four (spills) 7.2 ms.
// One row of the FP64 WLS (the body of k_wls_exact): WLS_OK with the
// gradient and Laplacian, WLS_NO_WEIGHT, or -1 when the row needs the
// ladder or plain divisions (k_wls_redo).
__device__ __forceinline__ int wls_exact_row(const double4* __restrict__ spos, const int* __restrict__ nbr,
                                             const int* __restrict__ cnt, int maxn, int i, double re, double l0,
                                             double (&g3)[3], double& L) {
  const double4 pi = spos[i];
  const int* col = nbr + (size_t)i * maxn;
  const int kn = cnt[i];
  const double h2 = 2 * l0;
  double M[45], B[9], ws = 0.0;
  int emin = 0, emax = 0;  // exponent range of the div_rn quotients
#pragma unroll
  for (int k = 0; k < 45; ++k) M[k] = 0.0;
#pragma unroll
  for (int a = 0; a < 9; ++a) B[a] = 0.0;
  // Prefetches past the row's end reload its last entry (no branches).
  double4 pn = kn > 0 ? spos[col[0]] : pi;
  int jn = col[kn > 1 ? 1 : 0];
  for (int k = 0; k < kn; ++k) {
    const double4 pj = pn;
    pn = spos[jn];
    jn = col[k + 2 < kn ? k + 2 : kn - 1];
    const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r <= 1e-12 || r >= re) continue;
    const double inv = 1.0 / r, h = h2 * r, g = l0 * r;
    const double tq = div_rn(re, r, inv, emin, emax) - 1.0;  // re / r >= 1: in range
    const double w = tq * tq;
    if (w == 0.0) continue;
    // h = (2 l0) r is exactly 2 g, so RN(1/h) = RN(1/g) / 2 while g and
    // 1/g are normal numbers (g's exponent joins the range check).
    const double yg = 1.0 / g, yh = 0.5 * yg;
    const int eg = __builtin_amdgcn_frexp_exp(g);
    emin = min(emin, eg);
    emax = max(emax, eg);
    double q[9];
    q[0] = dx * inv;
    q[1] = dy * inv;
    q[2] = dz * inv;
    q[3] = div_rn(dx * dx, h, yh, emin, emax);
    q[4] = div_rn(dy * dy, h, yh, emin, emax);
    q[5] = div_rn(dz * dz, h, yh, emin, emax);
    q[6] = div_rn(dx * dy, g, yg, emin, emax);
    q[7] = div_rn(dx * dz, g, yg, emin, emax);
    q[8] = div_rn(dy * dz, g, yg, emin, emax);
    const double df = pj.w - pi.w;
    ws += w;
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      const double wq = w * q[a];
      B[a] += div_rn(wq * df, r, inv, emin, emax);
#pragma unroll
      for (int c = 0; c <= a; ++c) M[tri(a, c)] += wq * q[c];
    }
  }
  if (ws <= 0) return WLS_NO_WEIGHT;
  const double yw = 1.0 / ws;
#pragma unroll
  for (int k = 0; k < 45; ++k) M[k] = div_rn(M[k], ws, yw, emin, emax);
#pragma unroll
  for (int a = 0; a < 9; ++a) B[a] = div_rn(B[a], ws, yw, emin, emax);
  const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // Quotients within [2^-900, 2^900] (a zero has exponent 0) and finite
  // sums (an overflow or a NaN input): else k_wls_redo divides.
  bool fin = isfinite(ws);
#pragma unroll
  for (int k = 0; k < 45; ++k) fin = fin && isfinite(M[k]);
#pragma unroll
  for (int a = 0; a < 9; ++a) fin = fin && isfinite(B[a]);
  const bool off = emin < -899 || emax > 900 || !fin;
  if (off || !chol9<double>(M, 1e-14, 0.0, zero9)) return -1;
  double sol[9];
  chol9_solve<double>(M, B, sol);
  g3[0] = sol[0];
  g3[1] = sol[1];
  g3[2] = sol[2];
  L = (sol[3] + sol[4] + sol[5]) / l0;
  return WLS_OK;
}

__global__ __launch_bounds__(WAVE) void k_wls_exact(const double4* __restrict__ spos,
                                                    const uint8_t* __restrict__ type, const int* __restrict__ nbr,
                                                    const int* __restrict__ cnt, int maxn, int n, double re, double l0,
                                                    double* __restrict__ grad, double* __restrict__ lap,
                                                    uint8_t* __restrict__ status, int* __restrict__ redo,
                                                    int* __restrict__ nredo, unsigned* __restrict__ bpart,
                                                    const int* __restrict__ out_perm) {
  const int i = (int)xcd_block(blockIdx.x, gridDim.x) * WAVE + threadIdx.x;
  int slot = -1;
  if (i < n) {
    const uint8_t t = type[i];
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    const int st = t != T_WALL && t != T_SPLASH ? wls_exact_row(spos, nbr, cnt, maxn, i, re, l0, g3, L) : WLS_SKIPPED;
    if (st < 0) {
      redo[atomicAdd(nredo, 1)] = i;
    } else {
      store_result(out_perm ? out_perm[i] : i, (uint8_t)st, g3, L, grad, lap, status);
      slot = st;
    }
  }
  count_status(slot, bpart);
}

// The rows k_wls_mixed left to FP64 (pend[s] = 1) as one list in
// ascending order: block b of the mixed grid (rows 64 b .. 64 b + 63) writes
// its flagged rows from offset poff[b], the exclusive scan of its count.
__global__ __launch_bounds__(WAVE) void k_wls_pend_list(const uint8_t* __restrict__ pend, int n,
                                                        const int* __restrict__ poff, int* __restrict__ list) {
  const int s = blockIdx.x * WAVE + threadIdx.x;
  const bool p = s < n && pend[s] != 0;
  const uint64_t mk = __builtin_amdgcn_ballot_w64(p);
  if (p) list[poff[blockIdx.x] + __popcll(mk & ((1ull << threadIdx.x) - 1ull))] = s;
}

// The listed rows (count list_len[0]) through the same row function as
// k_wls_exact, so they are the reference's bits; one row per lane, every
// wavefront a full batch of 64 (the FP64 row is a latency chain, so the
// time is the number of wavefronts over the ones resident at once). Rows
// that need the ladder go to the k_wls_redo list. Per block: rows finished
// here with WLS_OK, with WLS_NO_WEIGHT, and all of them.
__global__ __launch_bounds__(WAVE) void k_wls_pending(const double4* __restrict__ spos, const int* __restrict__ list,
                                                      const int* __restrict__ list_len, const int* __restrict__ nbr,
                                                      const int* __restrict__ cnt, int maxn, double re, double l0,
                                                      double* __restrict__ grad, double* __restrict__ lap,
                                                      uint8_t* __restrict__ status, int* __restrict__ redo,
                                                      int* __restrict__ nredo, unsigned* __restrict__ bpart,
                                                      const int* __restrict__ out_perm) {
  const int k = blockIdx.x * WAVE + threadIdx.x;
  unsigned n_ok = 0, n_nw = 0, n_done = 0;
  if (k < *list_len) {
    const int i = list[k];
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    const int st = wls_exact_row(spos, nbr, cnt, maxn, i, re, l0, g3, L);
    if (st < 0) {
      redo[atomicAdd(nredo, 1)] = i;
    } else {
      store_result(out_perm ? out_perm[i] : i, (uint8_t)st, g3, L, grad, lap, status);
      n_ok = st == WLS_OK ? 1u : 0u;
      n_nw = st == WLS_NO_WEIGHT ? 1u : 0u;
      n_done = 1;
    }
  }
  unsigned v[3] = {n_ok, n_nw, n_done};
  block_count_partials<3>(v, bpart, nullptr);
}

// The FP64 redo of the rows the FP32 kernel could not trust, eight lanes
// per row, with the same results as the reference's sequential ladder bit for bit:
//  - the per-neighbor values (weight, the nine scaled monomials, the field
//    difference and r) are computed by the row's lanes in parallel, a chunk
//    of RX_CHUNK neighbors at a time, into LDS, with the same expressions;
//  - each of the 55 sums (45 moments, 9 right-hand sides, the weight sum) is
//    owned by one lane, which adds the chunk's terms in neighbor order, so
//    every sum sees the same additions in the same order;
//  - the moments do not depend on the regularisation step, so the five
//    steps of the ladder are factorised side by side by lanes 0..4 and the
//    first that succeeds is taken.
// A row costs about a quarter of the sequential latency; the list is short
// (rows whose first FP64 factorisation failed), so that latency is the time.
constexpr int RX_LANES = 8, RX_ROWS = WAVE / RX_LANES, RX_CHUNK = 32, RX_REC = 12, RX_OWN = 7;

__global__ __launch_bounds__(WAVE) void k_wls_redo(const double4* __restrict__ spos, const int* __restrict__ nbr,
                                                   const int* __restrict__ cnt, int maxn, const int* __restrict__ list,
                                                   const int* __restrict__ nlist, double re, double l0, double reg,
                                                   double* __restrict__ grad, double* __restrict__ lap,
                                                   uint8_t* __restrict__ status,
                                                   unsigned long long* __restrict__ counts,
                                                   const int* __restrict__ out_perm) {
  __shared__ double rec[RX_ROWS][RX_CHUNK][RX_REC];  // w, q[0..8], df, r
  __shared__ int okn[RX_ROWS][RX_CHUNK];
  __shared__ double acc[RX_ROWS][56];
  const int lane = threadIdx.x, g = lane % RX_LANES, sub = lane / RX_LANES;
  const double h2 = 2 * l0;
  // Sums owned by this lane: j = g + 8 t (t < 7, j < 55); for a moment,
  // its monomial pair (a, c) with j = tri(a, c).
  int own_a[RX_OWN], own_c[RX_OWN];
#pragma unroll
  for (int t = 0; t < RX_OWN; ++t) {
    const int j = g + RX_LANES * t;
    int a = 0;
    while (a < 8 && tri(a + 1, 0) <= j) ++a;
    own_a[t] = j < 45 ? a : j < 54 ? j - 45 : -1;
    own_c[t] = j < 45 ? j - tri(a, 0) : -1;
  }
  const int nl = *nlist;
  for (int k0 = blockIdx.x * RX_ROWS; k0 < nl; k0 += gridDim.x * RX_ROWS) {
    const int k = k0 + sub;
    const bool live = k < nl;
    const int i = live ? list[k] : 0;
    const int kn = live ? cnt[i] : 0;
    const double4 pi = spos[i];
    const int* col = nbr + (size_t)i * maxn;
    double own[RX_OWN];
#pragma unroll
    for (int t = 0; t < RX_OWN; ++t) own[t] = 0.0;
    const int knmax = wave_max(kn);
    for (int c0 = 0; c0 < knmax; c0 += RX_CHUNK) {
      for (int u = g; u < RX_CHUNK; u += RX_LANES) {
        const int kk = c0 + u;
        int ok = 0;
        if (kk < kn) {
          const double4 pj = spos[col[kk]];
          const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
          const double r = sqrt(dx * dx + dy * dy + dz * dz);
          if (!(r <= 1e-12 || r >= re)) {
            const double tq = re / r - 1.0;
            const double w = tq * tq;
            if (w != 0.0) {
              const double inv = 1.0 / r, h = h2 * r, gg = l0 * r;
              double* o = rec[sub][u];
              o[0] = w;
              o[1] = dx * inv;
              o[2] = dy * inv;
              o[3] = dz * inv;
              o[4] = dx * dx / h;
              o[5] = dy * dy / h;
              o[6] = dz * dz / h;
              o[7] = dx * dy / gg;
              o[8] = dx * dz / gg;
              o[9] = dy * dz / gg;
              o[10] = pj.w - pi.w;
              o[11] = r;
              ok = 1;
            }
          }
        }
        okn[sub][u] = ok;
      }
      wave_sync();
      const int cend = kn - c0 < RX_CHUNK ? kn - c0 : RX_CHUNK;
      for (int u = 0; u < cend; ++u) {
        if (!okn[sub][u]) continue;
        const double* o = rec[sub][u];
        const double w = o[0];
#pragma unroll
        for (int t = 0; t < RX_OWN; ++t) {
          const int a = own_a[t], c = own_c[t];
          if (c >= 0) own[t] += (w * o[1 + a]) * o[1 + c];
          else if (a >= 0) own[t] += (w * o[1 + a]) * o[10] / o[11];
          else if (g + RX_LANES * t == 54) own[t] += w;
        }
      }
      wave_sync();
    }
#pragma unroll
    for (int t = 0; t < RX_OWN; ++t)
      if (g + RX_LANES * t < 55) acc[sub][g + RX_LANES * t] = own[t];
    wave_sync();
    // Lanes 0..4 of the row try the five steps of the ladder.
    uint8_t st = WLS_FAILED;
    double g3[3] = {0.0, 0.0, 0.0}, L = 0.0;
    const double ws = acc[sub][54];
    bool solved = false;
    if (ws <= 0) {
      st = WLS_NO_WEIGHT;
    } else if (g < 5) {
      const double add = g == 0 ? 0.0 : g == 1 ? reg * 1.0 : g == 2 ? reg * 10.0 : g == 3 ? reg * 100.0 : reg * 1000.0;
      if (g == 0 || add != 0.0) {
        double M[45], B[9];
#pragma unroll
        for (int j = 0; j < 45; ++j) M[j] = acc[sub][j] / ws;
#pragma unroll
        for (int a = 0; a < 9; ++a) B[a] = acc[sub][45 + a] / ws;
        if (add != 0.0) {
#pragma unroll
          for (int a = 0; a < 9; ++a) M[tri(a, a)] += add;
        }
        const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (chol9<double>(M, 1e-14, 0.0, zero9)) {
          double sol[9];
          chol9_solve<double>(M, B, sol);
          g3[0] = sol[0];
          g3[1] = sol[1];
          g3[2] = sol[2];
          L = (sol[3] + sol[4] + sol[5]) / l0;
          st = g == 0 ? WLS_OK : WLS_REGULARIZED;
          solved = true;
        }
      }
    }
    // The row's result: the first step that succeeded, else its lane 0.
    const unsigned mk = (unsigned)(__builtin_amdgcn_ballot_w64(live && solved) >> (sub * RX_LANES)) & 0xFFu;
    const int winner = mk ? __builtin_ctz(mk) : 0;
    if (live && g == winner) {
      store_result(out_perm ? out_perm[i] : i, st, g3, L, grad, lap, status);
      atomicAdd(&counts[st], 1ull);
      atomicAdd(&counts[WLS_REDO_SLOT], 1ull);
    }
    wave_sync();  // acc and rec are reused by the next rows
  }
}

// FP32 moments on offsets scaled by 1/re (so re -> 1, l0 -> l0/re); the
// solution vector is unchanged by the scaling because the monomials are
// dimensionless and df/r = (df/re)/(r/re).
//
// One wavefront per 64 consecutive sorted particles, in two halves of 32:
//  phase 1: 2 lanes per particle (32 particles at a time) stride through the
//           particle's neighbor row, four records in flight per lane; each
//           lane accumulates the 55 sums (45 lower-triangle moments, 9
//           right-hand sides, weight sum), the pair combines them with one
//           DPP step and writes them to LDS. (8 lanes per particle spent 165
//           DPP adds per particle on the combine and ran 0.81 ms at 3M;
//           4 lanes 0.62 ms, 2 lanes 0.58 ms, measured on MI355X.)
//  phase 2: one lane per particle factorises its 9x9 from LDS (FP32
//           Cholesky with reciprocal square roots) and solves. Rows with a
//           pivot at or below tau times its diagonal entry are flagged for
//           FP64 (pend[s] = 1; every row's flag is written).
constexpr int WLS_GROUP = 2;   // lanes per particle in phase 1
constexpr int WLS_NACC = 55;   // 45 moments + 9 RHS + weight sum
constexpr int WLS_STRIDE = 57; // LDS floats per particle (odd: no bank conflicts)
constexpr int WLS_HALF = 32;   // particles per phase-2 batch

// Offsets in FP32 from the FP64 differences, in the caller's length unit:
// the linear monomials d/r are scale-free, the quadratic ones d (d/r) lack
// the reference's constant factors (1/(2 l0) on the squares, 1/l0 on the
// products), so the moments are S^-1 M S^-1 for a constant diagonal S. The
// solution is then S times the reference's, undone on the Laplacian (whose
// l0 cancels); the pivot test is relative to each diagonal entry, which S
// leaves unchanged. (Normalising the offsets by re took four FP64 products
// per pair, 3-4% of the kernel at the same accuracy.) The sums stay in the
// FP32 range for re in [1e-12, 1e12]; outside it the caller flags every row
// for FP64 (tau = 1).
__device__ __forceinline__ void wls_accumulate(const double4 pj, const double4 pi, float re, float rmin2, float re2,
                                               float (&a)[WLS_NACC]) {
  const float dx = (float)(pj.x - pi.x), dy = (float)(pj.y - pi.y), dz = (float)(pj.z - pi.z);
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  // Out of range (or the row's own slot standing in for a missing
  // neighbor): weight 0 and finite monomials, so every sum adds 0.
  const bool ok = r2 > rmin2 && r2 < re2;
  const float inv = __builtin_amdgcn_rsqf(ok ? r2 : re2);
  const float tq = re * inv - 1.0f;
  const float w = ok ? tq * tq : 0.0f;
  float q[9];
  q[0] = dx * inv;
  q[1] = dy * inv;
  q[2] = dz * inv;
  q[3] = q[0] * dx;
  q[4] = q[1] * dy;
  q[5] = q[2] * dz;
  q[6] = q[0] * dy;
  q[7] = q[0] * dz;
  q[8] = q[1] * dz;
  const float dfr = (float)(pj.w - pi.w) * inv;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float wq = w * q[i];
    a[45 + i] = fmaf(wq, dfr, a[45 + i]);
#pragma unroll
    for (int c = 0; c <= i; ++c) a[tri(i, c)] = fmaf(wq, q[c], a[tri(i, c)]);
  }
  a[54] += w;
}

// FP32 Cholesky + solve from the 55 sums; false when a pivot is not
// > tau times its diagonal (or not finite).
__device__ __forceinline__ bool wls_solve_f32(const float* __restrict__ a, float tau, float (&sol)[9]) {
  const float ws = a[54];
  if (!(ws > 0.0f)) return false;
  const float iw = 1.0f / ws;
  float m[45], bb[9], il[9];
#pragma unroll
  for (int k = 0; k < 45; ++k) m[k] = a[k] * iw;
#pragma unroll
  for (int i = 0; i < 9; ++i) bb[i] = a[45 + i] * iw;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float d0 = m[tri(i, i)];
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      float sm = m[tri(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) sm = fmaf(-m[tri(i, k)], m[tri(j, k)], sm);
      if (i == j) {
        if (!(sm > tau * d0) || !isfinite(sm)) ok = false;
        const float r = __builtin_amdgcn_rsqf(sm);
        il[i] = r;
        m[tri(i, i)] = sm * r;
      } else {
        m[tri(i, j)] = sm * il[j];
      }
    }
  }
  if (!ok) return false;
  float yv[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float sm = bb[i];
#pragma unroll
    for (int k = 0; k < i; ++k) sm = fmaf(-m[tri(i, k)], yv[k], sm);
    yv[i] = sm * il[i];
  }
#pragma unroll
  for (int i = 8; i >= 0; --i) {
    float sm = yv[i];
#pragma unroll
    for (int k = i + 1; k < 9; ++k) sm = fmaf(-m[tri(k, i)], sol[k], sm);
    sol[i] = sm * il[i];
  }
  return true;
}

__global__ __launch_bounds__(WAVE) void k_wls_mixed(const double4* __restrict__ spos,
                                                    const uint8_t* __restrict__ type, const int* __restrict__ nbr,
                                                    const int* __restrict__ cnt, int maxn, int n, double re, float tau,
                                                    double* __restrict__ grad, double* __restrict__ lap,
                                                    uint8_t* __restrict__ status, uint8_t* __restrict__ pend,
                                                    unsigned* __restrict__ bpart, const int* __restrict__ out_perm) {
  __shared__ float sums[WLS_HALF * WLS_STRIDE];
  const int lane = threadIdx.x;
  const float re_f = (float)re, re2 = (float)(re * re);
  const float rmin2 = 1e-24f;  // the reference's r > 1e-12
  const int s0 = (int)xcd_block(blockIdx.x, gridDim.x) * WAVE;
  int n_ok = 0, n_skip = 0, n_pend = 0;
  for (int half = 0; half < WAVE / WLS_HALF; ++half) {
    // phase 1: WLS_GROUP lanes per particle, WAVE / WLS_GROUP particles per
    // pass, WLS_HALF / (WAVE / WLS_GROUP) passes (one with two lanes).
    const int g = lane % WLS_GROUP, sub = lane / WLS_GROUP;
    for (int pass = 0; pass < WLS_HALF / (WAVE / WLS_GROUP); ++pass) {
      const int local = pass * (WAVE / WLS_GROUP) + sub;  // 0..31 within the half
      const int s = s0 + half * WLS_HALF + local;
      float a[WLS_NACC];
#pragma unroll
      for (int k = 0; k < WLS_NACC; ++k) a[k] = 0.0f;
      if (s < n) {
        const uint8_t t = type[s];
        if (t != T_WALL && t != T_SPLASH) {
          const double4 pi = spos[s];
          const int* row = nbr + (size_t)s * maxn;
          const int kn = cnt[s];
          // Four neighbors per lane in flight: all four indices, then all
          // four records, then the four updates (the loads of one batch
          // do not wait on each other). Past the end of the row a lane
          // loads its own record, whose weight is 0: no branches around the
          // loads or the updates (0.58 -> 0.49 ms at 3M, the same sums).
          for (int k0 = g; k0 < kn; k0 += 4 * WLS_GROUP) {
            int j[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int q = k0 + u * WLS_GROUP;
              const int jq = row[q < kn ? q : kn - 1];
              j[u] = q < kn ? jq : s;
            }
            double4 pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pj[u] = spos[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) wls_accumulate(pj[u], pi, re_f, rmin2, re2, a);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < WLS_NACC; ++k) a[k] = group_sum<WLS_GROUP>(a[k]);
      // Lane g of the group stores its share of the sums.
      constexpr int share = (WLS_NACC + WLS_GROUP - 1) / WLS_GROUP;
#pragma unroll
      for (int k = 0; k < WLS_NACC; ++k)
        if (k / share == g) sums[local * WLS_STRIDE + k] = a[k];
    }
    __syncthreads();
    // phase 2: one lane per particle of this half.
    if (lane < WLS_HALF) {
      const int s = s0 + half * WLS_HALF + lane;
      if (s < n) {
        const uint8_t t = type[s];
        bool fp64 = false;
        if (t == T_WALL || t == T_SPLASH) {
          const double g0[3] = {0.0, 0.0, 0.0};
          store_result(out_perm ? out_perm[s] : s, WLS_SKIPPED, g0, 0.0, grad, lap, status);
          ++n_skip;
        } else {
          float sol[9];
          if (wls_solve_f32(sums + lane * WLS_STRIDE, tau, sol)) {
            // The reference's quadratic coefficients are the ones of the
            // unscaled monomials times 2 l0, and its Laplacian is their sum
            // over l0.
            const double g3[3] = {(double)sol[0], (double)sol[1], (double)sol[2]};
            const float lp = (sol[3] + sol[4] + sol[5]) * 2.0f;
            store_result(out_perm ? out_perm[s] : s, WLS_OK, g3, (double)lp, grad, lap, status);
            ++n_ok;
          } else {
            fp64 = true;
          }
        }
        pend[s] = fp64 ? 1 : 0;
        n_pend += fp64 ? 1 : 0;
      }
    }
    __syncthreads();
  }
  // Per-block status counts, and the count of rows flagged for FP64 (their
  // statuses are counted where they are computed), stored at the logical
  // block (rows s0 .. s0 + 63), which the list's offsets follow.
  unsigned v[3] = {(unsigned)n_ok, (unsigned)n_skip, (unsigned)n_pend};
  block_count_partials<3>(v, bpart, nullptr, (unsigned)s0 / WAVE);
}

// Quadratic field f = c0 + cx x + cy y + cz z + cxx x^2 + cyy y^2 + czz z^2
// + cxy x y + cxz x z + cyz y z, summed left to right with (c*u)*v
// products, which reproduces the reference's test-field expression bit for
// bit (src/lsmps.cpp:28).
__global__ __launch_bounds__(256) void k_field_quadratic(double4* __restrict__ spos, int n, QuadField q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  spos[i].w = quad_eval(q, spos[i].x, spos[i].y, spos[i].z);
}

// Errors of the sampled particles i = k*stride (original index) against the
// closed-form gradient and Laplacian of the quadratic field
// (src/lsmps.cpp:28); positions and results are read at the sorted slot.
__global__ __launch_bounds__(256) void k_wls_sample_errors(const double4* __restrict__ spos,
                                                           const uint8_t* __restrict__ type,
                                                           const double* __restrict__ grad,
                                                           const double* __restrict__ lap,
                                                           const uint8_t* __restrict__ status,
                                                           const int* __restrict__ iperm, int64_t stride, int ns,
                                                           QuadField q, double* __restrict__ ge,
                                                           double* __restrict__ le, uint8_t* __restrict__ st_out,
                                                           uint8_t* __restrict__ ty_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const int64_t io = (int64_t)k * stride;  // original index; all arrays here are in sorted order
  const int64_t i = iperm[io];
  const double a = spos[i].x, b = spos[i].y, c = spos[i].z;
  const double ex = q.c[1] + 2 * q.c[4] * a + q.c[7] * b + q.c[8] * c;
  const double ey = q.c[2] + 2 * q.c[5] * b + q.c[7] * a + q.c[9] * c;
  const double ez = q.c[3] + 2 * q.c[6] * c + q.c[8] * a + q.c[9] * b;
  const double d0 = grad[3 * i] - ex, d1 = grad[3 * i + 1] - ey, d2 = grad[3 * i + 2] - ez;
  ge[k] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  le[k] = fabs(lap[i] - q.lap_exact);
  st_out[k] = status[i];
  ty_out[k] = type[i];
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

// Field values given in particle order into the sorted records.
__global__ __launch_bounds__(256) void k_gather_field(const int* __restrict__ perm, const double* __restrict__ v, int n,
                                                      double4* __restrict__ spos) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) spos[s].w = v[perm[s]];
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_wls_unsort(const int* perm, const double* gs, const double* ls, const uint8_t* ss, int n, double* g,
                       double* l, uint8_t* st, hipStream_t stream) {
  if (n > 0) hipLaunchKernelGGL(k_wls_unsort, dim3(blocks(n, 256)), dim3(256), 0, stream, perm, gs, ls, ss, n, g, l, st);
}

void launch_gather_field(const int* perm, const double* v, int n, double4* spos, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_gather_field, dim3(blocks(n, 256)), dim3(256), 0, st, perm, v, n, spos);
}

void launch_field_quadratic(double4* spos, int n, const QuadField& q, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_field_quadratic, dim3(blocks(n, 256)), dim3(256), 0, st, spos, n, q);
}

void launch_wls_exact(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, double* grad, double* lap, uint8_t* status, int* redo,
                      int* nredo, unsigned long long* counts, unsigned* bpart, hipStream_t st, hipEvent_t ev_begin,
                      hipEvent_t ev_end, const int* out_perm) {
  if (n <= 0) return;
  const unsigned g = blocks(n, WAVE);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  hipLaunchKernelGGL(k_wls_exact, dim3(g), dim3(WAVE), 0, st, spos, type, nbr, cnt, maxn, n, re, l0, grad, lap, status,
                     redo, nredo, bpart, out_perm);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  CountSlots cs{WLS_REDO_SLOT, {0, 1, 2, 3, 4, 0, 0, 0}};
  launch_sum_counts(bpart, (int)g, cs, counts, st);
  // Rows whose first factorisation failed: the reference's ladder.
  const unsigned gr = blocks(n, RX_ROWS) < 1024u ? blocks(n, RX_ROWS) : 1024u;
  hipLaunchKernelGGL(k_wls_redo, dim3(gr), dim3(WAVE), 0, st, spos, nbr, cnt, maxn, redo, nredo, re, l0, reg, grad, lap,
                     status, counts, out_perm);
}

void launch_wls_mixed(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, float tau, const WlsMixedWork& w, double* grad, double* lap,
                      uint8_t* status, int* redo, int* nredo, unsigned long long* counts, unsigned* bpart,
                      hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end, const int* out_perm) {
  if (n <= 0) return;
  const unsigned gm = blocks(n, WAVE);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  hipLaunchKernelGGL(k_wls_mixed, dim3(gm), dim3(WAVE), 0, st, spos, type, nbr, cnt, maxn, n, re, tau, grad, lap,
                     status, w.pend, bpart, out_perm);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  // The flagged rows as one ascending list: offsets per mixed block, then
  // the list (its length lands in poff[gm]).
  launch_exclusive_scan(reinterpret_cast<const int*>(bpart + 2 * (size_t)gm), gm, w.poff, w.scan_scratch, st);
  hipLaunchKernelGGL(k_wls_pend_list, dim3(gm), dim3(WAVE), 0, st, w.pend, n, w.poff, w.list);
  // One row per lane over the list; blocks past its length only report
  // zero counts. Its three partial rows follow the mixed kernel's three,
  // and one launch adds all six (the flagged-row count into an unused slot).
  hipLaunchKernelGGL(k_wls_pending, dim3(gm), dim3(WAVE), 0, st, spos, w.list, w.poff + gm, nbr, cnt, maxn, re, l0,
                     grad, lap, status, redo, nredo, bpart + 3 * (size_t)gm, out_perm);
  CountSlots cs{6, {WLS_OK, WLS_SKIPPED, WLS_REDO_SLOT + 1, WLS_OK, WLS_NO_WEIGHT, WLS_REDO_SLOT, 0, 0}};
  launch_sum_counts(bpart, (int)gm, cs, counts, st);
  // Rows whose first FP64 factorisation failed: the reference's ladder.
  // The list length is only known on the device: a bounded grid strides
  // over it (typical lists are empty or a few hundred entries).
  const unsigned gr = blocks(n, RX_ROWS) < 1024u ? blocks(n, RX_ROWS) : 1024u;
  hipLaunchKernelGGL(k_wls_redo, dim3(gr), dim3(WAVE), 0, st, spos, nbr, cnt, maxn, redo, nredo, re, l0, reg, grad, lap,
                     status, counts, out_perm);
}

void launch_wls_sample_errors(const double4* spos, const uint8_t* type, const double* grad, const double* lap,
                              const uint8_t* status, const int* iperm, int64_t stride, int ns, const QuadField& q,
                              double* ge, double* le, uint8_t* st_out, uint8_t* ty_out, hipStream_t st) {
  if (ns > 0)
    hipLaunchKernelGGL(k_wls_sample_errors, dim3(blocks(ns, 256)), dim3(256), 0, st, spos, type, grad, lap, status,
                       iperm, stride, ns, q, ge, le, st_out, ty_out);
}

}  // namespace lsmps_hip

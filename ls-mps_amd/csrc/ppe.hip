// Pressure-Poisson (PPE) assembly on gfx950.
//
// Semantics of the reference's assemble_ppe (src/lsmps.cpp:29-64):
// unknowns are the first `cap` particles, in index order, of type Interior
// or NearWall (row rr = their rank); row 0 is the identity pin; every other
// row holds, in neighbor-list order, -a_ij = -w/(r*r) for each unknown
// neighbor with w > 0, then the diagonal sum(a_ij) last (1 with a
// failed-row count when it is not positive and finite); b = A x* summed
// left to right in that order. Optional extension: the diagonal is scaled
// by (1 + shift) after the check (shift = 0 is the reference).
//
// Layout. The matrix is written once, straight into the sliced-ELL form
// the solver streams: rows are the unknowns in cell-sorted order (row u =
// rank of the unknown among sorted slots), slices of 64 rows stored
// column-major and padded with (u, 0). Neighbor lists in sorted slots are
// ascending (the reference visits cells z, y, x-major and members by
// index, which is exactly the slot order), so each row's columns come out
// ascending with no sort; the diagonal is placed at its rank among them.
// The reference-order CSR (rows rr, diagonal last) is only materialised
// when a caller downloads it.
//
// Kernels: unknown flags over particle index, a scan (reference numbering
// rr), flags over sorted slots, a scan (sorted numbering u), the maps
// between them and x* per unknown slot; then a count pass (unknown
// neighbors per row -> slice widths) and the fill pass. Count and fill run
// one 512-thread block per slice, eight lanes per row: the lanes walk the
// row's ELL neighbor list eight entries at a time, and ordered ballots give
// each kept entry its position. The diagonal and b are sums in list order
// (each batch's terms staged in LDS and added by one lane per sum), so they
// are bit-identical to the reference (this file is built with
// -ffp-contract=off).
//
// WLS-Laplacian rows (k_ppe_wls_fill). The reference's C++ has only the
// graph Laplacian; its stated next step is PPE rows from the WLS operator
// (README.md:34), which its 2-D model builds (python/rain_plate_lsmps_cpu.py:
// 204-253). In 3-D with the reference's weight, basis and Cholesky ladder
// (src/lsmps.cpp:22-25,28): the WLS Laplacian of a field f at particle i is
// (x3 + x4 + x5) / l0 with x = M^-1 sum_j w_j q_j (f_j - f_i) / (r_j ws), so
// it is sum_j c_ij (f_j - f_i) with c_ij = (z . q_j) w_j / r_j / ws / l0 and
// z = M^-1 (0,0,0,1,1,1,0,0,0). Row i is minus that operator, the sign of
// the reference's graph Laplacian: -c_ij for each unknown neighbor (list
// order) and the diagonal d = sum_j c_ij over every neighbor with w > 0;
// neighbors that are not unknowns (FreeSurface, Splash, Wall) are p = 0
// (Dirichlet).
//
// A one-sided neighborhood (a surface particle the sector rule typed
// Interior, a plate edge) gives a stencil that is no Laplacian: d <= 0 or
// off-diagonals that outweigh it, which makes the matrix indefinite (on a
// wall film, Jacobi-scaled eigenvalues down to -32; BiCGStab and GMRES
// both fail). Such a row -- WLS failed, d not finite, d <= 0 or
// d < sum |c_ij| over the unknown neighbors -- takes the
// reference's graph-Laplacian row instead (-w/r^2 per unknown neighbor,
// their sum on the diagonal, 1 when that is not positive). One thread per
// row, FP64, no contraction: bit-identical to the oracle's restatement.
#include "kernels.hpp"
#include "wls_core.hpp"

namespace lsmps_hip {

constexpr int PPE_SLICE = 64;

__device__ __forceinline__ bool is_unknown_type(uint8_t t) { return t == T_INTERIOR || t == T_NEARWALL; }

// Seeded manufactured solution x*_k ~ U(-1, 1): the top 53 bits of the
// splitmix64 finaliser of seed * 2^64/phi + k, scaled to [0, 1) and mapped
// to [-1, 1). Counter-based on the reference row k, so every row draws its
// value independently of the layout.
__device__ __forceinline__ double xstar_random(uint32_t seed, int k) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + (uint64_t)k;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return -1.0 + 2.0 * ((double)(z >> 11) * 0x1.0p-53);
}

// flag[i] over particle index i, from the sorted-order types.
__global__ __launch_bounds__(256) void k_unknown_flags(const uint8_t* __restrict__ type_s,
                                                       const int* __restrict__ iperm, int n, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = is_unknown_type(type_s[iperm[i]]) ? 1 : 0;
}

// fs[s]: slot s holds an unknown (flagged and among the first cap).
__global__ __launch_bounds__(256) void k_unknown_slots(const int* __restrict__ perm, const int* __restrict__ flag,
                                                       const int* __restrict__ pos, int n, int cap,
                                                       int* __restrict__ fs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  fs[s] = flag[i] && pos[i] < cap ? 1 : 0;
}

// uid[s] = row u of slot s or -1; uslot[u] = s; urr[u] = reference row;
// rru[rr] = u; xs[s] = x* of the unknown at slot s (unset for other slots:
// the fill reads it for kept entries only), so the fill gathers one value
// per entry instead of hashing the reference row of every neighbor.
__global__ __launch_bounds__(256) void k_unknown_maps(const int* __restrict__ perm, const int* __restrict__ fs,
                                                      const int* __restrict__ upos, const int* __restrict__ pos, int n,
                                                      int* __restrict__ uid, int* __restrict__ uslot,
                                                      int* __restrict__ urr, int* __restrict__ rru,
                                                      const double4* __restrict__ spos, int xmode, uint32_t xseed,
                                                      double* __restrict__ xs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (!fs[s]) {
    uid[s] = -1;
    return;
  }
  const int u = upos[s], rr = pos[perm[s]];
  if (xmode == 0) {
    const double4 p = spos[s];
    xs[s] = p.x * p.x + p.y * p.y;  // x* = x^2 + y^2 (src/lsmps.cpp:29)
  } else {
    xs[s] = xstar_random(xseed, rr);
  }
  uid[s] = u;
  uslot[u] = s;
  urr[u] = rr;
  rru[rr] = u;
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

// Synthetic weight of the replay edge {lo, hi} (lo <= hi, particle indices):
// (0.5 + u) / l0^2 with u in [0, 1) from the splitmix64 finaliser of
// (lo * 2^32 + hi) xor (seed * 2^64/phi); symmetric in the edge.
__device__ __forceinline__ double replay_weight(uint32_t seed, int lo, int hi, double l0) {
  const uint64_t key = (((uint64_t)(unsigned)lo << 32) | (uint64_t)(unsigned)hi) ^ ((uint64_t)seed * 0x9E3779B97F4A7C15ull);
  return (0.5 + unit53(mix64(key))) / (l0 * l0);
}

// Eight lanes per row, eight rows per wavefront, one 64-row slice per
// block. The lanes of a row walk its neighbor list in batches of four
// chunks of eight entries: all of a batch's list entries are loaded first,
// then their unknown ids and positions, so a lane has four gathers in
// flight instead of one.
constexpr int PPE_GROUP = 8;
constexpr int PPE_THREADS = PPE_SLICE * PPE_GROUP;  // one slice per block
constexpr int PPE_BATCH = 4;                        // chunks of PPE_GROUP entries per batch
constexpr int PPE_BATCH_ENTRIES = PPE_BATCH * PPE_GROUP;

// Slice-local row of this thread's group and the slice index.
struct PpeRow {
  int slice, u, g, gbase, local;
  bool live;
};
__device__ __forceinline__ PpeRow ppe_row(int m) {
  PpeRow R;
  R.slice = (int)xcd_block(blockIdx.x, gridDim.x);
  R.g = threadIdx.x & (PPE_GROUP - 1);
  R.gbase = threadIdx.x & ~(PPE_GROUP - 1) & (WAVE - 1);
  R.local = (int)(threadIdx.x / PPE_GROUP);
  R.u = R.slice * PPE_SLICE + R.local;
  R.live = R.u < m;
  return R;
}

// Upper bound of each row's length (unknown neighbors + diagonal; exact
// except for coincident or boundary-rounded pairs, which become padding)
// and the slice widths wsz[slice] = 64 * longest row.
__global__ __launch_bounds__(PPE_THREADS) void k_ppe_count(const int* __restrict__ uslot, const int* __restrict__ urr,
                                                           const int* __restrict__ uid, const int* __restrict__ nbr,
                                                           const int* __restrict__ ncnt, int maxn, int m, int pin0,
                                                           int* __restrict__ wsz) {
  __shared__ int wmax;
  if (threadIdx.x == 0) wmax = 0;
  __syncthreads();
  const PpeRow R = ppe_row(m);
  int c = 0;
  if (R.live && (!pin0 || urr[R.u] != 0)) {
    const int s = uslot[R.u];
    const int* row = nbr + (size_t)s * maxn;
    const int k = ncnt[s];
    for (int c0 = 0; c0 < k; c0 += PPE_BATCH_ENTRIES) {
      int s2[PPE_BATCH];
#pragma unroll
      for (int j = 0; j < PPE_BATCH; ++j) {
        const int q = c0 + j * PPE_GROUP + R.g;
        s2[j] = q < k ? row[q] : -1;
      }
#pragma unroll
      for (int j = 0; j < PPE_BATCH; ++j) c += (s2[j] >= 0 && uid[s2[j]] >= 0) ? 1 : 0;
    }
  }
  c = group8_sum_int(c);
  if (R.live && R.g == 0) atomicMax(&wmax, c + 1);
  __syncthreads();
  if (threadIdx.x == 0) wsz[R.slice] = wmax * PPE_SLICE;
}

// The fill pass. Outputs in row order u: the sliced-ELL columns (sorted
// numbering) and FP64 values, the row lengths, the diagonal, x* and b.
// Counts per block into bpart (then counters): [0] failed rows, [1] NearWall
// rows except row 0, [2] nonzeros.
// Each lane computes its entries' a = w / r^2 and -a x*_j; ordered ballots
// place the kept entries; the batch's terms go to LDS in list order (an
// entry that is not kept contributes exact zeros, which leave the sums
// unchanged), and lanes 0 and 1 of the row add them to the diagonal and to
// b in that order, as the reference's sequential loop does.
// REPLAY: the edges of a replay graph (slots are particle indices) with
// replay_weight(wseed, ...) in place of w/r^2 from positions (re unused);
// self-edges are skipped as the reference skips r <= 1e-12.
template <bool REPLAY>
__global__ __launch_bounds__(PPE_THREADS) void k_ppe_fill(
    const double4* __restrict__ spos, const uint8_t* __restrict__ type_s, const int* __restrict__ uslot,
    const int* __restrict__ urr, const int* __restrict__ uid, const double* __restrict__ xs,
    const int* __restrict__ nbr, const int* __restrict__ ncnt, int maxn, int m, double re, double shift,
    uint32_t wseed, double l0, const int* __restrict__ sptr, int* __restrict__ col, double* __restrict__ val,
    int* __restrict__ rlen, double* __restrict__ diag, double* __restrict__ xe, double* __restrict__ b,
    unsigned* __restrict__ bpart) {
  __shared__ unsigned lds[3 * PPE_THREADS / WAVE];
  __shared__ double terms[PPE_SLICE][2][PPE_BATCH_ENTRIES];  // per row: a, then -a x*_j
  const PpeRow R = ppe_row(m);
  const unsigned below = (1u << R.g) - 1u;
  int failed = 0, neumann = 0, len = 0;
  if (R.live) {
    const int u = R.u;
    const int s = uslot[u];
    const int rr = urr[u];
    const double4 pi = spos[s];
    const int sb = sptr[R.slice];
    const int width = (sptr[R.slice + 1] - sb) / PPE_SLICE;
    const int base = sb + (u - R.slice * PPE_SLICE);
    const double xu = xs[s];
    double acc = 0.0;  // lane 0: the diagonal; lane 1: b (both in list order)
    int kept = 0, below_u = 0;
    if (rr != 0) {
      const int* row = nbr + (size_t)s * maxn;
      const int k = ncnt[s];
      double* tl = &terms[R.local][0][0];
      for (int c0 = 0; c0 < k; c0 += PPE_BATCH_ENTRIES) {
        int s2[PPE_BATCH], cj[PPE_BATCH];
        double4 pj[PPE_BATCH];
        double xjs[PPE_BATCH];
#pragma unroll
        for (int j = 0; j < PPE_BATCH; ++j) {
          const int q = c0 + j * PPE_GROUP + R.g;
          s2[j] = q < k ? row[q] : -1;
        }
#pragma unroll
        for (int j = 0; j < PPE_BATCH; ++j) cj[j] = s2[j] >= 0 ? uid[s2[j]] : -1;
#pragma unroll
        for (int j = 0; j < PPE_BATCH; ++j) xjs[j] = s2[j] >= 0 ? xs[s2[j]] : 0.0;
        if (!REPLAY) {
#pragma unroll
          for (int j = 0; j < PPE_BATCH; ++j) pj[j] = cj[j] >= 0 ? spos[s2[j]] : pi;
        }
#pragma unroll
        for (int j = 0; j < PPE_BATCH; ++j) {
          double a = 0.0, xj = 0.0;
          bool keep = false;
          if (cj[j] >= 0) {
            if (REPLAY) {
              keep = s2[j] != s;
              if (keep) {
                a = replay_weight(wseed, s < s2[j] ? s : s2[j], s < s2[j] ? s2[j] : s, l0);
                xj = xjs[j];
              }
            } else {
              keep = pair_coef(pj[j].x - pi.x, pj[j].y - pi.y, pj[j].z - pi.z, re, &a);
              if (keep) xj = xjs[j];
            }
          }
          const unsigned gm = group8_ballot(keep, R.gbase);
          const unsigned lt = group8_ballot(keep && cj[j] < u, R.gbase);
          if (keep) {
            // Entries with a column above u sit one slot later: the diagonal
            // takes slot below_u.
            const int slot = kept + __popc(gm & below) + (cj[j] > u ? 1 : 0);
            col[base + slot * PPE_SLICE] = cj[j];
            val[base + slot * PPE_SLICE] = -a;
          }
          kept += __popc(gm);
          below_u += __popc(lt);
          const int t = j * PPE_GROUP + R.g;
          tl[t] = keep ? a : 0.0;
          tl[PPE_BATCH_ENTRIES + t] = keep ? -a * xj : 0.0;
        }
        wave_sync();  // a row's terms are written and read by its own wavefront
        if (R.g < 2) {
          const double* src = tl + R.g * PPE_BATCH_ENTRIES;
          const int nt = k - c0 < PPE_BATCH_ENTRIES ? k - c0 : PPE_BATCH_ENTRIES;
          for (int t = 0; t < nt; ++t) acc += src[t];
        }
        wave_sync();  // read before the next batch overwrites them
      }
    }
    double d = __shfl(acc, R.gbase), bs = __shfl(acc, R.gbase + 1);
    if (rr != 0) {
      if (d <= 0 || !isfinite(d)) {
        failed = 1;
        d = 1.0;
      }
      if (shift != 0.0) d += shift * d;
      neumann = type_s[s] == T_NEARWALL ? 1 : 0;
    } else {
      d = 1.0;  // the identity pin of row 0
    }
    bs += d * xu;
    len = kept + 1;
    if (R.g == 0) {
      col[base + below_u * PPE_SLICE] = u;
      val[base + below_u * PPE_SLICE] = d;
      rlen[u] = len;
      diag[u] = d;
      xe[u] = xu;
      b[u] = bs;
    }
    for (int k2 = len + R.g; k2 < width; k2 += PPE_GROUP) {
      col[base + k2 * PPE_SLICE] = u;
      val[base + k2 * PPE_SLICE] = 0.0;
    }
  } else {
    // Rows past m in the last slice: every stored entry is defined, as
    // (0, 0), so passes over all stored entries stay in bounds.
    const int sb = sptr[R.slice];
    const int width = (sptr[R.slice + 1] - sb) / PPE_SLICE;
    const int base = sb + (R.u - R.slice * PPE_SLICE);
    for (int k2 = R.g; k2 < width; k2 += PPE_GROUP) {
      col[base + k2 * PPE_SLICE] = 0;
      val[base + k2 * PPE_SLICE] = 0.0;
    }
  }
  const bool leader = R.live && R.g == 0;
  unsigned v[3] = {leader ? (unsigned)failed : 0u, leader ? (unsigned)neumann : 0u, leader ? (unsigned)len : 0u};
  block_count_partials<3>(v, bpart, lds);
}

// z = M^-1 (0,0,0,1,1,1,0,0,0) for the particle at pi with the reference's
// moments (list order) and Cholesky ladder; false when the weight sum is not
// positive or every rung fails.
__device__ __noinline__ bool wls_lap_weights(const double4 pi, const int* __restrict__ row, int k,
                                             const double4* __restrict__ spos, double re, double l0, double reg,
                                             double (&z)[9], double* ws_out) {
  const double h2 = 2 * l0;
  for (int attempt = 0; attempt < 5; ++attempt) {
    double M[45], ws = 0.0;
#pragma unroll
    for (int e = 0; e < 45; ++e) M[e] = 0.0;
    // Each record is loaded one neighbor ahead and its index two ahead;
    // past the row's end they reload its last entry.
    double4 pn = k > 0 ? spos[row[0]] : pi;
    int jn = row[k > 1 ? 1 : 0];
    for (int q = 0; q < k; ++q) {
      const double4 pj = pn;
      pn = spos[jn];
      jn = row[q + 2 < k ? q + 2 : k - 1];
      const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r <= 1e-12 || r >= re) continue;
      const double tq = re / r - 1.0;
      const double w = tq * tq;
      if (w == 0.0) continue;
      const double inv = 1.0 / r, h = h2 * r, g = l0 * r;
      const double qv[9] = {dx * inv,    dy * inv,    dz * inv,    dx * dx / h, dy * dy / h,
                            dz * dz / h, dx * dy / g, dx * dz / g, dy * dz / g};
      ws += w;
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        const double wq = w * qv[a];
#pragma unroll
        for (int c = 0; c <= a; ++c) M[tri(a, c)] += wq * qv[c];
      }
    }
    if (ws <= 0) return false;
#pragma unroll
    for (int e = 0; e < 45; ++e) M[e] /= ws;
    const double add = attempt == 0 ? 0.0 : attempt == 1 ? reg * 1.0 : attempt == 2 ? reg * 10.0
                                                            : attempt == 3 ? reg * 100.0 : reg * 1000.0;
    if (attempt > 0 && add == 0.0) return false;
    if (add != 0.0) {
#pragma unroll
      for (int a = 0; a < 9; ++a) M[tri(a, a)] += add;
    }
    const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!chol9<double>(M, 1e-14, 0.0, zero9)) continue;
    const double e3[9] = {0, 0, 0, 1, 1, 1, 0, 0, 0};
    chol9_solve<double>(M, e3, z);
    *ws_out = ws;
    return true;
  }
  return false;
}

constexpr int PPE_WLS_THREADS = 128;
// A kept WLS row is weakly diagonally dominant with a positive diagonal, so
// every Gershgorin disc of D^-1 A lies in Re >= 0 (|z - 1| <= 1). A weaker
// bound (0.25) solved a 30k wall film but left the 3M industrial scene with
// eigenvalues in the left half-plane: BiCGStab broke down there.
constexpr double WLS_ROW_DOMINANCE = 1.0;
// One thread per stored row u < mpad (all rows of all slices). Counts per
// block into bpart: [0] rows whose diagonal was set to 1, [1] NearWall
// rows, [2] nonzeros, [3] rows that took the graph-Laplacian row.
__global__ __launch_bounds__(PPE_WLS_THREADS) void k_ppe_wls_fill(
    const double4* __restrict__ spos, const uint8_t* __restrict__ type_s, const int* __restrict__ uslot,
    const int* __restrict__ urr, const int* __restrict__ uid, const int* __restrict__ nbr,
    const int* __restrict__ ncnt, int maxn, int m, int mpad, double re, double l0, double reg, double shift,
    int xmode, uint32_t xseed, const int* __restrict__ sptr, int* __restrict__ col, double* __restrict__ val,
    int* __restrict__ rlen, double* __restrict__ diag, double* __restrict__ xe, double* __restrict__ b,
    unsigned* __restrict__ bpart) {
  __shared__ unsigned lds[4 * PPE_WLS_THREADS / WAVE];
  const int u = (int)xcd_block(blockIdx.x, gridDim.x) * PPE_WLS_THREADS + threadIdx.x;
  int failed = 0, neumann = 0, len = 0, graph = 0;
  if (u < mpad) {
    const int sl = u / PPE_SLICE;
    const int sb = sptr[sl];
    const int width = (sptr[sl + 1] - sb) / PPE_SLICE;
    const int base = sb + (u - sl * PPE_SLICE);
    if (u < m) {
      const int s = uslot[u];
      const int rr = urr[u];
      const double4 pi = spos[s];
      const double xu = xmode == 0 ? pi.x * pi.x + pi.y * pi.y : xstar_random(xseed, rr);
      const int* row = nbr + (size_t)s * maxn;
      const int k = ncnt[s];
      double z[9], ws = 0.0;
      bool ok = wls_lap_weights(pi, row, k, spos, re, l0, reg, z, &ws);
      double d = 0.0, bs = 0.0, offabs = 0.0;
      int kept = 0, below_u = 0;
      if (ok) {
        const double h2 = 2 * l0;
        double4 pn = k > 0 ? spos[row[0]] : pi;
        int sn = row[0], jn = row[k > 1 ? 1 : 0];
        for (int q = 0; q < k; ++q) {
          const int s2 = sn;
          const double4 pj = pn;
          pn = spos[jn];
          sn = jn;
          jn = row[q + 2 < k ? q + 2 : k - 1];
          const double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
          const double r = sqrt(dx * dx + dy * dy + dz * dz);
          if (r <= 1e-12 || r >= re) continue;
          const double tq = re / r - 1.0;
          const double w = tq * tq;
          if (w == 0.0) continue;
          const double inv = 1.0 / r, h = h2 * r, g = l0 * r;
          const double qv[9] = {dx * inv,    dy * inv,    dz * inv,    dx * dx / h, dy * dy / h,
                                dz * dz / h, dx * dy / g, dx * dz / g, dy * dz / g};
          double dot = 0.0;
#pragma unroll
          for (int a = 0; a < 9; ++a) dot += z[a] * qv[a];
          const double c = dot * w / r / ws / l0;
          d += c;
          const int cj = uid[s2];
          if (cj >= 0) {
            // Columns above u sit one slot later: the diagonal takes slot below_u.
            const int slot = kept + (cj > u ? 1 : 0);
            col[base + slot * PPE_SLICE] = cj;
            val[base + slot * PPE_SLICE] = -c;
            const double xj = xmode == 0 ? pj.x * pj.x + pj.y * pj.y : xstar_random(xseed, urr[cj]);
            bs += -c * xj;
            offabs += fabs(c);
            ++kept;
            below_u += cj < u ? 1 : 0;
          }
        }
        ok = isfinite(d) && d > 0.0 && d >= WLS_ROW_DOMINANCE * offabs;
      }
      if (!ok) {
        // The reference's graph-Laplacian row (src/lsmps.cpp:41-51).
        graph = 1;
        d = 0.0;
        bs = 0.0;
        kept = 0;
        below_u = 0;
        for (int q = 0; q < k; ++q) {
          const int s2 = row[q];
          const int cj = uid[s2];
          if (cj < 0) continue;
          const double4 pj = spos[s2];
          double a;
          if (!pair_coef(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, re, &a)) continue;
          const int slot = kept + (cj > u ? 1 : 0);
          col[base + slot * PPE_SLICE] = cj;
          val[base + slot * PPE_SLICE] = -a;
          const double xj = xmode == 0 ? pj.x * pj.x + pj.y * pj.y : xstar_random(xseed, urr[cj]);
          d += a;
          bs += -a * xj;
          ++kept;
          below_u += cj < u ? 1 : 0;
        }
        if (d <= 0 || !isfinite(d)) {
          failed = 1;
          d = 1.0;
        }
      }
      if (shift != 0.0) d += shift * d;
      neumann = type_s[s] == T_NEARWALL ? 1 : 0;
      bs += d * xu;
      len = kept + 1;
      col[base + below_u * PPE_SLICE] = u;
      val[base + below_u * PPE_SLICE] = d;
      rlen[u] = len;
      diag[u] = d;
      xe[u] = xu;
      b[u] = bs;
      for (int k2 = len; k2 < width; ++k2) {
        col[base + k2 * PPE_SLICE] = u;
        val[base + k2 * PPE_SLICE] = 0.0;
      }
    } else {
      for (int k2 = 0; k2 < width; ++k2) {
        col[base + k2 * PPE_SLICE] = 0;
        val[base + k2 * PPE_SLICE] = 0.0;
      }
    }
  }
  unsigned v[4] = {(unsigned)failed, (unsigned)neumann, (unsigned)len, (unsigned)graph};
  block_count_partials<4>(v, bpart, lds);
}

// ---- reference-order views, for downloads and uploads -----------------

// cnt_rr[rr] = rlen[rru[rr]] (CSR row lengths in reference order).
__global__ __launch_bounds__(256) void k_rr_lengths(const int* __restrict__ rru, const int* __restrict__ rlen, int m,
                                                    int* __restrict__ cnt_rr) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr < m) cnt_rr[rr] = rlen[rru[rr]];
}

// CSR row rr: the off-diagonal entries of sliced-ELL row u = rru[rr] in
// order with columns renumbered to reference rows, then the diagonal.
__global__ __launch_bounds__(256) void k_sell_to_csr(const int* __restrict__ rru, const int* __restrict__ urr,
                                                     const int* __restrict__ sptr, const int* __restrict__ scol,
                                                     const double* __restrict__ sval, const int* __restrict__ rlen,
                                                     const int* __restrict__ prow, int m, int* __restrict__ pcol,
                                                     double* __restrict__ pval) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= m) return;
  const int u = rru[rr];
  const int sl = u / PPE_SLICE;
  const int base = sptr[sl] + (u - sl * PPE_SLICE);
  const int len = rlen[u];
  int at = prow[rr];
  double dv = 0.0;
  for (int k = 0; k < len; ++k) {
    const int c = scol[base + k * PPE_SLICE];
    const double v = sval[base + k * PPE_SLICE];
    if (c == u) {
      dv = v;
      continue;
    }
    pcol[at] = urr[c];
    pval[at] = v;
    ++at;
  }
  pcol[at] = rr;
  pval[at] = dv;
}

// dst[rr] = src[rru[rr]] (row-order vector to reference order) and the
// particle index of each reference row.
__global__ __launch_bounds__(256) void k_to_rr(const int* __restrict__ rru, const double* __restrict__ src, int m,
                                               double* __restrict__ dst) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr < m) dst[rr] = src[rru[rr]];
}
__global__ __launch_bounds__(256) void k_rr_particle(const int* __restrict__ rru, const int* __restrict__ uslot,
                                                     const int* __restrict__ perm, int m, int* __restrict__ rev) {
  const int rr = blockIdx.x * blockDim.x + threadIdx.x;
  if (rr < m) rev[rr] = perm[uslot[rru[rr]]];
}
// dst[u] = src[urr[u]] (reference order to row order).
__global__ __launch_bounds__(256) void k_from_rr(const int* __restrict__ urr, const double* __restrict__ src, int m,
                                                 double* __restrict__ dst) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < m) dst[u] = src[urr[u]];
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

int ppe_slices(int m) { return (m + PPE_SLICE - 1) / PPE_SLICE; }
static int ppe_fill_blocks(int m) { return ppe_slices(m); }

void launch_unknown_flags(const uint8_t* type_s, const int* iperm, int n, int* flag, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unknown_flags, dim3(blocks(n, 256)), dim3(256), 0, st, type_s, iperm, n, flag);
}
void launch_unknown_slots(const int* perm, const int* flag, const int* pos, int n, int cap, int* fs, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unknown_slots, dim3(blocks(n, 256)), dim3(256), 0, st, perm, flag, pos, n, cap, fs);
}
void launch_unknown_maps(const int* perm, const int* fs, const int* upos, const int* pos, int n, int* uid, int* uslot,
                         int* urr, int* rru, const double4* spos, int xmode, uint32_t xseed, double* xs,
                         hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_unknown_maps, dim3(blocks(n, 256)), dim3(256), 0, st, perm, fs, upos, pos, n, uid, uslot, urr,
                       rru, spos, xmode, xseed, xs);
}
void launch_ppe_count(const int* uslot, const int* urr, const int* uid, const int* nbr, const int* ncnt, int maxn,
                      int m, bool pin0, int* wsz, hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_ppe_count, dim3(ppe_fill_blocks(m)), dim3(PPE_THREADS), 0, st, uslot, urr, uid, nbr, ncnt,
                       maxn, m, pin0 ? 1 : 0, wsz);
}
void launch_ppe_fill(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr, const int* uid,
                     const double* xs, const int* nbr, const int* ncnt, int maxn, int m, double re, double shift,
                     const int* sptr, int* col, double* val, int* rlen, double* diag, double* xe,
                     double* b, unsigned long long* counters, unsigned* bpart, hipStream_t st, hipEvent_t ev_begin,
                     hipEvent_t ev_end, bool replay, uint32_t wseed, double l0) {
  if (m <= 0) return;
  const int g = ppe_fill_blocks(m);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  if (replay)
    hipLaunchKernelGGL(k_ppe_fill<true>, dim3(g), dim3(PPE_THREADS), 0, st, spos, type_s, uslot, urr, uid, xs, nbr,
                       ncnt, maxn, m, re, shift, wseed, l0, sptr, col, val, rlen, diag, xe, b, bpart);
  else
    hipLaunchKernelGGL(k_ppe_fill<false>, dim3(g), dim3(PPE_THREADS), 0, st, spos, type_s, uslot, urr, uid, xs, nbr,
                       ncnt, maxn, m, re, shift, 0u, 0.0, sptr, col, val, rlen, diag, xe, b, bpart);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  launch_sum_counts(bpart, g, CountSlots{3, {0, 1, 2, 0, 0, 0, 0, 0}}, counters, st);
}
void launch_ppe_wls_fill(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr,
                         const int* uid, const int* nbr, const int* ncnt, int maxn, int m, double re, double l0,
                         double reg, double shift, int xmode, uint32_t xseed, const int* sptr, int* col, double* val,
                         int* rlen, double* diag, double* xe, double* b, unsigned long long* counters,
                         unsigned* bpart, hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
  if (m <= 0) return;
  const int mpad = ppe_slices(m) * PPE_SLICE;
  const int g = (int)blocks(mpad, PPE_WLS_THREADS);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  hipLaunchKernelGGL(k_ppe_wls_fill, dim3(g), dim3(PPE_WLS_THREADS), 0, st, spos, type_s, uslot, urr, uid, nbr, ncnt,
                     maxn, m, mpad, re, l0, reg, shift, xmode, xseed, sptr, col, val, rlen, diag, xe, b, bpart);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  launch_sum_counts(bpart, g, CountSlots{4, {0, 1, 2, 3, 0, 0, 0, 0}}, counters, st);
}
void launch_rr_lengths(const int* rru, const int* rlen, int m, int* cnt_rr, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_rr_lengths, dim3(blocks(m, 256)), dim3(256), 0, st, rru, rlen, m, cnt_rr);
}
void launch_sell_to_csr(const int* rru, const int* urr, const int* sptr, const int* scol, const double* sval,
                        const int* rlen, const int* prow, int m, int* pcol, double* pval, hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_sell_to_csr, dim3(blocks(m, 256)), dim3(256), 0, st, rru, urr, sptr, scol, sval, rlen, prow,
                       m, pcol, pval);
}
void launch_to_rr(const int* rru, const double* src, int m, double* dst, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_to_rr, dim3(blocks(m, 256)), dim3(256), 0, st, rru, src, m, dst);
}
void launch_rr_particle(const int* rru, const int* uslot, const int* perm, int m, int* rev, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_rr_particle, dim3(blocks(m, 256)), dim3(256), 0, st, rru, uslot, perm, m, rev);
}
void launch_from_rr(const int* urr, const double* src, int m, double* dst, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_from_rr, dim3(blocks(m, 256)), dim3(256), 0, st, urr, src, m, dst);
}

}  // namespace lsmps_hip

// BiCGStab for the PPE on gfx950, with device-resident scalars.
//
// Algorithm: van der Vorst's BiCGStab as the reference runs it
// (src/lsmps.cpp:65-67): rho = r0.r, beta = (rho/rho_prev)(alpha/omega),
// p = r + beta (p - omega v), v = A p, alpha = rho / r0.v, s = r - alpha v,
// t = A s, omega = t.s / t.t, x += alpha p + omega s, r = s - omega t, with
// the reference's 1e-30 breakdown guards and stop rule
// ||.|| <= tol * max(1, ||b||).
//
// The system solved is A S y = b, x = S y, where S is a diagonal scaling:
// identity (the reference), 1/diag(A) (right Jacobi) or a learned diagonal.
// The scaled matrix is made once per solve in the solve precision T by one
// streaming pass over the assembled FP64 matrix.
//
// The iteration matrix is stored sliced-ELL: rows in slices of 64 (one
// wavefront), each slice padded to its longest row and stored column-major,
// so one thread owns one row, value and column loads are fully coalesced,
// and neighboring rows (neighboring particles) gather neighboring vector
// entries. Padding entries hold value 0 and the row's own column.
//
// One iteration is five streaming launches and one one-block launch, with
// the scalars device-resident (no host round trip):
//   k_bicg_p   p = r + beta (p - omega v)
//   k_bicg_v   v = A p, partial r0.v
//   k_bicg_s   alpha = rho / r0.v, s = r - alpha v
//   k_bicg_t   t = A s (slices in reverse order), partial s.s, t.t, t.s
//   k_bicg_xr  omega = t.s / t.t, y += alpha p + omega s (FP64),
//              r = s - omega t, partial r.r, r0.r     -> k_fin_iter
// Every block of k_bicg_s and k_bicg_xr adds the previous launch's block
// partials itself, in the same fixed order, so all blocks hold the same
// alpha (omega) bit for bit and no one-block launch sits between the two
// streaming ones; block 0 stores the scalar for the later launches.
// so each matrix entry gathers one vector entry. Columns inside a row are
// stored in ascending order, so the 64 rows of a slice (64 neighboring
// particles) gather neighboring entries.
//
// Mixed precision (matrix type TM = float): the iteration matrix A S in
// FP32 (half the value bytes of every SpMV), the Krylov vectors, dots and y
// in FP64, and reliable updates: when the recurrence residual has dropped
// by `delta` since the last replacement, or when it meets the tolerance,
// k_bicg_true_residual recomputes r = b - A S y from the FP64 matrix and
// FP64 y, so the stopping test is on the true FP64 residual. With
// SOLVE_RESTART the Krylov space then restarts from the true residual
// (r0 = r, p = v = 0): each restart is one step of iterative refinement
// against the FP32 matrix. The PPE is a graph Laplacian pinned at one row,
// so its conditioning grows with the scene; with the vectors in FP32 as
// well, BiCGStab's recurrences there lose enough accuracy that iteration
// counts scatter over 1x-8x of the FP64 solver's and some solves never
// converge, while an FP32 matrix with FP64 vectors tracks the FP64 count.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "kernels.hpp"
#include "solver.hpp"

namespace lsmps_hip {

constexpr int SOLVER_THREADS = 256;

// Vector streams of the iteration (p, s, r, y, v, t and r0 in the streaming
// kernels) use non-temporal accesses, so they displace less of the
// iteration matrix from the memory-side cache (MALL, 256 MB): the two
// products of an iteration sweep the matrix in opposite directions
// (k_bicg_v forward, k_bicg_t backward), and each starts on the part the
// other streamed last. At 3M (FP32 matrix 0.52 GB, FP64 0.87 GB) this took
// 6% off an iteration in both precisions. The gathered vector entries stay
// ordinary loads (each is read ~30 times, from L2).
template <typename T>
__device__ __forceinline__ T nt_load(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void nt_store(T v, T* p) {
  __builtin_nontemporal_store(v, p);
}
constexpr int SLICE = 64;

// The state word is read by every block of k_bicg_s and k_bicg_xr while
// block 0 of the same launch may store a decision that every block reaches
// by itself; relaxed atomic accesses make that concurrent access defined.
__device__ __forceinline__ int load_state(const SolverScalars* sc) {
  return __hip_atomic_load(&sc->state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_state(SolverScalars* sc, int state) {
  __hip_atomic_store(&sc->state, state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool iter_active(const SolverScalars* sc) {
  return load_state(sc) == SOLVE_RUNNING && sc->it < sc->max_iter;
}

// A breakdown that the solve may answer with a restart from the true
// residual of its iterate (SOLVE_BREAK_RESTART, a bounded number of times);
// every block reads the same n_break (only k_fin_residual changes it).
__device__ __forceinline__ bool can_restart(const SolverScalars* sc) {
  return (sc->flags & SOLVE_BREAK_RESTART) && sc->n_break < SOLVER_MAX_BREAK_RESTARTS;
}

// After a step that leaves r updated: decide stop/continue and next beta.
// Runs in one thread of the last block.
__device__ void finish_iteration(SolverScalars* sc, double rr, double r0r) {
  const double rn = sqrt(rr);
  sc->rn = rn;
  sc->it += 1;
  sc->iterations = sc->it;
  if (!isfinite(rn) || !isfinite(r0r)) {
    sc->state = SOLVE_BREAKDOWN;
    return;
  }
  if (rn <= sc->thresh) {
    if (sc->flags & SOLVE_TRUE_CHECK) {
      sc->rho_prev = sc->rho;
      sc->replace = 2;
    } else {
      sc->state = SOLVE_CONVERGED;
    }
    return;
  }
  if (fabs(sc->omega) < 1e-30) {
    if (can_restart(sc)) sc->replace = 3;
    else sc->state = SOLVE_BREAKDOWN;
    return;
  }
  if ((sc->flags & SOLVE_RELIABLE) && rn < sc->delta * sc->rn_max) {
    sc->rho_prev = sc->rho;
    sc->replace = 1;
    return;
  }
  if (rn > sc->rn_max) sc->rn_max = rn;
  sc->rho_prev = sc->rho;
  sc->rho = r0r;
  if (fabs(r0r) < 1e-30) {
    if (can_restart(sc)) sc->replace = 3;
    else sc->state = SOLVE_BREAKDOWN;
    return;
  }
  sc->beta = (sc->rho / sc->rho_prev) * (sc->alpha / sc->omega);
  if (sc->it >= sc->max_iter) sc->state = SOLVE_MAXITER;
}

// Row `row` of a sliced-ELL matrix: first element index and width.
struct SellRow {
  int base, width;
};
__device__ __forceinline__ SellRow sell_row(const int* __restrict__ sptr, int row) {
  const int sl = row / SLICE;
  const int b = sptr[sl];
  return {b + (row - sl * SLICE), (sptr[sl + 1] - b) / SLICE};
}

// (A S y)_row with the FP64 matrix, S applied at the gathered column (S = 1
// when null); the products added in storage order, four entries per step
// with their loads first (a serial gather chain otherwise). ABS: also
// sum |a s y| over the row into *abs_sum (the scale of the row's product,
// against which an FP32 copy of A S rounds).
template <bool ABS = false>
__device__ __forceinline__ double sell_dot_scaled(const int* __restrict__ sptr, const int* __restrict__ col,
                                                  const double* __restrict__ val, const double* __restrict__ scale,
                                                  const double* __restrict__ y, int row, double* abs_sum = nullptr) {
  const SellRow R = sell_row(sptr, row);
  double acc = 0.0, ab = 0.0;
  int k = 0;
  for (; k + 4 <= R.width; k += 4) {
    int c[4];
    double v[4], yv[4], sv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = R.base + (k + j) * SLICE;
      c[j] = col[e];
      v[j] = val[e];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      yv[j] = y[c[j]];
      sv[j] = scale ? scale[c[j]] : 1.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double sy = scale ? sv[j] * yv[j] : yv[j];
      acc = fma(v[j], sy, acc);
      if (ABS) ab += fabs(v[j] * sy);
    }
  }
  for (; k < R.width; ++k) {
    const int e = R.base + k * SLICE;
    const int c = col[e];
    const double sy = scale ? scale[c] * y[c] : y[c];
    acc = fma(val[e], sy, acc);
    if (ABS) ab += fabs(val[e] * sy);
  }
  if (ABS) *abs_sum = ab;
  return acc;
}

__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_p(int m, const double* __restrict__ r,
                                                           const double* __restrict__ v, double* __restrict__ p,
                                                           const SolverScalars* sc) {
  if (!iter_active(sc)) return;
  const double beta = sc->beta, omega = sc->omega;
  for (int i = blockIdx.x * SOLVER_THREADS + threadIdx.x; i < m; i += gridDim.x * SOLVER_THREADS)
    nt_store(fma(beta, fma(-omega, nt_load(v + i), nt_load(p + i)), nt_load(r + i)), p + i);
}

// (A x)_row on the sliced-ELL matrix (values TM, FP64 x, FP64 sum), one
// thread per row.
template <typename TM>
__device__ __forceinline__ double sell_dot(const int* __restrict__ sptr, const int* __restrict__ col,
                                           const TM* __restrict__ val, const double* __restrict__ x, int row) {
  const SellRow R = sell_row(sptr, row);
  double acc = 0.0;
#pragma unroll 4
  for (int k = 0; k < R.width; ++k) {
    const int e = R.base + k * SLICE;
    acc = fma((double)val[e], x[col[e]], acc);
  }
  return acc;
}

// The same product with 16-bit packed columns: entry p of the slice is
// table[p >> 11] + (p & 2047), where the slice's table of up to 32 column
// bases sits in lanes 0..31 of `tab` (k_sell_pack). Same products, same
// order as sell_dot, so the same result; 2 bytes per entry instead of 4.
template <typename TM>
__device__ __forceinline__ double sell_dot_packed(const int* __restrict__ sptr, const uint16_t* __restrict__ pcol,
                                                  int tab, const TM* __restrict__ val, const double* __restrict__ x,
                                                  int row) {
  const SellRow R = sell_row(sptr, row);
  double acc = 0.0;
  int k = 0;
  if constexpr (sizeof(TM) == 4) {
    // FP32 values: eight entries per step, and the next step's column and
    // value loads issued before this step's gathers are waited on (386 ->
    // 361 us per mixed iteration at 3M; 4 or 12 entries per step measured
    // slower, and the FP64 product slower with either change, so it keeps
    // the plain loop below). Same products, same order.
    constexpr int NB = 8;
    if (R.width >= NB) {
      unsigned pk[NB];
      TM v[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int e = R.base + j * SLICE;
        pk[j] = pcol[e];
        v[j] = val[e];
      }
      for (;; k += NB) {
        double xv[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j)
          xv[j] = x[__shfl(tab, (int)(pk[j] >> SELL_PACK_OFF_BITS)) + (int)(pk[j] & SELL_PACK_OFF_MASK)];
        const bool more = k + 2 * NB <= R.width;
        unsigned pn[NB];
        TM vn[NB];
        if (more) {
#pragma unroll
          for (int j = 0; j < NB; ++j) {
            const int e = R.base + (k + NB + j) * SLICE;
            pn[j] = pcol[e];
            vn[j] = val[e];
          }
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) acc = fma((double)v[j], xv[j], acc);
        if (!more) break;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          pk[j] = pn[j];
          v[j] = vn[j];
        }
      }
      k += NB;
    }
  }
  // Four entries per step, loads first (the table exchange is a wavefront
  // operation, so the compiler does not unroll this loop by itself).
  for (; k + 4 <= R.width; k += 4) {
    unsigned pk[4];
    TM v[4];
    double xv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = R.base + (k + j) * SLICE;
      pk[j] = pcol[e];
      v[j] = val[e];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      xv[j] = x[__shfl(tab, (int)(pk[j] >> SELL_PACK_OFF_BITS)) + (int)(pk[j] & SELL_PACK_OFF_MASK)];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc = fma((double)v[j], xv[j], acc);
  }
  for (; k < R.width; ++k) {
    const int e = R.base + k * SLICE;
    const unsigned pk = pcol[e];
    const int c = __shfl(tab, (int)(pk >> SELL_PACK_OFF_BITS)) + (int)(pk & SELL_PACK_OFF_MASK);
    acc = fma((double)val[e], x[c], acc);
  }
  return acc;
}

// One slice (64 rows) per wavefront and loop step, every lane active (the
// last slice's rows past m are padding entries with value 0), so the packed
// columns' table exchange sees the whole wavefront. Each lane still owns the
// rows it owned row by row, so the partial sums keep their order.
template <typename TM>
__device__ __forceinline__ double sell_row_product(const int* __restrict__ sptr, const int* __restrict__ col,
                                                   const uint16_t* __restrict__ pcol, const int* __restrict__ ptab,
                                                   const TM* __restrict__ val, const double* __restrict__ x,
                                                   int row) {
  if (pcol) {
    const int slice = row / SLICE, lane = threadIdx.x & (WAVE - 1);
    const int tab = lane < SELL_PACK_IDS ? ptab[(size_t)slice * SELL_PACK_IDS + lane] : 0;
    if (__shfl(tab, 0) >= 0) return sell_dot_packed(sptr, pcol, tab, val, x, row);
  }
  return sell_dot(sptr, col, val, x, row);  // unpacked, or a slice too scattered to pack
}

template <typename TM>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_v(const int* __restrict__ sptr, const int* __restrict__ col,
                                                           const uint16_t* __restrict__ pcol,
                                                           const int* __restrict__ ptab, const TM* __restrict__ val,
                                                           int m, const double* __restrict__ p,
                                                           const double* __restrict__ r0, double* __restrict__ v,
                                                           SolverScalars* sc, double* __restrict__ partials) {
  if (!iter_active(sc)) return;
  __shared__ double lds[SOLVER_THREADS / WAVE];
  double dot = 0.0;
  const int mpad = (m + SLICE - 1) / SLICE * SLICE;
  for (int r0w = blockIdx.x * SOLVER_THREADS + (threadIdx.x & ~(WAVE - 1)); r0w < mpad;
       r0w += gridDim.x * SOLVER_THREADS) {
    const int row = r0w + (threadIdx.x & (WAVE - 1));
    const double acc = sell_row_product(sptr, col, pcol, ptab, val, p, row);
    if (row < m) {
      nt_store(acc, v + row);
      dot += nt_load(r0 + row) * acc;
    }
  }
  double vv[1] = {dot};
  block_partials<1>(vv, partials, lds);
}

// alpha from the r0.v block partials of k_bicg_v (every block adds them in
// the same order), then s = r - alpha v. A breakdown (r0.v ~ 0) is seen by
// every block alike; block 0 records it and alpha.
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_s(int m, const double* __restrict__ r,
                                                           const double* __restrict__ v, double* __restrict__ s,
                                                           SolverScalars* sc, const double* __restrict__ partials,
                                                           int nblk) {
  if (!iter_active(sc)) return;
  __shared__ double lds[SOLVER_THREADS / WAVE];
  double tot[1];
  finalize_sum<1>(partials, nblk, tot, lds);
  const double den = tot[0];
  const bool broke = fabs(den) < 1e-30 || !isfinite(den);
  const double alpha = broke ? 0.0 : sc->rho / den;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (broke && can_restart(sc)) {
      // The rest of this iteration is skipped; the true-residual launch
      // that follows restarts the Krylov space from the iterate.
      sc->replace = 3;
      store_state(sc, SOLVE_RESTARTING);
    } else if (broke) {
      store_state(sc, SOLVE_BREAKDOWN);
    } else {
      sc->alpha = alpha;
    }
  }
  if (broke) return;
  for (int i = blockIdx.x * SOLVER_THREADS + threadIdx.x; i < m; i += gridDim.x * SOLVER_THREADS)
    nt_store(fma(-alpha, nt_load(v + i), nt_load(r + i)), s + i);
}

template <typename TM>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_t(const int* __restrict__ sptr, const int* __restrict__ col,
                                                           const uint16_t* __restrict__ pcol,
                                                           const int* __restrict__ ptab, const TM* __restrict__ val,
                                                           int m, const double* __restrict__ s,
                                                           double* __restrict__ t, SolverScalars* sc,
                                                           double* __restrict__ partials) {
  if (!iter_active(sc)) return;
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  double ss = 0.0, tt = 0.0, ts = 0.0;
  const int mpad = (m + SLICE - 1) / SLICE * SLICE;
  for (int r0w = blockIdx.x * SOLVER_THREADS + (threadIdx.x & ~(WAVE - 1)); r0w < mpad;
       r0w += gridDim.x * SOLVER_THREADS) {
    // Slices in reverse order: the end of the matrix, which v = A p
    // streamed last, is the part still in the memory-side cache.
    const int row = (mpad - SLICE - r0w) + (threadIdx.x & (WAVE - 1));
    const double acc = sell_row_product(sptr, col, pcol, ptab, val, s, row);
    if (row < m) {
      const double si = nt_load(s + row);
      nt_store(acc, t + row);
      ss += si * si;
      tt += acc * acc;
      ts += acc * si;
    }
  }
  double vv[3] = {ss, tt, ts};
  block_partials<3>(vv, partials, lds);
}

// omega from the s.s, t.t, t.s block partials of k_bicg_t (every block
// adds them in the same order), then y += alpha p + omega s,
// r = s - omega t, and this launch's r.r and r0.r partials at
// partials + 3 * nblk (the first 3 * nblk are still being read by other
// blocks). When ||s|| already meets the tolerance (SOLVE_S_EXIT) the update
// is y += alpha p and the iteration ends there. Block 0 records ||s||,
// omega and the state for the later launches. The state read at entry is
// RUNNING, or S_CONVERGED when block 0 of this launch has already stored
// the decision every block reaches.
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_xr(int m, const double* __restrict__ p,
                                                            const double* __restrict__ s, const double* __restrict__ t,
                                                            const double* __restrict__ r0, double* __restrict__ r,
                                                            double* __restrict__ y, SolverScalars* sc,
                                                            double* __restrict__ partials, int nblk) {
  if (!(iter_active(sc) || load_state(sc) == SOLVE_S_CONVERGED)) return;
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  double tot[3];
  finalize_sum<3>(partials, nblk, tot, lds);
  const double sn = sqrt(tot[0]);
  const bool sconv = (sc->flags & SOLVE_S_EXIT) && sn <= sc->thresh;
  const bool broke = !sconv && (tot[1] < 1e-30 || !isfinite(tot[1]) || !isfinite(tot[2]));
  const double omega = sconv || broke ? 0.0 : tot[2] / tot[1];
  // On a breakdown every block returns without touching y (a block that
  // starts after block 0 stored the state returns at entry just the same);
  // with SOLVE_BREAK_RESTART the true-residual launch then restarts.
  const bool restart = broke && can_restart(sc);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->sn = sn;
    if (sconv) {
      store_state(sc, SOLVE_S_CONVERGED);
      sc->iterations = sc->it + 1;
    } else if (restart) {
      sc->replace = 3;
      store_state(sc, SOLVE_RESTARTING);
    } else if (broke) {
      store_state(sc, SOLVE_BREAKDOWN);
    } else {
      sc->omega = omega;
    }
  }
  if (broke) return;
  const double alpha = sc->alpha;
  double rr = 0.0, r0r = 0.0;
  for (int i = blockIdx.x * SOLVER_THREADS + threadIdx.x; i < m; i += gridDim.x * SOLVER_THREADS) {
    if (sconv) {
      y[i] += alpha * p[i];
    } else {
      const double si = nt_load(s + i);
      nt_store(nt_load(y + i) + (alpha * nt_load(p + i) + omega * si), y + i);
      const double ri = fma(-omega, nt_load(t + i), si);
      nt_store(ri, r + i);
      rr += ri * ri;
      r0r += nt_load(r0 + i) * ri;
    }
  }
  double vv[2] = {rr, r0r};
  block_partials<2>(vv, partials + 3 * (size_t)nblk, lds);
}

__global__ __launch_bounds__(SOLVER_THREADS) void k_fin_iter(const double* __restrict__ partials, int nblk,
                                                             SolverScalars* sc) {
  const bool sconv = sc->state == SOLVE_S_CONVERGED;
  if (!(sconv || iter_active(sc))) return;
  __shared__ double lds[2 * SOLVER_THREADS / WAVE];
  double tot[2];
  finalize_sum<2>(partials, nblk, tot, lds);
  if (threadIdx.x == 0) {
    if (sconv) {
      sc->state = SOLVE_CONVERGED;
      sc->it += 1;
    } else {
      finish_iteration(sc, tot[0], tot[1]);
    }
  }
}

// r = b - A S y with the FP64 matrix and FP64 y. INIT: also r0 = r, p = v
// = 0 and the scalars of a fresh solve. Otherwise runs only when a
// replacement is pending: sets rho/beta, stops when the true residual meets
// the tolerance, and with SOLVE_RESTART restarts the Krylov space.
template <bool INIT>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_true_residual(
    const int* __restrict__ sptr, const int* __restrict__ col, const double* __restrict__ val, int m,
    const double* __restrict__ b, const double* __restrict__ scale, const double* __restrict__ y,
    double* __restrict__ r, double* __restrict__ r0, double* __restrict__ p, double* __restrict__ v,
    SolverScalars* sc, double* __restrict__ partials) {
  if (!INIT && sc->replace == 0) return;
  const bool restart = !INIT && ((sc->flags & SOLVE_RESTART) || sc->replace == 3);
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  // bb: INIT, ||b||^2; otherwise the sum over rows of (sum |a s y|)^2.
  double rr = 0.0, r0r = 0.0, bb = 0.0;
  for (int row = blockIdx.x * SOLVER_THREADS + threadIdx.x; row < m; row += gridDim.x * SOLVER_THREADS) {
    double ab = 0.0;
    const double acc = y ? sell_dot_scaled<!INIT>(sptr, col, val, scale, y, row, &ab) : 0.0;
    if (!INIT) bb += ab * ab;
    const double bi = b[row];
    const double ri = bi - acc;
    r[row] = ri;
    if (INIT) {
      r0[row] = ri;
      p[row] = v[row] = 0.0;
      bb += bi * bi;
      rr += ri * ri;
    } else if (restart) {
      r0[row] = ri;
      p[row] = v[row] = 0.0;
      rr += ri * ri;
      r0r += ri * ri;
    } else {
      rr += ri * ri;
      r0r += r0[row] * ri;
    }
  }
  double vv[3] = {rr, r0r, bb};
  block_partials<3>(vv, partials, lds);
}

template <bool INIT>
__global__ __launch_bounds__(SOLVER_THREADS) void k_fin_residual(const double* __restrict__ partials, int nblk,
                                                                 SolverScalars* sc) {
  if (!INIT && sc->replace == 0) return;
  const bool restart = !INIT && ((sc->flags & SOLVE_RESTART) || sc->replace == 3);
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  double tot[3];
  finalize_sum<3>(partials, nblk, tot, lds);
  if (threadIdx.x == 0) {
    if (INIT) {
      sc->bnorm = sqrt(tot[2]);
      sc->thresh = sc->tol * (sc->bnorm > 1.0 ? sc->bnorm : 1.0);
      sc->init_res = sqrt(tot[0]);
      sc->rn = sc->rn_max = sc->init_res;
      sc->rn_best = sc->init_res;
      sc->fp32_floor = 0.0;
      sc->rho_prev = sc->alpha = sc->omega = 1.0;
      sc->rho = tot[0];
      sc->beta = sc->rho;
      sc->it = sc->iterations = 0;
      sc->replace = 0;
      sc->n_replace = 0;
      sc->n_break = 0;
      sc->state = SOLVE_RUNNING;
      if (sc->init_res <= sc->thresh && (sc->flags & SOLVE_TRUE_CHECK)) sc->state = SOLVE_CONVERGED;
      else if (fabs(sc->rho) < 1e-30) sc->state = SOLVE_BREAKDOWN;
      else if (sc->max_iter <= 0) sc->state = SOLVE_MAXITER;
    } else {
      const double rn = sqrt(tot[0]);
      const int kind = sc->replace;
      sc->replace = 0;
      sc->n_replace += 1;
      if (kind == 3) {
        sc->n_break += 1;
        sc->state = SOLVE_RUNNING;  // was RESTARTING when the breakdown came mid-iteration
      }
      sc->rn = rn;
      sc->rn_max = rn;
      if (!isfinite(rn) || !isfinite(tot[1])) {
        sc->state = SOLVE_BREAKDOWN;
        return;
      }
      // The true residual decides convergence whenever it is computed.
      if ((kind == 2 || (sc->flags & SOLVE_TRUE_CHECK)) && rn <= sc->thresh) {
        sc->state = SOLVE_CONVERGED;
        return;
      }
      // A reliable update that leaves the true residual above stall times
      // the best one so far: the FP32 iteration matrix no longer drives the
      // refinement (its error is as large as what is left to remove).
      if (kind == 1 && (sc->flags & SOLVE_STALL_CHECK) && !(rn <= sc->stall * sc->rn_best)) {
        sc->state = SOLVE_STALLED;
        return;
      }
      // The same hand-over when the tolerance lies far below what an FP32
      // copy of A S resolves on this iterate: rounding its entries moves the
      // product by up to 2^-24 sum |a s y| per row, and a target under
      // SOLVE_FP32_REACH times that norm takes one refinement restart per
      // decade or so, which costs more than FP64 iterations.
      sc->fp32_floor = 0x1p-24 * sqrt(tot[2]);
      if ((sc->flags & SOLVE_STALL_CHECK) && sc->thresh < SOLVE_FP32_REACH * sc->fp32_floor) {
        sc->state = SOLVE_STALLED;
        return;
      }
      if (rn < sc->rn_best) sc->rn_best = rn;
      sc->rho = tot[1];
      if (restart) sc->rho_prev = sc->alpha = sc->omega = 1.0;
      if (fabs(sc->rho) < 1e-30 || fabs(sc->omega) < 1e-30) {
        sc->state = SOLVE_BREAKDOWN;
        return;
      }
      sc->beta = (sc->rho / sc->rho_prev) * (sc->alpha / sc->omega);
      if (sc->it >= sc->max_iter) sc->state = SOLVE_MAXITER;
    }
  }
}

// x = S y; residual ||A x - b||, error against x*, max |x - x*|.
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_final(const int* __restrict__ sptr, const int* __restrict__ col,
                                                               const double* __restrict__ val, int m,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ scale,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ xe, double* __restrict__ x,
                                                               SolverScalars* sc, double* __restrict__ partials) {
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  double res = 0.0, se = 0.0, xn = 0.0, mx = 0.0;
  for (int row = blockIdx.x * SOLVER_THREADS + threadIdx.x; row < m; row += gridDim.x * SOLVER_THREADS) {
    const double acc = sell_dot_scaled(sptr, col, val, scale, y, row);
    const double d = acc - b[row];
    res += d * d;
    const double xi = scale ? scale[row] * y[row] : y[row];
    x[row] = xi;
    if (xe) {
      const double er = xi - xe[row];
      se += er * er;
      xn += xe[row] * xe[row];
      mx = fabs(er) > mx ? fabs(er) : mx;
    }
  }
  mx = wave_max(mx);
  if ((threadIdx.x & (WAVE - 1)) == 0 && mx > 0.0)
    atomicMax(reinterpret_cast<unsigned long long*>(&sc->max_err), (unsigned long long)__double_as_longlong(mx));
  double vv[3] = {res, se, xn};
  block_partials<3>(vv, partials, lds);
}

__global__ __launch_bounds__(SOLVER_THREADS) void k_fin_final(const double* __restrict__ partials, int nblk,
                                                              SolverScalars* sc) {
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  double tot[3];
  finalize_sum<3>(partials, nblk, tot, lds);
  if (threadIdx.x == 0) {
    sc->final_res = sqrt(tot[0]);
    sc->rel_l2 = sqrt(tot[1] / (tot[2] > 1e-30 ? tot[2] : 1e-30));
  }
}

// ---- iteration matrix from the assembled sliced-ELL -------------------
// val[e] = A[e] * S[col[e]] in precision T, over every stored entry
// (padding entries hold 0 and stay 0).
template <typename T>
__global__ __launch_bounds__(256) void k_sell_scale(const int* __restrict__ col, const double* __restrict__ val64,
                                                    const double* __restrict__ scale, int64_t nnz,
                                                    T* __restrict__ val) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    val[e] = (T)(scale ? val64[e] * scale[col[e]] : val64[e]);
}

__global__ __launch_bounds__(256) void k_inv_diag(const double* __restrict__ diag, int m, double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) s[i] = 1.0 / diag[i];
}

// y0 = x0 / S.
__global__ __launch_bounds__(256) void k_y_from_x(const double* __restrict__ x0, const double* __restrict__ scale, int m,
                                                  double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) y[i] = x0 ? (scale ? x0[i] / scale[i] : x0[i]) : 0.0;
}

__global__ void k_solver_setup(SolverScalars* sc, int max_iter, double tol, int flags, double delta, double stall) {
  if (threadIdx.x == 0) {
    sc->max_iter = max_iter;
    sc->tol = tol;
    sc->flags = flags;
    sc->delta = delta;
    sc->stall = stall;
    sc->max_err = 0.0;
    sc->state = SOLVE_RUNNING;
  }
}

// ---- 16-bit packed columns of the iteration matrix ---------------------
// Per slice (one wavefront): the distinct column blocks (col >> 11) of its
// stored entries go into a 64-slot LDS hash table; when there are at most
// 32, each gets an index and the slice's table holds their first columns,
// and every entry becomes (index << 11) | (col & 2047). Cell-sorted rows
// make a slice's columns fall into a few runs of neighboring cells (about
// 10 blocks at 3M particles). A slice with more than 32 blocks keeps its
// 32-bit columns: its table starts with -1.
__device__ __forceinline__ unsigned pack_hash(int key) { return ((unsigned)key * 2654435761u) >> 26; }

__global__ __launch_bounds__(256) void k_sell_pack(const int* __restrict__ sptr, const int* __restrict__ col,
                                                   int nslices, uint16_t* __restrict__ pcol, int* __restrict__ ptab) {
  __shared__ int keys[256 / WAVE][WAVE];
  __shared__ int ids[256 / WAVE][WAVE];
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int slice = blockIdx.x * (256 / WAVE) + w;
  if (slice >= nslices) return;  // whole wavefronts; no block barriers below
  keys[w][lane] = -1;
  wave_sync();
  const int sb = sptr[slice], width = (sptr[slice + 1] - sb) / SLICE;
  bool full = false;
  for (int k = 0; k < width; ++k) {
    const int key = col[sb + k * SLICE + lane] >> SELL_PACK_OFF_BITS;
    unsigned h = pack_hash(key);
    bool placed = false;
    for (int probe = 0; probe < WAVE && !placed; ++probe) {
      const int old = atomicCAS(&keys[w][h], -1, key);
      placed = old == -1 || old == key;
      h = (h + 1) & (WAVE - 1);
    }
    full = full || !placed;
  }
  wave_sync();
  const bool occ = keys[w][lane] != -1;
  const uint64_t mask = __builtin_amdgcn_ballot_w64(occ);
  const int id = __popcll(mask & ((1ull << lane) - 1ull));
  const bool packable = !__builtin_amdgcn_ballot_w64(full) && __popcll(mask) <= SELL_PACK_IDS;
  int* tab = ptab + (size_t)slice * SELL_PACK_IDS;
  if (!packable) {
    if (lane == 0) tab[0] = -1;
    return;
  }
  ids[w][lane] = id;
  if (occ) tab[id] = keys[w][lane] << SELL_PACK_OFF_BITS;
  if (lane >= __popcll(mask) && lane < SELL_PACK_IDS) tab[lane] = 0;
  wave_sync();
  for (int k = 0; k < width; ++k) {
    const int e = sb + k * SLICE + lane;
    const int c = col[e], key = c >> SELL_PACK_OFF_BITS;
    unsigned h = pack_hash(key);
    for (int probe = 0; probe < WAVE && keys[w][h] != key; ++probe) h = (h + 1) & (WAVE - 1);
    pcol[e] = (uint16_t)((ids[w][h] << SELL_PACK_OFF_BITS) | (c & SELL_PACK_OFF_MASK));
  }
}

void launch_sell_pack(const int* sptr, const int* col, int nslices, uint16_t* pcol, int* ptab, hipStream_t st) {
  if (nslices > 0)
    hipLaunchKernelGGL(k_sell_pack, dim3((unsigned)((nslices + 3) / 4)), dim3(256), 0, st, sptr, col, nslices, pcol,
                       ptab);
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

unsigned solver_grid(int m) {
  unsigned g = blocks(m, SOLVER_THREADS);
  return g < 1 ? 1 : (g > SOLVER_MAX_GRID ? SOLVER_MAX_GRID : g);
}

template <typename T>
void launch_sell_scale(const int* col, const double* val64, const double* scale, int64_t nnz, T* val,
                       hipStream_t st) {
  if (nnz <= 0) return;
  const int64_t want = (nnz + 255) / 256;
  const unsigned g = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(k_sell_scale<T>, dim3(g), dim3(256), 0, st, col, val64, scale, nnz, val);
}
template void launch_sell_scale<float>(const int*, const double*, const double*, int64_t, float*, hipStream_t);
template void launch_sell_scale<double>(const int*, const double*, const double*, int64_t, double*, hipStream_t);

void launch_inv_diag(const double* diag, int m, double* s, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_inv_diag, dim3(blocks(m, 256)), dim3(256), 0, st, diag, m, s);
}

template <typename TM>
void launch_solver_init(const SolverBufs<TM>& B, const double* x0, SolverScalars* sc, int max_iter, double tol,
                        int flags, double delta, double stall, hipStream_t st) {
  const int m = B.m;
  hipLaunchKernelGGL(k_solver_setup, dim3(1), dim3(64), 0, st, sc, max_iter, tol, flags, delta, stall);
  if (m > 0) hipLaunchKernelGGL(k_y_from_x, dim3(blocks(m, 256)), dim3(256), 0, st, x0, B.scale, m, B.y);
  const unsigned g = solver_grid(m);
  hipLaunchKernelGGL(k_bicg_true_residual<true>, dim3(g), dim3(SOLVER_THREADS), 0, st, B.sptr, B.col, B.val64, m,
                     B.b, B.scale, x0 ? (const double*)B.y : nullptr, B.r, B.r0, B.p, B.v, sc, B.partials);
  hipLaunchKernelGGL(k_fin_residual<true>, dim3(1), dim3(SOLVER_THREADS), 0, st, B.partials, (int)g, sc);
}

template <typename TM>
void launch_solver_iteration(const SolverBufs<TM>& B, bool with_replace, SolverScalars* sc, hipStream_t st) {
  const int m = B.m;
  const unsigned g = solver_grid(m);
  const dim3 one(1), blk(SOLVER_THREADS), grid(g);
  hipLaunchKernelGGL(k_bicg_p, grid, blk, 0, st, m, B.r, B.v, B.p, sc);
  hipLaunchKernelGGL((k_bicg_v<TM>), grid, blk, 0, st, B.sptr, B.col, B.pcol, B.ptab, B.val, m, B.p, B.r0, B.v, sc,
                     B.partials);
  hipLaunchKernelGGL(k_bicg_s, grid, blk, 0, st, m, B.r, B.v, B.s, sc, (const double*)B.partials, (int)g);
  hipLaunchKernelGGL((k_bicg_t<TM>), grid, blk, 0, st, B.sptr, B.col, B.pcol, B.ptab, B.val, m, B.s, B.t, sc,
                     B.partials);
  hipLaunchKernelGGL(k_bicg_xr, grid, blk, 0, st, m, B.p, B.s, B.t, B.r0, B.r, B.y, sc, B.partials, (int)g);
  hipLaunchKernelGGL(k_fin_iter, one, blk, 0, st, B.partials + 3 * (size_t)g, (int)g, sc);
  if (with_replace) {
    hipLaunchKernelGGL(k_bicg_true_residual<false>, grid, blk, 0, st, B.sptr, B.col, B.val64, m, B.b, B.scale,
                       (const double*)B.y, B.r, B.r0, B.p, B.v, sc, B.partials);
    hipLaunchKernelGGL(k_fin_residual<false>, one, blk, 0, st, B.partials, (int)g, sc);
  }
}

template <typename TM>
void launch_solver_final(const SolverBufs<TM>& B, const double* xe, double* x, SolverScalars* sc, hipStream_t st) {
  const int m = B.m;
  const unsigned g = solver_grid(m);
  hipLaunchKernelGGL(k_bicg_final, dim3(g), dim3(SOLVER_THREADS), 0, st, B.sptr, B.col, B.val64, m, B.b, B.scale,
                     (const double*)B.y, xe, x, sc, B.partials);
  hipLaunchKernelGGL(k_fin_final, dim3(1), dim3(SOLVER_THREADS), 0, st, B.partials, (int)g, sc);
}

template void launch_solver_init<float>(const SolverBufs<float>&, const double*, SolverScalars*, int, double, int, double,
                                        double, hipStream_t);
template void launch_solver_init<double>(const SolverBufs<double>&, const double*, SolverScalars*, int, double, int,
                                         double, double, hipStream_t);
template void launch_solver_iteration<float>(const SolverBufs<float>&, bool, SolverScalars*, hipStream_t);
template void launch_solver_iteration<double>(const SolverBufs<double>&, bool, SolverScalars*, hipStream_t);
template void launch_solver_final<float>(const SolverBufs<float>&, const double*, double*, SolverScalars*, hipStream_t);
template void launch_solver_final<double>(const SolverBufs<double>&, const double*, double*, SolverScalars*,
                                          hipStream_t);

}  // namespace lsmps_hip

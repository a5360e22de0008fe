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
// The scaled matrix is built once per solve in the solve precision T.
//
// One iteration is three launches, each ending in a deterministic grid
// reduction finished by the last block (no host round trip):
//   k_bicg_pv  p_new = r + beta (p_old - omega v_old) formed on the fly at
//              every column it gathers, v = A p_new, partial r0.v -> alpha
//   k_bicg_st  s = r - alpha v formed on the fly, t = A s, partial s.s,
//              t.t, t.s -> omega (or the early stop on ||s||)
//   k_bicg_xr  y += alpha p + omega s (FP64), r = s - omega t, partial
//              r.r, r0.r -> convergence, next rho and beta
// p and v are double-buffered because rows of k_bicg_pv read p_old/v_old
// at other rows while they write p_new/v_new.
//
// Mixed precision (T = float): matrix and Krylov vectors in FP32, every dot
// product accumulated in FP64, y in FP64, and reliable updates: when the
// recurrence residual has dropped 10x since the last replacement, or when
// it meets the tolerance, k_bicg_replace recomputes r = b - A S y from the
// FP64 matrix and FP64 y, so the stopping test is on the true FP64
// residual.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "solver.hpp"

namespace lsmps_hip {

constexpr int SOLVER_THREADS = 256;

__device__ __forceinline__ bool iter_active(const SolverScalars* sc) {
  return sc->state == SOLVE_RUNNING && sc->it < sc->max_iter;
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
    sc->state = SOLVE_BREAKDOWN;
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
    sc->state = SOLVE_BREAKDOWN;
    return;
  }
  sc->beta = (sc->rho / sc->rho_prev) * (sc->alpha / sc->omega);
  if (sc->it >= sc->max_iter) sc->state = SOLVE_MAXITER;
}

template <typename T, int G>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_pv(const int* __restrict__ rowp, const int* __restrict__ col,
                                                            const T* __restrict__ val, int m, const T* __restrict__ r,
                                                            const T* __restrict__ r0, const T* __restrict__ p_old,
                                                            const T* __restrict__ v_old, T* __restrict__ p_new,
                                                            T* __restrict__ v_new, SolverScalars* sc,
                                                            double* __restrict__ partials) {
  if (!iter_active(sc)) return;
  __shared__ double lds[SOLVER_THREADS / WAVE];
  const T beta = (T)sc->beta, omega = (T)sc->omega;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = SOLVER_THREADS / G;
  double dot = 0.0;
  for (int row = blockIdx.x * gpb + grp; row < m; row += gridDim.x * gpb) {
    T acc = 0;
    const int e = rowp[row + 1];
    for (int k = rowp[row] + lane; k < e; k += G) {
      const int c = col[k];
      const T pj = fma(beta, fma(-omega, v_old[c], p_old[c]), r[c]);
      acc = fma(val[k], pj, acc);
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      p_new[row] = fma(beta, fma(-omega, v_old[row], p_old[row]), r[row]);
      v_new[row] = acc;
      dot += (double)r0[row] * (double)acc;
    }
  }
  double v[1] = {dot}, tot[1];
  if (grid_sum_last_block<1>(v, partials, &sc->ticket[0], tot, lds) && threadIdx.x == 0) {
    const double den = tot[0];
    if (fabs(den) < 1e-30 || !isfinite(den)) {
      sc->state = SOLVE_BREAKDOWN;
    } else {
      sc->alpha = sc->rho / den;
    }
  }
}

template <typename T, int G>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_st(const int* __restrict__ rowp, const int* __restrict__ col,
                                                            const T* __restrict__ val, int m, const T* __restrict__ r,
                                                            const T* __restrict__ v, T* __restrict__ s,
                                                            T* __restrict__ t, SolverScalars* sc,
                                                            double* __restrict__ partials) {
  if (!iter_active(sc)) return;
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  const T alpha = (T)sc->alpha;
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = SOLVER_THREADS / G;
  double ss = 0.0, tt = 0.0, ts = 0.0;
  for (int row = blockIdx.x * gpb + grp; row < m; row += gridDim.x * gpb) {
    T acc = 0;
    const int e = rowp[row + 1];
    for (int k = rowp[row] + lane; k < e; k += G) {
      const int c = col[k];
      acc = fma(val[k], fma(-alpha, v[c], r[c]), acc);
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      const T si = fma(-alpha, v[row], r[row]);
      s[row] = si;
      t[row] = acc;
      ss += (double)si * (double)si;
      tt += (double)acc * (double)acc;
      ts += (double)acc * (double)si;
    }
  }
  double vv[3] = {ss, tt, ts}, tot[3];
  if (grid_sum_last_block<3>(vv, partials, &sc->ticket[1], tot, lds) && threadIdx.x == 0) {
    const double sn = sqrt(tot[0]);
    sc->sn = sn;
    if ((sc->flags & SOLVE_S_EXIT) && sn <= sc->thresh) {
      sc->state = SOLVE_S_CONVERGED;
      sc->iterations = sc->it + 1;
    } else if (tot[1] < 1e-30 || !isfinite(tot[1]) || !isfinite(tot[2])) {
      sc->state = SOLVE_BREAKDOWN;
    } else {
      sc->omega = tot[2] / tot[1];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_xr(int m, const T* __restrict__ p, const T* __restrict__ s,
                                                            const T* __restrict__ t, const T* __restrict__ r0,
                                                            T* __restrict__ r, double* __restrict__ y,
                                                            SolverScalars* sc, double* __restrict__ partials) {
  const int state = sc->state;
  const bool sconv = state == SOLVE_S_CONVERGED;
  if (!(sconv || iter_active(sc))) return;
  __shared__ double lds[2 * SOLVER_THREADS / WAVE];
  const double alpha = sc->alpha, omega = sc->omega;
  double rr = 0.0, r0r = 0.0;
  for (int i = blockIdx.x * SOLVER_THREADS + threadIdx.x; i < m; i += gridDim.x * SOLVER_THREADS) {
    if (sconv) {
      y[i] += alpha * (double)p[i];
    } else {
      const T si = s[i];
      y[i] += alpha * (double)p[i] + omega * (double)si;
      const T ri = fma((T)-omega, t[i], si);
      r[i] = ri;
      rr += (double)ri * (double)ri;
      r0r += (double)r0[i] * (double)ri;
    }
  }
  double vv[2] = {rr, r0r}, tot[2];
  if (grid_sum_last_block<2>(vv, partials, &sc->ticket[2], tot, lds) && threadIdx.x == 0) {
    if (sconv) {
      sc->state = SOLVE_CONVERGED;
      sc->it += 1;
    } else {
      finish_iteration(sc, tot[0], tot[1]);
    }
  }
}

// r = b - A S y with the FP64 matrix and FP64 y; optional r0 = r and zeroed
// p/v (INIT). In replace mode it then sets rho/beta (or stops when the
// true residual meets the tolerance).
template <typename T, int G, bool INIT>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_true_residual(
    const int* __restrict__ rowp, const int* __restrict__ col, const double* __restrict__ val, int m,
    const double* __restrict__ b, const double* __restrict__ scale, const double* __restrict__ y, T* __restrict__ r,
    T* __restrict__ r0, T* __restrict__ p0, T* __restrict__ p1, T* __restrict__ v0, T* __restrict__ v1,
    SolverScalars* sc, double* __restrict__ partials) {
  if (!INIT && sc->replace == 0) return;
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = SOLVER_THREADS / G;
  double rr = 0.0, r0r = 0.0, bb = 0.0;
  for (int row = blockIdx.x * gpb + grp; row < m; row += gridDim.x * gpb) {
    double acc = 0.0;
    if (y) {
      const int e = rowp[row + 1];
      for (int k = rowp[row] + lane; k < e; k += G) {
        const int c = col[k];
        const double yc = scale ? scale[c] * y[c] : y[c];
        acc = fma(val[k], yc, acc);
      }
      acc = group_sum<G>(acc);
    }
    if (lane == 0) {
      const double bi = b[row];
      const double ri = bi - acc;
      const T rt = (T)ri;
      r[row] = rt;
      if (INIT) {
        r0[row] = rt;
        p0[row] = p1[row] = v0[row] = v1[row] = (T)0;
        bb += bi * bi;
        rr += (double)rt * (double)rt;
      } else {
        rr += ri * ri;
        r0r += (double)r0[row] * (double)rt;
      }
    }
  }
  double vv[3] = {rr, r0r, bb}, tot[3];
  if (grid_sum_last_block<3>(vv, partials, &sc->ticket[3], tot, lds) && threadIdx.x == 0) {
    if (INIT) {
      sc->bnorm = sqrt(tot[2]);
      sc->thresh = sc->tol * (sc->bnorm > 1.0 ? sc->bnorm : 1.0);
      sc->init_res = sqrt(tot[0]);
      sc->rn = sc->rn_max = sc->init_res;
      sc->rho_prev = sc->alpha = sc->omega = 1.0;
      sc->rho = tot[0];
      sc->beta = sc->rho;
      sc->it = sc->iterations = 0;
      sc->replace = 0;
      sc->n_replace = 0;
      sc->state = SOLVE_RUNNING;
      if (sc->init_res <= sc->thresh && (sc->flags & SOLVE_TRUE_CHECK)) sc->state = SOLVE_CONVERGED;
      else if (fabs(sc->rho) < 1e-30) sc->state = SOLVE_BREAKDOWN;
      else if (sc->max_iter <= 0) sc->state = SOLVE_MAXITER;
    } else {
      const double rn = sqrt(tot[0]);
      const int kind = sc->replace;
      sc->replace = 0;
      sc->n_replace += 1;
      sc->rn = rn;
      sc->rn_max = rn;
      if (!isfinite(rn) || !isfinite(tot[1])) {
        sc->state = SOLVE_BREAKDOWN;
        return;
      }
      if (kind == 2 && rn <= sc->thresh) {
        sc->state = SOLVE_CONVERGED;
        return;
      }
      sc->rho = tot[1];
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
template <int G>
__global__ __launch_bounds__(SOLVER_THREADS) void k_bicg_final(const int* __restrict__ rowp, const int* __restrict__ col,
                                                               const double* __restrict__ val, int m,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ scale,
                                                               const double* __restrict__ y,
                                                               const double* __restrict__ xe, double* __restrict__ x,
                                                               SolverScalars* sc, double* __restrict__ partials) {
  __shared__ double lds[3 * SOLVER_THREADS / WAVE];
  const int lane = threadIdx.x % G, grp = threadIdx.x / G, gpb = SOLVER_THREADS / G;
  double res = 0.0, se = 0.0, xn = 0.0, mx = 0.0;
  for (int row = blockIdx.x * gpb + grp; row < m; row += gridDim.x * gpb) {
    double acc = 0.0;
    const int e = rowp[row + 1];
    for (int k = rowp[row] + lane; k < e; k += G) {
      const int c = col[k];
      acc = fma(val[k], scale ? scale[c] * y[c] : y[c], acc);
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
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
  }
  mx = wave_max(mx);
  if ((threadIdx.x & (WAVE - 1)) == 0 && mx > 0.0)
    atomicMax(reinterpret_cast<unsigned long long*>(&sc->max_err), (unsigned long long)__double_as_longlong(mx));
  double vv[3] = {res, se, xn}, tot[3];
  if (grid_sum_last_block<3>(vv, partials, &sc->ticket[4], tot, lds) && threadIdx.x == 0) {
    sc->final_res = sqrt(tot[0]);
    sc->rel_l2 = sqrt(tot[1] / (tot[2] > 1e-30 ? tot[2] : 1e-30));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale_matrix(const int* __restrict__ col, const double* __restrict__ val,
                                                      const double* __restrict__ scale, int64_t nnz, T* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) out[k] = (T)(scale ? val[k] * scale[col[k]] : val[k]);
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

__global__ void k_solver_setup(SolverScalars* sc, int max_iter, double tol, int flags, double delta) {
  if (threadIdx.x == 0) {
    sc->max_iter = max_iter;
    sc->tol = tol;
    sc->flags = flags;
    sc->delta = delta;
    sc->max_err = 0.0;
    sc->state = SOLVE_RUNNING;
    for (int k = 0; k < 8; ++k) sc->ticket[k] = 0u;
  }
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
constexpr int SPMV_G = 8;

unsigned solver_grid(int m) {
  const unsigned rows_per_block = SOLVER_THREADS / SPMV_G;
  unsigned g = blocks(m, rows_per_block);
  return g < 1 ? 1 : (g > SOLVER_MAX_GRID ? SOLVER_MAX_GRID : g);
}

void launch_inv_diag(const double* diag, int m, double* s, hipStream_t st) {
  if (m > 0) hipLaunchKernelGGL(k_inv_diag, dim3(blocks(m, 256)), dim3(256), 0, st, diag, m, s);
}

template <typename T>
void launch_scale_matrix(const int* col, const double* val, const double* scale, int64_t nnz, T* out, hipStream_t st) {
  if (nnz > 0) hipLaunchKernelGGL(k_scale_matrix<T>, dim3(blocks(nnz, 256)), dim3(256), 0, st, col, val, scale, nnz, out);
}
template void launch_scale_matrix<float>(const int*, const double*, const double*, int64_t, float*, hipStream_t);
template void launch_scale_matrix<double>(const int*, const double*, const double*, int64_t, double*, hipStream_t);

template <typename T>
void launch_solver_init(const SolverBufs<T>& B, const double* x0, SolverScalars* sc, int max_iter, double tol,
                        int flags, double delta, hipStream_t st) {
  const int m = B.m;
  hipLaunchKernelGGL(k_solver_setup, dim3(1), dim3(64), 0, st, sc, max_iter, tol, flags, delta);
  if (m > 0) hipLaunchKernelGGL(k_y_from_x, dim3(blocks(m, 256)), dim3(256), 0, st, x0, B.scale, m, B.y);
  hipLaunchKernelGGL((k_bicg_true_residual<T, SPMV_G, true>), dim3(solver_grid(m)), dim3(SOLVER_THREADS), 0, st,
                     B.rowp, B.col, B.val64, m, B.b, B.scale, x0 ? (const double*)B.y : nullptr, B.r, B.r0, B.p[0],
                     B.p[1], B.v[0], B.v[1], sc, B.partials);
}

template <typename T>
void launch_solver_iteration(const SolverBufs<T>& B, int parity, bool with_replace, SolverScalars* sc, hipStream_t st) {
  const int m = B.m;
  const unsigned g = solver_grid(m);
  hipLaunchKernelGGL((k_bicg_pv<T, SPMV_G>), dim3(g), dim3(SOLVER_THREADS), 0, st, B.rowp, B.col, B.val, m, B.r, B.r0,
                     B.p[parity], B.v[parity], B.p[parity ^ 1], B.v[parity ^ 1], sc, B.partials);
  hipLaunchKernelGGL((k_bicg_st<T, SPMV_G>), dim3(g), dim3(SOLVER_THREADS), 0, st, B.rowp, B.col, B.val, m, B.r,
                     B.v[parity ^ 1], B.s, B.t, sc, B.partials);
  hipLaunchKernelGGL((k_bicg_xr<T>), dim3(g), dim3(SOLVER_THREADS), 0, st, m, B.p[parity ^ 1], B.s, B.t, B.r0, B.r, B.y,
                     sc, B.partials);
  if (with_replace)
    hipLaunchKernelGGL((k_bicg_true_residual<T, SPMV_G, false>), dim3(g), dim3(SOLVER_THREADS), 0, st, B.rowp, B.col,
                       B.val64, m, B.b, B.scale, (const double*)B.y, B.r, B.r0, (T*)nullptr, (T*)nullptr,
                       (T*)nullptr, (T*)nullptr, sc, B.partials);
}

template <typename T>
void launch_solver_final(const SolverBufs<T>& B, const double* xe, double* x, SolverScalars* sc, hipStream_t st) {
  const int m = B.m;
  hipLaunchKernelGGL((k_bicg_final<SPMV_G>), dim3(solver_grid(m)), dim3(SOLVER_THREADS), 0, st, B.rowp, B.col, B.val64,
                     m, B.b, B.scale, (const double*)B.y, xe, x, sc, B.partials);
}

template void launch_solver_init<float>(const SolverBufs<float>&, const double*, SolverScalars*, int, double, int, double,
                                        hipStream_t);
template void launch_solver_init<double>(const SolverBufs<double>&, const double*, SolverScalars*, int, double, int,
                                         double, hipStream_t);
template void launch_solver_iteration<float>(const SolverBufs<float>&, int, bool, SolverScalars*, hipStream_t);
template void launch_solver_iteration<double>(const SolverBufs<double>&, int, bool, SolverScalars*, hipStream_t);
template void launch_solver_final<float>(const SolverBufs<float>&, const double*, double*, SolverScalars*, hipStream_t);
template void launch_solver_final<double>(const SolverBufs<double>&, const double*, double*, SolverScalars*,
                                          hipStream_t);

}  // namespace lsmps_hip

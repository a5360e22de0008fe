/*
 * CPU oracle for the LS-MPS per-step particle pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product path (ls-mps_amd/) never links, loads or calls it, and has no CPU
 * fallback.
 *
 * It restates, in plain single-threaded FP64 C++ (-O3 -ffp-contract=off),
 * the algorithm of the reference's src/lsmps.cpp:
 *   case generation     src/lsmps.cpp:68,72
 *   cell-list neighbors src/lsmps.cpp:16-18,26
 *   classification      src/lsmps.cpp:27
 *   WLS 9x9 operator    src/lsmps.cpp:22-25,28
 *   PPE CSR assembly    src/lsmps.cpp:29-64
 *   BiCGStab solve      src/lsmps.cpp:65-67
 *   run metrics         src/lsmps.cpp:73
 * plus the extensions that have no reference code: right-Jacobi BiCGStab,
 * per-particle WLS over an arbitrary sampled field, and the MLP
 * forward/backward used by the learned preconditioner.
 *
 * Pinning (how each function is checked, in tests/test_oracle.py):
 *   - generate/neighbors/classify/assemble/solve/wls_test: bit-exact against
 *     golden vectors written by the reference itself (tests/golden/), and
 *     against brute-force O(N^2) neighbor search and closed forms.
 *   - wls_particle: quadratic reproduction (closed-form gradient/Laplacian)
 *     and a textbook Cholesky (numpy) on the same moment matrix.
 *   - right-Jacobi BiCGStab: scipy.sparse.linalg.spsolve on small systems
 *     plus the discrete manufactured solution.
 *   - MLP: finite differences of the loss (gradient) and a numpy forward.
 */
#ifndef LSMPS_ORACLE_H
#define LSMPS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Statistics records; field meaning follows WlsStats / PpeStats in the
 * reference's include/lsmps.hpp:20-21 (size_t fields widened to uint64). */
typedef struct {
  uint64_t sampled, ok, regularized, failed, virtual_points;
  double max_gradient_error, rms_gradient_error, max_laplacian_error, rms_laplacian_error;
} orc_wls_stats;

typedef struct {
  uint64_t unknowns, nnz, dirichlet, neumann, failed_rows;
  int32_t iterations, converged;
  double initial_residual, final_residual, relative_l2_error, max_error;
  double assembly_ms, solve_ms;
  uint64_t graph_rows; /* WLS rows: rows that took the graph-Laplacian row */
} orc_ppe_stats;

typedef struct {
  uint64_t particles, pairs, truncated;
  uint64_t type_count[5];
  double neighbor_mean;
  int32_t p50, p90, p99;
  double neighbor_ms, classify_ms, wls_ms, wls_all_ms;
} orc_run_stats;

/* Case generation (src/lsmps.cpp:68,72). Writes xyz[3n] interleaved and
 * type[n] when the pointers are non-NULL; returns the particle count. */
uint64_t orc_generate_case(const char* case_name, uint64_t particles, uint32_t seed,
                           double l0, int film_layers, double* xyz, uint8_t* type);

typedef struct orc_state orc_state;
orc_state* orc_new(const double* xyz, const uint8_t* type, uint64_t n);
void orc_free(orc_state* s);

/* Cell-list neighbor search (src/lsmps.cpp:26). */
void orc_neighbors(orc_state* s, double re, int max_neighbors, uint64_t* pairs, uint64_t* truncated);
void orc_get_neighbors(const orc_state* s, int32_t* row_ptr, int32_t* col);
void orc_set_neighbors(orc_state* s, const int32_t* row_ptr, const int32_t* col);
/* Brute-force O(N^2) search with the same predicate and visiting order, for
 * pinning orc_neighbors on small inputs. */
void orc_neighbors_brute(orc_state* s, double re, int max_neighbors, uint64_t* pairs, uint64_t* truncated);

/* Classification in place (src/lsmps.cpp:27,73). */
void orc_classify(orc_state* s, double l0);
void orc_get_types(const orc_state* s, uint8_t* type);
void orc_set_types(orc_state* s, const uint8_t* type);

/* Status codes of a per-particle WLS solve. */
enum { ORC_WLS_OK = 0, ORC_WLS_REGULARIZED = 1, ORC_WLS_FAILED = 2, ORC_WLS_SKIPPED = 3, ORC_WLS_NO_WEIGHT = 4 };

/* WLS gradient and Laplacian of a field sampled at particles, for every
 * particle that is not Wall or Splash (src/lsmps.cpp:22-25,28). */
void orc_wls_all(const orc_state* s, const double* f, double re, double l0, double reg,
                 double* grad3, double* lap, uint8_t* status);
/* 9x9 moment matrix (divided by the weight sum) and RHS of one particle,
 * before Cholesky; returns the weight sum. */
double orc_wls_moments(const orc_state* s, const double* f, double re, double l0, uint64_t i,
                       double* m81, double* b9);
/* The reference's sampled quadratic test (src/lsmps.cpp:28). */
void orc_wls_test(const orc_state* s, double re, double l0, double reg, int wls_samples, orc_wls_stats* out);

/* The x* of row k under xmode 1: the splitmix64 finaliser of seed * 2^64/phi + k,
 * top 53 bits, mapped to [-1, 1). */
double orc_hashed_uniform(uint32_t seed, uint64_t k);
/* PPE assembly (src/lsmps.cpp:29-64). xmode 0: x* = x^2+y^2 (reference);
 * xmode 1: x* ~ U(-1,1) from the splitmix64 hash of (xseed, row). shift adds shift*diag to each
 * diagonal after assembly (0 = reference). */
void orc_assemble_ppe(orc_state* s, double re, int max_unknowns, int xmode, uint32_t xseed,
                      double shift, orc_ppe_stats* out);
/* WLS-Laplacian rows (no C++ reference; README.md:34, 2-D model
 * python/rain_plate_lsmps_cpu.py:204-253): same unknowns, x* and b = A x*;
 * row i holds -c_ij, c_ij = (z . q_ij) w_ij / r_ij / ws / l0, per unknown
 * neighbor in list order and the diagonal d = sum c_ij over all weighted
 * neighbors, with z = M^-1 (0,0,0,1,1,1,0,0,0) from the reference's moments
 * and ladder. A row whose WLS fails or with d <= 0 or d < sum |c_ij|
 * takes the graph-Laplacian row of orc_assemble_ppe (graph_rows). Read back
 * with orc_get_ppe. */
void orc_assemble_ppe_wls(orc_state* s, double re, double l0, double reg, int max_unknowns, int xmode,
                          uint32_t xseed, double shift, orc_ppe_stats* out);
void orc_get_ppe(const orc_state* s, int32_t* row, int32_t* col, double* val, double* b,
                 double* x_exact, int32_t* rev);
/* BiCGStab (src/lsmps.cpp:65-67). precond 0 = none (reference), 1 = right
 * Jacobi. x0 may be NULL (zero start). x[n] receives the solution. */
void orc_solve_ppe(orc_state* s, int max_iter, double tol, int precond, const double* x0,
                   orc_ppe_stats* out, double* x);

/* One full pipeline step like run_benchmark (src/lsmps.cpp:73), with stage
 * timers; wls_all_ms times the WLS over every particle for the field of the
 * reference's wls_test. */
void orc_run(orc_state* s, double l0, double re_ratio, int max_neighbors, int wls_samples,
             int max_unknowns, int max_iter, double tol, int do_solve,
             orc_run_stats* rs, orc_wls_stats* ws, orc_ppe_stats* ps);

/* Features of the learned preconditioner for every unknown of the
 * assembled PPE, reference row order, feat[16 m] (no reference
 * counterpart; definitions follow the GPU kernel's):
 *   0-4 one-hot type; 5 k/32; 6-8 mean r/re, mean (r/re)^2, min r/re over
 *   the neighbor list; 9 diag re^2/32; 10 sum of w/r^2 over unknown
 *   neighbors over the sum over all neighbors; 11 unknown neighbors / k;
 *   12 (b/diag)/s_b; 13 sum_j (a_ij/diag)(b_j/diag_j)/s_b (0 on pinned
 *   and failed rows, whose diagonal is not a sum of couplings); 14 Wall
 *   neighbors / k; 15 pinned or failed row.
 * s_b = sqrt(mean (b/diag)^2). */
void orc_features(const orc_state* s, double re, float* feat, double* rhs_scale);

/* MLP 16->64->64->2 (tanh hidden units), weights fp32 row-major
 * [W1(64x16) b1(64) W2(64x64) b2(64) W3(2x64) b3(2)]. Forward over n rows;
 * out[2n]. */
void orc_mlp_forward(const float* params, const float* feat, uint64_t n, float* out);
/* Loss 0.5/n * sum_k wk*(out_k - target_k)^2 and its gradient w.r.t. params. */
double orc_mlp_loss_grad(const float* params, const float* feat, const float* target,
                         const float* wk, uint64_t n, float* grad);

#ifdef __cplusplus
}
#endif
#endif

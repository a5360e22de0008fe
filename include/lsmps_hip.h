/*
 * C ABI of the MI355X (gfx950) LS-MPS per-step particle pipeline.
 *
 * Plain C: opaque context, raw host pointers and sizes, int status codes.
 * No exceptions, no C++ or torch types cross it. Each entry point replaces
 * one internal stage of the reference's run_benchmark (src/lsmps.cpp:73);
 * the C++ API of include/lsmps.hpp is implemented on top of it
 * (ls-mps_amd/csrc/host/lsmps_api.cpp), so the reference's callers keep
 * working unchanged.
 *
 * Data stays in HBM between calls: upload positions and types once, run the
 * stages, download only what a caller asks for. A context owns one HIP
 * stream and its device buffers; use one context per host thread.
 * Stage times are HIP-event times on the context's stream, in ms.
 */
#ifndef LSMPS_HIP_H
#define LSMPS_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSMPS_HIP_ABI_VERSION 8

enum {
  LSMPS_OK = 0,
  LSMPS_E_ARG = -1,    /* invalid argument or size */
  LSMPS_E_HIP = -2,    /* HIP runtime error (message in lsmps_last_error) */
  LSMPS_E_STATE = -3,  /* stage called before the stage it depends on */
  LSMPS_E_LIMIT = -4,  /* size beyond a documented limit (int32 indices, grid) */
  LSMPS_E_CHECK = -5,  /* a host-side self-check failed (LSMPS_CHECK_BUILD=1 at context creation:
                          after each neighbor build the cell list and neighbor lists are
                          downloaded and their invariants checked; message in lsmps_last_error) */
};

/* Particle types, numbered as the reference's ParticleType (include/lsmps.hpp:8). */
enum { LSMPS_INTERIOR = 0, LSMPS_FREE_SURFACE = 1, LSMPS_SPLASH = 2, LSMPS_NEAR_WALL = 3, LSMPS_WALL = 4 };
/* Per-particle WLS status. */
enum {
  LSMPS_WLS_OK = 0,
  LSMPS_WLS_REGULARIZED = 1, /* factorised after a diagonal shift of the ladder */
  LSMPS_WLS_FAILED = 2,      /* every rung of the ladder failed */
  LSMPS_WLS_SKIPPED = 3,     /* Wall or Splash particle */
  LSMPS_WLS_NO_WEIGHT = 4,   /* weight sum <= 0 (no usable neighbor) */
};

typedef struct lsmps_ctx lsmps_ctx;

int lsmps_abi_version(void);
/* Context on HIP device `device`. */
int lsmps_create(int device, lsmps_ctx** out);
void lsmps_destroy(lsmps_ctx* ctx);
/* Message of the last failing call on this context ("" if none). */
const char* lsmps_last_error(const lsmps_ctx* ctx);
/* Waits for all work queued on the context's stream. */
int lsmps_synchronize(lsmps_ctx* ctx);
/* Which HIP events the neighbor builds, lsmps_wls and lsmps_step record. No
 * reference counterpart (the reference times its stages with the host
 * clock): each event recorded on the stream costs it a few microseconds, so
 * a time-step loop that only needs results can switch them off.
 * LSMPS_TIMING_STAGES: the stages' ms; LSMPS_TIMING_KERNELS: their main
 * kernel's kernel_ms (the neighbor pass, the main WLS kernel). A field whose
 * events are off reads 0. Default: both. */
enum { LSMPS_TIMING_STAGES = 1, LSMPS_TIMING_KERNELS = 2 };
int lsmps_set_timing(lsmps_ctx* ctx, int32_t flags);
/* "name (arch, N CUs)" of the context's device into buf (NUL-terminated).
 * No reference counterpart: for reports written next to the reference's. */
int lsmps_device_name(lsmps_ctx* ctx, char* buf, int32_t cap);

/* Copies n particles (xyz interleaved, 3n doubles; type n bytes) into HBM.
 * Replaces the std::vector<Particle> handed to run_benchmark
 * (include/lsmps.hpp:10,29). Positions are stored SoA in FP64. */
int lsmps_upload_particles(lsmps_ctx* ctx, const double* xyz, const uint8_t* type, uint64_t n);

/* ---- cell-list neighbors: replaces neighbors() (src/lsmps.cpp:26) ------ */
typedef struct {
  uint64_t pairs;        /* sum of kept list lengths (NeighborData::pairs) */
  uint64_t truncated;    /* dropped pairs beyond max_neighbors (NeighborData::truncated) */
  double mean;           /* mean kept count (Metrics::neighbor_mean) */
  int32_t p50, p90, p99; /* quantiles as run_benchmark takes them (src/lsmps.cpp:73) */
  int32_t max_count;
  int32_t grid[3];       /* cells per axis of the dense grid */
  float ms;              /* GPU time of the whole stage */
  float kernel_ms;       /* GPU time of the neighbor pass kernel alone */
} lsmps_cnl_stats;
/* Lists are CSR (row_ptr int32[n+1], col int32[pairs]) in the reference's
 * visiting order, truncated to max_neighbors (<= 256). */
int lsmps_build_neighbors(lsmps_ctx* ctx, double re, int32_t max_neighbors, lsmps_cnl_stats* out);

/* Neighbors and classification in one pass over the candidates (the
 * classification reads exactly the offsets the neighbor test computes):
 * replaces neighbors() + the classify loop of run_benchmark
 * (src/lsmps.cpp:26-27,73). types->ms is 0 (time is in out->ms). */
typedef struct lsmps_type_stats lsmps_type_stats;
int lsmps_build_neighbors_classify(lsmps_ctx* ctx, double re, int32_t max_neighbors, double l0,
                                   lsmps_cnl_stats* out, lsmps_type_stats* types);

/* ---- classification: replaces classify() applied in run_benchmark
 * (src/lsmps.cpp:27,73). Overwrites the device types in place. ---------- */
struct lsmps_type_stats {
  uint64_t count[5]; /* per type, in LSMPS_* order */
  float ms;
};
int lsmps_classify(lsmps_ctx* ctx, double l0, lsmps_type_stats* out);

/* ---- WLS 9x9 operator: per-particle form of wls_test (src/lsmps.cpp:22-25,28) */
/* Field sampled at the particles. Quadratic form: coefficients of
 * 1, x, y, z, x^2, y^2, z^2, xy, xz, yz and its exact Laplacian (used for
 * error statistics). The reference's test field is
 * {1, 2, -3, .5, .7, -.2, .4, .3, -.1, .25} with Laplacian 1.8. */
typedef struct {
  double c[10];
  double lap_exact;
} lsmps_quad_field;
int lsmps_set_field_quadratic(lsmps_ctx* ctx, const lsmps_quad_field* f);
/* Arbitrary field values (n doubles). Error statistics are then unavailable. */
int lsmps_set_field_values(lsmps_ctx* ctx, const double* f, uint64_t n);

enum { LSMPS_WLS_FP64 = 0, LSMPS_WLS_MIXED = 1 };
typedef struct {
  double re, l0;
  double reg;        /* regularization base; the reference always uses 1e-8 */
  int32_t precision; /* LSMPS_WLS_FP64: bit-exact with the reference;
                        LSMPS_WLS_MIXED: FP32 moments, FP64 for weak rows */
  /* MIXED: a row whose FP32 Cholesky has a pivot at or below this fraction
   * of its diagonal entry is computed in FP64 instead (the reference's bits).
   * The FP32 error of a row grows like 1 / (smallest pivot ratio); 0 selects
   * LSMPS_WLS_FP32_PIVOT_DEFAULT. Must lie in [0, 1]. */
  double fp32_pivot_min;
} lsmps_wls_params;
#define LSMPS_WLS_FP32_PIVOT_DEFAULT 0.2
typedef struct {
  uint64_t ok, regularized, failed, skipped;
  uint64_t fp64_redo; /* MIXED: rows computed in FP64 */
  float ms;
  float kernel_ms;    /* GPU time of the main WLS kernel alone (no field pass, no FP64 rows of MIXED) */
} lsmps_wls_result;
/* Gradient and Laplacian for every particle except Wall and Splash. */
int lsmps_wls(lsmps_ctx* ctx, const lsmps_wls_params* p, lsmps_wls_result* out);

/* ---- one particle step: the body of run_benchmark's timed sequence
 * (neighbors, classify, wls_test; src/lsmps.cpp:26-28,73) as one call.
 * Same results as lsmps_build_neighbors_classify(re, max_neighbors, l0)
 * followed by lsmps_wls(wls), bit for bit, but both stages are queued before
 * the host waits once for their readbacks, so the GPU runs them back to
 * back. cnl->ms and wls->ms time each stage as the separate calls do.
 * wls.re must not exceed re. */
typedef struct {
  double re;             /* neighbor radius (lists and classification) */
  int32_t max_neighbors;
  double l0;             /* classification spacing */
  lsmps_wls_params wls;
} lsmps_step_params;
int lsmps_step(lsmps_ctx* ctx, const lsmps_step_params* p, lsmps_cnl_stats* cnl, lsmps_type_stats* types,
               lsmps_wls_result* wls);

/* lsmps_wls without allocating device memory. No counterpart in the
 * reference: for callers that cannot allocate (work captured into a HIP
 * graph, tight serving loops, harnesses that forbid hipMalloc). Runs the
 * same kernels as lsmps_wls on the context's current neighbor lists and
 * writes the gradient (3 per particle), Laplacian and status into the
 * caller's device arrays, in particle order (particles appended since the
 * last neighbor build are not written). The stage's work list, block
 * partials, counters and the mixed kernel's row flags and list live in a
 * fixed device array of about 19 MB per GPU, held until the process ends; a context with more than
 * LSMPS_NOALLOC_MAX_PARTICLES particles is refused with LSMPS_E_LIMIT.
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * (other threads, other streams) race. The context's own WLS results
 * (lsmps_download_wls, lsmps_wls_sampled_stats) are not changed. */
#define LSMPS_NOALLOC_MAX_PARTICLES 2097152
int lsmps_wls_noalloc(lsmps_ctx* ctx, const lsmps_wls_params* p, double* grad3_dev, double* lap_dev,
                      uint8_t* status_dev, lsmps_wls_result* out);

/* The reference's sampled statistics (WlsStats, include/lsmps.hpp:20):
 * stride max(1, n / wls_samples) from particle 0 until wls_samples are
 * counted, Wall/Splash skipped (src/lsmps.cpp:28). Needs a quadratic field. */
typedef struct {
  uint64_t sampled, ok, regularized, failed, virtual_points;
  double max_gradient_error, rms_gradient_error, max_laplacian_error, rms_laplacian_error;
} lsmps_wls_stats;
int lsmps_wls_sampled_stats(lsmps_ctx* ctx, int32_t wls_samples, lsmps_wls_stats* out);

/* ---- PPE assembly: replaces assemble_ppe (src/lsmps.cpp:29-64) --------- */
enum { LSMPS_XSTAR_QUADRATIC = 0, LSMPS_XSTAR_RANDOM = 1 };
typedef struct {
  double re;
  int32_t max_unknowns; /* Config::ppe_max_unknowns */
  int32_t xmode;        /* x* = x^2 + y^2 (reference) or U(-1,1) from a counter hash of (xseed, row) */
  uint32_t xseed;
  double shift;         /* diagonal *= 1 + shift after assembly; 0 = reference */
} lsmps_ppe_params;
typedef struct {
  uint64_t unknowns, nnz, dirichlet, neumann, failed_rows; /* PpeStats fields */
  float ms;
  float kernel_ms;     /* GPU time of the fill kernel alone */
  uint64_t padded_nnz; /* stored entries of the sliced-ELL matrix (nnz + padding) */
  uint64_t graph_rows; /* lsmps_assemble_ppe_wls: rows that took the graph-Laplacian row */
} lsmps_ppe_result;
int lsmps_assemble_ppe(lsmps_ctx* ctx, const lsmps_ppe_params* p, lsmps_ppe_result* out);
/* WLS-Laplacian rows in place of the graph Laplacian: the reference's
 * stated next PPE (README.md:34), built as its 2-D model builds it
 * (python/rain_plate_lsmps_cpu.py:204-253), in 3-D with the reference's
 * weight, basis and Cholesky ladder (src/lsmps.cpp:22-25, l0 and reg as in
 * wls_test). Same unknowns, layout, x* and b = A x* as lsmps_assemble_ppe.
 * Row i is minus the WLS Laplacian (the graph Laplacian's sign): -c_ij for
 * each unknown neighbor in list order, c_ij = (z . q_ij) w_ij / r_ij / ws /
 * l0 with z = M^-1 (0,0,0,1,1,1,0,0,0), then the diagonal d = sum_j c_ij
 * over all weighted neighbors (non-unknown neighbors are p = 0). A row
 * whose WLS fails, or with d <= 0 or d < sum |c_ij| over the unknown
 * neighbors (not weakly diagonally dominant, as one-sided stencils are),
 * takes the graph-Laplacian row of lsmps_assemble_ppe instead
 * (graph_rows); row 0 is not pinned. */
int lsmps_assemble_ppe_wls(lsmps_ctx* ctx, const lsmps_ppe_params* p, double l0, double reg,
                           lsmps_ppe_result* out);

/* ---- BiCGStab: replaces solve_ppe (src/lsmps.cpp:65-67) ---------------- */
enum {
  LSMPS_SOLVE_PLAIN = 0,   /* FP64, unpreconditioned, the reference's stopping rule */
  LSMPS_SOLVE_JACOBI = 1,  /* FP64, right Jacobi, true-residual check */
  LSMPS_SOLVE_MIXED = 2,   /* FP32 matrix, FP64 Krylov vectors, FP64 reliable
                              updates and restarts, right Jacobi; FP64 JACOBI
                              continues from its iterate when it does not
                              converge or a reliable update stalls (fallback) */
  LSMPS_SOLVE_LEARNED = 3, /* MIXED with the MLP diagonal and initial guess
                              (lsmps_mlp_set_params first); plain MIXED when
                              the guess raises the residual */
};
typedef struct {
  int32_t method;
  int32_t max_iter;
  double tol;          /* ||r|| <= tol * max(1, ||b||) */
  int32_t use_x0;      /* start from the vector given by lsmps_upload_x0 */
  int32_t chunk;       /* iterations per captured HIP graph (0 = default 16) */
} lsmps_solve_params;
typedef struct {
  int32_t iterations, converged, state, replacements;
  int32_t fallback;     /* MIXED, LEARNED: 1 when FP64 JACOBI finished the solve (LEARNED: also
                           when the guess raised the residual and plain MIXED ran) */
  double initial_residual, final_residual, relative_l2_error, max_error;
  float ms;             /* GPU time from setup to the final residual */
  int32_t breakdown_restarts; /* JACOBI, MIXED, LEARNED: restarts from the true residual
                                 after a rho, omega or r0.v breakdown (at most 32) */
} lsmps_solve_result;
int lsmps_solve_ppe(lsmps_ctx* ctx, const lsmps_solve_params* p, lsmps_solve_result* out);
int lsmps_upload_x0(lsmps_ctx* ctx, const double* x0, uint64_t n);

/* ---- time stepping (no counterpart in the reference's run_benchmark; the
 * step timeline of its Python models, python/rain_plate_lsmps_cpu.py:311-381:
 * injection, motion, outflow removal, neighbor rebuilds on a cadence) ------ */
/* Neighbor lists holding every pair within re + skin (cell size re + skin),
 * so they stay valid while particles move less than skin / 2; WLS, PPE
 * assembly and features take only pairs within re (their weights vanish
 * beyond it), and classification counts only neighbors within re. With
 * skin = 0 this is lsmps_build_neighbors[_classify]. max_neighbors bounds
 * the skin lists. */
int lsmps_build_neighbors_skin(lsmps_ctx* ctx, double re, double skin, int32_t max_neighbors, double l0,
                               int32_t classify, lsmps_cnl_stats* out, lsmps_type_stats* types);
/* Velocities of all particles, 3 per particle, caller order. */
int lsmps_set_velocities(lsmps_ctx* ctx, const double* uvw, uint64_t n);
/* Appends particles (xyz 3 per particle, types, velocities or NULL) after
 * the existing ones; they join the stages at the next neighbor build.
 * Velocities are required when the context has them (LSMPS_E_STATE
 * otherwise), and refused when the existing particles have none. A
 * field given as values (lsmps_set_field_values) has none for them, so it
 * must be set again; a quadratic field carries over. */
int lsmps_append_particles(lsmps_ctx* ctx, const double* xyz, const uint8_t* type, const double* uvw,
                           uint64_t count);
/* x += u dt for every particle (and the sorted copies the neighbor lists
 * refer to); WLS, PPE and solution become stale. */
int lsmps_advance(lsmps_ctx* ctx, double dt);
/* Removes the particles outside [lo, hi) (per axis), keeping the order of
 * the rest; the neighbor lists must be rebuilt afterwards. */
int lsmps_remove_outside(lsmps_ctx* ctx, const double* lo, const double* hi, uint64_t* removed);

/* ---- learned preconditioner (no reference counterpart) -----------------
 * A per-row MLP 16 -> 64 -> 64 -> 2 (tanh hidden units; BF16 weight
 * matrices, FP32 biases and accumulation) on features of the assembled PPE.
 * Output 0 times s_b (the RMS of b/diag) is the initial guess; output 1,
 * clamped to [-0.5, 0.5], is the log of a correction of the Jacobi
 * diagonal. LSMPS_SOLVE_LEARNED uses both inside the mixed solver and falls
 * back to the plain mixed method when the guess has a larger true residual
 * than the zero start; convergence is always decided on the true FP64
 * residual. Parameters are FP32, laid out as W1 (64x16, one row per output
 * unit), b1 (64), W2 (64x64), b2 (64), W3 (2x64), b3 (2). */
#define LSMPS_MLP_FEATURES 16
#define LSMPS_MLP_OUTPUTS 2
#define LSMPS_MLP_PARAMS 5378
/* Features of every unknown (reference row order, m x 16) and s_b; either
 * pointer may be NULL. Needs an assembled PPE. */
int lsmps_features(lsmps_ctx* ctx, float* feat, double* rhs_scale);
int lsmps_mlp_set_params(lsmps_ctx* ctx, const float* params, uint64_t n);
int lsmps_mlp_get_params(lsmps_ctx* ctx, float* params, uint64_t n);
/* The network on n host rows of features; out[2n]. */
int lsmps_mlp_predict(lsmps_ctx* ctx, const float* feat, uint64_t n, float* out);
/* Loss 0.5/n * sum_rows sum_k wk[k] (out_k - target_k)^2 and its gradient
 * (LSMPS_MLP_PARAMS floats), computed by the training kernels. */
int lsmps_mlp_loss_grad(lsmps_ctx* ctx, const float* feat, const float* target, uint64_t n, const float* wk,
                        float* grad, double* loss);
/* Appends every stride-th row of the current solved PPE to the context's
 * training set: features, and targets x / s_b and the log of the clamped
 * least-squares ratio sum x_j q_j / sum q_j^2 (q = b / diag) over the
 * row's columns. */
int lsmps_train_collect(lsmps_ctx* ctx, int32_t stride, uint64_t* total);
int lsmps_train_clear(lsmps_ctx* ctx);
typedef struct {
  int32_t steps, batch; /* Adam steps, rows per minibatch (drawn with replacement) */
  float lr, wk0, wk1;   /* learning rate, loss weights of the two outputs */
  uint32_t seed;
} lsmps_train_params;
typedef struct {
  double loss_first, loss_last; /* minibatch loss at the first and the last step */
  uint64_t samples;
  float ms;
} lsmps_train_result;
int lsmps_mlp_train(lsmps_ctx* ctx, const lsmps_train_params* p, lsmps_train_result* out);

/* ---- downloads (tests, parity, callers that want the arrays) ----------- */
int lsmps_download_types(lsmps_ctx* ctx, uint8_t* type);
/* Current positions, caller order, 3 per particle. */
int lsmps_download_positions(lsmps_ctx* ctx, double* xyz);
int lsmps_download_neighbors(lsmps_ctx* ctx, int32_t* row_ptr, int32_t* col);
int lsmps_download_wls(lsmps_ctx* ctx, double* grad3, double* lap, uint8_t* status);
/* Any pointer may be NULL. rev[k] = particle index of unknown k. */
int lsmps_download_ppe(lsmps_ctx* ctx, int32_t* row, int32_t* col, double* val, double* b, double* x_exact,
                       int32_t* rev);
int lsmps_download_solution(lsmps_ctx* ctx, double* x);

/* ---- synthetic CSR replay (python/workload_calibration.py:236-264) ------
 * The calibration tool's customer-profile proxy: a CSR neighbor graph
 * (csr_row_ptr int64[n+1], csr_col_idx int32[nnz]), particle_type u8[n] in
 * the tool's numbering (0 interior, 1 near wall, 2 free surface, 3 splash,
 * 4 wall/other; mapped to ParticleType) and wls_ill_conditioned bool[n]
 * (config/customer_car_rain_profile.example.json:5-18 describes the
 * profile). The graph has no positions: every edge gets a synthetic offset
 * of length [0.35, 0.95) re from a seeded hash, and flagged rows get
 * near-planar offsets, so their WLS takes the FP64 redo and the
 * regularisation ladder. Rows are sorted ascending on load; self-edges are
 * kept in the table and skipped by the stages. No counterpart in the
 * reference's C++: the tool writes these files for GPU kernel replay. */
typedef struct {
  uint64_t rows, edges;
  int32_t max_row;     /* longest row: the width of the neighbor table (<= 256) */
  int32_t pad;
  uint64_t self_edges;
  uint64_t types[5];   /* rows per ParticleType after the mapping */
  uint64_t flagged;    /* rows flagged wls_ill_conditioned */
  float ms;
} lsmps_replay_info;
/* Validates the arrays on the host (row_ptr[0] = 0, non-decreasing, nnz and
 * rows within int32, rows <= 256, columns in range, types <= 4) and replaces
 * the context's particles and neighbor table by the graph. */
int lsmps_replay_load(lsmps_ctx* ctx, const int64_t* csr_row_ptr, const int32_t* csr_col_idx,
                      const uint8_t* particle_type, const uint8_t* wls_ill_conditioned, uint64_t n,
                      uint32_t seed, lsmps_replay_info* out);
/* The WLS stage on the replay graph: FP32 moments and Cholesky per row on
 * the synthetic offsets with the reference's manufactured quadratic, FP64
 * redo with the ladder of rows whose FP32 pivots are not trusted. Results
 * download with lsmps_download_wls (gradient (2, -3, 0.5), Laplacian 1.8
 * where the fit is exact). */
int lsmps_replay_wls(lsmps_ctx* ctx, double re, double l0, double reg, lsmps_wls_result* out);
/* The graph-Laplacian PPE (src/lsmps.cpp:29-64 rules: unknowns Interior and
 * NearWall, row 0 pinned, diagonal last) on the replay graph with the
 * synthetic symmetric weight (0.5 + u) / l0^2 per edge, u in [0, 1) from
 * the splitmix64 finaliser of (lo << 32 | hi) ^ (seed * 0x9E3779B97F4A7C15);
 * x* ~ U(-1, 1) as LSMPS_XSTAR_RANDOM. Solve with lsmps_solve_ppe; tol = 0
 * runs exactly max_iter iterations (a fixed per-step count). */
int lsmps_replay_assemble(lsmps_ctx* ctx, int32_t max_unknowns, uint32_t xseed, double l0, lsmps_ppe_result* out);

#ifdef __cplusplus
}
#endif
#endif

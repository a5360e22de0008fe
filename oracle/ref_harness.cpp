// Exporter for golden vectors: compiles the reference's own src/lsmps.cpp
// (included from where it lies, path given by -DLSMPS_REF_SRC) in this
// translation unit so its internal functions can be called through a C ABI.
// Output goes to oracle/_ref/ only; nothing here is part of the product, and
// the reference source itself is never copied into this repository.
//
// Built by oracle/Makefile (target `ref`) only where the reference tree is
// present. Used by tests/golden/make_golden.py to write fixtures, and by
// bench.py's cpu_baseline leg (kind "reference") when the built library is
// present.
#ifndef LSMPS_REF_SRC
#error "define LSMPS_REF_SRC as the path of the reference src/lsmps.cpp"
#endif
#include LSMPS_REF_SRC

#include <chrono>
#include <cstring>

namespace {
struct RefState {
  lsmps::Config cfg;
  std::vector<lsmps::Particle> ps;
  lsmps::NeighborData nd;
  lsmps::CSR A;
  lsmps::PpeStats ppe;
};
RefState g;

void load_cfg(const char* case_name, uint64_t particles, uint32_t seed, double l0, double re_ratio,
              int film_layers, int maxn, int wls_samples, int max_unknowns, int max_iter, double tol) {
  g.cfg = lsmps::Config{};
  g.cfg.case_name = case_name;
  g.cfg.particles = particles;
  g.cfg.seed = seed;
  g.cfg.l0 = l0;
  g.cfg.re_ratio = re_ratio;
  g.cfg.film_layers = film_layers;
  g.cfg.max_neighbors = maxn;
  g.cfg.wls_samples = wls_samples;
  g.cfg.ppe_max_unknowns = max_unknowns;
  g.cfg.ppe_max_iter = max_iter;
  g.cfg.ppe_tol = tol;
}
}  // namespace

extern "C" {

// Sets the config and generates the case with the reference's generator.
uint64_t ref_setup(const char* case_name, uint64_t particles, uint32_t seed, double l0, double re_ratio,
                   int film_layers, int maxn, int wls_samples, int max_unknowns, int max_iter, double tol) {
  load_cfg(case_name, particles, seed, l0, re_ratio, film_layers, maxn, wls_samples, max_unknowns, max_iter, tol);
  g.ps = lsmps::generate_case(g.cfg);
  return g.ps.size();
}

// Sets the config and takes caller positions/types instead of generating.
void ref_setup_particles(const double* xyz, const uint8_t* type, uint64_t n, double l0, double re_ratio, int maxn,
                         int wls_samples, int max_unknowns, int max_iter, double tol) {
  load_cfg("custom", n, 0, l0, re_ratio, 4, maxn, wls_samples, max_unknowns, max_iter, tol);
  g.ps.assign(n, lsmps::Particle{});
  for (uint64_t i = 0; i < n; ++i) {
    g.ps[i].r = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    g.ps[i].type = (lsmps::ParticleType)type[i];
  }
}

void ref_get_particles(double* xyz, uint8_t* type) {
  for (size_t i = 0; i < g.ps.size(); ++i) {
    xyz[3 * i] = g.ps[i].r.x;
    xyz[3 * i + 1] = g.ps[i].r.y;
    xyz[3 * i + 2] = g.ps[i].r.z;
    type[i] = (uint8_t)g.ps[i].type;
  }
}

uint64_t ref_neighbors(uint64_t* truncated) {
  g.nd = lsmps::neighbors(g.ps, g.cfg.re_ratio * g.cfg.l0, g.cfg.max_neighbors);
  *truncated = g.nd.truncated;
  return g.nd.pairs;
}

void ref_get_neighbors(int32_t* row_ptr, int32_t* col) {
  int64_t at = 0;
  row_ptr[0] = 0;
  for (size_t i = 0; i < g.nd.list.size(); ++i) {
    for (int j : g.nd.list[i]) col[at++] = j;
    row_ptr[i + 1] = (int32_t)at;
  }
}

void ref_classify() {
  for (size_t i = 0; i < g.ps.size(); ++i) g.ps[i].type = lsmps::classify(g.ps[i], g.nd.list[i], g.ps, g.cfg.l0);
}

void ref_wls_test(uint64_t* counts5, double* err4) {
  lsmps::WlsStats st = lsmps::wls_test(g.cfg, g.ps, g.nd);
  counts5[0] = st.sampled;
  counts5[1] = st.ok;
  counts5[2] = st.regularized;
  counts5[3] = st.failed;
  counts5[4] = st.virtual_points;
  err4[0] = st.max_gradient_error;
  err4[1] = st.rms_gradient_error;
  err4[2] = st.max_laplacian_error;
  err4[3] = st.rms_laplacian_error;
}

// Per-particle gradient/Laplacian of the reference's test field computed
// with the reference's own weight/basis/cholesky/solve; status 0 ok,
// 1 regularized, 2 factorisation failed, 3 skipped type, 4 zero weight sum.
void ref_wls_particles(const uint64_t* idx, uint64_t count, double* grad3, double* lap, uint8_t* status) {
  const double re = g.cfg.re_ratio * g.cfg.l0, l0 = g.cfg.l0;
  auto f = [](lsmps::Vec3 r) {
    return 1 + 2 * r.x - 3 * r.y + .5 * r.z + .7 * r.x * r.x - .2 * r.y * r.y + .4 * r.z * r.z + .3 * r.x * r.y -
           .1 * r.x * r.z + .25 * r.y * r.z;
  };
  for (uint64_t s = 0; s < count; ++s) {
    const size_t i = idx[s];
    grad3[3 * s] = grad3[3 * s + 1] = grad3[3 * s + 2] = lap[s] = 0;
    if (g.ps[i].type == lsmps::ParticleType::Wall || g.ps[i].type == lsmps::ParticleType::Splash) {
      status[s] = 3;
      continue;
    }
    std::array<double, lsmps::K * lsmps::K> M{};
    std::array<double, lsmps::K> b{};
    double ws = 0, fi = f(g.ps[i].r);
    for (int j : g.nd.list[i]) {
      lsmps::Vec3 d = g.ps[j].r - g.ps[i].r;
      double r = lsmps::norm(d), w = lsmps::weight(r, re);
      if (!w) continue;
      auto q = lsmps::basis(d, r, l0);
      double df = f(g.ps[j].r) - fi;
      ws += w;
      for (int a = 0; a < lsmps::K; ++a) {
        b[a] += w * q[a] * df / r;
        for (int c = 0; c < lsmps::K; ++c) M[a * lsmps::K + c] += w * q[a] * q[c];
      }
    }
    if (ws <= 0) {
      status[s] = 4;
      continue;
    }
    for (double& x : M) x /= ws;
    for (double& x : b) x /= ws;
    bool reg = false;
    if (!lsmps::cholesky(M, g.cfg.wls_regularization, reg)) {
      status[s] = 2;
      continue;
    }
    auto x = lsmps::solve(M, b);
    grad3[3 * s] = x[0];
    grad3[3 * s + 1] = x[1];
    grad3[3 * s + 2] = x[2];
    lap[s] = (x[3] + x[4] + x[5]) / l0;
    status[s] = reg ? 1 : 0;
  }
}

// Assembles with the reference's assemble_ppe; returns nnz.
uint64_t ref_assemble(uint64_t* unknowns, uint64_t* failed_rows, uint64_t* neumann, uint64_t* dirichlet) {
  lsmps::PpeStats st;
  g.A = lsmps::assemble_ppe(g.cfg, g.ps, g.nd, st);
  *unknowns = st.unknowns;
  *failed_rows = st.failed_rows;
  *neumann = st.neumann;
  *dirichlet = st.dirichlet;
  return st.nnz;
}

void ref_get_ppe(int32_t* row, int32_t* col, double* val, double* b, double* x_exact) {
  std::copy(g.A.row.begin(), g.A.row.end(), row);
  std::copy(g.A.col.begin(), g.A.col.end(), col);
  std::copy(g.A.val.begin(), g.A.val.end(), val);
  std::copy(g.A.b.begin(), g.A.b.end(), b);
  std::copy(g.A.x_exact.begin(), g.A.x_exact.end(), x_exact);
}

// Runs the reference's solve_ppe (which re-assembles internally).
void ref_solve(int32_t* iterations, int32_t* converged, double* out5) {
  lsmps::PpeStats st = lsmps::solve_ppe(g.cfg, g.ps, g.nd);
  *iterations = st.iterations;
  *converged = st.converged ? 1 : 0;
  out5[0] = st.initial_residual;
  out5[1] = st.final_residual;
  out5[2] = st.relative_l2_error;
  out5[3] = st.max_error;
  out5[4] = st.solve_ms;
}

// Whole public-API step (load_config -> generate_case -> run_benchmark ->
// write_json/write_csv), timed; returns run_benchmark wall ms.
double ref_run_config(const char* cfg_path, const char* json_out, const char* csv_out) {
  lsmps::Config c = lsmps::load_config(cfg_path);
  auto ps = lsmps::generate_case(c);
  auto t0 = std::chrono::steady_clock::now();
  lsmps::Metrics m = lsmps::run_benchmark(c, ps);
  double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (json_out && *json_out) lsmps::write_json(json_out, c, m);
  if (csv_out && *csv_out) lsmps::write_csv(csv_out, c, m);
  return ms;
}

// Timed stages on the loaded particles (for the CPU baseline): neighbors,
// classify, sampled WLS, assembly (ms in out4).
void ref_time_stages(double* out4) {
  using C = std::chrono::steady_clock;
  auto ms = [](C::time_point a) { return std::chrono::duration<double, std::milli>(C::now() - a).count(); };
  auto t0 = C::now();
  g.nd = lsmps::neighbors(g.ps, g.cfg.re_ratio * g.cfg.l0, g.cfg.max_neighbors);
  out4[0] = ms(t0);
  t0 = C::now();
  for (size_t i = 0; i < g.ps.size(); ++i) g.ps[i].type = lsmps::classify(g.ps[i], g.nd.list[i], g.ps, g.cfg.l0);
  out4[1] = ms(t0);
  t0 = C::now();
  (void)lsmps::wls_test(g.cfg, g.ps, g.nd);
  out4[2] = ms(t0);
  lsmps::PpeStats st;
  t0 = C::now();
  g.A = lsmps::assemble_ppe(g.cfg, g.ps, g.nd, st);
  out4[3] = ms(t0);
}

}  // extern "C"

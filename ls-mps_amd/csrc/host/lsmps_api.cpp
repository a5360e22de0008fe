// The lsmps.hpp API (reference include/lsmps.hpp:27-32) on top of the
// gfx950 C ABI (lsmps_hip.h).
//
// load_config, write_json, write_csv and type_name keep the reference's
// behaviour and text formats (src/lsmps.cpp:70-71,74-75). run_benchmark
// runs every stage of the reference's step (src/lsmps.cpp:73) on the GPU:
// neighbors, classification (types written back into the caller's
// particles, as the reference does), the WLS test (computed for every
// particle; the reference's sampled statistics are taken on its subset),
// PPE assembly and BiCGStab. Stage times are GPU event times.
#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsmps.hpp"
#include "lsmps_hip.h"
#include "lsmps_host.h"

namespace lsmps {
namespace detail {
std::vector<Particle> make_case(const Config& c);
}

namespace {

std::string strip(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// One device context per host thread, created on first use.
struct ThreadCtx {
  lsmps_ctx* ctx = nullptr;
  ~ThreadCtx() {
    if (ctx) lsmps_destroy(ctx);
  }
};

lsmps_ctx* thread_ctx() {
  thread_local ThreadCtx holder;
  if (!holder.ctx) {
    const int rc = lsmps_create(0, &holder.ctx);
    if (rc != LSMPS_OK) throw std::runtime_error("lsmps: no usable HIP device (lsmps_create failed)");
  }
  return holder.ctx;
}

// What the GPU did in the last run_benchmark on this thread.
struct GpuReport {
  bool valid = false;
  std::string device;
  uint64_t n = 0, m = 0, nnz = 0, padded_nnz = 0, fp64_redo = 0;
  double k_mean = 0.0;
  float cnl_ms = 0, cnl_kernel_ms = 0, wls_ms = 0, wls_kernel_ms = 0, asm_ms = 0, asm_kernel_ms = 0, solve_ms = 0;
  int iterations = 0;
};
GpuReport& gpu_report() {
  thread_local GpuReport r;
  return r;
}

void check(lsmps_ctx* ctx, int rc, const char* stage) {
  if (rc != LSMPS_OK) throw std::runtime_error(std::string("lsmps ") + stage + ": " + lsmps_last_error(ctx));
}

}  // namespace

const char* type_name(ParticleType t) {
  switch (t) {
    case ParticleType::Interior: return "interior";
    case ParticleType::FreeSurface: return "free_surface";
    case ParticleType::Splash: return "splash";
    case ParticleType::NearWall: return "near_wall";
    case ParticleType::Wall: return "wall";
  }
  return "unknown";
}

// key = value lines, '#' starts a comment, surrounding blanks trimmed,
// unknown keys ignored; the same 13 keys as the reference are read.
Config load_config(const std::string& path) {
  Config c;
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open config: " + path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string key = strip(line.substr(0, eq)), val = strip(line.substr(eq + 1));
    if (key == "case") c.case_name = val;
    else if (key == "particles") c.particles = std::stoull(val);
    else if (key == "seed") c.seed = (uint32_t)std::stoul(val);
    else if (key == "l0") c.l0 = std::stod(val);
    else if (key == "re_ratio") c.re_ratio = std::stod(val);
    else if (key == "film_layers") c.film_layers = std::stoi(val);
    else if (key == "max_neighbors") c.max_neighbors = std::stoi(val);
    else if (key == "wls_samples") c.wls_samples = std::stoi(val);
    else if (key == "ppe_max_unknowns") c.ppe_max_unknowns = std::stoi(val);
    else if (key == "ppe_max_iter") c.ppe_max_iter = std::stoi(val);
    else if (key == "ppe_tol") c.ppe_tol = std::stod(val);
    else if (key == "boundary_constraint_weight") c.boundary_constraint_weight = std::stod(val);
    else if (key == "virtual_weight") c.virtual_weight = std::stod(val);
  }
  return c;
}

std::vector<Particle> generate_case(const Config& cfg) { return detail::make_case(cfg); }

Metrics run_benchmark(const Config& c, std::vector<Particle>& ps) {
  Metrics m;
  m.particles = ps.size();
  lsmps_ctx* ctx = thread_ctx();
  const size_t n = ps.size();
  std::vector<double> xyz(3 * n);
  std::vector<uint8_t> type(n);
  for (size_t i = 0; i < n; ++i) {
    xyz[3 * i] = ps[i].r.x;
    xyz[3 * i + 1] = ps[i].r.y;
    xyz[3 * i + 2] = ps[i].r.z;
    type[i] = (uint8_t)ps[i].type;
  }
  check(ctx, lsmps_upload_particles(ctx, xyz.data(), type.data(), n), "upload");
  const double re = c.re_ratio * c.l0;

  // Neighbors and classification share one pass over the candidates.
  lsmps_cnl_stats cs;
  lsmps_type_stats ts;
  check(ctx, lsmps_build_neighbors_classify(ctx, re, c.max_neighbors, c.l0, &cs, &ts), "neighbors");
  m.neighbor_pairs = cs.pairs;
  m.truncated_particles = cs.truncated;
  m.neighbor_ms = cs.ms;
  if (n) {
    m.neighbor_mean = cs.mean;
    m.neighbor_p50 = cs.p50;
    m.neighbor_p90 = cs.p90;
    m.neighbor_p99 = cs.p99;
  }

  m.classify_ms = ts.ms;
  check(ctx, lsmps_download_types(ctx, type.data()), "download types");
  for (size_t i = 0; i < n; ++i) ps[i].type = (ParticleType)type[i];
  m.interior = ts.count[0];
  m.free_surface = ts.count[1];
  m.splash = ts.count[2];
  m.near_wall = ts.count[3];
  m.wall = ts.count[4];

  // The reference's manufactured quadratic (src/lsmps.cpp:28).
  const lsmps_quad_field field = {{1, 2, -3, .5, .7, -.2, .4, .3, -.1, .25}, 1.8};
  check(ctx, lsmps_set_field_quadratic(ctx, &field), "field");
  lsmps_wls_params wp = {re, c.l0, c.wls_regularization, LSMPS_WLS_FP64, 0.0};
  lsmps_wls_result wr;
  check(ctx, lsmps_wls(ctx, &wp, &wr), "wls");
  lsmps_wls_stats ws;
  check(ctx, lsmps_wls_sampled_stats(ctx, c.wls_samples, &ws), "wls stats");
  m.wls_ms = wr.ms;
  m.wls.sampled = ws.sampled;
  m.wls.ok = ws.ok;
  m.wls.regularized = ws.regularized;
  m.wls.failed = ws.failed;
  m.wls.virtual_points = ws.virtual_points;
  m.wls.max_gradient_error = ws.max_gradient_error;
  m.wls.rms_gradient_error = ws.rms_gradient_error;
  m.wls.max_laplacian_error = ws.max_laplacian_error;
  m.wls.rms_laplacian_error = ws.rms_laplacian_error;

  lsmps_ppe_params pp = {re, c.ppe_max_unknowns, LSMPS_XSTAR_QUADRATIC, 0u, 0.0};
  lsmps_ppe_result pr;
  check(ctx, lsmps_assemble_ppe(ctx, &pp, &pr), "assemble");
  m.ppe.unknowns = pr.unknowns;
  m.ppe.nnz = pr.nnz;
  m.ppe.dirichlet = pr.dirichlet;
  m.ppe.neumann = pr.neumann;
  m.ppe.failed_rows = pr.failed_rows;
  m.ppe.assembly_ms = pr.ms;
  if (pr.unknowns) {
    lsmps_solve_params sp = {LSMPS_SOLVE_PLAIN, c.ppe_max_iter, c.ppe_tol, 0, 0};
    lsmps_solve_result sr;
    check(ctx, lsmps_solve_ppe(ctx, &sp, &sr), "solve");
    m.ppe.iterations = sr.iterations;
    m.ppe.converged = sr.converged != 0;
    m.ppe.initial_residual = sr.initial_residual;
    m.ppe.final_residual = sr.final_residual;
    m.ppe.relative_l2_error = sr.relative_l2_error;
    m.ppe.max_error = sr.max_error;
    m.ppe.solve_ms = sr.ms;
  }
  GpuReport& g = gpu_report();
  g = GpuReport{};
  char name[256] = {0};
  if (lsmps_device_name(ctx, name, (int32_t)sizeof(name)) == LSMPS_OK) g.device = name;
  g.n = n;
  g.m = pr.unknowns;
  g.nnz = pr.nnz;
  g.padded_nnz = pr.padded_nnz;
  g.fp64_redo = wr.fp64_redo;
  g.k_mean = cs.mean;
  g.cnl_ms = cs.ms;
  g.cnl_kernel_ms = cs.kernel_ms;
  g.wls_ms = wr.ms;
  g.wls_kernel_ms = wr.kernel_ms;
  g.asm_ms = pr.ms;
  g.asm_kernel_ms = pr.kernel_ms;
  g.solve_ms = (float)m.ppe.solve_ms;
  g.iterations = m.ppe.iterations;
  g.valid = true;
  return m;
}

// JSON text in the reference's layout (src/lsmps.cpp:74): fixed notation,
// 8 decimals, the same keys in the same order; open failures are silent.
void write_json(const std::string& path, const Config& c, const Metrics& m) {
  std::ofstream o(path);
  o << std::fixed << std::setprecision(8);
  o << "{\n  \"case\": \"" << c.case_name << "\",\n  \"particles\": " << m.particles << ",\n";
  o << "  \"types\": {\"interior\": " << m.interior << ", \"free_surface\": " << m.free_surface
    << ", \"splash\": " << m.splash << ", \"near_wall\": " << m.near_wall << ", \"wall\": " << m.wall << "},\n";
  o << "  \"neighbors\": {\"pairs\": " << m.neighbor_pairs << ", \"mean\": " << m.neighbor_mean
    << ", \"p50\": " << m.neighbor_p50 << ", \"p90\": " << m.neighbor_p90 << ", \"p99\": " << m.neighbor_p99
    << ", \"truncated\": " << m.truncated_particles << "},\n";
  o << "  \"wls\": {\"sampled\": " << m.wls.sampled << ", \"regularized\": " << m.wls.regularized
    << ", \"failed\": " << m.wls.failed << ", \"virtual_points\": " << m.wls.virtual_points
    << ", \"gradient_rms_error\": " << m.wls.rms_gradient_error
    << ", \"laplacian_rms_error\": " << m.wls.rms_laplacian_error << "},\n";
  o << "  \"ppe\": {\"unknowns\": " << m.ppe.unknowns << ", \"nnz\": " << m.ppe.nnz
    << ", \"dirichlet\": " << m.ppe.dirichlet << ", \"neumann\": " << m.ppe.neumann
    << ", \"failed_rows\": " << m.ppe.failed_rows << ", \"iterations\": " << m.ppe.iterations
    << ", \"converged\": " << (m.ppe.converged ? "true" : "false") << ", \"final_residual\": " << m.ppe.final_residual
    << ", \"relative_l2_error\": " << m.ppe.relative_l2_error << ", \"assembly_ms\": " << m.ppe.assembly_ms
    << ", \"solve_ms\": " << m.ppe.solve_ms << "}\n}\n";
}

// One header line and one row, default stream formatting (src/lsmps.cpp:75).
void write_csv(const std::string& path, const Config& c, const Metrics& m) {
  std::ofstream o(path);
  o << "case,particles,interior,free_surface,splash,near_wall,wall,neighbor_mean,p50,p90,p99,wls_failed,"
       "wls_virtual,ppe_unknowns,ppe_nnz,ppe_iterations,ppe_converged,ppe_residual,ppe_rel_l2,"
       "ppe_assembly_ms,ppe_solve_ms\n";
  const char s = ',';
  o << c.case_name << s << m.particles << s << m.interior << s << m.free_surface << s << m.splash << s
    << m.near_wall << s << m.wall << s << m.neighbor_mean << s << m.neighbor_p50 << s << m.neighbor_p90 << s
    << m.neighbor_p99 << s << m.wls.failed << s << m.wls.virtual_points << s << m.ppe.unknowns << s << m.ppe.nnz
    << s << m.ppe.iterations << s << m.ppe.converged << s << m.ppe.final_residual << s << m.ppe.relative_l2_error
    << s << m.ppe.assembly_ms << s << m.ppe.solve_ms << '\n';
}

}  // namespace lsmps

namespace {
void stage_json(std::ostream& o, const char* name, double ms, double kernel_ms, double bytes, bool last) {
  const double t = kernel_ms > 0 ? kernel_ms : ms;
  const double gbs = t > 0 ? bytes / (t * 1e-3) / 1e9 : 0.0;
  o << "    \"" << name << "\": {\"ms\": " << ms << ", \"kernel_ms\": " << kernel_ms
    << ", \"algorithmic_bytes\": " << bytes << ", \"GBs\": " << gbs << ", \"hbm_fraction\": " << gbs / 8000.0
    << "}" << (last ? "\n" : ",\n");
}
}  // namespace

// Algorithmic bytes per stage (inputs read once, outputs written once;
// k = mean neighbor count): neighbors + classification N (24 + 1 + 4k + 4
// + 1); WLS N (24 + 8 + 4 + 4k + 1 + 24 + 8 + 1); assembly N + m (4k + 4 +
// 24 + 4) + 12 nnz + 24 m; one BiCGStab iteration 24 nnz + 160 m (two FP64
// SpMVs over the stored matrix and the vector updates).
extern "C" int lsmps_write_gpu_report_c(const char* path) {
  const lsmps::GpuReport& g = lsmps::gpu_report();
  if (!g.valid || !path) return 1;
  std::ofstream o(path);
  if (!o) return 1;
  const double n = (double)g.n, m = (double)g.m, k = g.k_mean;
  o << std::setprecision(6);
  o << "{\n  \"device\": \"" << g.device << "\",\n  \"hbm_peak_GBs\": 8000,\n";
  o << "  \"particles\": " << g.n << ", \"unknowns\": " << g.m << ", \"nnz\": " << g.nnz
    << ", \"stored_entries\": " << g.padded_nnz << ", \"neighbor_mean\": " << k
    << ", \"wls_fp64_redo\": " << g.fp64_redo << ",\n";
  o << "  \"stages\": {\n";
  stage_json(o, "neighbors_classify", g.cnl_ms, g.cnl_kernel_ms, n * (24 + 1 + 4 * k + 4 + 1), false);
  stage_json(o, "wls", g.wls_ms, g.wls_kernel_ms, n * (24 + 8 + 4 + 4 * k + 1 + 24 + 8 + 1), false);
  stage_json(o, "ppe_assembly", g.asm_ms, g.asm_kernel_ms, n + m * (4 * k + 4 + 24 + 4) + 12.0 * g.nnz + 24 * m,
                    false);
  const double it = g.iterations > 0 ? g.iterations : 1;
  stage_json(o, "ppe_solve_per_iteration", g.solve_ms / it, 0.0, 24.0 * g.nnz + 160 * m, true);
  o << "  },\n  \"solve_iterations\": " << g.iterations << "\n}\n";
  return o ? 0 : 1;
}

extern "C" uint64_t lsmps_generate_case_c(const char* case_name, uint64_t particles, uint32_t seed, double l0,
                                          int film_layers, double* xyz, uint8_t* type) {
  lsmps::Config c;
  c.case_name = case_name ? case_name : "";
  c.particles = particles;
  c.seed = seed;
  c.l0 = l0;
  c.film_layers = film_layers;
  const std::vector<lsmps::Particle> p = lsmps::generate_case(c);
  for (size_t i = 0; i < p.size(); ++i) {
    if (xyz) {
      xyz[3 * i] = p[i].r.x;
      xyz[3 * i + 1] = p[i].r.y;
      xyz[3 * i + 2] = p[i].r.z;
    }
    if (type) type[i] = (uint8_t)p[i].type;
  }
  return p.size();
}

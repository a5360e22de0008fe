// lsmps-bench: the command line of the reference (src/main.cpp:5), same
// arguments, outputs and exit codes, running the step on the GPU.
//   lsmps-bench [config (default config/wall_film.cfg)] [out dir (default results)]
// Writes <out>/<case>.json and .csv, prints the summary, exits 0, 2 when the
// PPE did not converge or WLS failures exceed a fifth of the samples, 1 on
// error.
#include <filesystem>
#include <iomanip>
#include <iostream>

#include "lsmps.hpp"
#include "lsmps_host.h"

namespace {

int run(const std::string& cfg_path, const std::string& out_dir) {
  const lsmps::Config c = lsmps::load_config(cfg_path);
  std::vector<lsmps::Particle> ps = lsmps::generate_case(c);
  const lsmps::Metrics m = lsmps::run_benchmark(c, ps);
  std::filesystem::create_directories(out_dir);
  const std::string base = out_dir + "/" + c.case_name;
  lsmps::write_json(base + ".json", c, m);
  lsmps::write_csv(base + ".csv", c, m);
  lsmps_write_gpu_report_c((base + ".gpu.json").c_str());  // GPU stage report beside the reference's files
  std::ostream& o = std::cout;
  o << "LS-MPS industrial benchmark v0.2\n";
  o << "case=" << c.case_name << " particles=" << m.particles << "\n";
  o << "types interior=" << m.interior << " surface=" << m.free_surface << " splash=" << m.splash
    << " near_wall=" << m.near_wall << " wall=" << m.wall << "\n";
  o << std::fixed << std::setprecision(2);
  o << "neighbors mean=" << m.neighbor_mean << " p50=" << m.neighbor_p50 << " p90=" << m.neighbor_p90
    << " p99=" << m.neighbor_p99 << "\n";
  o << "WLS sampled=" << m.wls.sampled << " regularized=" << m.wls.regularized << " failed=" << m.wls.failed
    << " virtual=" << m.wls.virtual_points << " grad_rms=" << std::scientific << m.wls.rms_gradient_error
    << " lap_rms=" << m.wls.rms_laplacian_error << "\n";
  o << "PPE unknowns=" << m.ppe.unknowns << " nnz=" << m.ppe.nnz << " iterations=" << m.ppe.iterations
    << " converged=" << m.ppe.converged << " residual=" << m.ppe.final_residual
    << " rel_l2=" << m.ppe.relative_l2_error << "\n";
  o << "PPE timing(ms) assembly=" << std::fixed << m.ppe.assembly_ms << " solve=" << m.ppe.solve_ms << "\n";
  o << "output=" << base << ".{json,csv}\n";
  o.flush();
  const bool bad = !m.ppe.converged || m.wls.failed > m.wls.sampled / 5;
  return bad ? 2 : 0;
}

}  // namespace

extern "C" int lsmps_run_config_c(const char* cfg_path, const char* out_dir) {
  try {
    return run(cfg_path ? cfg_path : "config/wall_film.cfg", out_dir ? out_dir : "results");
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << '\n';
    return 1;
  }
}

#ifdef LSMPS_CLI_MAIN
int main(int argc, char** argv) {
  return lsmps_run_config_c(argc > 1 ? argv[1] : "config/wall_film.cfg", argc > 2 ? argv[2] : "results");
}
#endif

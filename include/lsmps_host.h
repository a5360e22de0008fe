/*
 * Host-side helpers of liblsmps_hip.so with a C ABI, for callers that do
 * not link C++ (ctypes, cgo-style bindings). They wrap the C++ API of
 * lsmps.hpp; none of them touches the GPU except lsmps_run_config_c, which
 * runs run_benchmark.
 */
#ifndef LSMPS_HOST_H
#define LSMPS_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* generate_case (reference src/lsmps.cpp:72) for a case name and the
 * Config fields it reads. Cases: the reference's wall_film, narrow_gap,
 * rain_injection and its lattice default (e.g. bulk_uniform), plus
 * bulk_lattice (50x50x20-style jittered block), rain_plate (inclined plate
 * with film and droplets) and industrial_mix (bulk + tilted wall film +
 * Poisson-disk splash). Writes xyz[3n] and type[n] when non-NULL; returns n. */
uint64_t lsmps_generate_case_c(const char* case_name, uint64_t particles, uint32_t seed, double l0,
                               int film_layers, double* xyz, uint8_t* type);

/* load_config -> generate_case -> run_benchmark -> write_json/write_csv,
 * like the CLI. Returns the CLI exit code (0, 1 on error, 2 when the PPE
 * did not converge or WLS failures exceed a fifth of the samples). */
int lsmps_run_config_c(const char* cfg_path, const char* out_dir);

/* Writes, as JSON, what the GPU did in the last run_benchmark of the
 * calling thread: device, and per stage the GPU time, the time of its main
 * kernel, the stage's algorithmic bytes (inputs read once, outputs written
 * once), the rate they imply and its fraction of the 8 TB/s HBM peak.
 * The reference's write_json / write_csv files are left as they are; this
 * report sits next to them (<case>.gpu.json from the command line).
 * Returns 0, or 1 when no run happened on this thread or the file cannot
 * be written. */
int lsmps_write_gpu_report_c(const char* path);

#ifdef __cplusplus
}
#endif
#endif

// Launchers of the gfx950 kernels of the LS-MPS step (one .hip file per
// stage). All take raw device pointers and a stream; no allocation happens
// inside a launcher.
#pragma once
#include <climits>
#include <cstdint>

#include "common.hpp"

namespace lsmps_hip {

// Particle types, numbered as the reference's ParticleType (include/lsmps.hpp:8).
enum : uint8_t { T_INTERIOR = 0, T_SURFACE = 1, T_SPLASH = 2, T_NEARWALL = 3, T_WALL = 4 };
// Per-particle WLS status.
enum : uint8_t { WLS_OK = 0, WLS_REGULARIZED = 1, WLS_FAILED = 2, WLS_SKIPPED = 3, WLS_NO_WEIGHT = 4 };
// counts[] layout of the WLS kernels: [status 0..4], [5] rows redone in FP64.
constexpr int WLS_REDO_SLOT = 5;

struct QuadField {
  double c[10];  // c0, cx, cy, cz, cxx, cyy, czz, cxy, cxz, cyz
  double lap_exact;
};

// scan.hip
size_t scan_scratch_elems(int64_t n);
void launch_exclusive_scan(const int* in, int64_t n, int* out, int64_t* scratch, hipStream_t st);

// cnl.hip
void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, hipStream_t st);
void launch_fill_int(int* p, int64_t n, int v, hipStream_t st);
void launch_cell_sort(const int* cx, const int* cy, const int* cz, int n, const int* box_host, int gx, int gy,
                      int gz, int* cell, int* slot, int* count, int* start, int64_t* scan_scratch, int* perm,
                      const double* x, const double* y, const double* z, double* sx, double* sy, double* sz,
                      hipStream_t st);
void launch_nbr_count(const double* sx, const double* sy, const double* sz, const int* perm, const int* start,
                      const int* cell, int n, int gx, int gy, int gz, double re2, int maxn, int* count,
                      unsigned long long* trunc_total, int* hist, hipStream_t st);
void launch_nbr_fill(const double* sx, const double* sy, const double* sz, const int* perm, const int* start,
                     const int* cell, int n, int gx, int gy, int gz, double re2, int maxn, const int* row_ptr, int* col,
                     hipStream_t st);
void launch_classify(const double* x, const double* y, const double* z, const int* row_ptr, const int* col, int n,
                     double zlim, uint8_t* type, unsigned long long* tcount, hipStream_t st);

// wls.hip
void launch_field_quadratic(const double* x, const double* y, const double* z, int n, const QuadField& q, double* f,
                            hipStream_t st);
void launch_wls_exact(const double* x, const double* y, const double* z, const double* f, const uint8_t* type,
                      const int* row_ptr, const int* col, int n, double re, double l0, double reg, double* grad,
                      double* lap, uint8_t* status, unsigned long long* counts, hipStream_t st);
void launch_wls_mixed(const double* x, const double* y, const double* z, const double* f, const uint8_t* type,
                      const int* row_ptr, const int* col, int n, double re, double l0, double reg, double* grad,
                      double* lap, uint8_t* status, int* redo, int* nredo, unsigned long long* counts,
                      hipStream_t st);
void launch_wls_sample_errors(const double* x, const double* y, const double* z, const uint8_t* type,
                              const double* grad, const double* lap, const uint8_t* status, int64_t stride, int ns,
                              const QuadField& q, double* ge, double* le, uint8_t* st_out, uint8_t* ty_out,
                              hipStream_t st);

// ppe.hip
void launch_unknown_flags(const uint8_t* type, int n, int* flag, hipStream_t st);
void launch_unknown_map(const int* flag, const int* pos, int n, int cap, int* map, int* rev, hipStream_t st);
void launch_ppe_row_count(const double* x, const double* y, const double* z, const int* nrow, const int* ncol,
                          const int* map, const int* rev, int m, double re, int* cnt, hipStream_t st);
void launch_ppe_row_fill(const double* x, const double* y, const double* z, const uint8_t* type, const int* nrow,
                         const int* ncol, const int* map, const int* rev, int m, double re, double shift, int xmode,
                         const int* prow, int* pcol, double* pval, double* xe, double* diag,
                         unsigned long long* counters, hipStream_t st);
void launch_ppe_rhs(const int* prow, const int* pcol, const double* pval, const double* xe, int m, double* b,
                    hipStream_t st);


}  // namespace lsmps_hip

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

// cnl.hip: buffers of the cell list (device pointers).
struct CnlBufs {
  int gx = 0, gy = 0, gz = 0;
  const double *x = nullptr, *y = nullptr, *z = nullptr;  // original order
  int *cx = nullptr, *cy = nullptr, *cz = nullptr, *cell = nullptr, *slot = nullptr;
  int *ccount = nullptr, *cstart = nullptr;               // per cell
  int64_t* scan_scratch = nullptr;
  int *perm = nullptr, *iperm = nullptr, *scell = nullptr;  // sorted slot <-> particle
  double4* spos = nullptr;                                  // sorted {x, y, z, field}
  int *nbr = nullptr, *cnt = nullptr, *cnt_orig = nullptr;  // ELL table
};
void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, hipStream_t st);
void launch_fill_int(int* p, int64_t n, int v, hipStream_t st);
void launch_cell_sort(const CnlBufs& B, int n, const int* box_host, hipStream_t st);
void launch_cnl_pass(const CnlBufs& B, int n, double re2, int maxn, bool classify, double zlim, uint8_t* type,
                     uint8_t* type_s, unsigned long long* trunc_total, int* hist, unsigned long long* tcount,
                     hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
void launch_classify(const CnlBufs& B, int n, int maxn, double zlim, uint8_t* type, uint8_t* type_s,
                     unsigned long long* tcount, hipStream_t st);
void launch_ell_to_csr(const CnlBufs& B, int n, int maxn, const int* row_ptr, int* col, hipStream_t st);
void launch_gather_types(const int* perm, const uint8_t* type, int n, uint8_t* type_s, hipStream_t st);

// wls.hip (sorted {x, y, z, field} records, types, neighbor rows and
// outputs in sorted order)
void launch_field_quadratic(double4* spos, int n, const QuadField& q, hipStream_t st);
void launch_gather_field(const int* perm, const double* v, int n, double4* spos, hipStream_t st);
void launch_wls_exact(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, double* grad, double* lap, uint8_t* status,
                      unsigned long long* counts, hipStream_t st, hipEvent_t ev_begin = nullptr,
                      hipEvent_t ev_end = nullptr);
void launch_wls_mixed(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, double* grad, double* lap, uint8_t* status, int* redo,
                      int* nredo, unsigned long long* counts, hipStream_t st, hipEvent_t ev_begin = nullptr,
                      hipEvent_t ev_end = nullptr);
void launch_wls_sample_errors(const double4* spos, const uint8_t* type, const double* grad, const double* lap,
                              const uint8_t* status, const int* iperm, int64_t stride, int ns, const QuadField& q,
                              double* ge, double* le, uint8_t* st_out, uint8_t* ty_out, hipStream_t st);
void launch_wls_unsort(const int* perm, const double* gs, const double* ls, const uint8_t* ss, int n, double* g,
                       double* l, uint8_t* st, hipStream_t stream);

// ppe.hip (rows in unknown order; positions and neighbor rows in sorted order)
void launch_unknown_flags(const uint8_t* type, int n, int* flag, hipStream_t st);
void launch_unknown_map(const int* flag, const int* pos, int n, int cap, const int* iperm, int* smap, int* rev,
                        hipStream_t st);
void launch_ppe_row_count(const double4* spos, const int* nbr, const int* ncnt, int maxn, const int* smap,
                          const int* rev, const int* iperm, int m, double re, int* cnt, hipStream_t st);
void launch_ppe_row_fill(const double4* spos, const uint8_t* type, const int* nbr, const int* ncnt, int maxn,
                         const int* smap, const int* rev, const int* iperm, int m, double re, double shift, int xmode,
                         uint32_t xseed, const int* prow, int* pcol, double* pval, double* xe, double* diag,
                         unsigned long long* counters, hipStream_t st);
void launch_ppe_rhs(const int* prow, const int* pcol, const double* pval, const double* xe, int m, double* b,
                    hipStream_t st);

}  // namespace lsmps_hip

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
// The quadratic at (a, b, c), in one fixed order of operations (the FP64
// oracle evaluates the same expression).
__device__ __forceinline__ double quad_eval(const QuadField& q, double a, double b, double c) {
  return q.c[0] + q.c[1] * a + q.c[2] * b + q.c[3] * c + q.c[4] * a * a + q.c[5] * b * b + q.c[6] * c * c +
         q.c[7] * a * b + q.c[8] * a * c + q.c[9] * b * c;
}
// Field the cell build writes into the sorted records' fourth word.
enum : int { FIELD_NONE = 0, FIELD_QUADRATIC = 1, FIELD_VALUES = 2 };

// scan.hip
size_t scan_scratch_elems(int64_t n);
void launch_exclusive_scan(const int* in, int64_t n, int* out, int64_t* scratch, hipStream_t st);
// Adds per-block count partials (block_count_partials) into out[slot[k]].
struct CountSlots {
  int k;
  int slot[8];
};
void launch_sum_counts(const unsigned* partials, int nblk, CountSlots slots, unsigned long long* out,
                       hipStream_t st);

// cnl.hip: buffers of the cell list (device pointers).
// Records allocated past n in sposf: the neighbor pass loads a whole block
// of candidates without masking the ones past the end of a run.
constexpr int CNL_PAD = 64;
// Blocks of the neighbor-count histogram; the hist buffer passed to
// launch_cnl_pass holds maxn + 1 bins plus (maxn + 6) * CNL_STATS_BLOCKS
// partial counts.
constexpr int CNL_STATS_BLOCKS = 1024;
// Grid-boundary flags of a particle's cell: a neighbor cell exists below /
// above it along x, y, z.
constexpr int CNL_XLO = 1, CNL_XHI = 2, CNL_YLO = 4, CNL_YHI = 8, CNL_ZLO = 16, CNL_ZHI = 32;
struct CnlBufs {
  int gx = 0, gy = 0, gz = 0;
  const double *x = nullptr, *y = nullptr, *z = nullptr;  // original order
  int *cx = nullptr, *cy = nullptr, *cz = nullptr, *cell = nullptr, *slot = nullptr;
  int* skey = nullptr;                                      // cell ids in sorted order
  int *ccount = nullptr, *cstart = nullptr;                 // per cell (cstart: ncells + 1)
  int64_t* scan_scratch = nullptr;
  int *perm = nullptr, *iperm = nullptr;                    // sorted slot <-> particle
  int* sflag = nullptr;                                     // sorted CNL_* grid-boundary flags of the cell
  const uint8_t* type = nullptr;                            // caller order
  uint8_t* type_s = nullptr;                                // sorted order
  double4* spos = nullptr;                                  // sorted {x, y, z, field}
  float4* sposf = nullptr;  // sorted {FP32 offsets from (ox, oy, oz), cell id as int bits}, n + CNL_PAD
  double ox = 0.0, oy = 0.0, oz = 0.0;                      // grid origin
  double extent = 0.0;  // span of the particles' cells (FP32 error bound); 0: the grid's own
  int *nbr = nullptr, *cnt = nullptr, *cnt_orig = nullptr;  // ELL table
};
// Also type[i] = type_s_old[iperm_old[i]] for i < n_unsort (the previous
// build's sorted types back in caller order).
void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, const int* iperm_old, const uint8_t* type_s_old, int n_unsort,
                        uint8_t* type, hipStream_t st);
void launch_fill_int(int* p, int64_t n, int v, hipStream_t st);
// A grid known before the build (the last build's box plus a margin): the
// cell coordinates, the box and the previous build's type unsort are made in
// the key pass instead of launch_cell_coords.
struct CellKeyDirect {
  double rl = 0.0;                       // cell size
  const int* iperm_old = nullptr;        // previous build's slots, for the type unsort
  const uint8_t* type_s_old = nullptr;
  int n_unsort = 0;
  uint8_t* type = nullptr;               // caller-order types (written for i < n_unsort)
  // The same particles as the last build (which left them in sorted order):
  // its sorted positions and slot -> particle map, read in that order. Null:
  // caller order. spos_old must not alias the build's output spos.
  const double4* spos_old = nullptr;
  const int* perm_old = nullptr;
};
// Stable sort by cell and the sorted copies: perm, skey, cstart, spos,
// sposf, sflag, iperm. Uses cx and cy as scratch (cx, cy, cz are consumed
// first). Zeroes the 16 build counters. fmode (FIELD_*): the field written
// into spos.w, from q or from values fv in caller order. box_host: the
// grid's first cell per axis; box[6] is set when a particle lies outside
// the grid (B.gx, B.gy, B.gz). With direct, box[0..5] is also computed here.
void launch_cell_sort(const CnlBufs& B, int n, const int* box_host, int fmode, const QuadField& q, const double* fv,
                      int* box, const CellKeyDirect* direct, unsigned long long* counters, hipStream_t st);
// Neighbor lists and classification work on sorted-order types only; the
// original-order arrays are produced on demand by the unsort launchers.
// re2: squared list radius; re2c: squared radius of the neighbors that
// classification counts (equal unless the lists carry a skin).
void launch_cnl_pass(const CnlBufs& B, int n, double re2, double re2c, int maxn, bool classify, double zlim,
                     uint8_t* type_s, unsigned long long* trunc_total, int* hist, unsigned long long* tcount,
                     hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
void launch_classify(const CnlBufs& B, int n, int maxn, double zlim, uint8_t* type_s, unsigned long long* tcount,
                     hipStream_t st);
void launch_unsort_u8(const int* iperm, const uint8_t* src, int n, uint8_t* dst, hipStream_t st);
void launch_unsort_int(const int* iperm, const int* src, int n, int* dst, hipStream_t st);
void launch_ell_to_csr(const CnlBufs& B, int n, int maxn, const int* row_ptr, int* col, hipStream_t st);

// wls.hip (sorted {x, y, z, field} records, types, neighbor rows and
// outputs in sorted order)
void launch_field_quadratic(double4* spos, int n, const QuadField& q, hipStream_t st);
void launch_gather_field(const int* perm, const double* v, int n, double4* spos, hipStream_t st);
// bpart: per-block count partials, 5 * ceil(n / 64) (exact) or
// 2 * ceil(n / 64) (mixed) words.
// FP64: rows whose first factorisation fails take the reference's ladder
// in k_wls_redo through the work list redo / nredo (*nredo zeroed first).
// Results go to grad / lap / status at the sorted slot, or at out_perm[slot]
// (the particle index) when out_perm is given.
void launch_wls_exact(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, double* grad, double* lap, uint8_t* status, int* redo,
                      int* nredo, unsigned long long* counts, unsigned* bpart, hipStream_t st,
                      hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr, const int* out_perm = nullptr);
// Work buffers of the mixed WLS's FP64 rows: n flags (sorted order), the
// exclusive scan of the per-64-row counts (ceil(n / 64) + 1 ints, the last
// the list length), its scan scratch (scan_scratch_elems(ceil(n / 64))) and
// the list (n ints).
struct WlsMixedWork {
  uint8_t* pend;
  int* poff;
  int64_t* scan_scratch;
  int* list;
};
// tau: a row whose FP32 Cholesky pivot is not above tau times its diagonal
// entry is computed in FP64 (k_wls_pending).
void launch_wls_mixed(const double4* spos, const uint8_t* type, const int* nbr, const int* cnt, int maxn, int n,
                      double re, double l0, double reg, float tau, const WlsMixedWork& w, double* grad, double* lap,
                      uint8_t* status, int* redo, int* nredo, unsigned long long* counts, unsigned* bpart,
                      hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                      const int* out_perm = nullptr);
void launch_wls_sample_errors(const double4* spos, const uint8_t* type, const double* grad, const double* lap,
                              const uint8_t* status, const int* iperm, int64_t stride, int ns, const QuadField& q,
                              double* ge, double* le, uint8_t* st_out, uint8_t* ty_out, hipStream_t st);
void launch_wls_unsort(const int* perm, const double* gs, const double* ls, const uint8_t* ss, int n, double* g,
                       double* l, uint8_t* st, hipStream_t stream);

// ppe.hip. Rows u are the unknowns in sorted-slot order; rr is the
// reference's row numbering (unknowns in particle index order).
int ppe_slices(int m);
void launch_unknown_flags(const uint8_t* type_s, const int* iperm, int n, int* flag, hipStream_t st);
void launch_unknown_slots(const int* perm, const int* flag, const int* pos, int n, int cap, int* fs, hipStream_t st);
void launch_unknown_maps(const int* perm, const int* fs, const int* upos, const int* pos, int n, int* uid, int* uslot,
                         int* urr, int* rru, const double4* spos, int xmode, uint32_t xseed, double* xs,
                         hipStream_t st);
// pin0: row 0 is the identity (graph-Laplacian rows), so its width is 1.
void launch_ppe_count(const int* uslot, const int* urr, const int* uid, const int* nbr, const int* ncnt, int maxn,
                      int m, bool pin0, int* wsz, hipStream_t st);
// replay: edges of a replay graph with synthetic weights (replay.hip).
void launch_ppe_fill(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr, const int* uid,
                     const double* xs, const int* nbr, const int* ncnt, int maxn, int m, double re, double shift,
                     const int* sptr, int* col, double* val, int* rlen, double* diag, double* xe,
                     double* b, unsigned long long* counters, unsigned* bpart, hipStream_t st,
                     hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr, bool replay = false,
                     uint32_t wseed = 0, double l0 = 0.0);
// WLS-Laplacian rows into the same sliced-ELL layout; counters as
// launch_ppe_fill, plus [3] rows that took the graph-Laplacian row. bpart:
// 4 * ceil(ceil(m / 64) / 2) words.
void launch_ppe_wls_fill(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr,
                         const int* uid, const int* nbr, const int* ncnt, int maxn, int m, double re, double l0,
                         double reg, double shift, int xmode, uint32_t xseed, const int* sptr, int* col, double* val,
                         int* rlen, double* diag, double* xe, double* b, unsigned long long* counters,
                         unsigned* bpart, hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
void launch_rr_lengths(const int* rru, const int* rlen, int m, int* cnt_rr, hipStream_t st);
void launch_sell_to_csr(const int* rru, const int* urr, const int* sptr, const int* scol, const double* sval,
                        const int* rlen, const int* prow, int m, int* pcol, double* pval, hipStream_t st);
void launch_to_rr(const int* rru, const double* src, int m, double* dst, hipStream_t st);
void launch_rr_particle(const int* rru, const int* uslot, const int* perm, int m, int* rev, hipStream_t st);
void launch_from_rr(const int* urr, const double* src, int m, double* dst, hipStream_t st);

// replay.hip: the synthetic CSR replay (calibration tool's .npy graph).
// type_in uses the tool's numbering; unknown values count in *bad_types.
void launch_replay_load(const int64_t* row_ptr, const int* col, const uint8_t* type_in, int n, int maxn,
                        uint32_t seed, int* nbr, int* cnt, uint8_t* type_s, double4* rec, unsigned* bad_types,
                        hipStream_t st);
// bpart: 3 * ceil(n / 128) words; counts[] as the WLS stage.
void launch_replay_wls(const int* nbr, const int* cnt, int maxn, const uint8_t* type, const uint8_t* bad,
                       const double4* rec, int n, double re, double l0, double reg, uint32_t seed, double* grad,
                       double* lap, uint8_t* status, int* redo, int* nredo, unsigned long long* counts,
                       unsigned* bpart, hipStream_t st, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);

}  // namespace lsmps_hip

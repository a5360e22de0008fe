// Cell-list neighbor search and particle classification on gfx950.
//
// Follows the semantics of the reference's neighbors() and classify()
// (src/lsmps.cpp:26-27) with a GPU data layout:
//  * a dense grid of cells of size re over the particles' bounding box,
//    linearised x-fastest, filled by a counting sort; inside a cell the
//    particles stay in ascending index order;
//  * every particle array is re-ordered into that cell-major order
//    ("sorted slots"), so all later kernels read neighbors from contiguous
//    runs of HBM that stay in L2;
//  * the neighbor table is ELL-shaped: row s (sorted slot) holds up to
//    max_neighbors sorted-slot indices at nbr[s * max_neighbors], with its
//    length in cnt[s]; one pass writes it (no count/fill double pass);
//    CSR in original indexing is built only when a caller downloads it.
//
// Bit-exactness: cell index (int)floor(x/re) with IEEE division, strict
// test ((dx*dx + dy*dy) + dz*dz) < re*re with no contraction (this file is
// built with -ffp-contract=off), and for every particle the candidates are
// visited cell z-major, then y, then x, each cell's members in ascending
// index, so truncation at max_neighbors keeps the reference's subset and
// each list keeps the reference's order. Because cells are x-fastest, the
// three x-neighbours of a cell are one contiguous run: 9 runs per particle.
// The classification (14 direction sectors of the kept neighbors, then the
// Wall / Splash / NearWall / FreeSurface / Interior rules) is fused into the
// same pass: it needs exactly the offsets the pass has just computed.
// Candidates are screened in FP32 first; only those whose FP32 distance is
// too close to re to decide take the FP64 test, so lists are unchanged.
#include <algorithm>
#include <cmath>

#include "kernels.hpp"

namespace lsmps_hip {

// box[0..2] min and box[3..5] max cell coordinates; box[6] is set when a
// particle falls outside the grid a build assumed (k_cell_key).
__global__ void k_init_box(int* box) {
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
  else if (threadIdx.x == 6) box[6] = 0;
}

// Cell coordinates of every particle and their bounding box: grid-stride,
// block reduction, one atomic per block and axis. For i < n_unsort it also
// brings the previous build's sorted types back to caller order
// (type[i] = type_s_old[iperm_old[i]]) before they are sorted again.
__global__ __launch_bounds__(256) void k_cell_coords(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int n, double re,
                                                     int* __restrict__ cx, int* __restrict__ cy,
                                                     int* __restrict__ cz, int* __restrict__ box,
                                                     const int* __restrict__ iperm_old,
                                                     const uint8_t* __restrict__ type_s_old, int n_unsort,
                                                     uint8_t* __restrict__ type) {
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  auto one = [&](double xv, double yv, double zv, int i) {
    const int a = (int)floor(xv / re), b = (int)floor(yv / re), c = (int)floor(zv / re);
    cx[i] = a;
    cy[i] = b;
    cz[i] = c;
    lo[0] = min(lo[0], a);
    hi[0] = max(hi[0], a);
    lo[1] = min(lo[1], b);
    hi[1] = max(hi[1], b);
    lo[2] = min(lo[2], c);
    hi[2] = max(hi[2], c);
    if (i < n_unsort) type[i] = type_s_old[iperm_old[i]];
  };
  // Four particles per thread and step, their twelve loads issued together.
  const int stride = gridDim.x * blockDim.x;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    double xv[4], yv[4], zv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[i + u * stride];
      yv[u] = y[i + u * stride];
      zv[u] = z[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) one(xv[u], yv[u], zv[u], i + u * stride);
  }
  for (; i < n; i += stride) one(x[i], y[i], z[i], i);
  __shared__ int red[6][256 / WAVE];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = wave_min(lo[k]);
    hi[k] = wave_max(hi[k]);
    if (lane == 0) {
      red[k][wid] = lo[k];
      red[3 + k][wid] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int v = red[threadIdx.x][0];
    for (int w = 1; w < 256 / WAVE; ++w)
      v = threadIdx.x < 3 ? min(v, red[threadIdx.x][w]) : max(v, red[threadIdx.x][w]);
    if (threadIdx.x < 3) atomicMin(&box[threadIdx.x], v);
    else atomicMax(&box[threadIdx.x], v);
  }
}

// Four items per thread: one 16-byte store when p is 16-byte aligned and the
// four are whole, else item by item.
__global__ void k_fill_int(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i + 4 <= n && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
    reinterpret_cast<int4*>(p + i)[0] = make_int4(v, v, v, v);
  } else {
    for (int64_t k = i; k < n && k < i + 4; ++k) p[k] = v;
  }
}

// Zeroes the cell counts (16 bytes per lane where aligned, as k_fill_int)
// and the 16 build counters, and resets the box unless box is null: one
// launch where three went at the start of a build.
__global__ void k_build_init(int* __restrict__ count, int64_t ncells, int* __restrict__ box,
                             unsigned long long* __restrict__ counters) {
  const int64_t i = 4 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (i + 4 <= ncells && (reinterpret_cast<uintptr_t>(count) & 15u) == 0) {
    reinterpret_cast<int4*>(count + i)[0] = make_int4(0, 0, 0, 0);
  } else {
    for (int64_t k = i; k < ncells && k < i + 4; ++k) count[k] = 0;
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < 16) counters[threadIdx.x] = 0ull;
    if (box) {
      if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
      else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
      else if (threadIdx.x == 6) box[6] = 0;
    }
  }
}

// Counting sort by cell, stable by particle index, in four launches:
//  k_cell_key    linear cell id (x fastest) and a slot inside the cell from
//                an atomic on the cell's count. Inputs often come in
//                lattice order, so consecutive particles share a cell: a run
//                of equal cells among consecutive lanes takes its slots with
//                one atomic from its first lane (about half the atomics).
//                With the grid known in advance (the last build's box plus a
//                margin), k_cell_key_direct computes the cell coordinates
//                and the box in the same pass (no k_cell_coords). When the
//                particles are those of the last build, it walks them in
//                that build's sorted order (SORTED): positions come as
//                32-byte records in order and the lanes of a wavefront
//                mostly share a cell, so a few atomics take all the slots.
//  (scan)        exclusive scan of the counts: the cells' first slots.
//  k_scatter     particle and cell id to start + slot (arrival order).
//  k_rank_gather each slot's rank among its cell's particle indices, so
//                perm holds every cell's particles in ascending index order,
//                and the sorted copies of the particle at its final slot.
__global__ __launch_bounds__(256) void k_cell_key(const int* __restrict__ cx, const int* __restrict__ cy,
                                                  const int* __restrict__ cz, int n, int bx, int by, int bz,
                                                  int gx, int gy, int gz, int* __restrict__ cell,
                                                  int* __restrict__ slot, int* __restrict__ count,
                                                  int* __restrict__ outside) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (WAVE - 1);
  const bool valid = i < n;
  int c = -1;
  if (valid) {
    const int u = cx[i] - bx, v = cy[i] - by, w = cz[i] - bz;
    // A grid taken from an earlier build may not hold every particle: flag
    // it (the build is then redone on the exact box) and use cell 0.
    if ((unsigned)u < (unsigned)gx && (unsigned)v < (unsigned)gy && (unsigned)w < (unsigned)gz) {
      c = (w * gy + v) * gx + u;
    } else {
      c = 0;
      atomicOr(outside, 1);
    }
  }
  const int prev = __shfl_up(c, 1);
  const bool head = valid && (lane == 0 || c != prev);
  const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
  const int nvalid = __popcll(__builtin_amdgcn_ballot_w64(valid));  // lanes 0 .. nvalid - 1
  // This lane's run starts at the last head at or below it and ends at the
  // next head (or the first invalid lane).
  const uint64_t upto = heads & (((uint64_t)2 << lane) - 1u);
  const int rs = valid ? 63 - __builtin_clzll(upto) : lane;
  const uint64_t later = heads & ~(((uint64_t)2 << rs) - 1u);
  const int rend = later ? __builtin_ctzll(later) : nvalid;
  int base = 0;
  if (head) base = atomicAdd(&count[c], rend - rs);
  base = __shfl(base, rs);
  if (valid) {
    cell[i] = c;
    slot[i] = base + (lane - rs);
  }
}

// k_cell_coords and k_cell_key in one pass for a grid known in advance
// (first cell bx, by, bz; extent gx, gy, gz): the cell coordinates
// (int)floor(x / rl), their bounding box (one atomic per block and axis),
// the type unsort of k_cell_coords, and the key and slot of k_cell_key.
// Blocks stride over chunks of 256 consecutive particles, so the lanes of
// a wavefront hold consecutive particles, as k_cell_key's runs need.
// SORTED: item i is the last build's slot i (position spos_old[i],
// particle perm_old[i]); cell and slot are indexed by that slot.
template <bool SORTED>
__global__ __launch_bounds__(256) void k_cell_key_direct(const double* __restrict__ x, const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         const double4* __restrict__ spos_old,
                                                         const int* __restrict__ perm_old, int n, double rl, int bx,
                                                         int by, int bz, int gx, int gy, int gz,
                                                         int* __restrict__ cell, int* __restrict__ slot,
                                                         int* __restrict__ count, int* __restrict__ box,
                                                         const int* __restrict__ iperm_old,
                                                         const uint8_t* __restrict__ type_s_old, int n_unsort,
                                                         uint8_t* __restrict__ type) {
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  const int lane = threadIdx.x & (WAVE - 1);
  for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
    const int i = base + threadIdx.x;
    const bool valid = i < n;
    int c = -1;
    if (valid) {
      double px, py, pz;
      if (SORTED) {
        const double4 p = spos_old[i];
        px = p.x;
        py = p.y;
        pz = p.z;
      } else {
        px = x[i];
        py = y[i];
        pz = z[i];
      }
      const int a = (int)floor(px / rl), b = (int)floor(py / rl), w0 = (int)floor(pz / rl);
      lo[0] = min(lo[0], a);
      hi[0] = max(hi[0], a);
      lo[1] = min(lo[1], b);
      hi[1] = max(hi[1], b);
      lo[2] = min(lo[2], w0);
      hi[2] = max(hi[2], w0);
      if (i < n_unsort) {
        if (SORTED) type[perm_old[i]] = type_s_old[i];
        else type[i] = type_s_old[iperm_old[i]];
      }
      const int u = a - bx, v = b - by, w = w0 - bz;
      if ((unsigned)u < (unsigned)gx && (unsigned)v < (unsigned)gy && (unsigned)w < (unsigned)gz) {
        c = (w * gy + v) * gx + u;
      } else {
        c = 0;
        atomicOr(&box[6], 1);
      }
    }
    const int prev = __shfl_up(c, 1);
    const bool head = valid && (lane == 0 || c != prev);
    const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
    const int nvalid = __popcll(__builtin_amdgcn_ballot_w64(valid));
    const uint64_t upto = heads & (((uint64_t)2 << lane) - 1u);
    const int rs = valid ? 63 - __builtin_clzll(upto) : lane;
    const uint64_t later = heads & ~(((uint64_t)2 << rs) - 1u);
    const int rend = later ? __builtin_ctzll(later) : nvalid;
    int sbase = 0;
    if (head) sbase = atomicAdd(&count[c], rend - rs);
    sbase = __shfl(sbase, rs);
    if (valid) {
      cell[i] = c;
      slot[i] = sbase + (lane - rs);
    }
  }
  __shared__ int red[6][256 / WAVE];
  const int wid = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = wave_min(lo[k]);
    hi[k] = wave_max(hi[k]);
    if (lane == 0) {
      red[k][wid] = lo[k];
      red[3 + k][wid] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int v = red[threadIdx.x][0];
    for (int w = 1; w < 256 / WAVE; ++w)
      v = threadIdx.x < 3 ? min(v, red[threadIdx.x][w]) : max(v, red[threadIdx.x][w]);
    if (threadIdx.x < 3) atomicMin(&box[threadIdx.x], v);
    else atomicMax(&box[threadIdx.x], v);
  }
}

// arrival[s]: the particle at slot s in arrival order; SORTED (keys indexed
// by the last build's slot j): also arrival_slot[s] = j.
template <bool SORTED>
__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ cell, const int* __restrict__ slot,
                                                 const int* __restrict__ start, const int* __restrict__ perm_old,
                                                 int n, int* __restrict__ arrival, int* __restrict__ arrival_slot,
                                                 int* __restrict__ skey) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int c = cell[i], s = start[c] + slot[i];
    if (SORTED) {
      arrival[s] = perm_old[i];
      arrival_slot[s] = i;
    } else {
      arrival[s] = i;
    }
    skey[s] = c;
  }
}

// Slot s of cell c holds arrival[s]; its place in ascending index order is
// p = start[c] + (number of the cell's particles with a smaller index). The
// cell's indices are read by every lane of the cell (contiguous, cached).
// The particle then goes to slot p: perm, the sorted copies of its position
// ({x, y, z, field} records; the field is evaluated or gathered here when
// one is set, else 0), {FP32 offsets from the grid origin, cell id} records
// (for the candidate prefilter), the cell's grid-boundary flags, its type,
// and the inverse permutation. arrival must not alias iperm (other lanes
// of the cell still read it). SORTED: the position comes from the last
// build's sorted record at arrival_slot[s] (the same values as x, y, z of
// the particle: both are moved by the same FP64 updates), and spos must not
// alias spos_old.
template <bool SORTED>
__global__ __launch_bounds__(256) void k_rank_gather(const int* __restrict__ arrival,
                                                     const int* __restrict__ arrival_slot,
                                                     const double4* __restrict__ spos_old,
                                                     const int* __restrict__ skey,
                                                     const int* __restrict__ start, int gx, int gy, int gz,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int n, double ox, double oy,
                                                     double oz, int* __restrict__ perm, double4* __restrict__ spos,
                                                     float4* __restrict__ sposf, int* __restrict__ sflag,
                                                     int* __restrict__ iperm, int fmode, QuadField q,
                                                     const double* __restrict__ fv, const uint8_t* __restrict__ type,
                                                     uint8_t* __restrict__ type_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int cl = skey[s], b = start[cl], e = start[cl + 1], i = arrival[s];
  int r = 0;
  for (int j = b; j < e; ++j) r += arrival[j] < i ? 1 : 0;
  const int p = b + r;
  perm[p] = i;
  double a, bb, c;
  if (SORTED) {
    const double4 po = spos_old[arrival_slot[s]];
    a = po.x;
    bb = po.y;
    c = po.z;
  } else {
    a = x[i];
    bb = y[i];
    c = z[i];
  }
  const double f = fmode == FIELD_QUADRATIC ? quad_eval(q, a, bb, c) : fmode == FIELD_VALUES ? fv[i] : 0.0;
  spos[p] = make_double4(a, bb, c, f);
  sposf[p] = make_float4((float)(a - ox), (float)(bb - oy), (float)(c - oz), __int_as_float(cl));
  const int u = cl % gx, t = cl / gx, v = t % gy, w = t / gy;
  sflag[p] = (u > 0 ? CNL_XLO : 0) | (u < gx - 1 ? CNL_XHI : 0) | (v > 0 ? CNL_YLO : 0) | (v < gy - 1 ? CNL_YHI : 0) |
             (w > 0 ? CNL_ZLO : 0) | (w < gz - 1 ? CNL_ZHI : 0);
  iperm[i] = p;
  type_s[p] = type[i];
}

// 14-sector bit of one offset: axis-dominant direction (ties x, then y,
// then z) and sign, plus the octant (src/lsmps.cpp:27).
__device__ __forceinline__ unsigned sector_bits(double dx, double dy, double dz) {
  const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
  int axis;
  if (ax >= ay && ax >= az) axis = dx >= 0 ? 0 : 1;
  else if (ay >= ax && ay >= az) axis = dy >= 0 ? 2 : 3;
  else axis = dz >= 0 ? 4 : 5;
  const int oct = 6 + (dx >= 0 ? 1 : 0) + (dy >= 0 ? 2 : 0) + (dz >= 0 ? 4 : 0);
  return (1u << axis) | (1u << oct);
}

// The same sectors from FP32 components each within eta of the exact
// offset, without branches. *sure is false when a sign (|d| <= es = 1.5 eta)
// or a comparison of two magnitudes (difference <= ec = 3 eta) is too close
// to call at that error; then the FP64 offset decides.
__device__ __forceinline__ unsigned sector_bits_f32(float dx, float dy, float dz, float es, float ec, bool* sure) {
  const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  *sure = fminf(fminf(ax, ay), az) > es && fminf(fminf(fabsf(ax - ay), fabsf(ax - az)), fabsf(ay - az)) > ec;
  // Sign bits (1 = negative). A component at or near zero makes the result
  // unsure above, so the sign of a zero never decides.
  const unsigned nx = __float_as_uint(dx) >> 31, ny = __float_as_uint(dy) >> 31, nz = __float_as_uint(dz) >> 31;
  // Octant bit 6 + (dx >= 0) + 2 (dy >= 0) + 4 (dz >= 0) = 13 - (nx + 2 ny + 4 nz).
  const unsigned oct = 0x2000u >> (nx | (ny << 1) | (nz << 2));
  // Dominant axis (ties x, then y): bit 0/1 for +x/-x, 2/3 for y, 4/5 for z.
  const bool xmax = ax >= ay && ax >= az, ymax = !xmax && ay >= az;
  const unsigned bx = 1u + nx, by = 4u << ny, bz = 16u << nz;
  const unsigned axis = xmax ? bx : (ymax ? by : bz);
  return axis | oct;
}

__device__ __forceinline__ uint8_t classify_rule(uint8_t t0, int kept, double z, double zlim, unsigned mask) {
  if (t0 == T_WALL) return T_WALL;
  if (kept < 10) return T_SPLASH;
  if (z < zlim) return T_NEARWALL;
  if (__popc(mask) < 11) return T_SURFACE;
  return T_INTERIOR;
}

// Eight lanes per particle, eight consecutive sorted particles (nearly
// always from the same few cells) per wavefront. For each of the nine
// candidate runs the wavefront stages the union of its particles' runs in
// LDS with one coalesced load per lane, and each group steps through its
// own run eight candidates at a time; a ballot gives every hit its rank
// among the group's hits, which keeps the reference's visiting order (and
// the cut at max_neighbors). Hits are staged in LDS; afterwards the group
// writes the row out with coalesced stores and, when CLASSIFY, takes the
// direction sectors of the kept neighbors only (~30) rather than of every
// candidate (~200).
//
// Candidate test: the squared distance in FP32 from 16-byte offset records
// decides every pair outside [lo, hi], a band around re^2 wider than the
// FP32 error bound (see cnl_prefilter_band); pairs inside it take the
// reference's FP64 test, so the result is exactly the FP64 predicate.
constexpr int CNL_GROUP = 8, CNL_EPI = 4;
static_assert(WAVE <= CNL_PAD + 1, "sposf padding covers one staged block");

// One chunk of candidates, one per lane: slot sc with FP32 offset record q,
// `take` false for slots outside the group's run and for the particle
// itself. The FP32 decision comes first, as wavefront masks; the rare
// pairs inside the band take the FP64 test in a branch that is uniform
// over the wavefront, so the common path has no per-candidate exec-mask
// branches. A ballot ranks every hit among its group's hits, which keeps
// the visiting order; hits go to the group's LDS row up to maxn.
__device__ __forceinline__ void cnl_chunk(const float4 q, int sc, bool take, const float4 pf, const double4 pi,
                                          const double4* __restrict__ spos, double re2, float lo, float mid, float hw,
                                          int lane, int gshift, unsigned below, int maxn, int* buf, int& found) {
  const float dx = q.x - pf.x, dy = q.y - pf.y, dz = q.z - pf.z;
  const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  const float dv = take ? d2 : __builtin_huge_valf();
  uint64_t hm = __builtin_amdgcn_ballot_w64(dv < lo);
  const uint64_t bm = __builtin_amdgcn_ballot_w64(fabsf(dv - mid) <= hw);
  if (bm) {
    bool f = false;
    if ((bm >> lane) & 1ull) {
      const double4 o = spos[sc];
      const double ddx = o.x - pi.x, ddy = o.y - pi.y, ddz = o.z - pi.z;
      f = ddx * ddx + ddy * ddy + ddz * ddz < re2;
    }
    hm = (hm & ~bm) | __builtin_amdgcn_ballot_w64(f);
  }
  const unsigned mk = (unsigned)(hm >> gshift) & 0xFFu;
  if (__builtin_amdgcn_inverse_ballot_w64(hm)) {
    const int pos = found + __popc(mk & below);
    if (pos < maxn) buf[pos] = sc;
  }
  found += __popc(mk);
}
template <bool CLASSIFY>
__global__ __launch_bounds__(256) void k_cnl_pass(const double4* __restrict__ spos, const float4* __restrict__ sposf,
                                                  const int* __restrict__ start, const int* __restrict__ sflag, int n,
                                                  int gx, int gy, int gz, double re2, double re2c, float lo,
                                                  float mid, float hw, float eta, float loc, float midc, float hwc,
                                                  int maxn,
                                                  int* __restrict__ nbr, int* __restrict__ cnt,
                                                  uint8_t* __restrict__ type_s, double zlim,
                                                  unsigned long long* __restrict__ trunc_total) {
  // LDS: (blockDim.x / CNL_GROUP) rows of maxn hits, then WAVE staged
  // candidate records per wavefront and WAVE records of tail padding.
  extern __shared__ int rowbuf[];
  const int lane = threadIdx.x & (WAVE - 1), g = lane & (CNL_GROUP - 1), sub = threadIdx.x / CNL_GROUP;
  const int per_block = blockDim.x / CNL_GROUP;
  const int s = (int)xcd_block(blockIdx.x, gridDim.x) * per_block + sub;
  const bool live = s < n;
  int* buf = rowbuf + sub * maxn;
  float4* stage = reinterpret_cast<float4*>(rowbuf + per_block * maxn) + (threadIdx.x & ~(WAVE - 1));
  int found = 0;
  double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
  float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
  int fl = 0;
  if (live) {
    pi = spos[s];
    pf = sposf[s];
    fl = sflag[s];
  }
  // Cells are x-fastest, so run r (rows dz = r / 3 - 1, dy = r % 3 - 1)
  // spans cells [c - 1, c + 1] + dz gx gy + dy gx, clipped at the grid's
  // x faces; rows outside the grid are empty. A dead lane (s >= n) has
  // nine empty runs.
  const unsigned c = (unsigned)__float_as_int(pf.w);
  const unsigned clo = c - ((fl & CNL_XLO) ? 1u : 0u), chi = c + ((fl & CNL_XHI) ? 2u : 1u);
  const unsigned below = (1u << g) - 1u;
  const int gshift = lane & ~(CNL_GROUP - 1);
  int rb[9], re_[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int dz = r / 3 - 1, dy = r % 3 - 1;
    const int need = (dz < 0 ? CNL_ZLO : dz > 0 ? CNL_ZHI : 0) | (dy < 0 ? CNL_YLO : dy > 0 ? CNL_YHI : 0);
    const bool ok = live && (fl & need) == need;
    const int* sr = start + ((ptrdiff_t)dz * gy + dy) * gx;
    rb[r] = ok ? sr[clo] : 0;
    re_[r] = ok ? sr[chi] : 0;
  }
  // The wavefront's eight particles are consecutive in cell order, so for
  // each run index their non-empty runs start and end in lane order and
  // their union is one slot range [u0, u1), from the first and the last of
  // them. When it holds at most WAVE slots (nearly always), the wavefront
  // stages it through LDS with one coalesced load, and each group steps
  // through its own run eight slots at a time. Otherwise (particles far
  // apart, in a sparse region) each group loads its own run.
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int b = rb[r], e = re_[r];
    const uint64_t ne = __builtin_amdgcn_ballot_w64(b < e);
    if (ne == 0) continue;
    const int u0 = __builtin_amdgcn_readlane(b, __builtin_ctzll(ne));
    const int u1 = __builtin_amdgcn_readlane(e, 63 - __builtin_clzll(ne));
    if (u1 - u0 <= WAVE) {
      // sposf holds CNL_PAD records past n, so this load needs no mask.
      stage[lane] = sposf[u0 + lane];
      wave_sync();
      // Lane g reads record (b - u0) + g + 8 j: below 2 WAVE for every
      // chunk the wavefront takes (b - u0 < WAVE, 8 j < WAVE), so reads
      // past this run land in the next wavefront's records or the padding.
      // An empty run takes no slot (rem <= 0), wherever it lies.
      const int bq = b < e ? b : u0;
      const float4* sq = stage + (bq - u0) + g;
      const int rem = (b < e ? e - b : 0) - g;
#pragma unroll
      for (int j = 0; j < WAVE / CNL_GROUP; ++j) {
        const bool inrun = j * CNL_GROUP < rem;
        if (__builtin_amdgcn_ballot_w64(inrun) == 0) break;
        const int sc = bq + j * CNL_GROUP + g;
        cnl_chunk(sq[j * CNL_GROUP], sc, inrun && (r != 4 || sc != s), pf, pi, spos, re2, lo, mid, hw, lane, gshift,
                  below, maxn, buf, found);
      }
      wave_sync();  // this run's reads are done before the next run's writes
    } else {
      for (int base = b; base < e; base += CNL_GROUP) {
        const int sc = base + g;
        cnl_chunk(sposf[sc], sc, sc < e && (r != 4 || sc != s), pf, pi, spos, re2, lo, mid, hw, lane, gshift, below,
                  maxn, buf, found);
      }
    }
  }
  // Each row was written by its own group, in this wavefront.
  wave_sync();
  const int kept = found < maxn ? found : maxn;
  // Classification counts and takes sectors of the kept neighbors within
  // the classification radius (the list radius unless the lists carry a
  // skin, in which case it drops the skin-only entries).
  // Both come from the FP32 offset records where the FP32 error bound eta
  // decides them (every sign and magnitude comparison, and the radius test
  // outside its band); the rest take the FP64 positions, so the result is
  // the FP64 one. Each lane takes four neighbors at a time, loads first.
  unsigned mask = 0;
  int kin = 0;
  if (live) {
    int* out = nbr + (size_t)s * maxn;
    const bool skin = re2c < re2;
    const float es = 1.5f * eta, ec = 3.0f * eta;
    for (int k0 = 0; k0 < kept; k0 += CNL_EPI * CNL_GROUP) {
      int s2[CNL_EPI];
      bool v[CNL_EPI];
#pragma unroll
      for (int u = 0; u < CNL_EPI; ++u) {
        const int k = k0 + u * CNL_GROUP + g;
        v[u] = k < kept;
        s2[u] = buf[v[u] ? k : 0];
      }
#pragma unroll
      for (int u = 0; u < CNL_EPI; ++u)
        if (v[u]) out[k0 + u * CNL_GROUP + g] = s2[u];
      if (CLASSIFY) {
        float4 q[CNL_EPI];
#pragma unroll
        for (int u = 0; u < CNL_EPI; ++u) q[u] = sposf[v[u] ? s2[u] : s];
#pragma unroll
        for (int u = 0; u < CNL_EPI; ++u) {
          const float dx = q[u].x - pf.x, dy = q[u].y - pf.y, dz = q[u].z - pf.z;
          bool sure;
          unsigned bits = sector_bits_f32(dx, dy, dz, es, ec, &sure);
          bool inside = true;
          if (skin) {
            const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            inside = d2 < loc;
            sure = sure && !(fabsf(d2 - midc) <= hwc);
          }
          if (v[u] && !sure) {
            const double4 o = spos[s2[u]];
            const double ex = o.x - pi.x, ey = o.y - pi.y, ez = o.z - pi.z;
            bits = sector_bits(ex, ey, ez);
            inside = ex * ex + ey * ey + ez * ez < re2c;
          }
          const bool take = v[u] && inside;
          mask |= take ? bits : 0u;
          kin += take ? 1 : 0;
        }
      }
    }
  }
  if (CLASSIFY) {
    mask = group8_or(mask);
    kin = group8_sum_int(kin);
  }
  const bool leader = live && g == 0;
  if (leader) {
    cnt[s] = kept;
    if (CLASSIFY) type_s[s] = classify_rule(type_s[s], kin, pi.z, zlim, mask);
  }
  unsigned long long tr = leader ? (unsigned long long)(found - kept) : 0ull;
  tr = wave_sum(tr);
  if (lane == 0 && tr) atomicAdd(trunc_total, tr);
}

// Histogram of the kept counts and (optionally) the type counts, from the
// sorted-slot arrays: LDS bins per block for the counts; the types (most
// particles share one) are counted with ballots per wavefront instead of
// same-address LDS atomics. Each block's bins go to its column of `part`
// (bin k of block b at part[k * gridDim.x + b]), then one block per bin adds
// its row. (Flushing with global atomics put ~1000 same-address atomics on
// each busy bin, serialised across the XCDs.)
__global__ __launch_bounds__(256) void k_cnl_stats(const int* __restrict__ cnt, const uint8_t* __restrict__ type_s,
                                                   int n, int maxn, bool types, int* __restrict__ part) {
  extern __shared__ int bins[];  // maxn + 1 count bins, then 5 type bins
  const int nb = maxn + 6;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) bins[k] = 0;
  __syncthreads();
  int tc[5] = {0, 0, 0, 0, 0};
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    atomicAdd(&bins[cnt[s]], 1);
    if (types) {
      const int t = type_s[s];
#pragma unroll
      for (int k = 0; k < 5; ++k) tc[k] += __popcll(__builtin_amdgcn_ballot_w64(t == k));
    }
  }
  if (types && (threadIdx.x & (WAVE - 1)) == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (tc[k]) atomicAdd(&bins[maxn + 1 + k], tc[k]);
  __syncthreads();
  for (int k = threadIdx.x; k < nb; k += blockDim.x) part[(size_t)k * gridDim.x + blockIdx.x] = bins[k];
}

__global__ __launch_bounds__(256) void k_cnl_stats_reduce(const int* __restrict__ part, int nblk, int maxn,
                                                          int* __restrict__ hist,
                                                          unsigned long long* __restrict__ tcount) {
  __shared__ unsigned long long lds[256 / WAVE];
  const int k = blockIdx.x;
  unsigned long long s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += (unsigned)part[(size_t)k * nblk + b];
  s = wave_sum(s);
  if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 256 / WAVE; ++w) t += lds[w];
    if (k <= maxn) hist[k] = (int)t;
    else tcount[k - maxn - 1] = t;
  }
}

// Stand-alone classification over the ELL table (src/lsmps.cpp:27,73),
// on sorted-order types.
__global__ __launch_bounds__(256) void k_classify(const double4* __restrict__ spos, const int* __restrict__ nbr,
                                                  const int* __restrict__ cnt, int n, int maxn, double zlim,
                                                  uint8_t* __restrict__ type_s,
                                                  unsigned long long* __restrict__ tcount) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int out = -1;
  if (s < n) {
    const int k = cnt[s];
    const int* row = nbr + (size_t)s * maxn;
    unsigned mask = 0;
    const double4 pi = spos[s];
    for (int q = 0; q < k; ++q) {
      const double4 o = spos[row[q]];
      mask |= sector_bits(o.x - pi.x, o.y - pi.y, o.z - pi.z);
    }
    out = classify_rule(type_s[s], k, pi.z, zlim, mask);
    type_s[s] = (uint8_t)out;
  }
  __shared__ unsigned c5[5];
  if (threadIdx.x < 5) c5[threadIdx.x] = 0;
  __syncthreads();
  if (out >= 0) atomicAdd(&c5[out], 1u);
  __syncthreads();
  if (threadIdx.x < 5 && c5[threadIdx.x]) atomicAdd(&tcount[threadIdx.x], (unsigned long long)c5[threadIdx.x]);
}

// Original-order copy of a sorted-order array: dst[i] = src[iperm[i]].
template <typename T>
__global__ __launch_bounds__(256) void k_unsort(const int* __restrict__ iperm, const T* __restrict__ src, int n,
                                                T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[iperm[i]];
}

// CSR in original indexing and reference order, from the ELL table.
__global__ __launch_bounds__(256) void k_ell_to_csr(const int* __restrict__ nbr, const int* __restrict__ cnt,
                                                    const int* __restrict__ iperm, const int* __restrict__ perm, int n,
                                                    int maxn, const int* __restrict__ row_ptr, int* __restrict__ col) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = iperm[i];
  const int* row = nbr + (size_t)s * maxn;
  int* out = col + row_ptr[i];
  const int k = cnt[s];
  for (int q = 0; q < k; ++q) out[q] = perm[row[q]];
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, const int* iperm_old, const uint8_t* type_s_old, int n_unsort,
                        uint8_t* type, hipStream_t st) {
  hipLaunchKernelGGL(k_init_box, dim3(1), dim3(64), 0, st, box);
  if (n > 0) {
    const unsigned g = blocks(n, 256) < 1024u ? blocks(n, 256) : 1024u;
    hipLaunchKernelGGL(k_cell_coords, dim3(g), dim3(256), 0, st, x, y, z, n, re, cx, cy, cz, box, iperm_old, type_s_old,
                       n_unsort, type);
  }
}

void launch_fill_int(int* p, int64_t n, int v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(blocks((n + 3) / 4, 256)), dim3(256), 0, st, p, n, v);
}

void launch_cell_sort(const CnlBufs& B, int n, const int* box_host, int fmode, const QuadField& q, const double* fv,
                      int* box, const CellKeyDirect* direct, unsigned long long* counters, hipStream_t st) {
  const int64_t ncells = (int64_t)B.gx * B.gy * B.gz;
  hipLaunchKernelGGL(k_build_init, dim3(std::max(1u, blocks((ncells + 3) / 4, 256))), dim3(256), 0, st, B.ccount,
                     ncells, direct ? box : nullptr, counters);
  const bool sorted = direct && direct->spos_old;
  if (direct) {
    if (n > 0) {
      const unsigned g = blocks(n, 256) < 1024u ? blocks(n, 256) : 1024u;
      if (sorted)
        hipLaunchKernelGGL(k_cell_key_direct<true>, dim3(g), dim3(256), 0, st, B.x, B.y, B.z, direct->spos_old,
                           direct->perm_old, n, direct->rl, box_host[0], box_host[1], box_host[2], B.gx, B.gy, B.gz,
                           B.cell, B.slot, B.ccount, box, direct->iperm_old, direct->type_s_old, direct->n_unsort,
                           direct->type);
      else
        hipLaunchKernelGGL(k_cell_key_direct<false>, dim3(g), dim3(256), 0, st, B.x, B.y, B.z, direct->spos_old,
                           direct->perm_old, n, direct->rl, box_host[0], box_host[1], box_host[2], B.gx, B.gy, B.gz,
                           B.cell, B.slot, B.ccount, box, direct->iperm_old, direct->type_s_old, direct->n_unsort,
                           direct->type);
    }
  } else if (n > 0) {
    hipLaunchKernelGGL(k_cell_key, dim3(blocks(n, 256)), dim3(256), 0, st, B.cx, B.cy, B.cz, n, box_host[0],
                       box_host[1], box_host[2], B.gx, B.gy, B.gz, B.cell, B.slot, B.ccount, box + 6);
  }
  launch_exclusive_scan(B.ccount, ncells, B.cstart, B.scan_scratch, st);
  if (n > 0) {
    // The arrival order goes to cx (and the last build's slots to cy; both
    // free once the keys are made): the rank pass reads it while it writes
    // iperm.
    if (sorted) {
      hipLaunchKernelGGL(k_scatter<true>, dim3(blocks(n, 256)), dim3(256), 0, st, B.cell, B.slot, B.cstart,
                         direct->perm_old, n, B.cx, B.cy, B.skey);
      hipLaunchKernelGGL(k_rank_gather<true>, dim3(blocks(n, 256)), dim3(256), 0, st, B.cx, B.cy, direct->spos_old,
                         B.skey, B.cstart, B.gx, B.gy, B.gz, B.x, B.y, B.z, n, B.ox, B.oy, B.oz, B.perm, B.spos,
                         B.sposf, B.sflag, B.iperm, fmode, q, fv, B.type, B.type_s);
    } else {
      hipLaunchKernelGGL(k_scatter<false>, dim3(blocks(n, 256)), dim3(256), 0, st, B.cell, B.slot, B.cstart,
                         nullptr, n, B.cx, nullptr, B.skey);
      hipLaunchKernelGGL(k_rank_gather<false>, dim3(blocks(n, 256)), dim3(256), 0, st, B.cx, nullptr, nullptr,
                         B.skey, B.cstart, B.gx, B.gy, B.gz, B.x, B.y, B.z, n, B.ox, B.oy, B.oz, B.perm, B.spos,
                         B.sposf, B.sflag, B.iperm, fmode, q, fv, B.type, B.type_s);
    }
  }
}

// Band [lo, hi] around re^2 outside which the FP32 squared distance of two
// offset records decides the FP64 predicate d^2 < re^2. Each offset is
// rounded once from FP64 (error <= E 2^-24 for offsets up to E, the grid
// extent), a difference adds |d| 2^-24 with |d| <= 2 re near the boundary,
// so each FP32 difference is within eta = 2 E 2^-24 + 2 re 2^-24 of the
// exact one; the squared norm (a product and two FMAs, three roundings)
// then errs by at most 2 sqrt(3) eta |d| + 3 eta^2 plus three roundings of
// ~re^2 2^-24 (the bound allows four). Twice that, plus a relative 1e-9 for
// the FP64 side, is the half-width.
static double cnl_eta(double re, double extent) { return 2.0 * extent * 0x1.0p-24 + 2.0 * re * 0x1.0p-24; }
static void cnl_prefilter_band(double re2, double extent, float* lo, float* hi) {
  const double u = 0x1.0p-24, re = std::sqrt(re2);
  const double eta = cnl_eta(re, extent);
  const double m = 2.0 * (2.0 * std::sqrt(3.0) * eta * (1.01 * re) + 3.0 * eta * eta + 4.0 * 1.1 * re2 * u) + 1e-9 * re2;
  *lo = (float)(re2 - m);
  *hi = (float)(re2 + m);
}

// Band test for the kernels: |d2 - mid| <= hw holds for every d2 in the
// band [lo, hi]: d2 - mid is exact there (Sterbenz), and hw covers half the
// width plus the rounding of mid.
static void cnl_band_test(double r2, double extent, float* lo, float* mid, float* hw) {
  float hi;
  cnl_prefilter_band(r2, extent, lo, &hi);
  const double mid_d = 0.5 * ((double)*lo + (double)hi);
  *mid = (float)mid_d;
  const double hw_d = 0.5 * ((double)hi - (double)*lo) + std::fabs((double)*mid - mid_d);
  *hw = (float)hw_d;
  if ((double)*hw < hw_d) *hw = std::nextafter(*hw, INFINITY);
}

void launch_cnl_pass(const CnlBufs& B, int n, double re2, double re2c, int maxn, bool classify, double zlim,
                     uint8_t* type_s, unsigned long long* trunc_total, int* hist, unsigned long long* tcount,
                     hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
  if (n <= 0) return;
  // 256 threads (32 particles) per block while the LDS rows stay <= 32 KB.
  const int threads = maxn <= 256 ? 256 : WAVE;
  const size_t lds = (size_t)(threads / CNL_GROUP) * (size_t)(maxn > 0 ? maxn : 1) * sizeof(int) +
                     (size_t)(threads + WAVE) * sizeof(float4);
  const unsigned grid = blocks(n, threads / CNL_GROUP);
  const double extent =
      B.extent > 0.0 ? B.extent : std::sqrt(re2) * (double)(std::max(B.gx, std::max(B.gy, B.gz)) + 1);
  float lo, mid, hw, loc, midc, hwc;
  cnl_band_test(re2, extent, &lo, &mid, &hw);
  cnl_band_test(std::min(re2c, re2), extent, &loc, &midc, &hwc);
  // Bound on the error of an FP32 offset component (for the sectors),
  // rounded up.
  const double eta_d = cnl_eta(std::sqrt(re2), extent);
  float eta = (float)eta_d;
  if ((double)eta < eta_d) eta = std::nextafter(eta, INFINITY);
  if (ev_begin) (void)hipEventRecord(ev_begin, st);
  if (classify)
    hipLaunchKernelGGL(k_cnl_pass<true>, dim3(grid), dim3(threads), lds, st, B.spos, B.sposf, B.cstart, B.sflag, n,
                       B.gx, B.gy, B.gz, re2, re2c, lo, mid, hw, eta, loc, midc, hwc, maxn, B.nbr, B.cnt, type_s, zlim, trunc_total);
  else
    hipLaunchKernelGGL(k_cnl_pass<false>, dim3(grid), dim3(threads), lds, st, B.spos, B.sposf, B.cstart, B.sflag, n,
                       B.gx, B.gy, B.gz, re2, re2c, lo, mid, hw, eta, loc, midc, hwc, maxn, B.nbr, B.cnt, type_s, zlim, trunc_total);
  if (ev_end) (void)hipEventRecord(ev_end, st);
  // hist holds maxn + 1 bins followed by (maxn + 6) * CNL_STATS_BLOCKS partials.
  const unsigned gs = std::min(blocks(n, 256), (unsigned)CNL_STATS_BLOCKS);
  int* part = hist + maxn + 1;
  hipLaunchKernelGGL(k_cnl_stats, dim3(gs), dim3(256), (size_t)(maxn + 6) * sizeof(int), st, B.cnt, type_s, n, maxn,
                     classify, part);
  hipLaunchKernelGGL(k_cnl_stats_reduce, dim3(classify ? maxn + 6 : maxn + 1), dim3(256), 0, st, part, (int)gs, maxn,
                     hist, tcount);
}

void launch_classify(const CnlBufs& B, int n, int maxn, double zlim, uint8_t* type_s, unsigned long long* tcount,
                     hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_classify, dim3(blocks(n, 256)), dim3(256), 0, st, B.spos, B.nbr, B.cnt, n, maxn, zlim, type_s,
                       tcount);
}

void launch_unsort_u8(const int* iperm, const uint8_t* src, int n, uint8_t* dst, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unsort<uint8_t>, dim3(blocks(n, 256)), dim3(256), 0, st, iperm, src, n, dst);
}

void launch_unsort_int(const int* iperm, const int* src, int n, int* dst, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_unsort<int>, dim3(blocks(n, 256)), dim3(256), 0, st, iperm, src, n, dst);
}

void launch_ell_to_csr(const CnlBufs& B, int n, int maxn, const int* row_ptr, int* col, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_ell_to_csr, dim3(blocks(n, 256)), dim3(256), 0, st, B.nbr, B.cnt, B.iperm, B.perm, n, maxn,
                       row_ptr, col);
}

}  // namespace lsmps_hip

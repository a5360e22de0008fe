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
#include "kernels.hpp"

namespace lsmps_hip {

__global__ void k_init_box(int* box) {
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
}

// Cell coordinates of every particle and their bounding box: grid-stride,
// block reduction, one atomic per block and axis.
__global__ __launch_bounds__(256) void k_cell_coords(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int n, double re,
                                                     int* __restrict__ cx, int* __restrict__ cy,
                                                     int* __restrict__ cz, int* __restrict__ box) {
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int a = (int)floor(x[i] / re), b = (int)floor(y[i] / re), c = (int)floor(z[i] / re);
    cx[i] = a;
    cy[i] = b;
    cz[i] = c;
    lo[0] = min(lo[0], a);
    hi[0] = max(hi[0], a);
    lo[1] = min(lo[1], b);
    hi[1] = max(hi[1], b);
    lo[2] = min(lo[2], c);
    hi[2] = max(hi[2], c);
  }
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

__global__ void k_fill_int(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Linear cell id (x fastest) and an arrival slot inside the cell.
__global__ __launch_bounds__(256) void k_cell_key(const int* __restrict__ cx, const int* __restrict__ cy,
                                                  const int* __restrict__ cz, int n, int bx, int by, int bz,
                                                  int gx, int gy, int* __restrict__ cell, int* __restrict__ slot,
                                                  int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = ((cz[i] - bz) * gy + (cy[i] - by)) * gx + (cx[i] - bx);
  cell[i] = c;
  slot[i] = atomicAdd(&count[c], 1);
}

__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ cell, const int* __restrict__ slot,
                                                 const int* __restrict__ start, int n, int* __restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[start[cell[i]] + slot[i]] = i;
}

// Restores ascending particle order inside each cell (the atomic slots are
// arrival order); cells hold ~10 particles, insertion sort per cell.
__global__ __launch_bounds__(256) void k_cell_sort(const int* __restrict__ start, int ncells, int* __restrict__ perm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int b = start[c], e = start[c + 1];
  for (int k = b + 1; k < e; ++k) {
    const int v = perm[k];
    int m = k - 1;
    while (m >= b && perm[m] > v) {
      perm[m + 1] = perm[m];
      --m;
    }
    perm[m + 1] = v;
  }
}

// Sorted-slot copies of positions and cell ids, and the inverse permutation.
__global__ __launch_bounds__(256) void k_gather_sorted(const int* __restrict__ perm, const int* __restrict__ cell,
                                                       const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, int n, double* __restrict__ sx,
                                                       double* __restrict__ sy, double* __restrict__ sz,
                                                       int* __restrict__ scell, int* __restrict__ iperm) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  sx[s] = x[i];
  sy[s] = y[i];
  sz[s] = z[i];
  scell[s] = cell[i];
  iperm[i] = s;
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

__device__ __forceinline__ uint8_t classify_rule(uint8_t t0, int kept, double z, double zlim, unsigned mask) {
  if (t0 == T_WALL) return T_WALL;
  if (kept < 10) return T_SPLASH;
  if (z < zlim) return T_NEARWALL;
  if (__popc(mask) < 11) return T_SURFACE;
  return T_INTERIOR;
}

// Block-aggregated small histograms (types, neighbor counts), flushed once.
struct PassCounters {
  unsigned long long* trunc_total;  // dropped pairs
  int* hist;                        // kept-count histogram, maxn + 1 bins
  unsigned long long* tcount;       // 5 type counts (classify only)
};

// One thread per sorted slot; a wavefront walks the same few candidate runs.
// Writes the ELL row, its length (sorted and original order) and, when
// CLASSIFY, the new type of the particle.
template <bool CLASSIFY>
__global__ __launch_bounds__(256) void k_cnl_pass(const double* __restrict__ sx, const double* __restrict__ sy,
                                                  const double* __restrict__ sz, const int* __restrict__ perm,
                                                  const int* __restrict__ start, const int* __restrict__ scell,
                                                  int n, int gx, int gy, int gz, double re2, int maxn,
                                                  int* __restrict__ nbr, int* __restrict__ cnt,
                                                  int* __restrict__ cnt_orig, uint8_t* __restrict__ type,
                                                  uint8_t* __restrict__ type_s, double zlim, PassCounters pc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int found = 0, kept = 0, newt = -1;
  if (s < n) {
    const int i = perm[s];
    const double xi = sx[s], yi = sy[s], zi = sz[s];
    const int c = scell[s];
    const int cxl = c % gx, t = c / gx, cyl = t % gy, czl = t / gy;
    const int x0 = cxl > 0 ? cxl - 1 : 0, x1 = cxl + 1 < gx ? cxl + 1 : gx - 1;
    int* out = nbr + (size_t)s * maxn;
    unsigned mask = 0;
    for (int dz = -1; dz <= 1; ++dz) {
      const int zz = czl + dz;
      if (zz < 0 || zz >= gz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = cyl + dy;
        if (yy < 0 || yy >= gy) continue;
        const int rowc = (zz * gy + yy) * gx;
        const int b = start[rowc + x0], e = start[rowc + x1 + 1];
        for (int s2 = b; s2 < e; ++s2) {
          if (s2 == s) continue;
          const double ddx = sx[s2] - xi, ddy = sy[s2] - yi, ddz = sz[s2] - zi;
          if (ddx * ddx + ddy * ddy + ddz * ddz < re2) {
            if (found < maxn) {
              out[found] = s2;
              if (CLASSIFY) mask |= sector_bits(ddx, ddy, ddz);
            }
            ++found;
          }
        }
      }
    }
    kept = found < maxn ? found : maxn;
    cnt[s] = kept;
    cnt_orig[i] = kept;
    if (CLASSIFY) {
      const uint8_t nt = classify_rule(type[i], kept, zi, zlim, mask);
      type[i] = nt;
      type_s[s] = nt;
      newt = nt;
    }
  }
  // Block-local histograms in LDS, flushed once per block.
  extern __shared__ int lhist[];  // maxn + 1 bins, then 5 type bins
  for (int k = threadIdx.x; k <= maxn + 5; k += blockDim.x) lhist[k] = 0;
  __syncthreads();
  if (s < n) {
    atomicAdd(&lhist[kept], 1);
    if (CLASSIFY) atomicAdd(&lhist[maxn + 1 + newt], 1);
  }
  unsigned long long tr = (unsigned long long)(found - kept);
  tr = wave_sum(tr);
  if ((threadIdx.x & (WAVE - 1)) == 0 && tr) atomicAdd(pc.trunc_total, tr);
  __syncthreads();
  for (int k = threadIdx.x; k <= maxn; k += blockDim.x)
    if (lhist[k]) atomicAdd(&pc.hist[k], lhist[k]);
  if (CLASSIFY && threadIdx.x < 5 && lhist[maxn + 1 + threadIdx.x])
    atomicAdd(&pc.tcount[threadIdx.x], (unsigned long long)lhist[maxn + 1 + threadIdx.x]);
}

// Stand-alone classification over the ELL table (src/lsmps.cpp:27,73).
__global__ __launch_bounds__(256) void k_classify(const double* __restrict__ sx, const double* __restrict__ sy,
                                                  const double* __restrict__ sz, const int* __restrict__ perm,
                                                  const int* __restrict__ nbr, const int* __restrict__ cnt, int n,
                                                  int maxn, double zlim, uint8_t* __restrict__ type,
                                                  uint8_t* __restrict__ type_s, unsigned long long* __restrict__ tcount) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int out = -1;
  if (s < n) {
    const int i = perm[s];
    const int k = cnt[s];
    const int* row = nbr + (size_t)s * maxn;
    unsigned mask = 0;
    const double xi = sx[s], yi = sy[s], zi = sz[s];
    for (int q = 0; q < k; ++q) {
      const int s2 = row[q];
      mask |= sector_bits(sx[s2] - xi, sy[s2] - yi, sz[s2] - zi);
    }
    out = classify_rule(type[i], k, zi, zlim, mask);
    type[i] = (uint8_t)out;
    type_s[s] = (uint8_t)out;
  }
  __shared__ unsigned c5[5];
  if (threadIdx.x < 5) c5[threadIdx.x] = 0;
  __syncthreads();
  if (out >= 0) atomicAdd(&c5[out], 1u);
  __syncthreads();
  if (threadIdx.x < 5 && c5[threadIdx.x]) atomicAdd(&tcount[threadIdx.x], (unsigned long long)c5[threadIdx.x]);
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

// Types in sorted order from the original-order array (after an upload or
// after a caller changes types without re-running classification).
__global__ __launch_bounds__(256) void k_gather_types(const int* __restrict__ perm, const uint8_t* __restrict__ type,
                                                      int n, uint8_t* __restrict__ type_s) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) type_s[s] = type[perm[s]];
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, hipStream_t st) {
  hipLaunchKernelGGL(k_init_box, dim3(1), dim3(64), 0, st, box);
  if (n > 0) {
    const unsigned g = blocks(n, 256) < 1024u ? blocks(n, 256) : 1024u;
    hipLaunchKernelGGL(k_cell_coords, dim3(g), dim3(256), 0, st, x, y, z, n, re, cx, cy, cz, box);
  }
}

void launch_fill_int(int* p, int64_t n, int v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(blocks(n, 256)), dim3(256), 0, st, p, n, v);
}

void launch_cell_sort(const CnlBufs& B, int n, const int* box_host, hipStream_t st) {
  const int64_t ncells = (int64_t)B.gx * B.gy * B.gz;
  launch_fill_int(B.ccount, ncells, 0, st);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_key, dim3(blocks(n, 256)), dim3(256), 0, st, B.cx, B.cy, B.cz, n, box_host[0],
                       box_host[1], box_host[2], B.gx, B.gy, B.cell, B.slot, B.ccount);
  launch_exclusive_scan(B.ccount, ncells, B.cstart, B.scan_scratch, st);
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(blocks(n, 256)), dim3(256), 0, st, B.cell, B.slot, B.cstart, n, B.perm);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks(ncells, 256)), dim3(256), 0, st, B.cstart, (int)ncells, B.perm);
    hipLaunchKernelGGL(k_gather_sorted, dim3(blocks(n, 256)), dim3(256), 0, st, B.perm, B.cell, B.x, B.y, B.z, n,
                       B.sx, B.sy, B.sz, B.scell, B.iperm);
  }
}

void launch_cnl_pass(const CnlBufs& B, int n, double re2, int maxn, bool classify, double zlim, uint8_t* type,
                     uint8_t* type_s, unsigned long long* trunc_total, int* hist, unsigned long long* tcount,
                     hipStream_t st) {
  if (n <= 0) return;
  PassCounters pc{trunc_total, hist, tcount};
  const size_t lds = (size_t)(maxn + 6) * sizeof(int);
  if (classify)
    hipLaunchKernelGGL(k_cnl_pass<true>, dim3(blocks(n, 256)), dim3(256), lds, st, B.sx, B.sy, B.sz, B.perm,
                       B.cstart, B.scell, n, B.gx, B.gy, B.gz, re2, maxn, B.nbr, B.cnt, B.cnt_orig, type, type_s,
                       zlim, pc);
  else
    hipLaunchKernelGGL(k_cnl_pass<false>, dim3(blocks(n, 256)), dim3(256), lds, st, B.sx, B.sy, B.sz, B.perm,
                       B.cstart, B.scell, n, B.gx, B.gy, B.gz, re2, maxn, B.nbr, B.cnt, B.cnt_orig, type, type_s,
                       zlim, pc);
}

void launch_classify(const CnlBufs& B, int n, int maxn, double zlim, uint8_t* type, uint8_t* type_s,
                     unsigned long long* tcount, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_classify, dim3(blocks(n, 256)), dim3(256), 0, st, B.sx, B.sy, B.sz, B.perm, B.nbr, B.cnt, n,
                       maxn, zlim, type, type_s, tcount);
}

void launch_ell_to_csr(const CnlBufs& B, int n, int maxn, const int* row_ptr, int* col, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_ell_to_csr, dim3(blocks(n, 256)), dim3(256), 0, st, B.nbr, B.cnt, B.iperm, B.perm, n, maxn,
                       row_ptr, col);
}

void launch_gather_types(const int* perm, const uint8_t* type, int n, uint8_t* type_s, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_gather_types, dim3(blocks(n, 256)), dim3(256), 0, st, perm, type, n, type_s);
}

}  // namespace lsmps_hip

// Cell-list neighbor search and particle classification on gfx950.
//
// Follows the semantics of the reference's neighbors() and classify()
// (src/lsmps.cpp:26-27) with a different data structure: a dense grid of
// cells of size re over the particles' bounding box, filled by a counting
// sort into cell-major order, and CSR neighbor lists in HBM.
//
// Bit-exactness: cell index (int)floor(x/re) with IEEE division, strict
// test ((dx*dx + dy*dy) + dz*dz) < re*re with no contraction (this file is
// built with -ffp-contract=off), and for every particle the candidates are
// visited cell z-major, then y, then x, each cell's members in ascending
// index, so truncation at max_neighbors keeps the reference's subset. Cells
// are linearised x-fastest, so the three x-neighbours of a cell are one
// contiguous run of the sorted array: 9 runs per particle instead of 27
// cells.
#include "kernels.hpp"

namespace lsmps_hip {

__global__ void k_init_box(int* box) {
  if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
}

// Cell coordinates of every particle and their bounding box.
__global__ __launch_bounds__(256) void k_cell_coords(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int n, double re,
                                                     int* __restrict__ cx, int* __restrict__ cy,
                                                     int* __restrict__ cz, int* __restrict__ box) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (i < n) {
    const int a = (int)floor(x[i] / re), b = (int)floor(y[i] / re), c = (int)floor(z[i] / re);
    cx[i] = a;
    cy[i] = b;
    cz[i] = c;
    lo[0] = hi[0] = a;
    lo[1] = hi[1] = b;
    lo[2] = hi[2] = c;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = wave_min(lo[k]);
    hi[k] = wave_max(hi[k]);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicMin(&box[k], lo[k]);
      atomicMax(&box[3 + k], hi[k]);
    }
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

// Positions in cell-major order, so candidate reads are contiguous runs.
__global__ __launch_bounds__(256) void k_gather_sorted(const int* __restrict__ perm, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       int n, double* __restrict__ sx, double* __restrict__ sy,
                                                       double* __restrict__ sz) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  sx[s] = x[i];
  sy[s] = y[i];
  sz[s] = z[i];
}

// One thread per particle, threads in cell-major order so a wavefront walks
// the same few candidate runs. COUNT pass: kept count per particle (original
// index), truncated total and a histogram of kept counts. FILL pass: writes
// the kept neighbor indices at row_ptr[i].
template <bool FILL>
__global__ __launch_bounds__(256) void k_nbr_pass(const double* __restrict__ sx, const double* __restrict__ sy,
                                                  const double* __restrict__ sz, const int* __restrict__ perm,
                                                  const int* __restrict__ start, const int* __restrict__ cell,
                                                  int n, int gx, int gy, int gz, double re2, int maxn,
                                                  int* __restrict__ count, const int* __restrict__ row_ptr,
                                                  int* __restrict__ col, unsigned long long* __restrict__ trunc_total,
                                                  int* __restrict__ hist) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int found = 0, kept = 0, i = 0;
  if (s < n) {
    i = perm[s];
    const double xi = sx[s], yi = sy[s], zi = sz[s];
    const int c = cell[i];
    const int cxl = c % gx, t = c / gx, cyl = t % gy, czl = t / gy;
    const int x0 = cxl > 0 ? cxl - 1 : 0, x1 = cxl + 1 < gx ? cxl + 1 : gx - 1;
    int* out = FILL ? col + row_ptr[i] : nullptr;
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
            if (FILL && found < maxn) out[found] = perm[s2];
            ++found;
          }
        }
      }
    }
    kept = found < maxn ? found : maxn;
    if (!FILL) count[i] = kept;
  }
  if (!FILL) {
    // Block-local histogram in LDS (maxn + 1 bins), flushed once per block.
    extern __shared__ int lhist[];
    for (int k = threadIdx.x; k <= maxn; k += blockDim.x) lhist[k] = 0;
    __syncthreads();
    if (s < n) atomicAdd(&lhist[kept], 1);
    unsigned long long tr = (unsigned long long)(found - kept);
    tr = wave_sum(tr);
    if ((threadIdx.x & (WAVE - 1)) == 0 && tr) atomicAdd(trunc_total, tr);
    __syncthreads();
    for (int k = threadIdx.x; k <= maxn; k += blockDim.x)
      if (lhist[k]) atomicAdd(&hist[k], lhist[k]);
  }
}

// Type from own type, kept neighbor count and the 14 direction sectors
// (src/lsmps.cpp:27): Wall stays, <10 neighbors Splash, z < 1.5 l0
// NearWall, fewer than 11 sectors FreeSurface, else Interior. Also counts
// the types.
__global__ __launch_bounds__(256) void k_classify(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const int* __restrict__ row_ptr,
                                                  const int* __restrict__ col, int n, double zlim,
                                                  uint8_t* __restrict__ type, unsigned long long* __restrict__ tcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int out = -1;
  if (i < n) {
    const uint8_t t0 = type[i];
    const int b = row_ptr[i], e = row_ptr[i + 1];
    if (t0 == T_WALL) {
      out = T_WALL;
    } else if (e - b < 10) {
      out = T_SPLASH;
    } else {
      const double xi = x[i], yi = y[i], zi = z[i];
      unsigned mask = 0;
      for (int k = b; k < e; ++k) {
        const int j = col[k];
        const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        const double ax = fabs(dx), ay = fabs(dy), az = fabs(dz);
        int axis;
        if (ax >= ay && ax >= az) axis = dx >= 0 ? 0 : 1;
        else if (ay >= ax && ay >= az) axis = dy >= 0 ? 2 : 3;
        else axis = dz >= 0 ? 4 : 5;
        const int oct = 6 + (dx >= 0 ? 1 : 0) + (dy >= 0 ? 2 : 0) + (dz >= 0 ? 4 : 0);
        mask |= (1u << axis) | (1u << oct);
      }
      if (zi < zlim) out = T_NEARWALL;
      else if (__popc(mask) < 11) out = T_SURFACE;
      else out = T_INTERIOR;
    }
    type[i] = (uint8_t)out;
  }
  __shared__ unsigned cnt[5];
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
  __syncthreads();
  if (out >= 0) atomicAdd(&cnt[out], 1u);
  __syncthreads();
  if (threadIdx.x < 5 && cnt[threadIdx.x]) atomicAdd(&tcount[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

static inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_cell_coords(const double* x, const double* y, const double* z, int n, double re, int* cx, int* cy,
                        int* cz, int* box, hipStream_t st) {
  hipLaunchKernelGGL(k_init_box, dim3(1), dim3(64), 0, st, box);
  if (n > 0) hipLaunchKernelGGL(k_cell_coords, dim3(blocks(n, 256)), dim3(256), 0, st, x, y, z, n, re, cx, cy, cz, box);
}

void launch_fill_int(int* p, int64_t n, int v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3(blocks(n, 256)), dim3(256), 0, st, p, n, v);
}

void launch_cell_sort(const int* cx, const int* cy, const int* cz, int n, const int* box_host, int gx, int gy,
                      int gz, int* cell, int* slot, int* count, int* start, int64_t* scan_scratch, int* perm,
                      const double* x, const double* y, const double* z, double* sx, double* sy, double* sz,
                      hipStream_t st) {
  const int64_t ncells = (int64_t)gx * gy * gz;
  launch_fill_int(count, ncells, 0, st);
  if (n > 0)
    hipLaunchKernelGGL(k_cell_key, dim3(blocks(n, 256)), dim3(256), 0, st, cx, cy, cz, n, box_host[0], box_host[1],
                       box_host[2], gx, gy, cell, slot, count);
  launch_exclusive_scan(count, ncells, start, scan_scratch, st);
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(blocks(n, 256)), dim3(256), 0, st, cell, slot, start, n, perm);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks(ncells, 256)), dim3(256), 0, st, start, (int)ncells, perm);
    hipLaunchKernelGGL(k_gather_sorted, dim3(blocks(n, 256)), dim3(256), 0, st, perm, x, y, z, n, sx, sy, sz);
  }
}

void launch_nbr_count(const double* sx, const double* sy, const double* sz, const int* perm, const int* start,
                      const int* cell, int n, int gx, int gy, int gz, double re2, int maxn, int* count,
                      unsigned long long* trunc_total, int* hist, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_nbr_pass<false>, dim3(blocks(n, 256)), dim3(256), (maxn + 1) * sizeof(int), st, sx, sy, sz, perm, start, cell, n,
                       gx, gy, gz, re2, maxn, count, (const int*)nullptr, (int*)nullptr, trunc_total, hist);
}

void launch_nbr_fill(const double* sx, const double* sy, const double* sz, const int* perm, const int* start,
                     const int* cell, int n, int gx, int gy, int gz, double re2, int maxn, const int* row_ptr, int* col,
                     hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_nbr_pass<true>, dim3(blocks(n, 256)), dim3(256), 0, st, sx, sy, sz, perm, start, cell, n,
                       gx, gy, gz, re2, maxn, (int*)nullptr, row_ptr, col, (unsigned long long*)nullptr, (int*)nullptr);
}

void launch_classify(const double* x, const double* y, const double* z, const int* row_ptr, const int* col, int n,
                     double zlim, uint8_t* type, unsigned long long* tcount, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_classify, dim3(blocks(n, 256)), dim3(256), 0, st, x, y, z, row_ptr, col, n, zlim, type,
                       tcount);
}

}  // namespace lsmps_hip

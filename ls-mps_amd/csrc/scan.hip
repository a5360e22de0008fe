// Exclusive prefix sum of int32 counts into offsets (CSR row pointers, cell
// starts). Three launches: tile sums, one-block scan of the tile sums, tile
// rescans. Each tile is 256 threads x 8 consecutive items.
#include <algorithm>

#include "kernels.hpp"

namespace lsmps_hip {

constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(const int* __restrict__ in, int64_t n,
                                                                 int64_t* __restrict__ tile_sum) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
  for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS) {
    const int64_t i = base + k;
    if (i < n) s += in[i];
  }
  __shared__ int64_t lds[SCAN_THREADS / WAVE];
  s = wave_sum(s);
  if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < SCAN_THREADS / WAVE; ++w) t += lds[w];
    tile_sum[blockIdx.x] = t;
  }
}

// One block: exclusive scan of the tile sums in place; writes the grand
// total to out[n] and to *total.
__global__ __launch_bounds__(1024) void k_scan_tile_offsets(int64_t* __restrict__ tile_sum, int ntiles,
                                                            int* __restrict__ out_last, int64_t* __restrict__ total) {
  __shared__ int64_t wsum[1024 / WAVE];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  for (int base = 0; base < ntiles; base += 1024) {
    const int i = base + threadIdx.x;
    int64_t v = i < ntiles ? tile_sum[i] : 0;
    int64_t inc = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      int64_t u = __shfl_up(inc, o);
      if (lane >= o) inc += u;
    }
    if (lane == WAVE - 1) wsum[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t acc = 0;
      for (int w = 0; w < 1024 / WAVE; ++w) {
        int64_t t = wsum[w];
        wsum[w] = acc;
        acc += t;
      }
    }
    __syncthreads();
    const int64_t excl = carry + wsum[wid] + inc - v;
    if (i < ntiles) tile_sum[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_last = (int)carry;
    *total = carry;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int* __restrict__ in, int64_t n,
                                                             const int64_t* __restrict__ tile_off,
                                                             int* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  // A lane's eight items are 32 contiguous bytes: two 16-byte accesses when
  // aligned and whole (the wavefront's accesses are then contiguous), else
  // item by item.
  const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0 &&
                   base + SCAN_ITEMS <= n;
  int v[SCAN_ITEMS];
  int tsum = 0;
  if (vec) {
    const int4 a = reinterpret_cast<const int4*>(in + base)[0], b = reinterpret_cast<const int4*>(in + base)[1];
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = (base + k < n) ? in[base + k] : 0;
  }
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) tsum += v[k];
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  int inc = wave_incl_scan(tsum);
  __shared__ int wsum[SCAN_THREADS / WAVE];
  if (lane == WAVE - 1) wsum[wid] = inc;
  __syncthreads();
  int wofs = 0;
  for (int w = 0; w < wid; ++w) wofs += wsum[w];
  int64_t run = tile_off[blockIdx.x] + wofs + inc - tsum;
  int o[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    o[k] = (int)run;
    run += v[k];
  }
  if (vec) {
    reinterpret_cast<int4*>(out + base)[0] = make_int4(o[0], o[1], o[2], o[3]);
    reinterpret_cast<int4*>(out + base)[1] = make_int4(o[4], o[5], o[6], o[7]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k < n) out[base + k] = o[k];
  }
}

size_t scan_scratch_elems(int64_t n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 2; }

// out[0..n] = exclusive prefix sums of in[0..n-1], out[n] = total; the
// 64-bit total also lands in scratch[ntiles] for overflow checks.
void launch_exclusive_scan(const int* in, int64_t n, int* out, int64_t* scratch, hipStream_t st) {
  const int ntiles = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
  if (ntiles > 0) hipLaunchKernelGGL(k_scan_tile_sums, dim3(ntiles), dim3(SCAN_THREADS), 0, st, in, n, scratch);
  hipLaunchKernelGGL(k_scan_tile_offsets, dim3(1), dim3(1024), 0, st, scratch, ntiles, out + n, scratch + ntiles);
  if (ntiles > 0) hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(SCAN_THREADS), 0, st, in, n, scratch, out);
}

// out[slot[k]] += sum over b of partials[k * nblk + b]. Row k is split over
// gridDim.y blocks, each adding its share with a few loads in flight per
// thread and one 64-bit atomic into out (a handful of atomics per launch,
// so no contention; integer sums, so the order does not matter). One block
// per row left the whole row to the load latency of 1024 threads (18 us for
// 47k partials).
constexpr int SUM_THREADS = 256;
__global__ __launch_bounds__(SUM_THREADS) void k_sum_counts(const unsigned* __restrict__ partials, int nblk,
                                                            CountSlots slots, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long lds[SUM_THREADS / WAVE];
  const int k = blockIdx.x;
  const unsigned* row = partials + (size_t)k * nblk;
  unsigned long long s0 = 0, s1 = 0;
  const int stride = SUM_THREADS * gridDim.y;
  int b = blockIdx.y * SUM_THREADS + threadIdx.x;
  for (; b + stride < nblk; b += 2 * stride) {
    s0 += row[b];
    s1 += row[b + stride];
  }
  if (b < nblk) s0 += row[b];
  const unsigned long long s = wave_sum(s0 + s1);
  if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < SUM_THREADS / WAVE; ++w) t += lds[w];
    if (t) atomicAdd(&out[slots.slot[k]], t);
  }
}

void launch_sum_counts(const unsigned* partials, int nblk, CountSlots slots, unsigned long long* out,
                       hipStream_t st) {
  if (nblk > 0 && slots.k > 0) {
    // About four partials per thread.
    const int per_block = 4 * SUM_THREADS;
    const unsigned gy = (unsigned)std::min(64, std::max(1, (nblk + per_block - 1) / per_block));
    hipLaunchKernelGGL(k_sum_counts, dim3(slots.k, gy), dim3(SUM_THREADS), 0, st, partials, nblk, slots, out);
  }
}

}  // namespace lsmps_hip

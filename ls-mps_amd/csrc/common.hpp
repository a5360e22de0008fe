// Shared device and host helpers for the gfx950 kernels of the LS-MPS step.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsmps_hip {

constexpr int WAVE = 64;  // CDNA wavefront width

// ---- wavefront reductions (64 lanes, butterfly over __shfl_xor) -----------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o);
    v = u < v ? u : v;
  }
  return v;
}
// Sum over aligned groups of G lanes (G a power of two <= 64); every lane of
// a group gets the group's sum.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Reductions over aligned 8-lane groups with DPP lane permutes (no LDS
// traffic): swap neighbours, swap pairs, then mirror inside the half-row.
// After the first two steps all four lanes of a quad hold the quad's value,
// so the half-row mirror pairs each quad with the other quad of its group.
constexpr int DPP_QUAD_SWAP1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_SWAP2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141;
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ float group8_sum(float v) {
  v += __int_as_float(dpp_i<DPP_QUAD_SWAP1>(__float_as_int(v)));
  v += __int_as_float(dpp_i<DPP_QUAD_SWAP2>(__float_as_int(v)));
  v += __int_as_float(dpp_i<DPP_HALF_MIRROR>(__float_as_int(v)));
  return v;
}
// Sum over aligned groups of G lanes (G = 1, 2, 4 or 8), by the same DPP steps.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
  static_assert(G == 1 || G == 2 || G == 4 || G == 8, "group of 1, 2, 4 or 8 lanes");
  if (G >= 2) v += __int_as_float(dpp_i<DPP_QUAD_SWAP1>(__float_as_int(v)));
  if (G >= 4) v += __int_as_float(dpp_i<DPP_QUAD_SWAP2>(__float_as_int(v)));
  if (G >= 8) v += __int_as_float(dpp_i<DPP_HALF_MIRROR>(__float_as_int(v)));
  return v;
}
__device__ __forceinline__ int group8_sum_int(int v) {
  v += dpp_i<DPP_QUAD_SWAP1>(v);
  v += dpp_i<DPP_QUAD_SWAP2>(v);
  v += dpp_i<DPP_HALF_MIRROR>(v);
  return v;
}
// Bits of `pred` over the aligned 8-lane group whose first lane (within the
// wavefront) is gbase, bit j for lane gbase + j.
__device__ __forceinline__ unsigned group8_ballot(bool pred, int gbase) {
  return (unsigned)(__ballot(pred) >> gbase) & 0xFFu;
}
__device__ __forceinline__ unsigned group8_or(unsigned v) {
  v |= (unsigned)dpp_i<DPP_QUAD_SWAP1>((int)v);
  v |= (unsigned)dpp_i<DPP_QUAD_SWAP2>((int)v);
  v |= (unsigned)dpp_i<DPP_HALF_MIRROR>((int)v);
  return v;
}

// Orders LDS accesses between the lanes of one wavefront (which execute in
// lockstep): everything before it is visible to every lane after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Block-wide sum of NV doubles; result valid in thread 0. Deterministic for a
// fixed block size. `lds` must hold NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nw + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[k * nw + w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// The GPU hands workgroup b to XCD b % 8, and each XCD has its own L2.
// Mapping the hardware block index to a logical one so that each XCD gets a
// contiguous range of logical blocks keeps spatially neighboring work (and
// the data it gathers) in one L2 instead of spreading it over all eight.
// A bijection of [0, nblk).
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblk) {
  const unsigned q = nblk / NUM_XCD, r = nblk % NUM_XCD, x = b % NUM_XCD, i = b / NUM_XCD;
  return x * q + (x < r ? x : r) + i;
}

// Event counts of a grid (statuses, failed rows, nonzeros). Same-address
// global atomics from every wavefront serialise across the XCDs, so each
// block writes its K counts to partials[k * gridDim.x + block] and one
// small launch (launch_sum_counts) adds them. v holds this thread's counts;
// lds needs K * (blockDim.x / 64) words.
template <int K>
__device__ __forceinline__ void block_count_partials(unsigned (&v)[K], unsigned* __restrict__ partials,
                                                     unsigned* lds, unsigned blk = blockIdx.x) {
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (nw == 1) {
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < K; ++k) partials[(size_t)k * gridDim.x + blk] = v[k];
    return;
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * nw + wid] = v[k];
  __syncthreads();
  if (threadIdx.x < K) {
    unsigned s = 0;
    for (int w = 0; w < nw; ++w) s += lds[threadIdx.x * nw + w];
    partials[(size_t)threadIdx.x * gridDim.x + blk] = s;
  }
}

// Grid-wide reductions are two launches: every block stores its partial
// sums (block_partials), then a one-block kernel adds them in block order
// (finalize_sum), so the result is the same on every run. (A last-block
// ticket with a device-scope fence in every block would save the second
// launch, but on this GPU each such fence writes back the XCD's L2 and costs
// far more than a launch.)
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* partials, double* lds) {
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// In a one-block kernel: totals of the NV partial arrays written by a grid
// of `nblk` blocks, valid in every thread.
template <int NV>
__device__ __forceinline__ void finalize_sum(const double* partials, int nblk, double (&tot)[NV], double* lds) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    acc[k] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc[k] += partials[(size_t)k * nblk + b];
  }
  block_sum<NV>(acc, lds);
  __shared__ double bc[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) bc[k] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = bc[k];
}

// splitmix64 finaliser (Steele, Lea, Flood 2014) and the top 53 bits of a
// hash as a double in [0, 1): the counter-based generators of the
// synthetic inputs.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit53(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

}  // namespace lsmps_hip

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

// Grid-wide reduction of NV doubles without a second launch: every block
// stores its partial sums, the last block to finish (found with an atomic
// ticket) adds the partials in block order, so the result is the same on
// every run. Returns true in all threads of the last block, which then find
// the totals in `tot` (valid in every thread). `ticket` must be zero on
// entry; the last block resets it.
template <int NV>
__device__ __forceinline__ bool grid_sum_last_block(double (&v)[NV], double* partials, unsigned* ticket,
                                                     double (&tot)[NV], double* lds) {
  block_sum<NV>(v, lds);
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    acc[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
      acc[k] += __hip_atomic_load(&partials[(size_t)k * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  block_sum<NV>(acc, lds);
  __shared__ double bc[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) bc[k] = acc[k];
    atomicExch(ticket, 0u);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = bc[k];
  return true;
}

}  // namespace lsmps_hip

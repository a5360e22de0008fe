// Learned preconditioner on gfx950: per-row features of the assembled PPE,
// and a 16 -> 64 -> 64 -> 2 MLP (tanh hidden units) with BF16 weight
// matrices, FP32 biases, FP32 accumulation, trained by minibatch Adam on
// FP32 master weights. There is no counterpart in the reference; the
// network is an extension of the solver (its output only changes the start
// vector and the diagonal scaling, never the stopping test).
//
// Features of unknown row u (particle at sorted slot s), all dimensionless:
//   0-4   one-hot particle type
//   5     k / 32, k = kept neighbor count
//   6-8   mean r/re, mean (r/re)^2, min r/re over the neighbor list
//   9     a_uu re^2 / 32 (the assembled diagonal)
//   10    coupling kept by the row: sum over unknown neighbors of w/r^2 over
//         the same sum over every neighbor (1 inside, < 1 next to
//         non-unknown particles)
//   11    unknown neighbors / k
//   12    (b_u / a_uu) / s_b, s_b = RMS of b/a over all rows
//   13    sum_j (a_uj / a_uu) (b_j / a_jj) / s_b over the row's off-diagonals
//         (0 on pinned and failed rows, whose diagonal is not a coupling sum)
//   14    Wall neighbors / k
//   15    1 for the pinned row 0 or a failed row, else 0
// Outputs: y0 -> initial guess x0 = s_b y0; y1 -> diagonal scaling
// S = exp(clamp(y1, -0.5, 0.5)) / a_uu.
#include <cmath>

#include "kernels.hpp"
#include "learn.hpp"

namespace lsmps_hip {

namespace {

constexpr int PPE_SLICE = 64, GROUP = 8, ROW_THREADS = PPE_SLICE * GROUP;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((unsigned)h << 16); }
// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const unsigned u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ double pair_w(double dx, double dy, double dz, double re, double* r) {
  const double rr = sqrt(dx * dx + dy * dy + dz * dz);
  *r = rr;
  if (rr <= 1e-12 || rr >= re) return 0.0;
  const double t = re / rr - 1.0;
  return t * t;
}

}  // namespace

// sum over rows of (b/a)^2 -> partials (one value per block).
__global__ __launch_bounds__(256) void k_rhs_scale_partials(const double* __restrict__ b, const double* __restrict__ d,
                                                            int m, double* __restrict__ partials) {
  __shared__ double lds[256 / WAVE];
  double s = 0.0;
  for (int u = blockIdx.x * 256 + threadIdx.x; u < m; u += gridDim.x * 256) {
    const double q = b[u] / d[u];
    s += q * q;
  }
  double v[1] = {s};
  block_partials<1>(v, partials, lds);
}
__global__ __launch_bounds__(256) void k_rhs_scale_final(const double* __restrict__ partials, int nblk, int m,
                                                         double* __restrict__ scal) {
  __shared__ double lds[256 / WAVE];
  double tot[1];
  finalize_sum<1>(partials, nblk, tot, lds);
  if (threadIdx.x == 0) {
    const double rms = m > 0 ? sqrt(tot[0] / m) : 0.0;
    scal[0] = rms > 0.0 && isfinite(rms) ? rms : 1.0;
  }
}

// One 512-thread block per slice of 64 rows, eight lanes per row.
__global__ __launch_bounds__(ROW_THREADS) void k_features(const double4* __restrict__ spos,
                                                          const uint8_t* __restrict__ type_s,
                                                          const int* __restrict__ uslot, const int* __restrict__ urr,
                                                          const int* __restrict__ uid, const int* __restrict__ nbr,
                                                          const int* __restrict__ ncnt, int maxn, int m, double re,
                                                          const double* __restrict__ diag,
                                                          const double* __restrict__ b,
                                                          const unsigned long long* __restrict__ failed_bits,
                                                          const double* __restrict__ scal, float* __restrict__ feat) {
  const int slice = (int)xcd_block(blockIdx.x, gridDim.x);
  const int g = threadIdx.x & (GROUP - 1);
  const int u = slice * PPE_SLICE + (int)(threadIdx.x / GROUP);
  const bool live = u < m;
  double sr = 0.0, sr2 = 0.0, rmin = 1e300, wu = 0.0, wall = 0.0, nb = 0.0;
  int nu = 0, nwall = 0, k = 0;
  double du = 1.0, qu = 0.0;
  if (live) {
    const int s = uslot[u];
    const double4 pi = spos[s];
    du = diag[u];
    k = ncnt[s];
    const int* row = nbr + (size_t)s * maxn;
    for (int q = g; q < k; q += GROUP) {
      const int s2 = row[q];
      const double4 pj = spos[s2];
      double r;
      const double w = pair_w(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, re, &r);
      const double rn = r / re;
      sr += rn;
      sr2 += rn * rn;
      rmin = rn < rmin ? rn : rmin;
      const double a = w > 0.0 ? w / (r * r) : 0.0;
      wall += a;
      nwall += type_s[s2] == T_WALL ? 1 : 0;
      const int cj = uid[s2];
      if (cj >= 0 && w > 0.0) {
        wu += a;
        ++nu;
        nb += a * (b[cj] / diag[cj]);
      }
    }
    qu = b[u] / du;
  }
  // Reductions over the row's eight lanes (xor butterflies stay in the group).
#pragma unroll
  for (int o = GROUP / 2; o > 0; o >>= 1) {
    sr += __shfl_xor(sr, o);
    sr2 += __shfl_xor(sr2, o);
    const double rm = __shfl_xor(rmin, o);
    rmin = rm < rmin ? rm : rmin;
    wu += __shfl_xor(wu, o);
    wall += __shfl_xor(wall, o);
    nb += __shfl_xor(nb, o);
    nu += __shfl_xor(nu, o);
    nwall += __shfl_xor(nwall, o);
  }
  if (!live || g != 0) return;
  const int s = uslot[u];
  const double sb = scal[0];
  const double ik = k > 0 ? 1.0 / k : 0.0;
  const bool pinned = urr[u] == 0 || ((failed_bits[u >> 6] >> (u & 63)) & 1ull);
  float f[LEARN_F];
  for (int t = 0; t < 5; ++t) f[t] = type_s[s] == t ? 1.0f : 0.0f;
  f[5] = (float)(k / 32.0);
  f[6] = (float)(sr * ik);
  f[7] = (float)(sr2 * ik);
  f[8] = (float)(k > 0 ? rmin : 0.0);
  f[9] = (float)(du * re * re / 32.0);
  f[10] = (float)(wall > 0.0 ? wu / wall : 0.0);
  f[11] = (float)(nu * ik);
  f[12] = (float)(qu / sb);
  f[13] = pinned ? 0.0f : (float)(nb / du / sb);
  f[14] = (float)(nwall * ik);
  f[15] = pinned ? 1.0f : 0.0f;
  float4* o = reinterpret_cast<float4*>(feat + (size_t)u * LEARN_F);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
  o[3] = make_float4(f[12], f[13], f[14], f[15]);
}

// Rows whose assembled diagonal was replaced by 1 (failed rows): a row with
// no off-diagonal entry other than the pin. Marked from the row lengths.
__global__ __launch_bounds__(256) void k_failed_bits(const int* __restrict__ rlen, const double* __restrict__ diag,
                                                     int m, unsigned long long* __restrict__ bits) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  const bool f = u < m && rlen[u] == 1 && diag[u] == 1.0;
  const unsigned long long mask = __ballot(f);
  if ((threadIdx.x & (WAVE - 1)) == 0 && u < m) bits[u >> 6] = mask;
}

// ---- the network ---------------------------------------------------------
// Block of 4 * MLP_ROWS threads: four threads per row, each owning 16 of
// the 64 hidden units; weight matrices staged in LDS as FP32 (converted
// from BF16), activations of the block's rows in LDS.
constexpr int MLP_ROWS = 32, MLP_THREADS = 4 * MLP_ROWS;
constexpr int PAD = LEARN_H + 1;  // row stride of activation tiles (bank spread)

struct MlpSmem {
  // Weight rows padded by one word, so the four threads of a row (16 output
  // units apart) read different banks.
  float w1[LEARN_H][LEARN_F + 1], b1[LEARN_H], w2[LEARN_H][LEARN_H + 1], b2[LEARN_H], w3[LEARN_O * LEARN_H],
      b3[LEARN_O];
  float x[MLP_ROWS][LEARN_F];
  float h1[MLP_ROWS][PAD], h2[MLP_ROWS][PAD];
  float g1[MLP_ROWS][PAD], g2[MLP_ROWS][PAD];
  float dy[MLP_ROWS][LEARN_O];
};

__device__ __forceinline__ void load_weights(MlpSmem& S, const uint16_t* __restrict__ wbf,
                                             const float* __restrict__ master) {
  for (int e = threadIdx.x; e < LEARN_H * LEARN_F; e += blockDim.x)
    S.w1[e / LEARN_F][e % LEARN_F] = bf16_to_f32(wbf[LEARN_OW1 + e]);
  for (int e = threadIdx.x; e < LEARN_H * LEARN_H; e += blockDim.x)
    S.w2[e / LEARN_H][e % LEARN_H] = bf16_to_f32(wbf[LEARN_OW2 + e]);
  for (int e = threadIdx.x; e < LEARN_O * LEARN_H; e += blockDim.x) S.w3[e] = bf16_to_f32(wbf[LEARN_OW3 + e]);
  for (int e = threadIdx.x; e < LEARN_H; e += blockDim.x) {
    S.b1[e] = master[LEARN_OB1 + e];
    S.b2[e] = master[LEARN_OB2 + e];
  }
  if (threadIdx.x < LEARN_O) S.b3[threadIdx.x] = master[LEARN_OB3 + threadIdx.x];
}

// Forward of the block's rows (x already in S.x); leaves h1, h2 in LDS and
// returns y (valid in the part-0 thread of each row).
// The 16 output units of this thread accumulate in registers; the inner
// loop reads one activation and 16 (broadcast) weights per input.
constexpr int QU = LEARN_H / 4;
__device__ __forceinline__ void mlp_forward_rows(MlpSmem& S, int r, int part, float y[LEARN_O]) {
  const int o0 = part * QU;
  float acc[QU];
#pragma unroll
  for (int q = 0; q < QU; ++q) acc[q] = S.b1[o0 + q];
#pragma unroll
  for (int k = 0; k < LEARN_F; ++k) {
    const float xk = S.x[r][k];
#pragma unroll
    for (int q = 0; q < QU; ++q) acc[q] = fmaf(S.w1[o0 + q][k], xk, acc[q]);
  }
#pragma unroll
  for (int q = 0; q < QU; ++q) S.h1[r][o0 + q] = tanhf(acc[q]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QU; ++q) acc[q] = S.b2[o0 + q];
  for (int k = 0; k < LEARN_H; ++k) {
    const float hk = S.h1[r][k];
#pragma unroll
    for (int q = 0; q < QU; ++q) acc[q] = fmaf(S.w2[o0 + q][k], hk, acc[q]);
  }
#pragma unroll
  for (int q = 0; q < QU; ++q) S.h2[r][o0 + q] = tanhf(acc[q]);
  __syncthreads();
  float p[LEARN_O];
#pragma unroll
  for (int o = 0; o < LEARN_O; ++o) {
    float a = 0.0f;
    for (int k = o0; k < o0 + LEARN_H / 4; ++k) a = fmaf(S.w3[o * LEARN_H + k], S.h2[r][k], a);
    p[o] = a;
  }
  // Sum the four parts (lanes 4r .. 4r+3 are adjacent in the wavefront).
#pragma unroll
  for (int o = 0; o < LEARN_O; ++o) {
    p[o] += __shfl_xor(p[o], 1);
    p[o] += __shfl_xor(p[o], 2);
    y[o] = p[o] + S.b3[o];
  }
}

// The parameters as the network uses them, in FP32: weight matrices
// rounded to BF16, biases as kept.
__global__ __launch_bounds__(256) void k_mlp_unpack(const uint16_t* __restrict__ wbf, const float* __restrict__ master,
                                                    float* __restrict__ wf) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= LEARN_P) return;
  const bool bias = (e >= LEARN_OB1 && e < LEARN_OW2) || (e >= LEARN_OB2 && e < LEARN_OW3) || e >= LEARN_OB3;
  wf[e] = bias ? master[e] : bf16_to_f32(wbf[e]);
}

// Inference: one row per lane, blocks striding over the rows. The block
// stages the FP32 parameters in LDS once; every weight read has the same
// address across the wavefront (an LDS broadcast, four weights per 16-byte
// read). The 64 first-layer activations stay in registers, and each
// second-layer unit is folded into the two outputs as soon as it is made.
// The same FP32 arithmetic as mlp_forward_rows (a unit's inputs summed in
// the same order; the outputs summed over the 64 units in order rather
// than as four partial sums).
constexpr int MLP_FWD_THREADS = 256;
__global__ __launch_bounds__(MLP_FWD_THREADS) void k_mlp_forward(const float* __restrict__ feat, int n,
                                                                 const float* __restrict__ wf,
                                                                 float* __restrict__ out) {
  __shared__ float4 W4[(LEARN_P + 3) / 4];
  float* W = reinterpret_cast<float*>(W4);
  for (int e = threadIdx.x; e < LEARN_P; e += MLP_FWD_THREADS) W[e] = wf[e];
  __syncthreads();
  const float4* w1 = reinterpret_cast<const float4*>(W + LEARN_OW1);
  const float4* w2 = reinterpret_cast<const float4*>(W + LEARN_OW2);
  for (int base = blockIdx.x * MLP_FWD_THREADS; base < n; base += gridDim.x * MLP_FWD_THREADS) {
    const int i = base + threadIdx.x;
    const bool live = i < n;
    float x[LEARN_F];
    {
      const float4* f4 = reinterpret_cast<const float4*>(feat + (size_t)(live ? i : 0) * LEARN_F);
#pragma unroll
      for (int q = 0; q < LEARN_F / 4; ++q) {
        const float4 v = f4[q];
        x[4 * q] = v.x;
        x[4 * q + 1] = v.y;
        x[4 * q + 2] = v.z;
        x[4 * q + 3] = v.w;
      }
    }
    float h1[LEARN_H];
#pragma unroll
    for (int o = 0; o < LEARN_H; ++o) {
      float a = W[LEARN_OB1 + o];
#pragma unroll
      for (int q = 0; q < LEARN_F / 4; ++q) {
        const float4 w = w1[o * (LEARN_F / 4) + q];
        a = fmaf(w.x, x[4 * q], a);
        a = fmaf(w.y, x[4 * q + 1], a);
        a = fmaf(w.z, x[4 * q + 2], a);
        a = fmaf(w.w, x[4 * q + 3], a);
      }
      h1[o] = tanhf(a);
      // Keep the scheduler from hoisting the whole layer's weight reads
      // (they would take every register).
      __builtin_amdgcn_sched_barrier(0);
    }
    float y0 = 0.0f, y1 = 0.0f;
    for (int o = 0; o < LEARN_H; ++o) {
      float a = W[LEARN_OB2 + o];
#pragma unroll
      for (int q = 0; q < LEARN_H / 4; ++q) {
        const float4 w = w2[o * (LEARN_H / 4) + q];
        a = fmaf(w.x, h1[4 * q], a);
        a = fmaf(w.y, h1[4 * q + 1], a);
        a = fmaf(w.z, h1[4 * q + 2], a);
        a = fmaf(w.w, h1[4 * q + 3], a);
      }
      const float h = tanhf(a);
      y0 = fmaf(W[LEARN_OW3 + o], h, y0);
      y1 = fmaf(W[LEARN_OW3 + LEARN_H + o], h, y1);
    }
    if (live) {
      out[(size_t)i * LEARN_O] = y0 + W[LEARN_OB3];
      out[(size_t)i * LEARN_O + 1] = y1 + W[LEARN_OB3 + 1];
    }
  }
}

// Loss 0.5 * sum_o wk[o] (y_o - t_o)^2 over the minibatch rows idx[0..nb)
// and its gradient: every block writes the sums over its MLP_ROWS rows to
// partials[block * (LEARN_P + 1) + e] (e = LEARN_P is the loss).
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_grad(const float* __restrict__ feat,
                                                          const float* __restrict__ target,
                                                          const int* __restrict__ idx, int nb,
                                                          const uint16_t* __restrict__ wbf,
                                                          const float* __restrict__ master, float wk0, float wk1,
                                                          float* __restrict__ partials) {
  __shared__ MlpSmem S;
  load_weights(S, wbf, master);
  const int r = threadIdx.x / 4, part = threadIdx.x & 3;
  const int j = blockIdx.x * MLP_ROWS + r;
  const bool live = j < nb;
  const int i = live ? idx[j] : 0;
  for (int k = part; k < LEARN_F; k += 4) S.x[r][k] = live ? feat[(size_t)i * LEARN_F + k] : 0.0f;
  __syncthreads();
  float y[LEARN_O];
  mlp_forward_rows(S, r, part, y);
  const float wk[LEARN_O] = {wk0, wk1};
  float loss = 0.0f;
  if (part == 0) {
#pragma unroll
    for (int o = 0; o < LEARN_O; ++o) {
      const float e = live ? y[o] - target[(size_t)i * LEARN_O + o] : 0.0f;
      S.dy[r][o] = wk[o] * e;
      loss += 0.5f * wk[o] * e * e;
    }
  }
  __syncthreads();
  const int o0 = part * (LEARN_H / 4);
  // g2 = dL/d(pre-activation of layer 2) = (W3^T dy) * (1 - h2^2).
  for (int o = o0; o < o0 + LEARN_H / 4; ++o) {
    float d = 0.0f;
#pragma unroll
    for (int q = 0; q < LEARN_O; ++q) d = fmaf(S.w3[q * LEARN_H + o], S.dy[r][q], d);
    const float h = S.h2[r][o];
    S.g2[r][o] = d * (1.0f - h * h);
  }
  __syncthreads();
  // g1 = (W2^T g2) * (1 - h1^2).
  for (int k = o0; k < o0 + LEARN_H / 4; ++k) {
    float d = 0.0f;
#pragma unroll 8
    for (int o = 0; o < LEARN_H; ++o) d = fmaf(S.w2[o][k], S.g2[r][o], d);
    const float h = S.h1[r][k];
    S.g1[r][k] = d * (1.0f - h * h);
  }
  __syncthreads();
  // Parameter gradients summed over the block's rows, one entry per thread
  // at a time, in the parameter layout.
  float* out = partials + (size_t)blockIdx.x * (LEARN_P + 1);
  for (int e = threadIdx.x; e < LEARN_P; e += MLP_THREADS) {
    float acc = 0.0f;
    if (e < LEARN_OB1) {
      const int o = e / LEARN_F, k = e % LEARN_F;
      for (int q = 0; q < MLP_ROWS; ++q) acc = fmaf(S.g1[q][o], S.x[q][k], acc);
    } else if (e < LEARN_OW2) {
      const int o = e - LEARN_OB1;
      for (int q = 0; q < MLP_ROWS; ++q) acc += S.g1[q][o];
    } else if (e < LEARN_OB2) {
      const int o = (e - LEARN_OW2) / LEARN_H, k = (e - LEARN_OW2) % LEARN_H;
      for (int q = 0; q < MLP_ROWS; ++q) acc = fmaf(S.g2[q][o], S.h1[q][k], acc);
    } else if (e < LEARN_OW3) {
      const int o = e - LEARN_OB2;
      for (int q = 0; q < MLP_ROWS; ++q) acc += S.g2[q][o];
    } else if (e < LEARN_OB3) {
      const int o = (e - LEARN_OW3) / LEARN_H, k = (e - LEARN_OW3) % LEARN_H;
      for (int q = 0; q < MLP_ROWS; ++q) acc = fmaf(S.dy[q][o], S.h2[q][k], acc);
    } else {
      const int o = e - LEARN_OB3;
      for (int q = 0; q < MLP_ROWS; ++q) acc += S.dy[q][o];
    }
    out[e] = acc;
  }
  // Block loss: rows are summed by their part-0 threads.
  __shared__ float lsum[MLP_THREADS / WAVE];
  float l = wave_sum(loss);
  if ((threadIdx.x & (WAVE - 1)) == 0) lsum[threadIdx.x / WAVE] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < MLP_THREADS / WAVE; ++w) t += lsum[w];
    out[LEARN_P] = t;
  }
}

// grad[e] = (sum over blocks of partials[b][e]) / nb, in block order.
__global__ __launch_bounds__(256) void k_mlp_grad_reduce(const float* __restrict__ partials, int nblk, int nb,
                                                         float* __restrict__ grad) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e > LEARN_P) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partials[(size_t)b * (LEARN_P + 1) + e];
  grad[e] = (float)(s / (nb > 0 ? nb : 1));
}

// Adam step on the FP32 master weights, then the BF16 copy of the matrices.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ master, float* __restrict__ m1,
                                              float* __restrict__ m2, const float* __restrict__ grad, float lr,
                                              float beta1, float beta2, float eps, float c1, float c2,
                                              uint16_t* __restrict__ wbf) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= LEARN_P) return;
  const float gr = grad[e];
  const float a = beta1 * m1[e] + (1.0f - beta1) * gr;
  const float v = beta2 * m2[e] + (1.0f - beta2) * gr * gr;
  m1[e] = a;
  m2[e] = v;
  const float w = master[e] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  master[e] = w;
  wbf[e] = f32_to_bf16(w);
}

__global__ __launch_bounds__(256) void k_mlp_cast(const float* __restrict__ master, uint16_t* __restrict__ wbf) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < LEARN_P) wbf[e] = f32_to_bf16(master[e]);
}

// Minibatch indices: idx[j] = hash(seed, step, j) mod n (with replacement).
__global__ __launch_bounds__(256) void k_batch_indices(uint32_t seed, uint32_t step, int nb, int n,
                                                       int* __restrict__ idx) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= nb) return;
  uint64_t z = ((uint64_t)seed << 32 | step) * 0x9E3779B97F4A7C15ull + (uint64_t)j;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[j] = (int)(z % (uint64_t)n);
}

// Training samples from the current system: rows u = start, start + stride,
// ... with features and two targets. t0 = x_u / s_b. t1 = ln of the
// least-squares ratio c = sum x_j q_j / sum q_j^2, q_j = b_j / a_jj, over
// the row's columns (itself and its unknown neighbors), clamped: the local
// scale that best maps one Jacobi step to the solution, averaged over the
// neighborhood so it varies smoothly in space (a per-row ratio x a / b is
// dominated by noise of the right-hand side).
__global__ __launch_bounds__(256) void k_collect(const float* __restrict__ feat, const double* __restrict__ x,
                                                 const double* __restrict__ b, const double* __restrict__ d,
                                                 const int* __restrict__ sptr, const int* __restrict__ scol,
                                                 const int* __restrict__ rlen, int m, int start, int stride,
                                                 const double* __restrict__ scal, int64_t base,
                                                 float* __restrict__ dfeat, float* __restrict__ dtarget) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int u = start + j * stride;
  if (u >= m) return;
  const int64_t o = base + j;
  for (int k = 0; k < LEARN_F; ++k) dfeat[o * LEARN_F + k] = feat[(size_t)u * LEARN_F + k];
  const double sb = scal[0];
  dtarget[o * LEARN_O] = (float)(x[u] / sb);
  const int sl = u / PPE_SLICE;
  const int e0 = sptr[sl] + (u - sl * PPE_SLICE);
  double xq = 0.0, qq = 0.0;
  for (int k = 0; k < rlen[u]; ++k) {
    const int c = scol[e0 + k * PPE_SLICE];
    const double q = b[c] / d[c];
    xq += x[c] * q;
    qq += q * q;
  }
  double t1 = 0.0;
  if (qq > 0.0) {
    const double c = xq / qq;
    t1 = c > 0.0 ? log(c) : -LEARN_DIAG_CLAMP;
  }
  t1 = t1 < -LEARN_DIAG_CLAMP ? -LEARN_DIAG_CLAMP : (t1 > LEARN_DIAG_CLAMP ? LEARN_DIAG_CLAMP : t1);
  dtarget[o * LEARN_O + 1] = (float)t1;
}

// The learned start and scaling: x0 = s_b y0, S = exp(clamp(y1)) / a.
__global__ __launch_bounds__(256) void k_learned_apply(const float* __restrict__ pred, const double* __restrict__ d,
                                                       int m, const double* __restrict__ scal,
                                                       double* __restrict__ x0, double* __restrict__ scale) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= m) return;
  const double y0 = pred[(size_t)u * LEARN_O], y1 = pred[(size_t)u * LEARN_O + 1];
  const double c = y1 < -LEARN_DIAG_CLAMP ? -LEARN_DIAG_CLAMP : (y1 > LEARN_DIAG_CLAMP ? LEARN_DIAG_CLAMP : y1);
  x0[u] = isfinite(y0) ? scal[0] * y0 : 0.0;
  scale[u] = (isfinite(c) ? exp(c) : 1.0) / d[u];
}

// ---- launchers -------------------------------------------------------------
static inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_rhs_scale(const double* b, const double* d, int m, double* partials, double* scal, hipStream_t st) {
  const unsigned g = m > 0 ? (nblocks(m, 256) < 1024u ? nblocks(m, 256) : 1024u) : 1u;
  hipLaunchKernelGGL(k_rhs_scale_partials, dim3(g), dim3(256), 0, st, b, d, m, partials);
  hipLaunchKernelGGL(k_rhs_scale_final, dim3(1), dim3(256), 0, st, partials, (int)g, m, scal);
}
void launch_features(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr, const int* uid,
                     const int* nbr, const int* ncnt, int maxn, int m, double re, const double* diag, const double* b,
                     const int* rlen, unsigned long long* failed_bits, const double* scal, float* feat,
                     hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_failed_bits, dim3(nblocks(m, 256)), dim3(256), 0, st, rlen, diag, m, failed_bits);
  hipLaunchKernelGGL(k_features, dim3(nblocks(m, PPE_SLICE)), dim3(ROW_THREADS), 0, st, spos, type_s, uslot, urr, uid,
                     nbr, ncnt, maxn, m, re, diag, b, failed_bits, scal, feat);
}
__global__ __launch_bounds__(256) void k_rows_to_rr(const int* __restrict__ rru, const float* __restrict__ src, int m,
                                                    int width, float* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)m * width) return;
  const int rr = (int)(t / width), k = (int)(t % width);
  dst[t] = src[(size_t)rru[rr] * width + k];
}
__global__ __launch_bounds__(256) void k_iota(int n, int* __restrict__ idx) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < n) idx[j] = j;
}
void launch_rows_to_rr(const int* rru, const float* src, int m, int width, float* dst, hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_rows_to_rr, dim3(nblocks((int64_t)m * width, 256)), dim3(256), 0, st, rru, src, m, width, dst);
}
void launch_iota(int n, int* idx, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3(nblocks(n, 256)), dim3(256), 0, st, n, idx);
}
void launch_mlp_cast(const float* master, uint16_t* wbf, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_cast, dim3(nblocks(LEARN_P, 256)), dim3(256), 0, st, master, wbf);
}
void launch_mlp_forward(const float* feat, int n, const uint16_t* wbf, const float* master, float* wf, float* out,
                        hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mlp_unpack, dim3(nblocks(LEARN_P, 256)), dim3(256), 0, st, wbf, master, wf);
  const unsigned g = nblocks(n, MLP_FWD_THREADS) < 2048u ? nblocks(n, MLP_FWD_THREADS) : 2048u;
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(MLP_FWD_THREADS), 0, st, feat, n, wf, out);
}
int mlp_grad_blocks(int nb) { return (int)nblocks(nb, MLP_ROWS); }
void launch_mlp_grad(const float* feat, const float* target, const int* idx, int nb, const uint16_t* wbf,
                     const float* master, float wk0, float wk1, float* partials, float* grad, hipStream_t st) {
  if (nb <= 0) return;
  const unsigned g = nblocks(nb, MLP_ROWS);
  hipLaunchKernelGGL(k_mlp_grad, dim3(g), dim3(MLP_THREADS), 0, st, feat, target, idx, nb, wbf, master, wk0, wk1,
                     partials);
  hipLaunchKernelGGL(k_mlp_grad_reduce, dim3(nblocks(LEARN_P + 1, 256)), dim3(256), 0, st, partials, (int)g, nb, grad);
}
void launch_adam(float* master, float* m1, float* m2, const float* grad, float lr, float beta1, float beta2, float eps,
                 int step, uint16_t* wbf, hipStream_t st) {
  const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblocks(LEARN_P, 256)), dim3(256), 0, st, master, m1, m2, grad, lr, beta1, beta2, eps,
                     c1, c2, wbf);
}
void launch_batch_indices(uint32_t seed, uint32_t step, int nb, int n, int* idx, hipStream_t st) {
  if (nb > 0) hipLaunchKernelGGL(k_batch_indices, dim3(nblocks(nb, 256)), dim3(256), 0, st, seed, step, nb, n, idx);
}
int collect_count(int m, int start, int stride) { return m > start ? (m - start + stride - 1) / stride : 0; }
void launch_collect(const float* feat, const double* x, const double* b, const double* d, const int* sptr,
                    const int* scol, const int* rlen, int m, int start, int stride, const double* scal, int64_t base,
                    float* dfeat, float* dtarget, hipStream_t st) {
  const int c = collect_count(m, start, stride);
  if (c > 0)
    hipLaunchKernelGGL(k_collect, dim3(nblocks(c, 256)), dim3(256), 0, st, feat, x, b, d, sptr, scol, rlen, m, start,
                       stride, scal, base, dfeat, dtarget);
}
void launch_learned_apply(const float* pred, const double* d, int m, const double* scal, double* x0, double* scale,
                          hipStream_t st) {
  if (m > 0)
    hipLaunchKernelGGL(k_learned_apply, dim3(nblocks(m, 256)), dim3(256), 0, st, pred, d, m, scal, x0, scale);
}

}  // namespace lsmps_hip

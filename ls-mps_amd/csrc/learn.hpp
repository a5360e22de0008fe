// Learned preconditioner: feature and MLP launchers (learn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsmps_hip {

constexpr int LEARN_F = 16, LEARN_H = 64, LEARN_O = 2;
// Parameter layout (FP32 master): W1 (64x16, row o = output unit), b1, W2
// (64x64), b2, W3 (2x64), b3.
constexpr int LEARN_OW1 = 0, LEARN_OB1 = LEARN_OW1 + LEARN_H * LEARN_F, LEARN_OW2 = LEARN_OB1 + LEARN_H,
              LEARN_OB2 = LEARN_OW2 + LEARN_H * LEARN_H, LEARN_OW3 = LEARN_OB2 + LEARN_H,
              LEARN_OB3 = LEARN_OW3 + LEARN_O * LEARN_H, LEARN_P = LEARN_OB3 + LEARN_O;
// |log| bound of the learned diagonal correction.
constexpr double LEARN_DIAG_CLAMP = 0.5;

void launch_rhs_scale(const double* b, const double* d, int m, double* partials, double* scal, hipStream_t st);
// failed_bits: ceil(m / 64) words.
void launch_features(const double4* spos, const uint8_t* type_s, const int* uslot, const int* urr, const int* uid,
                     const int* nbr, const int* ncnt, int maxn, int m, double re, const double* diag, const double* b,
                     const int* rlen, unsigned long long* failed_bits, const double* scal, float* feat,
                     hipStream_t st);
void launch_rows_to_rr(const int* rru, const float* src, int m, int width, float* dst, hipStream_t st);
void launch_mlp_cast(const float* master, uint16_t* wbf, hipStream_t st);
// wf: LEARN_P floats of scratch (the parameters unpacked to FP32).
void launch_mlp_forward(const float* feat, int n, const uint16_t* wbf, const float* master, float* wf, float* out,
                        hipStream_t st);
// partials: mlp_grad_blocks(nb) * (LEARN_P + 1) floats; grad: LEARN_P + 1
// (the last entry is the loss).
int mlp_grad_blocks(int nb);
void launch_mlp_grad(const float* feat, const float* target, const int* idx, int nb, const uint16_t* wbf,
                     const float* master, float wk0, float wk1, float* partials, float* grad, hipStream_t st);
void launch_adam(float* master, float* m1, float* m2, const float* grad, float lr, float beta1, float beta2, float eps,
                 int step, uint16_t* wbf, hipStream_t st);
void launch_batch_indices(uint32_t seed, uint32_t step, int nb, int n, int* idx, hipStream_t st);
void launch_iota(int n, int* idx, hipStream_t st);
int collect_count(int m, int start, int stride);
void launch_collect(const float* feat, const double* x, const double* b, const double* d, const int* sptr,
                    const int* scol, const int* rlen, int m, int start, int stride, const double* scal, int64_t base,
                    float* dfeat, float* dtarget, hipStream_t st);
void launch_learned_apply(const float* pred, const double* d, int m, const double* scal, double* x0, double* scale,
                          hipStream_t st);

}  // namespace lsmps_hip

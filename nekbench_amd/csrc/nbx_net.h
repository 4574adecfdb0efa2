// Learned element preconditioner: parameter layout and launchers (nbx_net.hip).
#pragma once
#include "nbx_internal.h"

namespace nbx {

// parameters (float, bf16-representable in the forward copy), in this order:
// Wr[27][32] residual taps, Wg[7][32] geometry, b1[32], W2[32][32] (in x out),
// b2[32], W3[27][32] output taps, b3
constexpr int kOffWr = 0, kOffWg = 864, kOffB1 = 1088, kOffW2 = 1120, kOffB2 = 2144, kOffW3 = 2176, kOffB3 = 3040;
constexpr int kNetParams = 3041;
constexpr int kNetParamsPadded = 3042;  // loss (double) follows, 8-byte aligned
constexpr int kNetPartial = 3044;       // floats per workgroup partial

int net_apply_launch(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                     const double* w, double* zl, const int* fallback, const int* skip, hipStream_t s);
// the same on the matrix cores (bf16 operands, FP32 accumulation;
// nbx_net_mma.hip): frag = the weight fragments of net_frag_launch, W the
// bf16-valued weights (for b3), feat16 the bf16 features; the assembly
// weight is 1/deg (when deg is given, equal to w bit for bit) or w
constexpr int kNetFragBytes = 64 * 176;
int net_frag_launch(const float* W, void* frag, hipStream_t s);
int net_features16_launch(int64_t n, const float* feat, uint16_t* feat16, hipStream_t s);
int net_apply_mma_launch(int32_t E, const void* frag, const float* W, const uint16_t* feat16, const double* r,
                         const double* invDiag, const double* w, const uint8_t* deg, double* zl,
                         const int* fallback, const int* skip, hipStream_t s);
// the same with the input residual (double)(res - Ad) of two FP32 vectors
// (the learned multigrid step), no fallback flag
int net_apply_mma_resid_launch(int32_t E, const void* frag, const float* W, const uint16_t* feat16,
                               const float* res, const float* Ad, const double* invDiag, const double* w,
                               const uint8_t* deg, double* zl, const int* skip, hipStream_t s);
size_t net_train_lds_bytes();
int net_train_launch(int32_t nsamples, int32_t chunk, const int2* samples, int64_t pairStride, const float* R,
                     const float* Ev, const double* invNorm, const float* W, const float* feat, const double* invDiag,
                     const double* G, float* partials, hipStream_t s);
int net_grad_sum_launch(int32_t nblocks, const float* partials, float* grad, double* loss, hipStream_t s);
int net_adam_launch(float* master, float* m, float* v, const float* grad, float* wbf, float lr, int step,
                    hipStream_t s);
int net_features_launch(int32_t E, const double* ggeo, const double* invScale, float* feat, hipStream_t s);

}  // namespace nbx

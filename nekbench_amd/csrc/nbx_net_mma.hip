// Inference of the learned element preconditioner on the matrix cores
// (bf16 operands, FP32 accumulation, v_mfma_f32_16x16x16_bf16).
//
// The three layers are products with the weights as the A operand and the
// activations, channels x nodes, as the B operand:
//   layer 1: (32 out) x (48 features: 7 geometry, 27 residual taps, 13 zero)
//            times (48 x 16 nodes), bias, ReLU
//   layer 2: (32 x 32) times (32 x 16 nodes), bias, ReLU
//   layer 3: (32 taps, 27 used) x 32 times (32 x 16 nodes)  -> T[tap][node]
// With this orientation the 16x16 FP32 output tile of one layer sits in each
// lane exactly as the next layer's B operand wants it (lane l holds rows
// 4(l/16)+i of column l%16), so layers chain in registers: ReLU and the bf16
// rounding happen in place, no LDS transpose.  The last step
// y_n = b3 + sum_tap T[tap][n + tap] and z = w invDiag (r + s y) are the FP32
// kernel's.  One workgroup of four waves per element, each wave 8 tiles of
// 16 nodes; 448 MFMAs per element.
#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kP = 512, kTaps = 27;

__device__ __forceinline__ short to_bf16(float x)
{
  uint32_t u = __float_as_uint(x);
  u = u + 0x7FFFu + ((u >> 16) & 1u);  // nearest even (finite values)
  return (short)(u >> 16);
}

__device__ __forceinline__ int nb_of(int n, int tap)
{
  const int i = n & 7, j = (n >> 3) & 7, k = n >> 6;
  const int ii = i + tap % 3 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap / 9 - 1;
  if ((unsigned)ii > 7u || (unsigned)jj > 7u || (unsigned)kk > 7u) return -1;
  return ii + (jj << 3) + (kk << 6);
}

// layer-1 weight for feature f (0..6 geometry, 7..33 residual taps, else 0) and output channel c
__device__ __forceinline__ float w1(const float* W, int f, int c)
{
  if (f < 7) return W[kOffWg + f * 32 + c];
  if (f < 34) return W[kOffWr + (f - 7) * 32 + c];
  return 0.f;
}

__global__ __launch_bounds__(256) void k_net_apply_mma(const int32_t E, const float* __restrict__ W,
                                                       const float* __restrict__ feat,
                                                       const double* __restrict__ r,
                                                       const double* __restrict__ invDiag,
                                                       const double* __restrict__ w, double* __restrict__ zl,
                                                       const int* __restrict__ fallback,
                                                       const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ float s_rh[kP];
  __shared__ float s_T[kTaps][kP];
  __shared__ double s_red[4];
  const int t = threadIdx.x, l = t & 63, wv = t >> 6, kg = l >> 4, col = l & 15;
  const int64_t e = blockIdx.x;
  double rn[2], di[2], wn[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t id = e * kP + t + 256 * h;
    rn[h] = r[id];
    di[h] = invDiag[id];
    wn[h] = w ? w[id] : 1.0;
  }
  if (fallback && *fallback) {
#pragma unroll
    for (int h = 0; h < 2; ++h) zl[e * kP + t + 256 * h] = wn[h] * (di[h] * rn[h]);
    return;
  }
  // element scale s = max |r| (exact)
  double m = fmax(fabs(rn[0]), fabs(rn[1]));
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) m = fmax(m, __shfl_xor(m, h, 64));
  if (l == 0) s_red[wv] = m;
  __syncthreads();
  const double s = fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]));
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) s_rh[t + 256 * h] = (float)(rn[h] * inv);
  // weight fragments: A[row = l%16][k = 4(l/16)+i]
  bf16x4 A1[2][3], A2[2][2], A3[2][2];
  float bias1[2][4], bias2[2][4];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const int c = mt * 16 + col;
#pragma unroll
    for (int ks = 0; ks < 3; ++ks)
#pragma unroll
      for (int i = 0; i < 4; ++i) A1[mt][ks][i] = to_bf16(w1(W, ks * 16 + 4 * kg + i, c));
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ci = ks * 16 + 4 * kg + i;
        A2[mt][ks][i] = to_bf16(W[kOffW2 + ci * 32 + c]);  // W2 is in x out
        A3[mt][ks][i] = c < kTaps ? to_bf16(W[kOffW3 + c * 32 + ci]) : (short)0;
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      bias1[mt][i] = W[kOffB1 + mt * 16 + 4 * kg + i];
      bias2[mt][i] = W[kOffB2 + mt * 16 + 4 * kg + i];
    }
  }
  __syncthreads();
  for (int j = 0; j < 8; ++j) {
    const int nd = (wv * 8 + j) * 16 + col;  // this lane's node (B column)
    // layer 1 operand: features 4kg+i of node nd, per k-step
    bf16x4 B1[3];
#pragma unroll
    for (int ks = 0; ks < 3; ++ks)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int f = ks * 16 + 4 * kg + i;
        float v = 0.f;
        if (f < 7) {
          v = feat[(e * 7 + f) * kP + nd];
        } else if (f < 34) {
          const int nb = nb_of(nd, f - 7);
          v = nb >= 0 ? s_rh[nb] : 0.f;
        }
        B1[ks][i] = to_bf16(v);
      }
    bf16x4 B2[2], B3[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 C = {bias1[mt][0], bias1[mt][1], bias1[mt][2], bias1[mt][3]};
#pragma unroll
      for (int ks = 0; ks < 3; ++ks) C = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1[mt][ks], B1[ks], C, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) B2[mt][i] = to_bf16(C[i] > 0.f ? C[i] : 0.f);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 C = {bias2[mt][0], bias2[mt][1], bias2[mt][2], bias2[mt][3]};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) C = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A2[mt][ks], B2[ks], C, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) B3[mt][i] = to_bf16(C[i] > 0.f ? C[i] : 0.f);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 C = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) C = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A3[mt][ks], B3[ks], C, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int tap = mt * 16 + 4 * kg + i;
        if (tap < kTaps) s_T[tap][nd] = C[i];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int n = t + 256 * h;
    float y = W[kOffB3];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
      const int nb = nb_of(n, tap);
      if (nb >= 0) y = y + s_T[tap][nb];
    }
    zl[e * kP + n] = wn[h] * (di[h] * (rn[h] + s * (double)y));
  }
}

}  // namespace

int net_apply_mma_launch(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                         const double* w, double* zl, const int* fallback, const int* skip, hipStream_t s)
{
  hipLaunchKernelGGL(k_net_apply_mma, dim3(E), dim3(256), 0, s, E, W, feat, r, invDiag, w, zl, fallback, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

}  // namespace nbx

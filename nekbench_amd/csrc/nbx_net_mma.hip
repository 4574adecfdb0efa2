// Inference of the learned element network on the matrix cores (bf16
// operands, FP32 accumulation; gfx950 v_mfma_f32_16x16x32_bf16).
//
// The three layers are products with the weights as the A operand and the
// activations, channels x nodes, as the B operand, for tiles of 16 nodes:
//   layer 1: (32 channels) x (34 features: 27 residual taps, 7 geometry
//            factors) -- one 16x16x32 product for the taps and the first
//            five factors, one 16x16x16 product for the last two
//   layer 2: (32 x 32) times (32 x 16 nodes)
//   layer 3: (32 taps, 27 used) x 32 times (32 x 16 nodes) -> T[tap][node]
// Per 16x16 output tile a lane holds rows 4g..4g+3 (g = lane / 16) of column
// lane % 16.  The next layer's B operand wants, per lane, eight K values of
// its node: slots 0..3 are taken as channels 4g..4g+3 (the lane's rows of the
// first output tile) and slots 4..7 as channels 16 + 4g..16 + 4g + 3 (of the
// second), and the A fragments list the weights in that same order, so the
// layers chain in registers (ReLU and bf16 rounding in place).  The
// contraction pairs slot j of lane group g in A with the same slot and group
// in B, whatever K index the hardware gives them, so any order shared by A
// and B contracts correctly.
//
// Weights: the fragments of every lane are built once per weight update
// (k_net_frag, 64 lanes x 176 bytes) and read by each workgroup with vector
// loads.  Residual taps: the element's r^ sits in a zero-padded 10^3 tile in
// LDS, so a tap is a fixed offset from the node; the geometric factors are
// staged in LDS as bf16.  One 512-lane workgroup per element, wave w takes
// z-plane w (four tiles); the 27-tap output sum y_n = b3 + sum_tap T[tap][n +
// tap] and z = w invDiag (r + s y) are the FP32 kernel's.
#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kP = 512, kTaps = 27, kPad = 10, kTRow = 516;  // T row stride: groups of 4 rows 16 banks apart

// one lane's fragments (176 bytes)
struct alignas(16) LaneFrag {
  s16x8 a1[2];   // layer 1, features 0..31: taps 0..26, factors 0..4 (slot j of group g: feature 8g + j)
  s16x8 a2[2];   // layer 2, input channel order (4g + j | 16 + 4g + j - 4)
  s16x8 a3[2];   // layer 3, rows = taps
  s16x4 a1t[2];  // layer 1 tail (16x16x16): features 32 + 4g + j (factors 5, 6, then zero)
  float b1[2][4];
  float b2[2][4];
};
static_assert(sizeof(LaneFrag) == 176, "fragment layout");

__device__ __forceinline__ short bfs(float x) { return (short)bf16_bits(x); }

// input channel of slot j (0..7) of lane group g for layers 2 and 3
__device__ __forceinline__ int chan_of(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

__global__ __launch_bounds__(64) void k_net_frag(const float* __restrict__ W, LaneFrag* __restrict__ frag)
{
  const int l = threadIdx.x, g = l >> 4, row = l & 15;
  LaneFrag f;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const int c = 16 * mt + row;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int ft = 8 * g + j;  // 0..31
      f.a1[mt][j] = bfs(ft < kTaps ? W[kOffWr + ft * 32 + c] : W[kOffWg + (ft - kTaps) * 32 + c]);
      const int ci = chan_of(g, j);
      f.a2[mt][j] = bfs(W[kOffW2 + ci * 32 + c]);  // W2 is in x out
      f.a3[mt][j] = c < kTaps ? bfs(W[kOffW3 + c * 32 + ci]) : (short)0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int ft = 32 + 4 * g + j;  // factors 5 and 6 at 32, 33
      f.a1t[mt][j] = ft < 34 ? bfs(W[kOffWg + (ft - kTaps) * 32 + c]) : (short)0;
      f.b1[mt][j] = W[kOffB1 + 16 * mt + 4 * g + j];
      f.b2[mt][j] = W[kOffB2 + 16 * mt + 4 * g + j];
    }
  }
  frag[l] = f;
}

__global__ __launch_bounds__(256) void k_feat16(const int64_t n, const float* __restrict__ feat,
                                                uint16_t* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bf16_bits(feat[i]);
}

__device__ __forceinline__ bf16x8 as_bf(s16x8 v) { return __builtin_bit_cast(bf16x8, v); }

__global__ __launch_bounds__(512) void k_net_apply_mma(const int32_t E, const LaneFrag* __restrict__ frag,
                                                       const float* __restrict__ W,
                                                       const uint16_t* __restrict__ feat16,
                                                       const double* __restrict__ r,
                                                       const double* __restrict__ invDiag,
                                                       const double* __restrict__ w, const uint8_t* __restrict__ deg,
                                                       double* __restrict__ zl, const int* __restrict__ fallback,
                                                       const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ float s_T[kTaps * kTRow];
  __shared__ float s_rp[kPad * kPad * kPad];
  __shared__ uint16_t s_g[7 * kP];
  __shared__ double s_red[8];
  const int t = threadIdx.x, l = t & 63, wv = t >> 6, g = l >> 4, col = l & 15;
  const int64_t e = blockIdx.x;
  const int64_t id = e * kP + t;
  const double rn = r[id];
  const double di = invDiag[id];
  const double wn = deg ? 1.0 / (double)deg[id] : (w ? w[id] : 1.0);
  if (fallback && *fallback) {
    zl[id] = wn * (di * rn);
    return;
  }
  for (int q = t; q < kPad * kPad * kPad; q += kP) s_rp[q] = 0.f;
#pragma unroll
  for (int q = 0; q < 7; ++q) s_g[q * kP + t] = feat16[(e * 7 + q) * kP + t];
  // element scale s = max |r| (exact)
  double m = fabs(rn);
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) m = fmax(m, __shfl_xor(m, h, 64));
  if (l == 0) s_red[wv] = m;
  __syncthreads();
  double s = s_red[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) s = fmax(s, s_red[q]);
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  {
    const int i = t & 7, j = (t >> 3) & 7, k = t >> 6;
    s_rp[(k + 1) * 100 + (j + 1) * 10 + (i + 1)] = (float)(rn * inv);
  }
  // this lane's fragments
  const LaneFrag f = frag[l];
  // offsets of the taps in slots 0..7 (feature 8g + j; taps < 27, factors after)
  int toff[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ft = 8 * g + j;
    toff[j] = ft < kTaps ? (ft / 9 - 1) * 100 + ((ft / 3) % 3 - 1) * 10 + (ft % 3 - 1) : -1;
  }
  __syncthreads();
#pragma unroll 1
  for (int tt = 0; tt < 4; ++tt) {
    const int nd = wv * 64 + tt * 16 + col;  // this lane's node (B column)
    const int base = (wv + 1) * 100 + ((nd >> 3 & 7) + 1) * 10 + ((nd & 7) + 1);
    s16x8 b1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int ft = 8 * g + j;
      b1[j] = ft < kTaps ? bfs(s_rp[base + toff[j]]) : (short)s_g[(ft - kTaps) * kP + nd];
    }
    s16x4 b1t;
#pragma unroll
    for (int j = 0; j < 4; ++j) b1t[j] = (g == 0 && j < 2) ? (short)s_g[(5 + j) * kP + nd] : (short)0;
    f32x4 C[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      C[mt] = f32x4{f.b1[mt][0], f.b1[mt][1], f.b1[mt][2], f.b1[mt][3]};
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a1[mt]), as_bf(b1), C[mt], 0, 0, 0);
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(f.a1t[mt], b1t, C[mt], 0, 0, 0);
    }
    s16x8 b2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b2[j] = bfs(C[0][j] > 0.f ? C[0][j] : 0.f);
      b2[4 + j] = bfs(C[1][j] > 0.f ? C[1][j] : 0.f);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      C[mt] = f32x4{f.b2[mt][0], f.b2[mt][1], f.b2[mt][2], f.b2[mt][3]};
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a2[mt]), as_bf(b2), C[mt], 0, 0, 0);
    }
    s16x8 b3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      b3[j] = bfs(C[0][j] > 0.f ? C[0][j] : 0.f);
      b3[4 + j] = bfs(C[1][j] > 0.f ? C[1][j] : 0.f);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f32x4 T = {0.f, 0.f, 0.f, 0.f};
      T = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a3[mt]), as_bf(b3), T, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int tap = 16 * mt + 4 * g + j;
        if (tap < kTaps) s_T[tap * kTRow + nd] = T[j];
      }
    }
  }
  __syncthreads();
  const int i = t & 7, j = (t >> 3) & 7, k = t >> 6;
  float y = W[kOffB3];
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int ii = i + tap % 3 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap / 9 - 1;
    if ((unsigned)ii <= 7u && (unsigned)jj <= 7u && (unsigned)kk <= 7u)
      y = y + s_T[tap * kTRow + ii + (jj << 3) + (kk << 6)];
  }
  zl[id] = wn * (di * (rn + s * (double)y));
}

}  // namespace

int net_frag_launch(const float* W, void* frag, hipStream_t s)
{
  hipLaunchKernelGGL(k_net_frag, dim3(1), dim3(64), 0, s, W, static_cast<LaneFrag*>(frag));
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_features16_launch(int64_t n, const float* feat, uint16_t* feat16, hipStream_t s)
{
  hipLaunchKernelGGL(k_feat16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, feat, feat16);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_apply_mma_launch(int32_t E, const void* frag, const float* W, const uint16_t* feat16, const double* r,
                         const double* invDiag, const double* w, const uint8_t* deg, double* zl,
                         const int* fallback, const int* skip, hipStream_t s)
{
  if (E <= 0) return NBX_SUCCESS;
  hipLaunchKernelGGL(k_net_apply_mma, dim3(E), dim3(kP), 0, s, E, static_cast<const LaneFrag*>(frag), W, feat16, r,
                     invDiag, w, deg, zl, fallback, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

}  // namespace nbx

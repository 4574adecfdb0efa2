// axhelm with the tensor contractions on the matrix cores: bf16 or fp16
// operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16 / _f16), N = 7, one
// field.
//
// One wave64 workgroup holds two elements.  Each 1-D contraction of an
// element, (D q)(a, c) = sum_m D[a][m] q(m, c) with a the index along the
// direction and c = 0..63 the other two, is a 8x8 by 8x64 product; the two
// elements are stacked into one 16x16x16 MFMA with the block-diagonal
// operand diag(D, D), four MFMAs per direction (16 columns each).  The
// geometric factors are applied in FP32 on the vector ALUs exactly as in the
// FP32 kernel, and the transposed contractions use diag(D^T, D^T):
//
//   q_h  = h(q),  D_h = h(D)            h = round to the operand type (nearest even)
//   qr   = D_h q_h along r (FP32 accumulate), qs, qt alike
//   Gr   = fma(G02, qt, fma(G01, qs, G00 qr))  (and Gs, Gt), non-BK * lambda0
//   Aq   = D_h^T h(Gr) + D_h^T h(Gs) + D_h^T h(Gt)   (summed r, s, t)
//          non-BK: + GWJ lambda1 q
//
// MFMA lane layout (16x16x16): lane l holds A[l%16][4(l/16)+i],
// B[4(l/16)+i][l%16] and C[4(l/16)+i][l%16], i = 0..3.
// bf16 keeps the FP32 exponent range, so small residuals do not underflow;
// fp16 carries 3 more significand bits but its normal range ends at 2^-14,
// so it serves operands of order one (the geometric factors stay FP32: on
// fine meshes GWJ falls below that range).
#include "nbx_internal.h"

namespace nbx {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_cvt_pk_bf16_f32: nearest even, the bits of bf16_bits for every non-NaN input
__device__ __forceinline__ short f2bf(float x) { return (short)bf16_bits_hw(x); }

// matrix-core operand types: storage element, 4-vector, rounding, product
struct OpBf16 {
  typedef short elem;
  typedef short vec4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ elem round(float x) { return f2bf(x); }
  static __device__ __forceinline__ f32x4 mma(const vec4& a, const vec4& b, const f32x4& c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};

struct OpF16 {
  typedef _Float16 elem;
  typedef _Float16 vec4 __attribute__((ext_vector_type(4)));
  // v_cvt_f16_f32: nearest even, fp16 subnormals kept
  static __device__ __forceinline__ elem round(float x) { return (_Float16)x; }
  static __device__ __forceinline__ f32x4 mma(const vec4& a, const vec4& b, const f32x4& c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
  }
};

__device__ __forceinline__ float ld_geo(const float* p) { return ldnt(p); }
__device__ __forceinline__ float ld_geo(const uint16_t* p) { return __uint_as_float((uint32_t)ldnt(p) << 16); }

// node of element-local index (a along the direction, c over the other two)
template <int DIR>
__device__ __forceinline__ int mma_node(int a, int c)
{
  if (DIR == 0) return a + 8 * c;                        // r: i = a, c = j + 8k
  if (DIR == 1) return (c & 7) + 8 * a + 64 * (c >> 3);  // s: j = a, c = i + 8k
  return c + 64 * a;                                     // t: k = a, c = i + 8j
}

template <typename OP, int DIR>
__device__ __forceinline__ void mma_forward(const typename OP::vec4& A, const typename OP::elem* __restrict__ src,
                                            float* __restrict__ dst, int l)
{
  const int kb = 4 * (l >> 4);
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    const int col = 16 * cc + (l & 15);
    typename OP::vec4 B;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kb + i;
      B[i] = src[(k >> 3) * 512 + mma_node<DIR>(k & 7, col)];
    }
    f32x4 C = {0.f, 0.f, 0.f, 0.f};
    C = OP::mma(A, B, C);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = kb + i;
      dst[(r >> 3) * 512 + mma_node<DIR>(r & 7, col)] = C[i];
    }
  }
}

template <typename OP, int DIR>
__device__ __forceinline__ void mma_transpose_acc(const typename OP::vec4& A, const typename OP::elem* __restrict__ src,
                                                  float* __restrict__ acc, int l)
{
  const int kb = 4 * (l >> 4);
#pragma unroll
  for (int cc = 0; cc < 4; ++cc) {
    const int col = 16 * cc + (l & 15);
    typename OP::vec4 B;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kb + i;
      B[i] = src[(k >> 3) * 512 + mma_node<DIR>(k & 7, col)];
    }
    f32x4 C = {0.f, 0.f, 0.f, 0.f};
    C = OP::mma(A, B, C);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = kb + i;
      const int n = (r >> 3) * 512 + mma_node<DIR>(r & 7, col);
      acc[n] = (DIR == 0) ? C[i] : acc[n] + C[i];
    }
  }
}

template <bool BK, typename G, typename OP>
__global__ __launch_bounds__(64) void k_axhelm_mma(const int32_t Nelements, const int32_t offset,
                                                   const G* __restrict__ ggeo, const float* __restrict__ D,
                                                   const G* __restrict__ lambda, const float* __restrict__ q,
                                                   float* __restrict__ Aq, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  typedef typename OP::elem H;
  __shared__ H s_q[1024];
  __shared__ float s_d[3][1024];
  __shared__ H s_g[3][1024];
  __shared__ float s_o[1024];
  const int l = threadIdx.x;
  const int64_t e0 = 2 * (int64_t)blockIdx.x;
  const bool two = e0 + 1 < Nelements;  // an odd count leaves the second slot idle (computes e0, stores nothing)
  typename OP::vec4 Af, At;
  {
    const int row = l & 15, kb = 4 * (l >> 4);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kb + i;
      const bool same = (row >> 3) == (k >> 3);
      Af[i] = same ? OP::round(D[(row & 7) * 8 + (k & 7)]) : (H)0;
      At[i] = same ? OP::round(D[(k & 7) * 8 + (row & 7)]) : (H)0;
    }
  }
  float qv[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int n = l + 64 * t;
    const int64_t e = (two && n >= 512) ? e0 + 1 : e0;
    qv[t] = ldnt(&q[e * 512 + (n & 511)]);
    s_q[n] = OP::round(qv[t]);
  }
  __syncthreads();
  mma_forward<OP, 0>(Af, s_q, s_d[0], l);
  mma_forward<OP, 1>(Af, s_q, s_d[1], l);
  mma_forward<OP, 2>(Af, s_q, s_d[2], l);
  __syncthreads();
#pragma unroll 4
  for (int t = 0; t < 16; ++t) {
    const int n = l + 64 * t, nn = n & 511;
    const int64_t e = (two && n >= 512) ? e0 + 1 : e0;
    const G* g = ggeo + e * (int64_t)(kNggeo * 512) + nn;
    const float qr = s_d[0][n], qs = s_d[1][n], qt = s_d[2][n];
    const float G00 = ld_geo(&g[kG00 * 512]), G01 = ld_geo(&g[kG01 * 512]), G02 = ld_geo(&g[kG02 * 512]);
    const float G11 = ld_geo(&g[kG11 * 512]), G12 = ld_geo(&g[kG12 * 512]), G22 = ld_geo(&g[kG22 * 512]);
    float Gr = __builtin_fmaf(G02, qt, __builtin_fmaf(G01, qs, G00 * qr));
    float Gs = __builtin_fmaf(G12, qt, __builtin_fmaf(G11, qs, G01 * qr));
    float Gt = __builtin_fmaf(G22, qt, __builtin_fmaf(G12, qs, G02 * qr));
    if (!BK) {
      const float lam0 = ld_geo(&lambda[e * 512 + nn]);
      Gr = lam0 * Gr;
      Gs = lam0 * Gs;
      Gt = lam0 * Gt;
    }
    s_g[0][n] = OP::round(Gr);
    s_g[1][n] = OP::round(Gs);
    s_g[2][n] = OP::round(Gt);
  }
  __syncthreads();
  mma_transpose_acc<OP, 0>(At, s_g[0], s_o, l);
  __syncthreads();
  mma_transpose_acc<OP, 1>(At, s_g[1], s_o, l);
  __syncthreads();
  mma_transpose_acc<OP, 2>(At, s_g[2], s_o, l);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int n = l + 64 * t, nn = n & 511;
    if (n >= 512 && !two) continue;
    const int64_t e = e0 + (n >> 9);
    float res = s_o[n];
    if (!BK) {
      const G* g = ggeo + e * (int64_t)(kNggeo * 512) + nn;
      res = __builtin_fmaf(ld_geo(&g[kGWJ * 512]) * ld_geo(&lambda[e * 512 + offset + nn]), qv[t], res);
    }
    stnt(res, &Aq[e * 512 + nn]);
  }
}

int axhelm_mma_launch(int N, int Nf, int bk, bool bf16_geometry, bool f16_operands, int32_t Nelements,
                      int32_t offset, const void* ggeo, const void* D, const void* lambda, const void* q, void* Aq,
                      const int* skip, hipStream_t s)
{
  const char* name = f16_operands ? "axhelm fp16-MFMA" : "axhelm bf16-MFMA";
  if (N != 7 || Nf != 1) return fail(NBX_ERR_UNSUPPORTED, "%s: N=7 and one field only", name);
  if (f16_operands && bf16_geometry) return fail(NBX_ERR_UNSUPPORTED, "%s: FP32 geometry only", name);
  if (Nelements <= 0) return Nelements == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "axhelm: Nelements < 0");
  const dim3 grid((unsigned)((Nelements + 1) / 2)), block(64);
#define NBX_MMA(BKV, GT, OPT)                                                                                \
  hipLaunchKernelGGL((k_axhelm_mma<BKV, GT, OPT>), grid, block, 0, s, Nelements, offset, (const GT*)ggeo,    \
                     (const float*)D, (const GT*)lambda, (const float*)q, (float*)Aq, skip)
  if (f16_operands) {
    if (bk) NBX_MMA(true, float, OpF16); else NBX_MMA(false, float, OpF16);
  } else if (bf16_geometry) {
    if (bk) NBX_MMA(true, uint16_t, OpBf16); else NBX_MMA(false, uint16_t, OpBf16);
  } else {
    if (bk) NBX_MMA(true, float, OpBf16); else NBX_MMA(false, float, OpBf16);
  }
#undef NBX_MMA
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

}  // namespace nbx

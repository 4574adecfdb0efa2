// axhelm with the tensor contractions on the matrix cores: bf16 or fp16
// operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16 / _f16).  N = 7 with
// one field uses k_axhelm_mma; N = 9, and N = 7 with three fields, use
// k_axhelm_mma16 (below).
//
// k_axhelm_mma: one wave64 workgroup holds two elements.  Each 1-D contraction of an
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

// General form for 8 <= Nq <= 16 and one or three fields (the N = 9 BP6
// operator among them): one element per 256-lane workgroup, each 1-D
// contraction of a field a (16 x 16) by (16 x Nq^2) product with D
// zero-padded to 16 x 16, columns in tiles of 16 (the padding rows and
// columns of D are zero, so the padded rows of the operand contribute
// nothing and the padded output rows are never read).  Operands sit in LDS
// in the MFMA's B order: for direction d the column c runs over the other two
// indices and the 16 values of the contracted index are contiguous, so a
// lane's four operands (rows 4g .. 4g + 3 of its column) are one 8-byte
// read and its four products one 16-byte write.  The geometric factors
// and lambda are read once per element into registers and serve every field;
// each field runs: stage h(q) in the three direction orders; forward
// products; G-scaling in FP32 (the FP32 kernel's fma order) rounded to the
// operand type in the order of its own direction; transposed products;
// Aq = (r + s) + t [+ GWJ lambda1 q].  The arithmetic is that of the N = 7
// kernel above, so one restatement (nbo_axhelm_*mma) covers both.
template <int NQ, int NF, bool BK, typename G, typename OP>
__global__ __launch_bounds__(256) void k_axhelm_mma16(const int32_t Nelements, const int32_t offset,
                                                      const G* __restrict__ ggeo, const float* __restrict__ D,
                                                      const G* __restrict__ lambda, const float* __restrict__ q,
                                                      float* __restrict__ Aq, const int* __restrict__ skip)
{
  static_assert(NQ >= 2 && NQ <= 16, "Nq in [2, 16]");
  constexpr int P = NQ * NQ * NQ, NQ2 = NQ * NQ, CT = (NQ2 + 15) / 16, NC = CT * 16;
  constexpr int BS = 20;  // operand column stride (elements): 40 bytes, 8-byte aligned reads
  constexpr int FS = 20;  // product column stride (floats): 80 bytes, 16-byte aligned writes
  constexpr int PT = (P + 255) / 256;
  typedef typename OP::elem H;
  typedef typename OP::vec4 V4;
  if (skip && *skip) return;
  __shared__ __align__(16) H s_b[3][NC * BS];      // operands per direction: h(q), then h(G)
  __shared__ __align__(16) float s_c[3][NC * FS];  // products per direction
  const int t = threadIdx.x, l = t & 63, w = t >> 6, kb = 4 * (l >> 4), lc = l & 15;
  const int64_t e = blockIdx.x;
  // zero operands once: the padding rows (contracted index >= Nq) stay zero
  for (int u = t; u < 3 * NC * BS; u += 256) (&s_b[0][0])[u] = (H)0;
  V4 Af, At;
  {
    const int row = lc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kb + i;
      const bool in = row < NQ && k < NQ;
      Af[i] = in ? OP::round(D[row * NQ + k]) : (H)0;
      At[i] = in ? OP::round(D[k * NQ + row]) : (H)0;
    }
  }
  // this lane's nodes: their places in the three direction orders and geometry
  int pb[PT][3], pc[PT][3];
  float gf[PT][6], l0[PT], wl1[PT];
#pragma unroll
  for (int u = 0; u < PT; ++u) {
    const int n = t + 256 * u, nn = n < P ? n : P - 1;
    const int i = nn % NQ, j = (nn / NQ) % NQ, k = nn / NQ2;
    pb[u][0] = (j + NQ * k) * BS + i;
    pb[u][1] = (i + NQ * k) * BS + j;
    pb[u][2] = (i + NQ * j) * BS + k;
    pc[u][0] = (j + NQ * k) * FS + i;
    pc[u][1] = (i + NQ * k) * FS + j;
    pc[u][2] = (i + NQ * j) * FS + k;
    const G* g = ggeo + e * (int64_t)(kNggeo * P) + nn;
    gf[u][0] = ld_geo(&g[kG00 * P]);
    gf[u][1] = ld_geo(&g[kG01 * P]);
    gf[u][2] = ld_geo(&g[kG02 * P]);
    gf[u][3] = ld_geo(&g[kG11 * P]);
    gf[u][4] = ld_geo(&g[kG12 * P]);
    gf[u][5] = ld_geo(&g[kG22 * P]);
    l0[u] = BK ? 1.f : ld_geo(&lambda[e * P + nn]);
    wl1[u] = BK ? 0.f : ld_geo(&g[kGWJ * P]) * ld_geo(&lambda[e * P + offset + nn]);
  }
  // a (direction, column tile) product per wave turn
  auto products = [&](const V4& A) {
    for (int tile = w; tile < 3 * CT; tile += 4) {
      const int d = tile / CT, col = (tile - d * CT) * 16 + lc;
      const V4 B = *reinterpret_cast<const V4*>(&s_b[d][col * BS + kb]);
      const f32x4 C = OP::mma(A, B, f32x4{0.f, 0.f, 0.f, 0.f});
      *reinterpret_cast<f32x4*>(&s_c[d][col * FS + kb]) = C;
    }
  };
  __syncthreads();
#pragma unroll 1
  for (int f = 0; f < NF; ++f) {
    const float* qf = q + (int64_t)f * offset + e * P;
    float qv[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int n = t + 256 * u;
      qv[u] = n < P ? ldnt(&qf[n]) : 0.f;
      if (n < P) {
        const H h = OP::round(qv[u]);
        s_b[0][pb[u][0]] = h;
        s_b[1][pb[u][1]] = h;
        s_b[2][pb[u][2]] = h;
      }
    }
    __syncthreads();
    products(Af);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      if (t + 256 * u >= P) continue;
      const float qr = s_c[0][pc[u][0]], qs = s_c[1][pc[u][1]], qt = s_c[2][pc[u][2]];
      float Gr = __builtin_fmaf(gf[u][2], qt, __builtin_fmaf(gf[u][1], qs, gf[u][0] * qr));
      float Gs = __builtin_fmaf(gf[u][4], qt, __builtin_fmaf(gf[u][3], qs, gf[u][1] * qr));
      float Gt = __builtin_fmaf(gf[u][5], qt, __builtin_fmaf(gf[u][4], qs, gf[u][2] * qr));
      if (!BK) {
        Gr = l0[u] * Gr;
        Gs = l0[u] * Gs;
        Gt = l0[u] * Gt;
      }
      s_b[0][pb[u][0]] = OP::round(Gr);
      s_b[1][pb[u][1]] = OP::round(Gs);
      s_b[2][pb[u][2]] = OP::round(Gt);
    }
    __syncthreads();
    products(At);
    __syncthreads();
    float* af = Aq + (int64_t)f * offset + e * P;
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int n = t + 256 * u;
      if (n >= P) continue;
      float res = (s_c[0][pc[u][0]] + s_c[1][pc[u][1]]) + s_c[2][pc[u][2]];
      if (!BK) res = __builtin_fmaf(wl1[u], qv[u], res);
      stnt(res, &af[n]);
    }
    // the next field's staging writes s_b only; its first barrier orders
    // these reads of s_c before its products overwrite them
  }
}

int axhelm_mma_launch(int N, int Nf, int bk, bool bf16_geometry, bool f16_operands, int32_t Nelements,
                      int32_t offset, const void* ggeo, const void* D, const void* lambda, const void* q, void* Aq,
                      const int* skip, hipStream_t s)
{
  const char* name = f16_operands ? "axhelm fp16-MFMA" : "axhelm bf16-MFMA";
  if (!((N == 7 || N == 9) && (Nf == 1 || Nf == 3)))
    return fail(NBX_ERR_UNSUPPORTED, "%s: N = 7 or 9, one or three fields", name);
  if (f16_operands && bf16_geometry) return fail(NBX_ERR_UNSUPPORTED, "%s: FP32 geometry only", name);
  if (Nelements <= 0) return Nelements == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "axhelm: Nelements < 0");
  if (Nf == 3 && offset < Nelements * (N + 1) * (N + 1) * (N + 1))
    return fail(NBX_ERR_INVALID, "axhelm: offset < Nelements * Np");
  if (!(N == 7 && Nf == 1)) {
#define NBX_MMA16(NQV, NFV, BKV, GT, OPT)                                                                     \
  hipLaunchKernelGGL((k_axhelm_mma16<NQV, NFV, BKV, GT, OPT>), dim3((unsigned)Nelements), dim3(256), 0, s,   \
                     Nelements, offset, (const GT*)ggeo, (const float*)D, (const GT*)lambda, (const float*)q, \
                     (float*)Aq, skip)
#define NBX_MMA16_BK(NQV, NFV, GT, OPT) \
  if (bk) NBX_MMA16(NQV, NFV, true, GT, OPT); else NBX_MMA16(NQV, NFV, false, GT, OPT);
#define NBX_MMA16_T(NQV, NFV)                                   \
  if (f16_operands) { NBX_MMA16_BK(NQV, NFV, float, OpF16) }     \
  else if (bf16_geometry) { NBX_MMA16_BK(NQV, NFV, uint16_t, OpBf16) } \
  else { NBX_MMA16_BK(NQV, NFV, float, OpBf16) }
    if (N == 7) { NBX_MMA16_T(8, 3) }
    else if (Nf == 1) { NBX_MMA16_T(10, 1) }
    else { NBX_MMA16_T(10, 3) }
#undef NBX_MMA16_T
#undef NBX_MMA16_BK
#undef NBX_MMA16
    NBX_HIP_LAUNCH();
    return NBX_SUCCESS;
  }
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

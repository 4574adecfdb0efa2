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
// loads.  Residual taps: the element's r^, rounded to bf16 once per node, sits
// in a zero-padded 10^3 tile in LDS, so a tap is a fixed offset from the node;
// the geometric factors' bf16 bits follow it, so the first layer's operands
// are LDS reads with no conversion.  Every operand read, T store and output tap is an
// unconditional LDS access at a selected address (no divergent branches in
// the inner loops).  A 512-lane workgroup computes one element at a time,
// wave w taking z-plane w (four tiles); the grid holds as many workgroups as
// the device runs at once, each looping over its elements and loading the
// next element's inputs while it computes the current one.  The 27-tap output
// sum y_n = b3 + sum_tap T[tap][n + tap] (taps outside the element add zero)
// and z = w invDiag (r + s y) are the FP32 kernel's.
#include <algorithm>

#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kP = 512, kTaps = 27, kPad = 10, kTRow = 516;  // T row stride: groups of 4 rows 16 banks apart
constexpr int kPadN = kPad * kPad * kPad;  // the zero-padded residual tile; the 7 factors follow it in LDS
constexpr int kFRow = 800;                 // a factor row: node (x, y, z) at 100 z + 10 y + x, as in the tile
constexpr int kTRows = kTaps + 1;          // row 27 takes the product of the padding tap 27
constexpr int kTGuard = 80;                // floats before the first and after the last T row (> 64 + 8 + 1)
constexpr int kTSize = kTGuard + kTRows * kTRow + kTGuard;

// one lane's fragments (144 bytes)
struct alignas(16) LaneFrag {
  s16x8 a1[2];   // layer 1, features 0..31: taps 0..26, factors 0..4 (slot j of group g: feature 8g + j)
  s16x8 a2[2];   // layer 2, input channel order (4g + j | 16 + 4g + j - 4)
  s16x8 a3[2];   // layer 3, rows = taps
  s16x8 a1t[2];  // layer 1 tail (a second 16x16x32 product): group 0's slots 0, 1, 2 are factors 5, 6 and
                 // the bias (times a constant-one input; bf16 parameters, so exact), every other slot zero
  float b2[2][4];
  short off[8];  // LDS place of the layer-1 operand of slot j, relative to the node's place 100 z + 10 y + x
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
    for (int j = 0; j < 8; ++j) {  // group 0: factors 5, 6, the bias
      const short v = j < 2 ? bfs(W[kOffWg + (5 + j) * 32 + c]) : j == 2 ? bfs(W[kOffB1 + c]) : (short)0;
      f.a1t[mt][j] = g == 0 ? v : (short)0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) f.b2[mt][j] = W[kOffB2 + 16 * mt + 4 * g + j];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {  // feature 8g + j: a residual tap (the node sits at 111 in the tile) or a factor
    const int ft = 8 * g + j;
    f.off[j] = (short)(ft < kTaps ? 111 + (ft / 9 - 1) * 100 + ((ft / 3) % 3 - 1) * 10 + (ft % 3 - 1)
                                  : kPadN + (ft - kTaps) * kFRow);
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

// the next layer's B operand from two output tiles: bf16 (two values per
// conversion), then relu on the bf16 bits as 16-bit integers (a negative
// value has the sign bit, so max(bits, 0) is relu; rounding keeps the sign,
// so this equals relu before rounding for every value that is not a NaN);
// slots 0..3 from the first tile, 4..7 from the second
__device__ __forceinline__ s16x8 relu_bf16(f32x4 c0, f32x4 c1)
{
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 p = {bf16x2_bits_hw(c0[0], c0[1]), bf16x2_bits_hw(c0[2], c0[3]), bf16x2_bits_hw(c1[0], c1[1]),
                   bf16x2_bits_hw(c1[2], c1[3])};
  return __builtin_elementwise_max(__builtin_bit_cast(s16x8, p), s16x8{0, 0, 0, 0, 0, 0, 0, 0});
}

// one step of a wave maximum over 64-bit values: each lane takes the larger of
// its value and the value the DPP pattern brings it (0 where it brings none)
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint64_t dpp_max_step(uint64_t v)
{
  const uint32_t lo = __builtin_amdgcn_update_dpp(0u, (uint32_t)v, kCtrl, kRowMask, 0xf, true);
  const uint32_t hi = __builtin_amdgcn_update_dpp(0u, (uint32_t)(v >> 32), kCtrl, kRowMask, 0xf, true);
  const uint64_t d = ((uint64_t)hi << 32) | lo;
  return d > v ? d : v;
}

// maximum over the wave's 64 lanes, in every lane: row shifts by 1, 2, 4 and
// 8 leave each row's maximum in its lane 15, the row broadcasts 15 and 31
// carry it on to lane 63, which is read back (no LDS round trips)
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v)
{
  v = dpp_max_step<0x111, 0xf>(v);  // row_shr:1
  v = dpp_max_step<0x112, 0xf>(v);  // row_shr:2
  v = dpp_max_step<0x114, 0xf>(v);  // row_shr:4
  v = dpp_max_step<0x118, 0xf>(v);  // row_shr:8
  v = dpp_max_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_max_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, 63), hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// 1 / d in double for a node degree d (the quotient the division gives,
// rounded at compile time): a table read instead of a division per node
struct InvDegreeTable {
  double v[256];
  constexpr InvDegreeTable() : v()
  {
    for (int d = 1; d < 256; ++d) v[d] = 1.0 / d;
  }
};
__constant__ InvDegreeTable kInvDegree = InvDegreeTable();

// RES: the input residual is (double)(res - Ad) of two FP32 vectors (the
// learned multigrid step's fine residual, formed here instead of in a pass
// of its own; the same value), otherwise the FP64 vector r
template <bool RES>
__global__ __launch_bounds__(512) void k_net_apply_mma(const int32_t E, const LaneFrag* __restrict__ frag,
                                                       const float* __restrict__ W,
                                                       const uint16_t* __restrict__ feat16,
                                                       const double* __restrict__ r,
                                                       const float* __restrict__ res, const float* __restrict__ Ad,
                                                       const double* __restrict__ invDiag,
                                                       const double* __restrict__ w, const uint8_t* __restrict__ deg,
                                                       double* __restrict__ zl, const int* __restrict__ fallback,
                                                       const int* __restrict__ skip)
{
  if (skip && *skip) return;
  auto rload = [&](int64_t n) { return RES ? (double)(res[n] - Ad[n]) : r[n]; };
  // one LDS block, the operand tile first, so that every read and store below
  // is a lane address plus an offset that fits the 16-bit LDS offset field
  struct Lds {
    uint16_t in[kPadN + 7 * kFRow];  // bf16 bits: r^ in a zero-padded 10^3 tile, then the 7 factors
    float T[kTSize];
    uint64_t red[8];
  };
  __shared__ Lds lds;
  uint16_t* const s_in = lds.in;
  float* const s_T = lds.T;
  const int t = threadIdx.x, l = t & 63, wv = t >> 6, g = l >> 4, col = l & 15;
  const int i = t & 7, j = (t >> 3) & 7, k = t >> 6;  // node t = (i, j, k)
  const int post = k * 100 + j * 10 + i;               // its place in the 10-strided layout
  if (fallback && *fallback) {  // Jacobi for every element of this workgroup
    for (int64_t e = blockIdx.x; e < E; e += gridDim.x) {
      const int64_t id = e * kP + t;
      const double wn = deg ? kInvDegree.v[deg[id]] : (w ? w[id] : 1.0);
      zl[id] = wn * (invDiag[id] * rload(id));
    }
    return;
  }
  // this lane's fragments, read once for all the workgroup's elements
  const LaneFrag f = frag[l];
  // the lane's node in tile tt is nd = 64 wv + 16 tt + col = (col & 7, 2 tt + (col >> 3), wv): its place
  // moves by 20 per tile in the tile and the factor rows alike, so every operand read below is one LDS
  // read at the lane's address plus a constant
  const int pos0 = wv * 100 + (col >> 3) * 10 + (col & 7);
  int adr[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) adr[q] = pos0 + f.off[q];
  const int tb0 = kTGuard + 4 * g * kTRow + wv * 64 + col;         // T rows 4g.. (first output tile)
  const int tb1 = kTGuard + (16 + 4 * g) * kTRow + wv * 64 + col;  // T rows 16 + 4g.. (second)
  const bool ox[3] = {i > 0, true, i < 7}, oy[3] = {j > 0, true, j < 7}, oz[3] = {k > 0, true, k < 7};
  const float b3 = W[kOffB3];
  // workgroup b takes elements b, b + G, b + 2G, ... (G = the grid); the
  // next element's inputs are loaded, reduced and staged while the current
  // one is computed, so each element takes two barriers
  int64_t e = blockIdx.x;
  if (e >= E) return;
  for (int q = t; q < kPadN; q += kP) s_in[q] = 0;  // the tile's border, once: elements rewrite the interior only
  // element scale s = max |r| (exact), as the largest bit pattern of |r|:
  // the bits of non-negative doubles order as their values
  auto wave_max = [](double v) { return wave_max_u64(__builtin_bit_cast(uint64_t, fabs(v))); };
  auto block_max = [&]() {  // lanes 0..7 of each wave hold the 8 wave maxima, the rest 0
    uint64_t m = l < 8 ? lds.red[l] : 0u;
    m = dpp_max_step<0x111, 0xf>(m);  // row_shr:1
    m = dpp_max_step<0x112, 0xf>(m);  // row_shr:2
    m = dpp_max_step<0x114, 0xf>(m);  // row_shr:4: lane 7 has all 8
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)m, 7), hi = __builtin_amdgcn_readlane((uint32_t)(m >> 32), 7);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  };
  // r^ = r / s rounded to bf16 once per node, and the factors, into the operand tile
  auto stage = [&](double rv, double sv, const uint16_t* fv) {
    const double inv = sv > 0.0 ? 1.0 / sv : 0.0;
    s_in[111 + post] = bf16_bits_hw((float)(rv * inv));
#pragma unroll
    for (int q = 0; q < 7; ++q) s_in[kPadN + q * kFRow + post] = fv[q];
  };
  double rn = rload(e * kP + t), di = invDiag[e * kP + t], wraw = w && !deg ? w[e * kP + t] : 1.0;
  uint32_t dg = deg ? deg[e * kP + t] : 0u;
  uint16_t fe[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) fe[q] = feat16[(e * 7 + q) * kP + t];
  {
    const uint64_t m = wave_max(rn);
    if (l == 0) lds.red[wv] = m;
  }
  __syncthreads();
  double s = block_max();
  stage(rn, s, fe);
  for (;;) {
    const int64_t id = e * kP + t;
    const double wn = deg ? kInvDegree.v[dg] : wraw;
    const int64_t en = e + gridDim.x;
    const bool more = en < E;
    // (past the last element the loads repeat this one: no branch, no use)
    const int64_t ef = more ? en : e;
    const double* const dp = invDiag + ef * kP;  // uniform bases, lane offsets fixed
    const uint16_t* const fp = feat16 + ef * 7 * kP;
    const double rn_n = rload(ef * kP + t), di_n = dp[t];
    const uint32_t dg_n = deg ? deg[ef * kP + t] : 0u;
    const double wraw_n = w && !deg ? w[ef * kP + t] : 1.0;
    uint16_t fe_n[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) fe_n[q] = fp[q * kP + t];
    __syncthreads();  // the operand tile holds this element; the previous element's gather is done
#pragma unroll
  for (int tt = 0; tt < 4; ++tt) {
    s16x8 b1;
#pragma unroll
    for (int q = 0; q < 8; ++q) b1[q] = (short)s_in[adr[q] + 20 * tt];
    // the tail's B operand: slots 0, 1, 2 hold factors 5, 6 and bf16 1.0 (the
    // A operand is zero past group 0's three slots, so the other lanes' values
    // do not count); both layer-1 products have the one 16x16x32 shape, so the
    // second may take the first's result as its accumulator
    s16x8 b1t = {(short)s_in[kPadN + 5 * kFRow + pos0 + 20 * tt], (short)s_in[kPadN + 6 * kFRow + pos0 + 20 * tt],
                 (short)0x3F80, 0, 0, 0, 0, 0};
    f32x4 C[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a1t[mt]), as_bf(b1t), f32x4{0.f, 0.f, 0.f, 0.f}, 0,
                                                      0, 0);
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a1[mt]), as_bf(b1), C[mt], 0, 0, 0);
    }
    const s16x8 b2 = relu_bf16(C[0], C[1]);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      C[mt] = f32x4{f.b2[mt][0], f.b2[mt][1], f.b2[mt][2], f.b2[mt][3]};
      C[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a2[mt]), as_bf(b2), C[mt], 0, 0, 0);
    }
    const s16x8 b3 = relu_bf16(C[0], C[1]);
    const f32x4 T0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a3[0]), as_bf(b3), f32x4{0.f, 0.f, 0.f, 0.f},
                                                             0, 0, 0);
    const f32x4 T1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(f.a3[1]), as_bf(b3), f32x4{0.f, 0.f, 0.f, 0.f},
                                                             0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) s_T[tb0 + q * kTRow + 16 * tt] = T0[q];  // taps 4g + q
    if (g < 3) {  // taps 16 + 4g + q (27 to the spare row); group 3's rows 28..31 are zero weights
#pragma unroll
      for (int q = 0; q < 4; ++q) s_T[tb1 + q * kTRow + 16 * tt] = T1[q];
    }
    if (tt == 2) {  // the next element's wave maxima, among this element's products
      const uint64_t m = wave_max(rn_n);
      if (l == 0) lds.red[wv] = m;
    }
  }
    __syncthreads();  // T is complete, and the next element's wave maxima are in place
    // the tile loop has read the operand tile for the last time: stage the next element
    const double s_n = more ? block_max() : 0.0;
    if (more) stage(rn_n, s_n, fe_n);
    // y_n = b3 + sum over the taps whose node n + tap is in the element of T[tap][n + tap]: every read
    // is at node t's address plus a constant and stays inside s_T (the guards cover the reach 64 + 8 + 1
    // of row 0 and row 26); the reads of taps that leave the element are discarded
    // (the zero padding is a bit mask on the value, not a select: a select of
    // a load can become a branch around the load, one LDS round trip per tap)
    float y = b3;
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
      const float v = s_T[kTGuard + tap * kTRow + t + (dx - 1) + 8 * (dy - 1) + 64 * (dz - 1)];
      const uint32_t keep = (ox[dx] && oy[dy] && oz[dz]) ? 0xFFFFFFFFu : 0u;
      y = y + __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & keep);
    }
    zl[id] = wn * (di * (rn + s * (double)y));
    if (!more) break;
    e = en;
    s = s_n;
    rn = rn_n;
    di = di_n;
    wraw = wraw_n;
    dg = dg_n;
    __builtin_memcpy(fe, fe_n, sizeof fe);
  }
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

template <bool RES>
static int net_apply_mma_launch_t(int32_t E, const void* frag, const float* W, const uint16_t* feat16,
                                  const double* r, const float* res, const float* Ad, const double* invDiag,
                                  const double* w, const uint8_t* deg, double* zl, const int* fallback,
                                  const int* skip, hipStream_t s)
{
  if (E <= 0) return NBX_SUCCESS;
  // as many workgroups as the device holds at once (each loops over its
  // elements), found once per device and kernel form
  static int resident[64] = {};
  int dev = 0;
  NBX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(NBX_ERR_UNSUPPORTED, "net_apply: device %d", dev);
  if (!resident[dev]) {
    int per_cu = 0, cus = 0;
    NBX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_net_apply_mma<RES>, kP, 0));
    NBX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    resident[dev] = std::max(1, per_cu) * std::max(1, cus);
  }
  const int grid = std::min<int64_t>(E, resident[dev]);
  hipLaunchKernelGGL(k_net_apply_mma<RES>, dim3(grid), dim3(kP), 0, s, E, static_cast<const LaneFrag*>(frag), W,
                     feat16, r, res, Ad, invDiag, w, deg, zl, fallback, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_apply_mma_launch(int32_t E, const void* frag, const float* W, const uint16_t* feat16, const double* r,
                         const double* invDiag, const double* w, const uint8_t* deg, double* zl,
                         const int* fallback, const int* skip, hipStream_t s)
{
  return net_apply_mma_launch_t<false>(E, frag, W, feat16, r, nullptr, nullptr, invDiag, w, deg, zl, fallback, skip,
                                       s);
}

int net_apply_mma_resid_launch(int32_t E, const void* frag, const float* W, const uint16_t* feat16,
                               const float* res, const float* Ad, const double* invDiag, const double* w,
                               const uint8_t* deg, double* zl, const int* skip, hipStream_t s)
{
  return net_apply_mma_launch_t<true>(E, frag, W, feat16, nullptr, res, Ad, invDiag, w, deg, zl, nullptr, skip, s);
}

}  // namespace nbx

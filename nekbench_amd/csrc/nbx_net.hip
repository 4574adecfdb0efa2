// Learned element preconditioner: a small 3-D convolutional network over the
// 8x8x8 GLL nodes of one element, inputs the residual and the 7 geometric
// factors, output a correction to Jacobi:
//
//   s      = max_n |r_n|                     (element scale; exact)
//   rh_n   = (float)(r_n / s)                (residual channel)
//   h1     = relu(b1 + Wg . g_n + sum_tap Wr[tap] rh_{n+tap})       32 channels
//   h2     = relu(b2 + W2^T h1)                                        32 channels
//   y_n    = b3 + sum_tap sum_c W3[tap][c] h2_{n+tap}[c]               1 channel
//   z_n    = invDiag_n (r_n + s y_n)          (assembled afterwards by gs)
//
// taps are the 27 neighbours (dz,dy,dx) in {-1,0,1}^3, index
// (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding at the element boundary.  With
// f = 0 the preconditioner is Jacobi exactly.  Weights are bf16 values (held
// in float), arithmetic FP32 with explicit fma, every sum in a fixed order;
// the CPU oracle restates the same order, so forward values and parameter
// gradients are reproduced bit for bit.
//
// One workgroup of 512 lanes per element, lane = node.  The last layer is
// evaluated as T[n][tap] = W3[tap] . h2_n (lane-local) followed by
// y_m = b3 + sum_tap T[m+tap][tap] (27 LDS reads), which keeps LDS traffic at
// 27 values per node.  Weight operands are uniform across the workgroup and
// come through scalar loads.
#include "nbx_internal.h"
#include "nbx_net.h"
#include <cmath>
#include <cstring>
#include <vector>

namespace nbx {

constexpr int kP = 512, kC = 32, kG = 7, kTaps = 27;
// row stride of the [channel][node] LDS tiles of the training kernel: odd, so
// the 32 lanes of a parameter gradient that read 32 channel rows at the same
// node hit 32 different banks
constexpr int kRow = kP + 1;

__device__ __forceinline__ int tap_neighbor(int n, int tap)
{
  const int i = n & 7, j = (n >> 3) & 7, k = n >> 6;
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int ii = i + dx, jj = j + dy, kk = k + dz;
  if ((unsigned)ii > 7u || (unsigned)jj > 7u || (unsigned)kk > 7u) return -1;
  return ii + (jj << 3) + (kk << 6);
}

struct NetFwd {
  float h1[kC];
  float h2[kC];
};

// forward of one node up to T (lane-local); rh in LDS; returns T[27] in T
__device__ __forceinline__ void net_forward_lane(const float* __restrict__ W, int n, const float* s_rh,
                                                 const float (&g)[kG], NetFwd& a, float (&T)[kTaps])
{
  const float* Wr = W + kOffWr;
  const float* Wg = W + kOffWg;
  const float* b1 = W + kOffB1;
  const float* W2 = W + kOffW2;
  const float* b2 = W + kOffB2;
  const float* W3 = W + kOffW3;
#pragma unroll
  for (int c = 0; c < kC; ++c) a.h1[c] = b1[c];
#pragma unroll
  for (int q = 0; q < kG; ++q)
#pragma unroll
    for (int c = 0; c < kC; ++c) a.h1[c] = __builtin_fmaf(Wg[q * kC + c], g[q], a.h1[c]);
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int nb = tap_neighbor(n, tap);
    if (nb >= 0) {
      const float v = s_rh[nb];
#pragma unroll
      for (int c = 0; c < kC; ++c) a.h1[c] = __builtin_fmaf(Wr[tap * kC + c], v, a.h1[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < kC; ++c) a.h1[c] = a.h1[c] > 0.f ? a.h1[c] : 0.f;
#pragma unroll
  for (int c = 0; c < kC; ++c) a.h2[c] = b2[c];
#pragma unroll
  for (int ci = 0; ci < kC; ++ci)
#pragma unroll
    for (int c = 0; c < kC; ++c) a.h2[c] = __builtin_fmaf(W2[ci * kC + c], a.h1[ci], a.h2[c]);
#pragma unroll
  for (int c = 0; c < kC; ++c) a.h2[c] = a.h2[c] > 0.f ? a.h2[c] : 0.f;
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    float t = 0.f;
#pragma unroll
    for (int c = 0; c < kC; ++c) t = __builtin_fmaf(W3[tap * kC + c], a.h2[c], t);
    T[tap] = t;
  }
}

// element scale s = max |r| over the element (block reduction, exact)
__device__ __forceinline__ double element_scale(double rn, double* s_red)
{
  double m = fabs(rn);
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) m = fmax(m, __shfl_xor(m, h, 64));
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  double s = s_red[0];
  for (int w = 1; w < kP / 64; ++w) s = fmax(s, s_red[w]);
  return s;
}

// Inference: zl[e*512+n] = w (invDiag (r + s y)), w the assembly weight
// (1/multiplicity, or 1 when w is null); fallback (device flag) turns the
// network off and leaves Jacobi.  Optionally skipped by *skip.
__global__ __launch_bounds__(512) void k_net_apply(const int32_t E, const float* __restrict__ W,
                                                   const float* __restrict__ feat, const double* __restrict__ r,
                                                   const double* __restrict__ invDiag, const double* __restrict__ w,
                                                   double* __restrict__ zl, const int* __restrict__ fallback,
                                                   const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ float s_rh[kP];
  __shared__ float s_T[kTaps][kP];
  __shared__ double s_red[kP / 64];
  const int n = threadIdx.x;
  const int64_t e = blockIdx.x;
  const int64_t id = e * kP + n;
  const double rn = r[id];
  const double di = invDiag[id];
  const double wn = w ? w[id] : 1.0;
  if (fallback && *fallback) {
    zl[id] = wn * (di * rn);
    return;
  }
  const double s = element_scale(rn, s_red);
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  s_rh[n] = (float)(rn * inv);
  float g[kG];
#pragma unroll
  for (int q = 0; q < kG; ++q) g[q] = feat[(e * kG + q) * kP + n];
  __syncthreads();
  NetFwd a;
  float T[kTaps];
  net_forward_lane(W, n, s_rh, g, a, T);
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) s_T[tap][n] = T[tap];
  __syncthreads();
  float y = W[kOffB3];
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int nb = tap_neighbor(n, tap);
    if (nb >= 0) y = y + s_T[tap][nb];
  }
  zl[id] = wn * (di * (rn + s * (double)y));
}

// Training: each workgroup runs a chunk of (pair, element) samples through
// forward and backward and accumulates the parameter gradient and the loss
//   L = sum_samples inv_norm[pair] sum_n (z_n - e_n)^2
// in lane-owned registers (parameter q is owned by lane q mod 512 of its
// tensor); one partial per workgroup goes to partials[block][kNetParams+1].
// With G non-null the loss is defined outside (energy norm of the assembled
// error) and G[e*512+n] holds its derivative with respect to the unweighted
// element output invDiag (r + s y) at local node n of element e; the kernel
// then back-propagates dL/dy = G invDiag s and reports no loss.
__global__ __launch_bounds__(512) void k_net_train(const int32_t nsamples, const int32_t chunk,
                                                   const int2* __restrict__ samples, const int64_t pairStride,
                                                   const float* __restrict__ R, const float* __restrict__ Ev,
                                                   const double* __restrict__ invNorm, const float* __restrict__ W,
                                                   const float* __restrict__ feat, const double* __restrict__ invDiag,
                                                   const double* __restrict__ G, float* __restrict__ partials)
{
  extern __shared__ float lds[];
  float* R1 = lds;                    // [32][kRow]: h2 -> h1 -> dpre1
  float* R2 = R1 + kC * kRow;         // [32][kRow]: T / dT -> dpre2
  float* s_rh = R2 + kC * kRow;       // [512]
  float* s_g = s_rh + kP;             // [7][512]
  float* s_dy = s_g + kG * kP;        // [512]
  double* s_red = (double*)(s_dy + kP);  // [8]
  const int n = threadIdx.x;
  const float* W3 = W + kOffW3;
  const float* W2 = W + kOffW2;
  float gW3[2] = {0.f, 0.f}, gW2[2] = {0.f, 0.f}, gWr[2] = {0.f, 0.f}, gWg = 0.f, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f;
  double loss = 0.0;
  const int32_t s0 = blockIdx.x * chunk, s1 = min(nsamples, s0 + chunk);
  for (int32_t si = s0; si < s1; ++si) {
    const int2 pe = samples[si];
    const int64_t e = pe.y;
    const int64_t off = (int64_t)pe.x * pairStride + e * kP + n;
    const double rn = (double)R[off];
    const double en = G ? 0.0 : (double)Ev[off];
    const double di = invDiag[e * kP + n];
    __syncthreads();  // previous sample done with LDS
    const double s = element_scale(rn, s_red);
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    s_rh[n] = (float)(rn * inv);
    float g[kG];
#pragma unroll
    for (int q = 0; q < kG; ++q) {
      g[q] = feat[(e * kG + q) * kP + n];
      s_g[q * kP + n] = g[q];
    }
    __syncthreads();
    NetFwd a;
    float T[kTaps];
    net_forward_lane(W, n, s_rh, g, a, T);
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) R2[tap * kRow + n] = T[tap];
    __syncthreads();
    float y = W[kOffB3];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
      const int nb = tap_neighbor(n, tap);
      if (nb >= 0) y = y + R2[tap * kRow + nb];
    }
    const double z = di * (rn + s * (double)y);
    const double res = z - en;
    const double wn = G ? 0.0 : invNorm[pe.x];
    // loss: lane values, fixed tree (per-wave halving, waves in order)
    if (!G) {
      double l = res * res * wn;
#pragma unroll
      for (int h = 32; h >= 1; h >>= 1) l += __shfl_xor(l, h, 64);
      __syncthreads();
      if ((n & 63) == 0) s_red[n >> 6] = l;
      __syncthreads();
      if (n == 0) {
        double t = s_red[0];
        for (int w = 1; w < kP / 64; ++w) t = t + s_red[w];
        loss = loss + t;
      }
    }
    const float dy = G ? (float)(G[e * kP + n] * di * s) : (float)(2.0 * res * wn * di * s);
    s_dy[n] = dy;
    __syncthreads();
    // dT[tap] = dy at the output node m = n - tap
    float dT[kTaps];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
      const int m = tap_neighbor(n, 26 - tap);  // n - delta(tap)
      dT[tap] = m >= 0 ? s_dy[m] : 0.f;
    }
    float dpre2[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      float acc = 0.f;
#pragma unroll
      for (int tap = 0; tap < kTaps; ++tap) acc = __builtin_fmaf(W3[tap * kC + c], dT[tap], acc);
      dpre2[c] = a.h2[c] > 0.f ? acc : 0.f;
    }
    __syncthreads();  // everyone has read T from R2
#pragma unroll
    for (int c = 0; c < kC; ++c) R1[c * kRow + n] = a.h2[c];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) R2[tap * kRow + n] = dT[tap];
    __syncthreads();
    // dW3[tap][c] = sum_n h2[n][c] dT[n][tap]
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = n + u * kP;
      if (q < kTaps * kC) {
        const int tap = q / kC, c = q % kC;
        float acc = 0.f;
        for (int m = 0; m < kP; ++m) acc = __builtin_fmaf(R1[c * kRow + m], R2[tap * kRow + m], acc);
        gW3[u] = gW3[u] + acc;
      }
    }
    if (n == 0) {
      float acc = 0.f;
      for (int m = 0; m < kP; ++m) acc = acc + s_dy[m];
      gb3 = gb3 + acc;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      R1[c * kRow + n] = a.h1[c];
      R2[c * kRow + n] = dpre2[c];
    }
    __syncthreads();
    // dW2[ci][c] = sum_n h1[n][ci] dpre2[n][c]; db2[c] = sum_n dpre2[n][c]
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = n + u * kP;
      const int ci = q / kC, c = q % kC;
      float acc = 0.f;
      for (int m = 0; m < kP; ++m) acc = __builtin_fmaf(R1[ci * kRow + m], R2[c * kRow + m], acc);
      gW2[u] = gW2[u] + acc;
    }
    if (n < kC) {
      float acc = 0.f;
      for (int m = 0; m < kP; ++m) acc = acc + R2[n * kRow + m];
      gb2 = gb2 + acc;
    }
    // dpre1 = relu'(h1) (W2 dpre2)
    float dpre1[kC];
#pragma unroll
    for (int ci = 0; ci < kC; ++ci) {
      float acc = 0.f;
#pragma unroll
      for (int c = 0; c < kC; ++c) acc = __builtin_fmaf(W2[ci * kC + c], dpre2[c], acc);
      dpre1[ci] = a.h1[ci] > 0.f ? acc : 0.f;
    }
    __syncthreads();  // R1 (h1) reads done
#pragma unroll
    for (int c = 0; c < kC; ++c) R1[c * kRow + n] = dpre1[c];
    __syncthreads();
    // dWr[tap][c] = sum_n dpre1[n][c] rh[n+tap]; dWg[q][c] = sum_n dpre1[n][c] g[n][q]; db1
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int q = n + u * kP;
      if (q < kTaps * kC) {
        const int tap = q / kC, c = q % kC;
        float acc = 0.f;
        for (int m = 0; m < kP; ++m) {
          const int nb = tap_neighbor(m, tap);
          if (nb >= 0) acc = __builtin_fmaf(R1[c * kRow + m], s_rh[nb], acc);
        }
        gWr[u] = gWr[u] + acc;
      }
    }
    if (n < kG * kC) {
      const int q = n / kC, c = n % kC;
      float acc = 0.f;
      for (int m = 0; m < kP; ++m) acc = __builtin_fmaf(R1[c * kRow + m], s_g[q * kP + m], acc);
      gWg = gWg + acc;
    }
    if (n < kC) {
      float acc = 0.f;
      for (int m = 0; m < kP; ++m) acc = acc + R1[n * kRow + m];
      gb1 = gb1 + acc;
    }
  }
  float* out = partials + (int64_t)blockIdx.x * kNetPartial;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = n + u * kP;
    if (q < kTaps * kC) {
      out[kOffW3 + q] = gW3[u];
      out[kOffWr + q] = gWr[u];
    }
    out[kOffW2 + q] = gW2[u];
  }
  if (n < kG * kC) out[kOffWg + n] = gWg;
  if (n < kC) {
    out[kOffB1 + n] = gb1;
    out[kOffB2 + n] = gb2;
  }
  if (n == 0) {
    out[kOffB3] = gb3;
    reinterpret_cast<double*>(out + kNetParamsPadded)[0] = loss;
  }
}

// grad[q] = sum over workgroups (in order) of partials[b][q]; loss likewise
__global__ __launch_bounds__(256) void k_net_grad_sum(const int32_t nblocks, const float* __restrict__ partials,
                                                      float* __restrict__ grad, double* __restrict__ loss)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < kNetParams) {
    float acc = 0.f;
    for (int b = 0; b < nblocks; ++b) acc = acc + partials[(int64_t)b * kNetPartial + q];
    grad[q] = acc;
  }
  if (q == 0) {
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b)
      acc = acc + reinterpret_cast<const double*>(partials + (int64_t)b * kNetPartial + kNetParamsPadded)[0];
    *loss = acc;
  }
}

__device__ __forceinline__ float round_bf16(float x) { return bf16_round(x); }

// Adam on FP32 master weights; the forward weights are their bf16 rounding
__global__ __launch_bounds__(256) void k_net_adam(const int32_t np, float* __restrict__ master,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ grad, float* __restrict__ wbf,
                                                  const float lr, const float b1, const float b2, const float eps,
                                                  const float c1, const float c2)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const float g = grad[q];
  const float mq = b1 * m[q] + (1.f - b1) * g;
  const float vq = b2 * v[q] + (1.f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  const float upd = lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  const float w = master[q] - upd;
  master[q] = w;
  wbf[q] = round_bf16(w);
}

// normalised geometric features: feat[e][q][n] = ggeo[e][q][n] / scale[q]
__global__ __launch_bounds__(256) void k_net_features(const int64_t n, const double* __restrict__ ggeo,
                                                      const double* __restrict__ invScale, float* __restrict__ feat)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int q = (int)((i / kP) % kG);
  feat[i] = (float)(ggeo[i] * invScale[q]);
}

int net_apply_launch(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                     const double* w, double* zl, const int* fallback, const int* skip, hipStream_t s)
{
  hipLaunchKernelGGL(k_net_apply, dim3(E), dim3(kP), 0, s, E, W, feat, r, invDiag, w, zl, fallback, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

size_t net_train_lds_bytes()
{
  return sizeof(float) * (2 * kC * kRow + kP + kG * kP + kP) + sizeof(double) * (kP / 64);
}

int net_train_launch(int32_t nsamples, int32_t chunk, const int2* samples, int64_t pairStride, const float* R,
                     const float* Ev, const double* invNorm, const float* W, const float* feat, const double* invDiag,
                     const double* G, float* partials, hipStream_t s)
{
  const int nblocks = (nsamples + chunk - 1) / chunk;
  const size_t lds = net_train_lds_bytes();
  static bool attr_set = false;
  if (!attr_set) {
    NBX_HIP(hipFuncSetAttribute((const void*)k_net_train, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_net_train, dim3(nblocks), dim3(kP), lds, s, nsamples, chunk, samples, pairStride, R, Ev,
                     invNorm, W, feat, invDiag, G, partials);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_grad_sum_launch(int32_t nblocks, const float* partials, float* grad, double* loss, hipStream_t s)
{
  hipLaunchKernelGGL(k_net_grad_sum, dim3((kNetParams + 255) / 256), dim3(256), 0, s, nblocks, partials, grad, loss);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_adam_launch(float* master, float* m, float* v, const float* grad, float* wbf, float lr, int step,
                    hipStream_t s)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_net_adam, dim3((kNetParams + 255) / 256), dim3(256), 0, s, kNetParams, master, m, v, grad, wbf,
                     lr, b1, b2, eps, c1, c2);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int net_features_launch(int32_t E, const double* ggeo, const double* invScale, float* feat, hipStream_t s)
{
  const int64_t n = (int64_t)E * kG * kP;
  hipLaunchKernelGGL(k_net_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, ggeo, invScale, feat);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

}  // namespace nbx

// Deterministic block reductions shared by the BLAS-1 and PCG kernels.
//
// Order (fixed, independent of timing and of the number of CUs):
//   tile b of 256 lanes x 8 values; lane t takes b*2048 + v*256 + t, v = 0..7,
//   accumulating with an fma chain in v order;
//   256-lane tree: halving tree inside each wave64 (lane 0 ends with
//   ((a0+a32)+(a16+a48))+...), then (S0 + S2) + (S1 + S3) over the waves;
//   a second one-block kernel combines the tile partials: lane t adds
//   partials t, t+256, ... in order, then the same 256-lane tree.
// (A single-pass "last block finishes" variant needs a device-scope fence
// per tile, which on this multi-XCD part writes back L2 each time; the
// kernel boundary is the cheaper synchronisation.)
// The CPU oracle restates this order, so GPU and CPU sums agree bit for bit.
#pragma once
#include "nbx_internal.h"

namespace nbx {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) v += __shfl_xor(v, h, 64);
  return v;
}

// Sum of `v` over a 256-lane block in the fixed tree order; valid in lane 0.
// s4 must hold NV*4 doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum256(double (&v)[NV], double* s4)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const double w = wave_sum(v[c]);
    if (lane == 0) s4[c * 4 + wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = (s4[c * 4 + 0] + s4[c * 4 + 2]) + (s4[c * 4 + 1] + s4[c * 4 + 3]);
  }
}

// Second phase: one 256-lane block combines the nb partials of each of the
// NV values (partials[c*nb + b]) in fixed order; totals end in lane 0 of v.
template <int NV>
__device__ __forceinline__ void combine_partials(double (&v)[NV], const double* __restrict__ partials,
                                                 unsigned int nb, double* s4)
{
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    // loads batched kBatch deep (a single block is latency-bound: 32 deep
    // takes a quarter of the round trips of 8); additions stay in i order
    constexpr int kBatch = 32;
    double acc = 0.0;
    const double* pc = partials + (size_t)c * nb;
    unsigned int i = threadIdx.x;
    for (; i + (kBatch - 1) * 256 < nb; i += kBatch * 256) {
      double t[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) t[u] = pc[i + u * 256];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) acc = acc + t[u];
    }
    for (; i < nb; i += 256) acc = acc + pc[i];
    v[c] = acc;
  }
  block_sum256<NV>(v, s4);
}

// First phase epilogue: lane 0 of each block stores its NV block sums.
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double* __restrict__ partials)
{
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) partials[(size_t)c * gridDim.x + blockIdx.x] = v[c];
  }
}

}  // namespace nbx

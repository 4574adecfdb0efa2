// Deterministic block reductions shared by the BLAS-1 and PCG kernels.
//
// Order (fixed, independent of timing and of the number of CUs):
//   tile b of 256 lanes x 8 values; lane t takes b*2048 + v*256 + t, v = 0..7,
//   accumulating with an fma chain in v order;
//   256-lane tree: halving tree inside each wave64 (lane 0 ends with
//   ((a0+a32)+(a16+a48))+...), then (S0 + S2) + (S1 + S3) over the waves;
//   the partials of all tiles are combined by the last tile to finish
//   (atomic ticket): lane t adds partials t, t+256, ... in order, then the
//   same 256-lane tree.
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

// Called by every block after it has its NV block sums in lane 0 of v.
// Writes them to partials[c*nblocks + b]; the last block to arrive combines all
// partials in fixed order and returns true with the totals in lane 0 of v.
// ticket must be 0 on entry and is 0 again after the last block.
template <int NV>
__device__ __forceinline__ bool last_block_combine(double (&v)[NV], double* __restrict__ partials,
                                                   unsigned int* __restrict__ ticket, double* s4)
{
  __shared__ bool s_last;
  const unsigned int nb = gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) partials[(size_t)c * nb + blockIdx.x] = v[c];
    __threadfence();
    const unsigned int prev = atomicAdd(ticket, 1u);
    s_last = (prev == nb - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    double acc = 0.0;
    for (unsigned int i = threadIdx.x; i < nb; i += 256)
      acc = acc + __hip_atomic_load(&partials[(size_t)c * nb + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v[c] = acc;
  }
  block_sum256<NV>(v, s4);
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
  return true;
}

}  // namespace nbx

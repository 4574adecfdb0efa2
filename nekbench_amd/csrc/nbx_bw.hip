// Memory-bandwidth probes of the reference's bw benchmark (bw/main.cpp,
// bw/kernel/bwKernels.okl) and the read and copy streams that bound the
// solver's kernels: the practical HBM and LDS ceilings their rooflines are
// read against.
//
//   triad  c = a + alpha b              (bwKernels.okl:1-11)
//   copy   c = a
//   read   sum of a (deterministic: fixed grid, fixed tree)
//   lds    per 128-lane block: s = a[block]; 128 rounds of s[(i+j)%128] += s0;
//          a[block] = s                 (bwKernels.okl:13-42, "smem")
//
// The streams move two doubles per lane per step with nontemporal loads and
// stores, as the solver's single-use streams do.
#include <cstdint>

#include "nbx_internal.h"

namespace nbx {

namespace {

typedef double double2_t __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_triad(const int32_t N, const double alpha, const double* __restrict__ a,
                                               const double* __restrict__ b, double* __restrict__ c)
{
  const int64_t i = 2 * ((int64_t)blockIdx.x * 256 + threadIdx.x);
  if (i + 1 < N) {
    const double2_t av = __builtin_nontemporal_load((const double2_t*)(a + i));
    const double2_t bv = __builtin_nontemporal_load((const double2_t*)(b + i));
    double2_t cv;
    cv.x = av.x + alpha * bv.x;
    cv.y = av.y + alpha * bv.y;
    __builtin_nontemporal_store(cv, (double2_t*)(c + i));
  } else if (i < N) {
    c[i] = a[i] + alpha * b[i];
  }
}

__global__ __launch_bounds__(256) void k_copy(const int32_t N, const double* __restrict__ a, double* __restrict__ c)
{
  const int64_t i = 2 * ((int64_t)blockIdx.x * 256 + threadIdx.x);
  if (i + 1 < N) {
    __builtin_nontemporal_store(__builtin_nontemporal_load((const double2_t*)(a + i)), (double2_t*)(c + i));
  } else if (i < N) {
    c[i] = a[i];
  }
}

constexpr int kReadBlocks = 2048;

// lane t of block b adds the pairs b*256+t, + kReadBlocks*256, ... in order
// (acc + (x0 + x1) per pair); per wave a halving tree, then (w0 + w2) + (w1 + w3)
__global__ __launch_bounds__(256) void k_read(const int32_t N, const double* __restrict__ a,
                                              double* __restrict__ partials)
{
  __shared__ double s_w[4];
  const int64_t stride = (int64_t)kReadBlocks * 256;
  double acc = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; 2 * p < N; p += stride) {
    const int64_t i = 2 * p;
    if (i + 1 < N) {
      const double2_t v = __builtin_nontemporal_load((const double2_t*)(a + i));
      acc = acc + (v.x + v.y);
    } else {
      acc = acc + a[i];
    }
  }
  for (int h = 32; h >= 1; h >>= 1) acc = acc + __shfl_down(acc, h, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (s_w[0] + s_w[2]) + (s_w[1] + s_w[3]);
}

__global__ __launch_bounds__(256) void k_read_finish(const double* __restrict__ partials, double* __restrict__ out)
{
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kReadBlocks; i += 256) acc = acc + partials[i];
  s[threadIdx.x] = acc;
  for (int h = 128; h >= 1; h >>= 1) {
    __syncthreads();
    if ((int)threadIdx.x < h) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + h];
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

__global__ __launch_bounds__(128) void k_lds_probe(double* __restrict__ a)
{
  __shared__ double s[128];
  const int i = threadIdx.x;
  const int64_t id = (int64_t)blockIdx.x * 128 + i;
  s[i] = a[id];
  __syncthreads();
  const double x = s[0];
  __syncthreads();
#pragma unroll 8
  for (int j = 0; j < 128; ++j) {
    s[(i + j) & 127] += x;
    __syncthreads();
  }
  a[id] = s[i];
}

}  // namespace

}  // namespace nbx

using namespace nbx;

extern "C" int nbx_triad(nbx_dlong N, double alpha, const double* a, const double* b, double* c, void* stream)
{
  if (N <= 0) return N == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "triad: N < 0");
  if (!a || !b || !c) return fail(NBX_ERR_INVALID, "triad: null array");
  if (((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) return fail(NBX_ERR_INVALID, "triad: arrays must be 16-byte aligned");
  const int64_t pairs = ((int64_t)N + 1) / 2;
  hipLaunchKernelGGL(k_triad, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, as_stream(stream), N, alpha, a, b,
                     c);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

extern "C" int nbx_copy(nbx_dlong N, const double* a, double* c, void* stream)
{
  if (N <= 0) return N == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "copy: N < 0");
  if (!a || !c) return fail(NBX_ERR_INVALID, "copy: null array");
  if (((uintptr_t)a | (uintptr_t)c) & 15) return fail(NBX_ERR_INVALID, "copy: arrays must be 16-byte aligned");
  const int64_t pairs = ((int64_t)N + 1) / 2;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, as_stream(stream), N, a, c);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

extern "C" int nbx_read_sum(nbx_dlong N, const double* a, double* work, void* stream)
{
  static_assert(kReadBlocks + 1 == NBX_READ_SUM_WORK, "read_sum workspace size");
  if (N <= 0 || !a || !work) return fail(NBX_ERR_INVALID, "read_sum: bad arguments");
  if ((uintptr_t)a & 15) return fail(NBX_ERR_INVALID, "read_sum: array must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_read, dim3(kReadBlocks), dim3(256), 0, s, N, a, work + 1);
  NBX_HIP_LAUNCH();
  hipLaunchKernelGGL(k_read_finish, dim3(1), dim3(256), 0, s, work + 1, work);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

extern "C" int nbx_lds_probe(nbx_dlong Nblocks, double* a, void* stream)
{
  if (Nblocks <= 0) return Nblocks == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "lds_probe: Nblocks < 0");
  if (!a) return fail(NBX_ERR_INVALID, "lds_probe: null array");
  hipLaunchKernelGGL(k_lds_probe, dim3(Nblocks), dim3(128), 0, as_stream(stream), a);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

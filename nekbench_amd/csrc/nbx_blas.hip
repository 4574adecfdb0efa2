// BLAS-1 kernels of the nekBone PCG loop (nekBone/kernel/utils.okl) with the
// deterministic reduction order of nbx_reduce.h.
#include "nbx_blas.h"
#include <algorithm>
#include <mutex>

namespace nbx {

// <x,y>_w = sum_id w[id] sum_fld x[id+fld*offset] y[id+fld*offset]
// (weightedMultipleInnerProduct2, nekBone/kernel/utils.okl:413-458); with
// XX also <x,x>_w (weightedInnerProductUpdate, utils.okl:687-793).
template <bool XX>
__global__ __launch_bounds__(256) void k_wdot(const int32_t N, const int32_t offset, const int Nf,
                                              const double* __restrict__ w, const double* __restrict__ x,
                                              const double* __restrict__ y, double* __restrict__ partials)
{
  constexpr int NV = XX ? 2 : 1;
  __shared__ double s4[NV * 4];
  double v[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) v[c] = 0.0;
  const int64_t b0 = (int64_t)blockIdx.x * kRedTile + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kRedPerLane; ++u) {
    const int64_t id = b0 + u * kRedLanes;
    if (id < N) {
      const double wid = w[id];
      for (int f = 0; f < Nf; ++f) {
        const double xf = x[id + (int64_t)f * offset];
        const double wx = wid * xf;
        v[0] = __builtin_fma(wx, y[id + (int64_t)f * offset], v[0]);
        if (XX) v[NV - 1] = __builtin_fma(wx, xf, v[NV - 1]);
      }
    }
  }
  block_sum256<NV>(v, s4);
  store_partials<NV>(v, partials);
}

template <int NV>
__global__ __launch_bounds__(256) void k_wdot_finish(const unsigned int nb, const double* __restrict__ partials,
                                                     double* __restrict__ out)
{
  __shared__ double s4[NV * 4];
  double v[NV];
  combine_partials<NV>(v, partials, nb, s4);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = v[c];
  }
}

__global__ __launch_bounds__(256) void k_scaled_add(const int32_t N, const double alpha,
                                                    const double* __restrict__ x, const double beta,
                                                    double* __restrict__ y)
{
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) y[n] = alpha * x[n] + beta * y[n];
}

// unweighted sum over the fields of the E-vector (sum, nekBone/kernel/utils.okl:212-262)
__global__ __launch_bounds__(256) void k_sum(const int32_t N, const int32_t offset, const int Nf,
                                             const double* __restrict__ x, double* __restrict__ partials)
{
  __shared__ double s4[4];
  double v[1] = {0.0};
  const int64_t b0 = (int64_t)blockIdx.x * kRedTile + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kRedPerLane; ++u) {
    const int64_t id = b0 + u * kRedLanes;
    if (id < N)
      for (int f = 0; f < Nf; ++f) v[0] = v[0] + x[id + (int64_t)f * offset];
  }
  block_sum256<1>(v, s4);
  store_partials<1>(v, partials);
}

// mean = total / count (BPZeroMean, nekBone/solve.cpp:129-131)
__global__ __launch_bounds__(256) void k_mean_finish(const unsigned int nb, const double* __restrict__ partials,
                                                     const double count, double* __restrict__ out)
{
  __shared__ double s4[4];
  double v[1];
  combine_partials<1>(v, partials, nb, s4);
  if (threadIdx.x == 0) out[0] = v[0] / count;
}

// y[n] += -(*mean) for n < N (addScalar, nekBone/kernel/utils.okl:29-36)
__global__ __launch_bounds__(256) void k_sub_scalar(const int64_t N, const double* __restrict__ mean,
                                                    double* __restrict__ y)
{
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) y[n] += -mean[0];
}

int launch_zero_mean(int32_t N, int32_t offset, int Nf, double count, double* q, ReduceWork& ws, hipStream_t s)
{
  if (N <= 0) return fail(NBX_ERR_INVALID, "zero_mean: N=%d", N);
  const unsigned int nb = (unsigned int)((N + kRedTile - 1) / kRedTile);
  if (nb > ws.capacity) return fail(NBX_ERR_INVALID, "zero_mean: workspace too small (%u < %u)", ws.capacity, nb);
  const int64_t Nall = Nf > 1 ? (int64_t)Nf * offset : (int64_t)std::max(N, offset);
  hipLaunchKernelGGL(k_sum, dim3(nb), dim3(256), 0, s, N, offset, Nf, q, ws.partials);
  hipLaunchKernelGGL(k_mean_finish, dim3(1), dim3(256), 0, s, nb, ws.partials, count, ws.result + 7);
  hipLaunchKernelGGL(k_sub_scalar, dim3((unsigned)((Nall + 255) / 256)), dim3(256), 0, s, Nall, ws.result + 7, q);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int launch_wdot(int32_t N, int32_t offset, int Nf, const double* w, const double* x, const double* y, bool xx,
                ReduceWork& ws, double* out_dev, hipStream_t s)
{
  if (N <= 0) return fail(NBX_ERR_INVALID, "wdot: N=%d", N);
  const unsigned int nb = (unsigned int)((N + kRedTile - 1) / kRedTile);
  if (nb > ws.capacity) return fail(NBX_ERR_INVALID, "wdot: workspace too small (%u < %u)", ws.capacity, nb);
  if (xx) {
    hipLaunchKernelGGL(k_wdot<true>, dim3(nb), dim3(256), 0, s, N, offset, Nf, w, x, y, ws.partials);
    hipLaunchKernelGGL(k_wdot_finish<2>, dim3(1), dim3(256), 0, s, nb, ws.partials, out_dev);
  } else {
    hipLaunchKernelGGL(k_wdot<false>, dim3(nb), dim3(256), 0, s, N, offset, Nf, w, x, y, ws.partials);
    hipLaunchKernelGGL(k_wdot_finish<1>, dim3(1), dim3(256), 0, s, nb, ws.partials, out_dev);
  }
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int ReduceWork::init(int64_t maxN, int nvals)
{
  capacity = (unsigned int)((maxN + kRedTile - 1) / kRedTile);
  if (capacity < 1) capacity = 1;
  NBX_HIP(hipMalloc(&partials, sizeof(double) * capacity * (size_t)nvals));
  NBX_HIP(hipMalloc(&result, sizeof(double) * 8));
  return NBX_SUCCESS;
}

void ReduceWork::release()
{
  if (partials) (void)hipFree(partials);
  if (result) (void)hipFree(result);
  partials = nullptr;
  result = nullptr;
  capacity = 0;
}

}  // namespace nbx

using namespace nbx;

// Workspace for the stand-alone reduction entry point; grown on demand, one
// per process.  Calls that overlap from several host threads are serialised.
static std::mutex g_wdot_mutex;
static ReduceWork g_wdot_ws;

static __global__ void k_store_zero(double* __restrict__ out) { *out = 0.0; }

extern "C" int nbx_weighted_inner_product(nbx_dlong N, nbx_dlong offset, int Nfields, const double* w,
                                          const double* x, const double* y, double* result_host, void* stream)
{
  if (N < 0 || Nfields < 1 || !result_host) return fail(NBX_ERR_INVALID, "weighted_inner_product: bad arguments");
  if (N == 0) {  // the empty sum (the reference's loops do nothing)
    *result_host = 0.0;
    return NBX_SUCCESS;
  }
  if (!w || !x || !y) return fail(NBX_ERR_INVALID, "weighted_inner_product: null array");
  if (Nfields > 1 && offset < N) return fail(NBX_ERR_INVALID, "weighted_inner_product: offset < N");
  std::lock_guard<std::mutex> lock(g_wdot_mutex);
  const unsigned int nb = (unsigned int)((N + kRedTile - 1) / kRedTile);
  if (nb > g_wdot_ws.capacity) {
    NBX_HIP(hipDeviceSynchronize());
    g_wdot_ws.release();
    int rc = g_wdot_ws.init(N, 2);
    if (rc) return rc;
  }
  hipStream_t s = as_stream(stream);
  int rc = launch_wdot(N, offset, Nfields, w, x, y, false, g_wdot_ws, g_wdot_ws.result, s);
  if (rc) return rc;
  NBX_HIP(hipMemcpyAsync(result_host, g_wdot_ws.result, sizeof(double), hipMemcpyDeviceToHost, s));
  NBX_HIP(hipStreamSynchronize(s));
  return NBX_SUCCESS;
}

extern "C" int nbx_scaled_add(nbx_dlong N, double alpha, const double* x, double beta, double* y, void* stream)
{
  if (N <= 0) return N == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "scaled_add: N < 0");
  if (!x || !y) return fail(NBX_ERR_INVALID, "scaled_add: null array");
  hipLaunchKernelGGL(k_scaled_add, dim3((N + 255) / 256), dim3(256), 0, as_stream(stream), N, alpha, x, beta, y);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

// ---- allocation-free weighted inner product ------------------------------
// The tile partials of nbx_weighted_inner_product_noalloc live in this fixed
// device array (one copy per GPU, held until the process ends): one double
// per tile of 2048 nodes, so N <= kNoAllocMaxTiles * 2048 = 134,217,728.  In
// host code `scratch` names a host-side stand-in, not device memory: kernels
// get the device address that hipGetSymbolAddress returns for the current
// device, looked up once per device on the first call there.
static constexpr unsigned int kNoAllocMaxTiles = 65536;  // 512 KiB of partials
static __device__ __align__(256) double scratch[kNoAllocMaxTiles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_dev[kMaxDevices] = {};

int scratch_for_current_device(double** out)
{
  int dev = -1;
  NBX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(NBX_ERR_UNSUPPORTED, "weighted_inner_product_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NBX_ERR_HIP, "weighted_inner_product_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return NBX_SUCCESS;
}
}  // namespace

extern "C" int nbx_weighted_inner_product_noalloc(nbx_dlong N, nbx_dlong offset, int Nfields, const double* w,
                                                  const double* x, const double* y, double* result_dev,
                                                  void* stream)
{
  if (N < 0 || Nfields < 1 || !result_dev)
    return fail(NBX_ERR_INVALID, "weighted_inner_product_noalloc: bad arguments");
  if (N == 0) {  // the empty sum, written on the stream like any result
    hipLaunchKernelGGL(k_store_zero, dim3(1), dim3(1), 0, as_stream(stream), result_dev);
    NBX_HIP_LAUNCH();
    return NBX_SUCCESS;
  }
  if (!w || !x || !y) return fail(NBX_ERR_INVALID, "weighted_inner_product_noalloc: null array");
  if (Nfields > 1 && offset < N) return fail(NBX_ERR_INVALID, "weighted_inner_product_noalloc: offset < N");
  const unsigned int nb = (unsigned int)((N + kRedTile - 1) / kRedTile);
  if (nb > kNoAllocMaxTiles)
    return fail(NBX_ERR_INVALID, "weighted_inner_product_noalloc: N = %d exceeds the maximum %u", N,
                kNoAllocMaxTiles * (unsigned int)kRedTile);
  double* partials = nullptr;
  int rc = scratch_for_current_device(&partials);
  if (rc) return rc;
  // every partial the finish pass reads is written by the tile pass first,
  // so nothing needs clearing between calls
  ReduceWork ws;
  ws.partials = partials;
  ws.capacity = kNoAllocMaxTiles;
  return launch_wdot(N, offset, Nfields, w, x, y, false, ws, result_dev, as_stream(stream));
}

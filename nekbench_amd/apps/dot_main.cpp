// Weighted inner product benchmark on one MI355X with the command line and
// output of the reference's dot driver (dot/main.cpp:27-171):
//
//   dot N numElements [NATIVE|OKL]+SERIAL|CUDA|OPENCL CPU|VOLTA [nRepetitions] [deviceId] [blockSize] [MPI]
//
// sum_n w[n] x[n] y[n] over Nelements*(N+1)^3 values through
// nbx_weighted_inner_product (deterministic tile tree, result read back to
// the host each call as the reference reads its block partials).  As in the
// reference, w, x and y all hold the first drand48 vector (dot/main.cpp:126-
// 132).  blockSize is accepted for compatibility; the library's reduction
// uses tiles of 2048 values.  Statistics with the reference's count:
// 3 words per node (dot/main.cpp:151-167).  One added line compares the
// result with a host sum.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "nbx.h"

namespace {

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    std::fprintf(stderr, "dot: %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 5) {
    std::printf(
        "Usage: ./dot N numElements [NATIVE|OKL]+SERIAL|CUDA|OPENCL CPU|VOLTA [nRepetitions] [deviceId] [blockSize] "
        "[MPI]\n");
    return 1;
  }
  const int N = std::atoi(argv[1]);
  const int Nelements = std::atoi(argv[2]);
  const int Ntests = argc >= 6 ? std::atoi(argv[5]) : 100;
  const int blockSize = argc >= 8 ? std::atoi(argv[7]) : 256;
  if (N < 1 || Nelements < 1 || Ntests < 1) {
    std::fprintf(stderr, "dot: need N >= 1, numElements >= 1, nRepetitions >= 1\n");
    return 1;
  }
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  const long long n = (long long)Nelements * Np;
  if (n > 0x7fffffffLL) {
    std::fprintf(stderr, "dot: numElements*(N+1)^3 exceeds the 32-bit dlong range\n");
    return 1;
  }
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  std::printf("word size: %zu bytes\n", sizeof(double));
  std::printf("active mode: HIP (%s, %s)\n", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName);
  std::vector<double> a(n);
  for (long long i = 0; i < n; ++i) a[i] = drand48();
  double* d;
  hip_check(hipMalloc(&d, n * sizeof(double)), "hipMalloc");
  hip_check(hipMemcpy(d, a.data(), n * sizeof(double), hipMemcpyHostToDevice), "copy");
  std::printf("blockSize: %lld\n", (n + 2047) / 2048);

  double wab = 0;
  int rc = nbx_weighted_inner_product((nbx_dlong)n, (nbx_dlong)n, 1, d, d, d, &wab, nullptr);
  if (rc < 0) {
    std::fprintf(stderr, "dot: nbx_weighted_inner_product failed (%d): %s\n", rc, nbx_last_error());
    return 1;
  }
  hip_check(hipDeviceSynchronize(), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  for (int t = 0; t < Ntests; ++t) nbx_weighted_inner_product((nbx_dlong)n, (nbx_dlong)n, 1, d, d, d, &wab, nullptr);
  hip_check(hipDeviceSynchronize(), "sync");
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / Ntests;

  double host = 0;
  for (long long i = 0; i < n; ++i) host += a[i] * a[i] * a[i];
  std::printf("Correctness check: |wxy - host sum| / |host sum| = %g\n", std::fabs(wab - host) / std::fabs(host));
  const double GDOFPerSecond = ((double)Nelements * N * N * N) / elapsed / 1.e9;
  const double bw = (3.0 * Np * sizeof(double) * Nelements / elapsed) / 1.e9;
  std::printf("MPItasks=1 OMPthreads=1 NRepetitions=%d N=%d Nelements=%d blockSize=%d elapsed time=%g GDOF/s=%g "
              "GB/s=%g\n",
              Ntests, N, Nelements, blockSize, elapsed, GDOFPerSecond, bw);
  hipFree(d);
  return 0;
}

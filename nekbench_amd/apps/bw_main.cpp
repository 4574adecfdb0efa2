// Memory-bandwidth benchmark on one MI355X with the sections and output lines
// of the reference's bw driver (bw/main.cpp:13-223):
//
//   bw HIP
//
//   triad stream at the reference's sizes (1 .. 16000*512 words, 2000
//     repetitions each, bwKernels.okl "triad"), bytes = 3 N 8
//   shared mem bw (bwKernels.okl "smem", 128M words), bytes = 128*2*8*N
//   D->H and H->D memcpy through pinned host memory, D->D memcpy
// The reference's largest triad (3 x 128 MB) fits the 256 MB Infinity Cache
// of this part, so one more section measures HBM itself with 2 GiB arrays:
// triad, copy and read streams (nbx_triad, nbx_copy, nbx_read_sum; the read
// includes its one-block finishing pass), timed with HIP events.
// Timing of the reference's sections as the reference's: host clock around
// the repetitions, the device drained before and after.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nbx.h"

namespace {

void die(const char* what, int rc)
{
  std::fprintf(stderr, "bw: %s failed (%d): %s\n", what, rc, nbx_last_error());
  std::exit(1);
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    std::fprintf(stderr, "bw: %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

template <typename F>
double event_time_per_call(int n, F&& f)
{
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  f();
  hip_check(hipEventRecord(e0, nullptr), "event");
  for (int i = 0; i < n; ++i) f();
  hip_check(hipEventRecord(e1, nullptr), "event");
  hip_check(hipEventSynchronize(e1), "event");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return ms * 1e-3 / n;
}

template <typename F>
double time_per_call(int n, F&& f)
{
  hip_check(hipDeviceSynchronize(), "sync");
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n; ++i) f();
  hip_check(hipDeviceSynchronize(), "sync");
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / n;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 2) {
    std::printf("Usage: ./bw SERIAL|CUDA|OPENCL\n");
    return 1;
  }
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  std::printf("active mode: HIP (%s, %s)\n\n", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName);

  {
    const int Ntests = 2000;
    const int N[] = {1, 1024, 8192, 16384, 64000, 256 * 512, 512 * 512, 1000 * 512, 2000 * 512, 4000 * 512,
                     8000 * 512, 16000 * 512};
    const int nWords = N[sizeof(N) / sizeof(N[0]) - 1];
    double *a, *b, *c;
    hip_check(hipMalloc(&a, nWords * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&b, nWords * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&c, nWords * sizeof(double)), "hipMalloc");
    hip_check(hipMemset(a, 0, nWords * sizeof(double)), "memset");
    hip_check(hipMemset(b, 0, nWords * sizeof(double)), "memset");
    for (int t = 0; t < Ntests; ++t) {
      const int rc = nbx_triad(1000, 1.0, a, b, c, nullptr);
      if (rc < 0) die("nbx_triad", rc);
    }
    for (int n : N) {
      const double bytes = 3.0 * n * sizeof(double);
      const double el = time_per_call(Ntests, [&] { nbx_triad(n, 1.0, a, b, c, nullptr); });
      std::printf("triad stream: %d words, %g s, %g bytes/s\n", n, el, bytes / el);
    }
    hipFree(a);
    hipFree(b);
    hipFree(c);
  }
  std::printf("\n");
  {
    const int Ntests = 50, blockSize = 128;
    const int N = blockSize * 1024 * 1024, Nblock = N / blockSize;
    double* a;
    hip_check(hipMalloc(&a, (size_t)N * sizeof(double)), "hipMalloc");
    hip_check(hipMemset(a, 0, (size_t)N * sizeof(double)), "memset");
    int rc = nbx_lds_probe(Nblock, a, nullptr);
    if (rc < 0) die("nbx_lds_probe", rc);
    const double el = time_per_call(Ntests, [&] { nbx_lds_probe(Nblock, a, nullptr); });
    const double bytes = (double)blockSize * 2 * sizeof(double) * N;
    std::printf("shared mem bw: %g s, %g bytes/s\n", el, bytes / el);
    hipFree(a);
  }
  std::printf("\n");
  {
    const int N[] = {1, 2000, 4000, 8000, 16000, 50000, 100000, 150000, 300000, 2000 * 512, 4000 * 512, 8000 * 512};
    const int nWords = N[sizeof(N) / sizeof(N[0]) - 1];
    double *h, *a;
    hip_check(hipHostMalloc(&h, nWords * sizeof(double)), "hipHostMalloc");
    hip_check(hipMalloc(&a, nWords * sizeof(double)), "hipMalloc");
    hip_check(hipMemset(a, 0, nWords * sizeof(double)), "memset");
    for (int n : N) {
      const int Ntests = n > 10000 ? 100 : 5000;
      const double bytes = (double)n * sizeof(double);
      const double el =
          time_per_call(Ntests, [&] { hip_check(hipMemcpy(h, a, bytes, hipMemcpyDeviceToHost), "D->H"); });
      std::printf("D->H memcpy %d words, %g s, %g bytes/s\n", n, el, bytes / el);
    }
    std::printf("\n");
    for (int n : N) {
      const int Ntests = n > 10000 ? 100 : 5000;
      const double bytes = (double)n * sizeof(double);
      const double el =
          time_per_call(Ntests, [&] { hip_check(hipMemcpy(a, h, bytes, hipMemcpyHostToDevice), "H->D"); });
      std::printf("H->D memcpy %d words, %g s, %g bytes/s\n", n, el, bytes / el);
    }
    hipHostFree(h);
    hipFree(a);
  }
  std::printf("\n");
  {
    const int Ntests = 20000;
    const int N[] = {1000 * 512, 2000 * 512, 4000 * 512, 8000 * 512};
    const int nWords = N[sizeof(N) / sizeof(N[0]) - 1];
    double *a, *b;
    hip_check(hipMalloc(&a, nWords * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&b, nWords * sizeof(double)), "hipMalloc");
    for (int n : N) {
      const double bytes = (double)n * sizeof(double);
      const double el = time_per_call(Ntests, [&] { hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nullptr); });
      std::printf("D->D memcpy %d words, %g s, %g bytes/s\n", n, el, 2 * bytes / el);
    }
    hipFree(a);
    hipFree(b);
  }
  std::printf("\n");
  {
    // HBM: 2 GiB per array, beyond the Infinity Cache
    const int n = 256 * 1024 * 1024, Ntests = 20;
    double *a, *b, *c;
    hip_check(hipMalloc(&a, (size_t)n * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&b, (size_t)n * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&c, (size_t)n * sizeof(double)), "hipMalloc");
    hip_check(hipMemset(a, 0, (size_t)n * sizeof(double)), "memset");
    hip_check(hipMemset(b, 0, (size_t)n * sizeof(double)), "memset");
    double el = event_time_per_call(Ntests, [&] { nbx_triad(n, 1.0, a, b, c, nullptr); });
    std::printf("HBM triad stream: %d words, %g s, %g bytes/s\n", n, el, 3.0 * n * sizeof(double) / el);
    el = event_time_per_call(Ntests, [&] { nbx_copy(n, a, c, nullptr); });
    std::printf("HBM copy stream: %d words, %g s, %g bytes/s\n", n, el, 2.0 * n * sizeof(double) / el);
    double* work;
    hip_check(hipMalloc(&work, NBX_READ_SUM_WORK * sizeof(double)), "hipMalloc");
    int rc = nbx_read_sum(n, a, work, nullptr);
    if (rc < 0) die("nbx_read_sum", rc);
    el = event_time_per_call(Ntests, [&] { nbx_read_sum(n, a, work, nullptr); });
    hipFree(work);
    std::printf("HBM read stream: %d words, %g s, %g bytes/s\n", n, el, 1.0 * n * sizeof(double) / el);
    hipFree(a);
    hipFree(b);
    hipFree(c);
  }
  return 0;
}

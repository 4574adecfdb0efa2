// axhelm benchmark on one MI355X, with the command line and output of the
// reference's driver (axhelm/main.cpp:26-186):
//
//   axhelm N Ndim numElements [NATIVE|OKL]+SERIAL|CUDA|HIP|OPENCL CPU|VOLTA [BKmode] [nRepetitions] [kernelVersion]
//
// The thread-model and arch arguments are accepted for compatibility; the
// kernel always runs on the GPU through nbx_axhelm_by_name with the
// reference's kernel name axhelm[_bk][_n<Ndim>]_v<k> (axhelm/main.cpp:105-109).
// Inputs are the reference's: ggeo, q and Aq filled by drand48 in that order
// (main.cpp:16-24,112-114), D from the GLL nodes, lambda0 = 1 and lambda1 =
// 1.1 (0 in BK mode) per node (main.cpp:121-127).  One warm-up launch, then
// the mean wall time of nRepetitions launches (main.cpp:130-142).
// Correctness check: the host evaluation lambda1 GWJ q + D^T (G D q) per
// element and field, as axhelm/axhelmReference.cpp:19-101 defines it,
// against the device result; max |difference| is printed.
// Statistics use the reference's counts (main.cpp:160-185).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nbx.h"

namespace {

void die(const char* what, int rc)
{
  std::fprintf(stderr, "axhelm: %s failed (%d): %s\n", what, rc, nbx_last_error());
  std::exit(1);
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    std::fprintf(stderr, "axhelm: %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

// Host evaluation of one field: Aq = lambda1 GWJ q + sum_ab D_a^T G_ab D_b q.
void host_axhelm(int Nq, int Nelements, double lambda1, const double* ggeo, const double* D, const double* q,
                 double* Aq)
{
  const int Np = Nq * Nq * Nq;
  std::vector<double> qr(Np), qs(Np), qt(Np);
  for (int e = 0; e < Nelements; ++e) {
    const double* g = ggeo + (size_t)e * 7 * Np;
    const double* qe = q + (size_t)e * Np;
    double* Ae = Aq + (size_t)e * Np;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double dr = 0, ds = 0, dt = 0;
          for (int m = 0; m < Nq; ++m) {
            dr += D[i * Nq + m] * qe[m + j * Nq + k * Nq * Nq];
            ds += D[j * Nq + m] * qe[i + m * Nq + k * Nq * Nq];
            dt += D[k * Nq + m] * qe[i + j * Nq + m * Nq * Nq];
          }
          const int n = i + j * Nq + k * Nq * Nq;
          const double G00 = g[1 * Np + n], G01 = g[2 * Np + n], G02 = g[3 * Np + n];
          const double G11 = g[4 * Np + n], G12 = g[5 * Np + n], G22 = g[6 * Np + n];
          qr[n] = G00 * dr + G01 * ds + G02 * dt;
          qs[n] = G01 * dr + G11 * ds + G12 * dt;
          qt[n] = G02 * dr + G12 * ds + G22 * dt;
        }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq * Nq;
          double a = lambda1 * g[n] * qe[n];
          for (int m = 0; m < Nq; ++m) {
            a += D[m * Nq + i] * qr[m + j * Nq + k * Nq * Nq];
            a += D[m * Nq + j] * qs[i + m * Nq + k * Nq * Nq];
            a += D[m * Nq + k] * qt[i + j * Nq + m * Nq * Nq];
          }
          Ae[n] = a;
        }
  }
}

std::vector<double> drand_fill(size_t n)
{
  std::vector<double> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = drand48();
  return v;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 6) {
    std::printf(
        "Usage: ./axhelm N Ndim numElements [NATIVE|OKL]+SERIAL|CUDA|HIP|OPENCL CPU|VOLTA [BKmode] [nRepetitions] "
        "[kernelVersion]\n");
    return 1;
  }
  const int N = std::atoi(argv[1]);
  const int Ndim = std::atoi(argv[2]);
  const int Nelements = std::atoi(argv[3]);
  const int BKmode = argc >= 7 ? std::atoi(argv[6]) : 1;
  const int Ntests = argc >= 8 ? std::atoi(argv[7]) : 100;
  const int kernelVersion = argc >= 9 ? std::atoi(argv[8]) : 0;
  if (N < 1 || N > 15 || (Ndim != 1 && Ndim != 3) || Nelements < 1 || Ntests < 1) {
    std::fprintf(stderr, "axhelm: need 1 <= N <= 15, Ndim 1 or 3, numElements >= 1, nRepetitions >= 1\n");
    return 1;
  }
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  const long long offset = (long long)Nelements * Np;
  if (offset * Ndim > 0x7fffffffLL) {
    std::fprintf(stderr, "axhelm: Ndim*Nelements*Np exceeds the 32-bit dlong range\n");
    return 1;
  }

  std::vector<double> z(Nq), w(Nq), D(Nq * Nq);
  int rc = nbx_gll(N, z.data(), w.data(), D.data());
  if (rc < 0) die("nbx_gll", rc);

  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  std::printf("word size: %zu bytes\n", sizeof(double));
  std::printf("active mode: HIP (%s, %s), thread model argument '%s', arch argument '%s'\n",
              prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName, argv[4], argv[5]);
  if (BKmode) std::printf("BK mode enabled\n");

  std::string kernelName = BKmode ? "axhelm_bk" : "axhelm";
  if (Ndim > 1) kernelName += "_n" + std::to_string(Ndim);
  kernelName += "_v" + std::to_string(kernelVersion);

  std::vector<double> ggeo = drand_fill((size_t)Np * Nelements * 7);
  std::vector<double> q = drand_fill((size_t)Ndim * Np * Nelements);
  std::vector<double> Aq = drand_fill((size_t)Ndim * Np * Nelements);
  const double lambda1 = BKmode ? 0.0 : 1.1;
  std::vector<double> lambda(2 * offset);
  for (long long i = 0; i < offset; ++i) {
    lambda[i] = 1.0;
    lambda[i + offset] = lambda1;
  }

  double *d_ggeo, *d_q, *d_Aq, *d_D, *d_lambda;
  hip_check(hipMalloc(&d_ggeo, ggeo.size() * sizeof(double)), "hipMalloc ggeo");
  hip_check(hipMalloc(&d_q, q.size() * sizeof(double)), "hipMalloc q");
  hip_check(hipMalloc(&d_Aq, Aq.size() * sizeof(double)), "hipMalloc Aq");
  hip_check(hipMalloc(&d_D, D.size() * sizeof(double)), "hipMalloc D");
  hip_check(hipMalloc(&d_lambda, lambda.size() * sizeof(double)), "hipMalloc lambda");
  hip_check(hipMemcpy(d_ggeo, ggeo.data(), ggeo.size() * sizeof(double), hipMemcpyHostToDevice), "copy ggeo");
  hip_check(hipMemcpy(d_q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice), "copy q");
  hip_check(hipMemcpy(d_Aq, Aq.data(), Aq.size() * sizeof(double), hipMemcpyHostToDevice), "copy Aq");
  hip_check(hipMemcpy(d_D, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice), "copy D");
  hip_check(hipMemcpy(d_lambda, lambda.data(), lambda.size() * sizeof(double), hipMemcpyHostToDevice),
            "copy lambda");

  auto launch = [&]() {
    const int r = nbx_axhelm_by_name(kernelName.c_str(), N, Nelements, (nbx_dlong)offset, nullptr, d_ggeo, d_D,
                                     d_lambda, d_q, d_Aq, nullptr);
    if (r < 0) die(kernelName.c_str(), r);
  };
  launch();
  hip_check(hipDeviceSynchronize(), "warm-up");
  const auto t0 = std::chrono::steady_clock::now();
  for (int t = 0; t < Ntests; ++t) launch();
  hip_check(hipDeviceSynchronize(), "timed launches");
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / Ntests;

  for (int n = 0; n < Ndim; ++n)
    host_axhelm(Nq, Nelements, lambda1, ggeo.data(), D.data(), q.data() + n * offset, Aq.data() + n * offset);
  std::vector<double> dev(Aq.size());
  hip_check(hipMemcpy(dev.data(), d_Aq, dev.size() * sizeof(double), hipMemcpyDeviceToHost), "copy Aq back");
  double maxDiff = 0;
  for (size_t i = 0; i < dev.size(); ++i) maxDiff = std::fmax(maxDiff, std::fabs(dev[i] - Aq[i]));
  std::printf("Correctness check: maxError = %g\n", maxDiff);

  const double GDOFPerSecond = ((double)Ndim * N * N * N * Nelements / elapsed) / 1.e9;
  long long bytesMoved = (long long)(Ndim * 2 * Np + 7 * Np) * (long long)sizeof(double);
  if (!BKmode) bytesMoved += 2LL * Np * (long long)sizeof(double);
  const double bw = ((double)bytesMoved * Nelements / elapsed) / 1.e9;
  double flopCount = (double)Np * 12 * Nq + 15.0 * Np;
  if (!BKmode) flopCount += 5.0 * Np;
  flopCount *= Ndim;
  const double gflops = (flopCount * Nelements / elapsed) / 1.e9;
  std::printf("MPItasks=1 OMPthreads=1 NRepetitions=%d Ndim=%d N=%d Nelements=%d elapsed time=%g GDOF/s=%g "
              "GB/s=%g GFLOPS/s=%g\n",
              Ntests, Ndim, N, Nelements, elapsed, GDOFPerSecond, bw, gflops);

  hipFree(d_ggeo);
  hipFree(d_q);
  hipFree(d_Aq);
  hipFree(d_D);
  hipFree(d_lambda);
  return 0;
}

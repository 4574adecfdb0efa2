// nekBone on one MI355X: the reference's program (nekBone/main.cpp:38-224,
// setup.cpp:36-130) driven by the same setup file, through the nbx C ABI.
//
//   nekBone setupfile                   run
//   nekBone --print-options setupfile   print the settings the run would use
//
// Setup-file keys read (nekBone/setup.ini): POLYNOMIAL DEGREE, BOX NX/NY/NZ
// (default 3), NFIELDS (1 or 3), BPMODE, LAMBDA (default 1.1), PRECONDITIONER
// (COPY, JACOBI), KRYLOV SOLVER (PCG; "FLEXIBLE" anywhere in the value selects
// the flexible beta), SOLVER TOLERANCE (default 1e-8), NREPETITIONS (default
// 10), MAXIMUM ITERATIONS, FIXED ITERATION COUNT, VERBOSE, PROFILING (TRUE
// prints the reference's "breakdown" block after the summary: device time
// per phase over the timed solves, nbx_bp_profile).  THREAD MODEL, ARCH,
// PLATFORM/DEVICE NUMBER, KERNEL ID, OVERLAP and TIMER SYNC are accepted and
// have no effect (one GPU, one rank, one kernel per case; the phase timers
// always wait for the device).
// Additions with no reference counterpart: PRECONDITIONER MULTIGRID with
// [MULTIGRID ORDERS] (comma list, default N, N/2, 1), [MULTIGRID DEGREE]
// (3), [MULTIGRID COARSE ITERATIONS] (20), [MULTIGRID PRECISION]
// (FP64|FP32|BF16, default FP64); [KERNEL SEQUENCE] REFERENCE|FUSED (default
// FUSED: same algorithm, p-update and p.Ap folded into the operator pass).
//
// Program flow as the reference's:
//   * the mesh is the periodic box [-1,1]^3 of NX x NY x NZ hexahedra
//     (main.cpp:61-66 fixes the extent; core/meshSetup.cpp:2480-2560);
//   * BPMODE TRUE selects the BK operator (BP5, or BP6 with 3 fields), the
//     COPY preconditioner and lambda1 = 0 (setup.cpp:43-53);
//   * rhs = (3 pi^2 + lambda1) JW cos(pi x) cos(pi y) cos(pi z), assembled
//     (setup.cpp:77-99); with lambda1 = 0 on the periodic box the problem is
//     all-Neumann and r, x are projected to zero mean (main.cpp:24-31);
//   * warm-up solve with the tolerance (at most 1000 iterations, main.cpp:13-22)
//     and the check max |x - cos(pi x) cos(pi y) cos(pi z)| (main.cpp:105-131);
//   * then NREPETITIONS timed solves; FIXED ITERATION COUNT TRUE makes them
//     run MAXIMUM ITERATIONS iterations with no convergence test
//     (main.cpp:134, solve.cpp:55-89): the warm-up solve changes only the
//     solver's copy of the options, so the file's flag decides;
//   * the summary uses the reference's counts (main.cpp:160-224).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "nbx.h"
#include "nbx_setup_ini.h"

namespace {

void die(const char* what, int rc)
{
  std::fprintf(stderr, "nekBone: %s failed (%d): %s\n", what, rc, nbx_last_error());
  std::exit(1);
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    std::fprintf(stderr, "nekBone: %s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

struct Settings {
  int N = 0, NX = 3, NY = 3, NZ = 3, Nfields = 1;
  int bpmode = 0, BPid = 0;
  double lambda1 = 1.1;
  nbx_precon precon = NBX_PRECON_COPY;
  int flexible = 0, verbose = 0, fixed = 0, fused = 1, profiling = 0;
  double tol = 1e-8;
  int Ntests = 10, maxIterFixed = 1000;
  std::vector<int> mgOrders;
  int mgDegree = 3, mgCoarseIters = 20;
  nbx_prec mgPrec = NBX_F64;
  std::string error;
};

Settings resolve(const nbx_app::SetupIni& o)
{
  Settings s;
  if (!o.get("POLYNOMIAL DEGREE", s.N)) s.error = "missing [POLYNOMIAL DEGREE]";
  o.get("BOX NX", s.NX);
  o.get("BOX NY", s.NY);
  o.get("BOX NZ", s.NZ);
  o.get("NFIELDS", s.Nfields);
  o.get("LAMBDA", s.lambda1);
  if (o.compare("PRECONDITIONER", "JACOBI")) s.precon = NBX_PRECON_JACOBI;
  if (o.compare("PRECONDITIONER", "MULTIGRID")) s.precon = NBX_PRECON_MG;
  if (o.compare("BPMODE", "TRUE")) {
    s.bpmode = 1;
    s.precon = NBX_PRECON_COPY;
    s.BPid = s.Nfields == 3 ? 6 : 5;
    s.lambda1 = 0.0;
  }
  if (o.has("KRYLOV SOLVER") && !o.compare("KRYLOV SOLVER", "PCG")) s.error = "only the PCG Krylov solver exists";
  s.flexible = o.compare("KRYLOV SOLVER", "FLEXIBLE") ? 1 : 0;
  s.verbose = o.compare("VERBOSE", "TRUE") ? 1 : 0;
  s.profiling = o.compare("PROFILING", "TRUE") ? 1 : 0;
  o.get("SOLVER TOLERANCE", s.tol);
  o.get("NREPETITIONS", s.Ntests);
  o.get("MAXIMUM ITERATIONS", s.maxIterFixed);
  s.fixed = o.compare("FIXED ITERATION COUNT", "TRUE") ? 1 : 0;
  if (o.compare("KERNEL SEQUENCE", "REFERENCE")) s.fused = 0;
  if (s.precon == NBX_PRECON_MG) {
    if (o.has("MULTIGRID ORDERS")) {
      std::stringstream ss(o.get("MULTIGRID ORDERS"));
      std::string item;
      while (std::getline(ss, item, ',')) s.mgOrders.push_back(std::atoi(item.c_str()));
    } else {
      s.mgOrders = {s.N};
      if (s.N / 2 > 1) s.mgOrders.push_back(s.N / 2);
      if (s.N > 1) s.mgOrders.push_back(1);
    }
    o.get("MULTIGRID DEGREE", s.mgDegree);
    o.get("MULTIGRID COARSE ITERATIONS", s.mgCoarseIters);
    if (o.compare("MULTIGRID PRECISION", "FP32")) s.mgPrec = NBX_F32;
    if (o.compare("MULTIGRID PRECISION", "BF16")) s.mgPrec = NBX_BF16;
  }
  if (s.error.empty() && (s.N < 1 || s.N > 15)) s.error = "POLYNOMIAL DEGREE must be in [1, 15]";
  if (s.error.empty() && (s.NX < 3 || s.NY < 3 || s.NZ < 3))
    s.error = "the periodic box needs at least 3 elements per direction";
  if (s.error.empty() && s.Nfields != 1 && s.Nfields != 3) s.error = "NFIELDS must be 1 or 3";
  if (s.error.empty() && s.Ntests < 1) s.error = "NREPETITIONS must be >= 1";
  return s;
}

const char* precon_name(nbx_precon p)
{
  switch (p) {
    case NBX_PRECON_COPY: return "COPY";
    case NBX_PRECON_JACOBI: return "JACOBI";
    case NBX_PRECON_MG: return "MULTIGRID";
    default: return "LEARNED";
  }
}

void print_settings(const Settings& s)
{
  std::printf("N=%d\nNX=%d\nNY=%d\nNZ=%d\nNfields=%d\nBPid=%d\nlambda1=%.17g\nprecon=%s\nflexible=%d\n", s.N, s.NX,
              s.NY, s.NZ, s.Nfields, s.BPid, s.lambda1, precon_name(s.precon), s.flexible);
  std::printf("tol=%.17g\nNtests=%d\nmaxIterFixed=%d\nfixed=%d\nverbose=%d\nfused=%d\n", s.tol, s.Ntests,
              s.maxIterFixed, s.fixed, s.verbose, s.fused);
  if (s.precon == NBX_PRECON_MG) {
    std::printf("mgOrders=");
    for (size_t i = 0; i < s.mgOrders.size(); ++i) std::printf(i ? ",%d" : "%d", s.mgOrders[i]);
    std::printf("\nmgDegree=%d\nmgCoarseIters=%d\nmgPrec=%d\n", s.mgDegree, s.mgCoarseIters, (int)s.mgPrec);
  }
}

// solve() of main.cpp:4-34: a positive tolerance stops at 1000 iterations at
// most; otherwise exactly maxIterFixed iterations run.
int solve(nbx_bp* bp, const double* lambda, double tol, int maxIterFixed, double* r, double* x, bool verbose)
{
  const int maxIter = tol > 0 ? 1000 : maxIterFixed;
  nbx_pcg_stats st;
  const int it = nbx_bp_pcg(bp, lambda, r, x, tol > 0 ? tol : 0.0, maxIter, &st);
  if (it < 0) die("nbx_bp_pcg", it);
  if (verbose) {
    std::vector<double> rr(it + 2), al(it + 2);
    const int n = nbx_bp_history(bp, rr.data(), al.data(), it + 2);
    for (int k = 0; k < n; ++k)
      std::printf("CG: it %d r norm %12.12le alpha = %le \n", k + 1, std::sqrt(rr[k]), k ? al[k - 1] : 0.0);
  }
  return it;
}

}  // namespace

int main(int argc, char** argv)
{
  const bool dump = argc == 3 && std::string(argv[1]) == "--print-options";
  if (argc != 2 && !dump) {
    std::printf("usage: ./nekBone setupfile\n");
    return 1;
  }
  nbx_app::SetupIni options;
  if (!options.load(argv[argc - 1])) {
    std::fprintf(stderr, "nekBone: cannot read %s\n", argv[argc - 1]);
    return 1;
  }
  const Settings s = resolve(options);
  if (!s.error.empty()) {
    std::fprintf(stderr, "nekBone: %s\n", s.error.c_str());
    return 1;
  }
  if (dump) {
    print_settings(s);
    return 0;
  }

  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
  std::printf("active mode: HIP (%s, %s)\n", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName);
  if (s.bpmode) std::printf("BP mode enabled\n");
  std::printf("overlap disabled\n");

  size_t free0 = 0, total = 0;
  hip_check(hipMemGetInfo(&free0, &total), "hipMemGetInfo");
  nbx_bp_options opt{};
  opt.box.N = s.N;
  opt.box.NX = s.NX;
  opt.box.NY = s.NY;
  opt.box.NZ = s.NZ;
  opt.box.periodic = 1;
  opt.box.perturb = 0.0;
  opt.box.zaspect = 0.0;
  opt.box.seed = 1;
  opt.box.x0 = opt.box.y0 = opt.box.z0 = -1.0;
  opt.box.x1 = opt.box.y1 = opt.box.z1 = 1.0;
  opt.Nfields = s.Nfields;
  opt.bpmode = s.bpmode;
  opt.lambda1 = s.lambda1;
  opt.dirichlet = 0;
  opt.precon = s.precon;
  opt.flexible = s.flexible;
  opt.verbose = s.verbose;
  opt.rhs_kind = 1;
  opt.fused = s.fused;
  nbx_bp* bp = nbx_bp_setup(&opt, nullptr);
  if (!bp) die("nbx_bp_setup", -1);
  if (s.precon == NBX_PRECON_MG) {
    const int rc = nbx_bp_mg_setup(bp, (int)s.mgOrders.size(), s.mgOrders.data(), s.mgDegree, s.mgCoarseIters, 0.1,
                                   1.1, s.mgPrec);
    if (rc < 0) die("nbx_bp_mg_setup", rc);
  }
  nbx_dlong E = 0, Np = 0, offset = 0, Nmasked = 0;
  nbx_bp_sizes(bp, &E, &Np, &offset, &Nmasked);
  const size_t Nall = (size_t)s.Nfields * offset;
  double *r = nullptr, *x = nullptr;
  hip_check(hipMalloc(&r, Nall * sizeof(double)), "hipMalloc r");
  hip_check(hipMalloc(&x, Nall * sizeof(double)), "hipMalloc x");
  size_t free1 = 0;
  hip_check(hipMemGetInfo(&free1, &total), "hipMemGetInfo");
  std::printf("setup done: bytes allocated = %lld\n", (long long)(free0 - free1));

  const double* rhs = static_cast<const double*>(nbx_bp_device_array(bp, "rhs"));
  const double* lambda = static_cast<const double*>(nbx_bp_device_array(bp, "lambda"));
  auto reset = [&]() {
    hip_check(hipMemset(x, 0, Nall * sizeof(double)), "reset x");
    hip_check(hipMemcpy(r, rhs, Nall * sizeof(double), hipMemcpyDeviceToDevice), "reset r");
    hip_check(hipDeviceSynchronize(), "reset");
  };

  // warm-up solve and correctness check against the exact solution
  double tol = s.tol;
  reset();
  int it = solve(bp, lambda, tol, s.maxIterFixed, r, x, s.verbose);
  {
    std::vector<double> xh(Nall), xn(E * (size_t)Np), yn(E * (size_t)Np), zn(E * (size_t)Np);
    hip_check(hipMemcpy(xh.data(), x, Nall * sizeof(double), hipMemcpyDeviceToHost), "copy x");
    const int rc = nbx_box_mesh(&opt.box, 0, xn.data(), yn.data(), zn.data(), nullptr, nullptr, nullptr);
    if (rc < 0) die("nbx_box_mesh", rc);
    double maxError = 0;
    for (int f = 0; f < s.Nfields; ++f)
      for (size_t id = 0; id < xn.size(); ++id) {
        const double exact = std::cos(M_PI * xn[id]) * std::cos(M_PI * yn[id]) * std::cos(M_PI * zn[id]);
        maxError = std::fmax(maxError, std::fabs(exact - xh[id + (size_t)f * offset]));
      }
    std::printf("correctness check: maxError = %g in %d iterations\n", maxError, it);
  }

  if (s.fixed) tol = 0;
  if (s.profiling) {  // phase timers over the timed solves (nekBone/main.cpp:181-193)
    int rc = nbx_bp_set_profiling(bp, 1);
    if (rc < 0) die("nbx_bp_set_profiling", rc);
  }
  std::printf("\nrunning solver ...");
  std::fflush(stdout);
  double elapsed = 0;
  for (int t = 0; t < s.Ntests; ++t) {
    reset();
    const auto t0 = std::chrono::steady_clock::now();
    it = solve(bp, lambda, tol, s.maxIterFixed, r, x, false);
    hip_check(hipDeviceSynchronize(), "solve");
    elapsed += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  std::printf(" done\n");
  elapsed /= s.Ntests;

  const int N = s.N, Nq = N + 1, Nf = s.Nfields;
  const double globalNdofs = (double)N * N * N * E;
  const double gDOFs = Nf * (it * (globalNdofs / elapsed)) / 1.e9;
  const double Nlocal = (double)Np * E;
  const double gbytesPrecon = Nf * Nlocal;
  const double gbytesScaledAdd = 2. * Nf * Nlocal;
  double gbytesAx = (7 + 2 * Nf) * Nlocal;
  if (s.BPid) gbytesAx += 2 * Nf * Nlocal;
  const double gbytesDot = (2 * Nf + 1) * Nlocal;
  const double gbytesPupdate = 4 * Nf * Nlocal;
  const double NGbytes =
      (gbytesPrecon + gbytesScaledAdd + gbytesAx + 3 * gbytesDot + gbytesPupdate) * (sizeof(double) / 1.e9);
  const double bw = (it * NGbytes) / elapsed;
  const double flopsScaledAdd = 2 * Nf * Nlocal;
  double flopsAx = Nf * Nlocal * (12 * Nq + 15);
  if (!s.BPid) flopsAx += 5 * Nf * Nlocal;
  const double flopsDot = 3 * Nf * Nlocal;
  const double flopsPupdate = 4 * Nf * Nlocal;
  const double flops = flopsScaledAdd + flopsAx + 3 * flopsDot + flopsPupdate;
  const double gFlops = (it * flops) / elapsed / 1e9;

  std::printf("\nsummary\n");
  std::printf("  MPItasks     : 1\n");
  std::printf("  polyN        : %d\n", N);
  std::printf("  Nelements    : %d\n", (int)E);
  std::printf("  Nfields      : %d\n", Nf);
  std::printf("  iterations   : %d\n", it);
  std::printf("  Nrepetitions : %d\n", s.Ntests);
  std::printf("  elapsed time : %g s\n", s.Ntests * elapsed);
  std::printf("  throughput   : %g GDOF/s/iter\n", gDOFs);
  std::printf("  bandwidth    : %g GB/s\n", bw);
  std::printf("  GFLOPS/s     : %g\n", gFlops);
  if (s.profiling) {
    double etime[5];
    const int rc = nbx_bp_profile(bp, etime, 0);
    if (rc < 0) die("nbx_bp_profile", rc);
    std::printf("\nbreakdown\n");
    std::printf("  local Ax  : %g s\n", etime[0]);
    std::printf("  gs        : %g s\n", etime[1]);
    std::printf("  updatePCG : %g s\n", etime[2]);
    std::printf("  dot       : %g s\n", etime[3]);
    std::printf("  preco     : %g s\n\n", etime[4]);
  }

  hipFree(r);
  hipFree(x);
  nbx_bp_free(bp);
  return 0;
}

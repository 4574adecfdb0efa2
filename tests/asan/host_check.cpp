// Host code of the library under AddressSanitizer: the GLL basis, the seeded
// stream, the box mesh producers (nekbench_amd/csrc/nbx_mesh.cpp) and the
// setup-file reader of the drivers (nekbench_amd/apps/nbx_setup_ini.h), on
// the orders, box shapes and setup texts the tests use.  Built and run by
// tests/test_asan.py with g++ -fsanitize=address,undefined; any memory error
// aborts with a report.  Test infrastructure only.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "nbx.h"
#include "nbx_setup_ini.h"

namespace nbx {
// the library's error sink (nbx_core.cpp) without the HIP runtime
int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vfprintf(stderr, fmt, ap);
  va_end(ap);
  std::fputc('\n', stderr);
  return code;
}
}  // namespace nbx

static int check(bool ok, const char* what)
{
  if (!ok) std::fprintf(stderr, "host_check: %s failed\n", what);
  return ok ? 0 : 1;
}

int main(int argc, char** argv)
{
  int bad = 0;
  for (int N = 1; N <= 15; ++N) {
    std::vector<double> z(N + 1), w(N + 1), D((N + 1) * (N + 1));
    bad |= check(nbx_gll(N, z.data(), w.data(), D.data()) == 0, "gll");
  }
  std::vector<double> u(10007);
  bad |= check(nbx_urand_fill(5, 3, (int64_t)u.size(), u.data()) == 0, "urand");
  struct Shape { int N, NX, NY, NZ, periodic; double perturb, zaspect; };
  const Shape shapes[] = {{1, 3, 3, 3, 1, 0.0, 0.0}, {3, 3, 2, 2, 0, 0.1, 0.0}, {7, 2, 2, 3, 0, 0.1, 10.0},
                          {9, 2, 2, 1, 0, 0.1, 0.0},  {5, 4, 3, 3, 1, 0.0, 0.0}, {2, 5, 1, 1, 0, 0.2, 3.0}};
  for (const Shape& s : shapes) {
    nbx_box b = {s.N, s.NX, s.NY, s.NZ, s.periodic, s.perturb, s.zaspect, 7, 0, 1, 0, 1, 0, 1};
    const size_t E = (size_t)s.NX * s.NY * s.NZ, Np = (size_t)(s.N + 1) * (s.N + 1) * (s.N + 1);
    std::vector<double> x(E * Np), y(E * Np), zc(E * Np), g(E * 7 * Np);
    std::vector<nbx_hlong> ids(E * Np);
    double minJ = 0;
    for (int mask = 0; mask < 2; ++mask)
      bad |= check(nbx_box_mesh(&b, mask, x.data(), y.data(), zc.data(), g.data(), ids.data(), &minJ) == 0 &&
                       minJ > 0,
                   "box_mesh");
    bad |= check(nbx_box_mesh(&b, 1, nullptr, nullptr, nullptr, g.data(), nullptr, nullptr) == 0, "box_mesh ggeo");
  }
  nbx_box badbox = {0, 0, 1, 1, 0, 0.0, 0.0, 1, 0, 1, 0, 1, 0, 1};
  bad |= check(nbx_box_mesh(&badbox, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) < 0, "box_mesh refusal");
  const char* texts[] = {
      "[POLYNOMIAL DEGREE]\n7\n[BOX NX]\n16 # comment\n[KRYLOV SOLVER]\nPCG+FLEXIBLE\n",
      "/* block\n comment */[A]\n\"quoted [text]\"\n[B]\n\\[literal\\]\n[C]",
      "", "[", "[]\n", "[KEY", "#only a comment", "[X]\n/* unterminated", "[Y]\n\"unterminated quote\n[Z]\n1\n"};
  for (const char* t : texts) {
    nbx_app::SetupIni ini;
    ini.parse(t);
    std::string v;
    (void)ini.get("POLYNOMIAL DEGREE", v);
    (void)ini.compare("KRYLOV SOLVER", "FLEXIBLE");
    (void)ini.has("A");
  }
  for (int a = 1; a < argc; ++a) {  // setup files given on the command line
    nbx_app::SetupIni ini;
    bad |= check(ini.load(argv[a]), "load setup file");
  }
  std::printf(bad ? "host_check: FAILED\n" : "host_check: ok\n");
  return bad;
}

// The reference's native-kernel symbols, served by the gfx950 operator.
//
// With THREAD MODEL NATIVE+SERIAL/OPENMP the reference compiles
// axhelm/kernel/CPU/axhelm.c once per order (-Dp_Nq, -Dp_Np) and loads its
// kernels by name (axhelm/kernelHelper.cpp:41-50); each is
//   extern "C" void axhelm_bk_v0(const dlong& Nelements, const dlong& offset,
//                                const dfloat* ggeo, const dfloat* D,
//                                const dfloat* lambda, const dfloat* q, dfloat* Aq)
// (axhelm/kernel/CPU/axhelm.c:229-235; references are pointers at the ABI
// level), and the Partial forms take `const dlong* elementList` after offset
// (axhelm/kernel/VOLTA/axhelm.okl:158-165).  This file is compiled the same
// way, once per order with -DNBX_NATIVE_N=<N>, into
// nekbench_amd/native/libaxhelm_N<N>.so, and exports those symbols with that
// signature; each one runs nbx_axhelm (FP64) on the null stream.  The arrays
// are device pointers (e.g. the OCCA HIP device's memory), dfloat = double,
// dlong = int32.  The reference's kernels return nothing; a failed call leaves
// its code in axhelm_native_status() and nbx_last_error() has the message.
#include <cstdint>
#include <cstdio>

#include "nbx.h"

#ifndef NBX_NATIVE_N
#error "compile once per order with -DNBX_NATIVE_N=<N>"
#endif

namespace {
thread_local int g_status = 0;

void run(int nf, int bk, const int32_t& E, const int32_t& offset, const int32_t* elementList, const double* ggeo,
         const double* D, const double* lambda, const double* q, double* Aq)
{
  g_status = nbx_axhelm(NBX_NATIVE_N, nf, bk, NBX_F64, E, offset, elementList, ggeo, D, lambda, q, Aq, nullptr);
  if (g_status != 0) std::fprintf(stderr, "axhelm (N=%d): %s\n", NBX_NATIVE_N, nbx_last_error());
}
}  // namespace

extern "C" {

// 0 after a successful call, else the NBX_ERR_* code of the last call on this thread
int axhelm_native_status(void) { return g_status; }
int axhelm_native_order(void) { return NBX_NATIVE_N; }

void axhelm_v0(const int32_t& E, const int32_t& offset, const double* ggeo, const double* D, const double* lambda,
               const double* q, double* Aq)
{
  run(1, 0, E, offset, nullptr, ggeo, D, lambda, q, Aq);
}

void axhelm_bk_v0(const int32_t& E, const int32_t& offset, const double* ggeo, const double* D, const double* lambda,
                  const double* q, double* Aq)
{
  run(1, 1, E, offset, nullptr, ggeo, D, lambda, q, Aq);
}

void axhelm_n3_v0(const int32_t& E, const int32_t& offset, const double* ggeo, const double* D, const double* lambda,
                  const double* q, double* Aq)
{
  run(3, 0, E, offset, nullptr, ggeo, D, lambda, q, Aq);
}

void axhelm_bk_n3_v0(const int32_t& E, const int32_t& offset, const double* ggeo, const double* D,
                     const double* lambda, const double* q, double* Aq)
{
  run(3, 1, E, offset, nullptr, ggeo, D, lambda, q, Aq);
}

void axhelmPartial_v0(const int32_t& E, const int32_t& offset, const int32_t* elementList, const double* ggeo,
                      const double* D, const double* lambda, const double* q, double* Aq)
{
  run(1, 0, E, offset, elementList, ggeo, D, lambda, q, Aq);
}

void axhelmPartial_bk_v0(const int32_t& E, const int32_t& offset, const int32_t* elementList, const double* ggeo,
                         const double* D, const double* lambda, const double* q, double* Aq)
{
  run(1, 1, E, offset, elementList, ggeo, D, lambda, q, Aq);
}

void axhelmPartial_n3_v0(const int32_t& E, const int32_t& offset, const int32_t* elementList, const double* ggeo,
                         const double* D, const double* lambda, const double* q, double* Aq)
{
  run(3, 0, E, offset, elementList, ggeo, D, lambda, q, Aq);
}

void axhelmPartial_bk_n3_v0(const int32_t& E, const int32_t& offset, const int32_t* elementList, const double* ggeo,
                            const double* D, const double* lambda, const double* q, double* Aq)
{
  run(3, 1, E, offset, elementList, ggeo, D, lambda, q, Aq);
}

}  // extern "C"

// axhelm instantiations for N = 1..15 with FP32 operands and arithmetic
// (kernels in nbx_axhelm_impl.h; the operator is described in nbx_axhelm.hip).
#include "nbx_axhelm_impl.h"

namespace nbx {

int axhelm_dispatch_f32(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                        const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                        const void* q, void* Aq, const int* skip, hipStream_t s)
{
  return dispatch_n<float>(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip, s);
}

int axhelm_pcg_dispatch_f32(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                            const void* D, const void* lambda, const double* z, double* p, double* Ap,
                            double* partials, const PcgFlags* S, hipStream_t s)
{
  return axhelm_pcg_launch_t<float, float>(N, Nf, bk, Nelements, offset, ggeo, D, lambda, z, p, Ap, partials, S, s);
}

}  // namespace nbx

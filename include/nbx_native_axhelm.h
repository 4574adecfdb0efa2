/*
 * nbx_native_axhelm.h -- the reference's native axhelm kernel symbols, served
 * by the gfx950 operator (nekbench_amd/native/libaxhelm_N<N>.so, one library
 * per polynomial order as the reference compiles axhelm/kernel/CPU/axhelm.c
 * per p_Nq).
 *
 * Each replaces the symbol of the same name that axhelm/kernelHelper.cpp:41-50
 * loads from the reference's compiled CPU kernel (NATIVE+SERIAL/OPENMP thread
 * models): axhelm_v0 (axhelm/kernel/CPU/axhelm.c:27-33), axhelm_n3_v0
 * (:113-119), axhelm_bk_v0 (:229-235), axhelm_bk_n3_v0 (:307-313), and the
 * Partial forms with the element list after offset
 * (axhelm/kernel/VOLTA/axhelm.okl:158-165).  Same argument order; the
 * reference's `const dlong&` is a pointer at the ABI level, so C callers pass
 * the address of the value.  Arrays are device pointers (FP64, dlong = int32);
 * the work is enqueued on the null stream.  Like the reference's kernels they
 * return nothing: axhelm_native_status() is 0 after a successful call, else
 * the NBX_ERR_* code (message in nbx_last_error()).
 */
#ifndef NBX_NATIVE_AXHELM_H
#define NBX_NATIVE_AXHELM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

int axhelm_native_status(void);
int axhelm_native_order(void); /* N of this library */

void axhelm_v0(const int32_t* Nelements, const int32_t* offset, const double* ggeo, const double* D,
               const double* lambda, const double* q, double* Aq);
void axhelm_bk_v0(const int32_t* Nelements, const int32_t* offset, const double* ggeo, const double* D,
                  const double* lambda, const double* q, double* Aq);
void axhelm_n3_v0(const int32_t* Nelements, const int32_t* offset, const double* ggeo, const double* D,
                  const double* lambda, const double* q, double* Aq);
void axhelm_bk_n3_v0(const int32_t* Nelements, const int32_t* offset, const double* ggeo, const double* D,
                     const double* lambda, const double* q, double* Aq);
void axhelmPartial_v0(const int32_t* Nelements, const int32_t* offset, const int32_t* elementList,
                      const double* ggeo, const double* D, const double* lambda, const double* q, double* Aq);
void axhelmPartial_bk_v0(const int32_t* Nelements, const int32_t* offset, const int32_t* elementList,
                         const double* ggeo, const double* D, const double* lambda, const double* q, double* Aq);
void axhelmPartial_n3_v0(const int32_t* Nelements, const int32_t* offset, const int32_t* elementList,
                         const double* ggeo, const double* D, const double* lambda, const double* q, double* Aq);
void axhelmPartial_bk_n3_v0(const int32_t* Nelements, const int32_t* offset, const int32_t* elementList,
                            const double* ggeo, const double* D, const double* lambda, const double* q, double* Aq);

#ifdef __cplusplus
}
#endif
#endif

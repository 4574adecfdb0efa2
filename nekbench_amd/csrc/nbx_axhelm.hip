// axhelm: tensor-product Helmholtz operator on hexahedral spectral elements.
//
// Reference semantics: axhelm/kernel/CPU/axhelm.c:27-413 (v0, n3_v0, bk_v0,
// bk_n3_v0), the Partial forms axhelm/kernel/VOLTA/axhelm.okl:158-290 and
// the host oracle axhelm/axhelmReference.cpp:19-101.
//
// Mapping to gfx950: one workgroup per element with Nq*Nq lanes (one wave64
// at N=7); lane (i,j) owns the node column k = 0..Nq-1 in registers.  Each
// k-slab of q goes through LDS for the r and s derivatives; the t derivative
// and the t-part of the transpose stay in the register column.  ggeo is
// streamed plane by plane, coalesced (64 consecutive doubles per load per
// factor at N=7).  The four D vectors a lane needs (row i, row j, column i,
// column j) sit in registers; D[k][:] is read from LDS as a broadcast.
//
// FP64 arithmetic order (identical to the CPU oracle, bit-exact):
//   qr = fma chain over m of D[i][m] q(m,j,k), qs and qt alike
//   Gqr = fma(G02,qt, fma(G01,qs, G00*qr)), Gqs, Gqt alike; non-BK * lambda0
//   Aq  = ((Aqr + Aqs) + Aqt), non-BK: fma(GWJ*lambda1, q, .)
#include "nbx_internal.h"
#include <cstring>
#include <cstdio>
#include <cstdlib>

namespace nbx {

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c);
template <>
__device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T, int NQ, int NF, bool BK, bool PARTIAL>
__global__ __launch_bounds__(NQ* NQ) void k_axhelm(const int32_t Nelements, const int32_t offset,
                                                   const int32_t* __restrict__ elementList,
                                                   const T* __restrict__ ggeo, const T* __restrict__ D,
                                                   const T* __restrict__ lambda, const T* __restrict__ q,
                                                   T* __restrict__ Aq, const int* __restrict__ skip)
{
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  if (skip && *skip) return;
  __shared__ T s_D[NQ2];
  __shared__ T s_q[NQ2];
  __shared__ T s_Gr[NQ2];
  __shared__ T s_Gs[NQ2];

  const int t = threadIdx.x;
  const int i = t % NQ;
  const int j = t / NQ;
  const int64_t e = PARTIAL ? (int64_t)elementList[blockIdx.x] : (int64_t)blockIdx.x;

  s_D[t] = D[t];
  __syncthreads();
  T Di[NQ], Dj[NQ], DTi[NQ], DTj[NQ];
#pragma unroll
  for (int m = 0; m < NQ; ++m) {
    Di[m] = s_D[i * NQ + m];
    Dj[m] = s_D[j * NQ + m];
    DTi[m] = s_D[m * NQ + i];
    DTj[m] = s_D[m * NQ + j];
  }

  const T* __restrict__ g = ggeo + e * (int64_t)(kNggeo * NP) + t;
  const T* __restrict__ l0 = lambda + e * NP + t;
  const T* __restrict__ l1 = lambda + e * NP + offset + t;

#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int64_t base = e * NP + (int64_t)f * offset + t;
    T rq[NQ], rAt[NQ], rArs[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      rq[k] = q[base + k * NQ2];
      rAt[k] = T(0);
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      __syncthreads();
      s_q[t] = rq[k];
      __syncthreads();
      T qr = T(0), qs = T(0), qt = T(0);
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        qr = fmaT(Di[m], s_q[j * NQ + m], qr);
        qs = fmaT(Dj[m], s_q[m * NQ + i], qs);
        qt = fmaT(s_D[k * NQ + m], rq[m], qt);
      }
      const int n = k * NQ2;
      const T G00 = g[kG00 * NP + n], G01 = g[kG01 * NP + n], G02 = g[kG02 * NP + n];
      const T G11 = g[kG11 * NP + n], G12 = g[kG12 * NP + n], G22 = g[kG22 * NP + n];
      T Gr = fmaT(G02, qt, fmaT(G01, qs, G00 * qr));
      T Gs = fmaT(G12, qt, fmaT(G11, qs, G01 * qr));
      T Gt = fmaT(G22, qt, fmaT(G12, qs, G02 * qr));
      if (!BK) {
        const T lam0 = l0[n];
        Gr = lam0 * Gr;
        Gs = lam0 * Gs;
        Gt = lam0 * Gt;
      }
#pragma unroll
      for (int m = 0; m < NQ; ++m) rAt[m] = fmaT(s_D[k * NQ + m], Gt, rAt[m]);
      s_Gr[t] = Gr;
      s_Gs[t] = Gs;
      __syncthreads();
      T ar = T(0), as = T(0);
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        ar = fmaT(DTi[m], s_Gr[j * NQ + m], ar);
        as = fmaT(DTj[m], s_Gs[m * NQ + i], as);
      }
      rArs[k] = ar + as;
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      T res = rArs[k] + rAt[k];
      if (!BK) {
        const int n = k * NQ2;
        res = fmaT(g[kGWJ * NP + n] * l1[n], rq[k], res);
      }
      Aq[base + k * NQ2] = res;
    }
  }
}

// Local Jacobi diagonal, nekBone/kernel/updateJacobi.okl:27-108, same
// expression order as the reference (plain mul/add, no contraction).
template <int NQ>
__global__ __launch_bounds__(NQ* NQ) void k_jacobi(const int32_t Nelements, const int32_t offset,
                                                   const int32_t* __restrict__ mapB,
                                                   const double* __restrict__ ggeo,
                                                   const double* __restrict__ D,
                                                   const double* __restrict__ lambda,
                                                   double* __restrict__ diag)
{
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  __shared__ double s_D[NQ2];
  const int t = threadIdx.x, i = t % NQ, j = t / NQ;
  const int64_t e = blockIdx.x;
  s_D[t] = D[t];
  __syncthreads();
  const double* g = ggeo + e * (int64_t)(kNggeo * NP);
  const double* l0 = lambda + e * NP;
  for (int k = 0; k < NQ; ++k) {
    const int n = i + j * NQ + k * NQ2;
    const int64_t id = e * NP + n;
    double r_q = 1.0;
    if (!(mapB && mapB[id] == 1)) {
      r_q = 0.0;
      const double lambda_0 = l0[n];
      r_q += 2.0 * g[kG01 * NP + n] * lambda_0 * s_D[i * NQ + i] * s_D[j * NQ + j];
      r_q += 2.0 * g[kG02 * NP + n] * lambda_0 * s_D[i * NQ + i] * s_D[k * NQ + k];
      r_q += 2.0 * g[kG12 * NP + n] * lambda_0 * s_D[j * NQ + j] * s_D[k * NQ + k];
      for (int m = 0; m < NQ; ++m) {
        const int nr = m + j * NQ + k * NQ2, ns = i + m * NQ + k * NQ2, nt = i + j * NQ + m * NQ2;
        r_q += g[kG00 * NP + nr] * l0[nr] * s_D[m * NQ + i] * s_D[m * NQ + i];
        r_q += g[kG11 * NP + ns] * l0[ns] * s_D[m * NQ + j] * s_D[m * NQ + j];
        r_q += g[kG22 * NP + nt] * l0[nt] * s_D[m * NQ + k] * s_D[m * NQ + k];
      }
      r_q += g[kGWJ * NP + n] * lambda[id + offset];
    }
    diag[id] = r_q;
  }
}

template <typename T, int NQ, int NF, bool BK, bool PARTIAL>
static int launch_axhelm_t(int32_t Nelements, int32_t offset, const int32_t* elementList,
                           const void* ggeo, const void* D, const void* lambda, const void* q, void* Aq,
                           const int* skip, hipStream_t s)
{
  hipLaunchKernelGGL((k_axhelm<T, NQ, NF, BK, PARTIAL>), dim3(Nelements), dim3(NQ * NQ), 0, s,
                     Nelements, offset, elementList, (const T*)ggeo, (const T*)D, (const T*)lambda,
                     (const T*)q, (T*)Aq, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

template <typename T, int NQ>
static int dispatch_flags(int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                          const int32_t* elementList, const void* ggeo, const void* D,
                          const void* lambda, const void* q, void* Aq, const int* skip, hipStream_t s)
{
#define NBX_AX_CASE(NFV, BKV, PV)                                                               \
  if (Nf == NFV && (bk != 0) == BKV && partial == PV)                                           \
    return launch_axhelm_t<T, NQ, NFV, BKV, PV>(Nelements, offset, elementList, ggeo, D, lambda, \
                                                 q, Aq, skip, s);
  NBX_AX_CASE(1, false, false)
  NBX_AX_CASE(1, true, false)
  NBX_AX_CASE(3, false, false)
  NBX_AX_CASE(3, true, false)
  NBX_AX_CASE(1, false, true)
  NBX_AX_CASE(1, true, true)
  NBX_AX_CASE(3, false, true)
  NBX_AX_CASE(3, true, true)
#undef NBX_AX_CASE
  return fail(NBX_ERR_UNSUPPORTED, "axhelm: Nfields=%d not supported (1 or 3)", Nf);
}

template <typename T>
static int dispatch_n(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                      const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                      const void* q, void* Aq, const int* skip, hipStream_t s)
{
  switch (N) {
#define NBX_N_CASE(NV) \
  case NV:             \
    return dispatch_flags<T, NV + 1>(Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip, s);
    NBX_N_CASE(1)
    NBX_N_CASE(2)
    NBX_N_CASE(3)
    NBX_N_CASE(4)
    NBX_N_CASE(5)
    NBX_N_CASE(6)
    NBX_N_CASE(7)
    NBX_N_CASE(8)
    NBX_N_CASE(9)
    NBX_N_CASE(10)
    NBX_N_CASE(11)
    NBX_N_CASE(12)
    NBX_N_CASE(13)
    NBX_N_CASE(14)
    NBX_N_CASE(15)
#undef NBX_N_CASE
    default:
      return fail(NBX_ERR_UNSUPPORTED, "axhelm: N=%d not supported (1..15)", N);
  }
}

int axhelm_launch(int N, int Nf, int bk, nbx_prec prec, int32_t Nelements, int32_t offset,
                  const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                  const void* q, void* Aq, const int* skip, hipStream_t s)
{
  const bool partial = elementList != nullptr;
  switch (prec) {
    case NBX_F64:
      return dispatch_n<double>(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip,
                                s);
    case NBX_F32:
      return dispatch_n<float>(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip,
                               s);
    default:
      return fail(NBX_ERR_UNSUPPORTED, "axhelm: precision %d not supported by nbx_axhelm", (int)prec);
  }
}

}  // namespace nbx

using namespace nbx;

extern "C" int nbx_axhelm(int N, int Nfields, int bk, nbx_prec prec, nbx_dlong Nelements,
                          nbx_dlong offset, const nbx_dlong* elementList, const void* ggeo,
                          const void* D, const void* lambda, const void* q, void* Aq, void* stream)
{
  if (Nelements < 0) return fail(NBX_ERR_INVALID, "axhelm: Nelements < 0");
  if (Nelements == 0) return NBX_SUCCESS;
  if (!ggeo || !D || !q || !Aq) return fail(NBX_ERR_INVALID, "axhelm: null array");
  if (!bk && !lambda) return fail(NBX_ERR_INVALID, "axhelm: lambda required when bk == 0");
  const int64_t Np = (int64_t)(N + 1) * (N + 1) * (N + 1);
  if (Nfields > 1 && (int64_t)offset < (int64_t)Nelements * Np && !elementList)
    return fail(NBX_ERR_INVALID, "axhelm: offset %d < Nelements*Np", offset);
  if (!bk && (int64_t)offset < (int64_t)Nelements * Np && !elementList)
    return fail(NBX_ERR_INVALID, "axhelm: lambda offset %d < Nelements*Np", offset);
  return axhelm_launch(N, Nfields, bk, prec, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, nullptr,
                       as_stream(stream));
}

extern "C" int nbx_axhelm_by_name(const char* kernelName, int N, nbx_dlong Nelements, nbx_dlong offset,
                                  const nbx_dlong* elementList, const double* ggeo, const double* D,
                                  const double* lambda, const double* q, double* Aq, void* stream)
{
  if (!kernelName) return fail(NBX_ERR_INVALID, "axhelm_by_name: null name");
  // grammar: axhelm[Partial][_bk][_n3]_v<k>   (nekBone/setup.cpp:313-320)
  const char* p = kernelName;
  if (strncmp(p, "axhelm", 6) != 0) return fail(NBX_ERR_INVALID, "unknown kernel '%s'", kernelName);
  p += 6;
  bool partial = false, bk = false;
  int nf = 1;
  if (strncmp(p, "Partial", 7) == 0) { partial = true; p += 7; }
  if (strncmp(p, "_bk", 3) == 0) { bk = true; p += 3; }
  if (strncmp(p, "_n", 2) == 0) {
    char* end = nullptr;
    nf = (int)strtol(p + 2, &end, 10);
    p = end;
  }
  if (strncmp(p, "_v", 2) != 0) return fail(NBX_ERR_INVALID, "kernel '%s': missing _v<k>", kernelName);
  char* end = nullptr;
  const long ver = strtol(p + 2, &end, 10);
  if (end == p + 2 || *end != '\0' || ver < 0) return fail(NBX_ERR_INVALID, "kernel '%s': bad version", kernelName);
  if (partial && !elementList) return fail(NBX_ERR_INVALID, "kernel '%s' needs an elementList", kernelName);
  if (!partial && elementList) return fail(NBX_ERR_INVALID, "kernel '%s' takes no elementList", kernelName);
  return nbx_axhelm(N, nf, bk ? 1 : 0, NBX_F64, Nelements, offset, elementList, ggeo, D, lambda, q, Aq,
                    stream);
}

extern "C" int nbx_update_jacobi_local(int N, nbx_dlong Nelements, nbx_dlong offset, const int32_t* mapB,
                                       const double* ggeo, const double* D, const double* lambda,
                                       double* diagA, void* stream)
{
  if (Nelements <= 0) return Nelements == 0 ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "jacobi: Nelements < 0");
  if (!ggeo || !D || !lambda || !diagA) return fail(NBX_ERR_INVALID, "jacobi: null array");
  hipStream_t s = as_stream(stream);
  switch (N) {
#define NBX_J_CASE(NV)                                                                              \
  case NV:                                                                                          \
    hipLaunchKernelGGL(k_jacobi<NV + 1>, dim3(Nelements), dim3((NV + 1) * (NV + 1)), 0, s, Nelements, \
                       offset, mapB, ggeo, D, lambda, diagA);                                       \
    NBX_HIP_LAUNCH();                                                                               \
    return NBX_SUCCESS;
    NBX_J_CASE(1)
    NBX_J_CASE(2)
    NBX_J_CASE(3)
    NBX_J_CASE(4)
    NBX_J_CASE(5)
    NBX_J_CASE(6)
    NBX_J_CASE(7)
    NBX_J_CASE(8)
    NBX_J_CASE(9)
    NBX_J_CASE(10)
    NBX_J_CASE(11)
    NBX_J_CASE(12)
    NBX_J_CASE(13)
    NBX_J_CASE(14)
    NBX_J_CASE(15)
#undef NBX_J_CASE
    default:
      return fail(NBX_ERR_UNSUPPORTED, "jacobi: N=%d not supported", N);
  }
}

// Kernels and per-precision launchers of the axhelm operator, shared by the
// translation units that instantiate them (nbx_axhelm.hip: FP64,
// nbx_axhelm_f32.hip: FP32, nbx_axhelm_bf16.hip: bf16 geometry), so that the
// instantiations for N = 1..15 compile in parallel.  See nbx_axhelm.hip for
// the operator, its mapping and its arithmetic order.
#pragma once
#include "nbx_internal.h"
#include "nbx_reduce.h"
#include <cstring>
#include <cstdio>
#include <type_traits>
#include <cstdlib>

namespace nbx {

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c);
template <>
__device__ __forceinline__ double fmaT<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fmaT<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Load of a geometric factor / lambda stored as G, converted to the math type T.
// bf16 is stored as its 16 high bits (uint16_t).
template <typename T, typename G>
__device__ __forceinline__ T ldg(const G* p);
template <>
__device__ __forceinline__ double ldg<double, double>(const double* p) { return ldnt(p); }
template <>
__device__ __forceinline__ float ldg<float, float>(const float* p) { return ldnt(p); }
template <>
__device__ __forceinline__ float ldg<float, uint16_t>(const uint16_t* p)
{
  return __uint_as_float((uint32_t)ldnt(p) << 16);
}
// NT = false: an ordinary (cached) load, for data that other lanes of the
// workgroup read again right away
template <typename T, typename G, bool NT>
__device__ __forceinline__ T ldgc(const G* p)
{
  if constexpr (NT) {
    return ldg<T, G>(p);
  } else if constexpr (sizeof(G) == 2) {
    return __uint_as_float((uint32_t)(*p) << 16);
  } else {
    return (T)(*p);
  }
}

// Elements per workgroup: one element (Nq*Nq lanes) from Nq = 8 up; small
// orders pack several elements (threadIdx.y = slot) so that a workgroup
// holds at least 128 lanes.
template <int NQ>
struct Epb {
  static constexpr int value = NQ * NQ >= 64 ? 1 : 128 / (NQ * NQ);
};

// Per-workgroup state of the operator: LDS slabs and the lane's D vectors.
template <typename T, int NQ>
struct AxLane {
  T Di[NQ], Dj[NQ], DTi[NQ], DTj[NQ];
  int t, i, j;
  __device__ __forceinline__ void init(const T* s_D)
  {
    t = threadIdx.x;
    i = t % NQ;
    j = t / NQ;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      Di[m] = s_D[i * NQ + m];
      Dj[m] = s_D[j * NQ + m];
      DTi[m] = s_D[m * NQ + i];
      DTj[m] = s_D[m * NQ + j];
    }
  }
};

// One field of one element: out[k] = (A q)(i,j,k) for the lane's column rq.
// g, l0, l1 point at the lane's node (k = 0) of ggeo, lambda0, lambda1.
// DL: read the lane's D vectors from LDS instead of registers (frees 4*NQ
// registers; used where that buys a fourth wave per SIMD)
// NTG: nontemporal loads of the geometric factors and lambda (false when
// other thread groups of the workgroup read the same values)
// EARLY: issue each slab's geometric-factor loads before the slab's
// barriers, so they are in flight while the slab goes through LDS
template <typename T, int NQ, bool BK, typename G = T, bool DL = false, bool NTG = true, bool EARLY = false>
__device__ __forceinline__ void axhelm_column(const AxLane<T, NQ>& L, const T (&rq)[NQ], T (&out)[NQ],
                                              const T* __restrict__ s_D, T* s_q, T* s_Gr, T* s_Gs,
                                              const G* __restrict__ g, const G* __restrict__ l0,
                                              const G* __restrict__ l1)
{
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  T rAt[NQ], rArs[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) rAt[k] = T(0);
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int n = k * NQ2;
    T Gk[6];
    auto load_g = [&]() {
      Gk[0] = ldgc<T, G, NTG>(&g[kG00 * NP + n]);
      Gk[1] = ldgc<T, G, NTG>(&g[kG01 * NP + n]);
      Gk[2] = ldgc<T, G, NTG>(&g[kG02 * NP + n]);
      Gk[3] = ldgc<T, G, NTG>(&g[kG11 * NP + n]);
      Gk[4] = ldgc<T, G, NTG>(&g[kG12 * NP + n]);
      Gk[5] = ldgc<T, G, NTG>(&g[kG22 * NP + n]);
    };
    if (EARLY) load_g();
    __syncthreads();
    s_q[L.t] = rq[k];
    __syncthreads();
    T qr = T(0), qs = T(0), qt = T(0);
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      qr = fmaT(DL ? s_D[L.i * NQ + m] : L.Di[m], s_q[L.j * NQ + m], qr);
      qs = fmaT(DL ? s_D[L.j * NQ + m] : L.Dj[m], s_q[m * NQ + L.i], qs);
      qt = fmaT(s_D[k * NQ + m], rq[m], qt);
    }
    if (!EARLY) load_g();
    const T G00 = Gk[0], G01 = Gk[1], G02 = Gk[2], G11 = Gk[3], G12 = Gk[4], G22 = Gk[5];
    T Gr = fmaT(G02, qt, fmaT(G01, qs, G00 * qr));
    T Gs = fmaT(G12, qt, fmaT(G11, qs, G01 * qr));
    T Gt = fmaT(G22, qt, fmaT(G12, qs, G02 * qr));
    if (!BK) {
      const T lam0 = ldgc<T, G, NTG>(&l0[n]);
      Gr = lam0 * Gr;
      Gs = lam0 * Gs;
      Gt = lam0 * Gt;
    }
#pragma unroll
    for (int m = 0; m < NQ; ++m) rAt[m] = fmaT(s_D[k * NQ + m], Gt, rAt[m]);
    s_Gr[L.t] = Gr;
    s_Gs[L.t] = Gs;
    __syncthreads();
    T ar = T(0), as = T(0);
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      ar = fmaT(DL ? s_D[m * NQ + L.i] : L.DTi[m], s_Gr[L.j * NQ + m], ar);
      as = fmaT(DL ? s_D[m * NQ + L.j] : L.DTj[m], s_Gs[m * NQ + L.i], as);
    }
    rArs[k] = ar + as;
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    T res = rArs[k] + rAt[k];
    if (!BK) {
      const int n = k * NQ2;
      res = fmaT(ldgc<T, G, NTG>(&g[kGWJ * NP + n]) * ldgc<T, G, NTG>(&l1[n]), rq[k], res);
    }
    out[k] = res;
  }
}

// the one-field BK kernel at N = 7 reads D from LDS and fits four waves per
// SIMD (128 VGPRs): 4096 elements, the 16^3 mesh, fill the chip in one round
template <int NQ, int NF, bool BK>
struct AxDL {
  static constexpr bool value = NQ == 8 && NF == 1 && BK;
};

// Thread groups of NQ*NQ lanes per workgroup: Epb elements at small orders;
// for the three-field Helmholtz (non-BK) operator with 8 <= Nq <= 12 and
// FP64 or FP32 geometry, one group per field of one element, reading the
// shared geometric factors and lambda with cached loads; each field's
// arithmetic is the one-field column's.  Measured on 32^3 N = 7: 0.805 ->
// 0.695 ms; the BK form and bf16 geometry are faster with the field loop and
// keep it.
template <int NQ, int NF, bool BK, typename G = double>
struct AxGroups {
  static constexpr bool fields = NF == 3 && !BK && Epb<NQ>::value == 1 && NQ <= 12 && sizeof(G) != 2;
  static constexpr int value = fields ? NF : Epb<NQ>::value;
};

// The FP64 field-group kernels read D from LDS as well (4 Nq fewer doubles
// in registers) and keep three waves per SIMD up to Nq = 10 (168 VGPRs; left
// alone they take 256 and run two); the one-field BK kernel at N = 7 keeps
// four (128 VGPRs).  The FP32-arithmetic kernels keep D in registers: their
// vectors take half the registers, and reading D from LDS made LDS the bound
// (one-field BK at N = 7 on 32^3: bf16 geometry 80 -> 65 us, FP32 89 -> 84 us
// with D in registers, the same bits).
template <typename T, int NQ, int NF, bool BK>
struct AxUseDL {
  static constexpr bool value = sizeof(T) == 8 && (AxDL<NQ, NF, BK>::value || AxGroups<NQ, NF, BK>::fields);
};
// Early geometric-factor loads (axhelm_column's EARLY), measured on 32^3,
// N = 7, one field: the FP64 BK PCG-form operator 230 -> 219 us (and no
// spill where it had 12 bytes), the FP32 operator 83.7 -> 82.3 us; bf16
// geometry got slower (63.2 -> 64.4 us), the stand-alone FP64 kernel spilled
// 124 bytes and the FP64 non-BK PCG form grew from 168 to 256 VGPRs (two
// waves per SIMD instead of three), so those keep the loads after the barriers.
template <typename T, int NQ, int NF, bool BK, typename G, bool PCG>
struct AxEarly {
  static constexpr bool value = NQ == 8 && NF == 1 && sizeof(G) != 2 && (sizeof(T) == 4 || (PCG && BK));
};
template <typename T, int NQ, int NF, bool BK>
struct AxMinWaves {
  static constexpr int value =
      AxDL<NQ, NF, BK>::value ? 4 : (AxGroups<NQ, NF, BK>::fields && sizeof(T) == 8 && NQ <= 10 ? 3 : 1);
};

template <typename T, int NQ, int NF, bool BK, bool PARTIAL, typename G = T>
__global__ __launch_bounds__((NQ * NQ * AxGroups<NQ, NF, BK, G>::value), (AxMinWaves<T, NQ, NF, BK>::value)) void k_axhelm(const int32_t Nelements, const int32_t offset,
                                                                  const int32_t* __restrict__ elementList,
                                                                  const G* __restrict__ ggeo,
                                                                  const T* __restrict__ D,
                                                                  const G* __restrict__ lambda,
                                                                  const T* __restrict__ q, T* __restrict__ Aq,
                                                                  const int* __restrict__ skip)
{
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int EPB = Epb<NQ>::value;
  constexpr bool FG = AxGroups<NQ, NF, BK, G>::fields;
  constexpr int GR = AxGroups<NQ, NF, BK, G>::value;
  if (skip && *skip) return;
  __shared__ T s_D[NQ2], s_q[GR][NQ2], s_Gr[GR][NQ2], s_Gs[GR][NQ2];
  const int slot = GR > 1 ? (int)threadIdx.y : 0;  // element slot, or field when FG
  const int32_t idx = FG ? (int32_t)blockIdx.x : (int32_t)blockIdx.x * EPB + slot;
  const bool active = EPB == 1 || idx < Nelements;
  // an idle slot works on the last element without storing: every lane
  // reaches the workgroup barriers
  const int32_t ie = EPB == 1 ? (int32_t)blockIdx.x : (active ? idx : Nelements - 1);
  const int64_t e = PARTIAL ? (int64_t)elementList[ie] : (int64_t)ie;
  if (slot == 0) s_D[threadIdx.x] = D[threadIdx.x];
  __syncthreads();
  AxLane<T, NQ> L;
  L.init(s_D);
  const G* g = ggeo + e * (int64_t)(kNggeo * NP) + L.t;
  const G* l0 = lambda + e * NP + L.t;
  const G* l1 = lambda + e * NP + offset + L.t;
  auto field = [&](int f) {
    const int64_t base = e * NP + (int64_t)f * offset + L.t;
    T rq[NQ], out[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) rq[k] = ldnt(&q[base + k * NQ2]);
    axhelm_column<T, NQ, BK, G, AxUseDL<T, NQ, NF, BK>::value, !FG, AxEarly<T, NQ, NF, BK, G, false>::value>(L, rq, out, s_D, s_q[slot], s_Gr[slot],
                                                                     s_Gs[slot], g, l0, l1);
    if (active) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) stnt(out[k], &Aq[base + k * NQ2]);
    }
  };
  if constexpr (FG) {
    field(slot);
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) field(f);
  }
}

// FP64 BK one-field N = 7 operator, two waves per element (defined in nbx_axhelm.hip).
__global__ __launch_bounds__(128, 8) void k_axhelm_bk8_split(const int32_t Nelements,
                                                             const double* ggeo,
                                                             const double* __restrict__ D,
                                                             const double* __restrict__ q, double* __restrict__ Aq,
                                                             const int* __restrict__ skip);

// PCG form (FP64): p = z + beta p (BPScaledAdd, nekBone/solve.cpp:100), then
// Ap_local = A p, and the element's share of pAp = sum_nodes p * Ap_local.
// Because p is continuous and zero on the mask, sum_local p (A_L p) equals
// <p, gs(A_L p)>_w exactly in real arithmetic (the weights 1/multiplicity
// undo the gather-scatter); this value is what BPAxOperator's dot returns.
// Lane (i,j) accumulates fma(p, Ap) over fields then k -- with one thread
// group per field (three fields, non-BK, 8 <= Nq <= 12) each field's chain runs over k and
// the lane takes (field0 + field1) + field2; the NQ*NQ lanes are summed by a
// halving tree over the next power of two (zero padded); the element
// partials go to partials[e].
template <int NQ, int NF, bool BK, typename T, typename G>
__global__ __launch_bounds__((NQ * NQ * AxGroups<NQ, NF, BK, G>::value), (AxMinWaves<T, NQ, NF, BK>::value)) void k_axhelm_pcg(const int32_t Nelements, const int32_t offset,
                                                                      const G* __restrict__ ggeo,
                                                                      const T* __restrict__ D,
                                                                      const G* __restrict__ lambda,
                                                                      const double* __restrict__ z,
                                                                      double* __restrict__ p,
                                                                      double* __restrict__ Ap,
                                                                      double* __restrict__ partials,
                                                                      const PcgFlags* __restrict__ S)
{
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int EPB = Epb<NQ>::value;
  constexpr bool FG = AxGroups<NQ, NF, BK, G>::fields;
  constexpr int GR = AxGroups<NQ, NF, BK, G>::value;
  constexpr int P2 = NQ2 <= 1 ? 1 : (NQ2 <= 2 ? 2 : (NQ2 <= 4 ? 4 : (NQ2 <= 8 ? 8 : (NQ2 <= 16 ? 16 : (NQ2 <= 32 ? 32 : (NQ2 <= 64 ? 64 : (NQ2 <= 128 ? 128 : 256)))))));
  if (S->done) return;
  const double beta = S->beta;
  __shared__ T s_D[NQ2], s_q[GR][NQ2], s_Gr[GR][NQ2], s_Gs[GR][NQ2];
  __shared__ double s_red[GR][P2];
  const int slot = GR > 1 ? (int)threadIdx.y : 0;  // element slot, or field when FG
  const int32_t idx = FG ? (int32_t)blockIdx.x : (int32_t)blockIdx.x * EPB + slot;
  const bool active = EPB == 1 || idx < Nelements;
  // idle slots compute on the last element and store nothing
  const int64_t e = EPB == 1 ? (int64_t)blockIdx.x : (active ? idx : Nelements - 1);
  if (slot == 0) s_D[threadIdx.x] = D[threadIdx.x];
  __syncthreads();
  AxLane<T, NQ> L;
  L.init(s_D);
  const G* g = ggeo + e * (int64_t)(kNggeo * NP) + L.t;
  const G* l0 = lambda + e * NP + L.t;
  const G* l1 = lambda + e * NP + offset + L.t;
  double acc = 0.0;
  auto field = [&](int f) {
    const int64_t base = e * NP + (int64_t)f * offset + L.t;
    double pk[NQ];
    T rq[NQ], out[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      pk[k] = __builtin_fma(beta, p[base + k * NQ2], ldnt(&z[base + k * NQ2]));
      rq[k] = (T)pk[k];
    }
    if (EPB > 1) __syncthreads();  // all slots read p before any writes it (idle slots alias the last element)
    if (active) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) p[base + k * NQ2] = pk[k];
    }
    axhelm_column<T, NQ, BK, G, AxUseDL<T, NQ, NF, BK>::value, !FG, AxEarly<T, NQ, NF, BK, G, true>::value>(L, rq, out, s_D, s_q[slot], s_Gr[slot],
                                                                     s_Gs[slot], g, l0, l1);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const double o = (double)out[k];
      if (active) stnt(o, &Ap[base + k * NQ2]);
      acc = __builtin_fma(pk[k], o, acc);
    }
  };
  if constexpr (FG) {
    field(slot);
    double* fa = s_red[slot];
    fa[L.t] = acc;
    __syncthreads();
    acc = (s_red[0][L.t] + s_red[1][L.t]) + s_red[2][L.t];
    __syncthreads();
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) field(f);
  }
  double sum;
  if (NQ2 == 64) {
    sum = wave_sum(acc);
  } else {
    double* red = s_red[slot];
    red[L.t] = acc;
    if (L.t + NQ2 < P2) red[L.t + NQ2] = 0.0;
    for (int h = P2 / 2; h >= 1; h >>= 1) {
      __syncthreads();
      if (L.t < h) red[L.t] = red[L.t] + red[L.t + h];
    }
    __syncthreads();
    sum = red[0];
  }
  if (L.t == 0 && active && (!FG || slot == 0)) partials[e] = sum;
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

template <typename T, int NQ, int NF, bool BK, bool PARTIAL, typename G = T>
static int launch_axhelm_t(int32_t Nelements, int32_t offset, const int32_t* elementList,
                           const void* ggeo, const void* D, const void* lambda, const void* q, void* Aq,
                           const int* skip, hipStream_t s)
{
  if constexpr (std::is_same<T, double>::value && std::is_same<G, double>::value && NQ == 8 && NF == 1 && BK &&
                !PARTIAL) {
    if (Nelements <= 0) return NBX_SUCCESS;
    hipLaunchKernelGGL(k_axhelm_bk8_split, dim3(Nelements), dim3(64, 2), 0, s, Nelements, (const double*)ggeo,
                       (const double*)D, (const double*)q, (double*)Aq, skip);
    NBX_HIP_LAUNCH();
    return NBX_SUCCESS;
  }
  constexpr int EPB = Epb<NQ>::value;
  constexpr int GR = AxGroups<NQ, NF, BK, G>::value;
  hipLaunchKernelGGL((k_axhelm<T, NQ, NF, BK, PARTIAL, G>), dim3((Nelements + EPB - 1) / EPB), dim3(NQ * NQ, GR), 0,
                     s, Nelements, offset, elementList, (const G*)ggeo, (const T*)D, (const G*)lambda,
                     (const T*)q, (T*)Aq, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

template <typename T, int NQ, typename G = T>
static int dispatch_flags(int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                          const int32_t* elementList, const void* ggeo, const void* D,
                          const void* lambda, const void* q, void* Aq, const int* skip, hipStream_t s)
{
  if (!std::is_same<T, G>::value) {  // bf16 geometry: one field, whole mesh
    if (Nf != 1 || partial) return fail(NBX_ERR_UNSUPPORTED, "axhelm bf16: one field, no element list");
    if (bk) return launch_axhelm_t<T, NQ, 1, true, false, G>(Nelements, offset, nullptr, ggeo, D, lambda, q, Aq, skip, s);
    return launch_axhelm_t<T, NQ, 1, false, false, G>(Nelements, offset, nullptr, ggeo, D, lambda, q, Aq, skip, s);
  } else {
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
  }
  return fail(NBX_ERR_UNSUPPORTED, "axhelm: Nfields=%d not supported (1 or 3)", Nf);
}

template <typename T, typename G = T>
static int dispatch_n(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                      const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                      const void* q, void* Aq, const int* skip, hipStream_t s)
{
  switch (N) {
#define NBX_N_CASE(NV) \
  case NV:             \
    return dispatch_flags<T, NV + 1, G>(Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip, s);
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

template <typename T, typename G>
static int axhelm_pcg_launch_t(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                               const void* D, const void* lambda, const double* z, double* p, double* Ap,
                               double* partials, const PcgFlags* S, hipStream_t s)
{
#define NBX_AXP(NV)                                                                                          \
  case NV:                                                                                                   \
    if (Nelements < 1) break;                                                                                \
    if (Nf == 1 && bk)                                                                                       \
      hipLaunchKernelGGL((k_axhelm_pcg<NV + 1, 1, true, T, G>), dim3((Nelements + Epb<NV + 1>::value - 1) / Epb<NV + 1>::value), dim3((NV + 1) * (NV + 1), Epb<NV + 1>::value), 0, \
                         s, Nelements, offset, (const G*)ggeo, (const T*)D, (const G*)lambda, z, p, Ap,       \
                         partials, S);                                                                       \
    else if (Nf == 1)                                                                                        \
      hipLaunchKernelGGL((k_axhelm_pcg<NV + 1, 1, false, T, G>), dim3((Nelements + Epb<NV + 1>::value - 1) / Epb<NV + 1>::value), dim3((NV + 1) * (NV + 1), Epb<NV + 1>::value),   \
                         0, s, Nelements, offset, (const G*)ggeo, (const T*)D, (const G*)lambda, z, p, Ap,    \
                         partials, S);                                                                       \
    else if (bk)                                                                                             \
      hipLaunchKernelGGL((k_axhelm_pcg<NV + 1, 3, true, T, G>), dim3((Nelements + Epb<NV + 1>::value - 1) / Epb<NV + 1>::value), dim3((NV + 1) * (NV + 1), AxGroups<NV + 1, 3, true, G>::value), 0, \
                         s, Nelements, offset, (const G*)ggeo, (const T*)D, (const G*)lambda, z, p, Ap,       \
                         partials, S);                                                                       \
    else                                                                                                     \
      hipLaunchKernelGGL((k_axhelm_pcg<NV + 1, 3, false, T, G>), dim3((Nelements + Epb<NV + 1>::value - 1) / Epb<NV + 1>::value), dim3((NV + 1) * (NV + 1), AxGroups<NV + 1, 3, false, G>::value),   \
                         0, s, Nelements, offset, (const G*)ggeo, (const T*)D, (const G*)lambda, z, p, Ap,    \
                         partials, S);                                                                       \
    break;
  if (Nf != 1 && Nf != 3) return fail(NBX_ERR_UNSUPPORTED, "axhelm_pcg: Nfields=%d", Nf);
  switch (N) {
    NBX_AXP(1)
    NBX_AXP(2)
    NBX_AXP(3)
    NBX_AXP(4)
    NBX_AXP(5)
    NBX_AXP(6)
    NBX_AXP(7)
    NBX_AXP(8)
    NBX_AXP(9)
    NBX_AXP(10)
    NBX_AXP(11)
    NBX_AXP(12)
    NBX_AXP(13)
    NBX_AXP(14)
    NBX_AXP(15)
    default:
      return fail(NBX_ERR_UNSUPPORTED, "axhelm_pcg: N=%d", N);
  }
#undef NBX_AXP
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}


// per-precision entry points of the dispatch (one translation unit each)
int axhelm_dispatch_f64(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                        const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                        const void* q, void* Aq, const int* skip, hipStream_t s);
int axhelm_dispatch_f32(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                        const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                        const void* q, void* Aq, const int* skip, hipStream_t s);
int axhelm_dispatch_bf16(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                         const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                         const void* q, void* Aq, const int* skip, hipStream_t s);
int axhelm_pcg_dispatch_f64(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                            const void* D, const void* lambda, const double* z, double* p, double* Ap,
                            double* partials, const PcgFlags* S, hipStream_t s);
int axhelm_pcg_dispatch_f32(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                            const void* D, const void* lambda, const double* z, double* p, double* Ap,
                            double* partials, const PcgFlags* S, hipStream_t s);
int axhelm_pcg_dispatch_bf16(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                             const void* D, const void* lambda, const double* z, double* p, double* Ap,
                             double* partials, const PcgFlags* S, hipStream_t s);

}  // namespace nbx

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
#include "nbx_axhelm_impl.h"

namespace nbx {

// FP64 BK one-field operator at N = 7 with each element split over two
// waves: wave w owns the slabs k = 4w .. 4w+3.  The element's q sits in LDS
// (4 KB), so the t derivative reads the column from there and no wave keeps
// a q column in registers; each wave walks four slabs (half the chain of
// dependent load rounds), and eight waves per SIMD (64 VGPRs) keep a 16^3
// mesh (8192 waves) in one round.  After the slab loop, Gt of every slab
// replaces q in LDS and each output node takes the t transpose as the
// one-wave column does, an fma chain over the slabs in increasing order:
// the arithmetic is the one-wave kernel's, bit for bit.
__global__ __launch_bounds__(128, 8) void k_axhelm_bk8_split(const int32_t Nelements,
                                                             const double* ggeo,
                                                             const double* __restrict__ D,
                                                             const double* __restrict__ q, double* __restrict__ Aq,
                                                             const int* __restrict__ skip)
{
  constexpr int NQ = 8, NQ2 = 64, NP = 512, H = 4;
  if (skip && *skip) return;
  __shared__ double s_D[NQ2], s_qt[NP], s_Gr[2][2][NQ2], s_Gs[2][2][NQ2];
  // w is uniform in each wave: D's row and column k come from scalar loads
  const int t = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y), i = t & 7, j = t >> 3;
  const int64_t e = blockIdx.x;
  if (w == 0) s_D[t] = D[t];
  // uniform element bases, one 32-bit lane offset: the loads take a scalar
  // base and a shared lane register rather than a 64-bit address each
  const double* qe = q + e * NP;
  const int o = H * w * NQ2 + t;
#pragma unroll
  for (int kk = 0; kk < H; ++kk) s_qt[o + kk * NQ2] = ldnt(&qe[o + kk * NQ2]);
  const double* g = ggeo + e * (int64_t)(kNggeo * NP);
  const double *g00 = g + kG00 * NP, *g01 = g + kG01 * NP, *g02 = g + kG02 * NP;
  const double *g11 = g + kG11 * NP, *g12 = g + kG12 * NP, *g22 = g + kG22 * NP;
  // the first slab's geometric factors are in flight across the barrier and
  // its derivatives (few registers are live here; later slabs load theirs
  // after the derivatives, inside the 64-register budget)
  double G0[6];
  G0[0] = ldnt(&g00[o]);
  G0[1] = ldnt(&g01[o]);
  G0[2] = ldnt(&g02[o]);
  G0[3] = ldnt(&g11[o]);
  G0[4] = ldnt(&g12[o]);
  G0[5] = ldnt(&g22[o]);
  __syncthreads();
  double Gt[H], ars[H];
#pragma unroll
  for (int kk = 0; kk < H; ++kk) {
    const int k = H * w + kk;
    double qr = 0.0, qs = 0.0, qt = 0.0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      qr = __builtin_fma(s_D[i * NQ + m], s_qt[k * NQ2 + j * NQ + m], qr);
      qs = __builtin_fma(s_D[j * NQ + m], s_qt[k * NQ2 + m * NQ + i], qs);
      qt = __builtin_fma(D[k * NQ + m], s_qt[m * NQ2 + t], qt);
      if (m % 2 == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the reads the scheduler batches
    }
    const int n = o + kk * NQ2;
    const double G00 = kk == 0 ? G0[0] : ldnt(&g00[n]), G01 = kk == 0 ? G0[1] : ldnt(&g01[n]);
    const double G02 = kk == 0 ? G0[2] : ldnt(&g02[n]), G11 = kk == 0 ? G0[3] : ldnt(&g11[n]);
    const double G12 = kk == 0 ? G0[4] : ldnt(&g12[n]), G22 = kk == 0 ? G0[5] : ldnt(&g22[n]);
    double* sr = s_Gr[w][kk & 1];
    double* ss = s_Gs[w][kk & 1];
    sr[t] = __builtin_fma(G02, qt, __builtin_fma(G01, qs, G00 * qr));
    ss[t] = __builtin_fma(G12, qt, __builtin_fma(G11, qs, G01 * qr));
    Gt[kk] = __builtin_fma(G22, qt, __builtin_fma(G12, qs, G02 * qr));
    __syncthreads();
    double ar = 0.0, as = 0.0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      ar = __builtin_fma(s_D[m * NQ + i], sr[j * NQ + m], ar);
      as = __builtin_fma(s_D[m * NQ + j], ss[m * NQ + i], as);
      if (m % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
    ars[kk] = ar + as;
  }
  // every q read of both waves precedes their last barrier above
#pragma unroll
  for (int kk = 0; kk < H; ++kk) s_qt[(H * w + kk) * NQ2 + t] = Gt[kk];
  __syncthreads();
  double* Ae = Aq + e * NP;
#pragma unroll
  for (int kk = 0; kk < H; ++kk) {
    const int k = H * w + kk;
    double at = 0.0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) at = __builtin_fma(D[m * NQ + k], s_qt[m * NQ2 + t], at);
    stnt(ars[kk] + at, &Ae[o + kk * NQ2]);
  }
}

int axhelm_dispatch_f64(int N, int Nf, int bk, bool partial, int32_t Nelements, int32_t offset,
                        const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                        const void* q, void* Aq, const int* skip, hipStream_t s)
{
  return dispatch_n<double>(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip, s);
}

int axhelm_pcg_dispatch_f64(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const void* ggeo,
                            const void* D, const void* lambda, const double* z, double* p, double* Ap,
                            double* partials, const PcgFlags* S, hipStream_t s)
{
  return axhelm_pcg_launch_t<double, double>(N, Nf, bk, Nelements, offset, ggeo, D, lambda, z, p, Ap, partials, S, s);
}

int axhelm_launch(int N, int Nf, int bk, nbx_prec prec, int32_t Nelements, int32_t offset,
                  const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                  const void* q, void* Aq, const int* skip, hipStream_t s)
{
  const bool partial = elementList != nullptr;
  switch (prec) {
    case NBX_F64:
      return axhelm_dispatch_f64(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip,
                                 s);
    case NBX_F32:
      return axhelm_dispatch_f32(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq, skip,
                                 s);
    case NBX_BF16:  // FP32 vectors and arithmetic, bf16 geometric factors and lambda
      return axhelm_dispatch_bf16(N, Nf, bk, partial, Nelements, offset, elementList, ggeo, D, lambda, q, Aq,
                                  skip, s);
    case NBX_BF16_MFMA:  // as NBX_BF16 with the contractions on the matrix cores
      if (partial) return fail(NBX_ERR_UNSUPPORTED, "axhelm bf16-MFMA: no element list");
      return axhelm_mma_launch(N, Nf, bk, true, false, Nelements, offset, ggeo, D, lambda, q, Aq, skip, s);
    case NBX_F16:  // fp16 matrix-core operands, FP32 geometry, vectors and accumulation
      if (partial) return fail(NBX_ERR_UNSUPPORTED, "axhelm fp16-MFMA: no element list");
      return axhelm_mma_launch(N, Nf, bk, false, true, Nelements, offset, ggeo, D, lambda, q, Aq, skip, s);
    default:
      return fail(NBX_ERR_UNSUPPORTED, "axhelm: precision %d not supported by nbx_axhelm", (int)prec);
  }
}

int axhelm_pcg_launch(int N, int Nf, int bk, nbx_prec prec, int32_t Nelements, int32_t offset, const void* ggeo,
                      const void* D, const void* lambda, const double* z, double* p, double* Ap,
                      double* partials, const PcgFlags* S, hipStream_t s)
{
  switch (prec) {
    case NBX_F64:
      return axhelm_pcg_dispatch_f64(N, Nf, bk, Nelements, offset, ggeo, D, lambda, z, p, Ap, partials,
                                                 S, s);
    case NBX_F32:
      return axhelm_pcg_dispatch_f32(N, Nf, bk, Nelements, offset, ggeo, D, lambda, z, p, Ap, partials,
                                               S, s);
    case NBX_BF16:
      return axhelm_pcg_dispatch_bf16(N, Nf, bk, Nelements, offset, ggeo, D, lambda, z, p, Ap,
                                                  partials, S, s);
    default:
      return fail(NBX_ERR_UNSUPPORTED, "axhelm_pcg: precision %d", (int)prec);
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

// Internal helpers shared by the nbx translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/nbx.h"

namespace nbx {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// true when p is device (or managed) memory the GPU may dereference; host
// memory that was never registered with HIP is refused before any launch
inline bool device_accessible(const void* p)
{
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky "invalid value" of an unknown pointer
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

constexpr int kGWJ = 0, kG00 = 1, kG01 = 2, kG02 = 3, kG11 = 4, kG12 = 5, kG22 = 6, kNggeo = 7;

// reduction tile: 256 lanes x 8 values (see nbx_blas.hip)
constexpr int kRedLanes = 256;
constexpr int kRedPerLane = 8;
constexpr int kRedTile = kRedLanes * kRedPerLane;

// streaming (single-use) loads and stores: nontemporal, so they do not evict
// reusable lines; measured ~13% more HBM bandwidth on read-heavy streams
template <typename T>
__device__ __forceinline__ T ldnt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void stnt(T v, T* p) { __builtin_nontemporal_store(v, p); }

// float -> bf16 bits (the upper half of the float) by round to nearest, ties
// to even: adding 0x7FFF (+1 when the kept part is odd) carries into bit 16
// exactly when the dropped half lies above the midpoint, or on it with an odd
// kept part; past the largest finite bf16 this carries into infinity.  A NaN
// stays a NaN (sign and leading payload kept, quiet bit set).
__host__ __device__ __forceinline__ uint16_t bf16_bits(float x)
{
  const uint32_t u = __builtin_bit_cast(uint32_t, x);
  const uint32_t rounded = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
  const uint32_t nan = (u >> 16) | 0x40u;
  // a select, not a branch: the conversions sit in matrix-core inner loops
  return (uint16_t)(((u & 0x7FFFFFFFu) > 0x7F800000u) ? nan : rounded);
}
// the same rounding, as a float value
__host__ __device__ __forceinline__ float bf16_round(float x)
{
  return __builtin_bit_cast(float, (uint32_t)bf16_bits(x) << 16);
}
// The hardware conversion (gfx950 v_cvt_pk_bf16_f32, one instruction for two
// values): for every input that is not a NaN it gives the bits of bf16_bits,
// checked over all 2^32 float patterns on the device; a NaN stays a NaN.
typedef __bf16 nbx_bf16x2 __attribute__((ext_vector_type(2)));
typedef float nbx_f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t bf16_bits_hw(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
// bf16 bits of (lo, hi) packed as lo | hi << 16
__device__ __forceinline__ uint32_t bf16x2_bits_hw(float lo, float hi)
{
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((nbx_f32x2{lo, hi}), nbx_bf16x2));
}

// Device scalar block of the PCG driver.  Kernels outside the driver only
// read `done` (skip all work once the solve has ended) and `beta`.
struct PcgFlags {
  int done;
  int pad_;
  double beta;
};

// launchers shared between translation units; `skip` (device int, may be
// null) turns the launch into a no-op when *skip != 0
// matrix-core axhelm (nbx_axhelm_mma.hip): N = 7, one field, FP32 vectors;
// bf16 operands with bf16 or FP32 geometry, or fp16 operands with FP32 geometry
int axhelm_mma_launch(int N, int Nf, int bk, bool bf16_geometry, bool f16_operands, int32_t Nelements,
                      int32_t offset, const void* ggeo, const void* D, const void* lambda, const void* q, void* Aq,
                      const int* skip, hipStream_t s);
int axhelm_launch(int N, int Nf, int bk, nbx_prec prec, int32_t Nelements, int32_t offset,
                  const int32_t* elementList, const void* ggeo, const void* D, const void* lambda,
                  const void* q, void* Aq, const int* skip, hipStream_t s);
int gs_launch(const nbx_ogs* o, void* v, int k, int32_t stride, nbx_type type, nbx_op op, const int* skip,
              hipStream_t s);
int mask_launch(int32_t Nmasked, const int32_t* maskIds, double* q, const int* skip, hipStream_t s);

// device mesh producers (nbx_mesh_dev.hip); equal to nbx_box_mesh bit for bit
int box_geometry_launch(const nbx_box& b, double* ggeo, double* x, double* y, double* z, hipStream_t s);
int box_ids_launch(const nbx_box& b, int mask_boundary, int64_t* ids, hipStream_t s);
int urand_fill_launch(uint64_t seed, uint64_t stream, int64_t n, double* out, hipStream_t s);
int mask_list_from_ids(int32_t N, const int64_t* ids, int32_t** maskIds, int32_t* Nmasked, hipStream_t s);
// p = z + beta p (FP64); Ap = A p (local) with the operator in precision prec
// (F64: double ggeo/D/lambda; F32: float; BF16: bf16 ggeo/lambda, float D and
// float arithmetic); partials[e] = element share of pAp (FP64)
int axhelm_pcg_launch(int N, int Nf, int bk, nbx_prec prec, int32_t Nelements, int32_t offset, const void* ggeo,
                      const void* D, const void* lambda, const double* z, double* p, double* Ap,
                      double* partials, const PcgFlags* S, hipStream_t s);
// double add gather-scatter of every field plus the Dirichlet mask, one launch
int gs_add_mask_launch(const nbx_ogs* o, double* v, int k, int32_t stride, int32_t Nmasked,
                       const int32_t* maskIds, const int* skip, hipStream_t s);
int gs_add_mask_launch(const nbx_ogs* o, float* v, int k, int32_t stride, int32_t Nmasked,
                       const int32_t* maskIds, const int* skip, hipStream_t s);
// group size per local node as bytes (null when a group exceeds 255); the
// weight 1.0/(double)deg equals invDegree bit for bit
const uint8_t* ogs_degree(const nbx_ogs* o);

// dot weight: invDegree from its 1-byte degree (exact: same IEEE division as
// the setup) or from the FP64 array
template <bool DEG>
__device__ __forceinline__ double weight(const double* __restrict__ w, const uint8_t* __restrict__ deg, int64_t id)
{
  if (DEG) return 1.0 / (double)deg[id];
  return ldnt(&w[id]);
}

}  // namespace nbx

#define NBX_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t nbx_e_ = (call);                                                              \
    if (nbx_e_ != hipSuccess)                                                                \
      return nbx::fail(NBX_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,               \
                       hipGetErrorString(nbx_e_));                                           \
  } while (0)

#define NBX_HIP_LAUNCH()                                                                     \
  do {                                                                                       \
    hipError_t nbx_e_ = hipGetLastError();                                                   \
    if (nbx_e_ != hipSuccess)                                                                \
      return nbx::fail(NBX_ERR_HIP, "%s:%d launch: %s", __FILE__, __LINE__,                 \
                       hipGetErrorString(nbx_e_));                                           \
  } while (0)

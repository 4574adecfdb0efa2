// ogs: gather-scatter assembly of E-vectors on one GPU.
//
// Setup restates the local part of ogsSetup (3rdParty/ogs/src/ogsSetup.cpp:137-212):
// nodes sorted by (|id|, local index), id 0 excluded; a group is a run of equal
// signed id in that order (so all-positive ids group by value); groups are
// numbered in order of their smallest local index; members listed ascending.
// The device path keeps only groups with two or more members (the reference
// kernel skips singletons, gatherScatterMany.okl:68), which for a hex box
// removes ~40-60% of the work without changing any result.
//
// Kernel: one lane per shared group, all k fields per lane (ids read once).
// Add/mul start from 0/1 and combine members in ascending order, min/max
// start from the first member (gatherScatterMany.okl:58-86, 148-176, 300-330,
// 455-485), so every result is bit-identical to the reference kernel's.
#include "nbx_internal.h"
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

struct nbx_ogs {
  int32_t N = 0;
  int32_t Nlocal = 0;
  int32_t NlocalGather = 0;
  int32_t NsharedGather = 0;
  int32_t Nshared = 0;
  std::vector<int32_t> h_offsets;  // full CSR (all groups)
  std::vector<int32_t> h_ids;
  int32_t* d_offsets = nullptr;    // shared groups of other sizes (CSR)
  int32_t* d_ids = nullptr;
  int32_t Ngen = 0;                // groups in the CSR above
  int32_t Nfix[3] = {0, 0, 0};     // groups of exactly 2, 4, 8 members
  int32_t* d_fix[3] = {nullptr, nullptr, nullptr};  // their members, DEG per group
  double* d_invDegree = nullptr;
  uint8_t* d_degree = nullptr;  // group size per node (1 if excluded); null if any group > 255
};

namespace nbx {

template <typename T, int OP>
struct GsOp;
template <typename T>
struct GsOp<T, NBX_ADD> {
  static constexpr bool first = false;
  __device__ static T init() { return T(0); }
  __device__ static T comb(T acc, T b) { return acc + b; }
};
template <typename T>
struct GsOp<T, NBX_MUL> {
  static constexpr bool first = false;
  __device__ static T init() { return T(1); }
  __device__ static T comb(T acc, T b) { return acc * b; }
};
template <typename T>
struct GsOp<T, NBX_MIN> {
  static constexpr bool first = true;
  __device__ static T init() { return T(0); }
  __device__ static T comb(T acc, T b) { return (b < acc) ? b : acc; }
};
template <typename T>
struct GsOp<T, NBX_MAX> {
  static constexpr bool first = true;
  __device__ static T init() { return T(0); }
  __device__ static T comb(T acc, T b) { return (b > acc) ? b : acc; }
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumbering so that XCD x runs the x-th contiguous eighth of the
// groups keeps neighbouring groups (which share cache lines of the same
// element faces) in one L2.  The grid is a multiple of 8 blocks.
__device__ __forceinline__ int32_t xcd_block(int32_t nblocks8)
{
  const int32_t per = nblocks8 >> 3;
  return (int32_t)(blockIdx.x & 7) * per + (int32_t)(blockIdx.x >> 3);
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_gs_many(const int32_t Ngroups, const int k, const int32_t stride,
                                                 const int32_t* __restrict__ offsets,
                                                 const int32_t* __restrict__ ids, T* __restrict__ q,
                                                 const int* __restrict__ skip)
{
  using O = GsOp<T, OP>;
  const int32_t g = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (g >= Ngroups || (skip && *skip)) return;
  const int32_t s = offsets[g], e = offsets[g + 1];
  for (int f = 0; f < k; ++f) {
    T* qf = q + (int64_t)f * stride;
    T acc = O::first ? qf[ids[s]] : O::init();
    for (int32_t n = s + (O::first ? 1 : 0); n < e; ++n) acc = O::comb(acc, qf[ids[n]]);
    for (int32_t n = s; n < e; ++n) qf[ids[n]] = acc;
  }
}

// groups of exactly DEG members, members stored contiguously (vector loads)
template <typename T, int OP, int DEG>
__global__ __launch_bounds__(256) void k_gs_fixed(const int32_t Ngroups, const int k, const int32_t stride,
                                                  const int32_t* __restrict__ ids, T* __restrict__ q,
                                                  const int* __restrict__ skip)
{
  using O = GsOp<T, OP>;
  const int32_t g = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (g >= Ngroups || (skip && *skip)) return;
  int32_t m[DEG];
  if (DEG == 2) {
    const int2 v = *reinterpret_cast<const int2*>(ids + (int64_t)g * 2);
    m[0] = v.x;
    m[1] = v.y;
  } else {
#pragma unroll
    for (int c = 0; c < DEG / 4; ++c) {
      const int4 v = *reinterpret_cast<const int4*>(ids + (int64_t)g * DEG + 4 * c);
      m[4 * c + 0] = v.x;
      m[4 * c + 1] = v.y;
      m[4 * c + 2] = v.z;
      m[4 * c + 3] = v.w;
    }
  }
  for (int f = 0; f < k; ++f) {
    T* qf = q + (int64_t)f * stride;
    T acc = O::first ? qf[m[0]] : O::init();
#pragma unroll
    for (int n = O::first ? 1 : 0; n < DEG; ++n) acc = O::comb(acc, qf[m[n]]);
#pragma unroll
    for (int n = 0; n < DEG; ++n) qf[m[n]] = acc;
  }
}

// All classes of a double-add gather-scatter and the Dirichlet mask in one
// launch.  Every class list (pairs, quads, octets, CSR groups, masked nodes)
// is cut into 8 equal chunks, chunk x of every class runs on XCD x (blocks
// b with b % 8 == x), so one L2 sees the same element range for every class.
// Each chunk is cut again into S equal slices, and an XCD runs slice s of
// every class before slice s + 1 of any: the class lists are ordered by
// their first member, so equal fractions cover about the same elements, and
// the lines a slice's pairs touch are still in L2 when its edges and corners
// touch them.  Masked nodes belong to no group: mask writes never touch a
// gathered value.  The order of groups does not change any group's sum.
struct GsClasses {
  int32_t n[5];          // pairs, quads, octets, CSR groups, masked nodes
  int32_t per[5];        // items per XCD chunk
  int32_t ps[5];         // items per slice of a chunk
  int32_t bstart[6];     // cumulative blocks of one slice, by class
  int32_t nslice;        // slices per chunk
  const int32_t* ids[5]; // members (pairs/quads/octets), CSR ids, mask ids
  const int32_t* off;    // CSR offsets
};

// one group of DEG members (ids[item*DEG ..]): the sum in ascending member
// order from zero, written back to every member.  DEG is a compile-time
// constant so that a lane issues the loads of all members before the first
// use (a run-time count compiled to one load per round trip to memory).
template <typename T, int DEG>
__device__ __forceinline__ void gs_group_sum(const int32_t* __restrict__ ids, const int32_t item, const int k,
                                             const int32_t stride, T* __restrict__ q)
{
  int32_t m[DEG];
  if (DEG == 2) {
    const int2 v = *reinterpret_cast<const int2*>(ids + (int64_t)item * 2);
    m[0] = v.x;
    m[1] = v.y;
  } else {
#pragma unroll
    for (int h = 0; h < DEG / 4; ++h) {
      const int4 v = *reinterpret_cast<const int4*>(ids + (int64_t)item * DEG + 4 * h);
      m[4 * h + 0] = v.x;
      m[4 * h + 1] = v.y;
      m[4 * h + 2] = v.z;
      m[4 * h + 3] = v.w;
    }
  }
  for (int f = 0; f < k; ++f) {
    T* qf = q + (int64_t)f * stride;
    T v[DEG];
#pragma unroll
    for (int n = 0; n < DEG; ++n) v[n] = qf[m[n]];
    T acc = T(0);
#pragma unroll
    for (int n = 0; n < DEG; ++n) acc = acc + v[n];
#pragma unroll
    for (int n = 0; n < DEG; ++n) qf[m[n]] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gs_add_mask(const GsClasses C, const int k, const int32_t stride,
                                                     T* __restrict__ q, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int32_t xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  const int32_t slice = local / C.bstart[5], lb = local - slice * C.bstart[5];
  int c = 0;
  while (c < 5 && lb >= C.bstart[c + 1]) ++c;
  if (c == 5 || slice >= C.nslice) return;
  const int32_t s0 = slice * C.ps[c];
  const int32_t item = xcd * C.per[c] + s0 + (lb - C.bstart[c]) * 256 + (int32_t)threadIdx.x;
  if (item >= min(C.n[c], xcd * C.per[c] + min(C.per[c], s0 + C.ps[c]))) return;
  if (c == 4) {
    const int32_t id = C.ids[4][item];
    for (int f = 0; f < k; ++f) q[id + (int64_t)f * stride] = T(0);
    return;
  }
  if (c == 3) {
    const int32_t s0 = C.off[item], e0 = C.off[item + 1];
    const int32_t* idg = C.ids[3];
    for (int f = 0; f < k; ++f) {
      T* qf = q + (int64_t)f * stride;
      T acc = T(0);
      for (int32_t n = s0; n < e0; ++n) acc = acc + qf[idg[n]];
      for (int32_t n = s0; n < e0; ++n) qf[idg[n]] = acc;
    }
    return;
  }
  if (c == 0) gs_group_sum<T, 2>(C.ids[0], item, k, stride, q);
  else if (c == 1) gs_group_sum<T, 4>(C.ids[1], item, k, stride, q);
  else gs_group_sum<T, 8>(C.ids[2], item, k, stride, q);
}

__global__ __launch_bounds__(256) void k_mask(const int32_t Nmasked, const int32_t* __restrict__ maskIds,
                                              double* __restrict__ q, const int* __restrict__ skip)
{
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (skip && *skip) return;
  if (n < Nmasked) q[maskIds[n]] = 0.0;
}

inline dim3 grid8(int32_t n) { return dim3((unsigned)(((n + 255) / 256 + 7) / 8 * 8)); }

template <typename T, int OP>
static void gs_launch_op(const nbx_ogs* o, T* v, int k, int32_t stride, const int* skip, hipStream_t s)
{
  if (o->Nfix[0])
    hipLaunchKernelGGL((k_gs_fixed<T, OP, 2>), grid8(o->Nfix[0]), dim3(256), 0, s, o->Nfix[0], k, stride, o->d_fix[0],
                       v, skip);
  if (o->Nfix[1])
    hipLaunchKernelGGL((k_gs_fixed<T, OP, 4>), grid8(o->Nfix[1]), dim3(256), 0, s, o->Nfix[1], k, stride, o->d_fix[1],
                       v, skip);
  if (o->Nfix[2])
    hipLaunchKernelGGL((k_gs_fixed<T, OP, 8>), grid8(o->Nfix[2]), dim3(256), 0, s, o->Nfix[2], k, stride, o->d_fix[2],
                       v, skip);
  if (o->Ngen)
    hipLaunchKernelGGL((k_gs_many<T, OP>), grid8(o->Ngen), dim3(256), 0, s, o->Ngen, k, stride, o->d_offsets, o->d_ids,
                       v, skip);
}

template <typename T>
static int gs_launch_t(const nbx_ogs* o, void* v, int k, int32_t stride, nbx_op op, const int* skip,
                       hipStream_t s)
{
  switch (op) {
    case NBX_ADD: gs_launch_op<T, NBX_ADD>(o, (T*)v, k, stride, skip, s); break;
    case NBX_MUL: gs_launch_op<T, NBX_MUL>(o, (T*)v, k, stride, skip, s); break;
    case NBX_MIN: gs_launch_op<T, NBX_MIN>(o, (T*)v, k, stride, skip, s); break;
    case NBX_MAX: gs_launch_op<T, NBX_MAX>(o, (T*)v, k, stride, skip, s); break;
    default: return fail(NBX_ERR_INVALID, "gs: unknown op %d", (int)op);
  }
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int gs_launch(const nbx_ogs* o, void* v, int k, int32_t stride, nbx_type type, nbx_op op, const int* skip,
              hipStream_t s)
{
  if (o->NsharedGather == 0) return NBX_SUCCESS;
  switch (type) {
    case NBX_DOUBLE: return gs_launch_t<double>(o, v, k, stride, op, skip, s);
    case NBX_FLOAT: return gs_launch_t<float>(o, v, k, stride, op, skip, s);
    case NBX_INT: return gs_launch_t<int32_t>(o, v, k, stride, op, skip, s);
    case NBX_LONG: return gs_launch_t<long long>(o, v, k, stride, op, skip, s);
    default: return fail(NBX_ERR_INVALID, "gs: unknown type %d", (int)type);
  }
}

template <typename T>
static int gs_add_mask_launch_t(const nbx_ogs* o, T* v, int k, int32_t stride, int32_t Nmasked,
                                const int32_t* maskIds, const int* skip, hipStream_t s)
{
  GsClasses C{};
  const int32_t n[5] = {o->Nfix[0], o->Nfix[1], o->Nfix[2], o->Ngen, Nmasked};
  const int32_t* ids[5] = {o->d_fix[0], o->d_fix[1], o->d_fix[2], o->d_ids, maskIds};
  int32_t items = 0;
  for (int c = 0; c < 5; ++c) {
    C.n[c] = n[c];
    C.per[c] = (n[c] + 7) / 8;
    C.ids[c] = ids[c];
    items += C.per[c];
  }
  if (items == 0) return NBX_SUCCESS;
  // slices of about kGsSliceItems items of an XCD's chunk: with pairs of
  // one-field N = 7 elements about 128 elements, whose 512 KB of values the
  // XCD's 4 MB L2 holds while the slice's blocks run
  constexpr int32_t kGsSliceItems = 16384;
  C.nslice = std::max<int32_t>(1, items / kGsSliceItems);
  C.bstart[0] = 0;
  for (int c = 0; c < 5; ++c) {
    C.ps[c] = (C.per[c] + C.nslice - 1) / C.nslice;
    C.bstart[c + 1] = C.bstart[c] + (C.ps[c] + 255) / 256;
  }
  const int64_t blocks = 8 * (int64_t)C.nslice * C.bstart[5];
  hipLaunchKernelGGL(k_gs_add_mask<T>, dim3((unsigned)blocks), dim3(256), 0, s, C, k, stride, v, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int gs_add_mask_launch(const nbx_ogs* o, double* v, int k, int32_t stride, int32_t Nmasked,
                       const int32_t* maskIds, const int* skip, hipStream_t s)
{
  return gs_add_mask_launch_t<double>(o, v, k, stride, Nmasked, maskIds, skip, s);
}

int gs_add_mask_launch(const nbx_ogs* o, float* v, int k, int32_t stride, int32_t Nmasked,
                       const int32_t* maskIds, const int* skip, hipStream_t s)
{
  return gs_add_mask_launch_t<float>(o, v, k, stride, Nmasked, maskIds, skip, s);
}

const uint8_t* ogs_degree(const nbx_ogs* o) { return o ? o->d_degree : nullptr; }

int mask_launch(int32_t Nmasked, const int32_t* maskIds, double* q, const int* skip, hipStream_t s)
{
  if (Nmasked <= 0) return NBX_SUCCESS;
  hipLaunchKernelGGL(k_mask, dim3((Nmasked + 255) / 256), dim3(256), 0, s, Nmasked, maskIds, q, skip);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

}  // namespace nbx

using namespace nbx;

static void ogs_release(nbx_ogs* o)
{
  if (!o) return;
  if (o->d_offsets) (void)hipFree(o->d_offsets);
  if (o->d_ids) (void)hipFree(o->d_ids);
  for (int c = 0; c < 3; ++c)
    if (o->d_fix[c]) (void)hipFree(o->d_fix[c]);
  if (o->d_invDegree) (void)hipFree(o->d_invDegree);
  if (o->d_degree) (void)hipFree(o->d_degree);
  delete o;
}

namespace nbx {
namespace {

// ---- device construction of the local CSR (ogsSetup.cpp:137-212, 325-352) ----
// 1. key = |id| (id 0 -> a sentinel past every id), value = local index;
//    stable radix sort by key gives the (|id|, local index) order.
// 2. a run starts where the signed id differs from its predecessor; runs are
//    numbered by an inclusive scan of the start flags.
// 3. a group's final number is the rank of its first (smallest) member among
//    all first members: flags over local indices, exclusive scan.
// 4. offsets = exclusive scan of the group sizes in final order; members,
//    1/multiplicity and the one-byte degree are scattered in one pass.
// 5. the groups with >= 2 members are split by size (2, 4, 8, other) in final
//    order with one exclusive scan per class.

constexpr int kOgsBlock = 256;

inline dim3 ogs_grid(int64_t n) { return dim3((unsigned)((n + kOgsBlock - 1) / kOgsBlock)); }

// largest |id| and number of nonzero ids (order-independent totals): a
// fixed grid strides over the ids, each block reduces with wave shuffles and
// LDS, one atomic pair per block (per-wave atomics on the two words took
// milliseconds at 16.7M ids)
constexpr int kOgsRangeBlocks = 1024;

__global__ __launch_bounds__(kOgsBlock) void k_ogs_range(int32_t N, const int64_t* __restrict__ ids,
                                                         unsigned long long* __restrict__ amax, int* __restrict__ count)
{
  __shared__ unsigned long long s_a[kOgsBlock / 64];
  __shared__ int s_c[kOgsBlock / 64];
  unsigned long long a = 0;
  int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kOgsBlock) {
    const int64_t v = ids[i];
    const unsigned long long u = (unsigned long long)(v < 0 ? -v : v);
    a = u > a ? u : a;
    c += v != 0 ? 1 : 0;
  }
  for (int h = 32; h >= 1; h >>= 1) {
    const unsigned long long ao = __shfl_xor(a, h, 64);
    a = ao > a ? ao : a;
    c += __shfl_xor(c, h, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_a[threadIdx.x >> 6] = a;
    s_c[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kOgsBlock / 64; ++w) {
      a = s_a[w] > a ? s_a[w] : a;
      c += s_c[w];
    }
    if (c > 0) {
      atomicMax(amax, a);
      atomicAdd(count, c);
    }
  }
}

__global__ void k_ogs_keys(int32_t N, const int64_t* __restrict__ ids, uint64_t sentinel, uint64_t* __restrict__ keys,
                           int32_t* __restrict__ vals)
{
  const int64_t i = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (i >= N) return;
  const int64_t v = ids[i];
  keys[i] = v == 0 ? sentinel : (uint64_t)(v < 0 ? -v : v);
  vals[i] = (int32_t)i;
}

__global__ void k_ogs_starts(int32_t Nlocal, const int64_t* __restrict__ ids, const int32_t* __restrict__ loc,
                             int32_t* __restrict__ start)
{
  const int64_t n = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (n >= Nlocal) return;
  start[n] = (n == 0 || ids[loc[n]] != ids[loc[n - 1]]) ? 1 : 0;
}

__global__ void k_ogs_runs(int32_t Nlocal, const int32_t* __restrict__ start, const int32_t* __restrict__ incl,
                           const int32_t* __restrict__ loc, int32_t* __restrict__ runStart,
                           int32_t* __restrict__ owner)
{
  const int64_t n = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (n >= Nlocal || !start[n]) return;
  runStart[incl[n] - 1] = (int32_t)n;
  owner[loc[n]] = 1;
}

__global__ void k_ogs_final(int32_t G, int32_t Nlocal, const int32_t* __restrict__ runStart,
                            const int32_t* __restrict__ loc, const int32_t* __restrict__ rank,
                            int32_t* __restrict__ finOfRun, int32_t* __restrict__ cntFin)
{
  const int64_t r = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (r >= G) return;
  const int32_t s = runStart[r], e = r + 1 < G ? runStart[r + 1] : Nlocal;
  const int32_t fin = rank[loc[s]];
  finOfRun[r] = fin;
  cntFin[fin] = e - s;
}

__global__ void k_ogs_members(int32_t Nlocal, const int32_t* __restrict__ incl, const int32_t* __restrict__ runStart,
                              const int32_t* __restrict__ finOfRun, const int32_t* __restrict__ offsets,
                              const int32_t* __restrict__ cntFin, const int32_t* __restrict__ loc,
                              int32_t* __restrict__ gids, double* __restrict__ invDeg, uint8_t* __restrict__ deg,
                              int* __restrict__ maxcnt)
{
  const int64_t n = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (n >= Nlocal) return;
  const int32_t r = incl[n] - 1;
  const int32_t m = (int32_t)n - runStart[r];
  const int32_t fin = finOfRun[r];
  const int32_t cnt = cntFin[fin];
  const int32_t l = loc[n];
  gids[offsets[fin] + m] = l;
  invDeg[l] = 1.0 / (double)cnt;
  deg[l] = (uint8_t)(cnt < 255 ? cnt : 255);
  if (m == 0 && cnt > 255) atomicMax(maxcnt, cnt);
}

__global__ void k_ogs_init_weights(int32_t N, double* __restrict__ invDeg, uint8_t* __restrict__ deg)
{
  const int64_t i = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (i >= N) return;
  invDeg[i] = 1.0;
  deg[i] = 1;
}

// class flags in final order: 2, 4, 8 members, other sizes >= 2, and the
// member count of the latter
__global__ void k_ogs_classes(int32_t G, const int32_t* __restrict__ cntFin, int32_t* __restrict__ f2,
                              int32_t* __restrict__ f4, int32_t* __restrict__ f8, int32_t* __restrict__ fg,
                              int32_t* __restrict__ gc)
{
  const int64_t g = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (g >= G) return;
  const int32_t c = cntFin[g];
  const bool gen = c > 1 && c != 2 && c != 4 && c != 8;
  f2[g] = c == 2;
  f4[g] = c == 4;
  f8[g] = c == 8;
  fg[g] = gen;
  gc[g] = gen ? c : 0;
}

__global__ void k_ogs_class_fill(int32_t G, const int32_t* __restrict__ offsets, const int32_t* __restrict__ cntFin,
                                 const int32_t* __restrict__ gids, const int32_t* __restrict__ p2,
                                 const int32_t* __restrict__ p4, const int32_t* __restrict__ p8,
                                 const int32_t* __restrict__ pg, const int32_t* __restrict__ goff,
                                 int32_t* __restrict__ fix2, int32_t* __restrict__ fix4, int32_t* __restrict__ fix8,
                                 int32_t* __restrict__ sOff, int32_t* __restrict__ sIds)
{
  const int64_t g = (int64_t)blockIdx.x * kOgsBlock + threadIdx.x;
  if (g >= G) return;
  const int32_t c = cntFin[g];
  if (c < 2) return;
  const int32_t* src = gids + offsets[g];
  int32_t* dst;
  if (c == 2) {
    dst = fix2 + 2 * (int64_t)p2[g];
  } else if (c == 4) {
    dst = fix4 + 4 * (int64_t)p4[g];
  } else if (c == 8) {
    dst = fix8 + 8 * (int64_t)p8[g];
  } else {
    sOff[pg[g]] = goff[g];
    dst = sIds + goff[g];
  }
  for (int32_t m = 0; m < c; ++m) dst[m] = src[m];
}

// device buffers of one setup, released together
struct OgsTmp {
  std::vector<void*> ptrs;
  template <typename T>
  int get(T** p, size_t n)
  {
    *p = nullptr;
    if (hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
      return fail(NBX_ERR_ALLOC, "ogs_setup: device allocation of %zu bytes failed", sizeof(T) * n);
    ptrs.push_back(*p);
    return NBX_SUCCESS;
  }
  ~OgsTmp()
  {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

template <typename T>
int exclusive_sum(OgsTmp& tmp, const T* in, T* out, int32_t n)
{
  size_t bytes = 0;
  NBX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n));
  void* w;
  if (int rc = tmp.get((char**)&w, bytes)) return rc;
  NBX_HIP(hipcub::DeviceScan::ExclusiveSum(w, bytes, in, out, n));
  return NBX_SUCCESS;
}

template <typename T>
int inclusive_sum(OgsTmp& tmp, const T* in, T* out, int32_t n)
{
  size_t bytes = 0;
  NBX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, n));
  void* w;
  if (int rc = tmp.get((char**)&w, bytes)) return rc;
  NBX_HIP(hipcub::DeviceScan::InclusiveSum(w, bytes, in, out, n));
  return NBX_SUCCESS;
}

int32_t read_i32(const int32_t* d)
{
  int32_t v = 0;
  (void)hipMemcpy(&v, d, sizeof(int32_t), hipMemcpyDeviceToHost);
  return v;
}

// ids: host array, or device array when on_device
int ogs_build(nbx_ogs* o, int32_t N, const nbx_hlong* ids, bool on_device)
{
  // |id| range and participating count
  uint64_t amax = 0;
  int32_t Nlocal = 0;
  if (!on_device) {
    for (int32_t i = 0; i < N; ++i) {
      const int64_t v = ids[i];
      if (v == 0) continue;
      const uint64_t a = (uint64_t)(v < 0 ? -v : v);
      amax = std::max(amax, a);
      ++Nlocal;
    }
  } else if (N > 0) {
    unsigned long long* d_r = nullptr;
    NBX_HIP(hipMalloc(&d_r, 2 * sizeof(unsigned long long)));
    (void)hipMemset(d_r, 0, 2 * sizeof(unsigned long long));
    hipLaunchKernelGGL(k_ogs_range, dim3((unsigned)std::min<int64_t>(ogs_grid(N).x, kOgsRangeBlocks)), dim3(kOgsBlock), 0,
                       nullptr, N, ids, d_r, (int*)(d_r + 1));
    unsigned long long h[2] = {0, 0};
    const hipError_t e1 = hipGetLastError(), e2 = hipMemcpy(h, d_r, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d_r);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(NBX_ERR_HIP, "ogs_setup: id range failed");
    amax = h[0];
    Nlocal = (int32_t)(h[1] & 0xffffffffu);
  }
  o->Nlocal = Nlocal;
  NBX_HIP(hipMalloc(&o->d_invDegree, sizeof(double) * std::max<int32_t>(N, 1)));
  NBX_HIP(hipMalloc(&o->d_degree, (size_t)std::max<int32_t>(N, 1)));
  if (N > 0) {
    hipLaunchKernelGGL(k_ogs_init_weights, ogs_grid(N), dim3(kOgsBlock), 0, nullptr, N, o->d_invDegree, o->d_degree);
    NBX_HIP_LAUNCH();
  }
  o->h_offsets.assign(1, 0);
  o->h_ids.assign(1, 0);
  if (Nlocal == 0) {
    o->NlocalGather = 0;
    NBX_HIP(hipMalloc(&o->d_offsets, sizeof(int32_t)));
    NBX_HIP(hipMemset(o->d_offsets, 0, sizeof(int32_t)));
    NBX_HIP(hipMalloc(&o->d_ids, sizeof(int32_t)));
    return NBX_SUCCESS;
  }
  int nbits = 1;
  while (nbits < 63 && (amax >> nbits) != 0) ++nbits;
  const uint64_t sentinel = 1ull << nbits;  // above every |id|; sorts the excluded nodes last

  OgsTmp tmp;
  const int64_t* d_id = ids;
  uint64_t *k0, *k1;
  int32_t *v0, *loc;
  int rc;
  if ((rc = tmp.get(&k0, N)) || (rc = tmp.get(&k1, N)) || (rc = tmp.get(&v0, N)) || (rc = tmp.get(&loc, N)))
    return rc;
  if (!on_device) {
    int64_t* up;
    if ((rc = tmp.get(&up, N))) return rc;
    NBX_HIP(hipMemcpy(up, ids, sizeof(int64_t) * N, hipMemcpyHostToDevice));
    d_id = up;
  }
  hipLaunchKernelGGL(k_ogs_keys, ogs_grid(N), dim3(kOgsBlock), 0, nullptr, N, d_id, sentinel, k0, v0);
  NBX_HIP_LAUNCH();
  {
    size_t bytes = 0;
    NBX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k0, k1, v0, loc, N, 0, nbits + 1));
    void* w;
    if ((rc = tmp.get((char**)&w, bytes))) return rc;
    NBX_HIP(hipcub::DeviceRadixSort::SortPairs(w, bytes, k0, k1, v0, loc, N, 0, nbits + 1));
  }
  // runs of equal signed id over the first Nlocal sorted nodes
  int32_t *start, *incl, *owner, *rank;
  if ((rc = tmp.get(&start, Nlocal)) || (rc = tmp.get(&incl, Nlocal)) || (rc = tmp.get(&owner, N)) ||
      (rc = tmp.get(&rank, N)))
    return rc;
  hipLaunchKernelGGL(k_ogs_starts, ogs_grid(Nlocal), dim3(kOgsBlock), 0, nullptr, Nlocal, d_id, loc, start);
  NBX_HIP_LAUNCH();
  if ((rc = inclusive_sum(tmp, start, incl, Nlocal))) return rc;
  const int32_t G = read_i32(incl + Nlocal - 1);
  o->NlocalGather = G;
  int32_t *runStart, *finOfRun, *cntFin, *offsets, *gids;
  if ((rc = tmp.get(&runStart, G)) || (rc = tmp.get(&finOfRun, G)) || (rc = tmp.get(&cntFin, (size_t)G + 1)) ||
      (rc = tmp.get(&offsets, (size_t)G + 1)) || (rc = tmp.get(&gids, Nlocal)))
    return rc;
  NBX_HIP(hipMemset(owner, 0, sizeof(int32_t) * N));
  hipLaunchKernelGGL(k_ogs_runs, ogs_grid(Nlocal), dim3(kOgsBlock), 0, nullptr, Nlocal, start, incl, loc, runStart,
                     owner);
  NBX_HIP_LAUNCH();
  if ((rc = exclusive_sum(tmp, owner, rank, N))) return rc;
  NBX_HIP(hipMemset(cntFin + G, 0, sizeof(int32_t)));
  hipLaunchKernelGGL(k_ogs_final, ogs_grid(G), dim3(kOgsBlock), 0, nullptr, G, Nlocal, runStart, loc, rank, finOfRun,
                     cntFin);
  NBX_HIP_LAUNCH();
  if ((rc = exclusive_sum(tmp, cntFin, offsets, G + 1))) return rc;
  int* maxcnt;
  if ((rc = tmp.get(&maxcnt, 1))) return rc;
  NBX_HIP(hipMemset(maxcnt, 0, sizeof(int)));
  hipLaunchKernelGGL(k_ogs_members, ogs_grid(Nlocal), dim3(kOgsBlock), 0, nullptr, Nlocal, incl, runStart, finOfRun,
                     offsets, cntFin, loc, gids, o->d_invDegree, o->d_degree, maxcnt);
  NBX_HIP_LAUNCH();
  // size classes of the shared groups
  int32_t* fl[5];
  int32_t* ps[5];
  for (int c = 0; c < 5; ++c)
    if ((rc = tmp.get(&fl[c], (size_t)G + 1)) || (rc = tmp.get(&ps[c], (size_t)G + 1))) return rc;
  for (int c = 0; c < 5; ++c) NBX_HIP(hipMemset(fl[c] + G, 0, sizeof(int32_t)));
  hipLaunchKernelGGL(k_ogs_classes, ogs_grid(G), dim3(kOgsBlock), 0, nullptr, G, cntFin, fl[0], fl[1], fl[2], fl[3],
                     fl[4]);
  NBX_HIP_LAUNCH();
  for (int c = 0; c < 5; ++c)
    if ((rc = exclusive_sum(tmp, fl[c], ps[c], G + 1))) return rc;
  int32_t tot[5];
  for (int c = 0; c < 5; ++c) tot[c] = read_i32(ps[c] + G);
  for (int c = 0; c < 3; ++c) {
    o->Nfix[c] = tot[c];
    if (tot[c]) NBX_HIP(hipMalloc(&o->d_fix[c], sizeof(int32_t) * (size_t)tot[c] * (2 << c)));
  }
  o->Ngen = tot[3];
  NBX_HIP(hipMalloc(&o->d_offsets, sizeof(int32_t) * ((size_t)tot[3] + 1)));
  NBX_HIP(hipMalloc(&o->d_ids, sizeof(int32_t) * std::max<size_t>((size_t)tot[4], 1)));
  NBX_HIP(hipMemcpy(o->d_offsets + tot[3], &tot[4], sizeof(int32_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_ogs_class_fill, ogs_grid(G), dim3(kOgsBlock), 0, nullptr, G, offsets, cntFin, gids, ps[0],
                     ps[1], ps[2], ps[3], ps[4], o->d_fix[0], o->d_fix[1], o->d_fix[2], o->d_offsets, o->d_ids);
  NBX_HIP_LAUNCH();
  o->NsharedGather = tot[0] + tot[1] + tot[2] + tot[3];
  o->Nshared = 2 * tot[0] + 4 * tot[1] + 8 * tot[2] + tot[4];
  // host copy of the full CSR (nbx_ogs_copy_csr)
  o->h_offsets.resize((size_t)G + 1);
  o->h_ids.resize(Nlocal);
  NBX_HIP(hipMemcpy(o->h_offsets.data(), offsets, sizeof(int32_t) * ((size_t)G + 1), hipMemcpyDeviceToHost));
  NBX_HIP(hipMemcpy(o->h_ids.data(), gids, sizeof(int32_t) * (size_t)Nlocal, hipMemcpyDeviceToHost));
  if (read_i32(maxcnt) > 255) {  // the one-byte degree cannot hold every group size
    NBX_HIP(hipFree(o->d_degree));
    o->d_degree = nullptr;
  }
  NBX_HIP(hipDeviceSynchronize());
  return NBX_SUCCESS;
}

}  // namespace
}  // namespace nbx

static nbx_ogs* ogs_setup_common(nbx_dlong N, const nbx_hlong* ids, int verbose, bool on_device)
{
  if (N < 0 || (N > 0 && !ids)) {
    fail(NBX_ERR_INVALID, "ogs_setup: bad arguments");
    return nullptr;
  }
  nbx_ogs* o = new nbx_ogs();
  o->N = N;
  const int rc = nbx::ogs_build(o, N, ids, on_device);
  if (rc != NBX_SUCCESS) {
    const std::string msg = nbx_last_error();
    ogs_release(o);
    fail(rc, "%s", msg.c_str());
    return nullptr;
  }
  if (verbose)
    fprintf(stderr, "nbx_ogs_setup: N=%d Nlocal=%d groups=%d shared groups=%d shared nodes=%d\n", N, o->Nlocal,
            o->NlocalGather, o->NsharedGather, o->Nshared);
  return o;
}

extern "C" nbx_ogs* nbx_ogs_setup(nbx_dlong N, const nbx_hlong* ids, int verbose)
{
  return ogs_setup_common(N, ids, verbose, false);
}

extern "C" nbx_ogs* nbx_ogs_setup_device(nbx_dlong N, const nbx_hlong* ids, int verbose)
{
  if (N > 0 && ids && !nbx::device_accessible(ids)) {
    fail(NBX_ERR_INVALID, "ogs_setup_device: ids must be device memory");
    return nullptr;
  }
  return ogs_setup_common(N, ids, verbose, true);
}

extern "C" void nbx_ogs_free(nbx_ogs* o) { ogs_release(o); }

extern "C" int nbx_ogs_gather_scatter_many(nbx_ogs* o, void* v, int k, nbx_dlong stride, nbx_type type, nbx_op op,
                                           void* stream)
{
  if (!o) return fail(NBX_ERR_INVALID, "gs: null handle");
  if (k < 1) return fail(NBX_ERR_INVALID, "gs: k=%d", k);
  if (k > 1 && stride < o->N) return fail(NBX_ERR_INVALID, "gs: stride %d < N %d", stride, o->N);
  if (o->NsharedGather == 0) return NBX_SUCCESS;
  if (!v) return fail(NBX_ERR_INVALID, "gs: null vector");
  return gs_launch(o, v, k, stride, type, op, nullptr, as_stream(stream));
}

extern "C" int nbx_ogs_gather_scatter(nbx_ogs* o, void* v, nbx_type type, nbx_op op, void* stream)
{
  return nbx_ogs_gather_scatter_many(o, v, 1, o ? o->N : 0, type, op, stream);
}

extern "C" const double* nbx_ogs_inv_degree(const nbx_ogs* o) { return o ? o->d_invDegree : nullptr; }

extern "C" int nbx_ogs_sizes(const nbx_ogs* o, nbx_dlong* N, nbx_dlong* Nlocal, nbx_dlong* NlocalGather,
                             nbx_dlong* NsharedGather, nbx_dlong* Nshared)
{
  if (!o) return fail(NBX_ERR_INVALID, "ogs_sizes: null handle");
  if (N) *N = o->N;
  if (Nlocal) *Nlocal = o->Nlocal;
  if (NlocalGather) *NlocalGather = o->NlocalGather;
  if (NsharedGather) *NsharedGather = o->NsharedGather;
  if (Nshared) *Nshared = o->Nshared;
  return NBX_SUCCESS;
}

extern "C" int nbx_ogs_copy_csr(const nbx_ogs* o, nbx_dlong* offsets, nbx_dlong* ids)
{
  if (!o) return fail(NBX_ERR_INVALID, "ogs_copy_csr: null handle");
  if (offsets) memcpy(offsets, o->h_offsets.data(), sizeof(int32_t) * ((size_t)o->NlocalGather + 1));
  if (ids && o->Nlocal) memcpy(ids, o->h_ids.data(), sizeof(int32_t) * (size_t)o->Nlocal);
  return NBX_SUCCESS;
}

extern "C" int nbx_mask(nbx_dlong Nmasked, const nbx_dlong* maskIds, double* q, void* stream)
{
  if (Nmasked < 0) return fail(NBX_ERR_INVALID, "mask: Nmasked < 0");
  if (Nmasked > 0 && (!maskIds || !q)) return fail(NBX_ERR_INVALID, "mask: null array");
  return mask_launch(Nmasked, maskIds, q, nullptr, as_stream(stream));
}

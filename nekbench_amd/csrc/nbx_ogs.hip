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
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
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
// Masked nodes belong to no group: mask writes never touch a gathered value.
struct GsClasses {
  int32_t n[5];          // pairs, quads, octets, CSR groups, masked nodes
  int32_t per[5];        // items per XCD chunk
  int32_t bstart[6];     // cumulative blocks per XCD
  const int32_t* ids[5]; // members (pairs/quads/octets), CSR ids, mask ids
  const int32_t* off;    // CSR offsets
};

template <typename T>
__global__ __launch_bounds__(256) void k_gs_add_mask(const GsClasses C, const int k, const int32_t stride,
                                                     T* __restrict__ q, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int32_t xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
  int c = 0;
  while (c < 5 && local >= C.bstart[c + 1]) ++c;
  if (c == 5) return;
  const int32_t item = xcd * C.per[c] + (local - C.bstart[c]) * 256 + (int32_t)threadIdx.x;
  if (item >= min(C.n[c], (xcd + 1) * C.per[c])) return;
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
  int32_t m[8];
  const int deg = 2 << c;
  if (c == 0) {
    const int2 v = *reinterpret_cast<const int2*>(C.ids[0] + (int64_t)item * 2);
    m[0] = v.x;
    m[1] = v.y;
  } else {
    for (int h = 0; h < deg / 4; ++h) {
      const int4 v = *reinterpret_cast<const int4*>(C.ids[c] + (int64_t)item * deg + 4 * h);
      m[4 * h + 0] = v.x;
      m[4 * h + 1] = v.y;
      m[4 * h + 2] = v.z;
      m[4 * h + 3] = v.w;
    }
  }
  for (int f = 0; f < k; ++f) {
    T* qf = q + (int64_t)f * stride;
    T acc = T(0);
    for (int n = 0; n < deg; ++n) acc = acc + qf[m[n]];
    for (int n = 0; n < deg; ++n) qf[m[n]] = acc;
  }
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
  C.bstart[0] = 0;
  for (int c = 0; c < 5; ++c) {
    C.n[c] = n[c];
    C.per[c] = (n[c] + 7) / 8;
    C.ids[c] = ids[c];
    C.bstart[c + 1] = C.bstart[c] + (C.per[c] + 255) / 256;
  }
  C.off = o->d_offsets;
  if (C.bstart[5] == 0) return NBX_SUCCESS;
  hipLaunchKernelGGL(k_gs_add_mask<T>, dim3((unsigned)(8 * C.bstart[5])), dim3(256), 0, s, C, k, stride, v, skip);
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

extern "C" nbx_ogs* nbx_ogs_setup(nbx_dlong N, const nbx_hlong* ids, int verbose)
{
  if (N < 0 || (N > 0 && !ids)) {
    fail(NBX_ERR_INVALID, "ogs_setup: bad arguments");
    return nullptr;
  }
  nbx_ogs* o = new nbx_ogs();
  o->N = N;
  // (|id|, local index) for every participating node, sorted (ogsSetup.cpp:42-54,156);
  // a group starts wherever the signed id differs from its predecessor
  // (ogsSetup.cpp:163), so equal |id| with mixed signs splits into runs
  struct Node {
    uint64_t a;
    int32_t local;
    int64_t id;
  };
  std::vector<Node> nodes;
  nodes.reserve(N);
  for (int32_t i = 0; i < N; ++i)
    if (ids[i] != 0) nodes.push_back(Node{(uint64_t)std::llabs(ids[i]), i, ids[i]});
  std::sort(nodes.begin(), nodes.end(),
            [](const Node& x, const Node& y) { return x.a < y.a || (x.a == y.a && x.local < y.local); });
  const int32_t Nlocal = (int32_t)nodes.size();
  o->Nlocal = Nlocal;
  // a group's final number is the rank of its smallest member among all
  // groups' smallest members (ogsSetup.cpp:177-188)
  std::vector<int32_t> runStart;
  for (int32_t n = 0; n < Nlocal; ++n)
    if (n == 0 || nodes[n].id != nodes[n - 1].id) runStart.push_back(n);
  const int32_t G = (int32_t)runStart.size();
  runStart.push_back(Nlocal);
  o->NlocalGather = G;
  std::vector<int32_t> rankOf(N > 0 ? N : 1, -1);
  {
    std::vector<char> owner(N > 0 ? N : 1, 0);
    for (int32_t g = 0; g < G; ++g) owner[nodes[runStart[g]].local] = 1;
    int32_t c = 0;
    for (int32_t i = 0; i < N; ++i)
      if (owner[i]) rankOf[i] = c++;
  }
  o->h_offsets.assign((size_t)G + 1, 0);
  std::vector<int32_t> order(G);  // run index of final group number
  for (int32_t g = 0; g < G; ++g) {
    const int32_t fin = rankOf[nodes[runStart[g]].local];
    order[fin] = g;
    o->h_offsets[fin + 1] = runStart[g + 1] - runStart[g];
  }
  for (int32_t g = 0; g < G; ++g) o->h_offsets[g + 1] += o->h_offsets[g];
  o->h_ids.resize(Nlocal > 0 ? Nlocal : 1);
  std::vector<double> invDeg(N > 0 ? N : 1, 1.0);
  std::vector<uint8_t> deg(N > 0 ? N : 1, 1);
  bool degFits = true;
  std::vector<int32_t> sOff(1, 0), sIds, fix[3];
  for (int32_t fin = 0; fin < G; ++fin) {
    const int32_t r = order[fin];
    const int32_t cnt = runStart[r + 1] - runStart[r];
    int32_t* dst = o->h_ids.data() + o->h_offsets[fin];
    for (int32_t m = 0; m < cnt; ++m) {
      dst[m] = nodes[runStart[r] + m].local;
      invDeg[dst[m]] = 1.0 / (double)cnt;  // ogsSetup.cpp:342-343
      deg[dst[m]] = (uint8_t)std::min(cnt, 255);
    }
    if (cnt > 255) degFits = false;
    if (cnt > 1) {
      o->NsharedGather++;
      o->Nshared += cnt;
      const int c = cnt == 2 ? 0 : cnt == 4 ? 1 : cnt == 8 ? 2 : -1;
      if (c >= 0) {
        fix[c].insert(fix[c].end(), dst, dst + cnt);
      } else {
        sIds.insert(sIds.end(), dst, dst + cnt);
        sOff.push_back((int32_t)sIds.size());
      }
    }
  }
  o->Ngen = (int32_t)sOff.size() - 1;
  if (verbose)
    fprintf(stderr, "nbx_ogs_setup: N=%d Nlocal=%d groups=%d shared groups=%d shared nodes=%d\n", N, Nlocal, G,
            o->NsharedGather, o->Nshared);
  if (hipMalloc(&o->d_offsets, sizeof(int32_t) * sOff.size()) != hipSuccess ||
      hipMalloc(&o->d_ids, sizeof(int32_t) * std::max<size_t>(sIds.size(), 1)) != hipSuccess ||
      hipMalloc(&o->d_invDegree, sizeof(double) * std::max<int32_t>(N, 1)) != hipSuccess) {
    fail(NBX_ERR_ALLOC, "ogs_setup: device allocation failed");
    ogs_release(o);
    return nullptr;
  }
  if (hipMemcpy(o->d_offsets, sOff.data(), sizeof(int32_t) * sOff.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (sIds.size() &&
       hipMemcpy(o->d_ids, sIds.data(), sizeof(int32_t) * sIds.size(), hipMemcpyHostToDevice) != hipSuccess) ||
      (N > 0 && hipMemcpy(o->d_invDegree, invDeg.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)) {
    fail(NBX_ERR_HIP, "ogs_setup: upload failed");
    ogs_release(o);
    return nullptr;
  }
  for (int c = 0; c < 3; ++c) {
    if (fix[c].empty()) continue;
    o->Nfix[c] = (int32_t)(fix[c].size() / (2 << c));
    if (hipMalloc(&o->d_fix[c], sizeof(int32_t) * fix[c].size()) != hipSuccess ||
        hipMemcpy(o->d_fix[c], fix[c].data(), sizeof(int32_t) * fix[c].size(), hipMemcpyHostToDevice) != hipSuccess) {
      fail(NBX_ERR_HIP, "ogs_setup: class upload failed");
      ogs_release(o);
      return nullptr;
    }
  }
  if (degFits && N > 0) {
    if (hipMalloc(&o->d_degree, (size_t)N) != hipSuccess ||
        hipMemcpy(o->d_degree, deg.data(), (size_t)N, hipMemcpyHostToDevice) != hipSuccess) {
      fail(NBX_ERR_HIP, "ogs_setup: degree upload failed");
      ogs_release(o);
      return nullptr;
    }
  }
  return o;
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

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
  int32_t* d_offsets = nullptr;    // shared groups only
  int32_t* d_ids = nullptr;
  double* d_invDegree = nullptr;
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

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_gs_many(const int32_t Ngroups, const int k, const int32_t stride,
                                                 const int32_t* __restrict__ offsets,
                                                 const int32_t* __restrict__ ids, T* __restrict__ q,
                                                 const int* __restrict__ skip)
{
  using O = GsOp<T, OP>;
  const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= Ngroups || (skip && *skip)) return;
  const int32_t s = offsets[g], e = offsets[g + 1];
  for (int f = 0; f < k; ++f) {
    T* qf = q + (int64_t)f * stride;
    T acc = O::first ? qf[ids[s]] : O::init();
    for (int32_t n = s + (O::first ? 1 : 0); n < e; ++n) acc = O::comb(acc, qf[ids[n]]);
    for (int32_t n = s; n < e; ++n) qf[ids[n]] = acc;
  }
}

__global__ __launch_bounds__(256) void k_mask(const int32_t Nmasked, const int32_t* __restrict__ maskIds,
                                              double* __restrict__ q, const int* __restrict__ skip)
{
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (skip && *skip) return;
  if (n < Nmasked) q[maskIds[n]] = 0.0;
}

template <typename T>
static int gs_launch_t(const nbx_ogs* o, void* v, int k, int32_t stride, nbx_op op, const int* skip,
                       hipStream_t s)
{
  const int32_t G = o->NsharedGather;
  const dim3 grid((G + 255) / 256), block(256);
  switch (op) {
    case NBX_ADD:
      hipLaunchKernelGGL((k_gs_many<T, NBX_ADD>), grid, block, 0, s, G, k, stride, o->d_offsets, o->d_ids, (T*)v, skip);
      break;
    case NBX_MUL:
      hipLaunchKernelGGL((k_gs_many<T, NBX_MUL>), grid, block, 0, s, G, k, stride, o->d_offsets, o->d_ids, (T*)v, skip);
      break;
    case NBX_MIN:
      hipLaunchKernelGGL((k_gs_many<T, NBX_MIN>), grid, block, 0, s, G, k, stride, o->d_offsets, o->d_ids, (T*)v, skip);
      break;
    case NBX_MAX:
      hipLaunchKernelGGL((k_gs_many<T, NBX_MAX>), grid, block, 0, s, G, k, stride, o->d_offsets, o->d_ids, (T*)v, skip);
      break;
    default:
      return fail(NBX_ERR_INVALID, "gs: unknown op %d", (int)op);
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
  if (o->d_invDegree) (void)hipFree(o->d_invDegree);
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
  std::vector<int32_t> sOff(1, 0), sIds;
  for (int32_t fin = 0; fin < G; ++fin) {
    const int32_t r = order[fin];
    const int32_t cnt = runStart[r + 1] - runStart[r];
    int32_t* dst = o->h_ids.data() + o->h_offsets[fin];
    for (int32_t m = 0; m < cnt; ++m) {
      dst[m] = nodes[runStart[r] + m].local;
      invDeg[dst[m]] = 1.0 / (double)cnt;
    }
    if (cnt > 1) {
      sIds.insert(sIds.end(), dst, dst + cnt);
      sOff.push_back((int32_t)sIds.size());
    }
  }
  o->NsharedGather = (int32_t)sOff.size() - 1;
  o->Nshared = (int32_t)sIds.size();
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

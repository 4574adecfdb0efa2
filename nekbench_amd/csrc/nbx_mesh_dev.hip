// Device producers of the box mesh inputs: trilinear nodes and geometric
// factors (core/meshSetup.cpp:2198-2230 and :266-360), global node ids
// (core/meshSetup.cpp:1638-1735), the seeded uniform fill and the Dirichlet
// mask list.  The arithmetic is the host producers' (nbx_mesh.cpp) term for
// term, compiled without contraction, so the device arrays equal the host
// ones bit for bit.  Vertex coordinates come from the host (box_vertices):
// (NX+1)(NY+1)(NZ+1) points, a few MB even for the largest meshes.
//
// Global ids in closed form.  The host labels a non-vertex node with
// 1 + Nnodes + (smallest local index of its GLL point).  Local indices are
// e * Np + n with e = ex + NX (ey + NY ez) increasing in each element
// coordinate, so the smallest copy of global GLL index I (per direction)
// sits in element I/N - 1 at i = N when I is a positive multiple of N, else
// in element I/N at i = I mod N; the three directions combine independently.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "nbx_internal.h"

namespace nbx {

void box_vertices(const nbx_box& b, std::vector<double>& V);
uint64_t urand_key(uint64_t seed, uint64_t stream);

namespace {

__device__ __forceinline__ uint64_t splitmix64_d(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <int NQ>
__global__ __launch_bounds__(NQ* NQ) void k_box_geom(const int NX, const int NY, const double* __restrict__ V,
                                                     const double* __restrict__ zr, const double* __restrict__ wr,
                                                     const double* __restrict__ D, double* __restrict__ ggeo,
                                                     double* __restrict__ xo, double* __restrict__ yo,
                                                     double* __restrict__ zo)
{
  constexpr int NQ2 = NQ * NQ, NP = NQ2 * NQ;
  __shared__ double s_x[NP], s_y[NP], s_z[NP];
  const int64_t e = blockIdx.x;
  const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
  const int t = threadIdx.x, i = t % NQ, j = t / NQ;
  double vx[8], vy[8], vz[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int I = ex + ((v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0);
    const int J = ey + ((v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0);
    const int K = ez + (v >= 4 ? 1 : 0);
    const int64_t vid = I + (int64_t)(NX + 1) * (J + (int64_t)(NY + 1) * K);
    vx[v] = V[3 * vid + 0];
    vy[v] = V[3 * vid + 1];
    vz[v] = V[3 * vid + 2];
  }
  for (int k = 0; k < NQ; ++k) {
    const int n = i + j * NQ + k * NQ2;
    const double rn = zr[i], sn = zr[j], tn = zr[k];
    const double c0 = 0.125 * (1 - rn) * (1 - sn) * (1 - tn), c1 = 0.125 * (1 + rn) * (1 - sn) * (1 - tn);
    const double c2 = 0.125 * (1 + rn) * (1 + sn) * (1 - tn), c3 = 0.125 * (1 - rn) * (1 + sn) * (1 - tn);
    const double c4 = 0.125 * (1 - rn) * (1 - sn) * (1 + tn), c5 = 0.125 * (1 + rn) * (1 - sn) * (1 + tn);
    const double c6 = 0.125 * (1 + rn) * (1 + sn) * (1 + tn), c7 = 0.125 * (1 - rn) * (1 + sn) * (1 + tn);
    s_x[n] = +c0 * vx[0] + c1 * vx[1] + c2 * vx[2] + c3 * vx[3] + c4 * vx[4] + c5 * vx[5] + c6 * vx[6] + c7 * vx[7];
    s_y[n] = +c0 * vy[0] + c1 * vy[1] + c2 * vy[2] + c3 * vy[3] + c4 * vy[4] + c5 * vy[5] + c6 * vy[6] + c7 * vy[7];
    s_z[n] = +c0 * vz[0] + c1 * vz[1] + c2 * vz[2] + c3 * vz[3] + c4 * vz[4] + c5 * vz[5] + c6 * vz[6] + c7 * vz[7];
    if (xo) xo[e * NP + n] = s_x[n];
    if (yo) yo[e * NP + n] = s_y[n];
    if (zo) zo[e * NP + n] = s_z[n];
  }
  if (!ggeo) return;
  __syncthreads();
  double* g = ggeo + e * (int64_t)7 * NP;
  for (int k = 0; k < NQ; ++k) {
    const int n = i + j * NQ + k * NQ2;
    double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr_ = 0, zs = 0, zt = 0;
    for (int m = 0; m < NQ; ++m) {
      const int idr = k * NQ2 + j * NQ + m, ids = k * NQ2 + m * NQ + i, idt = m * NQ2 + j * NQ + i;
      const double Di = D[i * NQ + m], Dj = D[j * NQ + m], Dk = D[k * NQ + m];
      xr += Di * s_x[idr];
      xs += Dj * s_x[ids];
      xt += Dk * s_x[idt];
      yr += Di * s_y[idr];
      ys += Dj * s_y[ids];
      yt += Dk * s_y[idt];
      zr_ += Di * s_z[idr];
      zs += Dj * s_z[ids];
      zt += Dk * s_z[idt];
    }
    const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr_ * (xs * yt - ys * xt);
    const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
    const double sx = -(yr * zt - zr_ * yt) / J, sy = (xr * zt - zr_ * xt) / J, sz = -(xr * yt - yr * xt) / J;
    const double tx = (yr * zs - zr_ * ys) / J, ty = -(xr * zs - zr_ * xs) / J, tz = (xr * ys - yr * xs) / J;
    const double JW = J * wr[i] * wr[j] * wr[k];
    g[0 * NP + n] = JW;
    g[1 * NP + n] = JW * (rx * rx + ry * ry + rz * rz);
    g[2 * NP + n] = JW * (rx * sx + ry * sy + rz * sz);
    g[3 * NP + n] = JW * (rx * tx + ry * ty + rz * tz);
    g[4 * NP + n] = JW * (sx * sx + sy * sy + sz * sz);
    g[5 * NP + n] = JW * (sx * tx + sy * ty + sz * tz);
    g[6 * NP + n] = JW * (tx * tx + ty * ty + tz * tz);
  }
}

// element coordinate and in-element index of the smallest local copy of
// global GLL index I along one direction
__device__ __forceinline__ void first_copy(int64_t I, int N, int64_t& ec, int& ic)
{
  if (I > 0 && I % N == 0) {
    ec = I / N - 1;
    ic = N;
  } else {
    ec = I / N;
    ic = (int)(I % N);
  }
}

__global__ __launch_bounds__(256) void k_box_ids(const int N, const int NX, const int NY, const int NZ,
                                                 const int periodic, const int mask_boundary, const int64_t total,
                                                 int64_t* __restrict__ ids)
{
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= total) return;
  const int Nq = N + 1, Nq2 = Nq * Nq, Np = Nq2 * Nq;
  const int64_t e = id / Np;
  const int n = (int)(id % Np);
  const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / Nq2;
  int64_t I = (int64_t)ex * N + i, J = (int64_t)ey * N + j, K = (int64_t)ez * N + k;
  const bool onbnd = I == 0 || J == 0 || K == 0 || I == (int64_t)NX * N || J == (int64_t)NY * N ||
                     K == (int64_t)NZ * N;
  const int64_t GX = periodic ? (int64_t)NX * N : (int64_t)NX * N + 1;
  const int64_t GY = periodic ? (int64_t)NY * N : (int64_t)NY * N + 1;
  const int64_t GZ = periodic ? (int64_t)NZ * N : (int64_t)NZ * N + 1;
  if (periodic) {
    I %= GX;
    J %= GY;
    K %= GZ;
  }
  const int64_t E = (int64_t)NX * NY * NZ;
  const int64_t Nnodes = periodic ? E : (int64_t)(NX + 1) * (NY + 1) * (NZ + 1);
  int64_t label;
  if ((i == 0 || i == N) && (j == 0 || j == N) && (k == 0 || k == N)) {
    const int64_t vi = ex + (i == N), vj = ey + (j == N), vk = ez + (k == N);
    label = 1 + (periodic ? (vi % NX) + (int64_t)NX * ((vj % NY) + (int64_t)NY * (vk % NZ))
                          : vi + (int64_t)(NX + 1) * (vj + (int64_t)(NY + 1) * vk));
  } else {
    int64_t cx, cy, cz;
    int ix, iy, iz;
    first_copy(I, N, cx, ix);
    first_copy(J, N, cy, iy);
    first_copy(K, N, cz, iz);
    const int64_t first = (cx + NX * (cy + (int64_t)NY * cz)) * Np + ix + Nq * (iy + Nq * iz);
    label = 1 + first + Nnodes;
  }
  ids[id] = (mask_boundary && !periodic && onbnd) ? 0 : label;
}

__global__ __launch_bounds__(256) void k_urand(const uint64_t key, const int64_t n, double* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (double)(splitmix64_d(key + (uint64_t)i) >> 11) * 0x1.0p-52 - 1.0;
}

struct IsZeroId {
  const int64_t* ids;
  __host__ __device__ bool operator()(const int32_t i) const { return ids[i] == 0; }
};

}  // namespace

int box_geometry_launch(const nbx_box& b, double* ggeo, double* x, double* y, double* z, hipStream_t s)
{
  const int N = b.N, Nq = N + 1;
  const int64_t E = (int64_t)b.NX * b.NY * b.NZ;
  std::vector<double> V, zr(Nq), wr(Nq), D((size_t)Nq * Nq);
  box_vertices(b, V);
  nbx_gll(N, zr.data(), wr.data(), D.data());
  double* dv = nullptr;
  NBX_HIP(hipMalloc(&dv, sizeof(double) * (V.size() + 2 * Nq + (size_t)Nq * Nq)));
  double *dzr = dv + V.size(), *dwr = dzr + Nq, *dD = dwr + Nq;
  int rc = NBX_SUCCESS;
  if (hipMemcpy(dv, V.data(), sizeof(double) * V.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dzr, zr.data(), sizeof(double) * Nq, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dwr, wr.data(), sizeof(double) * Nq, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dD, D.data(), sizeof(double) * Nq * Nq, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(NBX_ERR_HIP, "box geometry: upload failed");
  } else {
    switch (Nq) {
#define NBX_GEOM_CASE(NQV)                                                                                     \
  case NQV:                                                                                                    \
    hipLaunchKernelGGL(k_box_geom<NQV>, dim3((unsigned)E), dim3(NQV * NQV), 0, s, b.NX, b.NY, dv, dzr, dwr, dD, \
                       ggeo, x, y, z);                                                                         \
    break;
      NBX_GEOM_CASE(2)
      NBX_GEOM_CASE(3)
      NBX_GEOM_CASE(4)
      NBX_GEOM_CASE(5)
      NBX_GEOM_CASE(6)
      NBX_GEOM_CASE(7)
      NBX_GEOM_CASE(8)
      NBX_GEOM_CASE(9)
      NBX_GEOM_CASE(10)
      NBX_GEOM_CASE(11)
      NBX_GEOM_CASE(12)
      NBX_GEOM_CASE(13)
      NBX_GEOM_CASE(14)
      NBX_GEOM_CASE(15)
      NBX_GEOM_CASE(16)
#undef NBX_GEOM_CASE
      default:
        rc = fail(NBX_ERR_UNSUPPORTED, "box geometry: N=%d", N);
    }
    if (rc == NBX_SUCCESS && hipGetLastError() != hipSuccess) rc = fail(NBX_ERR_HIP, "box geometry: launch failed");
    if (rc == NBX_SUCCESS && hipStreamSynchronize(s) != hipSuccess) rc = fail(NBX_ERR_HIP, "box geometry: failed");
  }
  (void)hipFree(dv);
  return rc;
}

int box_ids_launch(const nbx_box& b, int mask_boundary, int64_t* ids, hipStream_t s)
{
  const int64_t total = (int64_t)b.NX * b.NY * b.NZ * (b.N + 1) * (b.N + 1) * (b.N + 1);
  if (total == 0) return NBX_SUCCESS;
  hipLaunchKernelGGL(k_box_ids, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, b.N, b.NX, b.NY, b.NZ,
                     b.periodic ? 1 : 0, mask_boundary ? 1 : 0, total, ids);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int urand_fill_launch(uint64_t seed, uint64_t stream, int64_t n, double* out, hipStream_t s)
{
  if (n <= 0) return NBX_SUCCESS;
  hipLaunchKernelGGL(k_urand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, urand_key(seed, stream), n, out);
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

// ascending list of the nodes with id 0 (device output, allocated here)
int mask_list_from_ids(int32_t N, const int64_t* ids, int32_t** maskIds, int32_t* Nmasked, hipStream_t s)
{
  int32_t *out = nullptr, *count = nullptr;
  void* tmp = nullptr;
  size_t bytes = 0;
  hipcub::CountingInputIterator<int32_t> it(0);
  IsZeroId pred{ids};
  NBX_HIP(hipcub::DeviceSelect::If(nullptr, bytes, it, out, count, N, pred, s));
  NBX_HIP(hipMalloc(&out, sizeof(int32_t) * std::max<int32_t>(N, 1)));
  if (hipMalloc(&count, sizeof(int32_t)) != hipSuccess || hipMalloc(&tmp, std::max<size_t>(bytes, 1)) != hipSuccess) {
    (void)hipFree(out);
    (void)hipFree(count);
    return fail(NBX_ERR_ALLOC, "mask list: allocation failed");
  }
  int rc = NBX_SUCCESS;
  if (hipcub::DeviceSelect::If(tmp, bytes, it, out, count, N, pred, s) != hipSuccess ||
      hipMemcpyAsync(Nmasked, count, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "mask list: selection failed");
  (void)hipFree(tmp);
  (void)hipFree(count);
  if (rc) {
    (void)hipFree(out);
    return rc;
  }
  *maskIds = out;
  return NBX_SUCCESS;
}

}  // namespace nbx

extern "C" int nbx_box_mesh_device(const nbx_box* box, int mask_boundary, double* ggeo, nbx_hlong* globalIds,
                                   void* stream)
{
  if (!box) return nbx::fail(NBX_ERR_INVALID, "box_mesh_device: null box");
  const nbx_box& b = *box;
  if (b.N < 1 || b.N > 15 || b.NX < 1 || b.NY < 1 || b.NZ < 1)
    return nbx::fail(NBX_ERR_INVALID, "box_mesh_device: bad sizes");
  if (b.periodic && (b.NX < 3 || b.NY < 3 || b.NZ < 3))
    return nbx::fail(NBX_ERR_INVALID, "box_mesh_device: periodic box needs >= 3 elements per direction");
  const int64_t Np = (int64_t)(b.N + 1) * (b.N + 1) * (b.N + 1);
  if ((int64_t)b.NX * b.NY * b.NZ * Np > INT32_MAX)
    return nbx::fail(NBX_ERR_INVALID, "box_mesh_device: E*Np exceeds dlong range");
  if ((ggeo && !nbx::device_accessible(ggeo)) || (globalIds && !nbx::device_accessible(globalIds)))
    return nbx::fail(NBX_ERR_INVALID, "box_mesh_device: outputs must be device memory");
  hipStream_t s = nbx::as_stream(stream);
  int rc;
  if (ggeo && (rc = nbx::box_geometry_launch(b, ggeo, nullptr, nullptr, nullptr, s))) return rc;
  if (globalIds && (rc = nbx::box_ids_launch(b, mask_boundary, globalIds, s))) return rc;
  NBX_HIP(hipStreamSynchronize(s));
  return NBX_SUCCESS;
}

// Learned preconditioner: network handle, checkpoint file, training-pair
// generation from the FP64 PCG, the training loop (Adam) and the attachment
// of a network to a BP handle.  The kernels are in nbx_net.hip.
#include "nbx_bp_internal.h"
#include "nbx_net.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct nbx_net {
  float* master = nullptr;  // FP32 master weights
  float* wbf = nullptr;     // bf16-rounded copy used by every forward
  void* frag = nullptr;     // matrix-core fragments of wbf (rebuilt after every weight change)
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  double* loss = nullptr;
  float* partials = nullptr;
  int partial_cap = 0;
  int2* samples = nullptr;
  int sample_cap = 0;
  int steps = 0;
  uint64_t seed = 0;
  hipStream_t stream = nullptr;
  // training pairs (residual, error) as float, one Nlocal vector each
  float* R = nullptr;
  float* Ev = nullptr;
  double* invNorm = nullptr;
  int npairs = 0, pair_cap = 0;
  int64_t pairStride = 0;
  std::vector<nbx_bp*> attached;  // handles whose bp->net points here
  int infer_mfma = 0;              // inference in the solve: 0 FP32 vector ALUs, 1 bf16 matrix cores
  // energy-norm training: three local vectors and 1 / e^T A e per pair
  double* work = nullptr;
  int64_t work_n = 0;
  double* invNormE = nullptr;
  int energy_pairs = 0;
};

namespace {

__host__ __device__ uint64_t mix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}


// Initial weights: uniform(-a, a) with a = 1/sqrt(fan_in) for Wr, Wg, W2
// (seeded splitmix64), zero biases, and a zero output layer so that the
// untrained network adds nothing: the preconditioner starts as Jacobi.
std::vector<float> init_params(uint64_t seed)
{
  using namespace nbx;
  std::vector<float> p(kNetParams, 0.f);
  const uint64_t key = mix64(seed ^ mix64(77));
  auto u = [&](uint64_t i) { return (float)((double)(mix64(key + i) >> 11) * 0x1.0p-52 - 1.0); };
  const float aR = 1.f / std::sqrt(27.f + 7.f), a2 = 1.f / std::sqrt(32.f);
  for (int q = 0; q < 864; ++q) p[kOffWr + q] = aR * u(q);
  for (int q = 0; q < 224; ++q) p[kOffWg + q] = aR * u(1000 + q);
  for (int q = 0; q < 1024; ++q) p[kOffW2 + q] = a2 * u(2000 + q);
  return p;
}

int upload_params(nbx_net* net, const float* host)
{
  using namespace nbx;
  std::vector<float> bf(kNetParams);
  for (int q = 0; q < kNetParams; ++q) bf[q] = bf16_round(host[q]);
  NBX_HIP(hipMemcpy(net->master, host, sizeof(float) * kNetParams, hipMemcpyHostToDevice));
  NBX_HIP(hipMemcpy(net->wbf, bf.data(), sizeof(float) * kNetParams, hipMemcpyHostToDevice));
  int rc = net_frag_launch(net->wbf, net->frag, nullptr);
  if (rc) return rc;
  NBX_HIP(hipDeviceSynchronize());
  NBX_HIP(hipMemset(net->m, 0, sizeof(float) * kNetParams));
  NBX_HIP(hipMemset(net->v, 0, sizeof(float) * kNetParams));
  net->steps = 0;
  return NBX_SUCCESS;
}

void net_release(nbx_net* n)
{
  if (!n) return;
  (void)hipDeviceSynchronize();
  for (nbx_bp* bp : n->attached) bp->net = nullptr;  // LEARNED solves now fail with an error
  for (void* p : {(void*)n->master, (void*)n->wbf, (void*)n->frag, (void*)n->m, (void*)n->v, (void*)n->grad, (void*)n->loss,
                  (void*)n->partials, (void*)n->samples, (void*)n->R, (void*)n->Ev, (void*)n->invNorm,
                  (void*)n->work, (void*)n->invNormE})
    if (p) (void)hipFree(p);
  delete n;
}

// sum of squares of each ggeo factor over the mesh, two phases: workgroup b
// takes elements [16 b, 16 b + 16), lane t runs an fma chain per factor over
// those elements and nodes t, t + 256, ...; the 256-lane tree gives one
// partial per workgroup and factor; one workgroup combines them in order.
constexpr int kFeatChunk = 16;

__global__ __launch_bounds__(256) void k_factor_sumsq(const int32_t E, const int Np, const double* __restrict__ ggeo,
                                                      double* __restrict__ partials)
{
  __shared__ double s4[7 * 4];
  double v[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) v[q] = 0.0;
  const int64_t e0 = (int64_t)blockIdx.x * kFeatChunk;
  const int64_t e1 = e0 + kFeatChunk < E ? e0 + kFeatChunk : E;
  for (int64_t e = e0; e < e1; ++e)
    for (int n = threadIdx.x; n < Np; n += 256) {
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const double g = ggeo[(e * 7 + q) * Np + n];
        v[q] = __builtin_fma(g, g, v[q]);
      }
    }
  nbx::block_sum256<7>(v, s4);
  nbx::store_partials<7>(v, partials);
}

__global__ __launch_bounds__(256) void k_factor_sumsq_finish(const unsigned int nb, const double* __restrict__ partials,
                                                             double* __restrict__ out)
{
  __shared__ double s4[7 * 4];
  double v[7];
  nbx::combine_partials<7>(v, partials, nb, s4);
  if (threadIdx.x == 0)
    for (int q = 0; q < 7; ++q) out[q] = v[q];
}

__global__ __launch_bounds__(256) void k_to_float_diff(const int64_t n, const double* __restrict__ a,
                                                       const double* __restrict__ b, float* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(b ? a[i] - b[i] : a[i]);
}

// sum of squares of a float vector (one block, fixed order), as double
__global__ __launch_bounds__(256) void k_sumsq_f(const int64_t n, const float* __restrict__ a, double* __restrict__ out)
{
  __shared__ double s4[4];
  double v[1] = {0.0};
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double x = (double)a[i];
    v[0] = __builtin_fma(x, x, v[0]);
  }
  nbx::block_sum256<1>(v, s4);
  if (threadIdx.x == 0) *out = 1.0 / v[0];
}

__global__ __launch_bounds__(256) void k_f2d(const int64_t n, const float* __restrict__ a, double* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)a[i];
}

// z <- z - e (e float)
__global__ __launch_bounds__(256) void k_sub_f(const int64_t n, double* __restrict__ z, const float* __restrict__ e)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = z[i] - (double)e[i];
}

// G = 2 wn w Ad in place of Ad (derivative of wn d^T A d per local copy)
__global__ __launch_bounds__(256) void k_energy_grad(const int64_t n, const double* __restrict__ wn,
                                                     const double* __restrict__ w, double* __restrict__ Ad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Ad[i] = 2.0 * wn[0] * (w[i] * Ad[i]);
}

// scale-invariant loss L = 1 - a^2 gE / b (a = <r,z>_w, b = <z,Az>_w, gE =
// 1 / e^T A e) and G = w dL/dz = w (2 a gE / b) ((a / b) Az - r), into Az
__global__ __launch_bounds__(256) void k_angle_grad(const int64_t n, const double* __restrict__ ab,
                                                    const double* __restrict__ gE, const double* __restrict__ w,
                                                    const double* __restrict__ r, double* __restrict__ Az,
                                                    double* __restrict__ loss)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = ab[0], b = ab[2], g = gE[0];
  if (i == 0) loss[0] = 1.0 - a * a * g / b;
  if (i < n) Az[i] = w[i] * ((2.0 * a * g / b) * ((a / b) * Az[i] - r[i]));
}

__global__ void k_scalar_mul(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = a[0] * b[0];
}

__global__ void k_scalar_inv(const double* __restrict__ a, double* __restrict__ out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = 1.0 / a[0];
}

// samples (p, e) for every element of pair p
__global__ __launch_bounds__(256) void k_pair_samples(const int32_t p, const int32_t E, int2* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < E) out[i] = make_int2(p, i);
}

// training batch: sample i = (mix64(key+c0+2i) mod npairs, mix64(key+c0+2i+1) mod E)
__global__ __launch_bounds__(256) void k_draw_samples(const int32_t batch, const uint64_t key, const uint64_t c0,
                                                      const int32_t npairs, const int32_t E, int2* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  const uint64_t a = mix64(key + c0 + 2ull * (uint64_t)i), b = mix64(key + c0 + 2ull * (uint64_t)i + 1ull);
  out[i] = make_int2((int)(a % (uint64_t)npairs), (int)(b % (uint64_t)E));
}

}  // namespace

namespace nbx {

int net_inference_mode(const nbx_net* net) { return net ? net->infer_mfma : -1; }

void net_detach(nbx_net* net, nbx_bp* bp)
{
  auto& a = net->attached;
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i] == bp) {
      a.erase(a.begin() + (long)i);
      break;
    }
}

int net_apply_learned(nbx_bp* bp, const double* r, double* out, const int* fallback, const int* skip)
{
  if (!bp->net || !bp->feat) return fail(NBX_ERR_INVALID, "learned preconditioner: no network attached");
  if (bp->net->infer_mfma)
    return net_apply_mma_launch(bp->E, bp->net->frag, bp->net->wbf, bp->feat16, r, bp->invDiagA,
                                nbx_ogs_inv_degree(bp->ogs), ogs_degree(bp->ogs), out, fallback, skip, bp->stream);
  return net_apply_launch(bp->E, bp->net->wbf, bp->feat, r, bp->invDiagA, nbx_ogs_inv_degree(bp->ogs), out, fallback,
                          skip, bp->stream);
}

// the learned multigrid step on the matrix cores with its input residual
// (double)(res - Ad) formed inside the network kernel
int net_apply_learned_resid(nbx_bp* bp, const float* res, const float* Ad, double* out, const int* skip)
{
  if (!bp->net || !bp->feat) return fail(NBX_ERR_INVALID, "learned preconditioner: no network attached");
  if (!bp->net->infer_mfma) return fail(NBX_ERR_UNSUPPORTED, "learned step: residual input on the matrix cores only");
  return net_apply_mma_resid_launch(bp->E, bp->net->frag, bp->net->wbf, bp->feat16, res, Ad, bp->invDiagA,
                                    nbx_ogs_inv_degree(bp->ogs), ogs_degree(bp->ogs), out, skip, bp->stream);
}

}  // namespace nbx

using namespace nbx;

extern "C" int nbx_net_num_params(void) { return kNetParams; }

extern "C" nbx_net* nbx_net_create(uint64_t seed, void* stream)
{
  nbx_net* n = new nbx_net();
  n->seed = seed;
  n->stream = as_stream(stream);
  if (hipMalloc(&n->master, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&n->wbf, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&n->frag, kNetFragBytes) != hipSuccess ||
      hipMalloc(&n->m, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&n->v, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&n->grad, sizeof(float) * kNetParams) != hipSuccess || hipMalloc(&n->loss, sizeof(double)) != hipSuccess) {
    fail(NBX_ERR_ALLOC, "net_create: allocation failed");
    net_release(n);
    return nullptr;
  }
  const std::vector<float> p = init_params(seed);
  if (upload_params(n, p.data()) != NBX_SUCCESS) {
    net_release(n);
    return nullptr;
  }
  return n;
}

extern "C" void nbx_net_free(nbx_net* n) { net_release(n); }

extern "C" int nbx_net_get_params(const nbx_net* n, float* host)
{
  if (!n || !host) return fail(NBX_ERR_INVALID, "net_get_params: bad arguments");
  NBX_HIP(hipMemcpy(host, n->master, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
  return NBX_SUCCESS;
}

extern "C" int nbx_net_set_params(nbx_net* n, const float* host)
{
  if (!n || !host) return fail(NBX_ERR_INVALID, "net_set_params: bad arguments");
  return upload_params(n, host);
}

// checkpoint: magic, version, parameter count, training seed, steps, params,
// FNV-1a checksum of the parameter bytes
extern "C" int nbx_net_save(const nbx_net* n, const char* path)
{
  if (!n || !path) return fail(NBX_ERR_INVALID, "net_save: bad arguments");
  std::vector<float> p(kNetParams);
  int rc = nbx_net_get_params(n, p.data());
  if (rc) return rc;
  uint64_t h = 1469598103934665603ull;
  const unsigned char* b = reinterpret_cast<const unsigned char*>(p.data());
  for (size_t i = 0; i < p.size() * 4; ++i) h = (h ^ b[i]) * 1099511628211ull;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(NBX_ERR_INVALID, "net_save: cannot open %s", path);
  const uint32_t hdr[4] = {0x4E425831u /*NBX1*/, 1u, (uint32_t)kNetParams, (uint32_t)n->steps};
  bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(&n->seed, 8, 1, f) == 1 &&
            fwrite(p.data(), 4, p.size(), f) == p.size() && fwrite(&h, 8, 1, f) == 1;
  ok = (fclose(f) == 0) && ok;
  return ok ? NBX_SUCCESS : fail(NBX_ERR_INVALID, "net_save: write failed");
}

extern "C" nbx_net* nbx_net_load(const char* path, void* stream)
{
  FILE* f = path ? fopen(path, "rb") : nullptr;
  if (!f) {
    fail(NBX_ERR_INVALID, "net_load: cannot open %s", path ? path : "(null)");
    return nullptr;
  }
  uint32_t hdr[4];
  uint64_t seed = 0, h = 0;
  std::vector<float> p(kNetParams);
  bool ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == 0x4E425831u && hdr[1] == 1u &&
            hdr[2] == (uint32_t)kNetParams && fread(&seed, 8, 1, f) == 1 &&
            fread(p.data(), 4, p.size(), f) == p.size() && fread(&h, 8, 1, f) == 1;
  fclose(f);
  uint64_t c = 1469598103934665603ull;
  const unsigned char* b = reinterpret_cast<const unsigned char*>(p.data());
  for (size_t i = 0; i < p.size() * 4; ++i) c = (c ^ b[i]) * 1099511628211ull;
  if (!ok || c != h) {
    fail(NBX_ERR_INVALID, "net_load: %s is not a valid checkpoint", path);
    return nullptr;
  }
  nbx_net* n = nbx_net_create(seed, stream);
  if (!n) return nullptr;
  if (nbx_net_set_params(n, p.data()) != NBX_SUCCESS) {
    net_release(n);
    return nullptr;
  }
  n->steps = (int)hdr[3];
  return n;
}

// Attach: features of this mesh (each ggeo factor divided by its RMS over the
// mesh), the element-output buffer, and precon LEARNED becomes usable.
extern "C" int nbx_bp_attach_net(nbx_bp* bp, nbx_net* net)
{
  if (!bp || !net) return fail(NBX_ERR_INVALID, "bp_attach_net: bad arguments");
  if (bp->N != 7) return fail(NBX_ERR_UNSUPPORTED, "learned preconditioner is defined on N=7 (8^3 nodes)");
  if (bp->Nf != 1) return fail(NBX_ERR_UNSUPPORTED, "learned preconditioner: one field");
  if (!bp->feat) NBX_HIP(hipMalloc(&bp->feat, sizeof(float) * (size_t)bp->E * 7 * bp->Np));
  const unsigned int nb = (unsigned int)((bp->E + kFeatChunk - 1) / kFeatChunk);
  double* d_s = nullptr;
  NBX_HIP(hipMalloc(&d_s, sizeof(double) * (14 + 7 * (size_t)nb)));
  double* d_part = d_s + 14;
  hipLaunchKernelGGL(k_factor_sumsq, dim3(nb), dim3(256), 0, bp->stream, bp->E, bp->Np, bp->ggeo, d_part);
  hipLaunchKernelGGL(k_factor_sumsq_finish, dim3(1), dim3(256), 0, bp->stream, nb, d_part, d_s);
  NBX_HIP_LAUNCH();
  double h[7];
  NBX_HIP(hipMemcpyAsync(h, d_s, sizeof(h), hipMemcpyDeviceToHost, bp->stream));
  NBX_HIP(hipStreamSynchronize(bp->stream));
  double inv[7];
  for (int q = 0; q < 7; ++q) {
    const double rms = std::sqrt(h[q] / ((double)bp->E * bp->Np));
    inv[q] = rms > 0 ? 1.0 / rms : 0.0;
  }
  NBX_HIP(hipMemcpyAsync(d_s + 7, inv, sizeof(inv), hipMemcpyHostToDevice, bp->stream));
  int rc = net_features_launch(bp->E, bp->ggeo, d_s + 7, bp->feat, bp->stream);
  if (rc) return rc;
  if (!bp->feat16) NBX_HIP(hipMalloc(&bp->feat16, sizeof(uint16_t) * (size_t)bp->E * 7 * bp->Np));
  if ((rc = net_features16_launch((int64_t)bp->E * 7 * bp->Np, bp->feat, bp->feat16, bp->stream))) return rc;
  NBX_HIP(hipStreamSynchronize(bp->stream));
  (void)hipFree(d_s);
  if (bp->net && bp->net != net) nbx::net_detach(bp->net, bp);
  if (bp->net != net) net->attached.push_back(bp);
  bp->net = net;
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_set_fallback(nbx_bp* bp, int stall_window, double stall_ratio)
{
  if (!bp || stall_window < 1 || !(stall_ratio > 0)) return fail(NBX_ERR_INVALID, "bp_set_fallback: bad arguments");
  bp->stall_window = stall_window;
  bp->stall_ratio = stall_ratio;
  return NBX_SUCCESS;
}

// element-local preconditioner output w invDiag (r + s f(r)) for tests; the
// Jacobi diagonal of the handle must be current (any PCG solve sets it)
extern "C" int nbx_net_apply(nbx_bp* bp, const double* r, double* zl, int weighted)
{
  if (!bp || !r || !zl || !bp->net) return fail(NBX_ERR_INVALID, "net_apply: bad arguments / no network");
  const double* w = weighted ? nbx_ogs_inv_degree(bp->ogs) : nullptr;
  int rc = bp->net->infer_mfma
               ? net_apply_mma_launch(bp->E, bp->net->frag, bp->net->wbf, bp->feat16, r, bp->invDiagA, w,
                                      weighted ? ogs_degree(bp->ogs) : nullptr, zl, nullptr, nullptr, bp->stream)
               : net_apply_launch(bp->E, bp->net->wbf, bp->feat, r, bp->invDiagA, w, zl, nullptr, nullptr, bp->stream);
  if (rc) return rc;
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

extern "C" int nbx_net_set_inference(nbx_net* n, int mode)
{
  if (!n || (mode != 0 && mode != 1)) return fail(NBX_ERR_INVALID, "net_set_inference: mode 0 or 1");
  n->infer_mfma = mode;
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_update_jacobi(nbx_bp* bp)
{
  if (!bp) return fail(NBX_ERR_INVALID, "bp_update_jacobi: null handle");
  double* r = bp->d;
  NBX_HIP(hipMemsetAsync(r, 0, sizeof(double) * bp->Nall, bp->stream));
  NBX_HIP(hipMemsetAsync(bp->b, 0, sizeof(double) * bp->Nall, bp->stream));
  return bp_solve_plain(bp, NBX_PRECON_JACOBI, r, bp->b, 1.0, 0, nullptr);
}

static int ensure_pairs(nbx_net* n, int cap, int64_t stride)
{
  if (n->pair_cap >= cap && n->pairStride == stride) return NBX_SUCCESS;
  for (void* p : {(void*)n->R, (void*)n->Ev, (void*)n->invNorm})
    if (p) (void)hipFree(p);
  n->R = n->Ev = nullptr;
  n->invNorm = nullptr;
  NBX_HIP(hipMalloc(&n->R, sizeof(float) * stride * cap));
  NBX_HIP(hipMalloc(&n->Ev, sizeof(float) * stride * cap));
  NBX_HIP(hipMalloc(&n->invNorm, sizeof(double) * cap));
  n->pair_cap = cap;
  n->pairStride = stride;
  n->npairs = 0;
  return NBX_SUCCESS;
}

// Training pairs from the plain FP64 Jacobi PCG on the handle's mesh: for
// each seed, x* = solve(b_seed) to tol_exact, then for each snapshot k the
// k-th PCG iterate gives (r_k, e_k = x* - x_k).  Pairs are stored as float
// with inv_norm = 1/sum e_k^2.  Returns the number of pairs.
extern "C" int nbx_net_generate_pairs(nbx_net* n, nbx_bp* bp, int nseeds, uint64_t seed0, const int* snapshots,
                                      int nsnap, double tol_exact)
{
  if (!n || !bp || nseeds < 1 || !snapshots || nsnap < 1 || bp->Nf != 1)
    return fail(NBX_ERR_INVALID, "net_generate_pairs: bad arguments");
  const int total = nseeds * nsnap;
  int rc = ensure_pairs(n, total, bp->Nlocal);
  if (rc) return rc;
  double *b = nullptr, *xs = nullptr, *r = nullptr, *x = nullptr;
  NBX_HIP(hipMalloc(&b, sizeof(double) * bp->Nall));
  NBX_HIP(hipMalloc(&xs, sizeof(double) * bp->Nall));
  NBX_HIP(hipMalloc(&r, sizeof(double) * bp->Nall));
  NBX_HIP(hipMalloc(&x, sizeof(double) * bp->Nall));
  const dim3 g((unsigned)((bp->Nlocal + 255) / 256));
  int k = 0;
  for (int sd = 0; sd < nseeds && rc == NBX_SUCCESS; ++sd) {
    rc = bp_make_rhs(bp, seed0 + (uint64_t)sd, b);
    if (rc) break;
    NBX_HIP(hipMemcpyAsync(r, b, sizeof(double) * bp->Nall, hipMemcpyDeviceToDevice, bp->stream));
    NBX_HIP(hipMemsetAsync(xs, 0, sizeof(double) * bp->Nall, bp->stream));
    if ((rc = bp_solve_plain(bp, NBX_PRECON_JACOBI, r, xs, tol_exact, 20000, nullptr))) break;
    for (int j = 0; j < nsnap; ++j) {
      NBX_HIP(hipMemcpyAsync(r, b, sizeof(double) * bp->Nall, hipMemcpyDeviceToDevice, bp->stream));
      NBX_HIP(hipMemsetAsync(x, 0, sizeof(double) * bp->Nall, bp->stream));
      if (snapshots[j] > 0 && (rc = bp_solve_plain(bp, NBX_PRECON_JACOBI, r, x, 0.0, snapshots[j], nullptr))) break;
      float* Rk = n->R + (int64_t)k * n->pairStride;
      float* Ek = n->Ev + (int64_t)k * n->pairStride;
      hipLaunchKernelGGL(k_to_float_diff, g, dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, r, nullptr, Rk);
      hipLaunchKernelGGL(k_to_float_diff, g, dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, xs, x, Ek);
      hipLaunchKernelGGL(k_sumsq_f, dim3(1), dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, Ek, n->invNorm + k);
      ++k;
    }
  }
  NBX_HIP(hipStreamSynchronize(bp->stream));
  for (void* p : {(void*)b, (void*)xs, (void*)r, (void*)x}) (void)hipFree(p);
  if (rc) return rc;
  n->npairs = k;
  n->energy_pairs = 0;
  return k;
}

// Training pairs for the learned multigrid step (nbx_bp_mg_set_learned): the
// residual and the error a V-cycle leaves behind.  For each seed, x* solves
// A x = b_seed by the multigrid PCG to tol_exact; for k = kfirst ..
// kfirst+kcount-1 the k-th iterate x_k of the same PCG gives r_k = b - A x_k
// and e_k = x* - x_k; one V-cycle without the learned step, z = M r_k, then
// leaves
//   r' = r_k - A z,   e' = e_k - z     (A e' = r'),
// stored as a float pair.  Returns the number of pairs.
extern "C" int nbx_net_generate_mg_pairs(nbx_net* n, nbx_bp* bp, int nseeds, uint64_t seed0, int kfirst,
                                         int kcount, double tol_exact)
{
  if (!n || !bp || nseeds < 1 || kfirst < 0 || kcount < 1 || bp->Nf != 1 || !(tol_exact > 0))
    return fail(NBX_ERR_INVALID, "net_generate_mg_pairs: bad arguments");
  const int maxk = kcount;
  if (bp->opt.precon != NBX_PRECON_MG || !bp->mg)
    return fail(NBX_ERR_INVALID, "net_generate_mg_pairs: the handle needs precon MG and nbx_bp_mg_setup");
  const int total = nseeds * maxk;
  int rc = ensure_pairs(n, total, bp->Nlocal);
  if (rc) return rc;
  const int was_learned = nbx_bp_mg_get_learned(bp);
  if (was_learned && (rc = nbx_bp_mg_set_learned(bp, 0))) return rc;
  double *b = nullptr, *xs = nullptr, *r = nullptr, *x = nullptr, *z = nullptr, *Az = nullptr;
  const size_t bytes = sizeof(double) * bp->Nall;
  NBX_HIP(hipMalloc(&b, bytes));
  NBX_HIP(hipMalloc(&xs, bytes));
  NBX_HIP(hipMalloc(&r, bytes));
  NBX_HIP(hipMalloc(&x, bytes));
  NBX_HIP(hipMalloc(&z, bytes));
  NBX_HIP(hipMalloc(&Az, bytes));
  const dim3 g((unsigned)((bp->Nlocal + 255) / 256));
  int k = 0;
  for (int sd = 0; sd < nseeds && rc == NBX_SUCCESS; ++sd) {
    if ((rc = bp_make_rhs(bp, seed0 + (uint64_t)sd, b))) break;
    NBX_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, bp->stream));
    NBX_HIP(hipMemsetAsync(xs, 0, bytes, bp->stream));
    if ((rc = bp_solve_plain(bp, NBX_PRECON_MG, r, xs, tol_exact, 1000, nullptr))) break;
    for (int j = kfirst; j < kfirst + maxk && rc == NBX_SUCCESS; ++j) {
      NBX_HIP(hipMemcpyAsync(r, b, bytes, hipMemcpyDeviceToDevice, bp->stream));
      NBX_HIP(hipMemsetAsync(x, 0, bytes, bp->stream));
      if (j > 0 && (rc = bp_solve_plain(bp, NBX_PRECON_MG, r, x, 0.0, j, nullptr))) break;
      // one V-cycle on r_k, then the residual and error it leaves
      if ((rc = mg_apply_direct(bp, r, z)) || (rc = bp_apply(bp, bp->lambda, z, Az, nullptr)) ||
          (rc = nbx_scaled_add((nbx_dlong)bp->Nall, -1.0, Az, 1.0, r, bp->stream)) ||
          (rc = nbx_scaled_add((nbx_dlong)bp->Nall, 1.0, z, 1.0, x, bp->stream)))
        break;
      float* Rk = n->R + (int64_t)k * n->pairStride;
      float* Ek = n->Ev + (int64_t)k * n->pairStride;
      hipLaunchKernelGGL(k_to_float_diff, g, dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, r, nullptr, Rk);
      hipLaunchKernelGGL(k_to_float_diff, g, dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, xs, x, Ek);
      hipLaunchKernelGGL(k_sumsq_f, dim3(1), dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, Ek, n->invNorm + k);
      NBX_HIP_LAUNCH();
      ++k;
    }
  }
  (void)hipStreamSynchronize(bp->stream);
  for (void* p : {(void*)b, (void*)xs, (void*)r, (void*)x, (void*)z, (void*)Az}) (void)hipFree(p);
  if (was_learned) (void)nbx_bp_mg_set_learned(bp, 1);
  if (rc) return rc;
  n->npairs = k;
  n->energy_pairs = 0;
  return k;
}

// Host copies of stored pairs first .. first+count-1 (Nlocal floats each).
extern "C" int nbx_net_copy_pairs(const nbx_net* n, int first, int count, float* R, float* Ev)
{
  if (!n || first < 0 || count < 0 || first + count > n->npairs)
    return fail(NBX_ERR_INVALID, "net_copy_pairs: pairs %d..%d of %d", first, first + count - 1, n ? n->npairs : 0);
  const size_t off = (size_t)first * n->pairStride, bytes = sizeof(float) * (size_t)count * n->pairStride;
  if (R) NBX_HIP(hipMemcpy(R, n->R + off, bytes, hipMemcpyDeviceToHost));
  if (Ev) NBX_HIP(hipMemcpy(Ev, n->Ev + off, bytes, hipMemcpyDeviceToHost));
  return NBX_SUCCESS;
}

// Store caller-provided pairs (device float vectors of length Nlocal each).
extern "C" int nbx_net_set_pairs(nbx_net* n, nbx_bp* bp, int npairs, const float* R, const float* Ev)
{
  if (!n || !bp || npairs < 1 || !R || !Ev) return fail(NBX_ERR_INVALID, "net_set_pairs: bad arguments");
  int rc = ensure_pairs(n, npairs, bp->Nlocal);
  if (rc) return rc;
  NBX_HIP(hipMemcpy(n->R, R, sizeof(float) * bp->Nlocal * (size_t)npairs, hipMemcpyDeviceToDevice));
  NBX_HIP(hipMemcpy(n->Ev, Ev, sizeof(float) * bp->Nlocal * (size_t)npairs, hipMemcpyDeviceToDevice));
  for (int k = 0; k < npairs; ++k)
    hipLaunchKernelGGL(k_sumsq_f, dim3(1), dim3(256), 0, bp->stream, (int64_t)bp->Nlocal,
                       n->Ev + (int64_t)k * n->pairStride, n->invNorm + k);
  NBX_HIP(hipStreamSynchronize(bp->stream));
  n->npairs = npairs;
  n->energy_pairs = 0;
  return NBX_SUCCESS;
}

static int ensure_work(nbx_net* n, int nsamples, int chunk)
{
  if (nsamples > n->sample_cap) {
    if (n->samples) (void)hipFree(n->samples);
    n->samples = nullptr;
    NBX_HIP(hipMalloc(&n->samples, sizeof(int2) * nsamples));
    n->sample_cap = nsamples;
  }
  const int nblocks = (nsamples + chunk - 1) / chunk;
  if (nblocks > n->partial_cap) {
    if (n->partials) (void)hipFree(n->partials);
    n->partials = nullptr;
    NBX_HIP(hipMalloc(&n->partials, sizeof(float) * kNetPartial * (size_t)nblocks));
    n->partial_cap = nblocks;
  }
  return NBX_SUCCESS;
}

// gradient over the nsamples samples already in n->samples (G: energy mode)
static int run_gradient(nbx_net* n, nbx_bp* bp, int nsamples, int chunk, const double* G = nullptr)
{
  const int nblocks = (nsamples + chunk - 1) / chunk;
  int rc = net_train_launch(nsamples, chunk, n->samples, n->pairStride, n->R, n->Ev, n->invNorm, n->wbf, bp->feat,
                            bp->invDiagA, G, n->partials, bp->stream);
  if (rc) return rc;
  return net_grad_sum_launch(nblocks, n->partials, n->grad, n->loss, bp->stream);
}

// Gradient of the training loss over the given (pair, element) samples, in
// the library's fixed order (chunk samples per workgroup); host outputs.
extern "C" int nbx_net_gradient(nbx_net* n, nbx_bp* bp, const int32_t* samples, int nsamples, int chunk,
                                const double* G, float* grad_host, double* loss_host)
{
  if (!n || !bp || !samples || nsamples < 1 || chunk < 1 || !bp->feat || n->npairs < 1)
    return fail(NBX_ERR_INVALID, "net_gradient: bad arguments (network attached, pairs set?)");
  for (int i = 0; i < nsamples; ++i)
    if (samples[2 * i] < 0 || samples[2 * i] >= n->npairs || samples[2 * i + 1] < 0 || samples[2 * i + 1] >= bp->E)
      return fail(NBX_ERR_INVALID, "net_gradient: sample %d out of range", i);
  int rc = ensure_work(n, nsamples, chunk);
  if (rc) return rc;
  NBX_HIP(hipMemcpy(n->samples, samples, sizeof(int2) * nsamples, hipMemcpyHostToDevice));
  if ((rc = run_gradient(n, bp, nsamples, chunk, G))) return rc;
  if (grad_host) NBX_HIP(hipMemcpyAsync(grad_host, n->grad, sizeof(float) * kNetParams, hipMemcpyDeviceToHost, bp->stream));
  if (loss_host) NBX_HIP(hipMemcpyAsync(loss_host, n->loss, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

// Adam training on the stored pairs: each step draws `batch` (pair, element)
// samples uniformly (seeded splitmix64), computes the gradient and updates.
// loss_hist (host, may be null) receives each step's loss / batch.
extern "C" int nbx_net_train(nbx_net* n, nbx_bp* bp, int steps, int batch, float lr, uint64_t seed,
                             double* loss_hist)
{
  if (!n || !bp || steps < 0 || batch < 1 || !bp->feat || n->npairs < 1 || !(lr > 0))
    return fail(NBX_ERR_INVALID, "net_train: bad arguments (network attached, pairs generated?)");
  const int chunk = 16;
  int rc = ensure_work(n, batch, chunk);
  if (rc) return rc;
  const uint64_t key = mix64(seed ^ mix64(91));
  for (int st = 0; st < steps; ++st) {
    hipLaunchKernelGGL(k_draw_samples, dim3((batch + 255) / 256), dim3(256), 0, bp->stream, batch, key,
                       2ull * (uint64_t)batch * (uint64_t)st, n->npairs, bp->E, n->samples);
    NBX_HIP_LAUNCH();
    if ((rc = run_gradient(n, bp, batch, chunk))) return rc;
    n->steps += 1;
    rc = net_adam_launch(n->master, n->m, n->v, n->grad, n->wbf, lr, n->steps, bp->stream);
    if (rc == NBX_SUCCESS) rc = net_frag_launch(n->wbf, n->frag, bp->stream);
    if (rc) return rc;
    if (loss_hist) {
      double l;
      NBX_HIP(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
      NBX_HIP(hipStreamSynchronize(bp->stream));
      loss_hist[st] = l / batch;
    }
  }
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

// Energy-norm training.  One pair per step, all its elements in one batch:
//   z = mask(gs(w invDiag (r + s f(r)))),  d = z - e,
//   L = d^T A d / e^T A e   (weighted dots over local nodes, FP64 operator),
// i.e. the A-norm of the error left by one preconditioned Richardson step,
// relative to the error before it.  dL/dz_local = 2 w A d / e^T A e is
// back-propagated through every element.  loss_hist (host, may be null)
// receives each step's L.
// mode 1: the scale-invariant form L = 1 - <r,z>^2 / (e^T A e <z,Az>), the
// A-norm error left by the best multiple of z (CG does not see the scale of
// the preconditioner).
extern "C" int nbx_net_train_energy(nbx_net* n, nbx_bp* bp, int steps, float lr, uint64_t seed, int mode,
                                    double* loss_hist)
{
  if (!n || !bp || steps < 0 || !bp->feat || n->npairs < 1 || !(lr > 0) || bp->Nf != 1 ||
      n->pairStride != bp->Nlocal || (mode != 0 && mode != 1))
    return fail(NBX_ERR_INVALID, "net_train_energy: bad arguments (network attached, pairs of this mesh?)");
  const int64_t Nall = bp->Nall, Nl = bp->Nlocal;
  if (n->work_n < Nall) {
    if (n->work) (void)hipFree(n->work);
    n->work = nullptr;
    NBX_HIP(hipMalloc(&n->work, sizeof(double) * 3 * Nall));
    n->work_n = Nall;
  }
  double* rd = n->work;
  double* zb = n->work + Nall;
  double* Ad = n->work + 2 * Nall;
  NBX_HIP(hipMemsetAsync(n->work, 0, sizeof(double) * 3 * Nall, bp->stream));
  const double* w = nbx_ogs_inv_degree(bp->ogs);
  const dim3 g((unsigned)((Nl + 255) / 256)), b(256);
  if (n->energy_pairs != n->npairs) {  // 1 / e^T A e for every pair
    if (n->invNormE) (void)hipFree(n->invNormE);
    n->invNormE = nullptr;
    NBX_HIP(hipMalloc(&n->invNormE, sizeof(double) * n->npairs));
    for (int p = 0; p < n->npairs; ++p) {
      hipLaunchKernelGGL(k_f2d, g, b, 0, bp->stream, Nl, n->Ev + (int64_t)p * n->pairStride, zb);
      int rc = bp_apply(bp, bp->lambda, zb, Ad, nullptr);
      if (rc) return rc;
      rc = launch_wdot(bp->Nlocal, bp->fieldOffset, 1, w, zb, Ad, false, bp->red, bp->red.result, bp->stream);
      if (rc) return rc;
      hipLaunchKernelGGL(k_scalar_inv, dim3(1), dim3(64), 0, bp->stream, bp->red.result, n->invNormE + p);
    }
    NBX_HIP_LAUNCH();
    n->energy_pairs = n->npairs;
  }
  const int chunk = 16;
  int rc = ensure_work(n, bp->E, chunk);
  if (rc) return rc;
  const uint64_t key = mix64(seed ^ mix64(93));
  for (int st = 0; st < steps; ++st) {
    const int p = (int)(mix64(key + (uint64_t)st) % (uint64_t)n->npairs);
    const float* Rp = n->R + (int64_t)p * n->pairStride;
    const float* Ep = n->Ev + (int64_t)p * n->pairStride;
    hipLaunchKernelGGL(k_f2d, g, b, 0, bp->stream, Nl, Rp, rd);
    if ((rc = net_apply_launch(bp->E, n->wbf, bp->feat, rd, bp->invDiagA, w, zb, nullptr, nullptr, bp->stream)))
      return rc;
    if ((rc = gs_add_mask_launch(bp->ogs, zb, 1, bp->fieldOffset, bp->Nmasked, bp->maskIds, nullptr, bp->stream)))
      return rc;
    if (mode == 0) {
      hipLaunchKernelGGL(k_sub_f, g, b, 0, bp->stream, Nl, zb, Ep);
      if ((rc = bp_apply(bp, bp->lambda, zb, Ad, nullptr))) return rc;
      if ((rc = launch_wdot(bp->Nlocal, bp->fieldOffset, 1, w, zb, Ad, false, bp->red, bp->red.result, bp->stream)))
        return rc;
      hipLaunchKernelGGL(k_energy_grad, g, b, 0, bp->stream, Nl, n->invNormE + p, w, Ad);
    } else {
      if ((rc = bp_apply(bp, bp->lambda, zb, Ad, nullptr))) return rc;
      if ((rc = launch_wdot(bp->Nlocal, bp->fieldOffset, 1, w, rd, zb, false, bp->red, bp->red.result, bp->stream)) ||
          (rc = launch_wdot(bp->Nlocal, bp->fieldOffset, 1, w, zb, Ad, false, bp->red, bp->red.result + 2,
                            bp->stream)))
        return rc;
      hipLaunchKernelGGL(k_angle_grad, g, b, 0, bp->stream, Nl, bp->red.result, n->invNormE + p, w, rd, Ad,
                         bp->red.result + 4);
    }
    hipLaunchKernelGGL(k_pair_samples, dim3((bp->E + 255) / 256), b, 0, bp->stream, p, bp->E, n->samples);
    NBX_HIP_LAUNCH();
    if ((rc = run_gradient(n, bp, bp->E, chunk, Ad))) return rc;
    if (mode == 0)
      hipLaunchKernelGGL(k_scalar_mul, dim3(1), dim3(64), 0, bp->stream, bp->red.result, n->invNormE + p, n->loss);
    else
      NBX_HIP(hipMemcpyAsync(n->loss, bp->red.result + 4, sizeof(double), hipMemcpyDeviceToDevice, bp->stream));
    n->steps += 1;
    if ((rc = net_adam_launch(n->master, n->m, n->v, n->grad, n->wbf, lr, n->steps, bp->stream))) return rc;
    if ((rc = net_frag_launch(n->wbf, n->frag, bp->stream))) return rc;
    if (loss_hist) {
      double l;
      NBX_HIP(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
      NBX_HIP(hipStreamSynchronize(bp->stream));
      loss_hist[st] = l;
    }
  }
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

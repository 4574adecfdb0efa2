// p-multigrid preconditioner for the BP PCG (precon NBX_PRECON_MG).
// (The cycle is enqueued once into a HIP graph and replayed: ~150 small
// kernels per cycle, most of them on the coarse levels, where launch gaps
// dominate.)
//
// Levels share the mesh: level l discretises the same box at polynomial
// order N_l (e.g. 7, 3, 1), each with its own BP handle (geometric factors,
// gather-scatter, Dirichlet mask and Jacobi diagonal at that order).  One
// application z = M r is a V-cycle:
//
//   level l < L-1:  Chebyshev-Jacobi pre-smoothing of degree k from x = 0
//                   (k operator applies, the residual is carried along),
//                   restriction  r_c = mask(gs_c(J^T (w res)))   per element,
//                   x_c = V-cycle(l+1, r_c),
//                   x += J x_c (element-local interpolation of the continuous
//                   coarse correction, itself continuous),
//                   Chebyshev-Jacobi post-smoothing of degree k.
//   level L-1:      a fixed number of Jacobi-PCG iterations from x = 0.
//
// J is the Lagrange interpolation from the coarse GLL nodes to the fine GLL
// nodes, applied as three 1-D contractions (x, then y, then z).  Chebyshev
// bounds are [lo, hi] * lambda_max(D^-1 A), lambda_max from a seeded power
// iteration at setup.  Every kernel and every sum follows a fixed order that
// the CPU oracle restates, so V-cycles and PCG iterates are reproduced bit for
// bit.  All work is enqueued on the handle's stream; every kernel honours the
// outer solver's done flag so the iterations enqueued past convergence cost
// nothing.
//
// Learned step (nbx_bp_mg_set_learned, a network attached with
// nbx_bp_attach_net): the fine level ends its cycle with one application of
// the element network to the residual the cycle leaves there,
//   rf = res - A d  (the fine residual after post-smoothing),
//   x += mask(gs(w invDiag (rf + s f(rf)))),
// a learned local correction trained on the error V-cycles leave behind.
//
// Precision: NBX_F64 runs the whole cycle in FP64.  NBX_F32 and NBX_BF16 run
// the smoothers and transfers of every level but the last in FP32 (vectors,
// Jacobi diagonal, interpolation and arithmetic), with the operator's ggeo and
// lambda in FP32 or bf16; the cycle reads the FP64 residual and returns an
// FP64 correction, and the coarse PCG stays FP64.
#include "nbx_bp_internal.h"
#include "nbx_axhelm_impl.h"
#include "nbx_blas.h"
#include <cmath>
#include <type_traits>
#include <vector>

struct nbx_mg_level {
  nbx_bp* bp = nullptr;  // level 0: the solver's handle (not owned)
  bool owned = false;
  int N = 0, Nq = 0;
  double lmax = 0.0;
  std::vector<double> c1, c2;  // Chebyshev step coefficients
  // vectors of the cycle's precision (double or float); the last level's r
  // and x are double (its PCG); level 0 reads r and writes x of the solver
  void *r = nullptr, *x = nullptr, *res = nullptr, *d = nullptr, *Ad = nullptr;
  double* J = nullptr;  // [Nq][Nq of level l+1]: interpolation from the next coarser level
  float* J32 = nullptr;
  float* invD32 = nullptr;  // FP32 Jacobi diagonal (FP32 / bf16 cycles)
};

struct nbx_mg {
  std::vector<nbx_mg_level> lv;
  int degree = 2, coarse_iters = 20;
  double lo = 0.1, hi = 1.1;
  nbx_prec prec = NBX_F64;
  nbx::PcgScalars* coarse_template = nullptr;  // device copy of the coarse solve's initial scalars
  // coarse CG in two kernels per iteration (coarse level of order 1, every
  // assembled node with at most 8 members): the member table of the
  // assembled nodes, the two passes' block partials and the rz slots
  bool cg_fused = false;
  int32_t cg_items = 0;
  int32_t* cg_members = nullptr;  // [items][8], -1 past the node's members
  double *cg_pa = nullptr, *cg_pb = nullptr, *cg_scal = nullptr;
  bool learned = false;        // fine level ends with the attached network's correction
  double* lrf = nullptr;       // learned step: the fine residual after the cycle (FP64 copy)
  double* lzl = nullptr;       // learned step: element-local network output
  std::vector<void*> owned;
  // the cycle as a HIP graph, captured for one (r, z, skip) and replayed
  bool use_graph = true;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_r = nullptr;
  void* g_z = nullptr;
  const int* g_skip = nullptr;
  const void* g_net = nullptr;  // the learned step's network and inference mode at capture:
  int g_infer = -1;             // the captured kernels depend on both
};

namespace nbx {

inline dim3 mg_grid(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename V>
__device__ __forceinline__ V fmaV(V a, V b, V c);
template <>
__device__ __forceinline__ double fmaV<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <>
__device__ __forceinline__ float fmaV<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Chebyshev-Jacobi step (n local nodes per field, fields at stride offset,
// blockIdx.y = field, one Jacobi diagonal for all fields), arithmetic in V; the
// level's right-hand side is R, its solution X (level 0: the solver's FP64
// residual and correction):
//   MODE 0 (first pre-smoothing step, x = 0):  res = (V)r; d = c2 D^-1 res; x = d
//   MODE 1 (general step):  res -= Ad; d = c1 d + c2 D^-1 res; x += d
//   MODE 2 (first post-smoothing step):  res -= Ad; d = c2 D^-1 res; x += d
template <int MODE, typename V, typename R, typename X>
__global__ __launch_bounds__(256) void k_cheb(const int32_t n, const int32_t offset, const V c1, const V c2,
                                              const V* __restrict__ invD, const R* __restrict__ r,
                                              V* __restrict__ res, const V* __restrict__ Ad, V* __restrict__ d,
                                              X* __restrict__ x, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;  // node; blockIdx.y = field
  if (i >= n) return;
  const int64_t id = (int64_t)blockIdx.y * offset + i;
  V ri;
  if (MODE == 0) ri = (V)r[id];
  else ri = res[id] - Ad[id];
  res[id] = ri;
  const V t = c2 * (invD[i] * ri);
  const V di = (MODE == 1) ? fmaV<V>(c1, d[id], t) : t;
  d[id] = di;
  x[id] = (MODE == 0) ? (X)di : x[id] + (X)di;
}

// Restriction of one element: res -= Ad (stored), u = w res, then
// rc[a,b,c] = sum_k J[k][c] sum_j J[j][b] sum_i J[i][a] u[i,j,k]
// (node index i + j Nq + k Nq^2; fma chains in ascending fine index).
// NQF, NQC: the orders at compile time (the loops unroll, the index
// arithmetic folds); 0 takes them from the arguments.  Either way each
// output is the same fma chain.  Launched with 256 lanes.
template <typename V, typename O, int NQF = 0, int NQC = 0>
__global__ __launch_bounds__(256) void k_mg_restrict(const int nqf, const int nqc, const int32_t offf,
                                                     const int32_t offc, const V* __restrict__ J,
                                                     const double* __restrict__ w, V* __restrict__ res,
                                                     const V* __restrict__ Ad, O* __restrict__ rc,
                                                     const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ V s_u[12 * 12 * 12];
  __shared__ V s_t[12 * 12 * 12];
  __shared__ V s_J[12 * 12];
  constexpr int kLanes = 256;
  const int Nqf = NQF ? NQF : nqf, Nqc = NQC ? NQC : nqc;
  const int64_t e = blockIdx.x;
  const int Npf = Nqf * Nqf * Nqf, Npc = Nqc * Nqc * Nqc;
  for (int q = threadIdx.x; q < Nqf * Nqc; q += kLanes) s_J[q] = J[q];
  const int64_t fo = (int64_t)blockIdx.y * offf, fc = (int64_t)blockIdx.y * offc;  // blockIdx.y = field
  for (int q = threadIdx.x; q < Npf; q += kLanes) {
    const int64_t id = e * Npf + q;
    const V rq = res[fo + id] - Ad[fo + id];
    res[fo + id] = rq;
    s_u[q] = (V)w[id] * rq;
  }
  __syncthreads();
  // x: t1[a + Nqc (j + Nqf k)]
  for (int q = threadIdx.x; q < Nqc * Nqf * Nqf; q += kLanes) {
    const int a = q % Nqc, jk = q / Nqc;
    V acc = V(0);
    for (int i = 0; i < Nqf; ++i) acc = fmaV<V>(s_J[i * Nqc + a], s_u[i + Nqf * jk], acc);
    s_t[q] = acc;
  }
  __syncthreads();
  // y: t2[a + Nqc (b + Nqc k)] into s_u
  for (int q = threadIdx.x; q < Nqc * Nqc * Nqf; q += kLanes) {
    const int a = q % Nqc, b = (q / Nqc) % Nqc, k = q / (Nqc * Nqc);
    V acc = V(0);
    for (int j = 0; j < Nqf; ++j) acc = fmaV<V>(s_J[j * Nqc + b], s_t[a + Nqc * (j + Nqf * k)], acc);
    s_u[q] = acc;
  }
  __syncthreads();
  // z: rc[a + Nqc (b + Nqc c)]
  for (int q = threadIdx.x; q < Npc; q += kLanes) {
    const int ab = q % (Nqc * Nqc), c = q / (Nqc * Nqc);
    V acc = V(0);
    for (int k = 0; k < Nqf; ++k) acc = fmaV<V>(s_J[k * Nqc + c], s_u[ab + Nqc * Nqc * k], acc);
    rc[fc + e * Npc + q] = (O)acc;
  }
}

// Prolongation of one element: v[i,j,k] = sum_c J[k][c] sum_b J[j][b] sum_a J[i][a] xc[a,b,c];
// d = v, x += v.
template <typename V, typename I, typename X, int NQF = 0, int NQC = 0>
__global__ __launch_bounds__(256) void k_mg_prolong(const int nqf, const int nqc, const int32_t offf,
                                                    const int32_t offc, const V* __restrict__ J,
                                                    const I* __restrict__ xc, V* __restrict__ d,
                                                    X* __restrict__ x, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ V s_u[12 * 12 * 12];
  __shared__ V s_t[12 * 12 * 12];
  __shared__ V s_J[12 * 12];
  constexpr int kLanes = 256;
  const int Nqf = NQF ? NQF : nqf, Nqc = NQC ? NQC : nqc;
  const int64_t e = blockIdx.x;
  const int Npf = Nqf * Nqf * Nqf, Npc = Nqc * Nqc * Nqc;
  for (int q = threadIdx.x; q < Nqf * Nqc; q += kLanes) s_J[q] = J[q];
  const int64_t fo = (int64_t)blockIdx.y * offf, fc = (int64_t)blockIdx.y * offc;  // blockIdx.y = field
  for (int q = threadIdx.x; q < Npc; q += kLanes) s_u[q] = (V)xc[fc + e * Npc + q];
  __syncthreads();
  // x: t1[i + Nqf (b + Nqc c)]
  for (int q = threadIdx.x; q < Nqf * Nqc * Nqc; q += kLanes) {
    const int i = q % Nqf, bc = q / Nqf;
    V acc = V(0);
    for (int a = 0; a < Nqc; ++a) acc = fmaV<V>(s_J[i * Nqc + a], s_u[a + Nqc * bc], acc);
    s_t[q] = acc;
  }
  __syncthreads();
  // y: t2[i + Nqf (j + Nqf c)] into s_u
  for (int q = threadIdx.x; q < Nqf * Nqf * Nqc; q += kLanes) {
    const int i = q % Nqf, j = (q / Nqf) % Nqf, c = q / (Nqf * Nqf);
    V acc = V(0);
    for (int b = 0; b < Nqc; ++b) acc = fmaV<V>(s_J[j * Nqc + b], s_t[i + Nqf * (b + Nqc * c)], acc);
    s_u[q] = acc;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < Npf; q += kLanes) {
    const int ij = q % (Nqf * Nqf), k = q / (Nqf * Nqf);
    V acc = V(0);
    for (int c = 0; c < Nqc; ++c) acc = fmaV<V>(s_J[k * Nqc + c], s_u[ij + Nqf * Nqf * c], acc);
    const int64_t id = fo + e * Npf + q;
    d[id] = acc;
    x[id] = x[id] + (X)acc;
  }
}

// learned step: rf = res - Ad as FP64 (the network's input)
template <typename V>
__global__ __launch_bounds__(256) void k_mg_resid(const int64_t n, const V* __restrict__ res, const V* __restrict__ Ad,
                                                  double* __restrict__ rf, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rf[i] = (double)(res[i] - Ad[i]);
}

// learned step: x += z (assembled network output)
__global__ __launch_bounds__(256) void k_mg_add(const int64_t n, const double* __restrict__ z, double* __restrict__ x,
                                                const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + z[i];
}

__global__ __launch_bounds__(256) void k_mg_to_f32(const int64_t n, const double* __restrict__ a,
                                                   float* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)a[i];
}

__global__ __launch_bounds__(256) void k_mg_fill0(const int64_t n, double* __restrict__ v,
                                                  const int* __restrict__ skip)
{
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

// coarse solve scalars <- template; the solve is skipped when the outer one is done
__global__ void k_mg_reset(PcgScalars* __restrict__ S, const PcgScalars* __restrict__ tmpl,
                           const int* __restrict__ skip)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  *S = *tmpl;
  if (skip && *skip) S->done = 1;
}

// dst = a * (invD ? invD src : src)
__global__ __launch_bounds__(256) void k_mg_scale(const int64_t n, const double a, const double* __restrict__ invD,
                                                  const double* __restrict__ src, double* __restrict__ dst)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = a * (invD ? invD[i] * src[i] : src[i]);
}

}  // namespace nbx

using namespace nbx;

namespace {

// Lagrange interpolation from nodes zc (Nqc) to nodes zf (Nqf): J[i][a] =
// prod_{m != a} (zf[i] - zc[m]) / (zc[a] - zc[m]), factors in ascending m.
std::vector<double> interp_matrix(int Nf, int Nc)
{
  std::vector<double> zf(Nf + 1), wf(Nf + 1), zc(Nc + 1), wc(Nc + 1);
  nbx_gll(Nf, zf.data(), wf.data(), nullptr);
  nbx_gll(Nc, zc.data(), wc.data(), nullptr);
  std::vector<double> J((size_t)(Nf + 1) * (Nc + 1));
  for (int i = 0; i <= Nf; ++i)
    for (int a = 0; a <= Nc; ++a) {
      double l = 1.0;
      for (int m = 0; m <= Nc; ++m)
        if (m != a) l = l * ((zf[i] - zc[m]) / (zc[a] - zc[m]));
      J[(size_t)i * (Nc + 1) + a] = l;
    }
  return J;
}

void cheb_coeffs(double lmax, double lo, double hi, int degree, std::vector<double>& c1, std::vector<double>& c2)
{
  const double a = lo * lmax, b = hi * lmax;
  const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
  double rho = 1.0 / sigma;
  c1.assign(degree, 0.0);
  c2.assign(degree, 0.0);
  c2[0] = 1.0 / theta;
  for (int s = 1; s < degree; ++s) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c1[s] = rn * rho;
    c2[s] = 2.0 * rn / delta;
    rho = rn;
  }
}

template <typename P>
int mg_alloc(nbx_mg* mg, P** p, int64_t n, size_t elem = sizeof(P))
{
  NBX_HIP(hipMalloc((void**)p, elem * (size_t)std::max<int64_t>(n, 1)));
  NBX_HIP(hipMemset(*p, 0, elem * (size_t)std::max<int64_t>(n, 1)));
  mg->owned.push_back(*p);
  return NBX_SUCCESS;
}

// lambda_max(D^-1 A) by power iteration on field 0 (the vector operator is
// block-diagonal with equal blocks): v0 = gs(mask(U[-1,1) seeded)),
// normalised in the weighted norm; v <- D^-1 A v / ||D^-1 A v||_w; the
// estimate is the last norm.  Norms are taken on the host (setup only).
int power_lmax(nbx_bp* bp, uint64_t seed, uint64_t stream_id, int iters, double* v, double* w, double* out)
{
  const int64_t n = bp->Nlocal;
  int rc = urand_fill_launch(seed, stream_id, n, v, bp->stream);
  if (rc) return rc;
  if ((rc = mask_launch(bp->Nmasked, bp->maskIds, v, nullptr, bp->stream))) return rc;
  if ((rc = gs_add_mask_launch(bp->ogs, v, 1, bp->fieldOffset, 0, nullptr, nullptr, bp->stream))) return rc;
  const double* wt = nbx_ogs_inv_degree(bp->ogs);
  double nrm = 0.0;
  auto norm2 = [&](const double* a) -> int {
    int r2 = launch_wdot(bp->Nlocal, bp->fieldOffset, 1, wt, a, a, false, bp->red, bp->red.result, bp->stream);
    if (r2) return r2;
    NBX_HIP(hipMemcpyAsync(&nrm, bp->red.result, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
    NBX_HIP(hipStreamSynchronize(bp->stream));
    return NBX_SUCCESS;
  };
  if ((rc = norm2(v))) return rc;
  if (!(nrm > 0)) return fail(NBX_ERR_INVALID, "mg: power iteration start vector is zero");
  hipLaunchKernelGGL(k_mg_scale, mg_grid(n), dim3(256), 0, bp->stream, n, 1.0 / std::sqrt(nrm), nullptr, v, v);
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    if ((rc = bp_apply(bp, bp->lambda, v, w, nullptr))) return rc;
    hipLaunchKernelGGL(k_mg_scale, mg_grid(n), dim3(256), 0, bp->stream, n, 1.0, bp->invDiagA, w, w);
    if ((rc = norm2(w))) return rc;
    lam = std::sqrt(nrm);
    hipLaunchKernelGGL(k_mg_scale, mg_grid(n), dim3(256), 0, bp->stream, n, 1.0 / lam, nullptr, w, v);
  }
  NBX_HIP_LAUNCH();
  *out = lam;
  return NBX_SUCCESS;
}

// ---- coarse CG in two kernels per iteration ------------------------------
// The coarse level (order 1) runs K Jacobi-PCG iterations from x = 0:
//   z = D^-1 r, rz_0 = <r, z>;  for k = 0..K-1:
//   beta = k ? rz_k / rz_{k-1} : 0;  p = z + beta p;  Ap_L = A_L p (element-local);
//   pAp = sum_e p_e . (A_L p)_e;  alpha = rz_k / pAp;
//   per assembled node g: Ap_g = sum of its members' Ap_L (member order of the
//   gather-scatter); x += alpha p, r -= alpha Ap_g, z = D^-1 r at every member;
//   rz_{k+1} = sum_g sum_fields r_g z_g;
//   a non-finite beta or alpha (the residual reached zero) is taken as 0.
// Pass A (k_cg_pass_a) does the p update, the element operator and pAp's block
// partials; pass B (k_cg_pass_b) the gather, the updates and rz's block
// partials.  Each pass combines the other pass's partials itself, every
// workgroup in the same fixed order, so no separate reduction kernel and no
// device-scope fence is needed; rz_k goes to slot k % 2 (written by pass A's
// workgroup 0, read by the next pass B and pass A).
//   Pass A: a quad of lanes per element (lane t = i + 2j of the 2 x 2 face,
//   columns k = 0, 1), 64 elements per 256-lane workgroup; the element
//   operator is axhelm_column's arithmetic for Nq = 2 with the other lanes'
//   values brought by DPP quad permutes instead of LDS, so no barrier; the
//   element share is the fused PCG's lane chain (fields, then k) and lane tree
//   (a0 + a2) + (a1 + a3); the partial is a halving tree over each wave's 16
//   shares (zero past the last element), then (S0 + S2) + (S1 + S3).
//   Pass B: 8 lanes per assembled node (lane n = member n), 32 nodes per
//   workgroup; Ap_g summed from 0 in member order, every member updated; the
//   node's value fma(r_g, z_g, .) over fields in its lane 0, the others 0,
//   summed by the 256-lane tree.
// Partials are combined as tile partials.  The oracle's coarse CG restates
// all of this order.
constexpr int kCgEpb = 64;   // elements per pass-A workgroup
constexpr int kCgNpb = 32;   // assembled nodes per pass-B workgroup

// the combined value of nb partials in the tile-partial order, in every lane
// of a 256-lane workgroup
__device__ __forceinline__ double cg_combine(const double* __restrict__ part, unsigned nb, double* s4)
{
  const int tid = (int)threadIdx.x;
  double acc = 0.0;
  unsigned i = (unsigned)tid;
  for (; i + 7 * 256 < nb; i += 8 * 256) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = part[i + u * 256];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + t[u];
  }
  for (; i < nb; i += 256) acc = acc + part[i];
  const double w = wave_sum(acc);
  if ((tid & 63) == 0) s4[tid >> 6] = w;
  __syncthreads();
  const double tot = (s4[0] + s4[2]) + (s4[1] + s4[3]);
  __syncthreads();  // s4 is reused
  return tot;
}

// the value of the quad lane that pattern Q (quad_perm) names for this lane
template <int Q>
__device__ __forceinline__ double quad(double v)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, Q, 0xf, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), Q, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// quad_perm patterns: lane t = i + 2j reads lane 2j (m = 0 along r), 2j + 1
// (m = 1 along r), i (m = 0 along s), i + 2 (m = 1 along s)
constexpr int kQr0 = 0 | 0 << 2 | 2 << 4 | 2 << 6, kQr1 = 1 | 1 << 2 | 3 << 4 | 3 << 6;
constexpr int kQs0 = 0 | 1 << 2 | 0 << 4 | 1 << 6, kQs1 = 2 | 3 << 2 | 2 << 4 | 3 << 6;

template <int NF, bool BK>
__global__ __launch_bounds__(256) void k_cg_pass_a(const int32_t E, const int32_t offset,
                                                   const double* __restrict__ ggeo, const double* __restrict__ D,
                                                   const double* __restrict__ lambda, const double* __restrict__ z,
                                                   double* __restrict__ p, double* __restrict__ Ap,
                                                   double* __restrict__ pa, const double* __restrict__ pb,
                                                   const unsigned nbB, double* __restrict__ scal, const int k,
                                                   const int* __restrict__ skip)
{
  constexpr int NP = 8;
  if (skip && *skip) return;
  __shared__ double s4[4];
  const int tid = (int)threadIdx.x, lane = tid & 63, t = tid & 3, i = t & 1, j = t >> 1;
  // workgroup w runs on XCD w % 8: XCD x takes the x-th eighth of the element
  // blocks, so its L2 holds the same part of the mesh in both passes
  const unsigned per = gridDim.x / 8, xb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int64_t e = (int64_t)xb * kCgEpb + (tid >> 2);
  if (xb * kCgEpb >= (unsigned)E) return;
  const bool active = e < E;  // an idle quad computes on zeros and stores nothing
  const double D00 = D[0], D01 = D[1], D10 = D[2], D11 = D[3];
  const double Di0 = i ? D10 : D00, Di1 = i ? D11 : D01, Dj0 = j ? D10 : D00, Dj1 = j ? D11 : D01;
  const double DTi0 = i ? D01 : D00, DTi1 = i ? D11 : D10, DTj0 = j ? D01 : D00, DTj1 = j ? D11 : D10;
  const double Dk[2][2] = {{D00, D01}, {D10, D11}};  // Dk[k][m] = D[k][m]
  double G[2][6], lam0[2], wl1[2];
  const double* g = ggeo + e * (int64_t)(kNggeo * NP) + t;
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const int n = kk * 4;
    G[kk][0] = active ? ldnt(&g[kG00 * NP + n]) : 0.0;
    G[kk][1] = active ? ldnt(&g[kG01 * NP + n]) : 0.0;
    G[kk][2] = active ? ldnt(&g[kG02 * NP + n]) : 0.0;
    G[kk][3] = active ? ldnt(&g[kG11 * NP + n]) : 0.0;
    G[kk][4] = active ? ldnt(&g[kG12 * NP + n]) : 0.0;
    G[kk][5] = active ? ldnt(&g[kG22 * NP + n]) : 0.0;
    lam0[kk] = (!BK && active) ? ldnt(&lambda[e * NP + n + t]) : 0.0;
    wl1[kk] = (!BK && active) ? ldnt(&g[kGWJ * NP + n]) * ldnt(&lambda[e * NP + offset + n + t]) : 0.0;
  }
  // this element's p and z (every field) are loaded before the partials are
  // combined, so their latency overlaps the combination's
  double pv[NF][2], zv[NF][2];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int64_t id = e * NP + (int64_t)f * offset + t + kk * 4;
      pv[f][kk] = active ? p[id] : 0.0;
      zv[f][kk] = active ? ldnt(&z[id]) : 0.0;
    }
  const double rz = cg_combine(pb, nbB, s4);
  double beta = k == 0 ? 0.0 : rz / scal[(k - 1) & 1];
  if (!isfinite(beta)) beta = 0.0;  // r reached zero: the iterate stays
  if (xb == 0 && tid == 0) scal[k & 1] = rz;
  double acc = 0.0;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int64_t base = e * NP + (int64_t)f * offset + t;
    double pk[2], out[2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      pk[kk] = active ? __builtin_fma(beta, pv[f][kk], zv[f][kk]) : 0.0;
      if (active) p[base + kk * 4] = pk[kk];
    }
    // axhelm_column for Nq = 2 (same operations in the same order)
    double rAt[2] = {0.0, 0.0}, rArs[2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const double q = pk[kk];
      double qr = 0.0, qs = 0.0, qt = 0.0;
      qr = __builtin_fma(Di0, quad<kQr0>(q), qr);
      qs = __builtin_fma(Dj0, quad<kQs0>(q), qs);
      qt = __builtin_fma(Dk[kk][0], pk[0], qt);
      qr = __builtin_fma(Di1, quad<kQr1>(q), qr);
      qs = __builtin_fma(Dj1, quad<kQs1>(q), qs);
      qt = __builtin_fma(Dk[kk][1], pk[1], qt);
      const double G00 = G[kk][0], G01 = G[kk][1], G02 = G[kk][2], G11 = G[kk][3], G12 = G[kk][4],
                   G22 = G[kk][5];
      double Gr = __builtin_fma(G02, qt, __builtin_fma(G01, qs, G00 * qr));
      double Gs = __builtin_fma(G12, qt, __builtin_fma(G11, qs, G01 * qr));
      double Gt = __builtin_fma(G22, qt, __builtin_fma(G12, qs, G02 * qr));
      if (!BK) {
        Gr = lam0[kk] * Gr;
        Gs = lam0[kk] * Gs;
        Gt = lam0[kk] * Gt;
      }
      rAt[0] = __builtin_fma(Dk[kk][0], Gt, rAt[0]);
      rAt[1] = __builtin_fma(Dk[kk][1], Gt, rAt[1]);
      double ar = 0.0, as = 0.0;
      ar = __builtin_fma(DTi0, quad<kQr0>(Gr), ar);
      as = __builtin_fma(DTj0, quad<kQs0>(Gs), as);
      ar = __builtin_fma(DTi1, quad<kQr1>(Gr), ar);
      as = __builtin_fma(DTj1, quad<kQs1>(Gs), as);
      rArs[kk] = ar + as;
    }
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      double res = rArs[kk] + rAt[kk];
      if (!BK) res = __builtin_fma(wl1[kk], pk[kk], res);
      out[kk] = res;
      if (active) stnt(res, &Ap[base + kk * 4]);
      acc = __builtin_fma(pk[kk], out[kk], acc);
    }
  }
  // element share (a0 + a2) + (a1 + a3); zero for an idle quad
  const double s2 = acc + quad<(2 | 3 << 2 | 0 << 4 | 1 << 6)>(acc);
  double sh = s2 + quad<(1 | 0 << 2 | 3 << 4 | 2 << 6)>(s2);
  sh = active ? sh : 0.0;
  // the wave's 16 shares (lanes 4s): halving tree s += s + h, h = 8, 4, 2, 1
#pragma unroll
  for (int h = 8; h >= 1; h >>= 1) sh = sh + __shfl(sh, (lane + 4 * h) & 63, 64);
  if (lane == 0) s4[tid >> 6] = sh;
  __syncthreads();
  if (tid == 0) pa[xb] = (s4[0] + s4[2]) + (s4[1] + s4[3]);
}

// pass B (UPD = false: the initial pass, z = D^-1 r and rz_0 only)
template <int NF, bool UPD>
__global__ __launch_bounds__(256) void k_cg_pass_b(const int32_t items, const int32_t* __restrict__ members,
                                                   const int32_t offset, const double* __restrict__ invD,
                                                   const double* __restrict__ p, const double* __restrict__ Ap,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, const double* __restrict__ pa,
                                                   const unsigned nbA, const double* __restrict__ scal, const int k,
                                                   double* __restrict__ pb, const int* __restrict__ skip)
{
  if (skip && *skip) return;
  __shared__ double s4[4];
  const int tid = (int)threadIdx.x, lane = tid & 63, g0 = lane & ~7, n = tid & 7;
  const unsigned per = gridDim.x / 8, xb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);  // as pass A
  if (xb * kCgNpb >= (unsigned)items) return;
  const int32_t item = (int32_t)xb * kCgNpb + (tid >> 3);
  const int32_t m = item < items ? members[(int64_t)item * 8 + n] : -1;
  const double di = m >= 0 ? invD[m] : 0.0;
  // the member's values (every field) before the partials are combined
  double rv[NF], av[NF], pv[NF], xv[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int64_t id = (int64_t)f * offset + (m >= 0 ? m : 0);
    rv[f] = m >= 0 ? r[id] : 0.0;
    av[f] = (UPD && m >= 0) ? Ap[id] : 0.0;
    pv[f] = (UPD && m >= 0) ? p[id] : 0.0;
    xv[f] = (UPD && m >= 0) ? x[id] : 0.0;
  }
  double alpha = 0.0;
  if (UPD) {
    alpha = scal[k & 1] / cg_combine(pa, nbA, s4);
    if (!isfinite(alpha)) alpha = 0.0;
  }
  double v[1] = {0.0};
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int64_t id = (int64_t)f * offset + (m >= 0 ? m : 0);
    double rr = rv[f];
    if (UPD) {
      const double a = av[f];
      double ag = 0.0;  // the gather-scatter's sum, in its member order
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double aq = __shfl(a, g0 + q, 64);
        const int32_t mq = __shfl(m, g0 + q, 64);
        if (mq >= 0) ag = ag + aq;
      }
      rr = __builtin_fma(-alpha, ag, rr);
      if (m >= 0) {
        x[id] = __builtin_fma(alpha, pv[f], xv[f]);
        r[id] = rr;
      }
    }
    const double zz = di * rr;
    if (m >= 0) z[id] = zz;
    if (n == 0) v[0] = __builtin_fma(rr, zz, v[0]);
  }
  block_sum256<1>(v, s4);
  if (tid == 0) pb[xb] = v[0];
}

// the member table of the coarse level's assembled nodes (gather-scatter
// group order; nodes with more than 8 members leave the fused CG off)
static int cg_setup(nbx_mg* mg, nbx_bp* b)
{
  nbx_dlong N = 0, Nl = 0, G = 0;
  if (nbx_ogs_sizes(b->ogs, &N, &Nl, &G, nullptr, nullptr) != NBX_SUCCESS)
    return fail(NBX_ERR_HIP, "mg: coarse ogs sizes");
  std::vector<nbx_dlong> off((size_t)G + 1), ids((size_t)std::max<nbx_dlong>(Nl, 1));
  if (nbx_ogs_copy_csr(b->ogs, off.data(), ids.data()) != NBX_SUCCESS) return fail(NBX_ERR_HIP, "mg: coarse csr");
  std::vector<int32_t> tab((size_t)std::max<nbx_dlong>(G, 1) * 8, -1);
  for (nbx_dlong g = 0; g < G; ++g) {
    const nbx_dlong n = off[g + 1] - off[g];
    if (n > 8) return NBX_SUCCESS;  // stays on the generic PCG
    for (nbx_dlong q = 0; q < n; ++q) tab[(size_t)g * 8 + q] = ids[off[g] + q];
  }
  int rc;
  const unsigned nbA = (unsigned)((b->E + kCgEpb - 1) / kCgEpb), nbB = (unsigned)((G + kCgNpb - 1) / kCgNpb);
  if ((rc = mg_alloc(mg, &mg->cg_members, (int64_t)tab.size())) || (rc = mg_alloc(mg, &mg->cg_pa, nbA)) ||
      (rc = mg_alloc(mg, &mg->cg_pb, nbB)) || (rc = mg_alloc(mg, &mg->cg_scal, 2)))
    return rc;
  NBX_HIP(hipMemcpy(mg->cg_members, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice));
  // masked nodes belong to no assembled node: their z and p stay zero
  NBX_HIP(hipMemset(b->z, 0, sizeof(double) * (size_t)b->Nall));
  NBX_HIP(hipMemset(b->p, 0, sizeof(double) * (size_t)b->Nall));
  mg->cg_items = (int32_t)G;
  mg->cg_fused = true;
  return NBX_SUCCESS;
}

static int coarse_cg(nbx_mg* mg, nbx_mg_level& L, const int* skip)
{
  nbx_bp* bp = L.bp;
  double* x = static_cast<double*>(L.x);
  double* r = static_cast<double*>(L.r);
  const unsigned nbA = (unsigned)((bp->E + kCgEpb - 1) / kCgEpb),
                 nbB = (unsigned)((mg->cg_items + kCgNpb - 1) / kCgNpb);
  hipLaunchKernelGGL(k_mg_fill0, mg_grid(bp->Nall), dim3(256), 0, bp->stream, bp->Nall, x, skip);
  if (mg->cg_items == 0 || nbA == 0) {  // every coarse node masked: the solution is x = 0
    NBX_HIP_LAUNCH();
    return NBX_SUCCESS;
  }
  const dim3 gA(8 * ((nbA + 7) / 8)), gB(8 * ((nbB + 7) / 8));  // eight XCD chunks of blocks
#define NBX_CGB(NFV, UPD, KK)                                                                                   \
  hipLaunchKernelGGL((k_cg_pass_b<NFV, UPD>), gB, dim3(256), 0, bp->stream, mg->cg_items, mg->cg_members,          \
                     bp->fieldOffset, bp->invDiagA, bp->p, bp->Ap, x, r, bp->z, mg->cg_pa, nbA, mg->cg_scal, KK,   \
                     mg->cg_pb, skip)
#define NBX_CGA(NFV, BKV, KK)                                                                                  \
  hipLaunchKernelGGL((k_cg_pass_a<NFV, BKV>), gA, dim3(256), 0, bp->stream, bp->E, bp->fieldOffset,                \
                     bp->ggeo, bp->D, bp->lambda, bp->z, bp->p, bp->Ap, mg->cg_pa, mg->cg_pb, nbB, mg->cg_scal, KK, \
                     skip)
  const bool bk = bp->opt.bpmode != 0;
  if (bp->Nf == 1) NBX_CGB(1, false, 0); else NBX_CGB(3, false, 0);
  for (int k = 0; k < mg->coarse_iters; ++k) {
    if (bp->Nf == 1) {
      if (bk) NBX_CGA(1, true, k); else NBX_CGA(1, false, k);
      NBX_CGB(1, true, k);
    } else {
      if (bk) NBX_CGA(3, true, k); else NBX_CGA(3, false, k);
      NBX_CGB(3, true, k);
    }
  }
#undef NBX_CGA
#undef NBX_CGB
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

int coarse_solve(nbx_mg* mg, nbx_mg_level& L, const int* skip)
{
  if (mg->cg_fused) return coarse_cg(mg, L, skip);
  nbx_bp* bp = L.bp;
  double* x = static_cast<double*>(L.x);
  hipLaunchKernelGGL(k_mg_fill0, mg_grid(bp->Nall), dim3(256), 0, bp->stream, bp->Nall, x, skip);
  hipLaunchKernelGGL(k_mg_reset, dim3(1), dim3(64), 0, bp->stream, bp->S, mg->coarse_template, skip);
  NBX_HIP_LAUNCH();
  for (int it = 0; it <= mg->coarse_iters; ++it) {
    int rc = bp_enqueue_iteration(bp, bp->lambda, static_cast<double*>(L.r), x);
    if (rc) return rc;
  }
  return NBX_SUCCESS;
}

// level operator Aq = mask(gs(A q)) in the cycle's precision
int level_apply(nbx_mg* mg, nbx_mg_level& L, const double* q, double* Aq, const int* skip)
{
  return bp_apply(L.bp, L.bp->lambda, q, Aq, skip);
}

int level_apply(nbx_mg* mg, nbx_mg_level& L, const float* q, float* Aq, const int* skip)
{
  nbx_bp* bp = L.bp;  // FP32 operator: Nf fields; bf16 geometry: one field (checked at setup)
  const bool bf = mg->prec == NBX_BF16 || mg->prec == NBX_BF16_MFMA;
  // matrix-core contractions at N = 7; the other levels use the bf16-geometry FP32 kernel
  const nbx_prec p = (mg->prec == NBX_BF16_MFMA && bp->N != 7) ? NBX_BF16 : mg->prec;
  int rc = axhelm_launch(bp->N, bp->Nf, bp->opt.bpmode ? 1 : 0, p, bp->E, bp->fieldOffset, nullptr,
                         bf ? (const void*)bp->ggeo16 : (const void*)bp->ggeo32, bp->D32,
                         bf ? (const void*)bp->lambda16 : (const void*)bp->lambda32, q, Aq, skip, bp->stream);
  if (rc) return rc;
  return gs_add_mask_launch(bp->ogs, Aq, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, skip, bp->stream);
}

template <typename V>
const V* level_J(const nbx_mg_level& L);
template <>
const double* level_J<double>(const nbx_mg_level& L) { return L.J; }
template <>
const float* level_J<float>(const nbx_mg_level& L) { return L.J32; }

template <typename V>
const V* level_invD(const nbx_mg_level& L);
template <>
const double* level_invD<double>(const nbx_mg_level& L) { return L.bp->invDiagA; }
template <>
const float* level_invD<float>(const nbx_mg_level& L) { return L.invD32; }

template <typename V>
int vcycle(nbx_mg* mg, int l, const void* r, void* x, const int* skip);

// one level of the cycle: right-hand side R, solution X, coarse vectors C
template <typename V, typename R, typename X, typename CV>
int vcycle_level(nbx_mg* mg, int l, const R* r, X* x, const int* skip)
{
  nbx_mg_level& L = mg->lv[l];
  nbx_mg_level& C = mg->lv[l + 1];
  nbx_bp* bp = L.bp;
  const int32_t n = bp->Nlocal, off = bp->fieldOffset;
  const dim3 g((unsigned)((n + 255) / 256), (unsigned)bp->Nf), b(256);
  const dim3 ge((unsigned)bp->E, (unsigned)bp->Nf);
  const int k = mg->degree;
  V* res = static_cast<V*>(L.res);
  V* d = static_cast<V*>(L.d);
  V* Ad = static_cast<V*>(L.Ad);
  const V* invD = level_invD<V>(L);
  int rc;
  // pre-smoothing from x = 0
  hipLaunchKernelGGL((k_cheb<0, V, R, X>), g, b, 0, bp->stream, n, off, V(0), (V)L.c2[0], invD, r, res,
                     (const V*)nullptr, d, x, skip);
  for (int s = 1; s < k; ++s) {
    if ((rc = level_apply(mg, L, d, Ad, skip))) return rc;
    hipLaunchKernelGGL((k_cheb<1, V, R, X>), g, b, 0, bp->stream, n, off, (V)L.c1[s], (V)L.c2[s], invD,
                       (const R*)nullptr, res, Ad, d, x, skip);
  }
  if ((rc = level_apply(mg, L, d, Ad, skip))) return rc;
  // restriction of the smoothed residual
  CV* rcv = static_cast<CV*>(C.r);
  {
    const double* wv = nbx_ogs_inv_degree(bp->ogs);
#define NBX_RESTRICT(F, C_)                                                                                  \
  hipLaunchKernelGGL((k_mg_restrict<V, CV, F, C_>), ge, b, 0, bp->stream, L.Nq, C.Nq, off, C.bp->fieldOffset, \
                     level_J<V>(L), wv, res, Ad, rcv, skip)
    if (L.Nq == 8 && C.Nq == 4) NBX_RESTRICT(8, 4);
    else if (L.Nq == 4 && C.Nq == 2) NBX_RESTRICT(4, 2);
    else if (L.Nq == 10 && C.Nq == 5) NBX_RESTRICT(10, 5);
    else if (L.Nq == 5 && C.Nq == 2) NBX_RESTRICT(5, 2);
    else NBX_RESTRICT(0, 0);
#undef NBX_RESTRICT
  }
  NBX_HIP_LAUNCH();
  nbx_bp* bc = C.bp;
  if ((rc = gs_add_mask_launch(bc->ogs, rcv, bc->Nf, bc->fieldOffset, bc->Nmasked, bc->maskIds, skip, bc->stream)))
    return rc;
  if ((rc = vcycle<V>(mg, l + 1, C.r, C.x, skip))) return rc;
  // coarse correction
#define NBX_PROLONG(F, C_)                                                                                 \
  hipLaunchKernelGGL((k_mg_prolong<V, CV, X, F, C_>), ge, b, 0, bp->stream, L.Nq, C.Nq, off, C.bp->fieldOffset, \
                     level_J<V>(L), static_cast<const CV*>(C.x), d, x, skip)
  if (L.Nq == 8 && C.Nq == 4) NBX_PROLONG(8, 4);
  else if (L.Nq == 4 && C.Nq == 2) NBX_PROLONG(4, 2);
  else if (L.Nq == 10 && C.Nq == 5) NBX_PROLONG(10, 5);
  else if (L.Nq == 5 && C.Nq == 2) NBX_PROLONG(5, 2);
  else NBX_PROLONG(0, 0);
#undef NBX_PROLONG
  NBX_HIP_LAUNCH();
  if ((rc = level_apply(mg, L, d, Ad, skip))) return rc;
  // post-smoothing
  hipLaunchKernelGGL((k_cheb<2, V, R, X>), g, b, 0, bp->stream, n, off, V(0), (V)L.c2[0], invD, (const R*)nullptr, res,
                     Ad, d, x, skip);
  for (int s = 1; s < k; ++s) {
    if ((rc = level_apply(mg, L, d, Ad, skip))) return rc;
    hipLaunchKernelGGL((k_cheb<1, V, R, X>), g, b, 0, bp->stream, n, off, (V)L.c1[s], (V)L.c2[s], invD,
                       (const R*)nullptr, res, Ad, d, x, skip);
  }
  NBX_HIP_LAUNCH();
  if constexpr (std::is_same<X, double>::value) {  // the fine level (its correction is FP64)
  if (l == 0 && mg->learned) {
    // learned step: rf = res - A d; x += mask(gs(w invDiag (rf + s f(rf))))
    if ((rc = level_apply(mg, L, d, Ad, skip))) return rc;
    const dim3 gn((unsigned)((n + 255) / 256));
    bool fused = false;  // matrix-core inference reads res and A d and forms rf itself
    if constexpr (std::is_same<V, float>::value) {
      if (net_inference_mode(bp->net) == 1) {
        if ((rc = net_apply_learned_resid(bp, res, Ad, mg->lzl, skip))) return rc;
        fused = true;
      }
    }
    if (!fused) {
      hipLaunchKernelGGL((k_mg_resid<V>), gn, b, 0, bp->stream, (int64_t)n, res, Ad, mg->lrf, skip);
      NBX_HIP_LAUNCH();
      if ((rc = net_apply_learned(bp, mg->lrf, mg->lzl, nullptr, skip))) return rc;
    }
    if ((rc = gs_add_mask_launch(bp->ogs, mg->lzl, 1, off, bp->Nmasked, bp->maskIds, skip, bp->stream))) return rc;
    hipLaunchKernelGGL(k_mg_add, gn, b, 0, bp->stream, (int64_t)n, mg->lzl, x, skip);
    NBX_HIP_LAUNCH();
  }
  }
  return NBX_SUCCESS;
}

template <typename V>
int vcycle(nbx_mg* mg, int l, const void* r, void* x, const int* skip)
{
  const int L = (int)mg->lv.size();
  if (l + 1 == L) return coarse_solve(mg, mg->lv[l], skip);
  const bool coarse_next = l + 2 == L;  // the next level is the FP64 coarse PCG
  if (l == 0) {
    const double* rr = static_cast<const double*>(r);
    double* xx = static_cast<double*>(x);
    return coarse_next ? vcycle_level<V, double, double, double>(mg, l, rr, xx, skip)
                       : vcycle_level<V, double, double, V>(mg, l, rr, xx, skip);
  }
  const V* rr = static_cast<const V*>(r);
  V* xx = static_cast<V*>(x);
  return coarse_next ? vcycle_level<V, V, V, double>(mg, l, rr, xx, skip)
                     : vcycle_level<V, V, V, V>(mg, l, rr, xx, skip);
}

}  // namespace

namespace nbx {

static void mg_drop_graph(nbx_mg* mg)
{
  if (mg->exec) (void)hipGraphExecDestroy(mg->exec);
  if (mg->graph) (void)hipGraphDestroy(mg->graph);
  mg->exec = nullptr;
  mg->graph = nullptr;
  mg->g_r = nullptr;
  mg->g_z = nullptr;
  mg->g_skip = nullptr;
  mg->g_net = nullptr;
  mg->g_infer = -1;
}

void mg_release(nbx_mg* mg)
{
  if (!mg) return;
  (void)hipDeviceSynchronize();
  mg_drop_graph(mg);
  if (mg->capture) (void)hipStreamDestroy(mg->capture);
  for (auto& L : mg->lv)
    if (L.owned && L.bp) nbx_bp_free(L.bp);
  for (void* p : mg->owned) (void)hipFree(p);
  if (mg->coarse_template) (void)hipFree(mg->coarse_template);
  delete mg;
}

static int mg_enqueue(nbx_mg* mg, const double* r, double* z, const int* skip)
{
  if (mg->prec == NBX_F64) return vcycle<double>(mg, 0, r, z, skip);
  return vcycle<float>(mg, 0, r, z, skip);
}

// Capture the cycle for (r, z, skip) on a private stream: every level's
// handle enqueues on it while capturing, then returns to its own stream.
static int mg_capture(nbx_mg* mg, nbx_bp* bp, const double* r, double* z, const int* skip)
{
  mg_drop_graph(mg);
  if (!mg->capture) NBX_HIP(hipStreamCreateWithFlags(&mg->capture, hipStreamNonBlocking));
  std::vector<hipStream_t> saved;
  for (auto& L : mg->lv) {
    saved.push_back(L.bp->stream);
    L.bp->stream = mg->capture;
  }
  int rc = NBX_SUCCESS;
  if (hipStreamBeginCapture(mg->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(NBX_ERR_HIP, "mg: stream capture could not start");
  } else {
    rc = mg_enqueue(mg, r, z, skip);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(mg->capture, &g);
    if (rc == NBX_SUCCESS && e != hipSuccess) rc = fail(NBX_ERR_HIP, "mg: stream capture failed");
    if (rc == NBX_SUCCESS && hipGraphInstantiate(&mg->exec, g, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(NBX_ERR_HIP, "mg: graph instantiation failed");
    mg->graph = g;
  }
  for (size_t l = 0; l < mg->lv.size(); ++l) mg->lv[l].bp->stream = saved[l];
  if (rc) {
    mg_drop_graph(mg);
    return rc;
  }
  mg->g_r = r;
  mg->g_z = z;
  mg->g_skip = skip;
  mg->g_net = mg->learned ? bp->net : nullptr;
  mg->g_infer = mg->learned ? net_inference_mode(bp->net) : -1;
  return NBX_SUCCESS;
}

int mg_apply_direct(nbx_bp* bp, const double* r, double* z)
{
  if (!bp->mg) return fail(NBX_ERR_INVALID, "multigrid preconditioner: not set up (nbx_bp_mg_setup)");
  return mg_enqueue(bp->mg, r, z, nullptr);
}

int mg_apply(nbx_bp* bp, const double* r, double* z, const int* skip)
{
  nbx_mg* mg = bp->mg;
  if (!mg) return fail(NBX_ERR_INVALID, "multigrid preconditioner: not set up (nbx_bp_mg_setup)");
  if (!mg->use_graph) return mg_enqueue(mg, r, z, skip);
  const bool net_changed = mg->learned && (mg->g_net != bp->net || mg->g_infer != net_inference_mode(bp->net));
  if (!mg->exec || mg->g_r != r || mg->g_z != z || mg->g_skip != skip || net_changed) {
    int rc = mg_capture(mg, bp, r, z, skip);
    if (rc) return rc;
  }
  NBX_HIP(hipGraphLaunch(mg->exec, bp->stream));
  return NBX_SUCCESS;
}

}  // namespace nbx

extern "C" int nbx_bp_mg_setup(nbx_bp* bp, int nlevels, const int* orders, int degree, int coarse_iters,
                               double lo, double hi, nbx_prec prec)
{
  if (!bp || nlevels < 2 || nlevels > 8 || !orders || degree < 1 || degree > 16 || coarse_iters < 1 ||
      !(lo > 0) || !(hi > lo))
    return fail(NBX_ERR_INVALID, "bp_mg_setup: bad arguments");
  if (prec != NBX_F64 && prec != NBX_F32 && prec != NBX_BF16 && prec != NBX_BF16_MFMA)
    return fail(NBX_ERR_UNSUPPORTED, "bp_mg_setup: precision %d (F64, F32, BF16 or BF16_MFMA)", (int)prec);
  if (bp->allNeumann)  // singular: the coarse problem has the constants in its null space
    return fail(NBX_ERR_UNSUPPORTED, "multigrid preconditioner: all-Neumann problem (no mask, lambda1 = 0)");
  if ((prec == NBX_BF16 || prec == NBX_BF16_MFMA) && bp->Nf != 1)
    return fail(NBX_ERR_UNSUPPORTED, "multigrid preconditioner: bf16 geometry with one field only");
  if (orders[0] != bp->N) return fail(NBX_ERR_INVALID, "bp_mg_setup: orders[0] must be the handle's N");
  for (int l = 1; l < nlevels; ++l)
    if (orders[l] < 1 || orders[l] >= orders[l - 1])
      return fail(NBX_ERR_INVALID, "bp_mg_setup: orders must decrease to >= 1");
  if (orders[0] > 11) return fail(NBX_ERR_UNSUPPORTED, "multigrid preconditioner: N <= 11");
  if (bp->mg) {
    mg_release(bp->mg);
    bp->mg = nullptr;
  }
  nbx_mg* mg = new nbx_mg();
  mg->degree = degree;
  mg->coarse_iters = coarse_iters;
  mg->lo = lo;
  mg->hi = hi;
  mg->prec = prec;
  mg->lv.resize(nlevels);
  const size_t es = prec == NBX_F64 ? sizeof(double) : sizeof(float);
  int rc = NBX_SUCCESS;
  for (int l = 0; l < nlevels && rc == NBX_SUCCESS; ++l) {
    nbx_mg_level& L = mg->lv[l];
    L.N = orders[l];
    L.Nq = orders[l] + 1;
    if (l == 0) {
      L.bp = bp;
    } else {
      nbx_bp_options o = bp->opt;
      o.box.N = orders[l];
      o.precon = NBX_PRECON_JACOBI;
      o.flexible = 0;
      o.fused = 1;
      o.verbose = 0;
      L.bp = nbx_bp_setup(&o, bp->stream);
      if (!L.bp) {
        rc = NBX_ERR_INVALID;
        break;
      }
      L.owned = true;
    }
    nbx_bp* b = L.bp;
    if ((rc = bp_update_jacobi(b, b->lambda))) break;
    const bool last = l + 1 == nlevels;
    if (!last && ((rc = mg_alloc(mg, &L.res, b->Nall, es)) || (rc = mg_alloc(mg, &L.d, b->Nall, es)) ||
                  (rc = mg_alloc(mg, &L.Ad, b->Nall, es))))
      break;
    if (l > 0 && ((rc = mg_alloc(mg, &L.r, b->Nall, last ? sizeof(double) : es)) ||
                  (rc = mg_alloc(mg, &L.x, b->Nall, last ? sizeof(double) : es))))
      break;
    if (!last) {
      const std::vector<double> J = interp_matrix(orders[l], orders[l + 1]);
      if ((rc = mg_alloc(mg, &L.J, (int64_t)J.size()))) break;
      if (hipMemcpy(L.J, J.data(), sizeof(double) * J.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(NBX_ERR_HIP, "bp_mg_setup: copy of J");
        break;
      }
      // lambda_max of the FP64 operator (the handle's z and Ax as scratch)
      if ((rc = power_lmax(b, bp->opt.box.seed, 17 + (uint64_t)l, 20, b->z, b->Ax, &L.lmax))) break;
      cheb_coeffs(L.lmax, lo, hi, degree, L.c1, L.c2);
      if (prec != NBX_F64) {
        if ((rc = bp_prepare_precision(b, prec == NBX_BF16_MFMA ? NBX_BF16 : prec, b->lambda))) break;
        std::vector<float> J32(J.size());
        for (size_t q = 0; q < J.size(); ++q) J32[q] = (float)J[q];
        if ((rc = mg_alloc(mg, &L.J32, (int64_t)J.size())) || (rc = mg_alloc(mg, &L.invD32, b->Nlocal))) break;
        if (hipMemcpy(L.J32, J32.data(), sizeof(float) * J32.size(), hipMemcpyHostToDevice) != hipSuccess) {
          rc = fail(NBX_ERR_HIP, "bp_mg_setup: copy of J");
          break;
        }
        hipLaunchKernelGGL(k_mg_to_f32, mg_grid(b->Nlocal), dim3(256), 0, b->stream, (int64_t)b->Nlocal,
                           b->invDiagA, L.invD32);
      }
    } else {
      // coarse solve: fixed-count Jacobi PCG with device-resident scalars
      if (b->hist_cap < coarse_iters + 2) {
        if (b->hist_rdotr) (void)hipFree(b->hist_rdotr);
        if (b->hist_alpha) (void)hipFree(b->hist_alpha);
        b->hist_rdotr = b->hist_alpha = nullptr;
        b->hist_cap = coarse_iters + 2;
        if (hipMalloc(&b->hist_rdotr, sizeof(double) * b->hist_cap) != hipSuccess ||
            hipMalloc(&b->hist_alpha, sizeof(double) * b->hist_cap) != hipSuccess) {
          rc = fail(NBX_ERR_ALLOC, "bp_mg_setup: coarse history");
          break;
        }
      }
      PcgScalars t{};
      t.rdotz1 = 1.0;
      t.rdotz2 = 0.0;
      t.tol2 = 0.0;
      t.abs_floor = 0.0;
      t.maxit = coarse_iters;
      t.fixed = 1;
      t.flexible = 0;
      if (hipMalloc(&mg->coarse_template, sizeof(PcgScalars)) != hipSuccess ||
          hipMemcpy(mg->coarse_template, &t, sizeof(PcgScalars), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(NBX_ERR_ALLOC, "bp_mg_setup: coarse scalars");
        break;
      }
      if (b->N == 1 && (rc = cg_setup(mg, b))) break;  // order 1: the two-pass coarse CG
    }
  }
  if (rc == NBX_SUCCESS && hipStreamSynchronize(bp->stream) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "bp_mg_setup: stream");
  if (rc != NBX_SUCCESS) {
    mg_release(mg);
    return rc;
  }
  bp->mg = mg;
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_mg_use_graph(nbx_bp* bp, int on)
{
  if (!bp || !bp->mg) return fail(NBX_ERR_INVALID, "bp_mg_use_graph: no multigrid set up");
  (void)hipStreamSynchronize(bp->stream);
  bp->mg->use_graph = on != 0;
  if (!on) mg_drop_graph(bp->mg);
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_mg_set_learned(nbx_bp* bp, int on)
{
  if (!bp || !bp->mg) return fail(NBX_ERR_INVALID, "bp_mg_set_learned: no multigrid set up");
  nbx_mg* mg = bp->mg;
  (void)hipStreamSynchronize(bp->stream);
  if (on) {
    if (!bp->net || !bp->feat)
      return fail(NBX_ERR_INVALID, "bp_mg_set_learned: no network attached (nbx_bp_attach_net)");
    if (bp->N != 7 || bp->Nf != 1)
      return fail(NBX_ERR_UNSUPPORTED, "bp_mg_set_learned: the network is built for N = 7, one field");
    if (!mg->lrf) {
      int rc;
      if ((rc = mg_alloc(mg, &mg->lrf, bp->Nall)) || (rc = mg_alloc(mg, &mg->lzl, bp->Nall))) return rc;
    }
  }
  if (mg->learned != (on != 0)) mg_drop_graph(mg);  // the cycle changes
  mg->learned = on != 0;
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_mg_get_learned(const nbx_bp* bp) { return bp && bp->mg && bp->mg->learned ? 1 : 0; }

extern "C" int nbx_bp_mg_info(const nbx_bp* bp, int level, int* N, double* lmax)
{
  if (!bp || !bp->mg || level < 0 || level >= (int)bp->mg->lv.size())
    return fail(NBX_ERR_INVALID, "bp_mg_info: no such level");
  if (N) *N = bp->mg->lv[level].N;
  if (lmax) *lmax = bp->mg->lv[level].lmax;
  return (int)bp->mg->lv.size();
}

// one V-cycle z = M r (device vectors of the handle's size), synchronous
extern "C" int nbx_bp_mg_apply(nbx_bp* bp, const double* r, double* z)
{
  if (!bp || !r || !z) return fail(NBX_ERR_INVALID, "bp_mg_apply: bad arguments");
  int rc = mg_apply(bp, r, z, nullptr);
  if (rc) return rc;
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

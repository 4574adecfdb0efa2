// BP driver: nekBone's setup and PCG (nekBone/setup.cpp:36-367,
// nekBone/solve.cpp:44-293) on one GPU.
//
// The PCG loop follows BPPCG (solve.cpp:44-113) step for step, with every
// scalar kept on the device: each reduction ends in a one-block kernel that
// also performs the scalar recurrences (TOL, beta, alpha and the stopping
// test), and every kernel of an iteration starts by reading the device
// `done` flag, so the host enqueues iterations in chunks and reads the
// scalars back once per chunk instead of twice per iteration.
//
// Per-iteration kernels (reference call in brackets):
//   1. x += alpha p, r -= alpha Ap of the previous iteration (BPUpdatePCG),
//      z = M r, tile partials of <r,z>_w, <r,r>_w [, <z,Ap>_w]
//      (BPPreconditioner + BPWeightedInnerProduct2 [+ flexible dot]);
//      one-block finish: TOL, stopping test, beta
//   2. p = z + beta p                          (BPScaledAdd)
//   3. Ap = axhelm(p)                          (BPAxOperator: kernel)
//   4. Ap = gs(Ap), Ap[mask] = 0               (BPAxOperator: oogs::start/finish; mask)
//   5. pAp = <p,Ap>_w, one-block finish: alpha = rdotz/pAp (BPWeightedInnerProduct)
// Moving the update into the next iteration's first pass changes no value:
// every element-wise operation is the same single fma in the same order.
// The dots use the fixed reduction tree of nbx_reduce.h.  The CPU oracle
// (nbo_bp_pcg) restates exactly this arithmetic.
#include "nbx_bp_internal.h"
#include "nbx_net.h"
#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace nbx {



__device__ __forceinline__ void pcg_after_dots(PcgScalars* S, double rdotz, double rdotr, double zdotAp,
                                               double* hist_rdotr)
{
  const int it = S->iter + 1;
  S->iter = it;
  if (it > S->maxit) {  // loop end (solve.cpp:76); the last update has been applied
    S->done = 1;
    S->iterations = S->maxit;
    return;
  }
  S->rdotz2 = S->rdotz1;
  S->rdotz1 = rdotz;
  S->rdotr = rdotr;
  S->zdotAp = zdotAp;
  hist_rdotr[it] = rdotr;
  if (it == 1) {
    S->rdotr0 = rdotr;
    S->TOL = fmax(S->tol2 * rdotr, S->abs_floor);  // solve.cpp:85 (abs_floor = tol^2 there)
  }
  if (!isfinite(rdotr)) {
    S->done = 1;
    S->breakdown = 1;
    S->iterations = it - 1;
    return;
  }
  if (rdotr <= S->TOL && !S->fixed) {  // solve.cpp:90
    S->done = 1;
    S->converged = 1;
    S->iterations = it - 1;
    return;
  }
  // learned preconditioner: Jacobi from the next iteration on when the
  // residual has not fallen below stall_ratio of its value stall_window
  // iterations ago (plain Jacobi PCG is the fallback)
  if (S->learned && !S->fallback && it > S->stall_window &&
      !(rdotr <= S->stall_ratio * hist_rdotr[it - S->stall_window])) {
    S->fallback = 1;
    S->fallback_iter = it;
  }
  if (S->flexible)
    S->beta = -S->alpha * S->zdotAp / S->rdotz2;  // solve.cpp:92-95
  else
    S->beta = (it == 1) ? 0.0 : S->rdotz1 / S->rdotz2;  // solve.cpp:97
}

// nodes per lane whose loads the update pass issues together
constexpr int kUpdateBatch = 8;

// 1. [x += alpha p, r -= alpha Ap of the previous iteration]; z = M r
//    (Jacobi: invDiag r; copy: z is r itself); tile partials of the dots
template <bool JAC, bool FLEX, bool DEG>
__global__ __launch_bounds__(256) void k_pcg_update_precon_dots(
    const int32_t N, const int32_t offset, const int Nf, const double* __restrict__ w, const uint8_t* __restrict__ deg,
    const double* __restrict__ invDiag, const double* __restrict__ p, const double* __restrict__ Ap,
    double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, double* __restrict__ partials,
    const PcgScalars* S)
{
  if (S->done) return;
  const bool update = S->iter > 0;  // iteration 1 starts from r = b - A x
  const double alpha = S->alpha;
  constexpr int NV = FLEX ? 3 : 2;
  __shared__ double s4[NV * 4];
  double v[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) v[c] = 0.0;
  const int64_t b0 = (int64_t)blockIdx.x * kRedTile + threadIdx.x;
  if (update && Nf == 1 && (int64_t)(blockIdx.x + 1) * kRedTile <= N) {
    // a full tile of one field after the first iteration (every block but the
    // last): each lane issues the loads of all its nodes before the first
    // store, instead of one node's loads per round trip to memory; the
    // arithmetic and the order of the sums are those of the general loop
    constexpr int kB = kUpdateBatch;
    for (int h = 0; h < kRedPerLane; h += kB) {
      double rr[kB], ap[kB], pp[kB], xx[kB], dg[kB], wi[kB];
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int64_t id = b0 + (h + u) * kRedLanes;
        rr[u] = ldnt(&r[id]);
        ap[u] = ldnt(&Ap[id]);
        pp[u] = ldnt(&p[id]);
        xx[u] = ldnt(&x[id]);
        dg[u] = JAC ? ldnt(&invDiag[id]) : 1.0;
        wi[u] = DEG ? (double)deg[id] : ldnt(&w[id]);  // the weight's source: degree or weight
      }
      // the divisions below wait for their degree loads: keep them after every load is issued
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < kB; ++u) {
        const int64_t id = b0 + (h + u) * kRedLanes;
        if (DEG) wi[u] = 1.0 / wi[u];  // weight<true>: the same IEEE division
        stnt(__builtin_fma(alpha, pp[u], xx[u]), &x[id]);
        const double rf = __builtin_fma(-alpha, ap[u], rr[u]);
        stnt(rf, &r[id]);
        double zf = rf;
        if (JAC) {
          zf = dg[u] * rf;
          stnt(zf, &z[id]);
        }
        v[0] = __builtin_fma(wi[u] * rf, zf, v[0]);
        v[1] = __builtin_fma(wi[u] * rf, rf, v[1]);
        if (FLEX) v[NV - 1] = __builtin_fma(wi[u] * zf, ap[u], v[NV - 1]);
      }
    }
    block_sum256<NV>(v, s4);
    store_partials<NV>(v, partials);
    return;
  }
#pragma unroll
  for (int u = 0; u < kRedPerLane; ++u) {
    const int64_t id = b0 + u * kRedLanes;
    if (id < N) {
      const double wid = weight<DEG>(w, deg, id);
      for (int f = 0; f < Nf; ++f) {
        const int64_t iid = id + (int64_t)f * offset;
        double rf = ldnt(&r[iid]);
        double apf = 0.0;
        if (update || FLEX) apf = ldnt(&Ap[iid]);
        if (update) {
          stnt(__builtin_fma(alpha, ldnt(&p[iid]), ldnt(&x[iid])), &x[iid]);  // BPUpdatePCG, utils.okl:795-807
          rf = __builtin_fma(-alpha, apf, rf);
          stnt(rf, &r[iid]);
        }
        double zf = rf;
        if (JAC) {
          zf = ldnt(&invDiag[id]) * rf;
          stnt(zf, &z[iid]);
        }
        v[0] = __builtin_fma(wid * rf, zf, v[0]);
        v[1] = __builtin_fma(wid * rf, rf, v[1]);
        if (FLEX) v[NV - 1] = __builtin_fma(wid * zf, apf, v[NV - 1]);
      }
    }
  }
  block_sum256<NV>(v, s4);
  store_partials<NV>(v, partials);
}

template <int NV>
__global__ __launch_bounds__(256) void k_pcg_finish_dots(const unsigned int nb, const double* __restrict__ partials,
                                                         PcgScalars* S, double* __restrict__ hist_rdotr)
{
  if (S->done) return;
  __shared__ double s4[NV * 4];
  double v[NV];
  combine_partials<NV>(v, partials, nb, s4);
  if (threadIdx.x == 0) pcg_after_dots(S, v[0], v[1], NV > 2 ? v[NV - 1] : 0.0, hist_rdotr);
}

// 2. p = z + beta p
__global__ __launch_bounds__(256) void k_pcg_pupdate(const int64_t Nall, const double* __restrict__ z,
                                                     double* __restrict__ p, const PcgScalars* S)
{
  if (S->done) return;
  const double beta = S->beta;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < Nall) p[n] = __builtin_fma(beta, p[n], z[n]);
}

// 5. pAp = <p,Ap>_w (tile partials), then alpha = rdotz/pAp
template <bool DEG>
__global__ __launch_bounds__(256) void k_pcg_pap(const int32_t N, const int32_t offset, const int Nf,
                                                 const double* __restrict__ w, const uint8_t* __restrict__ deg,
                                                 const double* __restrict__ p,
                                                 const double* __restrict__ Ap, double* __restrict__ partials,
                                                 const PcgScalars* S)
{
  if (S->done) return;
  __shared__ double s4[4];
  double v[1] = {0.0};
  const int64_t b0 = (int64_t)blockIdx.x * kRedTile + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kRedPerLane; ++u) {
    const int64_t id = b0 + u * kRedLanes;
    if (id < N) {
      const double wid = weight<DEG>(w, deg, id);
      for (int f = 0; f < Nf; ++f) {
        const int64_t iid = id + (int64_t)f * offset;
        v[0] = __builtin_fma(wid * ldnt(&p[iid]), ldnt(&Ap[iid]), v[0]);
      }
    }
  }
  block_sum256<1>(v, s4);
  store_partials<1>(v, partials);
}

__global__ __launch_bounds__(256) void k_pcg_finish_pap(const unsigned int nb, const double* __restrict__ partials,
                                                        PcgScalars* S, double* __restrict__ hist_alpha)
{
  if (S->done) return;
  __shared__ double s4[4];
  double v[1];
  combine_partials<1>(v, partials, nb, s4);
  if (threadIdx.x == 0) {
    S->pAp = v[0];
    const double alpha = S->rdotz1 / v[0];  // solve.cpp:104
    S->alpha = alpha;
    hist_alpha[S->iter] = alpha;
    if (!isfinite(alpha)) {
      S->done = 1;
      S->breakdown = 1;
      S->iterations = S->iter - 1;
    }
  }
}

// x += alpha p; r -= alpha Ap of the previous iteration (BPUpdatePCG)
__global__ __launch_bounds__(256) void k_pcg_update(const int64_t Nall, const double* __restrict__ p,
                                                    const double* __restrict__ Ap, double* __restrict__ x,
                                                    double* __restrict__ r, const PcgScalars* S)
{
  if (S->done || S->iter == 0) return;
  const double alpha = S->alpha;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < Nall) {
    x[n] = __builtin_fma(alpha, ldnt(&p[n]), ldnt(&x[n]));
    r[n] = __builtin_fma(-alpha, ldnt(&Ap[n]), ldnt(&r[n]));
  }
}

// tile partials of <r,z>_w, <r,r>_w, <z,Ap>_w for a given (assembled) z
template <bool DEG>
__global__ __launch_bounds__(256) void k_pcg_dots_z(const int32_t N, const int32_t offset, const int Nf,
                                                    const double* __restrict__ w, const uint8_t* __restrict__ deg,
                                                    const double* __restrict__ r, const double* __restrict__ z,
                                                    const double* __restrict__ Ap, double* __restrict__ partials,
                                                    const PcgScalars* S)
{
  if (S->done) return;
  __shared__ double s4[3 * 4];
  double v[3] = {0.0, 0.0, 0.0};
  const int64_t b0 = (int64_t)blockIdx.x * kRedTile + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kRedPerLane; ++u) {
    const int64_t id = b0 + u * kRedLanes;
    if (id < N) {
      const double wid = weight<DEG>(w, deg, id);
      for (int f = 0; f < Nf; ++f) {
        const int64_t iid = id + (int64_t)f * offset;
        const double rf = ldnt(&r[iid]);
        const double zf = ldnt(&z[iid]);
        v[0] = __builtin_fma(wid * rf, zf, v[0]);
        v[1] = __builtin_fma(wid * rf, rf, v[1]);
        v[2] = __builtin_fma(wid * zf, ldnt(&Ap[iid]), v[2]);
      }
    }
  }
  block_sum256<3>(v, s4);
  store_partials<3>(v, partials);
}

// vecInv (nekBone/kernel/utils.okl:605-613)
__global__ __launch_bounds__(256) void k_vec_inv(const int64_t N, double* __restrict__ v)
{
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) v[n] = 1.0 / v[n];
}

// operand copies for the low-precision operators: float, and bf16 as the
// round-to-nearest-even high half of the float (double -> float -> bf16)
__global__ __launch_bounds__(256) void k_cvt_f32(const int64_t n, const double* __restrict__ a, float* __restrict__ b)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ __launch_bounds__(256) void k_cvt_bf16(const int64_t n, const double* __restrict__ a,
                                                  uint16_t* __restrict__ b)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = bf16_bits((float)a[i]);
}

// x += d
__global__ __launch_bounds__(256) void k_axpy1(const int64_t n, const double* __restrict__ d, double* __restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + d[i];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace nbx

using namespace nbx;

static void bp_release(nbx_bp* bp)
{
  if (!bp) return;
  (void)hipDeviceSynchronize();
  for (void* p : bp->owned) (void)hipFree(p);
  for (void* p : {(void*)bp->ggeo32, (void*)bp->D32, (void*)bp->lambda32, (void*)bp->ggeo16, (void*)bp->lambda16})
    if (p) (void)hipFree(p);
  if (bp->maskIds) (void)hipFree(bp->maskIds);
  if (bp->S) (void)hipFree(bp->S);
  if (bp->hS) (void)hipHostFree(bp->hS);
  if (bp->hist_rdotr) (void)hipFree(bp->hist_rdotr);
  if (bp->hist_alpha) (void)hipFree(bp->hist_alpha);
  for (hipEvent_t ev : bp->prof_ev) (void)hipEventDestroy(ev);
  if (bp->ev0) (void)hipEventDestroy(bp->ev0);
  if (bp->ev1) (void)hipEventDestroy(bp->ev1);
  bp->red.release();
  if (bp->ogs) nbx_ogs_free(bp->ogs);
  if (bp->feat) (void)hipFree(bp->feat);
  if (bp->feat16) (void)hipFree(bp->feat16);
  if (bp->zl) (void)hipFree(bp->zl);
  if (bp->net) nbx::net_detach(bp->net, bp);
  if (bp->mg) nbx::mg_release(bp->mg);
  delete bp;
}

__global__ __launch_bounds__(256) void k_fill_lambda(const int64_t n, const double l1, double* __restrict__ lam)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  lam[i] = 1.0;
  lam[i + n] = l1;
}

static int bp_setup_impl(nbx_bp* bp)
{
  const nbx_bp_options& o = bp->opt;
  const nbx_box& b = o.box;
  bp->N = b.N;
  bp->Nq = b.N + 1;
  bp->Np = bp->Nq * bp->Nq * bp->Nq;
  bp->Nf = o.Nfields;
  if (bp->Nf != 1 && bp->Nf != 3) return fail(NBX_ERR_INVALID, "bp: Nfields must be 1 or 3");
  // JACOBI with Nfields = 3: nekBone stops here (setup.cpp:158-162); the
  // vector operator applies the same ggeo and lambda to every field, so the
  // one diagonal serves all three and the solve proceeds with it.
  const int64_t E = (int64_t)b.NX * b.NY * b.NZ;
  if (E * bp->Np * bp->Nf > INT32_MAX) return fail(NBX_ERR_INVALID, "bp: problem exceeds dlong range");
  bp->E = (int32_t)E;
  bp->Nlocal = (int32_t)(E * bp->Np);
  bp->fieldOffset = bp->Nlocal;  // Np*(Nelements + totalHaloPairs) on one rank (setup.cpp:73-74)
  bp->Nall = (int64_t)bp->Nf * bp->fieldOffset;
  const double lambda1 = o.bpmode ? 0.0 : o.lambda1;  // setup.cpp:51-53
  // the reference never masks, so it projects every lambda1 = 0 problem; with
  // the Dirichlet mask applied the operator is nonsingular and no projection
  // is made
  bp->allNeumann = lambda1 == 0.0 && !o.dirichlet;
  const bool manufactured = o.rhs_kind == 1;

  std::vector<double> zg(bp->Nq), wg(bp->Nq), D(bp->Nq * bp->Nq);
  nbx_gll(bp->N, zg.data(), wg.data(), D.data());
  int rc;
  if ((rc = bp->alloc(&bp->D, D.size())) || (rc = bp->alloc(&bp->ggeo, (int64_t)E * 7 * bp->Np)) ||
      (rc = bp->alloc(&bp->lambda, 2 * (int64_t)bp->fieldOffset)) || (rc = bp->alloc(&bp->invDiagA, bp->Nlocal)) ||
      (rc = bp->alloc(&bp->rhs, bp->Nall)) || (rc = bp->alloc(&bp->p, bp->Nall)) ||
      (rc = bp->alloc(&bp->z, bp->Nall)) || (rc = bp->alloc(&bp->Ap, bp->Nall)) || (rc = bp->alloc(&bp->Ax, bp->Nall)) ||
      (rc = bp->alloc(&bp->elem_partials, E)) || (rc = bp->alloc(&bp->b, bp->Nall)) ||
      (rc = bp->alloc(&bp->d, bp->Nall)))
    return rc;
  NBX_HIP(hipMemcpy(bp->D, D.data(), sizeof(double) * D.size(), hipMemcpyHostToDevice));
  // lambda0 = 1, lambda1 per node (setup.cpp:104-108)
  hipLaunchKernelGGL(k_fill_lambda, grid_for(bp->fieldOffset), dim3(256), 0, bp->stream, (int64_t)bp->fieldOffset,
                     lambda1, bp->lambda);
  NBX_HIP_LAUNCH();

  if (!manufactured) {
    // everything on the device: geometry, global ids, gather-scatter CSR,
    // Dirichlet list, seeded right-hand side (masked, assembled)
    if ((rc = box_geometry_launch(b, bp->ggeo, nullptr, nullptr, nullptr, bp->stream))) return rc;
    int64_t* d_ids = nullptr;
    NBX_HIP(hipMalloc(&d_ids, sizeof(int64_t) * std::max<int64_t>(bp->Nlocal, 1)));
    rc = box_ids_launch(b, o.dirichlet ? 1 : 0, d_ids, bp->stream);
    if (rc == NBX_SUCCESS && hipStreamSynchronize(bp->stream) != hipSuccess) rc = fail(NBX_ERR_HIP, "bp: ids failed");
    if (rc == NBX_SUCCESS) {
      bp->ogs = nbx_ogs_setup_device(bp->Nlocal, d_ids, o.verbose);
      if (!bp->ogs) rc = NBX_ERR_HIP;
    }
    if (rc == NBX_SUCCESS) rc = mask_list_from_ids(bp->Nlocal, d_ids, &bp->maskIds, &bp->Nmasked, bp->stream);
    (void)hipFree(d_ids);
    if (rc) return rc;
    for (int f = 0; f < bp->Nf && rc == NBX_SUCCESS; ++f) {
      double* rf = bp->rhs + (int64_t)f * bp->fieldOffset;
      rc = urand_fill_launch(b.seed, 3 + f, bp->Nlocal, rf, bp->stream);
      if (rc == NBX_SUCCESS) rc = mask_launch(bp->Nmasked, bp->maskIds, rf, nullptr, bp->stream);
    }
    if (rc) return rc;
    rc = gs_launch(bp->ogs, bp->rhs, bp->Nf, bp->fieldOffset, NBX_DOUBLE, NBX_ADD, nullptr, bp->stream);
    if (rc) return rc;
  } else {
    // the manufactured cosine needs node coordinates and the host cos:
    // host mesh, uploaded
    std::vector<double> ggeo((size_t)E * 7 * bp->Np), x(bp->Nlocal), y(bp->Nlocal), z(bp->Nlocal);
    std::vector<nbx_hlong> ids(bp->Nlocal);
    rc = nbx_box_mesh(&b, o.dirichlet ? 1 : 0, x.data(), y.data(), z.data(), ggeo.data(), ids.data(), nullptr);
    if (rc) return rc;
    NBX_HIP(hipMemcpy(bp->ggeo, ggeo.data(), sizeof(double) * ggeo.size(), hipMemcpyHostToDevice));
    bp->ogs = nbx_ogs_setup(bp->Nlocal, ids.data(), o.verbose);
    if (!bp->ogs) return NBX_ERR_HIP;
    std::vector<int32_t> mask;
    for (int32_t i = 0; i < bp->Nlocal; ++i)
      if (ids[i] == 0) mask.push_back(i);
    bp->Nmasked = (int32_t)mask.size();
    NBX_HIP(hipMalloc(&bp->maskIds, sizeof(int32_t) * std::max<size_t>(mask.size(), 1)));
    if (!mask.empty())
      NBX_HIP(hipMemcpy(bp->maskIds, mask.data(), sizeof(int32_t) * mask.size(), hipMemcpyHostToDevice));
    // right-hand side (setup.cpp:77-99), masked, then assembled
    std::vector<double> r(bp->Nall);
    for (int32_t e = 0; e < bp->E; ++e)
      for (int n = 0; n < bp->Np; ++n) {
        const int64_t id = (int64_t)e * bp->Np + n;
        const double JW = ggeo[(size_t)bp->Np * (e * 7 + 0) + n];
        const double v = (3. * M_PI * M_PI + lambda1) * JW * cos(M_PI * x[id]) * cos(M_PI * y[id]) *
                         cos(M_PI * z[id]);
        for (int f = 0; f < bp->Nf; ++f) r[id + (int64_t)f * bp->fieldOffset] = v;
      }
    for (int f = 0; f < bp->Nf; ++f)
      for (int32_t m : mask) r[m + (int64_t)f * bp->fieldOffset] = 0.0;
    NBX_HIP(hipMemcpy(bp->rhs, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice));
    rc = gs_launch(bp->ogs, bp->rhs, bp->Nf, bp->fieldOffset, NBX_DOUBLE, NBX_ADD, nullptr, bp->stream);
    if (rc) return rc;
  }
  if ((rc = bp->red.init(bp->Nlocal, 3))) return rc;
  NBX_HIP(hipMalloc(&bp->S, sizeof(PcgScalars)));
  NBX_HIP(hipHostMalloc((void**)&bp->hS, sizeof(PcgScalars), hipHostMallocDefault));
  NBX_HIP(hipEventCreate(&bp->ev0));
  NBX_HIP(hipEventCreate(&bp->ev1));
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

extern "C" nbx_bp* nbx_bp_setup(const nbx_bp_options* opt, void* stream)
{
  if (!opt) {
    fail(NBX_ERR_INVALID, "bp_setup: null options");
    return nullptr;
  }
  nbx_bp* bp = new nbx_bp();
  bp->opt = *opt;
  bp->stream = as_stream(stream);
  if (bp_setup_impl(bp) != NBX_SUCCESS) {
    std::string msg = nbx_last_error();
    bp_release(bp);
    fail(NBX_ERR_INVALID, "%s", msg.c_str());
    return nullptr;
  }
  return bp;
}

extern "C" void nbx_bp_free(nbx_bp* bp) { bp_release(bp); }

extern "C" int nbx_bp_sizes(const nbx_bp* bp, nbx_dlong* Nelements, nbx_dlong* Np, nbx_dlong* fieldOffset,
                            nbx_dlong* Nmasked)
{
  if (!bp) return fail(NBX_ERR_INVALID, "bp_sizes: null handle");
  if (Nelements) *Nelements = bp->E;
  if (Np) *Np = bp->Np;
  if (fieldOffset) *fieldOffset = bp->fieldOffset;
  if (Nmasked) *Nmasked = bp->Nmasked;
  return NBX_SUCCESS;
}

extern "C" const void* nbx_bp_device_array(const nbx_bp* bp, const char* name)
{
  if (!bp || !name) return nullptr;
  const std::string n(name);
  if (n == "ggeo") return bp->ggeo;
  if (n == "D") return bp->D;
  if (n == "lambda") return bp->lambda;
  if (n == "invDegree") return nbx_ogs_inv_degree(bp->ogs);
  if (n == "invDiagA") return bp->invDiagA;
  if (n == "rhs") return bp->rhs;
  if (n == "maskIds") return bp->maskIds;
  fail(NBX_ERR_INVALID, "bp_device_array: unknown array '%s'", name);
  return nullptr;
}

extern "C" nbx_ogs* nbx_bp_ogs(nbx_bp* bp) { return bp ? bp->ogs : nullptr; }

// Aq = mask(gs(axhelm(q)))  (BPAxOperator without the dot)
int bp_apply(nbx_bp* bp, const double* lambda, const double* q, double* Aq, const int* skip)
{
  const nbx_bp_options& o = bp->opt;
  int rc = axhelm_launch(bp->N, bp->Nf, o.bpmode ? 1 : 0, NBX_F64, bp->E, bp->fieldOffset, nullptr, bp->ggeo,
                         bp->D, lambda, q, Aq, skip, bp->stream);
  if (rc) return rc;
  return gs_add_mask_launch(bp->ogs, Aq, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, skip, bp->stream);
}

extern "C" int nbx_bp_ax_operator(nbx_bp* bp, const double* lambda, const double* q, double* Aq, double* pAp_host)
{
  if (!bp || !q || !Aq) return fail(NBX_ERR_INVALID, "bp_ax_operator: bad arguments");
  if (!lambda) lambda = bp->lambda;
  int rc = bp_apply(bp, lambda, q, Aq, nullptr);
  if (rc) return rc;
  if (pAp_host) {
    rc = launch_wdot(bp->Nlocal, bp->fieldOffset, bp->Nf, nbx_ogs_inv_degree(bp->ogs), q, Aq, false, bp->red,
                     bp->red.result, bp->stream);
    if (rc) return rc;
    NBX_HIP(hipMemcpyAsync(pAp_host, bp->red.result, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
  }
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

// updateJacobi (solve.cpp:146-165): local diagonal, assembled, inverted
int bp_update_jacobi(nbx_bp* bp, const double* lambda)
{
  int rc = nbx_update_jacobi_local(bp->N, bp->E, bp->fieldOffset, nullptr, bp->ggeo, bp->D, lambda, bp->invDiagA,
                                   bp->stream);
  if (rc) return rc;
  rc = gs_launch(bp->ogs, bp->invDiagA, 1, bp->fieldOffset, NBX_DOUBLE, NBX_ADD, nullptr, bp->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vec_inv, grid_for(bp->Nlocal), dim3(256), 0, bp->stream, (int64_t)bp->Nlocal, bp->invDiagA);
  NBX_HIP_LAUNCH();
  // masked nodes are not in any group: their diagonal stays local; they carry
  // r = 0, so z = 0 there whatever the value
  return NBX_SUCCESS;
}

int bp_enqueue_iteration(nbx_bp* bp, const double* lambda, double* r, double* x)
{
  const nbx_bp_options& o = bp->opt;
  const bool jac = o.precon == NBX_PRECON_JACOBI;
  const bool flex = o.flexible != 0;
  const unsigned int nb = (unsigned int)((bp->Nlocal + kRedTile - 1) / kRedTile);
  const double* w = nbx_ogs_inv_degree(bp->ogs);
  const uint8_t* deg = ogs_degree(bp->ogs);
  const double* zsrc = jac ? bp->z : r;
  if (o.precon == NBX_PRECON_LEARNED || o.precon == NBX_PRECON_MG) {
    // update; z = M r (learned: gs(w invDiag (r + s f(r))), Jacobi after a
    // stall; multigrid: one V-cycle); dots; finish
    hipLaunchKernelGGL(k_pcg_update, grid_for(bp->Nall), dim3(256), 0, bp->stream, bp->Nall, bp->p, bp->Ap, x, r,
                       bp->S);
    prof_mark(bp, kProfUpdate);
    int rc;
    if (o.precon == NBX_PRECON_MG) {
      if ((rc = mg_apply(bp, r, bp->z, &bp->S->done))) return rc;
    } else {
      if ((rc = net_apply_learned(bp, r, bp->z, &bp->S->fallback, &bp->S->done))) return rc;
      rc = gs_add_mask_launch(bp->ogs, bp->z, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, &bp->S->done,
                              bp->stream);
      if (rc) return rc;
    }
    prof_mark(bp, kProfPreco);
    if (deg)
      hipLaunchKernelGGL(k_pcg_dots_z<true>, dim3(nb), dim3(256), 0, bp->stream, bp->Nlocal, bp->fieldOffset, bp->Nf,
                         w, deg, r, bp->z, bp->Ap, bp->red.partials, bp->S);
    else
      hipLaunchKernelGGL(k_pcg_dots_z<false>, dim3(nb), dim3(256), 0, bp->stream, bp->Nlocal, bp->fieldOffset,
                         bp->Nf, w, deg, r, bp->z, bp->Ap, bp->red.partials, bp->S);
    hipLaunchKernelGGL(k_pcg_finish_dots<3>, dim3(1), dim3(256), 0, bp->stream, nb, bp->red.partials, bp->S,
                       bp->hist_rdotr);
    NBX_HIP_LAUNCH();
    prof_mark(bp, kProfDot);
    zsrc = bp->z;
  } else {
#define NBX_PD(J, F, DG)                                                                                         \
  hipLaunchKernelGGL((k_pcg_update_precon_dots<J, F, DG>), dim3(nb), dim3(256), 0, bp->stream, bp->Nlocal,        \
                     bp->fieldOffset, bp->Nf, w, deg, bp->invDiagA, bp->p, bp->Ap, x, r, bp->z, bp->red.partials,  \
                     bp->S)
#define NBX_PD2(J, F) \
  if (deg) NBX_PD(J, F, true); else NBX_PD(J, F, false);
  if (jac && flex) { NBX_PD2(true, true) }
  else if (jac) { NBX_PD2(true, false) }
  else if (flex) { NBX_PD2(false, true) }
  else { NBX_PD2(false, false) }
#undef NBX_PD2
#undef NBX_PD
  prof_mark(bp, kProfUpdate);  // x, r update, preconditioner and the dots' tile partials in one pass
  if (flex)
    hipLaunchKernelGGL(k_pcg_finish_dots<3>, dim3(1), dim3(256), 0, bp->stream, nb, bp->red.partials, bp->S,
                       bp->hist_rdotr);
  else
    hipLaunchKernelGGL(k_pcg_finish_dots<2>, dim3(1), dim3(256), 0, bp->stream, nb, bp->red.partials, bp->S,
                       bp->hist_rdotr);
  NBX_HIP_LAUNCH();
  prof_mark(bp, kProfDot);
  }
  if (o.fused || bp->op_prec != NBX_F64 || o.precon == NBX_PRECON_LEARNED || o.precon == NBX_PRECON_MG) {
    // p = z + beta p, Ap = A p and pAp element shares in one pass; then gs + mask
    const void* g = bp->ggeo;
    const void* Dm = bp->D;
    const void* lam = lambda;
    if (bp->op_prec == NBX_F32) {
      g = bp->ggeo32;
      Dm = bp->D32;
      lam = bp->lambda32;
    } else if (bp->op_prec == NBX_BF16) {
      g = bp->ggeo16;
      Dm = bp->D32;
      lam = bp->lambda16;
    }
    int rc = axhelm_pcg_launch(bp->N, bp->Nf, o.bpmode ? 1 : 0, bp->op_prec, bp->E, bp->fieldOffset, g, Dm, lam,
                               zsrc, bp->p, bp->Ap, bp->elem_partials, reinterpret_cast<const PcgFlags*>(bp->S),
                               bp->stream);
    if (rc) return rc;
    prof_mark(bp, kProfAx);  // p = z + beta p, A p and its pAp shares
    rc = gs_add_mask_launch(bp->ogs, bp->Ap, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, &bp->S->done,
                            bp->stream);
    if (rc) return rc;
    prof_mark(bp, kProfGs);
    hipLaunchKernelGGL(k_pcg_finish_pap, dim3(1), dim3(256), 0, bp->stream, (unsigned)bp->E, bp->elem_partials, bp->S,
                       bp->hist_alpha);
    NBX_HIP_LAUNCH();
    prof_mark(bp, kProfDot);
    return NBX_SUCCESS;
  }
  hipLaunchKernelGGL(k_pcg_pupdate, grid_for(bp->Nall), dim3(256), 0, bp->stream, bp->Nall, zsrc, bp->p, bp->S);
  NBX_HIP_LAUNCH();
  prof_mark(bp, kProfUpdate);
  int rc = axhelm_launch(bp->N, bp->Nf, bp->opt.bpmode ? 1 : 0, NBX_F64, bp->E, bp->fieldOffset, nullptr, bp->ggeo,
                         bp->D, lambda, bp->p, bp->Ap, &bp->S->done, bp->stream);
  if (rc) return rc;
  prof_mark(bp, kProfAx);
  rc = gs_add_mask_launch(bp->ogs, bp->Ap, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, &bp->S->done,
                          bp->stream);
  if (rc) return rc;
  prof_mark(bp, kProfGs);
  if (deg)
    hipLaunchKernelGGL(k_pcg_pap<true>, dim3(nb), dim3(256), 0, bp->stream, bp->Nlocal, bp->fieldOffset, bp->Nf, w,
                       deg, bp->p, bp->Ap, bp->red.partials, bp->S);
  else
    hipLaunchKernelGGL(k_pcg_pap<false>, dim3(nb), dim3(256), 0, bp->stream, bp->Nlocal, bp->fieldOffset, bp->Nf, w,
                       deg, bp->p, bp->Ap, bp->red.partials, bp->S);
  hipLaunchKernelGGL(k_pcg_finish_pap, dim3(1), dim3(256), 0, bp->stream, nb, bp->red.partials, bp->S,
                     bp->hist_alpha);
  NBX_HIP_LAUNCH();
  prof_mark(bp, kProfDot);
  return NBX_SUCCESS;
}

// PCG core (BPPCG, solve.cpp:44-113).  initial: r = r - A x first (x may be
// nonzero); otherwise r is taken as the residual of x = 0.  abs_floor is the
// absolute part of the stopping threshold (tol^2 in BPPCG, 0 for inner solves).
static int bp_pcg_core(nbx_bp* bp, const double* lambda, double* r, double* x, double tol, double abs_floor,
                       int maxit, bool initial, bool jacobi_setup, PcgScalars* out)
{
  if (maxit + 2 > bp->hist_cap) {
    NBX_HIP(hipStreamSynchronize(bp->stream));
    if (bp->hist_rdotr) (void)hipFree(bp->hist_rdotr);
    if (bp->hist_alpha) (void)hipFree(bp->hist_alpha);
    bp->hist_cap = maxit + 2;
    NBX_HIP(hipMalloc(&bp->hist_rdotr, sizeof(double) * bp->hist_cap));
    NBX_HIP(hipMalloc(&bp->hist_alpha, sizeof(double) * bp->hist_cap));
  }
  NBX_HIP(hipMemsetAsync(bp->hist_rdotr, 0, sizeof(double) * bp->hist_cap, bp->stream));
  NBX_HIP(hipMemsetAsync(bp->hist_alpha, 0, sizeof(double) * bp->hist_cap, bp->stream));
  PcgScalars s0{};
  s0.rdotz1 = 1.0;  // solve.cpp:58-60
  s0.rdotz2 = 0.0;
  s0.tol2 = tol * tol;
  s0.abs_floor = abs_floor;
  s0.maxit = maxit;
  s0.fixed = tol > 0 ? 0 : 1;  // solve() sets FIXED ITERATION COUNT when tol <= 0 (main.cpp:17-22)
  s0.flexible = bp->opt.flexible ? 1 : 0;
  if (bp->opt.precon == NBX_PRECON_LEARNED) {
    if (!bp->net) return fail(NBX_ERR_INVALID, "bp_pcg: LEARNED preconditioner without an attached network");
    s0.learned = 1;
    s0.flexible = 1;  // nonlinear, switching preconditioner: flexible beta (solve.cpp:92-95)
    s0.stall_window = bp->stall_window;
    s0.stall_ratio = bp->stall_ratio;
  }
  if (bp->opt.precon == NBX_PRECON_MG) {
    if (!bp->mg) return fail(NBX_ERR_INVALID, "bp_pcg: MG preconditioner without nbx_bp_mg_setup");
    s0.flexible = 1;  // the coarse solve is a fixed number of CG steps: flexible beta
  }
  *bp->hS = s0;
  NBX_HIP(hipMemcpyAsync(bp->S, bp->hS, sizeof(PcgScalars), hipMemcpyHostToDevice, bp->stream));
  int rc;
  prof_mark(bp, kProfNone);  // start of the timed phases
  if (initial) {  // Ax = A x; r = r - Ax  (solve.cpp:68-69)
    if ((rc = axhelm_launch(bp->N, bp->Nf, bp->opt.bpmode ? 1 : 0, NBX_F64, bp->E, bp->fieldOffset, nullptr,
                            bp->ggeo, bp->D, lambda, x, bp->Ax, nullptr, bp->stream)))
      return rc;
    prof_mark(bp, kProfAx);
    if ((rc = gs_add_mask_launch(bp->ogs, bp->Ax, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, nullptr,
                                 bp->stream)))
      return rc;
    prof_mark(bp, kProfGs);
    if ((rc = nbx_scaled_add((nbx_dlong)bp->Nall, -1.0, bp->Ax, 1.0, r, bp->stream))) return rc;
    prof_mark(bp, kProfNone);  // the reference does not time this BPScaledAdd
  }
  if (jacobi_setup && (bp->opt.precon == NBX_PRECON_JACOBI || bp->opt.precon == NBX_PRECON_LEARNED) &&
      (rc = bp_update_jacobi(bp, lambda)))
    return rc;  // solve.cpp:72
  prof_mark(bp, kProfPreco);  // updateJacobi is timed as "preco" (solve.cpp:71-73)
  // iterations in chunks; the device flag ends the work, the host only polls
  // each poll is a host wait; a multigrid iteration is long, so it is polled
  // more often (the kernels of iterations past convergence exit at once)
  const int chunk = bp->opt.precon == NBX_PRECON_MG ? 4 : 16;
  int enqueued = 0;
  for (;;) {
    const int n = std::min(chunk, maxit + 1 - enqueued);
    for (int c = 0; c < n; ++c)
      if ((rc = bp_enqueue_iteration(bp, lambda, r, x))) return rc;
    enqueued += n;
    NBX_HIP(hipMemcpyAsync(bp->hS, bp->S, sizeof(PcgScalars), hipMemcpyDeviceToHost, bp->stream));
    NBX_HIP(hipStreamSynchronize(bp->stream));
    if ((rc = prof_collect(bp))) return rc;
    if (bp->hS->done || enqueued > maxit) break;
  }
  *out = *bp->hS;
  bp->last_iters = out->iterations;
  return NBX_SUCCESS;
}

// q <- q - mean(q), the unweighted mean over the E-vector and its fields
// (BPZeroMean, nekBone/solve.cpp:115-134); stream-ordered, no host wait
extern "C" int nbx_bp_zero_mean(nbx_bp* bp, double* q)
{
  if (!bp || !q) return fail(NBX_ERR_INVALID, "bp_zero_mean: bad arguments");
  return launch_zero_mean(bp->Nlocal, bp->fieldOffset, bp->Nf, (double)bp->E * bp->Nf * bp->Np, q, bp->red,
                          bp->stream);
}

extern "C" int nbx_bp_all_neumann(const nbx_bp* bp) { return bp && bp->allNeumann ? 1 : 0; }

extern "C" int nbx_bp_pcg(nbx_bp* bp, const double* lambda, double* r, double* x, double tol, int maxit,
                          nbx_pcg_stats* stats)
{
  if (!bp || !r || !x) return fail(NBX_ERR_INVALID, "bp_pcg: bad arguments");
  if (maxit < 0) return fail(NBX_ERR_INVALID, "bp_pcg: maxit < 0");
  if (!lambda) lambda = bp->lambda;
  bp->op_prec = NBX_F64;
  NBX_HIP(hipEventRecord(bp->ev0, bp->stream));
  PcgScalars S{};
  int rc;
  if (bp->allNeumann && (rc = nbx_bp_zero_mean(bp, r))) return rc;  // main.cpp:24-25
  if ((rc = bp_pcg_core(bp, lambda, r, x, tol, tol * tol, maxit, true, true, &S))) return rc;
  if (bp->allNeumann && (rc = nbx_bp_zero_mean(bp, x))) return rc;  // main.cpp:30-31
  NBX_HIP(hipEventRecord(bp->ev1, bp->stream));
  NBX_HIP(hipEventSynchronize(bp->ev1));
  float ms = 0.f;
  NBX_HIP(hipEventElapsedTime(&ms, bp->ev0, bp->ev1));
  if (stats) {
    stats->iterations = S.iterations;
    stats->converged = S.converged;
    stats->rdotr0 = S.rdotr0;
    stats->rdotr = S.rdotr;
    stats->elapsed_s = ms * 1e-3;
    stats->cycles = 1;
    stats->fallback_iteration = S.fallback ? S.fallback_iter : 0;
  }
  if (S.breakdown) return fail(NBX_ERR_DIVERGED, "bp_pcg: breakdown at iteration %d", S.iter);
  return S.iterations;
}

int bp_prepare_precision(nbx_bp* bp, nbx_prec prec, const double* lambda)
{
  const int64_t ng = (int64_t)bp->E * 7 * bp->Np, nl = 2 * (int64_t)bp->fieldOffset;
  if (prec == NBX_F32 || prec == NBX_BF16) {
    if (!bp->D32) {
      NBX_HIP(hipMalloc(&bp->D32, sizeof(float) * bp->Nq * bp->Nq));
      hipLaunchKernelGGL(k_cvt_f32, grid_for(bp->Nq * bp->Nq), dim3(256), 0, bp->stream, (int64_t)bp->Nq * bp->Nq,
                         bp->D, bp->D32);
    }
  }
  if (prec == NBX_F32) {
    if (!bp->ggeo32) {
      NBX_HIP(hipMalloc(&bp->ggeo32, sizeof(float) * ng));
      NBX_HIP(hipMalloc(&bp->lambda32, sizeof(float) * nl));
      hipLaunchKernelGGL(k_cvt_f32, grid_for(ng), dim3(256), 0, bp->stream, ng, bp->ggeo, bp->ggeo32);
    }
    hipLaunchKernelGGL(k_cvt_f32, grid_for(nl), dim3(256), 0, bp->stream, nl, lambda, bp->lambda32);
  } else if (prec == NBX_BF16) {
    if (!bp->ggeo16) {
      NBX_HIP(hipMalloc(&bp->ggeo16, sizeof(uint16_t) * ng));
      NBX_HIP(hipMalloc(&bp->lambda16, sizeof(uint16_t) * nl));
      hipLaunchKernelGGL(k_cvt_bf16, grid_for(ng), dim3(256), 0, bp->stream, ng, bp->ggeo, bp->ggeo16);
    }
    hipLaunchKernelGGL(k_cvt_bf16, grid_for(nl), dim3(256), 0, bp->stream, nl, lambda, bp->lambda16);
  } else if (prec != NBX_F64) {
    return fail(NBX_ERR_UNSUPPORTED, "bp: operator precision %d", (int)prec);
  }
  NBX_HIP_LAUNCH();
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_pcg_mixed(nbx_bp* bp, const double* lambda, double* r, double* x, double tol, int maxit,
                                nbx_prec op_prec, double inner_tol, int max_cycles, nbx_pcg_stats* stats)
{
  if (!bp || !r || !x || tol <= 0 || inner_tol <= 0 || max_cycles < 1 || maxit < 0)
    return fail(NBX_ERR_INVALID, "bp_pcg_mixed: bad arguments");
  // The refinement stops on the true residual b - A x.  On the singular
  // all-Neumann problem only the part of b in the range of A can be removed,
  // and BPZeroMean's unweighted E-vector mean (nekBone/solve.cpp:115-134) is
  // not the null-space component of the weighted inner product, so that test
  // need never be met: refused, as the multigrid setup refuses it.
  if (bp->allNeumann)
    return fail(NBX_ERR_UNSUPPORTED, "bp_pcg_mixed: all-Neumann problem (no mask, lambda1 = 0); use nbx_bp_pcg");
  if (!lambda) lambda = bp->lambda;
  int rc = bp_prepare_precision(bp, op_prec, lambda);
  if (rc) return rc;
  NBX_HIP(hipEventRecord(bp->ev0, bp->stream));
  NBX_HIP(hipMemcpyAsync(bp->b, r, sizeof(double) * bp->Nall, hipMemcpyDeviceToDevice, bp->stream));
  const double* w = nbx_ogs_inv_degree(bp->ogs);
  int total = 0, cycles = 0;
  double rr0 = 0.0, rr = 0.0;
  bool conv = false, breakdown = false;
  for (;;) {
    // true residual in FP64: r = b - A x
    NBX_HIP(hipMemcpyAsync(r, bp->b, sizeof(double) * bp->Nall, hipMemcpyDeviceToDevice, bp->stream));
    if ((rc = bp_apply(bp, lambda, x, bp->Ax, nullptr))) return rc;
    if ((rc = nbx_scaled_add((nbx_dlong)bp->Nall, -1.0, bp->Ax, 1.0, r, bp->stream))) return rc;
    if ((rc = launch_wdot(bp->Nlocal, bp->fieldOffset, bp->Nf, w, r, r, false, bp->red, bp->red.result, bp->stream)))
      return rc;
    NBX_HIP(hipMemcpyAsync(&bp->hS->rdotr, bp->red.result, sizeof(double), hipMemcpyDeviceToHost, bp->stream));
    NBX_HIP(hipStreamSynchronize(bp->stream));
    rr = bp->hS->rdotr;
    if (cycles == 0) rr0 = rr;
    if (rr <= tol * tol * rr0) {
      conv = true;
      break;
    }
    if (cycles >= max_cycles || total >= maxit || breakdown) break;
    // inner solve A_lo d = r from d = 0 (operator in op_prec, vectors, dots and updates FP64)
    NBX_HIP(hipMemsetAsync(bp->d, 0, sizeof(double) * bp->Nall, bp->stream));
    bp->op_prec = op_prec;
    PcgScalars S{};
    rc = bp_pcg_core(bp, lambda, r, bp->d, inner_tol, 0.0, maxit - total, false, cycles == 0, &S);
    bp->op_prec = NBX_F64;
    if (rc) return rc;
    total += S.iterations;
    breakdown = S.breakdown != 0;
    hipLaunchKernelGGL(k_axpy1, grid_for(bp->Nall), dim3(256), 0, bp->stream, bp->Nall, bp->d, x);
    NBX_HIP_LAUNCH();
    ++cycles;
    if (S.iterations == 0) break;  // no progress possible
  }
  NBX_HIP(hipEventRecord(bp->ev1, bp->stream));
  NBX_HIP(hipEventSynchronize(bp->ev1));
  float ms = 0.f;
  NBX_HIP(hipEventElapsedTime(&ms, bp->ev0, bp->ev1));
  if (stats) {
    stats->iterations = total;
    stats->converged = conv ? 1 : 0;
    stats->rdotr0 = rr0;
    stats->rdotr = rr;
    stats->elapsed_s = ms * 1e-3;
    stats->cycles = cycles;
    stats->fallback_iteration = 0;
  }
  return total;
}

namespace nbx {

void prof_mark(nbx_bp* bp, int phase)
{
  if (!bp->prof) return;
  if (bp->prof_used == (int)bp->prof_ev.size()) {  // pool full: wait, credit, keep the last boundary
    if (hipStreamSynchronize(bp->stream) != hipSuccess || prof_collect(bp) != NBX_SUCCESS) return;
  }
  if (hipEventRecord(bp->prof_ev[bp->prof_used], bp->stream) != hipSuccess) return;
  bp->prof_tag[bp->prof_used] = phase;
  ++bp->prof_used;
}

int prof_collect(nbx_bp* bp)
{
  if (!bp->prof || bp->prof_used < 2) return NBX_SUCCESS;
  for (int k = 1; k < bp->prof_used; ++k) {
    if (bp->prof_tag[k] < 0) continue;
    float ms = 0.f;
    NBX_HIP(hipEventElapsedTime(&ms, bp->prof_ev[k - 1], bp->prof_ev[k]));
    bp->prof_t[bp->prof_tag[k]] += 1e-3 * ms;
  }
  std::swap(bp->prof_ev[0], bp->prof_ev[bp->prof_used - 1]);
  bp->prof_tag[0] = bp->prof_tag[bp->prof_used - 1];
  bp->prof_used = 1;
  return NBX_SUCCESS;
}

}  // namespace nbx

extern "C" int nbx_bp_set_profiling(nbx_bp* bp, int on)
{
  if (!bp) return fail(NBX_ERR_INVALID, "bp_set_profiling: null handle");
  NBX_HIP(hipStreamSynchronize(bp->stream));
  if (on && bp->prof_ev.empty()) {
    bp->prof_ev.resize(512);
    bp->prof_tag.assign(512, kProfNone);
    for (auto& ev : bp->prof_ev) NBX_HIP(hipEventCreate(&ev));
  }
  bp->prof = on != 0;
  bp->prof_used = 0;
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_profile(nbx_bp* bp, double* seconds, int reset)
{
  if (!bp || !seconds) return fail(NBX_ERR_INVALID, "bp_profile: bad arguments");
  NBX_HIP(hipStreamSynchronize(bp->stream));
  int rc = prof_collect(bp);
  if (rc) return rc;
  for (int q = 0; q < 5; ++q) seconds[q] = bp->prof_t[q];
  if (reset) {
    for (double& t : bp->prof_t) t = 0.0;
    bp->prof_used = 0;
  }
  return NBX_SUCCESS;
}

extern "C" int nbx_bp_history(const nbx_bp* bp, double* rdotr, double* alpha, int capacity)
{
  if (!bp || capacity < 0) return fail(NBX_ERR_INVALID, "bp_history: bad arguments");
  const int n = std::min(capacity, std::min(bp->last_iters + 2, bp->hist_cap));
  if (n <= 0) return 0;
  NBX_HIP(hipStreamSynchronize(bp->stream));
  if (rdotr) NBX_HIP(hipMemcpy(rdotr, bp->hist_rdotr, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (alpha) NBX_HIP(hipMemcpy(alpha, bp->hist_alpha, sizeof(double) * n, hipMemcpyDeviceToHost));
  return n;
}

int bp_solve_plain(nbx_bp* bp, nbx_precon precon, double* r, double* x, double tol, int maxit, int* iterations)
{
  const nbx_precon keep = bp->opt.precon;
  const int keepFlex = bp->opt.flexible;
  bp->opt.precon = precon;
  bp->opt.flexible = 0;
  bp->op_prec = NBX_F64;
  PcgScalars S{};
  const int rc = bp_pcg_core(bp, bp->lambda, r, x, tol, tol * tol, maxit, true, true, &S);
  bp->opt.precon = keep;
  bp->opt.flexible = keepFlex;
  if (iterations) *iterations = S.iterations;
  return rc;
}

extern "C" int nbx_bp_make_rhs(nbx_bp* bp, uint64_t seed, double* out)
{
  if (!bp || !out) return fail(NBX_ERR_INVALID, "bp_make_rhs: bad arguments");
  int rc = bp_make_rhs(bp, seed, out);
  if (rc) return rc;
  NBX_HIP(hipStreamSynchronize(bp->stream));
  return NBX_SUCCESS;
}

// drawn on the device (the same splitmix64 values as nbx_urand_fill), then
// assembled and masked: masked nodes belong to no group, so this equals the
// setup's mask-then-assemble order
int bp_make_rhs(nbx_bp* bp, uint64_t seed, double* out)
{
  for (int f = 0; f < bp->Nf; ++f) {
    const int rc = urand_fill_launch(seed, 3 + f, bp->Nlocal, out + (int64_t)f * bp->fieldOffset, bp->stream);
    if (rc) return rc;
  }
  return gs_add_mask_launch(bp->ogs, out, bp->Nf, bp->fieldOffset, bp->Nmasked, bp->maskIds, nullptr, bp->stream);
}
